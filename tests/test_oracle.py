"""Pins of the CPU oracle (CPU only).

Each check comes from outside the oracle: published known-answer vectors,
closed forms recomputed from first principles, brute force on small inputs,
finite differences, or an independent library (scikit-learn).
"""
import math

import numpy as np
import pytest

import graphgen
import orc


# ---- Philox4x32-10: Random123 known-answer vectors (kat_vectors, philox4x32 10 rounds)
@pytest.mark.parametrize("ctr,key,expect", [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
])
def test_philox_kat(ctr, key, expect):
    assert [int(x) for x in orc.philox(ctr, key)] == expect


# ---- CSR: brute force adjacency, neighbours in input order after undirecting
def test_csr_matches_brute_force():
    rng = np.random.default_rng(0)
    V, m = 50, 300
    src = rng.integers(0, V, m).astype(np.uint32)
    dst = rng.integers(0, V, m).astype(np.uint32)
    w = rng.random(m).astype(np.float32)
    rp, col, wo = orc.build_csr(V, src, dst, w, True)
    adj = [[] for _ in range(V)]
    for i in range(m):  # edge i then its reverse (include/graph.hpp:741-772)
        adj[src[i]].append((dst[i], w[i]))
        adj[dst[i]].append((src[i], w[i]))
    for v in range(V):
        got = list(zip(col[rp[v]:rp[v + 1]].tolist(), wo[rp[v]:rp[v + 1]].tolist()))
        assert got == [(int(a), float(b)) for a, b in adj[v]]


def test_csr_rejects_out_of_range():
    with pytest.raises(ValueError):
        orc.build_csr(4, np.array([0, 5], np.uint32), np.array([1, 2], np.uint32), np.ones(2, np.float32))


# ---- alias tables: exact identity P(j) = (t_j + sum_{alias_i = j} (2^32 - t_i)) / (2^32 n)
def alias_probs(thresh, alias_idx):
    n = len(thresh)
    t = thresh.astype(np.float64)
    keep = t / 2.0 ** 32
    p = keep / n
    np.add.at(p, alias_idx.astype(np.int64), (1.0 - keep) / n)
    return p


@pytest.mark.parametrize("case", ["uniform", "skewed", "zeros", "single", "pair", "equal", "huge"])
def test_alias_identity(case):
    rng = np.random.default_rng(hash(case) % 2 ** 32)
    w = {
        "uniform": lambda: rng.random(37),
        "skewed": lambda: rng.lognormal(0, 3, 200),
        "zeros": lambda: np.where(rng.random(50) < 0.5, 0.0, rng.random(50)),
        "single": lambda: np.array([0.3]),
        "pair": lambda: np.array([0.9, 0.1]),
        "equal": lambda: np.full(64, 0.25),
        "huge": lambda: rng.random(5000) ** 8,
    }[case]().astype(np.float32)
    rp = np.array([0, len(w)], np.uint64)
    t, a = orc.build_alias(1, rp, w)
    p = alias_probs(t, a)
    target = w.astype(np.float64) / w.astype(np.float64).sum()
    assert np.all(a < len(w))
    assert np.max(np.abs(p - target)) <= len(w) * 2.0 ** -32 + 1e-12


@pytest.mark.parametrize("n", [3, 64, 4097, 20_000])
def test_alias_integer_shares(n):
    """Every column's probability is its integer share q_j / (n 2^32) with
    q_j = floor(w_j n 2^32 / sum w): off w_j / sum w by less than one unit
    of 2^-32 / n, except the largest column, which also takes the rounding
    remainder (at most n units).  Rows longer than 4096 exercise the
    chunked sum."""
    rng = np.random.default_rng(n)
    w = rng.lognormal(0, 2, n).astype(np.float32)
    t, a = orc.build_alias(1, np.array([0, n], np.uint64), w)
    p = alias_probs(t, a)
    target = w.astype(np.float64) / w.astype(np.float64).sum()
    unit = 2.0 ** -32 / n
    jmax = int(np.argmax(target))
    err = np.abs(p - target)
    assert err[jmax] <= (n + 2) * unit
    err[jmax] = 0
    assert err.max() <= 1.01 * unit + 1e-14 * target.max()


def test_unigram_law():
    rng = np.random.default_rng(6)
    deg = rng.integers(1, 300, 400).astype(np.uint32)
    inv = (deg >= 5).astype(np.uint8)
    t, a = orc.unigram(deg, inv)
    target = np.where(inv == 1, deg.astype(np.float64) ** 0.75, 0.0)
    target /= target.sum()
    assert np.max(np.abs(alias_probs(t, a) - target)) <= 400 * 2.0 ** -32 + 1e-12


def test_alias_all_zero_is_uniform():
    w = np.zeros(9, np.float32)
    t, a = orc.build_alias(1, np.array([0, 9], np.uint64), w)
    assert np.allclose(alias_probs(t, a), 1 / 9, atol=1e-9)


def test_alias_sampling_chi_square():
    """Draw columns the way the walk does (hi32(r0*n), coin r1) from Philox
    and compare with the weights by a chi-square test."""
    rng = np.random.default_rng(4)
    w = rng.random(12).astype(np.float32) + 0.05
    t, a = orc.build_alias(1, np.array([0, 12], np.uint64), w)
    counts = np.zeros(12)
    N = 60_000
    for i in range(N):
        r = orc.philox([i, 0, 7, 1], [11, 0])
        j = (int(r[0]) * 12) >> 32
        counts[j if int(r[1]) < int(t[j]) else int(a[j])] += 1
    exp = N * w / w.sum()
    chi2 = float(((counts - exp) ** 2 / exp).sum())
    assert chi2 < 40.0  # 11 dof, p ~ 5e-5


# ---- walker statistics: Shannon entropy and Pearson R from the histogram
def closed_form_series(path):
    """H_n (bits) of the visit histogram and Pearson R of (H_i, S_i), i <= n."""
    Hs, Ss, Rs = [0.0], [1.0], []
    counts = {path[0]: 1}
    for n in range(2, len(path) + 1):
        counts[path[n - 1]] = counts.get(path[n - 1], 0) + 1
        H = -sum(c / n * math.log2(c / n) for c in counts.values())
        Hs.append(H)
        Ss.append(float(n))
        h, s = np.array(Hs), np.array(Ss)
        cov = ((h - h.mean()) * (s - s.mean())).mean()
        den = math.sqrt(((h - h.mean()) ** 2).mean() * ((s - s.mean()) ** 2).mean())
        Rs.append(cov / den if den > 0 else 0.0)
    return np.array(Hs[1:]), np.array(Rs)


def prior_visits(path):
    """Prior occurrences of path[i] in path[:i], for i >= 1."""
    return [path[:i].count(path[i]) for i in range(1, len(path))]


@pytest.mark.parametrize("seed", range(6))
def test_entropy_and_pearson_closed_form(seed):
    rng = np.random.default_rng(seed)
    path = [int(x) for x in rng.integers(0, 6 + 3 * seed, 60)]
    f = prior_visits(path)
    H, R = orc.corr_exact(f)
    Hc, Rc = closed_form_series(path)
    assert np.allclose(H, Hc, rtol=0, atol=1e-12)
    mask = np.abs(Rc) > 1e-6
    assert np.allclose(R[mask], Rc[mask], rtol=1e-9, atol=1e-9)
    Hr, Rr = orc.corr_reference(f)  # include/walk.hpp:296-366 as written
    assert np.allclose(Hr, H, rtol=0, atol=1e-11)
    assert np.allclose(Rr[mask], R[mask], rtol=1e-8, atol=1e-8)


def test_entropy_special_cases():
    # all distinct: H = log2(S)
    H, _ = orc.corr_exact([0] * 30)
    assert np.allclose(H, np.log2(np.arange(2, 32)), atol=1e-12)
    # back and forth between two vertices: H of (ceil(S/2), floor(S/2))
    path = [0, 1] * 15
    H, _ = orc.corr_exact(prior_visits(path))
    for i, h in enumerate(H):
        n = i + 2
        a, b = (n + 1) // 2, n // 2
        assert abs(h - (-(a / n) * math.log2(a / n) - (b / n) * math.log2(b / n))) < 1e-12


# ---- walks: graph invariants and the stop rule
@pytest.fixture(scope="module")
def er_graph():
    s, d, w = graphgen.erdos_renyi(2000, 8000, seed=3)
    rp, col, wo = orc.build_csr(2000, s, d, w, True)
    t, a = orc.build_alias(2000, rp, wo)
    return rp, col, wo, t, a


def test_walks_follow_edges_and_stop_rule(er_graph):
    rp, col, wo, t, a = er_graph
    off, tok, wid, steps = orc.walks(2000, rp, col, t, a, None, 2000, 20, 1000, 1, 99)
    edges = set()
    for v in range(2000):
        for u in col[rp[v]:rp[v + 1]]:
            edges.add((v, int(u)))
    assert steps == int(off[-1]) - len(wid)
    for i in range(0, len(wid), 37):
        p = tok[off[i]:off[i + 1]].tolist()
        assert p[0] == wid[i] % 2000
        assert all((p[j], p[j + 1]) in edges for j in range(len(p) - 1))
        _, R = closed_form_series(p)
        # first step > min_L whose R^2 < 0.999 or R < 0 ends the walk
        stops = [s for s in range(21, len(p)) if R[s - 1] ** 2 < 0.999 or R[s - 1] < 0]
        assert stops and stops[0] == len(p) - 1


def test_walks_skip_flagged_and_are_deterministic(er_graph):
    rp, col, wo, t, a = er_graph
    flags = np.zeros(2000, np.uint8)
    flags[::3] = 1
    off, tok, wid, _ = orc.walks(2000, rp, col, t, a, flags, 2000, 20, 1000, 2, 5, threads=4)
    assert not np.any(wid % 3 == 0)
    off2, tok2, wid2, _ = orc.walks(2000, rp, col, t, a, flags, 2000, 20, 1000, 2, 5, threads=1)
    assert np.array_equal(tok, tok2) and np.array_equal(off, off2)


def test_walk_edge_cases():
    # isolated vertex 2 -> length-1 walk; degree-2 vertex with itself first -> ends
    src = np.array([0, 1, 3], np.uint32)
    dst = np.array([1, 0, 3], np.uint32)
    w = np.ones(3, np.float32)
    rp, col, wo = orc.build_csr(4, src, dst, w, True)
    t, a = orc.build_alias(4, rp, wo)
    off, tok, wid, _ = orc.walks(4, rp, col, t, a, None, 4, 20, 1000, 1, 1)
    lens = np.diff(off.astype(np.int64))
    assert lens[2] == 1           # degree 0
    assert lens[3] == 1           # self loop only: degree 2, first neighbour itself
    assert lens[0] == 22 and lens[1] == 22
    # walk cap
    off, tok, wid, _ = orc.walks(4, rp, col, t, a, None, 4, 20, 7, 1, 1)
    assert np.diff(off.astype(np.int64))[0] == 7


# ---- the reference's common-neighbour transition (compat mode)
def test_tanh_restated_matches_libm():
    xs = np.concatenate([np.logspace(-12, np.log10(19.99), 3000), [0.0, 0.25, 0.2500001, 19.999, 20.0, 25.0, 1e300]])
    for x in np.concatenate([xs, -xs]):
        got, ref = orc.tanh_restated(x), math.tanh(x)
        if abs(x) >= 20.0:
            assert got == math.copysign(1.0, x)
        else:
            assert abs(got - ref) <= 4 * np.spacing(abs(ref)), (x, got, ref)


def _directed_star():
    """Vertex 0 -> 1..4 with edge data 0.5, 1.5, 2.5, 0.25 (common
    neighbours 0, 1, 2, 0); out-degrees of 1..4 are 1, 2, 3, 6 and none
    leads back to 0, so a walk revisits 0 at step 2 only by a trace."""
    src, dst, w = [0, 0, 0, 0], [1, 2, 3, 4], [0.5, 1.5, 2.5, 0.25]
    outs = {1: [5], 2: [5, 6], 3: [5, 6, 7], 4: [5, 6, 7, 8, 9, 10]}
    for v, us in outs.items():
        src += [v] * len(us)
        dst += us
        w += [1.0] * len(us)
    for v in range(5, 11):  # sinks walk on among themselves
        src += [v, v]
        dst += [5 + (v - 4) % 6, 5 + (v - 3) % 6]
        w += [1.0, 1.0]
    return (np.array(src, np.uint32), np.array(dst, np.uint32), np.array(w, np.float32), 11)


def test_compat_walk_first_step_matches_closed_form():
    """From vertex 0 (out-degree 4): the neighbour is uniform, and the walk
    traces back with probability 1 - tanh((1 / (4 - cn)) * max(4, d) / min(4, d))
    (include/walk.hpp:997-1019), tanh from libm here."""
    s, d, w, V = _directed_star()
    rp, col, wo = orc.build_csr(V, s, d, w, False)
    t, a = orc.build_alias(V, rp, wo)
    M = 40000
    flags = np.ones(V, np.uint8)
    flags[0] = 0
    off, tok, wid, _ = orc.walks(V, rp, col, t, a, flags, V * M, 20, 1000, 1, 77, cn_data=wo)
    assert len(wid) == M
    starts = off[:-1].astype(np.int64)
    first, second = tok[starts + 1], tok[starts + 2]
    deg = np.diff(rp.astype(np.int64))
    cn = {1: 0, 2: 1, 3: 2, 4: 0}
    counts = np.array([(first == j).sum() for j in (1, 2, 3, 4)])
    assert counts.sum() == M
    chi2 = ((counts - M / 4) ** 2 / (M / 4)).sum()
    assert chi2 < 16.3  # 3 degrees of freedom, p = 0.001
    for j in (1, 2, 3, 4):
        p = (1.0 / (4 - cn[j])) * max(4, deg[j]) / min(4, deg[j])
        q = 1.0 - math.tanh(p)
        n = (first == j).sum()
        k = (second[first == j] == 0).sum()
        sd = math.sqrt(q * (1 - q) / n)
        assert abs(k / n - q) < 4.5 * sd, (j, k / n, q)


def test_compat_walks_follow_edges_or_trace_and_stop_rule():
    """Random directed graph: every step is an edge or a trace back to the
    vertex two before; the stop rule ends each walk at its first allowed
    step; replays are thread-count independent."""
    rng = np.random.default_rng(8)
    V, m = 1500, 9000
    src = rng.integers(0, V, m).astype(np.uint32)
    dst = rng.integers(0, V, m).astype(np.uint32)
    w = (rng.random(m) * 4).astype(np.float32)
    rp, col, wo = orc.build_csr(V, src, dst, w, False)
    t, a = orc.build_alias(V, rp, wo)
    off, tok, wid, _ = orc.walks(V, rp, col, t, a, None, V, 20, 1000, 1, 3, threads=4, cn_data=wo)
    off1, tok1, wid1, _ = orc.walks(V, rp, col, t, a, None, V, 20, 1000, 1, 3, threads=1, cn_data=wo)
    assert np.array_equal(tok, tok1) and np.array_equal(off, off1)
    edges = set()
    for v in range(V):
        for u in col[rp[v]:rp[v + 1]]:
            edges.add((v, int(u)))
    traces = 0
    for i in range(0, len(wid), 7):
        p = tok[off[i]:off[i + 1]].tolist()
        for j in range(len(p) - 1):
            if (p[j], p[j + 1]) not in edges:
                assert j >= 1 and p[j + 1] == p[j - 1], (i, j)
                traces += 1
        if len(p) > 1 and deg_ok(rp, p[-1]):
            _, R = closed_form_series(p)
            stops = [s for s in range(21, len(p)) if R[s - 1] ** 2 < 0.999 or R[s - 1] < 0]
            assert stops and stops[0] == len(p) - 1
    assert traces > 0


def deg_ok(rp, v):
    """The walk ended by the stop rule (not by a degree-0 vertex)."""
    return rp[v + 1] > rp[v]


# ---- vocabulary, thresholds, initialisation
def test_vocab_order_and_thresholds():
    rng = np.random.default_rng(2)
    deg = rng.integers(0, 40, 500).astype(np.uint32)
    order, rank, inv, vs, tw = orc.vocab(deg, 5)
    assert np.array_equal(order, np.lexsort((np.arange(500), -deg.astype(np.int64))))
    assert np.array_equal(rank[order], np.arange(500))
    keep_mask = (deg >= 5) | (rank == 0)
    assert np.array_equal(inv.astype(bool), keep_mask)
    assert vs == keep_mask.sum() and tw == int(deg[keep_mask].sum())
    keep = orc.keep_thresholds(deg, 0.01, tw)
    sT = 0.01 * tw
    for v in range(500):
        c = int(deg[v])
        if c == 0:
            assert keep[v] == 0xFFFFFFFF
            continue
        prob = (math.sqrt(c / sT) + 1) * sT / c
        assert keep[v] == (0xFFFFFFFF if prob >= 1 else int(prob * 2 ** 32))


def test_init_syn0_lcg_closed_form():
    d, V = 4, 6
    rank = np.array([3, 0, 5, 1, 2, 4], np.uint32)
    got = orc.init_syn0(V, d, rank)
    x = 1
    order = np.argsort(rank)
    for a in range(V):
        for b in range(d):
            x = (x * 25214903917 + 11) % 2 ** 64
            expect = np.float32(((x & 0xFFFF) / 65536 - 0.5) / d)
            assert got[order[a], b] == expect


# ---- SGNS subsampling: the kept positions against Philox word 0
def test_sgns_oracle_subsample_draws():
    """Sentences of distinct tokens short enough for one tile: the tile's
    input rows are the kept tokens in order, which must be exactly the
    positions i whose draw (i, SGNS domain, sentence, round) has first word
    below the keep threshold (here 2^31: keep with probability 1/2)."""
    V, seed, rnd = 64, 0x1234_0000_0077, 3
    deg = np.full(V, 10, np.uint32)
    order, rank, inv, vs, tw = orc.vocab(deg, 5)
    ut, ua = orc.unigram(deg, inv)
    thr = 1 << 31
    tables = dict(in_vocab=inv, vocab_rank=rank, keep_thr=np.full(V, thr, np.uint32), uni_thr=ut, uni_alias=ua)
    rng = np.random.default_rng(5)
    sents = [rng.permutation(V)[:20].astype(np.uint32) for _ in range(6)]
    offsets = np.concatenate([[0], np.cumsum([len(x) for x in sents])]).astype(np.uint64)
    tokens = np.concatenate(sents)
    d = 4
    ref = orc.sgns_groups(tables, offsets, tokens, np.zeros((V, d)), np.zeros((V, d)), d=d, window=3, negative=2,
                          batch=0, alpha=0.0, rnd=rnd, seed=seed)
    key = [seed & 0xFFFFFFFF, seed >> 32]
    got = {int(g[0]): r[:int(g[3])].tolist() for g, r in zip(ref["groups"], ref["rows"])}
    n_checked = 0
    for sid, sent in enumerate(sents):
        kept = [int(t) for i, t in enumerate(sent) if int(orc.philox([i, 0x10000000, sid, rnd], key)[0]) < thr]
        if len(kept) < 2:
            assert sid not in got
            continue
        assert got[sid] == kept, sid
        n_checked += 1
    assert n_checked >= 4


# ---- SGNS FP64 oracle: gradients by finite differences of the SGNS loss
def test_sgns_oracle_gradient_finite_difference():
    rng = np.random.default_rng(8)
    V, d = 12, 4
    deg = np.full(V, 10, np.uint32)
    order, rank, inv, vs, tw = orc.vocab(deg, 5)
    ut, ua = orc.unigram(deg, inv)
    tables = dict(in_vocab=inv, vocab_rank=rank, keep_thr=np.full(V, 0xFFFFFFFF, np.uint32), uni_thr=ut,
                  uni_alias=ua)
    tokens = np.array([1, 4, 7, 4, 9, 2, 11, 3], np.uint32)
    offsets = np.array([0, len(tokens)], np.uint64)
    s0 = rng.standard_normal((V, d)) * 0.5
    s1 = rng.standard_normal((V, d)) * 0.5
    kw = dict(d=d, window=3, negative=3, batch=4, alpha=1.0, rnd=1, seed=77)
    ref = orc.sgns_groups(tables, offsets, tokens, s0, s1, mode=0, **kw)

    # enumerate (context, centre) and (context, negative) pairs independently
    # of the C code, from the Philox draws; the loss is
    # -sum log sigma(v_ctx . u_centre) - sum log sigma(-v_ctx . u_neg)
    m, w, k, B = len(tokens), 3, 3, 4
    pairs, negs = [], []
    ng = -(-m // B)
    neg = {}
    for g in range(ng):
        for j in range(k):
            r = orc.philox([g * k + j, 0x10000000, 0, 1], [77, 0])
            bk = (int(r[2]) * V) >> 32
            neg[(g, j)] = bk if int(r[3]) < int(ut[bk]) else int(ua[bk])
    for c in range(m):
        b = int(orc.philox([c, 0x10000000, 0, 1], [77, 0])[1]) % w
        for aa in range(b, 2 * w + 1 - b):
            p = c - w + aa
            if aa == w or p < 0 or p >= m:
                continue
            pairs.append((int(tokens[p]), int(tokens[c])))
            for j in range(k):
                n_ = neg[(c // B, j)]
                if n_ != tokens[c]:
                    negs.append((int(tokens[p]), n_))

    def L(A, Bm):
        sig = lambda x: 1 / (1 + np.exp(-x))
        return (-sum(np.log(sig(A[i] @ Bm[o])) for i, o in pairs)
                - sum(np.log(sig(-(A[i] @ Bm[o]))) for i, o in negs))

    assert ref["pairs"] == len(pairs)
    eps = 1e-6
    gV = np.zeros_like(s0)
    gU = np.zeros_like(s1)
    for v in range(V):
        for e in range(d):
            for M, G in ((s0, gV), (s1, gU)):
                M[v, e] += eps
                lp = L(s0, s1)
                M[v, e] -= 2 * eps
                lm = L(s0, s1)
                M[v, e] += eps
                G[v, e] = (lp - lm) / (2 * eps)
    # the oracle's tile deltas, summed per row, equal -alpha * dL (alpha = 1)
    dV = np.zeros_like(s0)
    dU = np.zeros_like(s1)
    for i, g in enumerate(ref["groups"]):
        n_in, n_out = int(g[3]), int(g[4])
        for r in range(n_in):
            dV[ref["rows"][i][r]] += ref["dv"][i][r]
        for o in range(n_out):
            dU[ref["rows"][i][64 + o]] += ref["du"][i][o]
    assert np.allclose(dV, -gV, atol=1e-6)
    assert np.allclose(dU, -gU, atol=1e-6)


# ---- AUC: Mann-Whitney with ties vs scikit-learn
def test_auc_matches_sklearn():
    from sklearn.metrics import roc_auc_score
    rng = np.random.default_rng(1)
    pos = np.round(rng.normal(0.3, 1, 500), 1).astype(np.float32)
    neg = np.round(rng.normal(0, 1, 700), 1).astype(np.float32)
    y = np.r_[np.ones(500), np.zeros(700)]
    assert abs(orc.auc(pos, neg) - roc_auc_score(y, np.r_[pos, neg])) < 1e-12


# ---- evaluation: brute-force cosine
def test_eval_brute_force(er_graph):
    rp, col, wo, t, a = er_graph
    rng = np.random.default_rng(5)
    V, d = 2000, 16
    base = rng.standard_normal(d)
    syn0 = (base + 0.7 * rng.standard_normal((V, d))).astype(np.float32)
    inv = np.ones(V, np.uint8)
    eva, flags, scores = orc.evaluate(syn0, rp, col, inv, np.zeros(V, np.uint8), 3, 1)
    for v in range(0, V, 17):
        nb = col[rp[v]:rp[v + 1]]
        if len(nb) > 30:
            continue
        x = syn0[v].astype(np.float64)
        cs = [x @ syn0[u] / np.linalg.norm(x) / np.linalg.norm(syn0[u]) for u in nb]
        assert abs(scores[v] - np.mean(cs)) < 1e-5
        assert bool(flags[v]) == (np.mean(cs) > 0.65)
    assert eva == int((scores <= 0.65).sum())


# ---- word2vec restatement: one pair, one negative disabled, exact sigmoid
def test_w2v_single_pair_update():
    V, d = 3, 4
    syn0 = np.array([[0.1, -0.2, 0.3, 0.0], [0.05, 0.1, -0.1, 0.2], [0.0, 0.0, 0.0, 0.0]], np.float32)
    syn1 = np.array([[0.2, 0.1, 0.0, -0.1], [0.3, -0.3, 0.2, 0.1], [0.0, 0.0, 0.0, 0.0]], np.float32)
    deg = np.array([10, 9, 8], np.uint32)
    order, rank, inv, vs, tw = orc.vocab(deg, 1)
    s0, s1 = syn0.copy(), syn1.copy()
    offsets = np.array([0, 2], np.uint64)
    tokens = np.array([0, 1], np.uint32)
    words, pairs = orc.w2v_train(s0, s1, offsets, tokens, deg, order, rank, inv, vs, tw, window=1, negative=0,
                                 sample=0.0, alpha=0.5, use_lut=False, threads=1)
    assert pairs == 2 and words == 2
    # pair (centre 0, context 1) then (centre 1, context 0), sequentially
    a0, a1 = syn0.astype(np.float64), syn1.astype(np.float64)
    for centre, ctx in ((0, 1), (1, 0)):
        f = a0[ctx] @ a1[centre]
        g = (1 - 1 / (1 + math.exp(-f))) * 0.5
        e = g * a1[centre]
        a1[centre] = a1[centre] + g * a0[ctx]
        a0[ctx] = a0[ctx] + e
    assert np.allclose(s0, a0, atol=1e-6) and np.allclose(s1, a1, atol=1e-6)


# ---- learned walk budget: FP64 network, features, visit diagnostics
import budget  # noqa: E402


def test_budget_mlp_gradient_finite_difference():
    rng = np.random.default_rng(3)
    p = budget.init_params(5).astype(np.float64)
    p[budget.N_PARAMS - 2:] = [0.3, -0.2]
    X = rng.standard_normal((40, 8))
    T = rng.standard_normal((40, 2))
    loss, g = budget.loss_and_grad(p, X, T)
    assert abs(loss - ((budget.forward(p, X) - T) ** 2).sum() / 40) < 1e-12
    eps = 1e-6
    for i in rng.choice(budget.N_PARAMS, 300, replace=False):
        q = p.copy()
        q[i] += eps
        lp = ((budget.forward(q, X) - T) ** 2).sum() / 40
        q[i] -= 2 * eps
        lm = ((budget.forward(q, X) - T) ** 2).sum() / 40
        assert abs((lp - lm) / (2 * eps) - g[i]) < 1e-6 * max(1.0, abs(g[i]))


def test_budget_features_by_hand():
    # star 0-{1,2,3} with weights 1,2,1 and an edge 3-4 (w 4)
    src = np.array([0, 0, 0, 3], np.uint32)
    dst = np.array([1, 2, 3, 4], np.uint32)
    w = np.array([1, 2, 1, 4], np.float32)
    rp, col, wo = orc.build_csr(5, src, dst, w, True)
    f = budget.features(rp, col, wo, np.array([1, 0, 0, 1, 0]))
    ws = np.array([1, 2, 1]) / 4.0
    assert np.allclose(f[0], [math.log(4), math.log(5), 2, 4 / 3, math.log(1 + 1 + 1 + 2), math.log(3),
                              -(ws * np.log(ws)).sum() / math.log(3), 1])
    # vertex 3: neighbours 0 (deg 3) and 4 (deg 1), weights 1 and 4
    ws = np.array([1, 4]) / 5.0
    assert np.allclose(f[3], [math.log(3), math.log(6), 4, 2.5, math.log(5), math.log(4),
                              -(ws * np.log(ws)).sum() / math.log(2), 1])
    assert np.allclose(f[4], [math.log(2), math.log(5), 4, 4, math.log(3), math.log(3), 0, 0])


def test_visit_stats_closed_form():
    deg = np.array([2, 1, 1, 0])
    visits = np.array([1, 1, 0, 2])
    kl, ent, unv = budget.visit_stats(visits, deg)
    # p = (1/2, 1/4, 1/4, 0), q = (1/4, 1/4, 0, 1/2)
    assert abs(kl - (0.5 * math.log(2) + 0.25 * math.log(1))) < 1e-15
    assert abs(ent - (2 * 0.25 * math.log(4) + 0.5 * math.log(2))) < 1e-15
    assert unv == 1


# ---- evaluation subset: Floyd's sampling of 30 of deg > 30 neighbours
@pytest.mark.parametrize("deg", [31, 45, 200, 5000])
def test_eval_subset_is_uniform(deg):
    """Each draw is 30 distinct columns in [0, deg); every column is included
    with probability 30/deg (chi-square over 20,000 draws)."""
    from scipy.stats import chi2
    n = 20_000
    counts = np.zeros(deg, np.int64)
    for k in range(n):
        s = orc.eval_subset(deg, k % 997, 1 + k // 997, 0xABCDEF)
        assert len(s) == 30 and len(np.unique(s)) == 30 and s.max() < deg
        counts[s] += 1
    expect = n * 30 / deg
    stat = ((counts - expect) ** 2 / expect).sum()
    assert stat < chi2.ppf(0.9999, deg - 1), (stat, deg)


def test_eval_subset_small_degree_takes_all_and_pairs_are_uniform():
    assert np.array_equal(orc.eval_subset(7, 3, 2, 9), np.arange(7))
    assert np.array_equal(orc.eval_subset(30, 3, 2, 9), np.arange(30))
    # deg 32: a subset is fixed by its two missing columns; all 496 pairs
    # must be equally likely
    from scipy.stats import chi2
    n = 49_600
    hist = {}
    for k in range(n):
        miss = tuple(sorted(set(range(32)) - set(orc.eval_subset(32, k, 5, 77).tolist())))
        hist[miss] = hist.get(miss, 0) + 1
    assert len(hist) == 496
    c = np.array(list(hist.values()))
    stat = ((c - 100.0) ** 2 / 100.0).sum()
    assert stat < chi2.ppf(0.9999, 495)


def test_eval_brute_force_hubs():
    """Vertices with more than 30 neighbours: the score is the mean cosine
    over exactly the columns Floyd's sampling picks."""
    rng = np.random.default_rng(8)
    V, d = 600, 16
    hubs = np.arange(6)
    src = np.concatenate([np.repeat(hubs, 90), rng.integers(6, V, 400)]).astype(np.uint32)
    dst = np.concatenate([rng.choice(np.arange(6, V), 540), rng.integers(6, V, 400)]).astype(np.uint32)
    keep = src != dst
    src, dst = src[keep], dst[keep]
    rp, col, _ = orc.build_csr(V, src, dst, np.ones(len(src), np.float32), True)
    deg = np.diff(rp.astype(np.int64))
    assert (deg > 30).sum() >= 6
    syn0 = (rng.standard_normal(d) + 0.8 * rng.standard_normal((V, d))).astype(np.float32)
    eva, flags, scores = orc.evaluate(syn0, rp, col, np.ones(V, np.uint8), np.zeros(V, np.uint8), 4, 0x77)
    for v in range(V):
        if deg[v] == 0:
            continue
        cols = orc.eval_subset(int(deg[v]), v, 4, 0x77)
        nb = col[rp[v] + cols.astype(np.uint64)]
        x = syn0[v].astype(np.float64)
        ref = np.mean([x @ syn0[u] / np.linalg.norm(x) / np.linalg.norm(syn0[u]) for u in nb])
        assert abs(scores[v] - ref) < 1e-5
        assert bool(flags[v]) == (scores[v] > 0.65)


# ---- word2vec restatement: unigram table, sigmoid table, negative draws
def test_w2v_unigram_table_law():
    """Slot a holds index i until a / table_size passes the cumulative
    share D_i of cn^0.75 (src/word2vec.cu:674-694), the slot that passes it
    included, so indices 0..i hold floor(table_size * D_i) + 2 slots."""
    rng = np.random.default_rng(4)
    cn = np.sort(rng.zipf(1.8, 3000).clip(1, 50_000))[::-1].astype(np.uint32)
    T = 2_000_000
    table = orc.w2v_unigram_table(cn, T)
    assert table.min() == 0 and table.max() == len(cn) - 1
    assert np.all(np.diff(table) >= 0) and np.all(np.diff(table) <= 1)
    counts = np.bincount(table, minlength=len(cn))
    share = np.cumsum(cn.astype(np.float64) ** 0.75) / (cn.astype(np.float64) ** 0.75).sum()
    gap = np.cumsum(counts)[:-1] - T * share[:-1]
    assert gap.min() >= 1.0 - 1e-6 and gap.max() <= 2.0 + 1e-6


def test_w2v_sigmoid_table_and_index():
    t = orc.w2v_exp_table()
    x = (np.arange(1000) / 1000.0 * 2 - 1) * 6
    assert np.abs(t[:1000] - 1 / (1 + np.exp(-x))).max() < 2e-7
    assert t[1000] == t[999]
    # (int)((f + 6) * 83): the integer division 1000/6/2 is 83
    for f, idx in ((0.0, 498), (1.0, 581), (-6.0, 0), (5.999, 995), (-0.5, 456), (2.5, 705), (-5.99, 0)):
        assert orc.w2v_lut_index(f) == idx, f


def test_w2v_negative_draw_hand_traced():
    table = np.array([0, 0, 1, 1, 1, 2, 3, 3], np.int32)
    vocab = 4
    for seed in (0, 1, 12345, 2 ** 63 + 17, 987654321987):
        x = (seed * 25214903917 + 11) % (1 << 64)
        tr = int(table[(x >> 16) % 8])
        if tr == 0:
            tr = x % (vocab - 1) + 1
        got, nxt = orc.w2v_negative(seed, table, vocab)
        assert (got, nxt) == (tr, x)


def _w2v_fp64(syn0, syn1, sent, order, window, negative, alpha, table, vocab_size):
    """skip_gram_kernel (src/word2vec.cu:442-536) per pair, sequentially, in
    FP64, with the sigmoid table and the kernel's two LCGs; sentences of
    vertex ids, block index = sentence index."""
    s0 = syn0.astype(np.float64)
    s1 = syn1.astype(np.float64)
    et = orc.w2v_exp_table().astype(np.float64)
    M = (1 << 64) - 1
    for blk, words in enumerate(sent):
        nr = blk
        for pos, word in enumerate(words):
            nr = (nr * 2514903917 + 11) & M
            b = nr % window
            for a in range(b, 2 * window + 1 - b):
                if a == window:
                    continue
                c = pos - window + a
                if c < 0 or c >= len(words):
                    continue
                last = words[c]
                neu1e = np.zeros(s0.shape[1])
                for dd in range(negative + 1):
                    if dd == 0:
                        target, label = word, 1.0
                    else:
                        nr = (nr * 25214903917 + 11) & M
                        tr = int(table[(nr >> 16) % len(table)])
                        if tr == 0:
                            tr = nr % (vocab_size - 1) + 1
                        target = int(order[tr])
                        if target == word:
                            continue
                        label = 0.0
                    f = s0[last] @ s1[target]
                    if f > 6:
                        g = (label - 1) * alpha
                    elif f < -6:
                        g = label * alpha
                    else:
                        g = (label - et[int((f + 6) * 83)]) * alpha
                    neu1e += g * s1[target]
                    s1[target] = s1[target] + g * s0[last]
                s0[last] = s0[last] + neu1e
    return s0, s1


def test_w2v_train_with_negatives_and_table_matches_fp64_restatement():
    """orc_w2v_train with negatives and the sigmoid table (one thread, no
    subsampling) against an FP64 restatement driven by the same LCG draws;
    logits near a table-cell boundary may pick the neighbouring cell, so the
    check is 1e-4 on every element."""
    rng = np.random.default_rng(21)
    V, d = 12, 8
    deg = np.array([9, 8, 8, 7, 6, 6, 5, 5, 4, 3, 3, 2], np.uint32)
    order, rank, inv, vs, tw = orc.vocab(deg, 1)
    syn0 = (rng.standard_normal((V, d)) * 0.4).astype(np.float32)
    syn1 = (rng.standard_normal((V, d)) * 0.4).astype(np.float32)
    sent = [rng.integers(0, V, 9).tolist(), rng.integers(0, V, 14).tolist(), rng.integers(0, V, 6).tolist()]
    offsets = np.cumsum([0] + [len(s) for s in sent]).astype(np.uint64)
    tokens = np.array(sum(sent, []), np.uint32)
    T = 1000
    table = orc.w2v_unigram_table(deg[order], T)
    s0, s1 = syn0.copy(), syn1.copy()
    words, pairs = orc.w2v_train(s0, s1, offsets, tokens, deg, order, rank, inv, vs, tw, window=2, negative=3,
                                 sample=0.0, alpha=0.05, use_lut=True, table_size=T, threads=1)
    assert words == len(tokens) and pairs > 0
    r0, r1 = _w2v_fp64(syn0, syn1, sent, order, 2, 3, 0.05, table, vs)
    assert np.abs(s0 - r0).max() < 1e-4 and np.abs(s1 - r1).max() < 1e-4
    assert np.abs(r1 - syn1).max() > 1e-3  # the negatives did train


@pytest.mark.parametrize("mode", ["skip-gram", "reuse", "cbow"])
def test_w2v_trainer_modes_match_fp64_restatement(mode):
    """orc_w2v_train_mode (one thread, no subsampling, sigmoid table) against
    the FP64 restatement of the same trainer with the same draws."""
    import w2v_fp64
    rng = np.random.default_rng(31)
    V, d = 14, 8
    deg = np.array([9, 8, 8, 7, 6, 6, 5, 5, 4, 3, 3, 2, 2, 2], np.uint32)
    order, rank, inv, vs, tw = orc.vocab(deg, 1)
    syn0 = (rng.standard_normal((V, d)) * 0.4).astype(np.float32)
    syn1 = (rng.standard_normal((V, d)) * 0.4).astype(np.float32)
    sent = [rng.integers(0, V, 9).tolist(), rng.integers(0, V, 14).tolist(), rng.integers(0, V, 6).tolist()]
    offsets = np.cumsum([0] + [len(s) for s in sent]).astype(np.uint64)
    tokens = np.array(sum(sent, []), np.uint32)
    T = 1000
    table = orc.w2v_unigram_table(deg[order], T)
    m = {"skip-gram": orc.W2V_SKIPGRAM, "reuse": orc.W2V_REUSE, "cbow": orc.W2V_CBOW}[mode]
    s0, s1 = syn0.copy(), syn1.copy()
    seed, rnd = 0x1234567890AB, 3
    words, pairs = orc.w2v_train_mode(s0, s1, offsets, tokens, deg, order, rank, inv, vs, tw, mode=m, window=2,
                                      negative=3, sample=0.0, alpha=0.05, table_size=T, threads=1, seed=seed,
                                      rnd=rnd)
    assert words == len(tokens) and pairs > 0
    negs = w2v_fp64.reuse_negatives(len(sent), 3, table, vs, order, seed, rnd) if mode == "reuse" else None
    r0, r1 = w2v_fp64.train(mode, syn0, syn1, sent, range(len(sent)), order, 2, 3, 0.05, table, vs, negs)
    assert np.abs(s0 - r0).max() < 1e-4 and np.abs(s1 - r1).max() < 1e-4
    assert np.abs(r1 - syn1).max() > 1e-3 and np.abs(r0 - syn0).max() > 1e-3


# ---- -classes k-means restatement
def test_kmeans_by_hand_and_separated_clusters():
    """One round by hand on four rows; on well-separated clusters each round
    moves every row to the best unit centroid of the previous round's classes
    and no true cluster is split."""
    syn0 = np.array([[1, 0], [0.9, 0.1], [0, 1], [0.2, 0.8]], np.float32)
    order = np.arange(4, dtype=np.uint32)
    # start: classes 0,1,0,1; centroid 0 = (row0+row2)/3, centroid 1 = (row1+row3)/3, unit length
    c0 = (syn0[0] + syn0[2]) / 3
    c1 = (syn0[1] + syn0[3]) / 3
    c0, c1 = c0 / np.linalg.norm(c0), c1 / np.linalg.norm(c1)
    want = [int(np.argmax([x @ c0, x @ c1])) for x in syn0]
    assert orc.kmeans_classes(syn0, order, 4, 2, iters=1).tolist() == want
    rng = np.random.default_rng(2)
    centres = rng.standard_normal((6, 16)) * 5
    labels = rng.integers(0, 6, 3000)
    x = (centres[labels] + 0.05 * rng.standard_normal((3000, 16))).astype(np.float32)
    perm = rng.permutation(3000).astype(np.uint32)
    # round 10 moves each row to the best unit centroid of round 9's classes
    c9 = orc.kmeans_classes(x, perm, 3000, 6, iters=9)
    c10 = orc.kmeans_classes(x, perm, 3000, 6, iters=10)
    rows = x[perm].astype(np.float64)
    cent = np.stack([rows[c9 == k].sum(0) / ((c9 == k).sum() + 1) for k in range(6)])
    with np.errstate(invalid="ignore"):
        cent /= np.linalg.norm(cent, axis=1, keepdims=True)  # an empty class: 0/0, never chosen
    dots = np.nan_to_num(rows @ cent.T, nan=-np.inf)
    best = np.where(dots.max(1) > -10, dots.argmax(1), 0)
    assert np.array_equal(c10, best)
    # rows of one true cluster never split
    for lab in range(6):
        assert len(np.unique(c10[labels[perm] == lab])) == 1
