"""GPU checks at the full sizes of the three large workloads.

  Chung-Lu 2.24M vertices / 35M edges, BF16 tables: alias tables and the
      first round of walks bit for bit against the oracle, BF16 SGNS tiles
      of 1,000 walks against FP64, the evaluation on a vertex sample;
  R-MAT 3.07M / 117M edges, LogNormal(0, 1) weights: maximum degree above
      30,000, alias tables bit for bit, the walks of the first 200,000
      walkers bit for bit, and the hub's transitions against its weights
      (chi-square);
  R-MAT 41.6M / 620M edges (1.24B CSR entries), BF16 tables, a learned
      budget: every step is an edge, every vertex starts exactly its count
      of walks of exactly its length, the corpus is the sum of the budgets,
      no buffer overflows, a full BF16 training round stays finite.
The graphs come from the device generators (graphgen.*_device), run in a
child process that writes the reference's binary edge file; the test loads
it with felog_load_edges."""
import os
import tempfile
import time

import numpy as np
import pytest

import felog
import graphgen
import orc
import walk_budget

pytestmark = pytest.mark.gpu
SEED = 0xC0DE5


def log(msg):
    print(f"[{time.strftime('%H:%M:%S')}] {msg}", flush=True)


def built(gen, n, m, seed, weights="uniform", **kw):
    """Generate the graph into a binary edge file (RAM-backed when /dev/shm
    is writable), load it through the C ABI, build."""
    opts = dict(size=128, window=10, negative=5, sample=1e-3, min_L=20, min_R=2, seed=SEED)
    opts.update(kw)
    log(f"[case] {gen} V={n} m={m} {weights} {opts}")
    base = "/dev/shm" if os.access("/dev/shm", os.W_OK) else None
    fd, path = tempfile.mkstemp(suffix=".bin", dir=base)
    os.close(fd)
    try:
        graphgen.device_graph_file(gen, n, m, seed, weights, path)
        f = felog.Felog(**opts)
        f.load_edges(path, n, make_undirected=True)
    finally:
        os.unlink(path)
    f.build()
    log(f"loaded {f.E // 2} edges; built in {f.stats()['build_ms']:.1f} ms")
    return f


def test_chunglu_2_24m_bf16():
    n = 2_240_000
    f = built("chung_lu_device", n, 35_000_000, 3, precision=felog.BF16)
    try:
        rp, col, wo = f.csr()
        deg = np.diff(rp.astype(np.int64))
        log(f"max degree {deg.max()}")
        t, a = f.alias()
        t2, a2 = orc.build_alias(n, rp, wo)
        assert np.array_equal(t, t2) and np.array_equal(a, a2)
        log("alias tables equal")
        c = f.walk_round(1)
        off, tok, wid = f.corpus()
        off2, tok2, wid2, steps2 = orc.walks(n, rp, col, t, a, None, n, 20, 1000, 1, SEED, threads=16)
        assert c["n_walks"] == len(wid2) == n and c["n_steps"] == steps2
        assert np.array_equal(off, off2) and np.array_equal(tok, tok2)
        log(f"round 1 walks equal: {c['n_tokens']} tokens")
        # BF16 tiles of the first 1,000 walks against FP64 on the same bf16 tables
        rng = np.random.default_rng(6)
        s0 = (rng.standard_normal((n, 128)) * 0.3).astype(np.float32)
        s1 = (rng.standard_normal((n, 128)) * 0.3).astype(np.float32)
        f.set_embeddings(s0, s1)
        b0, b1 = f.embeddings()
        k = 1000
        off_s, tok_s = off[: k + 1], tok[: int(off[k])]
        got = f.debug_sgns(off_s, tok_s, 1, 0.0025, 2 * k)
        order, vdeg = f.vocab()
        _, rank, inv, vs, tw = orc.vocab(vdeg, 5)
        ut, ua = f.unigram()
        tables = dict(in_vocab=inv, vocab_rank=rank, keep_thr=orc.keep_thresholds(vdeg, 1e-3, tw), uni_thr=ut,
                      uni_alias=ua)
        ref = orc.sgns_groups(tables, off_s, tok_s, b0, b1, d=128, window=10, negative=5, batch=0, alpha=0.0025,
                              rnd=1, seed=SEED, mode=0, max_groups=2 * k)
        assert len(got["groups"]) == ref["n_groups"] >= k
        key = {(int(g[0]), int(g[1])): i for i, g in enumerate(ref["groups"])}
        worst = {"logits": 0.0, "dv": 0.0, "du": 0.0}
        for i, g in enumerate(got["groups"]):
            j = key[(int(g[0]), int(g[1]))]
            assert np.array_equal(g, ref["groups"][j]) and np.array_equal(got["rows"][i], ref["rows"][j])
            n_in, n_out = int(g[3]), int(g[4])
            for name, shape in (("logits", (n_in, n_out)), ("dv", (n_in, 128)), ("du", (n_out, 128))):
                x = got[name][i][: shape[0], : shape[1]].astype(np.float64)
                y = ref[name][j][: shape[0], : shape[1]]
                worst[name] = max(worst[name], np.linalg.norm(x - y) / max(np.linalg.norm(y), 1e-30))
        log(f"bf16 tiles vs FP64 over {len(got['groups'])} tiles: worst relative error {worst}")
        assert worst["logits"] < 1e-5 and worst["dv"] < 3e-5 and worst["du"] < 3e-5
        # train the round, then the evaluation on a sample of vertices (all of
        # the 2,000 largest hubs and 20,000 others)
        f.set_embeddings(orc.init_syn0(n, 128, rank), np.zeros((n, 128), np.float32))
        f.train_round()
        flags0, _ = f.flags()
        f.evaluate()
        _, scores = f.flags()
        e0, _ = f.embeddings()
        sample = np.zeros(n, np.uint8)
        sample[np.argsort(deg)[-2000:]] = 1
        sample[rng.choice(n, 20_000, replace=False)] = 1
        sample &= inv
        _, _, sc2 = orc.evaluate(e0, rp, col, sample, flags0, f.stats()["rounds_done"], SEED)
        m = sample.astype(bool)
        err = np.abs(scores[m] - sc2[m]).max()
        log(f"evaluation on {int(m.sum())} vertices ({int((deg[m] > 30).sum())} with deg > 30): worst {err:.2e}")
        assert err < 1e-5
    finally:
        f.close()


def test_rmat_3_07m_lognormal():
    n = 3_070_000
    f = built("rmat_device", n, 117_000_000, 4, "lognormal")
    try:
        rp, col, wo = f.csr()
        deg = np.diff(rp.astype(np.int64))
        hub = int(np.argmax(deg))
        log(f"max degree {deg[hub]} at vertex {hub}")
        assert deg[hub] > 30_000
        t, a = f.alias()
        t2, a2 = orc.build_alias(n, rp, wo)
        assert np.array_equal(t, t2) and np.array_equal(a, a2)
        log("alias tables equal")
        c = f.walk_round(1)
        off, tok, wid = f.corpus()
        k = 200_000
        off2, tok2, wid2, _ = orc.walks(n, rp, col, t, a, None, k, 20, 1000, 1, SEED, threads=16)
        assert np.array_equal(wid[:k], wid2) and np.array_equal(off[: k + 1], off2)
        assert np.array_equal(tok[: int(off[k])], tok2)
        log(f"first {k} walks equal; round has {c['n_tokens']} tokens")
        # every move out of the hub in the round against its weights
        b, e = int(rp[hub]), int(rp[hub + 1])
        nb, wt = col[b:e], wo[b:e].astype(np.float64)
        pos = np.nonzero(tok == hub)[0]
        end = off[np.searchsorted(off, pos, side="right")].astype(np.int64)
        pos = pos[pos + 1 < end]
        nxt = tok[pos + 1]
        order = np.argsort(nb)
        idx = order[np.searchsorted(nb[order], nxt)]
        assert np.array_equal(nb[idx], nxt)  # every move is an edge of the hub
        p = wt / wt.sum()
        bins = np.minimum((np.cumsum(p) - p / 2) * 40, 39).astype(np.int64)
        obs = np.bincount(bins[idx], minlength=40)
        exp = np.bincount(bins, weights=p, minlength=40) * len(idx)
        stat = ((obs - exp) ** 2 / exp).sum()
        log(f"{len(idx)} moves out of the hub; chi-square {stat:.1f} over 40 weight bins")
        from scipy.stats import chi2
        assert len(idx) > 4000 and stat < chi2.ppf(0.9999, 39)
    finally:
        f.close()


def _budget_model_of_degree():
    """ln count = 0.15 ln(1+deg), ln length = 0.05 ln(1+deg) + ln 18: counts
    1 .. ~8 and lengths 19 .. ~36 vertices, growing with the degree."""
    p = np.zeros(felog.BUDGET_PARAMS, np.float32)
    # W1 (64x8) row 0 picks feature 0; W2 (64x64) row 0 passes it on
    p[0] = 1.0
    oW2 = 64 * 8 + 64
    p[oW2] = 1.0
    oW3 = oW2 + 64 * 64 + 64
    p[oW3] = 0.15
    p[oW3 + 64] = 0.05
    p[oW3 + 128 + 1] = np.log(18.0)
    big = np.full(8, 1e30)
    return walk_budget.make_model(p, np.zeros(8), np.ones(8), -big, big, [0.0, 0.0], [1.0, 1.0])


def test_rmat_41_6m_learned_budget_bf16():
    n = 41_600_000
    f = built("rmat_device", n, 620_000_000, 5, precision=felog.BF16, max_walk_len=64, overflow_capacity=1 << 28)
    try:
        st = f.stats()
        assert st["e_num"] > 1_200_000_000
        f.budget_set(_budget_model_of_degree())
        cnt, ln, rounds = f.budget()
        assert cnt.min() >= 1 and 3 <= rounds <= 12
        log(f"budget: counts {np.bincount(cnt).tolist()}, lengths {int(ln.min())}..{int(ln.max())}")
        starts = np.zeros(n, np.int64)
        rng = np.random.default_rng(9)
        for r in range(1, rounds + 1):
            c = f.walk_round(r)
            off, tok, wid = f.corpus()
            v = wid.astype(np.int64) % n
            want = np.nonzero(cnt >= r)[0]
            assert np.array_equal(np.sort(v), want), r
            lens = np.diff(off.astype(np.int64))
            assert np.array_equal(lens, ln[v].astype(np.int64)), r
            assert c["n_tokens"] == int(ln[want].astype(np.int64).sum())
            assert np.array_equal(tok[off[:-1].astype(np.int64)], v.astype(np.uint32))
            starts[v] += 1
            log(f"round {r}: {c['n_walks']} walks, {c['n_tokens']} tokens, counts and lengths as budgeted")
            if r == 1:
                rp, col, _ = f.csr(weights=False)
                log("CSR copied out")
                # 100,000 random steps are edges (each source row scanned once)
                i = rng.integers(0, len(tok) - 1, 100_000)
                end = off[np.searchsorted(off, i, side="right")].astype(np.int64)
                i = i[i + 1 < end]
                u, x = tok[i].astype(np.int64), tok[i + 1]
                order = np.argsort(u, kind="stable")
                u, x = u[order], x[order]
                cut = np.flatnonzero(np.diff(u)) + 1
                for us, xs in zip(np.split(u, cut), np.split(x, cut)):
                    row = col[int(rp[us[0]]):int(rp[us[0] + 1])]
                    assert np.isin(xs, row).all(), int(us[0])
                log(f"{len(i)} sampled steps are edges ({len(cut) + 1} distinct sources)")
                del rp, col
                f.train_round()
                st = f.stats()
                log(f"BF16 training round: {st['train_pairs']} pairs in {st['train_kernel_ms']:.1f} ms")
        assert np.array_equal(starts, cnt.astype(np.int64))
        # every syn0 row is finite after training: with the budget removed the
        # evaluation scores every vocabulary vertex against its neighbours
        f.budget_set(None)
        f.evaluate()
        _, scores = f.flags()
        assert np.isfinite(scores).all() and (scores != 0).any()
        log("tables finite after training")
    finally:
        f.close()
