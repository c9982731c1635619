"""GPU parity of the HIP path against the CPU oracle, through the C ABI.

Graph construction, alias tables, vocabulary, table initialisation and walks
must match bit for bit; SGNS arithmetic must match the FP64 oracle to 1e-5
relative (norm-wise) on race-free groups; the evaluation to 1e-5 on scores.
"""
import numpy as np
import pytest

import felog
import graphgen
import orc

pytestmark = pytest.mark.gpu

SEED = 0x1234ABCD


def small_graph(n=10_000, m=50_000, seed=1, weights="uniform"):
    return graphgen.erdos_renyi(n, m, seed=seed, weights=weights)


def make_ctx(src, dst, w, n, **kw):
    opts = dict(size=128, window=10, negative=5, sample=1e-3, alpha=0.025, min_L=20, min_R=2, seed=SEED)
    opts.update(kw)
    print(f"[case] ctx V={n} m={len(src)} opts={opts}", flush=True)
    f = felog.Felog(**opts)
    f.set_edges(src, dst, w, n, make_undirected=True)
    f.build()
    return f


@pytest.fixture(scope="module")
def er():
    src, dst, w = small_graph()
    f = make_ctx(src, dst, w, 10_000)
    yield f, (src, dst, w)
    f.close()


def test_csr_bitexact(er):
    f, (src, dst, w) = er
    rp, col, wo = f.csr()
    rp2, col2, wo2 = orc.build_csr(10_000, src, dst, w, True)
    assert np.array_equal(rp, rp2)
    assert np.array_equal(col, col2)
    assert np.array_equal(wo.view(np.uint32), wo2.view(np.uint32))


def test_alias_bitexact(er):
    f, _ = er
    rp, col, wo = f.csr()
    t, a = f.alias()
    t2, a2 = orc.build_alias(10_000, rp, wo)
    assert np.array_equal(t, t2)
    assert np.array_equal(a, a2)


def test_vocab_tables_bitexact(er):
    f, _ = er
    order, deg = f.vocab()
    o2, rank2, inv2, vs, tw = orc.vocab(deg, 5)
    assert np.array_equal(order, o2)
    ut, ua = f.unigram()
    ut2, ua2 = orc.unigram(deg, inv2)
    assert np.array_equal(ut, ut2) and np.array_equal(ua, ua2)
    s0, s1 = f.embeddings()
    assert np.array_equal(s0.view(np.uint32), orc.init_syn0(10_000, 128, rank2).view(np.uint32))
    assert not s1.any()


@pytest.mark.parametrize("rnd", [1, 2])
def test_walks_bitexact(er, rnd):
    f, _ = er
    c = f.walk_round(rnd)
    off, tok, wid = f.corpus()
    rp, col, _ = f.csr()
    t, a = f.alias()
    off2, tok2, wid2, steps2 = orc.walks(10_000, rp, col, t, a, None, 10_000, 20, 1000, rnd, SEED)
    assert c["n_walks"] == len(wid2) and c["n_steps"] == steps2
    assert np.array_equal(wid, wid2)
    assert np.array_equal(off, off2)
    assert np.array_equal(tok, tok2)


@pytest.mark.parametrize("undirected", [False, True])
def test_compat_walks_bitexact(undirected):
    """The reference's common-neighbour transition (compat_walk): uniform
    neighbour, tanh acceptance, trace steps.  Edge data in [0, 4) gives
    common-neighbour counts 0..3; two rounds against the host replay."""
    n, m = 10_000, 50_000
    rng = np.random.default_rng(12)
    src = rng.integers(0, n, m).astype(np.uint32)
    dst = rng.integers(0, n, m).astype(np.uint32)
    w = (rng.random(m) * 4).astype(np.float32)
    opts = dict(size=64, window=10, negative=5, min_L=20, min_R=2, seed=SEED, compat_walk=1)
    print(f"[case] compat walks V={n} m={m} undirected={undirected}", flush=True)
    f = felog.Felog(**opts)
    try:
        f.set_edges(src, dst, w, n, make_undirected=undirected)
        f.build()
        rp, col, wo = f.csr()
        t, a = f.alias()
        for rnd in (1, 2):
            c = f.walk_round(rnd)
            off, tok, wid = f.corpus()
            off2, tok2, wid2, steps2 = orc.walks(n, rp, col, t, a, None, n, 20, 1000, rnd, SEED, cn_data=wo)
            assert c["n_walks"] == len(wid2) and c["n_steps"] == steps2
            assert np.array_equal(wid, wid2) and np.array_equal(off, off2) and np.array_equal(tok, tok2)
    finally:
        f.close()


def _random_tables(V, d, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((V, d)) * 0.3).astype(np.float32), (rng.standard_normal((V, d)) * 0.3).astype(np.float32)


def _tables_for_oracle(f):
    order, deg = f.vocab()
    _, rank, inv, vs, tw = orc.vocab(deg, 5)
    ut, ua = f.unigram()
    keep = orc.keep_thresholds(deg, 1e-3, tw)
    return dict(in_vocab=inv, vocab_rank=rank, keep_thr=keep, uni_thr=ut, uni_alias=ua)


def test_sgns_groups_vs_fp64(er):
    f, _ = er
    f.walk_round(1)
    off, tok, _ = f.corpus()
    n = 300
    off_s, tok_s = off[: n + 1], tok[: int(off[n])]
    s0, s1 = _random_tables(10_000, 128, 5)
    f.set_embeddings(s0, s1)
    got = f.debug_sgns(off_s, tok_s, 1, 0.025, 4096)
    ref = orc.sgns_groups(_tables_for_oracle(f), off_s, tok_s, s0, s1, d=128, window=10, negative=5, batch=0,
                          alpha=0.025, rnd=1, seed=SEED, mode=0)
    assert len(got["groups"]) == ref["n_groups"] > 0
    key = {(int(g[0]), int(g[1])): i for i, g in enumerate(ref["groups"])}
    worst = 0.0
    for i, g in enumerate(got["groups"]):
        j = key[(int(g[0]), int(g[1]))]
        assert np.array_equal(g, ref["groups"][j])
        assert np.array_equal(got["rows"][i], ref["rows"][j])
        n_in, n_out = int(g[3]), int(g[4])
        for name, shape in (("logits", (n_in, n_out)), ("dv", (n_in, 128)), ("du", (n_out, 128))):
            x = got[name][i][: shape[0], : shape[1]].astype(np.float64)
            y = ref[name][j][: shape[0], : shape[1]]
            err = np.linalg.norm(x - y) / max(np.linalg.norm(y), 1e-30)
            worst = max(worst, err)
            assert err < 1e-5, (name, i, err)
    print(f"worst relative error vs FP64: {worst:.3e}")


@pytest.mark.parametrize("rows", [32, 48, 64])
def test_sgns_long_sentences_vs_fp64(er, rows):
    """Sentences of 30-120 random tokens make tiles of up to 41 distinct
    input rows: with the default 32-row split those over 32 rows run on the
    48-row kernel in a second launch.  Every tile shape matches FP64."""
    f, _ = er
    rng = np.random.default_rng(17)
    lens = rng.integers(30, 121, 60)
    off_s = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
    tok_s = rng.integers(0, 10_000, int(off_s[-1])).astype(np.uint32)
    s0, s1 = _random_tables(10_000, 128, 9)
    f.set_embeddings(s0, s1)
    f.set_tile_shape(rows, 0, 0)
    try:
        got = f.debug_sgns(off_s, tok_s, 2, 0.025, 4096)
    finally:
        f.set_tile_shape()
    ref = orc.sgns_groups(_tables_for_oracle(f), off_s, tok_s, s0, s1, d=128, window=10, negative=5, batch=0,
                          alpha=0.025, rnd=2, seed=SEED, mode=0)
    assert len(got["groups"]) == ref["n_groups"] > 0
    assert max(int(g[3]) for g in got["groups"]) > 32  # tiles for the second kernel
    key = {(int(g[0]), int(g[1])): i for i, g in enumerate(ref["groups"])}
    for i, g in enumerate(got["groups"]):
        j = key[(int(g[0]), int(g[1]))]
        assert np.array_equal(g, ref["groups"][j])
        n_in, n_out = int(g[3]), int(g[4])
        for name, shape in (("logits", (n_in, n_out)), ("dv", (n_in, 128)), ("du", (n_out, 128))):
            x = got[name][i][: shape[0], : shape[1]].astype(np.float64)
            y = ref[name][j][: shape[0], : shape[1]]
            assert np.linalg.norm(x - y) <= 1e-5 * max(np.linalg.norm(y), 1e-30), (rows, name, i)


def test_sgns_paired_sentences_vs_fp64(er):
    """Sentences of at most 32 tokens are prepared two per wavefront (one per
    half).  With 21 centres per group, lengths 23-32 give most such
    sentences a second tile, starting past centre 0 with its own negatives
    and window shrinks; 2-22 give one tile; an odd count of sentences leaves
    the last wavefront half empty, and a few sentences of 33-60 tokens send
    their pair to the whole-wavefront path.  Every tile matches FP64."""
    f, _ = er
    rng = np.random.default_rng(23)
    lens = np.concatenate([rng.integers(23, 33, 40), rng.integers(2, 23, 16), rng.integers(33, 61, 5)])
    rng.shuffle(lens)
    assert len(lens) % 2 == 1
    off_s = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
    tok_s = rng.integers(0, 10_000, int(off_s[-1])).astype(np.uint32)
    s0, s1 = _random_tables(10_000, 128, 11)
    f.set_embeddings(s0, s1)
    got = f.debug_sgns(off_s, tok_s, 3, 0.025, 4096)
    ref = orc.sgns_groups(_tables_for_oracle(f), off_s, tok_s, s0, s1, d=128, window=10, negative=5, batch=0,
                          alpha=0.025, rnd=3, seed=SEED, mode=0)
    assert len(got["groups"]) == ref["n_groups"] > len(lens)  # second tiles exist
    assert sum(1 for g in got["groups"] if int(g[1]) > 0) >= 20  # tiles starting past centre 0
    key = {(int(g[0]), int(g[1])): i for i, g in enumerate(ref["groups"])}
    for i, g in enumerate(got["groups"]):
        j = key[(int(g[0]), int(g[1]))]
        assert np.array_equal(g, ref["groups"][j]), (i, g, ref["groups"][j])
        assert np.array_equal(got["rows"][i], ref["rows"][j]), i
        n_in, n_out = int(g[3]), int(g[4])
        for name, shape in (("logits", (n_in, n_out)), ("dv", (n_in, 128)), ("du", (n_out, 128))):
            x = got[name][i][: shape[0], : shape[1]].astype(np.float64)
            y = ref[name][j][: shape[0], : shape[1]]
            assert np.linalg.norm(x - y) <= 1e-5 * max(np.linalg.norm(y), 1e-30), (name, i)


def test_sgns_groups_bf16_vs_fp64():
    """bf16 tables: X on bf16 MFMA from the exact bf16 rows; the gradient
    products split G into bf16 hi + lo (relative residue below 2^-16), so the
    deltas match FP64 over the same bf16 tables to 3e-5, the logits to 1e-5."""
    src, dst, w = small_graph()
    f = make_ctx(src, dst, w, 10_000, precision=felog.BF16)
    try:
        f.walk_round(1)
        off, tok, _ = f.corpus()
        n = 300
        off_s, tok_s = off[: n + 1], tok[: int(off[n])]
        s0, s1 = _random_tables(10_000, 128, 7)
        f.set_embeddings(s0, s1)
        b0, b1 = f.embeddings()  # the bf16 values the kernel reads
        assert not np.array_equal(b0, s0)
        got = f.debug_sgns(off_s, tok_s, 1, 0.025, 4096)
        ref = orc.sgns_groups(_tables_for_oracle(f), off_s, tok_s, b0, b1, d=128, window=10, negative=5, batch=0,
                              alpha=0.025, rnd=1, seed=SEED, mode=0)
        assert len(got["groups"]) == ref["n_groups"] > 0
        key = {(int(g[0]), int(g[1])): i for i, g in enumerate(ref["groups"])}
        worst = {"logits": 0.0, "dv": 0.0, "du": 0.0}
        tol = {"logits": 1e-5, "dv": 3e-5, "du": 3e-5}
        for i, g in enumerate(got["groups"]):
            j = key[(int(g[0]), int(g[1]))]
            assert np.array_equal(g, ref["groups"][j])
            n_in, n_out = int(g[3]), int(g[4])
            for name, shape in (("logits", (n_in, n_out)), ("dv", (n_in, 128)), ("du", (n_out, 128))):
                x = got[name][i][: shape[0], : shape[1]].astype(np.float64)
                y = ref[name][j][: shape[0], : shape[1]]
                err = np.linalg.norm(x - y) / max(np.linalg.norm(y), 1e-30)
                worst[name] = max(worst[name], err)
                assert err < tol[name], (name, i, err)
        print(f"bf16 worst relative error vs FP64: {worst}")
    finally:
        f.close()


def test_eval_vs_oracle(er):
    f, _ = er
    rng = np.random.default_rng(3)
    base = rng.standard_normal((1, 128)).astype(np.float32)
    s0 = (base + 0.8 * rng.standard_normal((10_000, 128))).astype(np.float32)
    f.set_embeddings(s0, None)
    flags0, _ = f.flags()
    nu, stop = f.evaluate()
    flags, scores = f.flags()
    rp, col, _ = f.csr()
    order, deg = f.vocab()
    _, _, inv, _, _ = orc.vocab(deg, 5)
    eva, flags2, scores2 = orc.evaluate(s0, rp, col, inv, flags0, f.stats()["rounds_done"], SEED)
    assert np.allclose(scores, scores2, atol=1e-5)
    near = np.abs(scores2 - 0.65) < 1e-5
    assert np.array_equal(flags[~near], flags2[~near])
    assert abs(int(nu) - eva) <= int(near.sum())
