"""GPU parity of the steps either side of the path, through the C ABI:
common-neighbour counts (pairs, every CSR entry, the gcn tool), the
reference's transition reading those counts, the device link-prediction AUC,
and text edge loading."""
import os
import subprocess

import numpy as np
import pytest

import felog
import graphgen
import orc

pytestmark = pytest.mark.gpu
SEED = 0x6C6E
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def make(src, dst, w, n, **kw):
    opts = dict(size=64, window=10, negative=5, min_L=20, min_R=100, seed=SEED)
    opts.update(kw)
    print(f"[case] V={n} m={len(src)} {opts}", flush=True)
    f = felog.Felog(**opts)
    f.set_edges(src, dst, w, n, make_undirected=True)
    f.build()
    return f


def test_common_neighbours_match_oracle():
    """Pairs and every CSR entry on an R-MAT graph with hubs (long sorted
    lists, both orders of the two lists)."""
    n = 20_000
    src, dst, w = graphgen.rmat(n, 150_000, seed=51)
    f = make(src, dst, w, n)
    try:
        rp, col, _ = f.csr()
        got = f.common_neighbours(src, dst)
        ref = orc.common_neighbours(n, rp, col, src, dst)
        assert np.array_equal(got, ref) and ref.max() > 50
        rng = np.random.default_rng(1)
        u, v = rng.integers(0, n, 5000), rng.integers(0, n, 5000)
        assert np.array_equal(f.common_neighbours(u, v), orc.common_neighbours(n, rp, col, u, v))
        f.use_common_neighbours(True)
        rows = np.repeat(np.arange(n), np.diff(rp.astype(np.int64)))
        assert np.array_equal(f.cn(), orc.common_neighbours(n, rp, col, rows, col).astype(np.float32))
    finally:
        f.close()


def test_compat_walks_read_device_counts():
    """The reference's transition with counts from felog_use_common_neighbours
    (not the edge data): walks bit for bit as the host replay fed the
    oracle's counts."""
    n = 10_000
    src, dst, w = graphgen.rmat(n, 60_000, seed=52)
    f = make(src, dst, w, n, compat_walk=1)
    try:
        f.use_common_neighbours(True)
        rp, col, _ = f.csr()
        t, a = f.alias()
        rows = np.repeat(np.arange(n), np.diff(rp.astype(np.int64)))
        cn = orc.common_neighbours(n, rp, col, rows, col).astype(np.float32)
        for r in (1, 2):
            c = f.walk_round(r)
            off, tok, wid = f.corpus()
            off2, tok2, wid2, steps2 = orc.walks(n, rp, col, t, a, None, n, 20, 1000, r, SEED, cn_data=cn)
            assert c["n_steps"] == steps2
            assert np.array_equal(off, off2) and np.array_equal(tok, tok2)
        # switching back reads the edge data again
        f.use_common_neighbours(False)
        f.walk_round(3)
        off, tok, _ = f.corpus()
        _, wo = f.csr()[1:]
        off3, tok3, _, _ = orc.walks(n, rp, col, t, a, None, n, 20, 1000, 3, SEED, cn_data=wo)
        assert np.array_equal(tok, tok3)
    finally:
        f.close()


def test_gcn_tool_writes_the_oracle_counts(tmp_path):
    n = 5000
    src, dst, w = graphgen.rmat(n, 30_000, seed=53)
    txt, out = str(tmp_path / "g.txt"), str(tmp_path / "cn.txt")
    felog.write_edges(txt, src, dst, None, binary=False)
    exe = os.path.join(ROOT, "info-centric_ge_amd", "gcn")
    r = subprocess.run([exe, "-i", txt, "-o", out, "-v", str(n), "-s", "unweighted", "--make-undirected", "-t", "4"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rows = np.loadtxt(out, dtype=np.int64).reshape(-1, 3)
    rp, col, _ = orc.build_csr(n, src, dst, w, True)
    assert np.array_equal(rows[:, 0], src) and np.array_equal(rows[:, 1], dst)
    assert np.array_equal(rows[:, 2], orc.common_neighbours(n, rp, col, src, dst))


def test_link_auc_matches_oracle():
    n = 20_000
    src, dst, w = graphgen.rmat(n, 100_000, seed=54)
    for prec in (felog.FP32, felog.BF16):
        f = make(src, dst, w, n, precision=prec)
        try:
            rng = np.random.default_rng(2)
            s0 = rng.standard_normal((n, 64)).astype(np.float32)
            f.set_embeddings(s0, None)
            e0, _ = f.embeddings()
            pairs = rng.integers(0, n, (30_000, 2))
            pairs[:2000] = pairs[2000:4000]  # ties
            labels = (rng.random(30_000) < 0.5).astype(np.uint8)
            auc, scores = f.link_auc(pairs, labels)
            a, b = e0[pairs[:, 0]].astype(np.float64), e0[pairs[:, 1]].astype(np.float64)
            ref_s = (a * b).sum(1) / np.sqrt((a * a).sum(1) * (b * b).sum(1))
            assert np.abs(scores - ref_s).max() < 1e-5
            ref = orc.auc(scores[labels == 1], scores[labels == 0])
            print(f"precision {prec}: AUC device {auc:.9f} oracle on the device scores {ref:.9f}")
            assert abs(auc - ref) < 1e-9
        finally:
            f.close()


def test_text_edges_load_like_binary(tmp_path):
    n = 3000
    src, dst, w = graphgen.erdos_renyi(n, 12_000, seed=55)
    txt = str(tmp_path / "g.txt")
    felog.write_edges(txt, src, dst, w, binary=False)
    _, _, wt = felog.read_edges_txt(txt, 3)
    fa = make(src, dst, wt, n)
    fb = felog.Felog(size=64, window=10, negative=5, min_L=20, min_R=100, seed=SEED)
    try:
        fb.load_edges_txt(txt, n, make_undirected=True, weighted=True)
        fb.build()
        for x, y in zip(fa.csr(), fb.csr()):
            assert np.array_equal(x, y)
        assert np.array_equal(fa.alias()[0], fb.alias()[0])
    finally:
        fa.close()
        fb.close()
