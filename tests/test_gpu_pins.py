"""GPU parity on the branches the small Erdos-Renyi fixture never reaches.

  - evaluation of vertices with more than 30 neighbours (Floyd's subset of
    30 columns) on an R-MAT graph full of such vertices;
  - walker selection gated by convergence flags that an evaluation set;
  - the learned budget's min_R recheck: exactly the budgeted vertices whose
    score is not above 0.65 get one more round;
  - error paths of the binary edge loader.
Every comparison is against the CPU oracle through the C ABI.
"""
import os

import numpy as np
import pytest

import felog
import graphgen
import orc
import walk_budget

pytestmark = pytest.mark.gpu
SEED = 0x51A7E


def make_ctx(src, dst, w, n, **kw):
    opts = dict(size=64, window=10, negative=5, sample=1e-3, min_L=20, min_R=2, seed=SEED)
    opts.update(kw)
    print(f"[case] V={n} m={len(src)} {opts}", flush=True)
    f = felog.Felog(**opts)
    f.set_edges(src, dst, w, n, make_undirected=True)
    f.build()
    return f


@pytest.fixture(scope="module")
def rmat():
    n = 20_000
    src, dst, w = graphgen.rmat(n, 100_000, seed=31)
    f = make_ctx(src, dst, w, n)
    yield f, n
    f.close()


def clustered_rows(n, d, frac, seed):
    """A fraction `frac` of the vertices share one direction (plus a little
    noise), the rest are random: a vertex's mean cosine to its neighbours is
    then about the share of them in the cluster, so scores fall on both
    sides of 0.65."""
    rng = np.random.default_rng(seed)
    base = rng.standard_normal(d)
    base *= 3.0 / np.linalg.norm(base)
    rows = rng.standard_normal((n, d)) * 0.3
    cl = rng.random(n) < frac
    rows[cl] = base + 0.05 * rng.standard_normal((int(cl.sum()), d))
    return rows.astype(np.float32)


def test_eval_vs_oracle_hubs(rmat):
    """Scores of vertices with more than 30 neighbours (the Floyd branch)
    within 1e-5 of the oracle's, flags equal away from the threshold."""
    f, n = rmat
    rp, col, _ = f.csr()
    deg = np.diff(rp.astype(np.int64))
    hubs = int((deg > 30).sum())
    assert hubs > 1000, hubs
    s0 = clustered_rows(n, 64, 0.7, 3)
    f.set_embeddings(s0, None)
    flags0, _ = f.flags()
    assert not flags0.any()
    nu, _ = f.evaluate()
    flags, scores = f.flags()
    order, vdeg = f.vocab()
    _, _, inv, _, _ = orc.vocab(vdeg, 5)
    eva, flags2, scores2 = orc.evaluate(s0, rp, col, inv, flags0, f.stats()["rounds_done"], SEED)
    big = deg > 30
    err = np.abs(scores - scores2)
    print(f"{hubs} vertices with deg > 30; worst score error {err[big].max():.2e} (hubs), {err.max():.2e} (all); "
          f"{int(flags2[big].sum())} hubs flagged, {int((~flags2.astype(bool) & big).sum())} not")
    assert err.max() < 1e-5
    near = np.abs(scores2 - 0.65) < 1e-5
    assert np.array_equal(flags[~near], flags2[~near])
    assert abs(int(nu) - eva) <= int(near.sum())
    assert flags2[big].any() and (~flags2[big].astype(bool)).any()
    # reset the flags for the other tests of this module
    f.set_embeddings(np.zeros_like(s0), None)


def test_flag_gated_walks_match_oracle():
    """An evaluation flags part of the vertices; the next round starts
    walkers only at the others, bit for bit as the host replay with the
    same flags does."""
    n = 20_000
    src, dst, w = graphgen.rmat(n, 100_000, seed=32)
    f = make_ctx(src, dst, w, n)
    try:
        f.walk_round(1)
        f.set_embeddings(clustered_rows(n, 64, 0.7, 4), None)
        f.evaluate()
        flags, _ = f.flags()
        print(f"{int(flags.sum())} of {n} vertices flagged")
        assert 1000 < flags.sum() < n - 1000
        c = f.walk_round(2)
        off, tok, wid = f.corpus()
        rp, col, _ = f.csr()
        t, a = f.alias()
        off2, tok2, wid2, steps2 = orc.walks(n, rp, col, t, a, flags, n, 20, 1000, 2, SEED)
        assert c["n_walks"] == len(wid2) == n - int(flags.sum())
        assert c["n_steps"] == steps2
        assert not flags[wid % n].any()
        assert np.array_equal(wid, wid2) and np.array_equal(off, off2) and np.array_equal(tok, tok2)
    finally:
        f.close()


def test_budget_recheck_extends_exactly_the_unconverged():
    """Every vertex budgeted for 2 rounds; at round min_R = 2 the recheck
    scores them with the evaluation's test.  The counts raised to 3 must be
    exactly those of in-vocabulary vertices whose oracle score on the
    round-2 tables is not above 0.65 (scores within 1e-4 of the threshold
    excepted), and both sets must be non-empty."""
    n = 20_000
    src, dst, w = graphgen.rmat(n, 100_000, seed=33)
    f = make_ctx(src, dst, w, n, max_rounds=2)
    try:
        f.set_embeddings(clustered_rows(n, 64, 0.7, 5), None)
        f.budget_set(walk_budget.constant_model(2, 22))
        f.budget_recheck(True)
        cnt0, _, r0 = f.budget()
        assert (cnt0 == 2).all() and r0 == 2
        f.run_budget()
        assert f.stats()["rounds_done"] == 2
        cnt, _, rounds = f.budget()
        s0, _ = f.embeddings()
        rp, col, _ = f.csr()
        order, vdeg = f.vocab()
        _, _, inv, _, _ = orc.vocab(vdeg, 5)
        flags, _ = f.flags()
        assert not flags.any()
        _, _, scores = orc.evaluate(s0, rp, col, inv, np.zeros(n, np.uint8), 2, SEED)
        ext = cnt == 3
        want = inv.astype(bool) & ~(scores > 0.65)
        near = np.abs(scores - 0.65) < 1e-4
        print(f"recheck: {int(ext.sum())} extended, {int((inv.astype(bool) & ~ext).sum())} kept, "
              f"{int(near.sum())} near the threshold; rounds reported {rounds}")
        assert set(np.unique(cnt).tolist()) <= {2, 3}
        assert np.array_equal(ext[~near], want[~near])
        assert ext.any() and (inv.astype(bool) & ~ext).any()
        assert rounds == 3  # the largest count after the recheck
    finally:
        f.close()


def test_edge_loader_errors(tmp_path):
    """A file that is not a whole number of 12-byte records is an I/O error,
    a vertex id >= v_num an argument error, a missing file an I/O error;
    nothing is loaded and the context stays usable."""
    src = np.array([0, 1, 2], np.uint32)
    dst = np.array([1, 2, 3], np.uint32)
    w = np.ones(3, np.float32)
    good = str(tmp_path / "g.bin")
    graphgen.write_bin(good, src, dst, w)
    raw = open(good, "rb").read()
    short = str(tmp_path / "short.bin")
    open(short, "wb").write(raw[:-5])
    ragged = str(tmp_path / "ragged.bin")
    open(ragged, "wb").write(raw + b"\x01\x02")
    f = felog.Felog(size=32, window=2, negative=1, min_L=4, seed=1)
    try:
        for path, status in ((short, -2), (ragged, -2), (str(tmp_path / "missing.bin"), -2)):
            with pytest.raises(felog.FelogError) as e:
                f.load_edges(path, 4)
            assert e.value.status == status, path
        with pytest.raises(felog.FelogError) as e:
            f.load_edges(good, 3)  # vertex 3 is out of range
        assert e.value.status == -1 and "v_num" in str(e.value)
        with pytest.raises(felog.FelogError) as e:
            f.build()
        assert e.value.status == -6  # nothing was loaded
        f.load_edges(good, 4)
        f.build()
        rp, col, _ = f.csr()
        assert rp.tolist() == [0, 1, 3, 5, 6] and sorted(col.tolist()) == [0, 1, 1, 2, 2, 3]
    finally:
        f.close()
