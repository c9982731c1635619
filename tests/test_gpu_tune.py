"""GPU tests of the SGNS tile shapes and their autotuner (felog_tune_sgns).

A tile shape changes how the tile kernel is laid on the chip (input rows
48 or 64, the register budget it is compiled for, persistent workgroups per
CU, centre windows counted per wavefront, columns per wavefront item of the
update products), never what a tile computes.  On a graph whose tiles share
no rows (disjoint cycles, one walker per cycle, one tile per walk, no
negatives) every FP32 shape must therefore train the same tables bit for
bit.  (BF16 tables are rounded stochastically with a stream per thread, and
the dim split changes which thread stores a column.)"""
import numpy as np
import pytest

import felog
import graphgen

pytestmark = pytest.mark.gpu
SEED = 0x7A11


def cycles(k=512, length=24):
    j = np.repeat(np.arange(k), length)
    t = np.tile(np.arange(length), k)
    src = (j + k * t).astype(np.uint32)
    dst = (j + k * ((t + 1) % length)).astype(np.uint32)
    return src, dst, np.ones(len(src), np.float32), k * length


def ctx(src, dst, w, n, **kw):
    opts = dict(size=64, window=10, negative=5, sample=1e-3, min_L=20, min_R=10, seed=SEED)
    opts.update(kw)
    print(f"[case] V={n} m={len(src)} {opts}", flush=True)
    f = felog.Felog(**opts)
    f.set_edges(src, dst, w, n, make_undirected=True)
    f.build()
    return f


def shapes(f, prec):
    lo = 2 if prec == felog.FP32 else 3
    out = []
    for rows in (32, 48, 64):
        for wpe in (lo, lo + 1):
            f.set_tile_shape(rows, wpe, 0)
            for g in range(1, f.tile_shape()["groups_per_cu"] + 1):
                out.append((rows, wpe, g))
    f.set_tile_shape()
    return out


def window_split_shapes(dp):
    """windows per wavefront x dim split, the default (8, dp) first"""
    out = [(8, dp)]
    for wpw in (1, 2, 4, 8):
        for ds in (dp // 4, dp // 2, dp):
            if (wpw, ds) != out[0]:
                out.append((wpw, ds))
    return out


def test_every_tile_shape_trains_the_same_tables():
    *g, n = cycles()
    kw = dict(walkers=512, max_walk_len=22, negative=0, min_count=1)
    probe = ctx(*g, n, **kw)
    try:
        cand = shapes(probe, felog.FP32)
    finally:
        probe.close()
    cand = [(rows, wpe, gpc, 8, 64) for rows, wpe, gpc in cand]
    # windows per wavefront and dim split on the default rows / budget
    cand += [(32, 3, 0, wpw, ds) for wpw, ds in window_split_shapes(64)[1:]]
    print(f"{len(cand)} shapes: {cand}")
    assert len(cand) >= 15
    ref = None
    for rows, wpe, gpc, wpw, ds in cand:
        f = ctx(*g, n, **kw)
        try:
            f.set_tile_shape(rows, wpe, gpc, wpw, ds)
            got = f.tile_shape()
            assert (got["rows"], got["waves_per_simd"], got["windows_per_wave"], got["dim_split"]) == \
                (rows, wpe, wpw, ds)
            assert gpc == 0 or got["groups_per_cu"] == gpc
            for r in (1, 2):
                f.walk_round(r)
                f.train_round()
            tabs = f.embeddings()
            assert f.stats()["train_pairs"] > 0
        finally:
            f.close()
        if ref is None:
            ref = tabs
        else:
            assert np.array_equal(tabs[0], ref[0]) and np.array_equal(tabs[1], ref[1]), (rows, wpe, gpc, wpw, ds)


@pytest.mark.parametrize("prec", [felog.FP32, felog.BF16])
def test_tuner_restores_tables_and_installs_the_fastest(prec):
    n = 20_000
    f = ctx(*graphgen.rmat(n, 100_000, seed=41), n, precision=prec)
    try:
        expect = shapes(f, prec)
        f.walk_round(1)
        before = f.embeddings()
        res = f.tune_sgns(sample=5000, reps=2)
        after = f.embeddings()
        for r in res:
            print(r)
        key = lambda r: tuple(r[k] for k in felog.SHAPE_FIELDS)  # noqa: E731
        first = res[:len(expect)]
        # first sweep: every rows x budget x workgroups at the default
        # windows per wavefront (8) and dim split (dp = 64)
        assert [key(r) for r in first] == [e + (8, 64) for e in expect]
        # second sweep: windows x split at the first sweep's fastest
        b = min(first, key=lambda r: r["tile_ms"])
        assert [key(r) for r in res[len(expect):]] == \
            [(b["rows"], b["waves_per_simd"], b["groups_per_cu"], wpw, ds) for wpw, ds in window_split_shapes(64)[1:]]
        assert all(r["tile_ms"] > 0 for r in res)
        assert np.array_equal(before[0], after[0]) and np.array_equal(before[1], after[1])
        best = min(res, key=lambda r: r["tile_ms"])
        assert key(f.tile_shape()) == key(best)
        f.train_round()
        assert f.stats()["train_pairs"] > 0 and np.isfinite(f.embeddings()[0]).all()
    finally:
        f.close()
