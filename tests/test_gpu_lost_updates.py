"""GPU tests of the lost-update tracking of the tile trainer's write-back
(felog_track_lost_updates).

  - On disjoint cycles (one walker per cycle, no negatives) no two tiles
    share a row: each walk is one tile that writes back its distinct
    vertices once as input rows and once as output rows, none is lost, and
    the trained tables equal those of a run without the tracking bit for bit
    (the tracking only reads and counts).
  - On a 20k R-MAT graph hub rows are shared by many concurrent tiles: some
    write-backs replace another tile's update, never more than were written,
    in both table precisions; with the tracking off the counts are zero."""
import numpy as np
import pytest

import felog
import graphgen

pytestmark = pytest.mark.gpu
SEED = 0x105E


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


def test_disjoint_tiles_lose_nothing_and_tracking_changes_nothing():
    *g, n = cycles()
    kw = dict(walkers=512, max_walk_len=22, negative=0, min_count=1)
    tabs = {}
    for track in (False, True):
        f = ctx(*g, n, **kw)
        try:
            if track:
                f.track_lost_updates(True)
            for r in (1, 2):
                c = f.walk_round(r)
                off, tok, _ = f.corpus()
                f.train_round()
                st = f.stats()
                if track:
                    # one tile per walk: its distinct vertices as input and as output rows
                    assert c["n_walks"] == 512 and (np.diff(off.astype(np.int64)) == 22).all()
                    expect = 2 * sum(len(set(tok[off[i]:off[i + 1]].tolist())) for i in range(512))
                    assert st["train_rows_written"] == expect, (st["train_rows_written"], expect)
                    assert st["train_rows_lost"] == 0
                else:
                    assert st["train_rows_written"] == 0 and st["train_rows_lost"] == 0
            tabs[track] = f.embeddings()
        finally:
            f.close()
    assert np.array_equal(tabs[False][0], tabs[True][0]) and np.array_equal(tabs[False][1], tabs[True][1])


@pytest.mark.parametrize("prec", [felog.FP32, felog.BF16])
def test_shared_hub_rows_lose_some_updates(prec):
    n = 20_000
    f = ctx(*graphgen.rmat(n, 100_000, seed=41), n, precision=prec)
    try:
        f.track_lost_updates(True)
        f.walk_round(1)
        f.train_round()
        st = f.stats()
        written, lost = st["train_rows_written"], st["train_rows_lost"]
        print(f"precision {prec}: {written} rows written back, {lost} replaced another tile's update "
              f"({lost / max(written, 1):.4f})")
        assert written > 0 and 0 < lost < written
        # captured rounds count the same way
        f.run_graph(1)
        st = f.stats()
        assert 0 < st["train_rows_lost"] < st["train_rows_written"]
        f.track_lost_updates(False)
        f.walk_round(3)
        f.train_round()
        st = f.stats()
        assert st["train_rows_written"] == 0 and st["train_rows_lost"] == 0
    finally:
        f.close()
