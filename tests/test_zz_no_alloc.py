"""felog_train_corpus_noalloc: the corpus staging and the SGNS tile
descriptors in the library's static device scratch instead of allocations.

Two calls in a row, the second with a smaller corpus, must train exactly the
tables felog_train_corpus trains from the same start.  Tiles are race free
(disjoint cycles, one walker per cycle, no negatives), so the comparison is
bit for bit; a corpus beyond the scratch is refused."""
import ctypes as C

import numpy as np
import pytest

import felog

pytestmark = pytest.mark.gpu
SEED = 0x0A110C


def cycles(k=512, length=24):
    j = np.repeat(np.arange(k), length)
    t = np.tile(np.arange(length), k)
    src = (j + k * t).astype(np.uint32)
    dst = (j + k * ((t + 1) % length)).astype(np.uint32)
    return src, dst, np.ones(len(src), np.float32), k * length


def ctx(src, dst, w, n):
    f = felog.Felog(size=64, window=10, negative=0, sample=1e-3, min_L=20, seed=SEED, walkers=512, max_walk_len=22,
                    min_count=1)
    f.set_edges(src, dst, w, n, make_undirected=True)
    f.build()
    return f


def noalloc(f, offsets, tokens, rnd):
    fn = f.lib.felog_train_corpus_noalloc
    fn.restype = C.c_int32
    fn.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint32]
    offsets = np.ascontiguousarray(offsets, np.uint64)
    tokens = np.ascontiguousarray(tokens, np.uint32)
    return fn(f.ctx, offsets.ctypes.data, tokens.ctypes.data, len(offsets) - 1, rnd)


def test_noalloc_training_matches_the_allocating_entry_point():
    *g, n = cycles()
    fa, fb = ctx(*g, n), ctx(*g, n)
    try:
        fa.walk_round(1)
        off, tok, _ = fa.corpus()
        print(f"[case] corpus {len(off) - 1} walks, {len(tok)} tokens; then the first 100", flush=True)
        rng = np.random.default_rng(2)
        s0 = (rng.standard_normal((n, 64)) * 0.3).astype(np.float32)
        s1 = (rng.standard_normal((n, 64)) * 0.3).astype(np.float32)
        fa.set_embeddings(s0, s1)
        fb.set_embeddings(s0, s1)
        k = 100
        for rnd, (o, t) in enumerate(((off, tok), (off[: k + 1], tok[: int(off[k])])), start=1):
            fa.train_corpus(o, t, rnd)
            st = noalloc(fb, o, t, rnd)
            assert st == 0, fb.lib.felog_last_error(fb.ctx).decode()
            a0, a1 = fa.embeddings()
            b0, b1 = fb.embeddings()
            assert np.array_equal(a0.view(np.uint32), b0.view(np.uint32)), rnd
            assert np.array_equal(a1.view(np.uint32), b1.view(np.uint32)), rnd
            assert fb.stats()["train_pairs"] == fa.stats()["train_pairs"] > 0
            assert not np.array_equal(a1, s1)
        # a corpus beyond the scratch is refused with a capacity error
        big = 3_000_000
        o = np.arange(0, 2 * big + 1, 2, dtype=np.uint64)
        t = np.zeros(2 * big, np.uint32)
        st = noalloc(fb, o, t, 3)
        assert st == -7 and "scratch" in fb.lib.felog_last_error(fb.ctx).decode()
    finally:
        fa.close()
        fb.close()
