"""Device memory comes back when a context is destroyed: the build keeps
the blocks it has finished with (edge input, scratch) for the round
buffers and tile descriptors instead of freeing them, so this checks that
every such block, used or not, is freed with the context.

Free memory is read with hipMemGetInfo from the HIP runtime the library
itself loaded (the same libamdhip64 it links), never through a second
runtime in the same process."""
import ctypes

import pytest

import felog
import graphgen

pytestmark = pytest.mark.gpu


def free_bytes():
    felog.load_library()  # the library's HIP runtime is loaded first
    hip = ctypes.CDLL("libamdhip64.so.7", mode=ctypes.RTLD_GLOBAL)
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def cycle(g, n, **kw):
    f = felog.Felog(size=64, window=5, negative=5, min_L=20, min_R=2, seed=5, **kw)
    try:
        f.set_edges(*g, n, make_undirected=True)
        f.build()
        f.walk_round(1)
        f.train_round()
        f.run_graph(1)
    finally:
        f.close()


def test_contexts_return_their_memory():
    n = 200_000
    g = graphgen.rmat(n, 1_000_000, seed=3)
    print(f"[case] R-MAT {n} vertices, create / build / train / destroy", flush=True)
    cycle(g, n)  # loads the code objects and the runtime's own buffers
    free0 = free_bytes()
    for prec in (felog.FP32, felog.BF16, felog.FP32):
        cycle(g, n, precision=prec)
    free1 = free_bytes()
    print(f"free before {free0 / 2**20:.0f} MiB, after three contexts {free1 / 2**20:.0f} MiB")
    # one build at this size holds several hundred MiB; a leaked block would show
    assert free0 - free1 < 32 << 20
