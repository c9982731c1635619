"""Device memory comes back when a context is destroyed: the build keeps
the blocks it has finished with (edge input, scratch) for the round
buffers and tile descriptors instead of freeing them, so this checks that
every such block, used or not, is freed with the context."""
import numpy as np
import pytest

import felog
import graphgen

pytestmark = pytest.mark.gpu


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
    import torch
    n = 200_000
    g = graphgen.rmat(n, 1_000_000, seed=3)
    print(f"[case] R-MAT {n} vertices, create / build / train / destroy", flush=True)
    cycle(g, n)  # loads the code objects and the runtime's own buffers
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    for prec in (felog.FP32, felog.BF16, felog.FP32):
        cycle(g, n, precision=prec)
    torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info()[0]
    print(f"free before {free0 / 2**20:.0f} MiB, after three contexts {free1 / 2**20:.0f} MiB")
    # one build at this size holds several hundred MiB; a leaked block would show
    assert free0 - free1 < 32 << 20
