"""Seeded synthetic graphs, the binary edge format and the link-prediction split.

Generators (all undirected, no self loops, no duplicate edges; every vertex
gets degree >= 1 because the reference asserts it, include/walk.hpp:778):
  erdos_renyi(n, m)          G(n, m) uniform pairs
  rmat(n, m, a, b, c)        recursive-matrix edges, ids >= n rejected
  chung_lu(n, m, gamma)      endpoints drawn with probability ~ expected
                             degree, expected degrees ~ i^(-1/(gamma-1))
  rmat_device, chung_lu_device   the same laws generated on the GPU (torch)
                             for graphs of tens of millions of edges
Weights: "uniform" -> U(0,1], "lognormal" -> LogNormal(0, 1).

Binary graph file (include/storage.hpp:40-61,107-115; src/tools/gconverter.cpp:69-76):
packed little-endian records {u32 src, u32 dst, f32 data}, no header.

Link-prediction split (src/tools/gtrain.cpp:292-343): a fraction of the
edges is held out as positives, each only while both endpoints keep degree
>= 2 in the remaining graph; the same number of vertex pairs that are not
edges (and not repeated) are drawn as negatives.
"""
from __future__ import annotations

import os

import numpy as np

EDGE_DTYPE = np.dtype([("src", "<u4"), ("dst", "<u4"), ("data", "<f4")])


def _weights(rng, m, law):
    if law == "uniform":
        return (1.0 - rng.random(m)).astype(np.float32)  # (0, 1]
    if law == "lognormal":
        return rng.lognormal(0.0, 1.0, m).astype(np.float32)
    raise ValueError(law)


def _canon(src, dst, n):
    lo = np.minimum(src, dst).astype(np.uint64)
    hi = np.maximum(src, dst).astype(np.uint64)
    keep = lo != hi
    key = lo[keep] * np.uint64(n) + hi[keep]
    return key


def _dedupe_keys(keys, rng):
    u, first = np.unique(keys, return_index=True)
    # restore generation order of first occurrences
    order = np.argsort(first, kind="stable")
    return u[order]


def _fit_patches(keys, n, m):
    """Leading generated edges such that, with one patch edge per isolated
    vertex, the graph has about m edges."""
    keys = keys[:m]
    seen = np.zeros(n, bool)
    seen[(keys // np.uint64(n)).astype(np.int64)] = True
    seen[(keys % np.uint64(n)).astype(np.int64)] = True
    iso = n - int(seen.sum())
    return keys[:max(m - iso, 1)]


def _finish(keys, n, rng, weights):
    src = (keys // np.uint64(n)).astype(np.uint32)
    dst = (keys % np.uint64(n)).astype(np.uint32)
    # random orientation, like an edge list read from a file
    flip = rng.random(len(src)) < 0.5
    s = np.where(flip, dst, src)
    d = np.where(flip, src, dst)
    s, d = _patch_isolated(s, d, n, rng)
    w = _weights(rng, len(s), weights)
    return s.astype(np.uint32), d.astype(np.uint32), w


def _patch_isolated(src, dst, n, rng):
    deg = np.bincount(src, minlength=n) + np.bincount(dst, minlength=n)
    iso = np.nonzero(deg == 0)[0].astype(np.uint32)
    if len(iso) == 0:
        return src, dst
    nonzero = np.nonzero(deg > 0)[0]
    if len(nonzero) == 0:
        partner = (iso + 1) % n
    else:
        partner = nonzero[rng.integers(0, len(nonzero), len(iso))].astype(np.uint32)
    return np.concatenate([src, iso]), np.concatenate([dst, partner])


def erdos_renyi(n, m, seed=1, weights="uniform"):
    rng = np.random.default_rng(seed)
    keys = np.empty(0, np.uint64)
    while len(keys) < m:
        need = int((m - len(keys)) * 1.1) + 16
        k = _canon(rng.integers(0, n, need, dtype=np.uint64), rng.integers(0, n, need, dtype=np.uint64), n)
        keys = _dedupe_keys(np.concatenate([keys, k]), rng)
    return _finish(_fit_patches(keys, n, m), n, rng, weights)


def _rmat_batch(rng, count, scale, a, b, c):
    src = np.zeros(count, np.uint64)
    dst = np.zeros(count, np.uint64)
    for _ in range(scale):
        r = rng.random(count)
        sbit = r >= a + b            # quadrants c, d
        dbit = ((r >= a) & (r < a + b)) | (r >= a + b + c)  # quadrants b, d
        src = (src << np.uint64(1)) | sbit.astype(np.uint64)
        dst = (dst << np.uint64(1)) | dbit.astype(np.uint64)
    return src, dst


def rmat(n, m, a=0.57, b=0.19, c=0.19, seed=1, weights="uniform"):
    rng = np.random.default_rng(seed)
    scale = int(np.ceil(np.log2(max(n, 2))))
    keys = np.empty(0, np.uint64)
    while len(keys) < m:
        need = int((m - len(keys)) * 1.3) + 1024
        s, d = _rmat_batch(rng, need, scale, a, b, c)
        ok = (s < n) & (d < n)
        keys = _dedupe_keys(np.concatenate([keys, _canon(s[ok], d[ok], n)]), rng)
    return _finish(_fit_patches(keys, n, m), n, rng, weights)


def _device_edges(n, m, g, seed, weights, draw, log=None):
    """Distinct undirected edges from `draw(count)` (device endpoint pairs,
    self loops and ids >= n rejected) until m exist; a random m of them
    kept, isolated vertices patched, random orientation.  Host arrays."""
    import torch
    dev = torch.device("cuda")
    keys = torch.empty(0, dtype=torch.int64, device=dev)
    while keys.numel() < m:
        need = int((m - keys.numel()) * 1.4) + 1024
        parts = [keys]
        for src, dst in draw(need):
            ok = (src < n) & (dst < n) & (src != dst)
            lo_, hi_ = torch.minimum(src, dst)[ok], torch.maximum(src, dst)[ok]
            parts.append(lo_ * n + hi_)
            del src, dst, ok, lo_, hi_
        keys = torch.unique(torch.cat(parts))
        del parts
        if log:
            log(f"device generator: {keys.numel()} distinct edges of {m}")
    keys = keys[torch.randperm(keys.numel(), generator=g, device=dev)[:m]]
    src, dst = keys // n, keys % n
    del keys
    flip = torch.rand(src.numel(), generator=g, device=dev) < 0.5
    src, dst = torch.where(flip, dst, src), torch.where(flip, src, dst)
    deg = torch.bincount(src, minlength=n) + torch.bincount(dst, minlength=n)
    iso = torch.nonzero(deg == 0).flatten()
    if iso.numel():
        nz = torch.nonzero(deg > 0).flatten()
        partner = nz[torch.randint(0, nz.numel(), (iso.numel(),), generator=g, device=dev)]
        src, dst = torch.cat([src, iso]), torch.cat([dst, partner])
    out_s = src.to(torch.int32).cpu().numpy().view(np.uint32)
    out_d = dst.to(torch.int32).cpu().numpy().view(np.uint32)
    del src, dst, deg
    torch.cuda.empty_cache()
    rng = np.random.default_rng(seed)
    return out_s, out_d, _weights(rng, len(out_s), weights)


def rmat_device(n, m, a=0.57, b=0.19, c=0.19, seed=1, weights="uniform", chunk=1 << 27, log=None):
    """R-MAT on the GPU (torch) for graphs whose host generation would take
    minutes: the same quadrant law as rmat(), self loops and ids >= n
    rejected, duplicates removed, a random m of the distinct edges kept,
    isolated vertices patched, random orientation.  Edge order is random
    (not generation order).  Returns host numpy arrays."""
    import torch
    dev = torch.device("cuda")
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    scale = int(np.ceil(np.log2(max(n, 2))))

    def draw(need):
        for lo in range(0, need, chunk):
            cnt = min(chunk, need - lo)
            src = torch.zeros(cnt, dtype=torch.int64, device=dev)
            dst = torch.zeros(cnt, dtype=torch.int64, device=dev)
            for _ in range(scale):
                r = torch.rand(cnt, generator=g, device=dev, dtype=torch.float64)
                sbit = r >= a + b
                dbit = ((r >= a) & (r < a + b)) | (r >= a + b + c)
                src = (src << 1) | sbit.to(torch.int64)
                dst = (dst << 1) | dbit.to(torch.int64)
            yield src, dst

    return _device_edges(n, m, g, seed, weights, draw, log)


def chung_lu_device(n, m, gamma=2.2, seed=1, weights="uniform", chunk=1 << 27, log=None):
    """Chung-Lu on the GPU (torch): the law of chung_lu() (endpoints drawn
    in proportion to expected degrees i^(-1/(gamma-1)), hubs at random ids),
    with rmat_device()'s dedupe, subset, patching and orientation."""
    import torch
    dev = torch.device("cuda")
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    ew = torch.arange(1, n + 1, dtype=torch.float64, device=dev).pow(-1.0 / (gamma - 1.0))
    cdf = torch.cumsum(ew / ew.sum(), 0)
    cdf[-1] = 1.0
    perm = torch.randperm(n, generator=g, device=dev)

    def draw(need):
        for lo in range(0, need, chunk):
            cnt = min(chunk, need - lo)
            ends = []
            for _ in range(2):
                r = torch.rand(cnt, generator=g, device=dev, dtype=torch.float64)
                ends.append(perm[torch.searchsorted(cdf, r, right=True).clamp(0, n - 1)])
            yield ends[0], ends[1]

    return _device_edges(n, m, g, seed, weights, draw, log)


def chung_lu(n, m, gamma=2.2, seed=1, weights="uniform"):
    rng = np.random.default_rng(seed)
    ew = np.power(np.arange(1, n + 1, dtype=np.float64), -1.0 / (gamma - 1.0))
    p = ew / ew.sum()
    cdf = np.cumsum(p)
    cdf[-1] = 1.0
    perm = rng.permutation(n).astype(np.uint64)  # hubs at random ids
    keys = np.empty(0, np.uint64)
    while len(keys) < m:
        need = int((m - len(keys)) * 1.3) + 1024
        s = perm[np.searchsorted(cdf, rng.random(need), side="right").clip(0, n - 1)]
        d = perm[np.searchsorted(cdf, rng.random(need), side="right").clip(0, n - 1)]
        keys = _dedupe_keys(np.concatenate([keys, _canon(s, d, n)]), rng)
    return _finish(_fit_patches(keys, n, m), n, rng, weights)


def write_bin(path, src, dst, w):
    rec = np.empty(len(src), EDGE_DTYPE)
    rec["src"], rec["dst"], rec["data"] = src, dst, w
    rec.tofile(path)


def read_bin(path):
    rec = np.fromfile(path, EDGE_DTYPE)
    return rec["src"].copy(), rec["dst"].copy(), rec["data"].copy()


def lp_split(src, dst, w, n, frac=0.1, seed=7):
    """Held-out positives and sampled non-edges (src/tools/gtrain.cpp:292-343).

    Returns (train_src, train_dst, train_w, pos (k,2), neg (k,2))."""
    rng = np.random.default_rng(seed)
    m = len(src)
    k = int(frac * m)
    deg = np.bincount(src, minlength=n) + np.bincount(dst, minlength=n)
    order = rng.permutation(m)
    held = np.zeros(m, bool)
    got = 0
    for e in order:
        if got == k:
            break
        u, v = src[e], dst[e]
        if deg[u] <= 1 or deg[v] <= 1:
            continue
        deg[u] -= 1
        deg[v] -= 1
        held[e] = True
        got += 1
    pos = np.stack([src[held], dst[held]], 1)
    edge_keys = set(_canon(src, dst, n).tolist())
    neg = []
    seen = set()
    while len(neg) < got:
        u, v = (int(x) for x in rng.integers(0, n, 2))
        if u == v:
            continue
        key = min(u, v) * n + max(u, v)
        if key in edge_keys or key in seen:
            continue
        seen.add(key)
        neg.append((u, v))
    keep = ~held
    return src[keep], dst[keep], w[keep], pos.astype(np.uint32), np.asarray(neg, np.uint32).reshape(-1, 2)


def device_graph_file(gen, n, m, seed, weights, path, timeout=900):
    """Generate with a device generator ("rmat_device" or "chung_lu_device")
    in a child process and write the binary edge file to `path`.  The child
    owns the GPU only while it generates, so a process that already drives
    the GPU through libfelog_hip.so never loads a second HIP runtime."""
    import subprocess
    import sys
    here = os.path.dirname(os.path.abspath(__file__))
    subprocess.run([sys.executable, os.path.join(here, "graphgen.py"), gen, str(n), str(m), str(seed), weights,
                    path], check=True, timeout=timeout)
    return path


if __name__ == "__main__":
    import sys
    gen_name, n_, m_, seed_, law_, out_ = sys.argv[1:7]
    t0 = __import__("time").perf_counter()
    s_, d_, w_ = globals()[gen_name](int(n_), int(m_), seed=int(seed_), weights=law_,
                                     log=lambda t: print(f"[graphgen] {t}", flush=True))
    write_bin(out_, s_, d_, w_)
    print(f"[graphgen] {gen_name} n={n_}: {len(s_)} edges written to {out_} in "
          f"{__import__('time').perf_counter() - t0:.1f} s", flush=True)
