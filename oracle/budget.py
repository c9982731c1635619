"""FP64 restatement of the learned walk budget and the visit diagnostics.

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, always as the reference results are compared
against, never as the product.

The network is the library's 8-64-64-2 ReLU MLP with the same flat parameter
layout (row-major, out x in): W1 (64x8), b1, W2 (64x64), b2, W3 (2x64), b3.
"""
from __future__ import annotations

import numpy as np

F, H, O = 8, 64, 2
SIZES = [("W1", (H, F)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (O, H)), ("b3", (O,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SIZES)  # 4866


def unpack(p):
    p = np.asarray(p, np.float64)
    out, o = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = p[o:o + n].reshape(shape)
        o += n
    return out


def init_params(seed):
    """He-uniform weights, zero biases."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, shape in SIZES:
        if name.startswith("W"):
            lim = np.sqrt(6.0 / shape[1])
            parts.append(rng.uniform(-lim, lim, shape).ravel())
        else:
            parts.append(np.zeros(shape).ravel())
    return np.concatenate(parts).astype(np.float32)


def forward(p, X):
    """Outputs (n x 2) of the network for standardised features X (n x 8)."""
    q = unpack(p)
    X = np.asarray(X, np.float64)
    h1 = np.maximum(X @ q["W1"].T + q["b1"], 0.0)
    h2 = np.maximum(h1 @ q["W2"].T + q["b2"], 0.0)
    return h2 @ q["W3"].T + q["b3"]


def loss_and_grad(p, X, T):
    """Mean over samples of the squared error summed over the two outputs,
    and its gradient (flat, FP64) by backpropagation."""
    q = unpack(p)
    X = np.asarray(X, np.float64)
    T = np.asarray(T, np.float64)
    n = len(X)
    z1 = X @ q["W1"].T + q["b1"]
    h1 = np.maximum(z1, 0.0)
    z2 = h1 @ q["W2"].T + q["b2"]
    h2 = np.maximum(z2, 0.0)
    y = h2 @ q["W3"].T + q["b3"]
    e = y - T
    loss = float((e ** 2).sum() / n)
    dy = 2.0 * e / n
    g = {"W3": dy.T @ h2, "b3": dy.sum(0)}
    d2 = (dy @ q["W3"]) * (z2 > 0)
    g["W2"] = d2.T @ h1
    g["b2"] = d2.sum(0)
    d1 = (d2 @ q["W2"]) * (z1 > 0)
    g["W1"] = d1.T @ X
    g["b1"] = d1.sum(0)
    return loss, np.concatenate([g[name].ravel() for name, _ in SIZES])


def features(row_ptr, col, w, in_vocab):
    """Raw per-vertex features, as the library computes them."""
    V = len(row_ptr) - 1
    deg = np.diff(np.asarray(row_ptr, np.int64))
    out = np.zeros((V, F), np.float64)
    for v in range(V):
        b, e = int(row_ptr[v]), int(row_ptr[v + 1])
        ws = np.asarray(w[b:e], np.float64)
        nd = deg[np.asarray(col[b:e], np.int64)].astype(np.float64)
        s = ws.sum()
        out[v, 0] = np.log1p(deg[v])
        out[v, 1] = np.log1p(s)
        out[v, 2] = ws.max() if e > b else 0.0
        out[v, 3] = s / deg[v] if deg[v] else 0.0
        out[v, 4] = np.log1p(nd.sum())
        out[v, 5] = np.log1p(nd.max()) if e > b else 0.0
        if deg[v] > 1 and s > 0:
            pr = ws / s
            pr = pr[pr > 0]
            out[v, 6] = -(pr * np.log(pr)).sum() / np.log(deg[v])
        out[v, 7] = 1.0 if in_vocab[v] else 0.0
    return out


def visit_stats(visits, deg):
    """KL(p || q) over visited vertices, Shannon entropy of q (nats),
    vertices with degree and no visit; p = degree share, q = visit share
    (include/walk.hpp:754-788)."""
    visits = np.asarray(visits, np.float64)
    deg = np.asarray(deg, np.float64)
    q = visits / visits.sum()
    p = deg / deg.sum()
    vis = q > 0
    both = vis & (p > 0)
    kl = float((p[both] * (np.log(p[both]) - np.log(q[both]))).sum())
    ent = float(-(q[vis] * np.log(q[vis])).sum())
    return kl, ent, int(((p > 0) & ~vis).sum())


def standardise(A):
    """Column means and standard deviations (1 where a column is constant)."""
    A = np.asarray(A, np.float64)
    m = A.mean(0)
    s = A.std(0)
    s[s == 0] = 1.0
    return m, s
