"""Host side of the learned walk budget.

The plain information-centric rule spends, per vertex v, count(v) walks (one
per round until v converges or the run stops, src/word2vec.cu:2016-2080) of
mean length length(v) vertices (the entropy / Pearson stop rule,
include/walk.hpp:945-951,1056-1062).  `plain_budgets` runs that rule on a
built context and records both; `fit` trains the library's 8-64-64-2 MLP on
them (standardised features -> standardised ln count, ln length) with the
library's HIP kernels; `make_model` packs the result for `Felog.budget_set`.
"""
from __future__ import annotations

import time

import numpy as np

import felog

# Standard deviations beyond the training range still trusted.  The network
# is piecewise linear past its data; 2 keeps 25 of the 2.24M Chung-Lu
# vertices on the plain rule when trained on the 10k Erdos-Renyi and 1.1M
# R-MAT graphs, whose largest hub (~13k neighbours) is far below the
# Chung-Lu graph's (~530k): at 0.25, 1.64M vertices fell back on that one
# feature.
BOX_MARGIN = 2.0


def plain_budgets(f, max_rounds=20, evaluate=True):
    """Run the plain rule on built context `f` (rounds until its stop rule or
    max_rounds); return per-vertex walk counts and mean walk lengths, the
    corpora and the wall time of the rounds."""
    V = f.V
    cnt = np.zeros(V, np.int64)
    lsum = np.zeros(V, np.int64)
    corpora = []
    t0 = time.perf_counter()
    for r in range(1, max_rounds + 1):
        c = f.walk_round(r)
        if c["n_walks"] == 0:
            break
        off, tok, _ = f.corpus()
        f.train_round()
        corpora.append((off, tok))
        starts = tok[off[:-1].astype(np.int64)]
        np.add.at(cnt, starts, 1)
        np.add.at(lsum, starts, np.diff(off.astype(np.int64)))
        if evaluate and r >= f.opt.min_R:
            _, stop = f.evaluate()
            if stop:
                break
    wall = time.perf_counter() - t0
    length = np.where(cnt > 0, lsum / np.maximum(cnt, 1), 0.0)
    return cnt, length, corpora, wall


def training_set(features, counts, lengths):
    """Samples (vertices that walked at least once): raw features and the
    targets ln count, ln length."""
    keep = counts > 0
    X = np.asarray(features, np.float64)[keep]
    T = np.stack([np.log(counts[keep].astype(np.float64)), np.log(np.maximum(lengths[keep], 1.0))], 1)
    return X, T


def fit(f, X, T, epochs=30, batch=4096, lr=2e-3, seed=1, params=None):
    """Train on raw features X and targets T (both n x k, any float);
    returns (model, per-epoch loss).  Standardisation statistics and the
    trusted feature box come from X."""
    fm, fs = X.mean(0), X.std(0)
    fs[fs == 0] = 1.0
    tm, ts = T.mean(0), T.std(0)
    ts[ts == 0] = 1.0
    Xs = ((X - fm) / fs).astype(np.float32)
    Ts = ((T - tm) / ts).astype(np.float32)
    if params is None:
        rng = np.random.default_rng(seed)
        parts = []
        for name, shape in [("W1", (64, 8)), ("b1", (64,)), ("W2", (64, 64)), ("b2", (64,)), ("W3", (2, 64)),
                            ("b3", (2,))]:
            lim = np.sqrt(6.0 / shape[1]) if name[0] == "W" else 0.0
            parts.append(rng.uniform(-lim, lim, shape).ravel() if lim else np.zeros(shape).ravel())
        params = np.concatenate(parts).astype(np.float32)
    p, losses = f.budget_train(Xs, Ts, params, epochs, batch=batch, lr=lr, seed=seed)
    lo = Xs.min(0) - BOX_MARGIN
    hi = Xs.max(0) + BOX_MARGIN
    return make_model(p, fm, fs, lo, hi, tm, ts), losses


def make_model(params, feat_mean, feat_std, feat_lo, feat_hi, targ_mean, targ_std):
    m = felog.BudgetModel()
    m.params[:] = np.asarray(params, np.float32).tolist()
    for name, val in (("feat_mean", feat_mean), ("feat_std", feat_std), ("feat_lo", feat_lo),
                      ("feat_hi", feat_hi), ("targ_mean", targ_mean), ("targ_std", targ_std)):
        getattr(m, name)[:] = np.asarray(val, np.float32).tolist()
    return m


def constant_model(count, length, box=None):
    """A model predicting the same budget for every vertex (W3 = 0): for
    tests and as the method without the network."""
    p = np.zeros(felog.BUDGET_PARAMS, np.float32)
    lo, hi = (np.full(8, -1e30), np.full(8, 1e30)) if box is None else box
    return make_model(p, np.zeros(8), np.ones(8), lo, hi, [np.log(count), np.log(length)], [1.0, 1.0])
