"""GPU tests of the captured round (felog_run_graph): every replay of the
graph against the eager calls it stands for.

Walks depend only on (seed, round, convergence flags), so while no vertex has
converged each replay's corpus must equal the eager walk_round's bit for
bit.  Training is Hogwild (tiles that share rows race), so tables are
compared bit for bit on a graph whose tiles share no rows: disjoint cycles,
one walker per cycle, walks of one tile and no negatives."""
import numpy as np
import pytest

import felog
import graphgen
import walk_budget

pytestmark = pytest.mark.gpu
SEED = 0x6A4F


def ctx(src, dst, w, n, **kw):
    opts = dict(size=64, window=10, negative=5, sample=1e-3, min_L=20, seed=SEED)
    opts.update(kw)
    print(f"[case] V={n} m={len(src)} {opts}", flush=True)
    f = felog.Felog(**opts)
    f.set_edges(src, dst, w, n, make_undirected=True)
    f.build()
    return f


def cycles(k=512, length=24):
    """k disjoint cycles of `length` vertices; cycle j holds vertices
    j, j + k, j + 2k, ..., so walkers 0 .. k-1 start one per cycle."""
    j = np.repeat(np.arange(k), length)
    t = np.tile(np.arange(length), k)
    src = (j + k * t).astype(np.uint32)
    dst = (j + k * ((t + 1) % length)).astype(np.uint32)
    return src, dst, np.ones(len(src), np.float32), k * length


# one tile per walk (22 vertices, window 10), no shared negative rows
RACE_FREE = dict(walkers=512, max_walk_len=22, negative=0, min_count=1)


@pytest.fixture(scope="module")
def rmat():
    n = 20_000
    return graphgen.rmat(n, 100_000, seed=31) + (n,)


def test_every_replay_walks_the_eager_corpus(rmat):
    """Four replays, one call each, with copies and host work between them:
    each round's corpus (offsets, tokens, walker ids) equals eager
    walk_round(r) on a second context.  min_R = 10 keeps the flags clear."""
    *g, n = rmat
    fg = ctx(*g, n, min_R=10)
    fe = ctx(*g, n, min_R=10)
    try:
        for r in range(1, 5):
            fg.run_graph(1)
            og, tg, wg = fg.corpus()
            fe.walk_round(r)
            fe.train_round()
            oe, te, we = fe.corpus()
            print(f"round {r}: {len(og) - 1} walks, {int(og[-1])} tokens", flush=True)
            assert fg.stats()["rounds_done"] == r
            assert np.array_equal(og, oe) and np.array_equal(tg, te) and np.array_equal(wg, we)
    finally:
        fg.close()
        fe.close()


def test_replays_train_like_the_eager_rounds():
    """Three replays against three eager rounds on race-free tiles: the
    tables agree bit for bit (same corpora, learning rates and random
    draws), and so do the training counters."""
    *g, n = cycles()
    kw = dict(min_R=10, **RACE_FREE)
    fg = ctx(*g, n, **kw)
    fe = ctx(*g, n, **kw)
    try:
        fe_init = fe.embeddings()
        fg.run_graph(3)
        for r in (1, 2, 3):
            fe.walk_round(r)
            fe.train_round()
        sg, se = fg.stats(), fe.stats()
        for k in ("train_words", "train_kept", "train_pairs", "train_groups", "walk_tokens", "walk_steps"):
            assert sg[k] == se[k], k
        eg, ug = fg.embeddings()
        ee, ue = fe.embeddings()
        print(f"pairs {sg['train_pairs']}; rows changed {(eg != fe_init[0]).any(1).sum()}; "
              f"graph {sg['graph_ms_total']:.2f} ms, eager {se['walk_ms_total'] + se['train_ms_total']:.2f} ms")
        assert sg["train_pairs"] > 0 and (ug != fe_init[1]).any()
        assert np.array_equal(eg, ee) and np.array_equal(ug, ue)
    finally:
        fg.close()
        fe.close()


def test_stop_rule_runs_on_the_device():
    """The plain rule to its stop on race-free tiles: the captured run stops
    in the same round as the eager felog_run, with the same unconverged
    count, convergence flags and tables."""
    *g, n = cycles()
    kw = dict(min_R=2, max_rounds=12, alpha=0.025, **RACE_FREE)
    fg = ctx(*g, n, **kw)
    fe = ctx(*g, n, **kw)
    try:
        fg.run_graph()
        fe.run()
        sg, se = fg.stats(), fe.stats()
        print(f"captured: {sg['rounds_done']} rounds, {sg['n_unconverged']} unconverged; "
              f"eager: {se['rounds_done']} rounds, {se['n_unconverged']} unconverged")
        assert sg["rounds_done"] == se["rounds_done"]
        assert sg["n_unconverged"] == se["n_unconverged"]
        fl_g, sc_g = fg.flags()
        fl_e, sc_e = fe.flags()
        assert np.array_equal(fl_g, fl_e) and np.array_equal(sc_g, sc_e)
        assert np.array_equal(fg.embeddings()[0], fe.embeddings()[0])
    finally:
        fg.close()
        fe.close()


def test_budget_rounds_replay(rmat):
    """A constant budget (count 2, length 7, every vertex): the captured run
    takes two rounds, and its visit histogram equals the eager run's."""
    *g, n = rmat
    fg = ctx(*g, n, min_R=2)
    fe = ctx(*g, n, min_R=2)
    try:
        m = walk_budget.constant_model(2, 7)
        for f in (fg, fe):
            f.budget_set(m)
            f.visits_reset()
        fg.run_graph()
        fe.run_budget()
        vg, ve = fg.visit_stats(), fe.visit_stats()
        print(f"budget: captured {fg.stats()['rounds_done']} rounds, {vg['visits']} visits; "
              f"eager {fe.stats()['rounds_done']} rounds, {ve['visits']} visits")
        assert fg.stats()["rounds_done"] == fe.stats()["rounds_done"] == 2
        assert vg["visits"] == ve["visits"] and vg["unvisited"] == ve["unvisited"]
        assert abs(vg["kl"] - ve["kl"]) <= 1e-12 * max(1.0, abs(ve["kl"]))
        assert abs(vg["entropy"] - ve["entropy"]) <= 1e-12 * ve["entropy"]
        og, tg, _ = fg.corpus()
        oe, te, _ = fe.corpus()
        assert np.array_equal(og, oe) and np.array_equal(tg, te)
        assert (np.diff(og.astype(np.int64)) == 7).all()
    finally:
        fg.close()
        fe.close()


def test_budget_recheck_replays():
    """The budget recheck on race-free tiles (every vertex budgeted: count 2,
    full-length walks, min_R 2): the captured run and the eager
    felog_run_budget extend the same counts to 3, walk the same third round
    and train the same tables; flags are left alone."""
    *g, n = cycles()
    kw = dict(min_R=2, **RACE_FREE)
    fg = ctx(*g, n, **kw)
    fe = ctx(*g, n, **kw)
    try:
        m = walk_budget.constant_model(2, 22)
        for f in (fg, fe):
            f.budget_set(m)
            f.budget_recheck(True)
            f.visits_reset()
        fg.run_graph()
        fe.run_budget()
        cg, _, _ = fg.budget()
        ce, _, _ = fe.budget()
        walked = np.arange(RACE_FREE["walkers"])  # walker i starts at vertex i
        print(f"recheck: {int((cg == 3).sum())} of {n} counts extended, "
              f"{int((cg[walked] == 3).sum())} of the walked vertices; rounds {fg.stats()['rounds_done']}")
        assert np.array_equal(cg, ce)
        assert set(np.unique(cg).tolist()) <= {2, 3}
        want = 3 if (ce[walked] == 3).any() else 2
        assert fg.stats()["rounds_done"] == fe.stats()["rounds_done"] == want
        vg, ve = fg.visit_stats(), fe.visit_stats()
        assert vg["visits"] == ve["visits"]
        assert np.array_equal(fg.embeddings()[0], fe.embeddings()[0])
        assert not fg.flags()[0].any() and not fe.flags()[0].any()
    finally:
        fg.close()
        fe.close()


def test_bf16_replays(rmat):
    """BF16 tables: two replays walk the eager corpora and train tables of
    the same scale."""
    *g, n = rmat
    fg = ctx(*g, n, min_R=10, precision=felog.BF16)
    fe = ctx(*g, n, min_R=10, precision=felog.BF16)
    try:
        fg.run_graph(2)
        for r in (1, 2):
            fe.walk_round(r)
            fe.train_round()
        og, tg, _ = fg.corpus()
        oe, te, _ = fe.corpus()
        assert np.array_equal(og, oe) and np.array_equal(tg, te)
        eg, _ = fg.embeddings()
        ee, _ = fe.embeddings()
        assert np.isfinite(eg).all()
        rg, re_ = np.linalg.norm(eg, axis=1).mean(), np.linalg.norm(ee, axis=1).mean()
        cos = (eg * ee).sum(1) / (np.linalg.norm(eg, axis=1) * np.linalg.norm(ee, axis=1) + 1e-12)
        print(f"bf16: mean row norm {rg:.4f} vs {re_:.4f}; median row cosine {np.median(cos):.4f}")
        assert abs(rg - re_) < 0.02 * re_
        assert np.median(cos) > 0.99
    finally:
        fg.close()
        fe.close()
