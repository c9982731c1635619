#!/usr/bin/env python3
"""Benchmark of the FeLoG walk -> SGNS hot path on one MI355X.

One step = one round of the path: a walk of every walker over the resident
graph (walker select, persistent walk kernel, corpus compaction), then
skip-gram training over that round's corpus (SGNS kernels).  The graph, alias
tables and both embedding tables are built in HBM before the timed region;
nothing crosses PCIe inside it.  The timed steps run as captured rounds
(felog_run_graph: one hipGraph replay per round, one host wait per round);
the per-kernel times come from as many eager rounds before them, with HIP
events around each launch.

Workload (default): Chung-Lu power law (exponent 2.2), 2.24M vertices,
~35M undirected edges, weights U(0,1], info-centric walks (min_L=20), plain
FP32 HogBatch SGNS with d=128, window 10, negative 5, batch 21.

Prints one JSON line.  `value` is corpus tokens walked and trained per
second of the whole round (plain FP32 form, captured rounds); the eager
rounds' rate, walk steps/s and SGNS words/s are reported beside it.  The roofline is that of the kernel that takes the
most time, from HIP events recorded on the library's stream around each
launch.  The `e2e` object holds the end-to-end comparison the metric names:
on the same graph with 10% of its edges held out, the plain FP32 form run
to the plain rule's stop (felog_run) against the BF16 form with the learned
walk budget (felog_run_budget; network trained on the plain rule's budgets
of the 10k Erdos-Renyi and 1.1M R-MAT graphs), with held-out AUC, corpus
size and visit entropy of both.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "info-centric_ge_amd"))

import felog  # noqa: E402
import graphgen  # noqa: E402

METRIC = "walk steps/s + SGNS words/s on 1 GPU; bf16+surrogate vs plain fp32 at equal AUC"
HBM_PEAK_GBS = 8000.0
# dense MFMA peaks: FP32 tables multiply on FP32 MFMA, BF16 tables on BF16 MFMA
MFMA_PEAK_TFS = {"fp32": 157.3, "bf16": 2500.0}

WORKLOADS = {
    # name: (generator, vertices, undirected edges, weight law)
    "chunglu-2.24M": ("chung_lu", 2_240_000, 35_000_000, "uniform"),
    "rmat-1.1M": ("rmat", 1_100_000, 3_000_000, "uniform"),
    # uneven-batching stress: degree max > 30k, lognormal(0, 1) weights
    "rmat-3.07M": ("rmat", 3_070_000, 117_000_000, "lognormal"),
    # learned-budget scale: ~1.2B directed edges, generated on the GPU
    "rmat-41.6M": ("rmat_device", 41_600_000, 600_000_000, "uniform"),
    "er-10k": ("erdos_renyi", 10_000, 50_000, "uniform"),
}


def make_graph(name, seed):
    gen, n, m, law = WORKLOADS[name]
    fn = getattr(graphgen, gen)
    if gen == "rmat_device":
        return fn(n, m, seed=seed, weights=law, log=lambda t: print(f"[bench] {t}", file=sys.stderr, flush=True)), n, m
    return fn(n, m, seed=seed, weights=law), n, m


def sgns_bytes_per_word(d, k, batch, esize):
    # target output row r+w, input row r+w (resident for its windows),
    # negatives shared by `batch` centres, the corpus token
    return 2 * d * esize + 2 * d * esize + 2 * k * d * esize / batch + 4


def measured_traffic(args):
    """HBM bytes per launch from the committed rocprofv3 counter summary of
    this workload (FETCH_SIZE and WRITE_SIZE passes scaled by a plain-copy
    calibration), or {} when there is none for it."""
    tag = {"rmat-1.1M": "rmat1.1M", "chunglu-2.24M": "chunglu2.24M", "rmat-3.07M": "rmat3.07M"}.get(args.workload)
    path = os.path.join(ROOT, "profiles", f"traffic_{tag}_{'f32' if args.precision == 'fp32' else 'bf16'}.json")
    if tag is None or args.dim != 128 or not os.path.exists(path):
        return {}
    with open(path) as fh:
        kernels = json.load(fh)["kernels"]
    return {name: round(v["hbm_bytes"]) for name, v in kernels.items()}


def _auc(f, pos, neg):
    """Held-out AUC on the device (felog_link_auc): cosine of syn0 rows,
    Mann-Whitney with ties at half weight."""
    pairs = np.concatenate([pos, neg]).astype(np.uint32)
    labels = np.concatenate([np.ones(len(pos), np.uint8), np.zeros(len(neg), np.uint8)])
    return f.link_auc(pairs, labels)[0]


def train_budget_model(args):
    """The learned walk budget: plain-rule budgets of the 10k Erdos-Renyi
    (alpha 0.01) and 1.1M R-MAT graphs, then Adam on the library's kernels."""
    import walk_budget
    Xs, Ts, info = [], [], {}
    t0 = time.perf_counter()
    for name, (gen, n, m), kw in (("er-10k", ("erdos_renyi", 10_000, 50_000), dict(alpha=0.01)),
                                  ("rmat-1.1M", ("rmat", 1_100_000, 3_000_000), {})):
        src, dst, w = getattr(graphgen, gen)(n, m, seed=args.seed if gen == "rmat" else 1, weights="uniform")
        f = felog.Felog(size=args.dim, window=10, negative=5, sample=1e-3, min_L=20, min_R=2, seed=args.seed, **kw)
        f.set_edges(src, dst, w, n, make_undirected=True)
        f.build()
        cnt, length, _, _ = walk_budget.plain_budgets(f, max_rounds=12)
        X, T = walk_budget.training_set(f.budget_features(), cnt, length)
        Xs.append(X)
        Ts.append(T)
        info[name] = {"count_histogram": np.bincount(cnt).tolist()[:8], "mean_length": float(length[cnt > 0].mean())}
        if name == "rmat-1.1M":
            model, losses = walk_budget.fit(f, np.concatenate(Xs), np.concatenate(Ts), epochs=20, batch=4096,
                                            lr=2e-3, seed=args.seed)
        f.close()
    info["samples"] = int(sum(len(x) for x in Xs))
    info["loss_first_last_epoch"] = [float(losses[0]), float(losses[-1])]
    info["seconds"] = time.perf_counter() - t0
    return model, info


def end_to_end(src, dst, w, n, args, model):
    """Plain FP32 run (felog_run's rule) against BF16 + learned budget
    (felog_run_budget's rule) on the graph with 10% of its edges held out;
    both as captured rounds (felog_run_graph).  With --e2e-no-auc the whole
    graph is trained and no AUC is measured (the host split is too slow for
    the 41.6M graph)."""
    if args.e2e_no_auc:
        ts, td, tw, pos, neg = src, dst, w, None, None
    else:
        ts, td, tw, pos, neg = graphgen.lp_split(src, dst, w, n, frac=0.1, seed=3)
    out = {}
    forms = (("plain_fp32", felog.FP32, None), ("bf16_surrogate", felog.BF16, model))
    if args.e2e_ablation:
        # the same budgeted run with the network replaced by one constant
        # budget (count min_R, length min_L + 2): what the network itself buys
        import walk_budget
        forms += (("bf16_constant_budget", felog.BF16, walk_budget.constant_model(2, 22)),)
    for form, prec, budget_model in forms:
        opts = dict(size=args.dim, window=10, negative=5, batch_size=21, sample=1e-3, min_L=20, min_R=2,
                    precision=prec, seed=args.seed)
        # tile shape for this precision, timed on a first-round corpus of a
        # separate context (a one-time calibration, not part of the run)
        f = felog.Felog(**opts)
        f.set_edges(ts, td, tw, n, make_undirected=True)
        f.build()
        f.walk_round(1)
        f.tune_sgns(sample=0, reps=1)
        shape = f.tile_shape()
        f.close()
        f = felog.Felog(**opts)
        f.set_edges(ts, td, tw, n, make_undirected=True)  # host edge list to the device: input I/O, not timed
        f.set_tile_shape(*(shape[k] for k in felog.SHAPE_FIELDS))
        extra = {}
        # timed: CSR + alias + vocabulary + tables, the budget's prediction,
        # every round to the stop
        t0 = time.perf_counter()
        f.build()
        t_build = time.perf_counter() - t0
        f.visits_reset()
        if budget_model is not None:
            f.budget_set(budget_model)
            # one convergence test at min_R for vertices whose budget ends there
            f.budget_recheck(True)
        t1 = time.perf_counter()
        f.run_graph()  # felog_run_budget / felog_run as captured rounds
        t2 = time.perf_counter()
        if budget_model is not None:
            cnt, ln, rounds = f.budget()
            extra = {"budgeted_vertices": int((cnt > 0).sum()), "plain_rule_fallback": int((cnt == 0).sum()),
                     "max_count": int(rounds), "recheck": True, "budget_set_ms": (t1 - t0 - t_build) * 1e3}
        st = f.stats()
        vs = f.visit_stats()
        auc = None
        if pos is not None:
            auc = _auc(f, pos, neg)
        print(f"[bench] e2e {form}: build {t_build * 1e3:.1f} ms + {st['rounds_done']} rounds "
              f"{(t2 - t1) * 1e3:.1f} ms, AUC {auc}", file=sys.stderr, flush=True)
        out[form] = dict(wall_ms=(t2 - t0) * 1e3, build_ms=t_build * 1e3, rounds_ms=(t2 - t1) * 1e3,
                         rounds=st["rounds_done"], tokens=vs["visits"], tokens_per_s=vs["visits"] / (t2 - t0),
                         auc=auc, visit_kl=vs["kl"], visit_entropy=vs["entropy"],
                         rounds_gpu_ms=st["graph_ms_total"],
                         tile_shape={k: shape[k] for k in felog.SHAPE_FIELDS}, **extra)
        f.close()
    p, b = out["plain_fp32"], out["bf16_surrogate"]
    out["throughput_ratio"] = p["wall_ms"] / b["wall_ms"]
    out["throughput_ratio_rounds_only"] = p["rounds_ms"] / b["rounds_ms"]
    out["auc_gap"] = None if p["auc"] is None else b["auc"] - p["auc"]
    out["definition"] = ("throughput_ratio = wall time of the plain FP32 run / wall time of the BF16 + "
                         "learned-budget run, each from felog_build (CSR, alias, vocabulary, tables; the "
                         "budget's prediction) to its stop on the same graph, graph input excluded; the "
                         "budgeted run scores the vertices whose budget ends at min_R once and walks those "
                         "not converged one more round; auc_gap = held-out AUC difference")
    return out


def auc_study(src, dst, w, n, args):
    """Held-out AUC of the trainers on the same corpora: two rounds of every
    walker on the graph with 10% of its edges held out, then the cosine AUC
    of syn0 (device scorer).  The walks depend only on the graph and the
    seed, so every trainer sees the same two corpora:
      tile trainer (HogBatch tiles), the reference's per-pair
      skip_gram_kernel rule on the GPU, and the CPU restatement of that
      kernel (the oracle, with the sigmoid table and the 1e8-entry unigram
      table), each at the tile trainer's base rate 0.0025 and at the
      reference's 0.025 (ExponentialDecayLR(0.025, 0.1))."""
    ts, td, tw, pos, neg = graphgen.lp_split(src, dst, w, n, frac=0.1, seed=3)
    pairs = np.concatenate([pos, neg]).astype(np.uint32)
    labels = np.concatenate([np.ones(len(pos), np.uint8), np.zeros(len(neg), np.uint8)])
    base = dict(size=args.dim, window=10, negative=5, batch_size=21, sample=1e-3, min_L=20, min_R=1 << 30,
                seed=args.seed)
    forms = (("tile_alpha_0.0025", dict(alpha=0.0025)), ("tile_alpha_0.025", dict(compat_lr=1)),
             ("skipgram_per_pair_gpu_alpha_0.0025", dict(trainer=felog.TRAIN_SKIPGRAM, alpha=0.0025)),
             ("skipgram_per_pair_gpu_alpha_0.025", dict(trainer=felog.TRAIN_SKIPGRAM, compat_lr=1)))
    out, corpora, rp = {}, [], None
    for name, kw in forms:
        f = felog.Felog(**base, **kw)
        f.set_edges(ts, td, tw, n, make_undirected=True)
        f.build()
        t0 = time.perf_counter()
        for r in (1, 2):
            f.walk_round(r)
            if len(corpora) < r:
                corpora.append(f.corpus()[:2])
            f.train_round()
        f.embeddings()  # waits for the device
        out[name] = {"auc": f.link_auc(pairs, labels)[0], "seconds": time.perf_counter() - t0}
        if rp is None:
            rp, _, _ = f.csr(weights=False)
        f.close()
        print(f"[bench] auc study {name}: {out[name]['auc']:.4f}", file=sys.stderr, flush=True)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc
    deg = np.diff(rp.astype(np.int64)).astype(np.uint32)
    order, rank, inv, vs, tw_ = orc.vocab(deg, 5)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    for base_rate in (0.0025, 0.025):
        c0 = orc.init_syn0(n, args.dim, rank)
        c1 = np.zeros((n, args.dim), np.float32)
        t0 = time.perf_counter()
        for r, (off, tok) in enumerate(corpora, start=1):
            alpha = np.float32(base_rate * np.exp(-0.1 * (r - 1)))
            orc.w2v_train_mode(c0, c1, off, tok, deg, order, rank, inv, vs, tw_, mode=orc.W2V_SKIPGRAM, window=10,
                               negative=5, sample=1e-3, alpha=alpha, table_size=100_000_000, threads=threads,
                               seed=args.seed, rnd=r)
        a, b = c0[pairs[:, 0]].astype(np.float64), c0[pairs[:, 1]].astype(np.float64)
        cos = (a * b).sum(1) / (np.linalg.norm(a, axis=1) * np.linalg.norm(b, axis=1) + 1e-30)
        name = f"skipgram_cpu_restatement_alpha_{base_rate}"
        out[name] = {"auc": orc.auc(cos[labels == 1].astype(np.float32), cos[labels == 0].astype(np.float32)),
                     "seconds": time.perf_counter() - t0, "threads": threads}
        print(f"[bench] auc study {name}: {out[name]['auc']:.4f}", file=sys.stderr, flush=True)
    # what the held-out pairs look like: degrees of positive and negative endpoints
    out["mean_endpoint_degree"] = {"held_out_edges": float(deg[pos].mean()), "non_edges": float(deg[neg].mean())}
    out["definition"] = ("two rounds of every walker on the graph with 10% of edges held out (and as many "
                         "non-edges), cosine of syn0, Mann-Whitney AUC; same corpora for every trainer")
    return out


def cpu_model():
    """CPU model, sockets and logical CPUs of this host (from lscpu)."""
    info = {}
    try:
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in txt.splitlines():
            k, _, v = line.partition(":")
            info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return {"model": info.get("Model name"), "sockets": info.get("Socket(s)"), "cpus": info.get("CPU(s)"),
            "threads_per_core": info.get("Thread(s) per core")}


def cpu_baseline(src, dst, w, n, args, seed, threads):
    """The oracle (CPU restatement) timed on the host cores over a bounded
    sample of walkers: walks, then word2vec skip-gram on their corpus; the
    median of three runs of each."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc
    rp, col, wo = orc.build_csr(n, src, dst, w, True)
    th, al = orc.build_alias(n, rp, wo)
    deg = np.diff(rp.astype(np.int64)).astype(np.uint32)
    order, rank, inv, vs, tw = orc.vocab(deg, 5)
    sample = min(n, args.cpu_walkers)
    t_walks, t_trains = [], []
    for rep in range(3):
        t0 = time.perf_counter()
        off, tok, wid, steps = orc.walks(n, rp, col, th, al, None, sample, 20, 1000, 1, seed, threads=threads)
        t_walks.append(time.perf_counter() - t0)
        rng = np.random.default_rng(0)
        syn0 = ((rng.random((n, args.dim)) - 0.5) / args.dim).astype(np.float32)
        syn1 = np.zeros((n, args.dim), np.float32)
        t0 = time.perf_counter()
        words, pairs = orc.w2v_train(syn0, syn1, off, tok, deg, order, rank, inv, vs, tw, window=10, negative=5,
                                     sample=1e-3, alpha=0.025, use_lut=True, threads=threads)
        t_trains.append(time.perf_counter() - t0)
        del syn0, syn1
    t_walk, t_train = float(np.median(t_walks)), float(np.median(t_trains))
    tokens = int(off[-1])
    return {
        "value": tokens / (t_walk + t_train),
        "unit": "tokens/s (walk + SGNS, end to end)",
        "cores": threads,
        "kind": "port",
        "sample": f"{sample} walkers of the same graph ({tokens} tokens); walk {steps / t_walk:.4g} steps/s, "
                  f"word2vec skip-gram {words / t_train:.4g} words/s; median of 3 runs",
        "walk_steps_per_s": steps / t_walk,
        "sgns_words_per_s": words / t_train,
        "runs_s": {"walk": t_walks, "word2vec": t_trains},
        "host": cpu_model(),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="chunglu-2.24M", choices=sorted(WORKLOADS))
    ap.add_argument("--precision", default="fp32", choices=["fp32", "bf16"])
    ap.add_argument("--dim", type=int, default=128)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--cpu-walkers", type=int, default=200_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-e2e", action="store_true", help="skip the end-to-end plain vs BF16 + budget runs")
    ap.add_argument("--e2e-no-auc", action="store_true", help="end-to-end runs on the whole graph, no AUC")
    ap.add_argument("--e2e-ablation", action="store_true",
                    help="also run the budgeted form with a constant budget instead of the network")
    ap.add_argument("--lost-updates", action="store_true",
                    help="after the timed rounds, one more round with the tile trainer's lost-update tracking on")
    ap.add_argument("--auc-study", action="store_true",
                    help="also compare the trainers' held-out AUC on the same corpora (tile at 0.0025 and 0.025, "
                         "per-pair skip-gram on the GPU and the CPU restatement at 0.025)")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default tile shape (profiling runs: one tile kernel variant)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the whole graph and both tables live in one GPU's HBM)", file=sys.stderr)
        sys.exit(2)

    t0 = time.perf_counter()
    (src, dst, w), n, m = make_graph(args.workload, args.seed)
    print(f"[bench] graph {args.workload}: {n} vertices, {len(src)} edges ({time.perf_counter() - t0:.1f}s)",
          file=sys.stderr, flush=True)
    prec = felog.BF16 if args.precision == "bf16" else felog.FP32
    esize = 2 if prec == felog.BF16 else 4
    # every round walks every vertex: no convergence evaluation in the steps
    f = felog.Felog(size=args.dim, window=10, negative=5, batch_size=21, sample=1e-3, min_L=20,
                    min_R=1 << 30, precision=prec, seed=args.seed)
    f.set_edges(src, dst, w, n, make_undirected=True)
    f.build()
    st = f.stats()
    print(f"[bench] build {st['build_ms']:.1f} ms, CSR entries {st['e_num']}", file=sys.stderr, flush=True)

    def one_round(r):
        c = f.walk_round(r)
        f.train_round()
        return c

    for r in range(1, args.warmup + 1):
        one_round(r)
    if args.warmup == 0 and not args.no_tune:
        f.walk_round(1)  # a corpus for the autotuner; round 1 is walked again when timed
    # the autotuner times every tile shape on the last warm-up corpus (tables
    # restored after each pass) and installs the fastest
    default_shape = f.tile_shape()
    t0 = time.perf_counter()
    tuned = [] if args.no_tune else f.tune_sgns(sample=0, reps=2)
    tune_s = time.perf_counter() - t0
    chosen = f.tile_shape()
    print(f"[bench] tile autotune {len(tuned)} shapes in {tune_s:.1f}s: chosen {chosen}", file=sys.stderr,
          flush=True)
    tok = steps = words = kept = pairs = 0
    walk_ms = comp_ms = train_ms = prep_ms = tile_ms = 0.0
    t0 = time.perf_counter()
    for r in range(args.warmup + 1, args.warmup + args.steps + 1):
        c = one_round(r)
        s = f.stats()
        tok += c["n_tokens"]
        steps += c["n_steps"]
        words += s["train_words"]
        kept += s["train_kept"]
        pairs += s["train_pairs"]
        walk_ms += s["walk_kernel_ms"]
        comp_ms += s["compact_kernel_ms"]
        train_ms += s["train_kernel_ms"]
        prep_ms += s["sgns_prep_ms"]
        tile_ms += s["sgns_tile_ms"]
    eager_elapsed = time.perf_counter() - t0
    eager_tok = tok

    # timed region: K captured rounds (inputs resident; run_graph returns
    # after its last round's host wait)
    f.run_graph(1)  # capture + one warm replay
    tok0 = f.stats()["walk_tokens_total"]
    t0 = time.perf_counter()
    f.run_graph(args.steps)
    elapsed = time.perf_counter() - t0
    tok = f.stats()["walk_tokens_total"] - tok0

    K = args.steps
    walk_gbs = steps * 24 / (walk_ms / 1e3) / 1e9
    # the walk's access against what the device returns for it: independent
    # random 32-byte reads over a table the size of the edge records
    erec_bytes = f.stats()["e_num"] * 32
    probe_rate = felog.probe_random_reads(erec_bytes, 64_000_000, 32)
    walk_rate = steps / (walk_ms / 1e3)
    print(f"[bench] random 32-byte reads over {erec_bytes / 1e9:.2f} GB: {probe_rate / 1e9:.1f}G/s; "
          f"walk {walk_rate / 1e9:.1f}G steps/s", file=sys.stderr, flush=True)
    # the tile kernel moves the rows; the prep kernel reads the 4-byte tokens
    tile_bytes_word = sgns_bytes_per_word(args.dim, 5, 21, esize) - 4
    sgns_gbs = kept * tile_bytes_word / (tile_ms / 1e3) / 1e9
    sgns_tfs = pairs * 6 * 6 * args.dim / (tile_ms / 1e3) / 1e12
    traffic = measured_traffic(args)
    if tile_ms >= walk_ms:
        roof = {"kernel": "k_sgns_tile", "bound": "hbm", "achieved": sgns_gbs, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": sgns_gbs / HBM_PEAK_GBS,
                "traffic": traffic.get("k_sgns_tile"),
                "algorithmic_bytes_per_launch": kept / K * tile_bytes_word,
                "avg_launch_ms": tile_ms / K}
    else:
        roof = {"kernel": "k_walk", "bound": "hbm", "achieved": walk_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": walk_gbs / HBM_PEAK_GBS, "traffic": traffic.get("k_walk"),
                "algorithmic_bytes_per_launch": steps / K * 24, "avg_launch_ms": walk_ms / K}
    out = {
        "metric": METRIC,
        "value": tok / elapsed,
        "unit": "tokens/s (walk + SGNS, end to end)",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": elapsed / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if prec == felog.FP32 else "bf16",
        "data": "synthetic (seeded generator; random-init tables)",
        "config": {"workload": f"{args.workload} {args.precision}", "vertices": n, "undirected_edges": int(len(src)),
                   "dim": args.dim, "window": 10, "negative": 5, "batch": 21, "min_L": 20, "sample": 1e-3},
        "eager_tokens_per_s": eager_tok / eager_elapsed,
        "eager_ms_per_step": eager_elapsed / K * 1e3,
        "walk_steps_per_s": steps / (walk_ms / 1e3),
        "sgns_words_per_s": words / (train_ms / 1e3),
        "walk_kernel_ms": walk_ms / K,
        "compact_kernel_ms": comp_ms / K,
        "sgns_kernel_ms": train_ms / K,
        "sgns_prep_kernel_ms": prep_ms / K,
        "sgns_tile_kernel_ms": tile_ms / K,
        "tokens_per_step": tok / K,
        "eager_tokens_per_step": eager_tok / K,
        "pairs_per_step": pairs / K,
        "walk_roofline": {"bound": "hbm", "achieved": walk_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": walk_gbs / HBM_PEAK_GBS, "bytes_per_step": 24,
                          "random_reads_per_s": probe_rate, "random_read_table_bytes": erec_bytes,
                          "steps_per_random_read": walk_rate / probe_rate},
        "sgns_roofline": {"hbm_GBs": sgns_gbs, "hbm_frac": sgns_gbs / HBM_PEAK_GBS, "mfma_TFs": sgns_tfs,
                          "mfma_peak_TFs": MFMA_PEAK_TFS[args.precision],
                          "mfma_frac": sgns_tfs / MFMA_PEAK_TFS[args.precision],
                          "bytes_per_kept_word": sgns_bytes_per_word(args.dim, 5, 21, esize),
                          "flop_per_pair": 6 * 6 * args.dim},
        "roofline": roof,
        "tile_autotune": {
            "chosen": {k: chosen[k] for k in felog.SHAPE_FIELDS},
            "default": {k: default_shape[k] for k in felog.SHAPE_FIELDS},
            "seconds": tune_s,
            "shapes": [[t[k] for k in felog.SHAPE_FIELDS] + [round(t["tile_ms"], 3)] for t in tuned],
            "fields": ", ".join(felog.SHAPE_FIELDS) + ", tile kernel ms on the warm-up corpus"},
        "cpu_baseline": None,
    }
    if args.lost_updates:
        # one more round with per-row versions: write-backs that replaced
        # another tile's update (not part of the timed region)
        f.track_lost_updates(True)
        r = args.warmup + args.steps + 2
        f.walk_round(r)
        f.train_round()
        s = f.stats()
        out["lost_updates"] = {"round": r, "rows_written": s["train_rows_written"], "rows_lost": s["train_rows_lost"],
                               "fraction": s["train_rows_lost"] / max(s["train_rows_written"], 1),
                               "tile_ms_with_tracking": s["sgns_tile_ms"]}
        f.track_lost_updates(False)
    f.close()
    if not args.no_e2e:
        t0 = time.perf_counter()
        model, info = train_budget_model(args)
        print(f"[bench] budget model trained ({info['seconds']:.1f}s)", file=sys.stderr, flush=True)
        out["e2e"] = end_to_end(src, dst, w, n, args, model)
        out["e2e"]["budget_model"] = info
        print(f"[bench] end-to-end comparison ({time.perf_counter() - t0:.1f}s)", file=sys.stderr, flush=True)
    if args.auc_study:
        out["auc_study"] = auc_study(src, dst, w, n, args)
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        out["cpu_baseline"] = cpu_baseline(src, dst, w, n, args, args.seed, threads)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
