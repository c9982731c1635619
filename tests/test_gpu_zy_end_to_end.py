"""End to end on the 2.24M-vertex Chung-Lu workload: the plain FP32 rule to
its stop against BF16 tables with the learned walk budget and the min_R
recheck, on the graph with 10% of its edges held out, each timed from
felog_build to its stop (the comparison bench.py reports in its `e2e`
object, run through the same functions).

Asserts: held-out AUC within 0.005 of the plain rule's either side, the
budgeted form's wall time below the plain form's by at least 1.3x (the
measured ratio is printed; bench.py records it per run), a smaller corpus,
and visit entropy within 1%.  The network is trained first on the plain
rule's budgets of the 10k Erdos-Renyi and 1.1M R-MAT graphs."""
import os
import sys
import types

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

pytestmark = pytest.mark.gpu


def test_bf16_learned_budget_end_to_end_chunglu_2_24m():
    import bench
    args = types.SimpleNamespace(dim=128, seed=2024, e2e_no_auc=False, e2e_ablation=False)
    print("[case] chunglu-2.24M plain FP32 vs BF16 + learned budget + recheck", flush=True)
    (src, dst, w), n, _ = bench.make_graph("chunglu-2.24M", args.seed)
    model, info = bench.train_budget_model(args)
    print(f"budget model: {info['samples']} samples, loss {info['loss_first_last_epoch']}", flush=True)
    out = bench.end_to_end(src, dst, w, n, args, model)
    p, b = out["plain_fp32"], out["bf16_surrogate"]
    print(f"plain: {p['rounds']} rounds, {p['wall_ms']:.1f} ms, AUC {p['auc']:.4f}, {p['tokens']} tokens; "
          f"budget: {b['rounds']} rounds, {b['wall_ms']:.1f} ms, AUC {b['auc']:.4f}, {b['tokens']} tokens; "
          f"ratio {out['throughput_ratio']:.3f}, gap {out['auc_gap']:+.4f}", flush=True)
    assert abs(out["auc_gap"]) <= 0.005
    assert out["throughput_ratio"] >= 1.3
    assert b["tokens"] < p["tokens"]
    assert abs(b["visit_entropy"] - p["visit_entropy"]) <= 0.01 * p["visit_entropy"]
