"""Error of each precision form against the FP64 oracle, over several training steps.

For every form (FP32 table; BF16 table with FP32 or FP16 adjoint rows; BF16 table + FP32
master with FP32 or FP16 adjoint rows) a d = 256 problem of the 50M-node workload's
shape (S = 20, K = 20, up to 32 features, B = 64) runs N explicit-sample steps on the GPU
(sparse Adam, lr 0.01).  Reported, as normwise relative errors max|x - x64| / max|x64|:

* loss and merged gradient of every step against the FP64 oracle evaluated on the
  parameters the kernels read at that step (the BF16 rows for the BF16 forms): the
  arithmetic error of the step (FP32 accumulation, FP16 adjoint rows);
* loss and merged gradient of every step against the FP64 oracle of an all-FP64 run from
  the same initial table on the same samples: the arithmetic plus the storage error
  (BF16 rows) accumulated over the steps;
* the final embeddings (the FP32 table, the FP32 master, or the BF16 table widened)
  against the all-FP64 run's, over the touched rows.

Bars: FP32 form <= 1e-5 on loss and gradients; BF16 forms <= 1e-3 on
gradients against the oracle on the same parameters.  With EP_ERROR_REPORT=<path> the
table is also written there (markdown).
"""
import os

import numpy as np
import pytest

from cloverleaf_amd import _native as N
from oracle import ep_oracle as O
from tests.problems import Problem, dense_grad, make_problem

pytestmark = pytest.mark.gpu

FORMS = [
    ("fp32", N.EP_PREC_FP32, 0),
    ("bf16 (fp32 adjoints)", N.EP_PREC_BF16, 0),
    ("bf16 (fp16 adjoints)", N.EP_PREC_BF16, 1),
    ("bf16-master (fp32 adjoints)", N.EP_PREC_BF16_MASTER, 0),
    ("bf16-master (fp16 adjoints)", N.EP_PREC_BF16_MASTER, 1),
]


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30))


def widen_bf16(u16):
    return (u16.astype(np.uint32) << 16).view(np.float32)


def run_form(prec, fp16_adj, steps=5):
    D, S, K, B, lr = 256, 20, 20, 64, 0.01
    prob = make_problem(seed=2600, n=500, V=1500, D=D, avg_deg=40, fmax=32, zipf=1.1)
    cfg = N.default_config()
    cfg.dims, cfg.batch_size, cfg.negatives, cfg.gamma = D, B, K, 0.5
    cfg.max_nodes_kind, cfg.max_nodes = N.EP_SAMPLE_FIXED, S
    cfg.precision, cfg.fp16_adjoints = prec, fp16_adj
    cfg.valid_pct = 0.0
    t64 = prob.table.astype(np.float64)  # the all-FP64 run
    m64, v64 = np.zeros_like(t64), np.zeros_like(t64)
    e = {"loss_same": 0.0, "grad_same": 0.0, "loss_fp64run": 0.0, "grad_fp64run": 0.0}
    touched = set()
    rng = np.random.default_rng(7)
    with N.Context(cfg) as ctx:
        ctx.set_graph(prob.indptr, prob.indices)
        ctx.set_features(prob.f_indptr, prob.f_ids, prob.V)
        ctx.set_table(prob.table)
        for step in range(steps):
            # what the forward reads at this step
            cur = ctx.get_table() if prec == N.EP_PREC_FP32 else widen_bf16(ctx.get_table_bf16())
            anchors = rng.choice(prob.n, size=B, replace=False).astype(np.uint32)
            cnt, nbr, neg = ctx.sample_batch(anchors, 1)
            out = ctx.step_explicit(anchors, cnt, nbr, neg, 1, lr, apply_update=True)
            Gg = np.zeros((prob.V, D))
            Gg[out["feat"]] = out["grad"]
            keep = [i for i in range(B) if cnt[i] > 0]
            args = ([int(anchors[i]) for i in keep], [list(map(int, nbr[i, :cnt[i]])) for i in keep],
                    [1.0] * len(keep), [list(map(int, neg[i])) for i in keep], 0.5)
            # oracle on the parameters the kernels read (no update of its own)
            ts = cur.astype(np.float64)
            res, merged = O.ep_batch_step(ts, np.zeros_like(ts), np.zeros_like(ts), prob.f_indptr, prob.f_ids,
                                          *args, 0.0, 1, "first_leaf")
            loss_ref = np.zeros(B)
            for i, r in zip(keep, res):
                loss_ref[i] = r.loss if r.contributes else 0.0
            e["loss_same"] = max(e["loss_same"], rel(out["loss"], loss_ref))
            e["grad_same"] = max(e["grad_same"], rel(Gg, dense_grad(merged, prob.V, D)))
            # the all-FP64 run
            res64, merged64 = O.ep_batch_step(t64, m64, v64, prob.f_indptr, prob.f_ids, *args, lr, 1, "first_leaf")
            loss64 = np.zeros(B)
            for i, r in zip(keep, res64):
                loss64[i] = r.loss if r.contributes else 0.0
            e["loss_fp64run"] = max(e["loss_fp64run"], rel(out["loss"], loss64))
            e["grad_fp64run"] = max(e["grad_fp64run"], rel(Gg, dense_grad(merged64, prob.V, D)))
            touched |= set(merged64)
        final = ctx.get_table() if prec != N.EP_PREC_BF16 else widen_bf16(ctx.get_table_bf16())
    rows = sorted(touched)
    e["emb_fp64run"] = rel(final[rows], t64[rows])
    return e


def test_error_against_fp64_per_form():
    lines = ["| form | loss (same parameters) | gradient (same parameters) | loss (FP64 run) | gradient (FP64 run) "
             "| embeddings after 5 steps (FP64 run) |", "|---|---|---|---|---|---|"]
    results = {}
    for name, prec, fp16_adj in FORMS:
        e = run_form(prec, fp16_adj)
        results[name] = e
        lines.append(f"| {name} | {e['loss_same']:.2e} | {e['grad_same']:.2e} | {e['loss_fp64run']:.2e} | "
                     f"{e['grad_fp64run']:.2e} | {e['emb_fp64run']:.2e} |")
        print(f"[errors] {lines[-1]}", flush=True)
    path = os.environ.get("EP_ERROR_REPORT")
    if path:
        with open(path, "w") as fh:
            fh.write("\n".join(lines) + "\n")
    fp32 = results["fp32"]
    assert fp32["loss_same"] <= 1e-5 and fp32["grad_same"] <= 1e-5
    for name, e in results.items():
        if name != "fp32":
            assert e["loss_same"] <= 1e-5, name
            assert e["grad_same"] <= 1e-3, name
