"""ep_step_explicit_noalloc (step arrays in the library's per-GPU scratch, no device
allocation in the call) against ep_step_explicit with apply_update = 1 on a twin context:
the same explicit samples through both, two calls in a row (the second with fewer
anchors, so it runs over what the first left in the scratch), and the per-example loss,
table (FP32, BF16 and master rows) and Adam moments must agree bit for bit after each.
Also: no change of free device memory over a call, the capacity error for a step whose
arrays exceed the scratch (nothing launched, table unchanged) and the state error before
a table is set.
"""
import numpy as np
import pytest
import torch

from cloverleaf_amd import _native as N
from tests.problems import make_problem

pytestmark = pytest.mark.gpu

FORMS = [
    ("fp32 d=64", N.EP_PREC_FP32, 0, 64, 0),
    ("bf16-master fp16-adjoints d=256", N.EP_PREC_BF16_MASTER, 1, 256, 0),
    ("bf16 d=128 exact gradients", N.EP_PREC_BF16, 0, 128, 1),
]


def make_config(prec, fp16_adj, D, grad_mode, batch=64):
    cfg = N.default_config()
    cfg.dims, cfg.batch_size, cfg.negatives, cfg.gamma = D, batch, 20, 0.5
    cfg.max_nodes_kind, cfg.max_nodes = N.EP_SAMPLE_FIXED, 20
    cfg.precision, cfg.fp16_adjoints, cfg.grad_mode = prec, fp16_adj, grad_mode
    cfg.valid_pct = 0.0
    return cfg


def open_ctx(cfg, prob, table=True):
    ctx = N.Context(cfg)
    ctx.set_graph(prob.indptr, prob.indices)
    ctx.set_features(prob.f_indptr, prob.f_ids, prob.V)
    if table:
        ctx.set_table(prob.table)
    return ctx


def state(ctx, prec):
    m, v = ctx.get_adam()
    out = {"table": ctx.get_table(), "m": m, "v": v}
    if prec != N.EP_PREC_FP32:
        out["bf16"] = ctx.get_table_bf16()
    return out


@pytest.mark.parametrize("name,prec,fp16_adj,D,grad_mode", FORMS, ids=[f[0] for f in FORMS])
def test_noalloc_matches_step_explicit(name, prec, fp16_adj, D, grad_mode):
    prob = make_problem(seed=4100 + D, n=600, V=1500, D=D, avg_deg=40, fmax=32, zipf=1.1)
    cfg = make_config(prec, fp16_adj, D, grad_mode)
    rng = np.random.default_rng(D)
    with open_ctx(cfg, prob) as ref, open_ctx(cfg, prob) as na:
        for call, n in enumerate((64, 23)):
            print(f"[noalloc] {name}: call {call} with {n} anchors", flush=True)
            anchors = rng.choice(prob.n, size=n, replace=False).astype(np.uint32)
            cnt, nbr, neg = ref.sample_batch(anchors, call + 1)
            out = ref.step_explicit(anchors, cnt, nbr, neg, call + 1, 0.01, apply_update=True)
            loss = na.step_explicit_noalloc(anchors, cnt, nbr, neg, call + 1, 0.01)
            assert np.count_nonzero(out["loss"]) > n // 2
            assert np.array_equal(loss, out["loss"]), call
            a, b = state(ref, prec), state(na, prec)
            for k in a:
                assert np.array_equal(a[k], b[k]), (call, k)
            moved = np.flatnonzero(np.any(a["table"] != prob.table, axis=1))
            assert len(moved) > 100, call
        # a third call: the free device memory does not change over it
        anchors = rng.choice(prob.n, size=40, replace=False).astype(np.uint32)
        cnt, nbr, neg = ref.sample_batch(anchors, 3)
        torch.cuda.synchronize()
        free0, _ = torch.cuda.mem_get_info(0)
        na.step_explicit_noalloc(anchors, cnt, nbr, neg, 3, 0.01)
        torch.cuda.synchronize()
        free1, _ = torch.cuda.mem_get_info(0)
        assert free1 == free0


def test_noalloc_errors():
    D = 256
    prob = make_problem(seed=4200, n=300, V=800, D=D, avg_deg=40, fmax=32, zipf=1.1)
    cfg = make_config(N.EP_PREC_FP32, 0, D, 0, batch=4096)
    with open_ctx(cfg, prob, table=False) as ctx:
        anchors = np.arange(8, dtype=np.uint32)
        cnt = np.zeros(8, np.uint32)
        nbr = np.zeros((8, 20), np.uint32)
        neg = np.zeros((8, 20), np.uint32)
        with pytest.raises(N.NativeError, match="EP_ESTATE"):
            ctx.step_explicit_noalloc(anchors, cnt, nbr, neg, 1, 0.01)
        ctx.set_table(prob.table)
        # 4096 anchors of up to 32 * 41 feature rows: far more than the scratch holds
        anchors = np.arange(4096, dtype=np.uint32) % prob.n
        cnt = np.full(4096, 20, np.uint32)
        nbr = np.zeros((4096, 20), np.uint32)
        neg = np.zeros((4096, 20), np.uint32)
        with pytest.raises(N.NativeError, match="EP_ECAPACITY.*bytes of scratch"):
            ctx.step_explicit_noalloc(anchors, cnt, nbr, neg, 1, 0.01)
        assert np.array_equal(ctx.get_table(), prob.table)
