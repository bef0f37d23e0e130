"""GPU parity of the EP step at the shapes of the large workloads, against the FP64 oracle.

* d = 256 mean model at the 10M-node workload's shape (S = 32, K = 20, up to 128
  features per node, FP32, one batch and two concurrent batches) and at the 50M-node
  workload's shape (S = 20, K = 20, up to 64 features, BF16 table + FP32 master), both
  gradient modes.  These instantiate k_forward<256>, k_emit, k_seg_update<256, ...>,
  k_pieces<256> and k_long_combine<256> (hot rows past the 128-entry window).
  Reference: model.rs:257-301 (constructions), loss.rs:59-73, mod.rs:329-378 (first-leaf
  extraction), mod.rs:223-233 (merge), optimizer.rs:43-81 (Adam) at dims = 256.
* The attention model with blocks of 33-64 items (H = 4, d_k = 32, d_model = 128, up to
  64 features per node), FP32 and BF16, both modes: the two-tile path of the attention
  forward and backward (attention.rs:100-131, 171-191, model.rs:306-446).
* The 1M-node workload's model shape (d = 128, B = 1024, K = 10, S = 20) on its graph
  generator at 1% of the nodes, edges and vocabulary.

Samples are drawn by the GPU sampler and the same lists are fed to the oracle.  Bars as
in the other parity tests: FP32 loss, constructions and merged gradient <= 1e-5
(normwise relative); Adam m <= 1e-5, v <= 1e-4; theta within 1e-5 where the gradient is
above FP32 noise.  The BF16 forms are compared with the oracle run on the bf16 rows the
kernels read (FP32 accumulation: <= 1e-5).
"""
import numpy as np
import pytest

from cloverleaf_amd import _native as N
from tests.problems import Problem, dense_grad, make_problem, oracle_group_step, oracle_step

pytestmark = pytest.mark.gpu

TOL = 1e-5


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30))


def bf16_round_nearest(x):
    u = np.asarray(x, dtype=np.float32).view(np.uint32)
    return ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).view(np.float32)


def make_ctx(prob, S, K, B, gamma=0.5, grad_mode=N.EP_GRAD_FIRST_LEAF, precision=N.EP_PREC_FP32, concurrent=1,
             attention=None, seg_tile=(0, 0), fp16_adjoints=0):
    cfg = N.default_config()
    cfg.dims, cfg.batch_size, cfg.negatives, cfg.gamma = prob.D, B, K, gamma
    if attention is not None:
        heads, dk, dm = attention
        cfg.dims, cfg.attention_heads, cfg.attention_dk = dm, heads, dk
    cfg.max_nodes_kind, cfg.max_nodes = N.EP_SAMPLE_FIXED, S
    cfg.grad_mode, cfg.precision = grad_mode, precision
    cfg.concurrent_batches = concurrent
    cfg.seg_vpl, cfg.seg_unroll = seg_tile
    cfg.fp16_adjoints = fp16_adjoints
    cfg.valid_pct = 0.0
    ctx = N.Context(cfg)
    ctx.set_graph(prob.indptr, prob.indices)
    ctx.set_features(prob.f_indptr, prob.f_ids, prob.V)
    ctx.set_table(prob.table)
    return ctx


def mode_name(grad_mode):
    return "first_leaf" if grad_mode == N.EP_GRAD_FIRST_LEAF else "exact"


def check_adam(ctx, prob, merged, t64, m64, v64, table=None):
    new = ctx.get_table() if table is None else table
    m, v = ctx.get_adam()
    G64 = dense_grad(merged, prob.V, prob.D)
    rows = sorted(merged)
    assert rel(m[rows], m64[rows]) <= 1e-5
    assert rel(v[rows], v64[rows]) <= 1e-4
    mask = np.abs(G64[rows]) > 1e-3 * np.abs(G64).max()
    assert np.max(np.abs(new[rows] - t64[rows])[mask]) < 1e-5
    return rows


def long_rows(merged_counts, window=128):
    return sum(1 for c in merged_counts if c > window)


@pytest.mark.parametrize("grad_mode", [N.EP_GRAD_FIRST_LEAF, N.EP_GRAD_EXACT])
@pytest.mark.parametrize("fmax", [32, 128])
def test_d256_ten_million_node_shape(grad_mode, fmax):
    """S = 32, K = 20, d = 256, FP32, one batch (k_emit runs when F_max (1+S+K) needs the
    8192-entry first-leaf hash, i.e. fmax = 128)."""
    D, S, K, B = 256, 32, 20, 64
    prob = make_problem(seed=256 + fmax, n=500, V=2500, D=D, avg_deg=48, fmax=fmax, zipf=1.05)
    ctx = make_ctx(prob, S, K, B, grad_mode=grad_mode)
    anchors = np.random.default_rng(3).choice(prob.n, size=B, replace=False).astype(np.uint32)
    cnt, nbr, neg = ctx.sample_batch(anchors, 1)
    assert np.any(cnt == S)
    print(f"[shapes] d=256 S={S} K={K} fmax={fmax} mode={grad_mode}", flush=True)
    out = ctx.step_explicit(anchors, cnt, nbr, neg, pass_=1, lr=0.01, apply_update=True, want_cons=True)
    res, merged, t64, m64, v64 = oracle_step(prob, anchors, cnt, nbr, neg, 0.5, 0.01, 1, mode_name(grad_mode))
    ref_loss = np.array([r.loss if (r is not None and r.contributes) else 0.0 for r in res])
    assert np.count_nonzero(ref_loss) > B // 2
    assert rel(out["loss"], ref_loss) <= TOL
    for i, r in enumerate(res):
        if r is not None:
            assert rel(out["cons"][i], np.stack(r.vectors)) <= TOL
    G64 = dense_grad(merged, prob.V, prob.D)
    Gg = np.zeros_like(G64)
    assert len(set(out["feat"].tolist())) == len(out["feat"])
    assert set(out["feat"].tolist()) == set(merged)
    Gg[out["feat"]] = out["grad"]
    err = rel(Gg, G64)
    print(f"[shapes] grad rel err {err:.2e} over {len(merged)} rows", flush=True)
    assert err <= TOL
    check_adam(ctx, prob, merged, t64, m64, v64)
    ctx.close()


@pytest.mark.parametrize("grad_mode", [N.EP_GRAD_FIRST_LEAF, N.EP_GRAD_EXACT])
def test_d256_two_concurrent_batches(grad_mode):
    """Two batches of the 10M-node shape in one step: both read the step-start table, the
    per-(row, batch) gradients are merged separately and applied in batch order; hot
    rows run past the 128-entry window (long-row pieces at d = 256)."""
    D, S, K, B, P = 256, 32, 20, 40, 2
    prob = make_problem(seed=2562, n=400, V=600, D=D, avg_deg=48, fmax=128, zipf=1.05)
    ctx = make_ctx(prob, S, K, B, grad_mode=grad_mode, concurrent=P)
    anchors = np.random.default_rng(4).choice(prob.n, size=P * B - 3, replace=False).astype(np.uint32)
    cnt, nbr, neg = ctx.sample_batch(anchors, 1)
    print(f"[shapes] d=256 P={P} mode={grad_mode}", flush=True)
    out = ctx.step_explicit(anchors, cnt, nbr, neg, 1, 0.01, apply_update=True)
    merged, t64, m64, v64 = oracle_group_step(prob, anchors, cnt, nbr, neg, 0.5, 0.01, 1, mode_name(grad_mode), B)
    rows_g = out["feat"] & ((1 << 27) - 1)
    bat_g = out["feat"] >> 27
    for k in range(P):
        Gk = dense_grad(merged[k], prob.V, prob.D)
        sel = bat_g == k
        assert set(rows_g[sel].tolist()) == set(merged[k])
        Gg = np.zeros_like(Gk)
        Gg[rows_g[sel]] = out["grad"][sel]
        assert rel(Gg, Gk) <= TOL
    assert len(set(merged[0]) & set(merged[1])) > 50
    new = ctx.get_table()
    m, v = ctx.get_adam()
    rows = sorted(set(merged[0]) | set(merged[1]))
    assert rel(m[rows], m64[rows]) <= 1e-5
    assert rel(v[rows], v64[rows]) <= 1e-4
    gmax = max(np.abs(dense_grad(mk, prob.V, prob.D)).max() for mk in merged)
    mask = np.ones((len(rows), prob.D), dtype=bool)
    for mk in merged:
        Gk = dense_grad(mk, prob.V, prob.D)[rows]
        present = np.array([r in mk for r in rows])[:, None]
        mask &= ~present | (np.abs(Gk) > 1e-3 * gmax)
    assert mask.sum() > 1000
    assert np.max(np.abs(new[rows] - t64[rows])[mask]) < 1e-5
    untouched = np.setdiff1d(np.arange(prob.V), rows)
    assert np.array_equal(new[untouched], prob.table[untouched])
    ctx.close()


@pytest.mark.parametrize("grad_mode", [N.EP_GRAD_FIRST_LEAF, N.EP_GRAD_EXACT])
@pytest.mark.parametrize("precision,fp16_adj", [(N.EP_PREC_BF16_MASTER, 0), (N.EP_PREC_BF16, 0),
                                                (N.EP_PREC_BF16_MASTER, 1), (N.EP_PREC_BF16, 1)])
def test_d256_fifty_million_node_shape_bf16(grad_mode, precision, fp16_adj):
    """S = 20, K = 20, d = 256, up to 64 features per node, BF16 table (+ FP32 master).
    With FP16 adjoint rows the merged gradient keeps the BF16 form's 1e-3 bar (each
    adjoint value is rounded to 11 significant bits relative to its row's largest)."""
    D, S, K, B = 256, 20, 20, 64
    prob = make_problem(seed=2564, n=500, V=2000, D=D, avg_deg=40, fmax=64, zipf=1.1)
    ctx = make_ctx(prob, S, K, B, grad_mode=grad_mode, precision=precision, fp16_adjoints=fp16_adj)
    anchors = np.random.default_rng(5).choice(prob.n, size=B, replace=False).astype(np.uint32)
    cnt, nbr, neg = ctx.sample_batch(anchors, 1)
    print(f"[shapes] d=256 bf16 precision={precision} mode={grad_mode}", flush=True)
    out = ctx.step_explicit(anchors, cnt, nbr, neg, 1, 0.01, apply_update=True)
    tb = bf16_round_nearest(prob.table)
    pb = Problem(prob.indptr, prob.indices, prob.f_indptr, prob.f_ids, prob.V, prob.D, tb)
    res, merged, t64, m64, v64 = oracle_step(pb, anchors, cnt, nbr, neg, 0.5, 0.01, 1, mode_name(grad_mode))
    ref_loss = np.array([r.loss if (r is not None and r.contributes) else 0.0 for r in res])
    assert rel(out["loss"], ref_loss) <= TOL
    G64 = dense_grad(merged, prob.V, prob.D)
    Gg = np.zeros_like(G64)
    assert set(out["feat"].tolist()) == set(merged)
    Gg[out["feat"]] = out["grad"]
    err = rel(Gg, G64)
    print(f"[shapes] bf16 grad rel err {err:.2e} (bf16 rows, FP32 accumulation, fp16 adjoints {fp16_adj})", flush=True)
    assert err <= (1e-3 if fp16_adj else TOL)
    if fp16_adj:
        assert err > 1e-7  # the FP16 rows are really used
    rows = sorted(merged)
    m, v = ctx.get_adam()
    assert rel(m[rows], m64[rows]) <= (2e-3 if fp16_adj else 1e-5)
    assert rel(v[rows], v64[rows]) <= (4e-3 if fp16_adj else 1e-4)
    tbf = (ctx.get_table_bf16().astype(np.uint32) << 16).view(np.float32)
    untouched = np.setdiff1d(np.arange(prob.V), rows)
    assert np.array_equal(tbf[untouched], tb[untouched])
    mask = np.abs(G64[rows]) > (1e-2 if fp16_adj else 1e-3) * np.abs(G64).max()
    if precision == N.EP_PREC_BF16_MASTER:
        master = ctx.get_table()
        # master = FP32 start row + the Adam delta (the delta depends only on the gradient)
        expect = prob.table.astype(np.float64) + (t64 - tb.astype(np.float64))
        assert np.max(np.abs(master[rows] - expect[rows])[mask]) < (1e-4 if fp16_adj else 1e-5)
        diff = np.abs(tbf[rows] - master[rows])
        assert np.all(diff <= np.abs(master[rows]) * 2.0 ** -7 + 1e-30)
    else:
        # the stored value is a bf16 neighbour of the exact update: within one bf16 ulp
        # (of the update's or the stored value's binade, whichever is larger) of it, up to
        # the FP32 error of the Adam step itself (theta + step can cancel to a value far
        # below both, and the step's sensitivity to the gradient grows as |g| nears eps:
        # 1e-5 of the step size lr = 0.01)
        x = t64[rows][mask]
        got = tbf[rows][mask].astype(np.float64)
        ulp = np.exp2(np.floor(np.log2(np.maximum(np.maximum(np.abs(x), np.abs(got)), 1e-30))) - 7)
        over = np.maximum(np.abs(got - x) - 1e-7, 0.0) / ulp
        worst = int(np.argmax(over))
        print(f"[shapes] bf16 update: worst {over[worst]:.3f} ulp (x {x[worst]:.8f}, stored {got[worst]:.8f})",
              flush=True)
        assert over[worst] <= 1.0 + 1e-3
    ctx.close()


def att_items_problem(seed, heads=4, dk=32, dm=128, n=240, V=5000, fmax=64):
    """Nodes of 1..64 near-distinct features (flat Zipf over a large vocabulary), so that
    about half of the attention blocks hold 33-64 items."""
    W = heads * (2 * dk + dm)
    return make_problem(seed=seed, n=n, V=V, D=W, avg_deg=10, fmax=fmax, zipf=0.3, table_scale=0.6)


@pytest.mark.parametrize("grad_mode", [N.EP_GRAD_FIRST_LEAF, N.EP_GRAD_EXACT])
@pytest.mark.parametrize("precision", [N.EP_PREC_FP32, N.EP_PREC_BF16])
def test_attention_blocks_of_33_to_64_items(grad_mode, precision):
    heads, dk, dm = 4, 32, 128
    S, K, B = 5, 4, 24
    prob = att_items_problem(seed=64 + grad_mode + 2 * precision)
    sizes = np.diff(prob.f_indptr.astype(np.int64))
    assert np.count_nonzero(sizes > 32) > prob.n // 3
    ctx = make_ctx(prob, S, K, B, grad_mode=grad_mode, precision=precision, attention=(heads, dk, dm))
    anchors = np.random.default_rng(6).choice(prob.n, size=B, replace=False).astype(np.uint32)
    cnt, nbr, neg = ctx.sample_batch(anchors, 1)
    print(f"[shapes] attention 33-64 items precision={precision} mode={grad_mode}", flush=True)
    out = ctx.step_explicit(anchors, cnt, nbr, neg, 1, 0.01, apply_update=True, want_cons=True)
    table = prob.table if precision == N.EP_PREC_FP32 else bf16_round_nearest(prob.table)
    pb = Problem(prob.indptr, prob.indices, prob.f_indptr, prob.f_ids, prob.V, prob.D, table)
    res, merged, t64, m64, v64 = oracle_step(pb, anchors, cnt, nbr, neg, 0.5, 0.01, 1, mode_name(grad_mode),
                                             attention=(heads, dk))
    # blocks past 32 items: h(v) / negatives over distinct features, neighbours over listed ones
    big = sum(1 for v in anchors if len(set(prob.f_ids[prob.f_indptr[v]:prob.f_indptr[v + 1]].tolist())) > 32)
    assert big > 4
    ref_loss = np.array([r.loss if (r is not None and r.contributes) else 0.0 for r in res])
    assert np.count_nonzero(ref_loss) > 4
    assert rel(out["loss"], ref_loss) <= TOL
    for i, r in enumerate(res):
        if r is not None:
            assert rel(out["cons"][i], np.stack(r.vectors)) <= TOL
    G64 = dense_grad(merged, prob.V, prob.D)
    Gg = np.zeros_like(G64)
    assert set(out["feat"].tolist()) == set(merged)
    Gg[out["feat"]] = out["grad"]
    err = rel(Gg, G64)
    print(f"[shapes] attention grad rel err {err:.2e} over {len(merged)} rows", flush=True)
    assert err <= TOL
    if precision == N.EP_PREC_FP32:
        check_adam(ctx, prob, merged, t64, m64, v64)
    ctx.close()


@pytest.mark.parametrize("grad_mode", [N.EP_GRAD_FIRST_LEAF, N.EP_GRAD_EXACT])
def test_one_million_node_shape_scaled(grad_mode):
    """d = 128, B = 1024, K = 10, S = 20 on the bipartite query-item generator at 1%
    (3k queries, 7k items, 200k edges, 10k vocabulary, features U{1..32})."""
    from cloverleaf_amd import synth

    w = synth.make_workload("ep-cfg2-1m", scale=0.01)
    table = np.random.default_rng(9).uniform(-1, 1, size=(w.vocab, w.dims)).astype(np.float32)
    prob = Problem(w.indptr, w.indices, w.f_indptr, w.f_ids, w.vocab, w.dims, table)
    ctx = make_ctx(prob, w.max_nodes, w.negatives, w.batch, gamma=w.gamma, grad_mode=grad_mode)
    anchors = np.random.default_rng(10).choice(prob.n, size=w.batch, replace=False).astype(np.uint32)
    cnt, nbr, neg = ctx.sample_batch(anchors, 1)
    print(f"[shapes] 1M-node shape at 1%: B={w.batch} S={w.max_nodes} K={w.negatives} mode={grad_mode}", flush=True)
    out = ctx.step_explicit(anchors, cnt, nbr, neg, 1, w.alpha, apply_update=True)
    res, merged, t64, m64, v64 = oracle_step(prob, anchors, cnt, nbr, neg, w.gamma, w.alpha, 1, mode_name(grad_mode))
    ref_loss = np.array([r.loss if (r is not None and r.contributes) else 0.0 for r in res])
    assert rel(out["loss"], ref_loss) <= TOL
    G64 = dense_grad(merged, prob.V, prob.D)
    Gg = np.zeros_like(G64)
    assert set(out["feat"].tolist()) == set(merged)
    Gg[out["feat"]] = out["grad"]
    err = rel(Gg, G64)
    print(f"[shapes] grad rel err {err:.2e} over {len(merged)} rows", flush=True)
    assert err <= TOL
    check_adam(ctx, prob, merged, t64, m64, v64)
    ctx.close()
