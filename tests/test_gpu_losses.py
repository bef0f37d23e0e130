"""GPU parity of the EP step under the reference's other losses (Loss::compute,
loss.rs:56-151) and the PPR positives (random walks, loss.rs:191-226), through the
C ABI, against the FP64 oracle (oracle.ep_oracle.loss_adjoints).  Bars as the margin
tests: loss, constructions and merged gradient within 1e-5 (normwise relative)."""
import numpy as np
import pytest

from cloverleaf_amd import _native as N
from tests.problems import dense_grad, make_problem, oracle_step

pytestmark = pytest.mark.gpu

TOL = 1e-5


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30))


def loss_ctx(prob, kind, params, S=6, K=5, B=48, attention=None, precision=N.EP_PREC_FP32):
    cfg = N.default_config()
    cfg.dims, cfg.batch_size, cfg.negatives = (attention[2] if attention else prob.D), B, K
    if attention:
        cfg.attention_heads, cfg.attention_dk = attention[0], attention[1]
    cfg.loss_kind = N.LOSS_KINDS[kind]
    cfg.gamma = params[0]
    cfg.loss_param2 = params[1] if len(params) > 1 else 0.0
    cfg.max_nodes_kind, cfg.max_nodes = N.EP_SAMPLE_FIXED, S
    cfg.precision = precision
    cfg.valid_pct = 0.0
    ctx = N.Context(cfg)
    ctx.set_graph(prob.indptr, prob.indices)
    ctx.set_features(prob.f_indptr, prob.f_ids, prob.V)
    ctx.set_table(prob.table)
    return ctx


CASES = [("starspace", (0.4,)), ("contrastive", (0.6, 0.1)), ("rank", (0.5,)), ("rankspace", (0.5,)),
         ("ppr", (0.5, 0.3))]


def check(prob, ctx, kind, params, attention=None, seed=0):
    rng = np.random.default_rng(seed)
    anchors = rng.choice(prob.n, size=48, replace=False).astype(np.uint32)
    cnt, nbr, neg = ctx.sample_batch(anchors, 1)
    out = ctx.step_explicit(anchors, cnt, nbr, neg, pass_=1, lr=0.01, apply_update=True, want_cons=True)
    res, merged, t64, m64, v64 = oracle_step(prob, anchors, cnt, nbr, neg, params[0], 0.01, 1, "first_leaf",
                                             attention=attention[:2] if attention else None,
                                             loss=(kind, params))
    ref_loss = np.array([r.loss if (r is not None and r.contributes) else 0.0 for r in res])
    print(f"[losses] {kind}{params} att={attention}: {np.count_nonzero(ref_loss)} contributing examples", flush=True)
    assert np.count_nonzero(ref_loss) > 5
    assert rel(out["loss"], ref_loss) <= TOL
    for i, r in enumerate(res):
        if r is not None:
            assert rel(out["cons"][i], np.stack(r.vectors)) <= TOL
    G64 = dense_grad(merged, prob.V, prob.D)
    Gg = np.zeros_like(G64)
    assert set(out["feat"].tolist()) == set(merged.keys())
    Gg[out["feat"]] = out["grad"]
    assert rel(Gg, G64) <= TOL
    new = ctx.get_table()
    m, v = ctx.get_adam()
    rows = sorted(merged)
    assert rel(m[rows], m64[rows]) <= 1e-5
    assert rel(v[rows], v64[rows]) <= 1e-4
    return cnt, nbr


@pytest.mark.parametrize("kind,params", CASES)
def test_loss_matches_oracle_mean_model(kind, params):
    prob = make_problem(seed=50 + len(kind), n=150, V=200, D=32)
    ctx = loss_ctx(prob, kind, params)
    check(prob, ctx, kind, params)
    ctx.close()


@pytest.mark.parametrize("kind,params", CASES)
def test_loss_matches_oracle_attention_model(kind, params):
    heads, dk, dm = 2, 8, 16
    prob = make_problem(seed=60 + len(kind), n=150, V=120, D=heads * (2 * dk + dm), fmax=10, zipf=1.2,
                        table_scale=0.6)
    ctx = loss_ctx(prob, kind, params, attention=(heads, dk, dm))
    check(prob, ctx, kind, params, attention=(heads, dk, dm))
    ctx.close()


def test_ppr_walk_endpoints():
    # loss.rs:195-226: K walks from v; endpoints are never v while v has neighbours; a node
    # without neighbours is its own positive; endpoint frequencies follow the walk process
    from tests.problems import Problem
    n = 6
    # a path 0-1-2-3-4 (undirected) plus an isolated node 5
    edges = {(0, 1), (1, 0), (1, 2), (2, 1), (2, 3), (3, 2), (3, 4), (4, 3)}
    indptr = np.zeros(n + 1, dtype=np.uint64)
    for a, _ in edges:
        indptr[a + 1] += 1
    indptr = np.cumsum(indptr).astype(np.uint64)
    indices = np.array([b for a in range(n) for b in sorted(x for (y, x) in edges if y == a)], dtype=np.uint32)
    f_indptr = np.arange(n + 1, dtype=np.uint64)
    f_ids = np.arange(n, dtype=np.uint32)
    prob = Problem(indptr, indices, f_indptr, f_ids, n, 8, np.ones((n, 8), np.float32))
    K, restart = 8, 0.5
    ctx = loss_ctx(prob, "ppr", (0.5, restart), K=K, B=4096)
    anchors = np.array([1] * 4000 + [5] * 96, dtype=np.uint32)
    cnt, nbr, _ = ctx.sample_batch(anchors, 1)
    assert np.all(cnt[4000:] == 1) and np.all(nbr[4000:, 0] == 5)
    assert np.all(cnt[:4000] == K)
    ends = nbr[:4000].ravel()
    assert not np.any(ends == 1)
    # the walk process on the host (vectorised Monte Carlo with numpy), same rules
    rng = np.random.default_rng(0)
    M = 400000
    deg = np.diff(indptr.astype(np.int64))
    node = np.full(M, 1, dtype=np.int64)
    live = np.ones(M, dtype=bool)
    for i in range(1, 11):
        live &= deg[node] > 0
        pick = indptr[node].astype(np.int64) + (rng.random(M) * np.maximum(deg[node], 1)).astype(np.int64)
        node = np.where(live, indices[np.minimum(pick, len(indices) - 1)].astype(np.int64), node)
        stop = (rng.random(M) < restart) & (node != 1) if i > 1 else np.zeros(M, dtype=bool)
        live &= ~stop
    back = node == 1
    node[back] = indices[int(indptr[1]) + rng.integers(0, deg[1], back.sum())]
    sim = np.bincount(node, minlength=n).astype(np.float64)
    expect = sim / sim.sum() * len(ends)
    got = np.bincount(ends, minlength=n).astype(np.float64)
    keep = expect > 0
    chi2 = float((((got[keep] - expect[keep]) ** 2) / expect[keep]).sum())
    print(f"[losses] ppr endpoints {got.tolist()} expected {np.round(expect).tolist()} chi2 {chi2:.1f}")
    assert chi2 < 25  # 3 degrees of freedom
    ctx.close()
