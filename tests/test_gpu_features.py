"""GPU parity of the EP step with max_features sampling (feature dropout,
utils.rs:193-216 + model.rs:257-278) through the C ABI.  The kept features of every node
of an example are recomputed on the host from the GPU's selection stream
(tests/feature_rules.py) and given to the FP64 oracle.  Bars as the other step tests:
loss, constructions, merged gradient and Adam m within 1e-5 (normwise relative)."""
import numpy as np
import pytest

from cloverleaf_amd import _native as N
from tests import feature_rules as R
from tests.problems import dense_grad, make_problem, oracle_step

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30))


def dropout_ctx(prob, kind, k, p, S, K, B, grad_mode=N.EP_GRAD_FIRST_LEAF, seed=20222022, wavg_cdf=None,
                loss=None):
    cfg = N.default_config()
    cfg.dims, cfg.batch_size, cfg.negatives, cfg.gamma = prob.D, B, K, 0.5
    cfg.max_nodes_kind, cfg.max_nodes = N.EP_SAMPLE_FIXED, S
    cfg.max_features_kind, cfg.max_features, cfg.max_features_p = kind, k, p
    cfg.grad_mode = grad_mode
    cfg.valid_pct = 0.0
    cfg.seed = seed
    if wavg_cdf is not None:
        cfg.weighted_neighbor_averaging = 1
    if loss is not None:
        cfg.loss_kind = N.LOSS_KINDS[loss[0]]
        cfg.gamma = loss[1][0]
        cfg.loss_param2 = loss[1][1] if len(loss[1]) > 1 else 0.0
    ctx = N.Context(cfg)
    ctx.set_graph(prob.indptr, prob.indices)
    if wavg_cdf is not None:
        ctx.set_edge_weights(wavg_cdf)
    ctx.set_features(prob.f_indptr, prob.f_ids, prob.V)
    ctx.set_table(prob.table)
    return ctx, cfg


def check_dropout_step(prob, ctx, cfg, anchors, mode, pass_=1, weights=None, loss=None, tol=1e-5):
    cnt, nbr, neg = ctx.sample_batch(anchors, pass_)
    if weights is not None:
        weights = ctx.sample_weights(len(anchors))
        ctx.set_explicit_weights(weights)
    out = ctx.step_explicit(anchors, cnt, nbr, neg, pass_=pass_, lr=0.01, apply_update=True, want_cons=True)
    sels = [R.example_selections(cfg.max_features_kind, cfg.max_features, cfg.max_features_p, prob.f_indptr,
                                 cfg.seed, i, pass_, int(anchors[i]), nbr[i, :cnt[i]], neg[i])
            for i in range(len(anchors))]
    dropped = sum(len(pos) < int(prob.f_indptr[u + 1] - prob.f_indptr[u])
                  for i in range(len(anchors))
                  for (pos, _), u in zip(sels[i], [anchors[i]] + list(nbr[i, :cnt[i]]) + list(neg[i])))
    if weights is None:
        # the max_nodes scale deg / min(S, deg) (model.rs:355-360); PPR endpoints weigh 1
        # (loss.rs:171-177).  With unequal feature scales it no longer cancels.
        deg = np.diff(prob.indptr.astype(np.int64))[anchors]
        sc = np.float32(1.0) if loss is not None and loss[0] == "ppr" else None
        weights = np.array([[sc if sc is not None else np.float32(d) / np.float32(min(cfg.max_nodes, d))] *
                            cnt.max(initial=1) for d in deg], dtype=np.float32)
    res, merged, t64, m64, v64 = oracle_step(prob, anchors, cnt, nbr, neg, cfg.gamma, 0.01, pass_, mode,
                                             nbr_weights=weights.astype(np.float64), feat_sels=sels, loss=loss)
    ref_loss = np.array([r.loss if (r is not None and r.contributes) else 0.0 for r in res])
    print(f"[features] kind={cfg.max_features_kind} k={cfg.max_features} p={cfg.max_features_p:.2f} {mode}: "
          f"{dropped} nodes subsampled, {np.count_nonzero(ref_loss)} examples", flush=True)
    assert dropped > 20 and np.count_nonzero(ref_loss) > 5
    assert rel(out["loss"], ref_loss) <= tol
    for i, r in enumerate(res):
        if r is not None:
            assert rel(out["cons"][i], np.stack(r.vectors)) <= tol
    G64 = dense_grad(merged, prob.V, prob.D)
    Gg = np.zeros_like(G64)
    assert set(out["feat"].tolist()) == set(merged.keys())
    Gg[out["feat"]] = out["grad"]
    assert rel(Gg, G64) <= tol
    m, v = ctx.get_adam()
    rows = sorted(merged)
    assert rel(m[rows], m64[rows]) <= tol


@pytest.mark.parametrize("kind,k,p", [(R.FIXED, 3, 0.0), (R.FIXED, 1, 0.0), (R.PROBABILITY, 0, 0.4),
                                      (R.PROBABILITY, 0, 0.1)])
@pytest.mark.parametrize("mode", ["first_leaf", "exact"])
def test_feature_sampling_step_matches_oracle(kind, k, p, mode):
    prob = make_problem(seed=30 + k + int(10 * p), n=150, V=300, D=32, fmax=8)
    ctx, cfg = dropout_ctx(prob, kind, k, p, S=6, K=5, B=48,
                           grad_mode=N.EP_GRAD_EXACT if mode == "exact" else N.EP_GRAD_FIRST_LEAF)
    anchors = np.random.default_rng(k + 3).choice(prob.n, 48, replace=False).astype(np.uint32)
    check_dropout_step(prob, ctx, cfg, anchors, mode)
    ctx.close()


def test_feature_sampling_with_weighted_averaging():
    from tests.test_gpu_sampling import weighted_problem
    prob, cdf = weighted_problem(21)
    ctx, cfg = dropout_ctx(prob, R.PROBABILITY, 0, 0.5, S=6, K=5, B=48, wavg_cdf=cdf)
    anchors = np.concatenate([[0, 1, 2], np.random.default_rng(5).choice(np.arange(3, prob.n), 45, replace=False)])
    check_dropout_step(prob, ctx, cfg, anchors.astype(np.uint32), "first_leaf", weights=True)
    ctx.close()


def test_feature_sampling_with_ppr_and_starspace():
    prob = make_problem(seed=44, n=150, V=300, D=32, fmax=8)
    for loss in (("ppr", (0.5, 0.3)), ("starspace", (0.4,))):
        ctx, cfg = dropout_ctx(prob, R.FIXED, 2, 0.0, S=6, K=5, B=48, loss=loss)
        anchors = np.random.default_rng(9).choice(prob.n, 48, replace=False).astype(np.uint32)
        check_dropout_step(prob, ctx, cfg, anchors, "first_leaf", loss=loss)
        ctx.close()


def test_feature_sampling_large_examples_unstaged():
    # hub anchors with 300 sampled neighbours of up to 64 features: more occurrences than
    # the forward's LDS stage holds, so the fallback gathers with masks
    prob = make_problem(seed=51, n=400, V=500, D=64, fmax=64)
    rng = np.random.default_rng(51)
    rows = [set(prob.indices[int(prob.indptr[i]):int(prob.indptr[i + 1])].tolist()) for i in range(prob.n)]
    for h in range(4):
        rows[h] |= set(rng.choice(np.arange(4, prob.n), size=320, replace=False).tolist())
    prob.indptr = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.uint64)
    prob.indices = np.array([j for r in rows for j in sorted(r)], dtype=np.uint32)
    ctx, cfg = dropout_ctx(prob, R.FIXED, 40, 0.0, S=300, K=3, B=16, grad_mode=N.EP_GRAD_EXACT)
    anchors = np.concatenate([[0, 1, 2, 3], rng.choice(np.arange(4, prob.n), 12, replace=False)]).astype(np.uint32)
    check_dropout_step(prob, ctx, cfg, anchors, "exact")
    ctx.close()


def test_feature_sampling_learns_through_facade():
    import cloverleaf_amd as C
    from cloverleaf_amd.synth import make_workload
    w = make_workload("ep-cfg1-10k-comm")
    g = C.Graph.from_csr(w.indptr, w.indices)
    fs = C.FeatureSet.from_csr(g, w.f_indptr, w.f_ids, w.vocab)
    for mf in (5, 0.5):
        ep = C.EmbeddingPropagator(alpha=0.05, loss=C.EPLoss.margin(0.5, 5), batch_size=256, dims=32, passes=3,
                                   max_nodes=10, max_features=mf, indicator=False)
        emb = ep.learn_features(g, fs)
        assert np.all(np.isfinite(emb.embeddings))
        assert ep.history[-1]["train"] < ep.history[0]["train"]
