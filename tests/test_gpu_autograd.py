"""The hand-written backward of the HIP step wrapped in a torch.autograd.Function (test
harness only) and checked by central finite differences of the HIP forward.

The Function's forward runs ep_step_explicit (no update) on the given table and returns
the batch loss sum_e L_e (the contributing examples' losses, loss.rs:59-73, as the
reference's per-example tapes, mod.rs:183-214); its backward returns the merged gradient
the kernels produced (exact gradient mode: every construction's contribution, the true
derivative of the batch loss).  The checks:

* autograd plumbing: loss.backward() leaves exactly the kernels' merged gradient in
  table.grad;
* directional central differences: for directions u mixing the gradient's direction
  with a random one over the touched rows, (L(t + h u) - L(t - h u)) / 2h against the
  gradient's prediction, with the forward evaluated by the HIP kernels in FP32
  (per-example losses summed in float64);
* coordinate central differences on the largest gradient entries.

The margin gamma = 10 keeps every negative active, so the batch loss is smooth (no hinge
between the evaluation points) and every construction's adjoint is exercised; the
three-point estimates at h and h/2 must agree.  Tolerance: FP32 losses of about 10 per
example carry about 1e-6 of rounding each, so with h = 0.02 the difference quotient has
about 1e-4 of noise against a derivative of O(|grad|); the bar is 2e-3 relative
(directions) and 1e-2 (coordinates).  Mean model and attention model
(attention.rs:100-296).
"""
import numpy as np
import pytest
import torch

from cloverleaf_amd import _native as N
from tests.problems import make_problem

pytestmark = pytest.mark.gpu


class HipStep(torch.autograd.Function):
    """sum_e L_e of one explicit-sample batch, forward and backward on the GPU kernels."""

    @staticmethod
    def forward(fctx, table, ep, lists):
        anchors, cnt, nbr, neg = lists
        ep.set_table(table.detach().cpu().numpy().astype(np.float32))
        out = ep.step_explicit(anchors, cnt, nbr, neg, 1, 0.0, apply_update=False)
        G = np.zeros(tuple(table.shape), dtype=np.float64)
        G[out["feat"]] = out["grad"]
        fctx.save_for_backward(torch.from_numpy(G))
        return torch.tensor(float(out["loss"].astype(np.float64).sum()), dtype=torch.float64)

    @staticmethod
    def backward(fctx, gout):
        (G,) = fctx.saved_tensors
        return gout * G, None, None


def hip_loss(table, ep, lists):
    return float(HipStep.apply(torch.from_numpy(table), ep, lists))


GAMMA = 10.0


def make_ep(prob, S, K, B, attention=None):
    cfg = N.default_config()
    cfg.dims, cfg.batch_size, cfg.negatives, cfg.gamma = prob.D, B, K, GAMMA
    if attention is not None:
        heads, dk, dm = attention
        cfg.dims, cfg.attention_heads, cfg.attention_dk = dm, heads, dk
    cfg.max_nodes_kind, cfg.max_nodes = N.EP_SAMPLE_FIXED, S
    cfg.grad_mode = N.EP_GRAD_EXACT
    cfg.valid_pct = 0.0
    ep = N.Context(cfg)
    ep.set_graph(prob.indptr, prob.indices)
    ep.set_features(prob.f_indptr, prob.f_ids, prob.V)
    ep.set_table(prob.table)
    return ep


@pytest.mark.parametrize("model", ["mean", "attention"])
def test_autograd_function_matches_central_differences(model):
    if model == "mean":
        prob = make_problem(seed=80, n=120, V=160, D=32, fmax=6)
        att = None
    else:
        heads, dk, dm = 2, 8, 16
        prob = make_problem(seed=81, n=120, V=160, D=heads * (2 * dk + dm), fmax=6, table_scale=0.6)
        att = (heads, dk, dm)
    S, K, B = 5, 4, 16
    ep = make_ep(prob, S, K, B, att)
    rng = np.random.default_rng(3)
    anchors = rng.choice(prob.n, size=B, replace=False).astype(np.uint32)
    lists = (anchors, *ep.sample_batch(anchors, 1))
    table0 = prob.table.astype(np.float64)

    # autograd plumbing
    t = torch.from_numpy(table0.copy()).requires_grad_(True)
    loss = HipStep.apply(t, ep, lists)
    loss.backward()
    G = t.grad.numpy().copy()
    touched = np.flatnonzero(np.any(G != 0, axis=1))
    assert float(loss) > 0 and len(touched) > 20
    ep.set_table(prob.table)
    out = ep.step_explicit(*lists, 1, 0.0, apply_update=False)
    G2 = np.zeros_like(G)
    G2[out["feat"]] = out["grad"]
    assert np.array_equal(G, G2)

    def central(u, h):
        """Central differences at h and h/2 over the FP32 tables the kernels read, and the
        gradient's prediction <G, t+ - t-> / 2h at h."""
        est, pred = [], None
        for hh in (h, h / 2):
            tp = (table0 + hh * u).astype(np.float32).astype(np.float64)
            tm = (table0 - hh * u).astype(np.float32).astype(np.float64)
            est.append((hip_loss(tp, ep, lists) - hip_loss(tm, ep, lists)) / (2 * hh))
            if pred is None:
                pred = float((G * (tp - tm)).sum()) / (2 * hh)
        return est[0], est[1], pred

    # directional central differences: the gradient's direction and mixes with random ones
    gdir = G / np.linalg.norm(G)
    errs = []
    for k in range(7):
        u = gdir.copy()
        if k:
            xi = np.zeros_like(table0)
            xi[touched] = rng.standard_normal((len(touched), prob.D))
            u += xi / np.linalg.norm(xi)
            u /= np.linalg.norm(u)
        fd, fd2, g = central(u, 0.02)
        assert abs(fd - fd2) <= 2e-3 * abs(fd), (k, fd, fd2)  # smooth between the points
        errs.append(abs(fd - g) / abs(g))
    print(f"[autograd] {model}: directional max rel err {max(errs):.2e} (|grad| {np.linalg.norm(G):.3f})",
          flush=True)
    assert max(errs) <= 2e-3

    # coordinate differences on the largest entries
    cerrs = []
    for idx in np.argsort(-np.abs(G).ravel())[:8]:
        r, col = divmod(int(idx), prob.D)
        e = np.zeros_like(table0)
        e[r, col] = 1.0
        fd, _, g = central(e, 0.02)
        cerrs.append(abs(fd - g) / abs(g))
    print(f"[autograd] {model}: coordinate max rel err {max(cerrs):.2e}", flush=True)
    assert max(cerrs) <= 1e-2
    ep.close()
