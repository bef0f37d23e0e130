"""GPU checks of the learning-rate schedule and the validation forward of ep_learn.

* LR: LRScheduler::cos_decay as built at mod.rs:132-137 (scheduler.rs:37-51): warm-up
  steps = floor(T/5), floor alpha/100, cosine decay after the warm-up; the step counter
  starts at 1 and advances once per reference batch (mod.rs:234).  ep_learn's progress
  callback reports the learning rate of the last batch of each pass, read from the
  device-resident step state; with one batch per pass that is every step 1..T, so the
  warm-up end and the step after it (the schedule's discontinuity) are both checked
  against oracle.cos_decay_lr.
* Validation (mod.rs:267-283): every pass ends with a forward-only pass over the
  validation ids, negatives drawn from the validation ids with the seed minus one, and
  reports sum(loss) / |valid| without the loss weighting.  The test draws the same
  neighbour and negative lists through a second context (seed - 1, pool = the valid ids,
  pass 0) and recomputes the losses with the FP64 oracle on the trained table.
"""
import numpy as np
import pytest

from cloverleaf_amd import _native as N
from oracle import ep_oracle as O
from tests.problems import make_problem, oracle_step

pytestmark = pytest.mark.gpu


def make_cfg(D, B, S, K, passes, alpha, concurrent=1, lw=N.EP_LW_NONE, seed=20222022):
    cfg = N.default_config()
    cfg.dims, cfg.batch_size, cfg.negatives, cfg.gamma, cfg.alpha = D, B, K, 0.5, alpha
    cfg.max_nodes_kind, cfg.max_nodes = N.EP_SAMPLE_FIXED, S
    cfg.passes = passes
    cfg.concurrent_batches = concurrent
    cfg.loss_weighting = lw
    cfg.seed = seed
    return cfg


@pytest.mark.parametrize("passes,alpha", [(25, 0.05), (11, 0.9)])
def test_device_lr_matches_cos_decay(passes, alpha):
    prob = make_problem(seed=70, n=120, V=150, D=32)
    B = 64
    train = np.arange(B, dtype=np.uint32)  # one batch per pass: T = passes
    cfg = make_cfg(32, B, 5, 3, passes, alpha)
    hist = []
    with N.Context(cfg) as ctx:
        ctx.set_graph(prob.indptr, prob.indices)
        ctx.set_features(prob.f_indptr, prob.f_ids, prob.V)
        ctx.set_table(prob.table)
        ctx.set_split(train, np.zeros(0, dtype=np.uint32))
        ctx.learn(lambda p, step, tr, va, lr: hist.append((p, step, lr)))
    T = passes
    warm = int(T / 5.0)
    assert [h[0] for h in hist] == list(range(1, passes + 1))
    for p, step, lr in hist:
        assert step == p + 1  # the counter after the pass's only batch
        ref = O.cos_decay_lr(p, alpha, T)
        assert abs(lr - ref) <= 2e-6 * abs(ref) + 1e-9, (p, lr, ref)
    # the steps the schedule turns on
    lrs = {p: lr for p, _, lr in hist}
    assert abs(lrs[1] - O.cos_decay_lr(1, alpha, T)) <= 2e-6 * alpha
    assert abs(lrs[warm] - alpha) <= 2e-6 * alpha  # warm-up end reaches alpha
    assert lrs[warm + 1] < lrs[warm]  # cosine branch from the next step on
    assert abs(lrs[T] - alpha / 100.0) <= 2e-6 * alpha  # floor at T


def test_device_lr_concurrent_batches():
    # two batches per step (P = 2): the group's batches take consecutive step numbers
    prob = make_problem(seed=71, n=200, V=150, D=32)
    B, P, passes, alpha = 50, 2, 10, 0.2
    train = np.arange(2 * B, dtype=np.uint32)  # 2 reference batches per pass: T = 20
    cfg = make_cfg(32, B, 5, 3, passes, alpha, concurrent=P)
    hist = []
    with N.Context(cfg) as ctx:
        ctx.set_graph(prob.indptr, prob.indices)
        ctx.set_features(prob.f_indptr, prob.f_ids, prob.V)
        ctx.set_table(prob.table)
        ctx.set_split(train, np.zeros(0, dtype=np.uint32))
        ctx.learn(lambda p, step, tr, va, lr: hist.append((p, step, lr)))
    T = 2 * passes
    for p, step, lr in hist:
        assert step == 2 * p + 1
        ref = O.cos_decay_lr(2 * p, alpha, T)  # the pass's last batch
        assert abs(lr - ref) <= 2e-6 * abs(ref) + 1e-9, (p, lr, ref)


@pytest.mark.parametrize("lw", [N.EP_LW_NONE, N.EP_LW_LOG])
def test_validation_loss_matches_oracle(lw):
    prob = make_problem(seed=72, n=300, V=200, D=32, avg_deg=10)
    S, K, B, seed = 5, 4, 64, 12345
    rng = np.random.default_rng(0)
    ids = rng.permutation(prob.n).astype(np.uint32)
    valid, train = ids[:70], ids[70:]
    cfg = make_cfg(32, B, S, K, 2, 0.05, lw=lw, seed=seed)
    hist = []
    with N.Context(cfg) as ctx:
        ctx.set_graph(prob.indptr, prob.indices)
        ctx.set_features(prob.f_indptr, prob.f_ids, prob.V)
        ctx.set_table(prob.table)
        ctx.set_split(train, valid)
        ctx.learn(lambda p, step, tr, va, lr: hist.append((p, tr, va)))
        trained = ctx.get_table()
    # the validation lists: seed - 1, negatives from the valid ids, pass 0, positions 0..
    cfg2 = make_cfg(32, B, S, K, 1, 0.05, lw=lw, seed=seed - 1)
    with N.Context(cfg2) as c2:
        c2.set_graph(prob.indptr, prob.indices)
        c2.set_features(prob.f_indptr, prob.f_ids, prob.V)
        c2.set_table(prob.table)
        c2.set_split(valid, np.zeros(0, dtype=np.uint32))
        cnt, nbr, neg = c2.sample_batch(valid, 0)
    assert np.all(np.isin(neg, valid))
    from tests.problems import Problem
    pt = Problem(prob.indptr, prob.indices, prob.f_indptr, prob.f_ids, prob.V, prob.D, trained)
    res, *_ = oracle_step(pt, valid, cnt, nbr, neg, 0.5, 0.0, 1, "first_leaf")
    # unweighted forward loss of every validation anchor (loss weighting applies to training only)
    losses = np.array([r.loss if r is not None else 0.0 for r in res])
    expect = losses.sum() / len(valid)
    got = hist[-1][2]
    print(f"[schedule] validation loss {got:.6f} vs oracle {expect:.6f}", flush=True)
    assert expect > 0
    assert abs(got - expect) <= 1e-5 * expect
