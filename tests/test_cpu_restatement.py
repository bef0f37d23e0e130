"""The C++ CPU restatement of the reference loop (oracle/ep_cpu.cpp: the timed CPU
baseline and the CPU side of the AUC comparison) against the FP64 oracle.

ep_cpu --explicit runs one batch sequentially on given neighbour and negative lists:
every example's first-leaf gradients (mod.rs:329-378) merged in anchor order
(mod.rs:223-233), then one sparse Adam update (optimizer.rs:43-81).  The same lists go
through oracle.ep_batch_step.  Bars (FP32 restatement against FP64): per-example loss
and merged gradient 1e-5 normwise relative, Adam m 1e-5, v 1e-4, theta 1e-5 where the
gradient is above FP32 noise; mean model and attention model.
"""
import os
import subprocess

import numpy as np
import pytest

from oracle import ep_oracle as O
from tests.problems import make_problem

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "oracle", "_build", "ep_cpu")


@pytest.fixture(scope="module")
def ep_cpu():
    if not os.path.exists(EXE):
        from oracle import build

        build.build()
    return EXE


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30))


def sample_lists(prob, anchors, S, K, rng):
    ip = prob.indptr.astype(np.int64)
    n = len(anchors)
    cnt = np.zeros(n, np.uint32)
    nbr = np.zeros((n, S), np.uint32)
    for i, v in enumerate(anchors):
        row = prob.indices[ip[v]:ip[v + 1]]
        k = min(len(row), S)
        cnt[i] = k
        nbr[i, :k] = row if len(row) <= S else rng.choice(row, size=S, replace=False)
    neg = rng.integers(0, prob.n, size=(n, K)).astype(np.uint32)
    return cnt, nbr, neg


@pytest.mark.parametrize("model", ["mean", "attention"])
@pytest.mark.parametrize("seed", [0, 1])
def test_ep_cpu_batch_matches_oracle(ep_cpu, tmp_path, model, seed):
    S, K, B, gamma, lr, pass_ = 6, 4, 40, 0.5, 0.02, 2
    if model == "mean":
        prob = make_problem(seed=90 + seed, n=150, V=120, D=32, fmax=8, zipf=1.2)
        att, extra = None, []
    else:
        heads, dk, dm = 2, 8, 16
        prob = make_problem(seed=95 + seed, n=150, V=120, D=heads * (2 * dk + dm), fmax=6, zipf=1.2, table_scale=0.6)
        att, extra = (heads, dk), ["--heads", str(heads), "--dk", str(dk)]
    dm = prob.D if att is None else 16
    rng = np.random.default_rng(seed)
    anchors = rng.choice(prob.n, size=B, replace=False).astype(np.uint32)
    cnt, nbr, neg = sample_lists(prob, anchors, S, K, rng)
    d = str(tmp_path)
    for name, arr in (("indptr", prob.indptr.astype(np.uint64)), ("indices", prob.indices.astype(np.uint32)),
                      ("findptr", prob.f_indptr.astype(np.uint64)), ("fids", prob.f_ids.astype(np.uint32)),
                      ("table", prob.table.astype(np.float32)), ("anchors", anchors), ("cnt", cnt), ("nbr", nbr),
                      ("neg", neg)):
        arr.tofile(os.path.join(d, name + ".bin"))
    subprocess.run([ep_cpu, "--dir", d, "--explicit", d, "--nodes", str(prob.n), "--vocab", str(prob.V),
                    "--dims", str(dm), "--negatives", str(K), "--max-nodes", str(S), "--gamma", str(gamma),
                    "--lr", str(lr), "--pass", str(pass_), *extra], check=True)
    W = prob.D
    loss = np.fromfile(os.path.join(d, "out_loss.bin"), dtype=np.float32)
    grad = np.fromfile(os.path.join(d, "out_grad.bin"), dtype=np.float32).reshape(prob.V, W)
    table = np.fromfile(os.path.join(d, "out_table.bin"), dtype=np.float32).reshape(prob.V, W)
    m = np.fromfile(os.path.join(d, "out_m.bin"), dtype=np.float32).reshape(prob.V, W)
    v = np.fromfile(os.path.join(d, "out_v.bin"), dtype=np.float32).reshape(prob.V, W)

    t64 = prob.table.astype(np.float64)
    m64, v64 = np.zeros_like(t64), np.zeros_like(t64)
    keep = [i for i in range(B) if cnt[i] > 0]
    res, merged = O.ep_batch_step(t64, m64, v64, prob.f_indptr, prob.f_ids, [int(anchors[i]) for i in keep],
                                  [list(map(int, nbr[i, :cnt[i]])) for i in keep], [1.0] * len(keep),
                                  [list(map(int, neg[i])) for i in keep], gamma, lr, pass_, "first_leaf",
                                  None, None if att is None else [att] * len(keep))
    ref_loss = np.zeros(B)
    for i, r in zip(keep, res):
        ref_loss[i] = r.loss if r.contributes else 0.0
    assert np.count_nonzero(ref_loss) > 10
    assert rel(loss, ref_loss) <= 1e-5
    G64 = np.zeros((prob.V, W))
    for f, g in merged.items():
        G64[f] = g
    assert set(np.flatnonzero(np.any(grad != 0, axis=1)).tolist()) <= set(merged)
    assert rel(grad, G64) <= 1e-5
    rows = sorted(merged)
    assert rel(m[rows], m64[rows]) <= 1e-5
    assert rel(v[rows], v64[rows]) <= 1e-4
    mask = np.abs(G64[rows]) > 1e-3 * np.abs(G64).max()
    assert np.max(np.abs(table[rows] - t64[rows])[mask]) < 1e-5
    untouched = np.setdiff1d(np.arange(prob.V), rows)
    assert np.array_equal(table[untouched], prob.table[untouched])
