"""Full-size training runs of the 10M-node and 50M-node workloads on the GPU.

The graphs, feature stores, vocabularies (5M and 20M rows of d = 256), batch sizes,
neighbour samples and negative counts are the full workloads' (200M directed / 1B
undirected edges).  Training runs three passes over a train split of
40 batches per pass (the full pass is 1221 / 3052 batches), with a held-out validation set,
through ep_learn (mod.rs:92-287): shuffle, concurrent batch groups, hipGraph-replayed
steps, validation forward each pass.  Properties checked: no capacity error, every pass
loss finite, the train error of the last pass below the first, finite validation loss.
"""
import math

import numpy as np
import pytest

from cloverleaf_amd import _native as N

pytestmark = pytest.mark.gpu

FORMS = {
    "fp32-plain": (N.EP_PREC_FP32, 1, 0),
    "fp32-batched": (N.EP_PREC_FP32, 4, 1),
    "bf16-master-batched": (N.EP_PREC_BF16_MASTER, 4, 1),
    "bf16-batched": (N.EP_PREC_BF16, 4, 1),
}


def run_form(w, form, passes=3, batches_per_pass=40, n_valid=16384):
    prec, P, use_graph = FORMS[form]
    cfg = N.default_config()
    cfg.dims, cfg.batch_size, cfg.negatives, cfg.gamma, cfg.alpha = w.dims, w.batch, w.negatives, w.gamma, w.alpha
    cfg.max_nodes_kind, cfg.max_nodes = N.EP_SAMPLE_FIXED, w.max_nodes
    cfg.precision, cfg.concurrent_batches, cfg.use_graph = prec, P, use_graph
    cfg.passes = passes
    rng = np.random.default_rng(11)
    ids = rng.choice(w.n_nodes, size=batches_per_pass * w.batch + n_valid, replace=False).astype(np.uint32)
    hist = []
    with N.Context(cfg) as ctx:
        ctx.set_graph(w.indptr, w.indices)
        ctx.set_features(w.f_indptr, w.f_ids, w.vocab)
        ctx.set_table(None)
        ctx.set_split(ids[n_valid:], ids[:n_valid])
        ctx.learn(lambda p, step, tr, va, lr: hist.append((p, step, tr, va, lr)))
    return hist


@pytest.mark.timeout(900)
@pytest.mark.parametrize("name,forms", [
    ("ep-cfg3-10m", ["fp32-plain", "fp32-batched"]),
    ("ep-cfg4-50m", ["fp32-plain", "bf16-master-batched", "bf16-batched"]),
])
def test_full_size_training(name, forms):
    from cloverleaf_amd import synth

    w = synth.make_workload(name, log=lambda m: print(f"[fullsize] {m}", flush=True))
    print(f"[fullsize] {name}: {w.n_nodes} nodes, {len(w.indices)} CSR entries, vocab {w.vocab}", flush=True)
    for form in forms:
        hist = run_form(w, form)
        print(f"[fullsize] {name} {form}: " + ", ".join(f"pass {p}: train {tr:.4f} valid {va:.4f} lr {lr:.5f}"
                                                        for p, _, tr, va, lr in hist), flush=True)
        assert len(hist) == 3
        assert all(math.isfinite(tr) and math.isfinite(va) for _, _, tr, va, _ in hist)
        assert hist[-1][2] < hist[0][2], hist
        assert hist[0][2] > 0
