"""End-to-end quality: holdout link-prediction AUC of the GPU path matches the CPU
restatement of the reference loop (oracle/ep_cpu.cpp) within 0.005, on the
community-structured 10k-node workload (mean over 8 training seeds each)."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from cloverleaf_amd import quality as Q
from cloverleaf_amd import synth

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SEEDS = [20222022 + i for i in range(8)]


def cpu_tables(w, h):
    exe = os.path.join(ROOT, "oracle", "_build", "ep_cpu")
    assert os.path.exists(exe), "build the CPU restatement first (__graft_entry__.build())"
    with tempfile.TemporaryDirectory() as td:
        h.indptr.astype(np.uint64).tofile(os.path.join(td, "indptr.bin"))
        h.indices.astype(np.uint32).tofile(os.path.join(td, "indices.bin"))
        w.f_indptr.astype(np.uint64).tofile(os.path.join(td, "findptr.bin"))
        w.f_ids.astype(np.uint32).tofile(os.path.join(td, "fids.bin"))
        for seed in SEEDS:
            out = os.path.join(td, "t.bin")
            subprocess.run([exe, "--dir", td, "--nodes", str(w.n_nodes), "--vocab", str(w.vocab), "--dims",
                            str(w.dims), "--batch", str(w.batch), "--negatives", str(w.negatives), "--max-nodes",
                            str(w.max_nodes), "--gamma", str(w.gamma), "--alpha", str(w.alpha), "--threads", "8",
                            "--passes", str(w.passes), "--seed", str(seed), "--table-out", out, "--json"],
                           check=True, capture_output=True, timeout=300)
            yield np.fromfile(out, dtype=np.float32).reshape(w.vocab, w.dims)


@pytest.fixture(scope="module")
def setup():
    w = synth.make_workload("ep-cfg1-10k-comm")
    h = Q.holdout_edges(w.indptr, w.indices, 0.05, seed=7)

    def auc(tab):
        return Q.holdout_auc(Q.node_embeddings_averaged(tab, w.f_indptr, w.f_ids), h)

    cpu = [auc(t) for t in cpu_tables(w, h)]
    return w, h, auc, cpu


# (precision, hipGraph replay, reference batches per step)
FORMS = [("fp32", False, 1), ("bf16", True, 1), ("bf16-master", True, 1), ("fp32", True, 4), ("bf16", True, 4),
         ("bf16", True, 8)]


@pytest.mark.parametrize("precision,use_graph,concurrent", FORMS)
def test_holdout_auc_matches_cpu_restatement(setup, precision, use_graph, concurrent):
    w, h, auc, cpu = setup
    print(f"[quality] {precision} graph={use_graph} P={concurrent} holdout pairs {len(h.pos)}", flush=True)
    gpu = [auc(Q.train_gpu_table(h.indptr, h.indices, w.f_indptr, w.f_ids, w.vocab, w.dims, w.batch, w.negatives,
                                 w.max_nodes, w.gamma, w.alpha, w.passes, s, precision, use_graph,
                                 concurrent=concurrent)) for s in SEEDS]
    print(f"[quality] gpu {np.mean(gpu):.4f} +- {np.std(gpu):.4f}  cpu {np.mean(cpu):.4f} +- {np.std(cpu):.4f}")
    assert np.mean(cpu) > 0.8  # the workload is learnable
    assert abs(np.mean(gpu) - np.mean(cpu)) <= 0.005
