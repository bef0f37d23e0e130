"""Holdout link-prediction quality of learned feature embeddings.

Protocol: hold out a fraction of the undirected edges (both directions removed, both
endpoints keep degree >= 1), train on the rest, build node embeddings with the Averaged
aggregator (aggregator.rs:34-47: mean of the node's feature rows), score each held-out pair
and an equal number of uniformly random node pairs by cosine similarity, and report the
ROC-AUC (Mann-Whitney, ties averaged).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class Holdout:
    indptr: np.ndarray      # training graph (held-out edges removed)
    indices: np.ndarray
    pos: np.ndarray         # (m, 2) held-out node pairs
    neg: np.ndarray         # (m, 2) random node pairs


def holdout_edges(indptr: np.ndarray, indices: np.ndarray, frac: float = 0.01, seed: int = 7) -> Holdout:
    """Removes `frac` of the undirected edges u<v of a symmetric CSR, keeping degrees >= 1."""
    rng = np.random.default_rng(seed)
    ip = indptr.astype(np.int64)
    n = len(ip) - 1
    src = np.repeat(np.arange(n, dtype=np.int64), np.diff(ip))
    dst = indices.astype(np.int64)
    und = np.flatnonzero(src < dst)
    want = max(1, int(len(und) * frac))
    deg = np.diff(ip).copy()
    chosen = []
    for e in rng.permutation(und):
        u, v = src[e], dst[e]
        if deg[u] > 1 and deg[v] > 1:
            deg[u] -= 1
            deg[v] -= 1
            chosen.append((u, v))
            if len(chosen) == want:
                break
    pos = np.array(chosen, dtype=np.int64)
    drop = set(map(tuple, pos.tolist())) | set(map(tuple, pos[:, ::-1].tolist()))
    keep = np.fromiter(((s, t) not in drop for s, t in zip(src.tolist(), dst.tolist())), dtype=bool, count=len(src))
    src_k, dst_k = src[keep], dst[keep]
    new_ip = np.zeros(n + 1, dtype=np.uint64)
    np.add.at(new_ip, src_k + 1, 1)
    new_ip = np.cumsum(new_ip).astype(np.uint64)
    neg = rng.integers(0, n, size=(len(pos), 2))
    return Holdout(new_ip, dst_k.astype(np.uint32), pos, neg)


def cosine_scores(emb: np.ndarray, pairs: np.ndarray) -> np.ndarray:
    a = emb[pairs[:, 0]].astype(np.float64)
    b = emb[pairs[:, 1]].astype(np.float64)
    na = np.linalg.norm(a, axis=1)
    nb = np.linalg.norm(b, axis=1)
    den = np.where((na > 0) & (nb > 0), na * nb, 1.0)
    return (a * b).sum(axis=1) / den


def roc_auc(pos_scores: np.ndarray, neg_scores: np.ndarray) -> float:
    """Mann-Whitney U / (n_pos n_neg) with average ranks for ties."""
    from scipy.stats import rankdata

    allv = np.concatenate([pos_scores, neg_scores])
    r = rankdata(allv, method="average")
    npos, nneg = len(pos_scores), len(neg_scores)
    return float((r[:npos].sum() - npos * (npos + 1) / 2.0) / (npos * nneg))


def node_embeddings_averaged(table: np.ndarray, f_indptr: np.ndarray, f_ids: np.ndarray) -> np.ndarray:
    """Averaged aggregate on the host (numpy); the GPU version is ep_embed_averaged."""
    ip = f_indptr.astype(np.int64)
    counts = np.diff(ip)
    sums = np.add.reduceat(table[f_ids.astype(np.int64)].astype(np.float64), ip[:-1], axis=0) \
        if len(f_ids) else np.zeros((len(counts), table.shape[1]))
    sums[counts == 0] = 0.0
    return sums / np.maximum(counts, 1)[:, None]


def holdout_auc(node_emb: np.ndarray, h: Holdout) -> float:
    return roc_auc(cosine_scores(node_emb, h.pos), cosine_scores(node_emb, h.neg))


def train_gpu_table(graph_indptr, graph_indices, f_indptr, f_ids, vocab, dims, batch, negatives, max_nodes, gamma,
                    alpha, passes, seed, precision="fp32", use_graph=False, scatter="sort", concurrent=1):
    """Trains with the GPU path (ep_learn) and returns the FP32 feature table."""
    from . import _native as N

    cfg = N.default_config()
    cfg.dims, cfg.batch_size, cfg.negatives, cfg.gamma, cfg.alpha = dims, batch, negatives, gamma, alpha
    cfg.passes, cfg.valid_pct, cfg.seed = passes, 0.0, seed
    cfg.max_nodes_kind, cfg.max_nodes = N.EP_SAMPLE_FIXED, max_nodes
    cfg.precision = {"fp32": N.EP_PREC_FP32, "bf16": N.EP_PREC_BF16, "bf16-master": N.EP_PREC_BF16_MASTER}[precision]
    cfg.use_graph = 1 if use_graph else 0
    cfg.scatter = N.EP_SCATTER_ATOMIC if scatter == "atomic" else N.EP_SCATTER_SORT
    cfg.concurrent_batches = concurrent
    with N.Context(cfg) as ctx:
        ctx.set_graph(graph_indptr, graph_indices)
        ctx.set_features(f_indptr, f_ids, vocab)
        ctx.set_table(None)
        ctx.learn()
        return ctx.get_table()
