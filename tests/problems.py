"""Small seeded EP problems shared by the parity tests and the golden-fixture script."""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from oracle import ep_oracle as O


@dataclass
class Problem:
    indptr: np.ndarray   # u64
    indices: np.ndarray  # u32
    f_indptr: np.ndarray
    f_ids: np.ndarray
    V: int
    D: int
    table: np.ndarray    # f32 V x D

    @property
    def n(self):
        return len(self.indptr) - 1


def make_problem(seed=0, n=120, V=300, D=32, avg_deg=6, fmax=8, zipf=1.1, zero_deg=0, table_scale=1.0):
    """Random undirected graph (dedup, sorted columns) + Zipf features + uniform table.

    ``zero_deg`` nodes at the end of the id range get no edges (directed dead ends)."""
    rng = np.random.default_rng(seed)
    live = n - zero_deg
    edges = set()
    for i in range(live):
        for j in rng.integers(0, live, size=avg_deg // 2 + 1):
            j = int(j)
            if i != j:
                edges.add((i, j))
                edges.add((j, i))
    rows, cols, _ = O.csr_from_edges([(a, b, 1.0) for a, b in sorted(edges)], deduplicate=True)
    indptr = np.zeros(n + 1, dtype=np.uint64)
    indptr[:len(rows)] = rows
    indptr[len(rows):] = rows[-1]
    counts = rng.integers(1, fmax + 1, size=n)
    f_indptr = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint64)
    p = 1.0 / np.arange(1, V + 1) ** zipf
    f_ids = rng.choice(V, size=int(f_indptr[-1]), p=p / p.sum()).astype(np.uint32)
    table = (rng.uniform(-1, 1, size=(V, D)) * table_scale).astype(np.float32)
    return Problem(indptr, cols.astype(np.uint32), f_indptr, f_ids, V, D, table)


def oracle_step(prob: Problem, anchors, cnt, nbr, neg, gamma, lr, pass_, grad_mode, table=None, mom=None, var=None,
                loss_divs=None, attention=None, loss=None, nbr_weights=None, feat_sels=None):
    """FP64 oracle of one explicit-sample step; returns (results, merged, table, mom, var).
    attention = (heads, d_k[, window[, random blocks]]) selects the attention model (a list:
    one tuple per anchor); loss = (kind, params) the loss;
    nbr_weights[i][j] is neighbour j's weight (weighted averaging; default uniform);
    feat_sels[i] the max_features selections of example i (see example_grads)."""
    t = (prob.table if table is None else table).astype(np.float64).copy()
    m = np.zeros_like(t) if mom is None else mom.astype(np.float64).copy()
    v = np.zeros_like(t) if var is None else var.astype(np.float64).copy()
    B = len(anchors)
    nbr_lists = [list(map(int, nbr[i, :cnt[i]])) for i in range(B)]
    results, merged = [], {}
    # anchors with no neighbourhood are skipped (no example)
    keep = [i for i in range(B) if cnt[i] > 0]
    if isinstance(attention, list):
        attention = [attention[i] for i in keep]
    res_k, merged = O.ep_batch_step(t, m, v, prob.f_indptr, prob.f_ids, [int(anchors[i]) for i in keep],
                                    [nbr_lists[i] for i in keep],
                                    [1.0] * len(keep) if nbr_weights is None
                                    else [list(map(float, nbr_weights[i][:cnt[i]])) for i in keep],
                                    [list(map(int, neg[i])) for i in keep], gamma, lr, pass_, grad_mode,
                                    None if loss_divs is None else [loss_divs[i] for i in keep], attention, loss,
                                    None if feat_sels is None else [feat_sels[i] for i in keep])
    it = iter(res_k)
    results = [next(it) if i in keep else None for i in range(B)]
    return results, merged, t, m, v


def oracle_group_step(prob: Problem, anchors, cnt, nbr, neg, gamma, lr, pass_, grad_mode, batch_size,
                      attention=None):
    """FP64 oracle of one explicit-sample step of consecutive batches of `batch_size`
    anchors run as one concurrent group (every batch reads the group-start table, updates
    in batch order, all with learning rate lr).  Returns (per-batch merged, table, mom, var)."""
    t = prob.table.astype(np.float64).copy()
    m = np.zeros_like(t)
    v = np.zeros_like(t)
    B = len(anchors)
    keep = [i for i in range(B) if cnt[i] > 0]  # anchors with no neighbourhood give no example
    nb = (B + batch_size - 1) // batch_size
    _, merged = O.ep_group_step(t, m, v, prob.f_indptr, prob.f_ids, [int(anchors[i]) for i in keep],
                                [list(map(int, nbr[i, :cnt[i]])) for i in keep], [1.0] * len(keep),
                                [list(map(int, neg[i])) for i in keep], gamma, [lr] * nb,
                                [i // batch_size for i in keep], pass_, grad_mode, None, attention)
    return merged, t, m, v


def dense_grad(merged, V, D):
    G = np.zeros((V, D))
    for f, g in merged.items():
        G[f] = g
    return G
