"""Pins the FP64 oracle (oracle/ep_oracle.py) before anything is compared against it.

Every check is independent of the oracle's own formulas: reference known-answer
tests, closed forms worked by hand, central finite differences, brute force with
per-construction autograd leaves (the reference's extraction mechanism), and
library routines.
"""
import math

import numpy as np
import pytest
import torch

from oracle import ep_oracle as O

# --------------------------------------------------------------------------------------
# Graph construction: reference KATs src/graph.rs:696-776
# --------------------------------------------------------------------------------------
KAT_EDGES = [(0, 1, 1.0), (1, 1, 3.0), (1, 2, 2.0), (2, 0, 2.5), (1, 0, 10.0)]  # graph.rs:697-705


def test_csr_kat_insertion_order():
    # graph.rs:707-713 expects rows [0,1,4,5], columns [1,1,2,0,0], weights [1,3,2,10,2.5];
    # those are the insertion-order results (deduplicate = false)
    rows, cols, w = O.csr_from_edges(KAT_EDGES, deduplicate=False)
    assert rows.tolist() == [0, 1, 4, 5]
    assert cols.tolist() == [1, 1, 2, 0, 0]
    assert w.tolist() == [1.0, 3.0, 2.0, 10.0, 2.5]


def test_csr_dedup_sorts_and_merges():
    # with deduplicate=true the edges are sorted by (from, to) first (graph.rs:120-142)
    rows, cols, w = O.csr_from_edges(KAT_EDGES, deduplicate=True)
    assert rows.tolist() == [0, 1, 4, 5]
    assert cols.tolist() == [1, 0, 1, 2, 0]
    assert w.tolist() == [1.0, 10.0, 3.0, 2.0, 2.5]
    rows, cols, w = O.csr_from_edges([(0, 1, 1.0), (0, 1, 2.0), (1, 0, 4.0), (0, 1, 0.5)], deduplicate=True)
    assert rows.tolist() == [0, 1, 2] and cols.tolist() == [1, 0] and w.tolist() == [3.5, 4.0]


def test_cumcsr_cdf_kat():
    # graph.rs:744-752: weights [1, 3/15, 5/15, 15/15, 1]
    rows, cols, w = O.csr_from_edges(KAT_EDGES, deduplicate=False)
    cdf = O.cumcsr_convert(rows, w)
    assert np.allclose(cdf, [1.0, 3 / 15, 5 / 15, 1.0, 1.0], atol=1e-12)


def test_cdf_to_p_kat():
    # graph.rs:755-768
    rows, cols, w = O.csr_from_edges(KAT_EDGES, deduplicate=False)
    cdf = O.cumcsr_convert(rows, w)
    assert np.allclose(O.cdf_to_p(cdf[1:4]), [3 / 15, 2 / 15, 10 / 15], atol=1e-7)


def test_cdf_edge_cases():
    assert O.cdf_convert(np.zeros(0)).size == 0  # dead-end node (graph.rs:668-670)
    assert np.allclose(O.cdf_convert(np.zeros(4)), [0.25, 0.5, 0.75, 1.0])  # zero weights -> uniform
    assert O.cdf_convert(np.array([1.0, 1.0, 1.0]))[-1] == 1.0


# --------------------------------------------------------------------------------------
# Sample, distance, constructions
# --------------------------------------------------------------------------------------
def test_sample_spec():
    assert O.sample_spec("fixed", 10, 25) == (10, 2.5)
    assert O.sample_spec("fixed", 10, 4) == (4, 1.0)
    assert O.sample_spec("all", 0, 7) == (7, 1.0)


def test_euclidean_kat():
    # loss.rs:250-255: distance([1,3],[3,5]) = sqrt(8)
    assert O.euclidean(np.array([1.0, 3.0]), np.array([3.0, 5.0])) == math.sqrt(8.0)


def test_construct_mean_closed_form():
    table = np.array([[1.0, 0.0], [0.0, 2.0], [4.0, 4.0]])
    # node with features [0, 1, 1]: counts 2,4 (weight 1 + scale 1 per occurrence) -> (2*[1,0] + 4*[0,2]) / 6
    f_indptr = np.array([0, 3, 4])
    f_ids = np.array([0, 1, 1, 2])
    c = O.build_anchor_construction(f_indptr, f_ids, 0)
    assert c.counts == {0: 2.0, 1: 4.0}
    assert np.allclose(O.construct_mean(table, c), [1 / 3, 4 / 3])
    # order of features does not matter
    c2 = O.build_anchor_construction(np.array([0, 3]), np.array([1, 0, 1]), 0)
    assert np.allclose(O.construct_mean(table, c2), O.construct_mean(table, c))


def test_neighbour_construction_is_feature_pooled():
    # ~h(v) pools all neighbour features (model.rs:398-414), not a mean of per-node means
    table = np.array([[0.0], [10.0], [1.0]])
    f_indptr = np.array([0, 1, 4])  # node 0: [0]; node 1: [1, 2, 2]
    f_ids = np.array([0, 1, 2, 2])
    c = O.build_neighbour_construction(f_indptr, f_ids, [0, 1], scale_weight=3.0)
    pooled = (0.0 + 10.0 + 1.0 + 1.0) / 4.0
    assert np.allclose(O.construct_mean(table, c), [pooled])
    assert not np.isclose(pooled, (0.0 + (10 + 1 + 1) / 3) / 2)
    # the per-neighbour weight (scale + 1) is uniform and cancels
    c1 = O.build_neighbour_construction(f_indptr, f_ids, [0, 1], scale_weight=1.0)
    assert np.allclose(O.construct_mean(table, c1), O.construct_mean(table, c))


def test_single_feature_complete_graph_gives_zero_distance():
    # every node has the same single feature: h(v) == ~h(v)
    table = np.array([[0.3, -0.4]])
    f_indptr = np.array([0, 1, 2, 3])
    f_ids = np.array([0, 0, 0])
    a = O.construct_mean(table, O.build_anchor_construction(f_indptr, f_ids, 0))
    p = O.construct_mean(table, O.build_neighbour_construction(f_indptr, f_ids, [1, 2], 1.0))
    assert O.euclidean(a, p) == 0.0


# --------------------------------------------------------------------------------------
# Margin loss
# --------------------------------------------------------------------------------------
def test_margin_loss_closed_form():
    a = np.array([0.0, 0.0])
    p = np.array([3.0, 4.0])          # D0 = 5
    n1 = np.array([3.0, 0.0])         # D1 = 4 -> l = 1 + 5 - 4 = 2 (active)
    n2 = np.array([3.0, 14.0])        # D2 = 10 -> l = -4 (inactive)
    L, act, ga, gp, gn = O.margin_loss_adjoints(a, p, [n1, n2], gamma=1.0)
    assert L == 2.0 and act.tolist() == [True, False]
    assert np.allclose(ga, (a - p) / 5.0)
    assert np.allclose(gn[0], (p - n1) / 4.0) and gn[1] is None
    assert np.allclose(gp, (p - a) / 5.0 - (p - n1) / 4.0)
    L0, act0, *_ = O.margin_loss_adjoints(a, p, [n2], gamma=1.0)
    assert L0 == 0.0 and not act0.any()


def test_margin_loss_translation_invariant():
    rng = np.random.default_rng(0)
    a, p, n1, n2 = rng.normal(size=(4, 5))
    t = rng.normal(size=5)
    L1 = O.margin_loss_adjoints(a, p, [n1, n2], 2.0)[0]
    L2 = O.margin_loss_adjoints(a + t, p + t, [n1 + t, n2 + t], 2.0)[0]
    assert abs(L1 - L2) < 1e-12 and L1 >= 0


def test_margin_adjoints_central_fd():
    rng = np.random.default_rng(1)
    d, K = 6, 4
    a, p = rng.normal(size=(2, d))
    ns = list(rng.normal(size=(K, d)))
    gaps = sorted(O.euclidean(p, n) - O.euclidean(p, a) for n in ns)
    gamma = 0.5 * (gaps[K // 2 - 1] + gaps[K // 2])  # half the negatives active, away from the kink
    L, act, ga, gp, gn = O.margin_loss_adjoints(a, p, ns, gamma)
    assert act.any() and not act.all()

    def f(a_, p_, ns_):
        return O.margin_loss_adjoints(a_, p_, ns_, gamma)[0]

    eps = 1e-6
    for vec, g in [("a", ga), ("p", gp)] + [(i, gn[i]) for i in range(K)]:
        for j in range(d):
            def pert(s):
                a_, p_, ns_ = a.copy(), p.copy(), [x.copy() for x in ns]
                if vec == "a":
                    a_[j] += s
                elif vec == "p":
                    p_[j] += s
                else:
                    ns_[vec][j] += s
                return f(a_, p_, ns_)
            fd = (pert(eps) - pert(-eps)) / (2 * eps)
            exp = 0.0 if g is None else g[j]
            assert abs(fd - exp) < 1e-7, (vec, j, fd, exp)


# --------------------------------------------------------------------------------------
# Gradients of one example: exact mode vs FD, first-leaf mode vs per-leaf autograd
# --------------------------------------------------------------------------------------
def _toy_example(seed=3, V=12, D=5, n_nodes=10, K=4):
    rng = np.random.default_rng(seed)
    counts = rng.integers(1, 5, size=n_nodes)
    f_indptr = np.concatenate([[0], np.cumsum(counts)])
    f_ids = rng.integers(0, V, size=f_indptr[-1])  # heavy feature sharing
    table = rng.normal(size=(V, D))
    nbrs = [1, 2, 3]
    negs = list(rng.integers(0, n_nodes, size=K))
    return table, f_indptr, f_ids, 0, nbrs, negs


def _example_loss(table, f_indptr, f_ids, v, nbrs, negs, gamma):
    return O.example_grads(table, f_indptr, f_ids, v, nbrs, 1.0, negs, gamma, "exact").loss


@pytest.mark.parametrize("seed", [3, 4, 5])
def test_exact_grads_central_fd(seed):
    table, f_indptr, f_ids, v, nbrs, negs = _toy_example(seed)
    gamma = 2.0
    r = O.example_grads(table, f_indptr, f_ids, v, nbrs, 1.0, negs, gamma, "exact")
    assert r.contributes
    eps = 1e-6
    for f in range(table.shape[0]):
        for j in range(table.shape[1]):
            tp, tm = table.copy(), table.copy()
            tp[f, j] += eps
            tm[f, j] -= eps
            fd = (_example_loss(tp, f_indptr, f_ids, v, nbrs, negs, gamma)
                  - _example_loss(tm, f_indptr, f_ids, v, nbrs, negs, gamma)) / (2 * eps)
            g = r.grads[f][j] if f in r.grads else 0.0
            assert abs(fd - g) < 1e-6, (f, j, fd, g)


def torch_first_leaf(table, f_indptr, f_ids, v, nbrs, negs, gamma, nbr_weights=None):
    """The reference's mechanism (model.rs:257-278, mod.rs:329-378), in torch fp64:
    one autograd leaf per (construction, feature); keep the first leaf's gradient per
    feature in the order a, p, n_1..K; leaves the loss does not reach have no grad."""
    def spec(nodes, weights):
        out = []
        for u, weight in zip(nodes, weights):
            for f in f_ids[f_indptr[u]:f_indptr[u + 1]]:
                out.append((int(f), weight + 1.0))
        return out

    nw = [1.0] * len(nbrs) if nbr_weights is None else list(nbr_weights)
    specs = [spec([v], [1.0]), spec(nbrs, nw)] + [spec([u], [1.0]) for u in negs]
    leaves, vecs = [], []
    for sp in specs:
        d = {}
        for f, w in sp:
            if f in d:
                d[f][1] += w
            else:
                d[f] = [torch.tensor(table[f], dtype=torch.float64, requires_grad=True), w]
        num = sum(leaf * c for leaf, c in d.values())
        den = sum(c for _, c in d.values())
        vecs.append(num / den)
        leaves.append(d)
    a, p, ns = vecs[0], vecs[1], vecs[2:]
    d1 = gamma + torch.sqrt(((p - a) ** 2).sum())
    ls = [d1 - torch.sqrt(((p - n) ** 2).sum()) for n in ns]
    pos = [l for l in ls if l.item() > 0]
    if not pos:
        return 0.0, {}
    L = sum(pos) / len(pos)
    L.backward()
    grads = {}
    for d in leaves:
        for f, (leaf, _) in d.items():
            if f in grads or leaf.grad is None:
                continue
            g = leaf.grad.numpy()
            if np.all(np.isfinite(g)):
                grads[f] = g.copy()
    return L.item(), grads


@pytest.mark.parametrize("seed", [3, 4, 5, 6, 7])
def test_first_leaf_matches_per_leaf_autograd(seed):
    table, f_indptr, f_ids, v, nbrs, negs = _toy_example(seed)
    gamma = 2.0
    r = O.example_grads(table, f_indptr, f_ids, v, nbrs, 1.0, negs, gamma, "first_leaf")
    L, g = torch_first_leaf(table, f_indptr, f_ids, v, nbrs, negs, gamma)
    assert abs(r.loss - L) < 1e-12
    assert set(r.grads) == set(g)
    for f in g:
        assert np.allclose(r.grads[f], g[f], atol=1e-12)


def test_weighted_neighbour_construction_closed_form():
    # model.rs:357-361 + 264-276: neighbour j's occurrences count w_j + 1
    f_indptr = np.array([0, 2, 3], dtype=np.uint64)
    f_ids = np.array([0, 1, 1], dtype=np.uint32)
    con = O.build_neighbour_construction(f_indptr, f_ids, [0, 1], [3.0, 0.5])
    assert con.counts == {0: 4.0, 1: 5.5}
    table = np.array([[1.0, 0.0], [0.0, 2.0]])
    assert np.allclose(O.construct_mean(table, con), np.array([4.0, 11.0]) / 9.5)
    # one weight for every neighbour: the unweighted path
    assert (O.build_neighbour_construction(f_indptr, f_ids, [0, 1], [2.0, 2.0]).counts
            == O.build_neighbour_construction(f_indptr, f_ids, [0, 1], 2.0).counts)


def test_neighbour_weights_kat():
    f = np.float32
    cdf = np.array([0.2, 0.5, 1.0], dtype=np.float32)
    # Fixed(2) of 3: scale 3/2; weighted: scale * p_j * deg
    w = O.neighbour_weights(cdf, [0, 2], 2, True)
    assert w == [float(f(1.5) * f(0.2) * f(3)), float(f(1.5) * (cdf[2] - cdf[1]) * f(3))]
    assert np.allclose(w, [0.9, 2.25])
    # every neighbour taken: scale 1
    assert np.allclose(O.neighbour_weights(cdf, [0, 1, 2], 5, True), [0.6, 0.9, 1.5])
    # unweighted averaging: the scale only
    assert O.neighbour_weights(cdf, [1, 2], 2, False) == [1.5, 1.5]
    # a uniform CDF gives every neighbour the scale
    cdf_u = (np.arange(1, 5) / 4).astype(np.float32)
    assert np.allclose(O.neighbour_weights(cdf_u, [0, 3], 2, True), [2.0, 2.0])


@pytest.mark.parametrize("seed", [3, 4])
def test_weighted_neighbours_exact_fd_and_first_leaf(seed):
    table, f_indptr, f_ids, v, nbrs, negs = _toy_example(seed)
    ws = list(np.random.default_rng(seed + 10).uniform(0.2, 3.0, size=len(nbrs)))
    gamma = 2.0
    r = O.example_grads(table, f_indptr, f_ids, v, nbrs, ws, negs, gamma, "exact")
    assert r.contributes
    eps = 1e-6
    loss = lambda t: O.example_grads(t, f_indptr, f_ids, v, nbrs, ws, negs, gamma, "exact").loss
    for fr in range(table.shape[0]):
        for j in range(table.shape[1]):
            tp, tm = table.copy(), table.copy()
            tp[fr, j] += eps
            tm[fr, j] -= eps
            fd = (loss(tp) - loss(tm)) / (2 * eps)
            g = r.grads[fr][j] if fr in r.grads else 0.0
            assert abs(fd - g) < 1e-6, (fr, j, fd, g)
    fl = O.example_grads(table, f_indptr, f_ids, v, nbrs, ws, negs, gamma, "first_leaf")
    L, g = torch_first_leaf(table, f_indptr, f_ids, v, nbrs, negs, gamma, nbr_weights=ws)
    assert abs(fl.loss - L) < 1e-12 and set(fl.grads) == set(g)
    for fr in g:
        assert np.allclose(fl.grads[fr], g[fr], atol=1e-12)


def test_first_leaf_differs_from_exact_only_on_shared_features():
    table, f_indptr, f_ids, v, nbrs, negs = _toy_example(3)
    fl = O.example_grads(table, f_indptr, f_ids, v, nbrs, 1.0, negs, 2.0, "first_leaf").grads
    ex = O.example_grads(table, f_indptr, f_ids, v, nbrs, 1.0, negs, 2.0, "exact").grads
    assert set(fl) == set(ex)
    assert any(not np.allclose(fl[f], ex[f]) for f in fl)  # the toy shares features


def test_first_leaf_equals_exact_when_constructions_are_disjoint():
    V, D = 9, 3
    rng = np.random.default_rng(2)
    table = rng.normal(size=(V, D))
    f_indptr = np.array([0, 2, 4, 6, 8])
    f_ids = np.arange(8)
    fl = O.example_grads(table, f_indptr, f_ids, 0, [1], 1.0, [2, 3], 5.0, "first_leaf").grads
    ex = O.example_grads(table, f_indptr, f_ids, 0, [1], 1.0, [2, 3], 5.0, "exact").grads
    assert set(fl) == set(ex)
    for f in fl:
        assert np.allclose(fl[f], ex[f])


def test_zero_distance_drops_anchor_and_positive():
    # a == p (D0 = 0): their adjoints are NaN and dropped (mod.rs:370-375); the negative keeps its
    # gradient, and a feature shared with the negative is taken from the negative (first finite leaf)
    table = np.array([[1.0, 0.0], [0.0, 1.0], [0.5, 0.5]])
    f_indptr = np.array([0, 1, 2, 4])  # node0: [0], node1: [0], node2: [1, 0]
    f_ids = np.array([0, 0, 1, 0])
    r = O.example_grads(table, f_indptr, f_ids, 0, [1], 1.0, [2], 1.0, "first_leaf")
    assert r.contributes
    assert not np.all(np.isfinite(r.adjoints[0])) and not np.all(np.isfinite(r.adjoints[1]))
    L, g = torch_first_leaf(table, f_indptr, f_ids, 0, [1], [2], 1.0)
    assert set(r.grads) == set(g) == {0, 1}
    for f in g:
        assert np.allclose(r.grads[f], g[f])


def test_batch_merge_is_dense_scatter_sum():
    table, f_indptr, f_ids, v, nbrs, negs = _toy_example(5)
    rs = [O.example_grads(table, f_indptr, f_ids, v, nbrs, 1.0, negs, 2.0, "first_leaf"),
          O.example_grads(table, f_indptr, f_ids, 4, [5, 6], 1.0, negs[::-1], 2.0, "first_leaf")]
    merged = O.merge_batch(rs, table.shape[1])
    dense = np.zeros_like(table)
    for r in rs:
        for f, g in r.grads.items():
            dense[f] += g
    for f in range(table.shape[0]):
        assert np.allclose(dense[f], merged.get(f, np.zeros(table.shape[1])))


# --------------------------------------------------------------------------------------
# Adam and the LR schedule
# --------------------------------------------------------------------------------------
def test_adam_first_update_closed_form():
    V, D = 3, 4
    rng = np.random.default_rng(4)
    table = rng.normal(size=(V, D))
    t0 = table.copy()
    m, v = np.zeros((V, D)), np.zeros((V, D))
    g = rng.normal(size=D)
    lr = 0.05
    O.adam_update(table, m, v, {1: g}, lr, pass_=1)  # t = 2
    b1, b2 = 0.9, 0.999
    step = lr * (g * (1 - b1) / (1 - b1 ** 2)) / (np.abs(g) * math.sqrt((1 - b2) / (1 - b2 ** 2)) + 1e-8)
    assert np.allclose(table[1], t0[1] - step, atol=1e-14)
    assert np.array_equal(table[0], t0[0]) and np.array_equal(table[2], t0[2])  # lazy: untouched rows stay


def test_adam_skips_nan_rows():
    table = np.ones((2, 2))
    m, v = np.zeros((2, 2)), np.zeros((2, 2))
    O.adam_update(table, m, v, {0: np.array([np.nan, 1.0]), 1: np.array([1.0, 1.0])}, 0.1, 1)
    assert np.array_equal(table[0], [1.0, 1.0]) and not np.array_equal(table[1], [1.0, 1.0])
    assert np.array_equal(m[0], [0.0, 0.0])


# --------------------------------------------------------------------------------------
# Concurrent batches (mod.rs:174 par_iter over chunks): stale reads, updates in batch order
# --------------------------------------------------------------------------------------
def _group_problem(seed, V, disjoint):
    rng = np.random.default_rng(seed)
    n_nodes, D = 12, 4
    if disjoint:  # nodes 0-5 use features [0, V/2), nodes 6-11 features [V/2, V)
        f_ids, f_indptr = [], [0]
        for i in range(n_nodes):
            lo, hi = (0, V // 2) if i < 6 else (V // 2, V)
            k = int(rng.integers(1, 4))
            f_ids += list(rng.integers(lo, hi, size=k))
            f_indptr.append(len(f_ids))
        f_ids, f_indptr = np.array(f_ids), np.array(f_indptr)
    else:
        counts = rng.integers(1, 4, size=n_nodes)
        f_indptr = np.concatenate([[0], np.cumsum(counts)])
        f_ids = rng.integers(0, V, size=f_indptr[-1])
    table = rng.normal(size=(V, D))
    # batch 0: anchors 0, 1 (graph part 0-5); batch 1: anchors 6, 7 (graph part 6-11)
    anchors = [0, 1, 6, 7]
    nbrs = [[2, 3], [4, 5], [8, 9], [10, 11]]
    negs = [[3, 4], [5, 2], [9, 10], [11, 8]]
    return table, f_indptr, f_ids, anchors, nbrs, negs


def test_group_of_one_batch_is_the_batch_step():
    table, f_indptr, f_ids, anchors, nbrs, negs = _group_problem(1, 10, False)
    t1, m1, v1 = table.copy(), np.zeros_like(table), np.zeros_like(table)
    t2, m2, v2 = table.copy(), np.zeros_like(table), np.zeros_like(table)
    O.ep_batch_step(t1, m1, v1, f_indptr, f_ids, anchors, nbrs, [1.0] * 4, negs, 3.0, 0.05, 1)
    O.ep_group_step(t2, m2, v2, f_indptr, f_ids, anchors, nbrs, [1.0] * 4, negs, 3.0, [0.05], [0, 0, 0, 0], 1)
    assert np.array_equal(t1, t2) and np.array_equal(m1, m2) and np.array_equal(v1, v2)


def test_group_on_disjoint_rows_equals_sequential_batches():
    # no row read by batch 1 is written by batch 0: stale and fresh reads coincide
    table, f_indptr, f_ids, anchors, nbrs, negs = _group_problem(2, 10, True)
    t1, m1, v1 = table.copy(), np.zeros_like(table), np.zeros_like(table)
    O.ep_batch_step(t1, m1, v1, f_indptr, f_ids, anchors[:2], nbrs[:2], [1.0] * 2, negs[:2], 3.0, 0.05, 1)
    O.ep_batch_step(t1, m1, v1, f_indptr, f_ids, anchors[2:], nbrs[2:], [1.0] * 2, negs[2:], 3.0, 0.02, 1)
    t2, m2, v2 = table.copy(), np.zeros_like(table), np.zeros_like(table)
    O.ep_group_step(t2, m2, v2, f_indptr, f_ids, anchors, nbrs, [1.0] * 4, negs, 3.0, [0.05, 0.02], 2, 1)
    assert np.allclose(t1, t2, atol=1e-15) and np.allclose(m1, m2, atol=1e-15)
    assert not np.array_equal(t1, table)


def test_group_shared_row_gets_two_adam_steps_from_stale_gradients():
    table, f_indptr, f_ids, anchors, nbrs, negs = _group_problem(3, 6, False)
    snap = table.copy()
    # per-batch merged gradients, both from the group-start parameters
    g = []
    for lo in (0, 2):
        rs = [O.example_grads(snap, f_indptr, f_ids, anchors[i], nbrs[i], 1.0, negs[i], 3.0, "first_leaf")
              for i in (lo, lo + 1)]
        g.append(O.merge_batch(rs, 4))
    shared = sorted(set(g[0]) & set(g[1]))
    assert shared
    t, m, v = table.copy(), np.zeros_like(table), np.zeros_like(table)
    O.ep_group_step(t, m, v, f_indptr, f_ids, anchors, nbrs, [1.0] * 4, negs, 3.0, [0.05, 0.02], 2, 2)
    b1, b2, eps = 0.9, 0.999, 1e-8
    bc1, bc2 = 1 - b1 ** 3, 1 - b2 ** 3  # t = pass + 1 = 3
    for f in shared:
        g1, g2 = g[0][f], g[1][f]
        mm, vv = (1 - b1) * g1, (1 - b2) * g1 ** 2
        th = snap[f] - 0.05 * (mm / bc1) / (np.sqrt(vv / bc2) + eps)
        mm, vv = b1 * mm + (1 - b1) * g2, b2 * vv + (1 - b2) * g2 ** 2
        th = th - 0.02 * (mm / bc1) / (np.sqrt(vv / bc2) + eps)
        assert np.allclose(t[f], th, atol=1e-14) and np.allclose(m[f], mm, atol=1e-15)


def test_noise_schedule_closed_form():
    # sd(0) = noise, geometric decay, 1e-9 at the last step; no-op at or below 1e-9
    noise, T = 0.1, 50
    assert O.noise_sd(0, noise, T) == pytest.approx(noise, rel=1e-12)
    assert O.noise_sd(T, noise, T) == pytest.approx(1e-9, rel=1e-9)
    r = [O.noise_sd(k + 1, noise, T) / O.noise_sd(k, noise, T) for k in range(5)]
    assert max(r) - min(r) < 1e-12 and r[0] == pytest.approx((1e-9 / noise) ** (1.0 / T))
    assert O.noise_sd(3, 1e-9, T) == 0.0 and O.noise_sd(3, 0.0, T) == 0.0


def test_cos_decay_schedule_closed_form():
    alpha, T = 0.01, 100
    lo = alpha / 100
    warm = 20
    assert math.isclose(O.cos_decay_lr(1, alpha, T), lo + (1 / warm) * (alpha - lo))
    assert math.isclose(O.cos_decay_lr(warm, alpha, T), alpha)
    # no warm-up offset after warm-up: jump to the cosine at step/T (scheduler.rs:39-43)
    assert math.isclose(O.cos_decay_lr(warm + 1, alpha, T), lo + 0.5 * (alpha - lo) * (1 + math.cos(math.pi * 21 / 100)))
    assert math.isclose(O.cos_decay_lr(T, alpha, T), lo)


def test_init_rows_unit_norm():
    rng = np.random.default_rng(0)
    x = O.normalize_rows(rng.uniform(size=(100, 16)))
    assert np.allclose(np.linalg.norm(x, axis=1), 1.0)
    assert abs(x.mean()) < 0.05


# --------------------------------------------------------------------------------------
# Attention: reference KATs src/algos/ep/attention.rs:303-425
# --------------------------------------------------------------------------------------
ATT_ROWS = np.array([[-1.0, -1.0, 1.0, 1.0], [0.0, 0.0, 2.0, 2.0], [1.0, 1.0, -1.0, -1.0]])


def test_attention_matrix_kat():
    Q, K, V, dm = O.attention_split(ATT_ROWS, heads=1, d_k=1)
    S = O.attention_matrix_full(Q[0], K[0], np.ones(3))
    assert S.tolist() == [[1.0, -0.0, -1.0], [-0.0, 0.0, 0.0], [-1.0, 0.0, 1.0]]


def test_attention_softmax_kat():
    Q, K, V, dm = O.attention_split(ATT_ROWS, heads=1, d_k=1)
    S = O.attention_matrix_full(Q[0], K[0], np.ones(3))
    A = np.stack([O.softmax_exact(S[i]) for i in range(3)])
    exp = [[0.66524096, 0.24472847, 0.09003057], [1 / 3, 1 / 3, 1 / 3], [0.09003057, 0.24472847, 0.66524096]]
    assert np.allclose(A, exp, atol=1e-7)


def test_attention_reweighted_kat():
    Q, K, V, dm = O.attention_split(ATT_ROWS, heads=1, d_k=1)
    S = O.attention_matrix_full(Q[0], K[0], np.ones(3))
    A = np.stack([O.softmax_exact(S[i]) for i in range(3)])
    rows = A @ V[0]
    exp = [[1.0647, 1.0647], [0.6667, 0.6667], [-0.0858, -0.0858]]
    assert np.all(np.abs(rows - exp) < 1e-4)
    # attention_mean = (1/n) sum_i rows_i, mean over heads (attention.rs:122-130)
    assert np.allclose(O.attention_mean(ATT_ROWS, np.ones(3), 1, 1), rows.sum(axis=0) / 3)


def test_attention_single_feature_returns_head0_value():
    rows = np.array([[1.0, 2.0, 3.0, 4.0, 5.0, 6.0, 7.0, 8.0]])  # H=2, d_k=1, d_model=2
    assert O.attention_mean(rows, np.ones(1), 2, 1).tolist() == [5.0, 6.0]


def test_attention_count_scaling_kat():
    # logits multiplied by c_i c_j when >= 1 (attention.rs:183-186)
    Q, K, V, dm = O.attention_split(ATT_ROWS, heads=1, d_k=1)
    S = O.attention_matrix_full(Q[0], K[0], np.array([2.0, 2.0, 2.0]))
    assert np.allclose(S, 4.0 * (Q[0] @ K[0].T))
    S2 = O.attention_matrix_full(Q[0], K[0], np.array([0.5, 0.5, 0.5]))
    assert np.allclose(S2, Q[0] @ K[0].T)


@pytest.mark.parametrize("n,heads,dk,dm", [(5, 2, 3, 4), (3, 1, 2, 2), (7, 4, 2, 3)])
def test_attention_backward_vs_fd_and_torch(n, heads, dk, dm):
    rng = np.random.default_rng(n * 10 + heads)
    W = heads * (2 * dk + dm)
    rows = rng.normal(size=(n, W))
    counts = rng.integers(1, 3, size=n).astype(np.float64)
    g = rng.normal(size=dm)
    grad = O.attention_mean_backward(rows, counts, heads, dk, g)
    # torch fp64 autograd on an independent formulation
    R = torch.tensor(rows, dtype=torch.float64, requires_grad=True)
    c = torch.tensor(counts, dtype=torch.float64)
    out = 0
    for h in range(heads):
        Qh = R[:, h * dk:(h + 1) * dk]
        Kh = R[:, heads * dk + h * dk: heads * dk + (h + 1) * dk]
        Vh = R[:, 2 * heads * dk + h * dm: 2 * heads * dk + (h + 1) * dm]
        num = torch.outer(c, c)
        S = (Qh @ Kh.T) * torch.where(num >= 1, num, torch.ones_like(num)) / math.sqrt(dk)
        A = torch.softmax(S, dim=1)
        out = out + (A.sum(0) @ Vh) / n
    out = out / heads
    (out @ torch.tensor(g)).backward()
    assert np.allclose(grad, R.grad.numpy(), atol=1e-10)
    assert np.allclose(O.attention_mean(rows, counts, heads, dk), out.detach().numpy(), atol=1e-12)
    eps = 1e-6
    for i in range(n):
        for j in range(W):
            rp, rm = rows.copy(), rows.copy()
            rp[i, j] += eps
            rm[i, j] -= eps
            fd = (O.attention_mean(rp, counts, heads, dk) @ g - O.attention_mean(rm, counts, heads, dk) @ g) / (2 * eps)
            assert abs(fd - grad[i, j]) < 1e-7


# --------------------------------------------------------------------------------------
# Attention model example: model.rs:158-249, 306-335, 419-446
# --------------------------------------------------------------------------------------
def _att_toy(seed, V=10, heads=2, dk=2, dm=3, n_nodes=10, K=3):
    rng = np.random.default_rng(seed)
    counts = rng.integers(1, 5, size=n_nodes)
    counts[4] = 1  # a single-feature node (attention.rs:113-115)
    f_indptr = np.concatenate([[0], np.cumsum(counts)])
    f_ids = rng.integers(0, V, size=f_indptr[-1])  # repeated features inside nodes and across them
    W = heads * (2 * dk + dm)
    table = 0.7 * rng.normal(size=(V, W))
    negs = [4] + list(rng.integers(0, n_nodes, size=K - 1))
    return table, f_indptr, f_ids, 0, [1, 2, 3], negs


def test_attention_item_rules():
    f_indptr = np.array([0, 3])
    f_ids = np.array([5, 7, 5])
    # h(v): one item per distinct feature, each occurrence adds 1 + 1 (model.rs:267-268)
    assert O.attention_node_items(f_indptr, f_ids, 0) == ([5, 7], [4.0, 2.0])
    # ~h(v) block: the list as listed, count = multiplicity in the neighbour (model.rs:426-443)
    assert O.attention_neighbour_items(f_indptr, f_ids, 0) == ([5, 7, 5], [2.0, 1.0, 2.0])
    # inference: every listed feature, count 1 (aggregator.rs:141-146)
    assert O.attention_inference_items([5, 7, 5]) == ([5, 7, 5], [1.0, 1.0, 1.0])


def _torch_attention(leaf_rows, counts, heads, dk):
    n = len(leaf_rows)
    R = torch.stack(leaf_rows)
    dm = (R.shape[1] - 2 * heads * dk) // heads
    if n == 1:
        return R[0, 2 * heads * dk: 2 * heads * dk + dm]
    c = torch.tensor(counts, dtype=torch.float64)
    num = torch.outer(c, c)
    m = torch.where(num >= 1, num, torch.ones_like(num))
    out = 0
    for h in range(heads):
        Qh = R[:, h * dk:(h + 1) * dk]
        Kh = R[:, heads * dk + h * dk: heads * dk + (h + 1) * dk]
        Vh = R[:, 2 * heads * dk + h * dm: 2 * heads * dk + (h + 1) * dm]
        A = torch.softmax((Qh @ Kh.T) * m / math.sqrt(dk), dim=1)
        out = out + (A @ Vh).sum(0) / n
    return out / heads


def torch_attention_first_leaf(table, f_indptr, f_ids, v, nbrs, negs, gamma, heads, dk):
    """The reference's mechanism in torch fp64: one leaf per (construction, distinct
    feature); h(v)/h(u) attend over the map entries with count 2 x multiplicity; ~h(v)
    runs one attention per neighbour over its listed features (counts = multiplicity
    in that neighbour), all reading the construction's leaves, then the plain mean."""
    def leaf(f):
        return torch.tensor(table[f], dtype=torch.float64, requires_grad=True)

    def node_con(u):
        d = {}
        for f in f_ids[f_indptr[u]:f_indptr[u + 1]]:
            f = int(f)
            if f in d:
                d[f][1] += 2.0
            else:
                d[f] = [leaf(f), 2.0]
        return d, _torch_attention([x[0] for x in d.values()], [x[1] for x in d.values()], heads, dk)

    a_map, a = node_con(v)
    p_map, outs = {}, []
    for u in nbrs:
        fs = [int(f) for f in f_ids[f_indptr[u]:f_indptr[u + 1]]]
        for f in fs:
            if f not in p_map:
                p_map[f] = [leaf(f), 0.0]
        mult = {f: fs.count(f) for f in fs}
        outs.append(_torch_attention([p_map[f][0] for f in fs], [float(mult[f]) for f in fs], heads, dk))
    p = sum(outs) / len(outs)
    neg = [node_con(u) for u in negs]
    d1 = gamma + torch.sqrt(((p - a) ** 2).sum())
    ls = [d1 - torch.sqrt(((p - x) ** 2).sum()) for _, x in neg]
    pos = [l for l in ls if l.item() > 0]
    if not pos:
        return 0.0, {}
    L = sum(pos) / len(pos)
    L.backward()
    grads = {}
    for d in [a_map, p_map] + [m for m, _ in neg]:
        for f, (lf, _) in d.items():
            if f in grads or lf.grad is None:
                continue
            g = lf.grad.numpy()
            if np.all(np.isfinite(g)):
                grads[f] = g.copy()
    return L.item(), grads


@pytest.mark.parametrize("seed,heads,dk,dm,gamma", [(1, 2, 2, 3, 1.0), (2, 1, 3, 2, 0.5), (3, 3, 2, 2, 1.0),
                                                     (4, 2, 2, 3, 0.2)])
def test_attention_first_leaf_matches_per_leaf_autograd(seed, heads, dk, dm, gamma):
    table, f_indptr, f_ids, v, nbrs, negs = _att_toy(seed, heads=heads, dk=dk, dm=dm)
    r = O.attention_example_grads(table, f_indptr, f_ids, v, nbrs, negs, gamma, heads, dk, "first_leaf")
    L, g = torch_attention_first_leaf(table, f_indptr, f_ids, v, nbrs, negs, gamma, heads, dk)
    assert abs(r.loss - L) < 1e-12
    assert set(r.grads) == set(g)
    for f in g:
        assert np.allclose(r.grads[f], g[f], atol=1e-12)


@pytest.mark.parametrize("seed", [5, 6])
def test_attention_exact_grads_central_fd(seed):
    heads, dk, dm, gamma = 2, 2, 3, 6.0  # a wide margin keeps every negative active under +-eps
    table, f_indptr, f_ids, v, nbrs, negs = _att_toy(seed, heads=heads, dk=dk, dm=dm)
    r = O.attention_example_grads(table, f_indptr, f_ids, v, nbrs, negs, gamma, heads, dk, "exact")
    assert r.contributes and all(x is not None for x in r.adjoints)
    eps = 1e-6
    for f in range(table.shape[0]):
        for j in range(table.shape[1]):
            tp, tm = table.copy(), table.copy()
            tp[f, j] += eps
            tm[f, j] -= eps
            lp = O.attention_example_grads(tp, f_indptr, f_ids, v, nbrs, negs, gamma, heads, dk, "exact").loss
            lm = O.attention_example_grads(tm, f_indptr, f_ids, v, nbrs, negs, gamma, heads, dk, "exact").loss
            got = r.grads[f][j] if f in r.grads else 0.0
            assert abs((lp - lm) / (2 * eps) - got) < 1e-7


def test_attention_single_feature_node_and_permutation():
    heads, dk, dm = 2, 2, 3
    table, f_indptr, f_ids, v, nbrs, negs = _att_toy(7, heads=heads, dk=dk, dm=dm)
    r = O.attention_example_grads(table, f_indptr, f_ids, v, nbrs, negs, 0.5, heads, dk, "exact")
    # negative 0 is node 4 with one feature: its construction is that row's head-0 value
    f4 = int(f_ids[f_indptr[4]])
    assert np.array_equal(r.vectors[2], table[f4, 2 * heads * dk: 2 * heads * dk + dm])
    # Full attention over the anchor's distinct features does not depend on their order
    fs = f_ids[f_indptr[0]:f_indptr[1]][::-1].copy()
    f2 = f_ids.copy()
    f2[f_indptr[0]:f_indptr[1]] = fs
    r2 = O.attention_example_grads(table, f_indptr, f2, v, nbrs, negs, 0.5, heads, dk, "exact")
    assert np.allclose(r.vectors[0], r2.vectors[0], atol=1e-14)


# --------------------------------------------------------------------------------------
# AUC
# --------------------------------------------------------------------------------------
def test_roc_auc_matches_sklearn():
    from sklearn.metrics import roc_auc_score

    rng = np.random.default_rng(0)
    pos = np.round(rng.normal(1.0, 1.0, 300), 1)
    neg = np.round(rng.normal(0.0, 1.0, 400), 1)  # rounding creates ties
    y = np.concatenate([np.ones(300), np.zeros(400)])
    assert abs(O.roc_auc(pos, neg) - roc_auc_score(y, np.concatenate([pos, neg]))) < 1e-12
    assert O.roc_auc(np.array([2.0, 3.0]), np.array([0.0, 1.0])) == 1.0
    assert O.roc_auc(np.array([1.0]), np.array([1.0])) == 0.5


def test_averaged_node_embeddings_closed_form():
    table = np.array([[1.0, 1.0], [3.0, -1.0]])
    out = O.averaged_node_embeddings(table, np.array([0, 2, 2, 3]), np.array([0, 1, 1]))
    assert np.allclose(out, [[2.0, 0.0], [0.0, 0.0], [3.0, -1.0]])


# --------------------------------------------------------------------------------------
# SIF-weighted inference aggregate: aggregator.rs:82-116
# --------------------------------------------------------------------------------------
def test_weighted_aggregate_closed_form():
    table = np.array([[1.0, 0.0], [0.0, 3.0], [5.0, 5.0]])
    p_w = np.array([0.5, 0.25, 0.25])
    f_indptr = np.array([0, 3, 3, 4])  # node 0: [0, 1, 1]; node 1: none; node 2: [2]
    f_ids = np.array([0, 1, 1, 2])
    alpha = 0.5
    w0, w1 = 0.5 / 1.0, 0.5 / 0.75  # alpha / (alpha + p)
    out = O.weighted_node_embeddings(table, f_indptr, f_ids, p_w, alpha)
    assert np.allclose(out[0], (w0 * table[0] + 2 * w1 * table[1]) / (w0 + 2 * w1))
    assert np.all(out[1] == 0)            # no features: the row stays zero
    assert np.allclose(out[2], table[2])  # one feature: its row
    # a large alpha makes the weights equal: the plain mean of the listed rows
    big = O.weighted_node_embeddings(table, f_indptr, f_ids, p_w, 1e12)
    assert np.allclose(big, O.averaged_node_embeddings(table, f_indptr, f_ids))


# --------------------------------------------------------------------------------------
# Other losses: loss.rs:56-151
# --------------------------------------------------------------------------------------
def test_il2norm_kat():
    # loss.rs:258-264: il2norm([1, 3]) = [1, 3] / sqrt(10)
    u, n = O._unit(np.array([1.0, 3.0]))
    assert np.allclose(u, np.array([1.0, 3.0]) / math.sqrt(10.0)) and n == math.sqrt(10.0)


def _torch_loss(kind, a, p, negs, params):
    """The reference's Loss::compute written again in torch (fp64), gradients by autograd."""
    A = torch.tensor(a, dtype=torch.float64, requires_grad=True)
    Pp = torch.tensor(p, dtype=torch.float64, requires_grad=True)
    Ns = [torch.tensor(n, dtype=torch.float64, requires_grad=True) for n in negs]
    unit = lambda x: x / torch.sqrt((x * x).sum())

    def starspace(gamma):
        rd = unit(Pp) @ unit(A)
        ls = [gamma - (rd - unit(A) @ unit(n)) for n in Ns]
        ls = [l for l in ls if l.item() > 0]
        return sum(ls) / len(ls) if ls else None

    def rank(tau):
        ds = torch.stack([n @ A for n in Ns] + [A @ Pp])
        sm = torch.softmax(ds, dim=0)
        return -torch.log(sm[-1]) if sm[-1].item() < tau else None

    if kind == "starspace":
        L = starspace(params[0])
    elif kind == "rank":
        L = rank(params[0])
    elif kind == "rankspace":
        parts = [x for x in (starspace(params[0]), rank(params[0])) if x is not None]
        L = sum(parts) if parts else None
    else:  # contrastive
        pos = params[0] - unit(Pp) @ unit(A)
        ts = [t for t in [unit(A) @ unit(n) - params[1] for n in Ns] if t.item() > 0]
        if pos.item() > 0:
            ts.append(pos)
        L = sum(ts) / len(ts) if ts else None
    if L is None:
        return 0.0, [None] * (len(negs) + 2)
    L.backward()
    return L.item(), [None if x.grad is None else x.grad.numpy() for x in [A, Pp] + Ns]


LOSS_CASES = [("starspace", (0.3,)), ("starspace", (1.2,)), ("contrastive", (0.6, 0.1)), ("contrastive", (-0.9, -0.2)),
              ("rank", (0.9,)), ("rank", (0.05,)), ("rankspace", (0.4,)), ("rankspace", (0.9,))]


@pytest.mark.parametrize("kind,params", LOSS_CASES)
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_loss_adjoints_vs_torch_and_fd(kind, params, seed):
    rng = np.random.default_rng(seed)
    d, K = 5, 4
    a, p = rng.normal(size=(2, d))
    negs = list(rng.normal(size=(K, d)))
    L, reached, ga, gp, gn = O.loss_adjoints(kind, a, p, negs, params)
    Lt, gt = _torch_loss(kind, a, p, negs, params)
    assert abs(L - Lt) < 1e-12
    for g, r, t in zip([ga, gp] + gn, reached, gt):
        assert (g is None) == (t is None) == (not r)
        if g is not None:
            assert np.allclose(g, t, atol=1e-12)
    if L > 0:
        eps = 1e-6
        vecs = [a, p] + negs
        for vi in range(len(vecs)):
            for j in range(d):
                vp = [v.copy() for v in vecs]
                vm = [v.copy() for v in vecs]
                vp[vi][j] += eps
                vm[vi][j] -= eps
                fd = (O.loss_adjoints(kind, vp[0], vp[1], vp[2:], params)[0]
                      - O.loss_adjoints(kind, vm[0], vm[1], vm[2:], params)[0]) / (2 * eps)
                g = ([ga, gp] + gn)[vi]
                assert abs(fd - (0.0 if g is None else g[j])) < 1e-6


def test_loss_closed_forms():
    a, p = np.array([1.0, 0.0]), np.array([0.0, 2.0])     # a^.p^ = 0
    n1, n2 = np.array([3.0, 0.0]), np.array([-1.0, 0.0])  # a^.n^ = 1, -1
    # starspace: l_i = gamma - (0 - a^.n_i^) -> 0.5 + 1, 0.5 - 1 (inactive)
    L, r, *_ = O.loss_adjoints("starspace", a, p, [n1, n2], (0.5,))
    assert L == 1.5 and r == [True, True, True, False]
    # contrastive: pos = 0.7 - 0; negatives a^.n^ - 0.2 -> 0.8 (kept), -1.2 (dropped)
    L, r, *_ = O.loss_adjoints("contrastive", a, p, [n1, n2], (0.7, 0.2))
    assert abs(L - (0.8 + 0.7) / 2) < 1e-15 and r == [True, True, True, False]
    # rank: ds = (3, -1, 0); s_pos = 1 / (e^3 + e^-1 + 1)
    L, r, *_ = O.loss_adjoints("rank", a, p, [n1, n2], (0.5,))
    assert abs(L - math.log(math.exp(3) + math.exp(-1) + 1)) < 1e-12 and all(r)
    L, r, *_ = O.loss_adjoints("rank", a, p, [n1, n2], (0.01,))  # s_pos ~ 0.045 >= tau: no loss
    assert L == 0 and not any(r)


# --------------------------------------------------------------------------------------
# NeighborhoodAligner (alignment.rs:25-75)
# --------------------------------------------------------------------------------------
def _align_toy():
    # 0 -> 1 (0.25), 0 -> 2 (0.75); 1 -> 0 (1.0); 2 has no edges
    indptr = np.array([0, 2, 3, 3])
    indices = np.array([1, 2, 0])
    cdf = np.array([0.25, 1.0, 1.0], dtype=np.float32)
    emb = np.array([[1.0, 0.0], [0.0, 2.0], [4.0, 4.0]])
    return indptr, indices, cdf, emb


def test_align_closed_form():
    indptr, indices, cdf, emb = _align_toy()
    out = O.neighborhood_align(indptr, indices, cdf, emb, 0.5, None)
    assert np.allclose(out[0], 0.5 * (0.25 * emb[1] + 0.75 * emb[2]) + 0.5 * emb[0])   # [2.0, 1.75]
    assert np.allclose(out[0], [2.0, 1.75])
    assert np.allclose(out[1], 0.5 * emb[0] + 0.5 * emb[1])
    assert np.allclose(out[2], 0.5 * emb[2])  # no edges
    # alpha 1: unchanged; alpha 0: the weighted neighbourhood sum
    assert np.allclose(O.neighborhood_align(indptr, indices, cdf, emb, 1.0, None), emb)
    assert np.allclose(O.neighborhood_align(indptr, indices, cdf, emb, 0.0, None)[0], [3.0, 3.5])


def test_align_max_neighbors_does_not_renormalise():
    indptr, indices, cdf, emb = _align_toy()
    out = O.neighborhood_align(indptr, indices, cdf, emb, 0.0, 1)
    assert np.allclose(out[0], 0.25 * emb[1])  # only the first edge, weight 0.25


def test_align_adaptive_alpha():
    indptr, indices, cdf, emb = _align_toy()
    out = O.neighborhood_align(indptr, indices, cdf, emb, None, None)
    a0 = 1.0 / math.log(3.0)  # degree 2: 0.910
    assert np.allclose(out[0], (1 - a0) * np.array([3.0, 3.5]) + a0 * emb[0])
    # degree 1: 1 / ln 2 > 1 -> 1; degree 0: 1 / ln 1 = inf -> 1
    assert np.allclose(out[1], emb[1]) and np.allclose(out[2], emb[2])
    # the cut keeps the full degree for alpha
    cut = O.neighborhood_align(indptr, indices, cdf, emb, None, 1)
    assert np.allclose(cut[0], (1 - a0) * 0.25 * emb[1] + a0 * emb[0])


def test_align_uniform_weights_give_neighbour_mean():
    rng = np.random.default_rng(4)
    n = 30
    rows, cols, _ = O.csr_from_edges([(int(a), int(b), 1.0) for a, b in rng.integers(0, n, (120, 2)) if a != b],
                                     deduplicate=True)
    indptr = np.zeros(n + 1, dtype=np.int64)
    indptr[:len(rows)] = rows
    indptr[len(rows):] = rows[-1]
    cdf = np.concatenate([np.arange(1, d + 1, dtype=np.float32) / np.float32(d) for d in np.diff(indptr)])
    emb = rng.normal(size=(n, 5))
    out = O.neighborhood_align(indptr, cols, cdf, emb, 0.0, None)
    for v in range(n):
        s, e = indptr[v], indptr[v + 1]
        want = emb[cols[s:e]].mean(axis=0) if e > s else np.zeros(5)
        assert np.allclose(out[v], want, atol=1e-6)


# --------------------------------------------------------------------------------------
# max_features sampling (utils.rs:193-216, model.rs:257-278)
# --------------------------------------------------------------------------------------
def test_probability_sample_kat():
    # Binomial draw k = 0 with at_least_one: one item, scale (1 - p^n) / p (as written)
    k, s = O.probability_sample(0.25, 3, 0)
    assert k == 1 and s == float((np.float32(1) - np.float32(0.25) ** np.float32(3)) / np.float32(0.25))
    assert abs(s - 3.9375) < 1e-7
    assert O.probability_sample(0.25, 3, 2) == (2, 4.0)
    assert O.sample_spec("fixed", 2, 5) == (2, 2.5) and O.sample_spec("fixed", 7, 5) == (5, 1.0)


def test_feature_selection_constructions_closed_form():
    # node 0: features [0, 1, 2, 3]; node 1: [1, 4]; node 2: [2]
    f_indptr = np.array([0, 4, 6, 7], dtype=np.uint64)
    f_ids = np.array([0, 1, 2, 3, 1, 4, 2], dtype=np.uint32)
    # h(v) keeps positions {1, 3} with scale 2: counts 1 + 2 each -> plain mean of rows 1, 3
    a = O.build_anchor_construction(f_indptr, f_ids, 0, ([3, 1], 2.0))
    assert a.counts == {1: 3.0, 3: 3.0}
    # ~h(v) over nodes 1 (weight 1.5, keeps position 1 = feature 4, scale 2) and 2 (weight
    # 1.5, all, scale 1): counts 1.5 + 2 for feature 4 and 1.5 + 1 for feature 2
    p = O.build_neighbour_construction(f_indptr, f_ids, [1, 2], 1.5, [([1], 2.0), ([0], 1.0)])
    assert p.counts == {4: 3.5, 2: 2.5}
    table = np.eye(5)
    assert np.allclose(O.construct_mean(table, p), (3.5 * table[4] + 2.5 * table[2]) / 6.0)
    # taking every position with scale 1 is the unsampled path
    full = O.build_neighbour_construction(f_indptr, f_ids, [0, 1], 1.0, [([0, 1, 2, 3], 1.0), ([0, 1], 1.0)])
    assert full.counts == O.build_neighbour_construction(f_indptr, f_ids, [0, 1], 1.0).counts


@pytest.mark.parametrize("seed", [3, 4])
def test_feature_selection_exact_grads_fd(seed):
    table, f_indptr, f_ids, v, nbrs, negs = _toy_example(seed)
    rng = np.random.default_rng(seed + 20)
    sel = []
    for u in [v] + list(nbrs) + list(negs):
        n = int(f_indptr[u + 1] - f_indptr[u])
        r = int(rng.integers(1, n + 1))
        sel.append((sorted(rng.choice(n, r, replace=False).tolist()), n / r))
    gamma = 2.0
    r = O.example_grads(table, f_indptr, f_ids, v, nbrs, 1.0, negs, gamma, "exact", feat_sel=sel)
    assert r.contributes
    eps = 1e-6
    loss = lambda t: O.example_grads(t, f_indptr, f_ids, v, nbrs, 1.0, negs, gamma, "exact", feat_sel=sel).loss
    for fr in range(table.shape[0]):
        for j in range(table.shape[1]):
            tp, tm = table.copy(), table.copy()
            tp[fr, j] += eps
            tm[fr, j] -= eps
            g = r.grads[fr][j] if fr in r.grads else 0.0
            assert abs((loss(tp) - loss(tm)) / (2 * eps) - g) < 1e-6


# --------------------------------------------------------------------------------------
# Sliding attention (attention.rs:194-218; model.rs:320-329)
# --------------------------------------------------------------------------------------
def test_sliding_attention_matrix_kat():
    # attention.rs test_attention_matrix_cw: window 1 drops (0, 2) and (2, 0); window 10 is full
    Q, K, V, dm = O.attention_split(ATT_ROWS, heads=1, d_k=1)
    S, mask, _ = O.attention_scores(Q[0], K[0], np.ones(3), 1)
    assert mask.tolist() == [[True, True, False], [True, True, True], [False, True, True]]
    assert np.where(mask, S, 0.0).tolist() == [[1.0, -0.0, 0.0], [-0.0, 0.0, 0.0], [0.0, 0.0, 1.0]]
    S10, mask10, _ = O.attention_scores(Q[0], K[0], np.ones(3), 10)
    assert mask10.all() and np.array_equal(S10, O.attention_matrix_full(Q[0], K[0], np.ones(3)))


def test_sliding_attention_ignores_counts_and_softmaxes_the_window():
    Q, K, V, dm = O.attention_split(ATT_ROWS, heads=1, d_k=1)
    counts = np.array([2.0, 3.0, 2.0])
    S, mask, m = O.attention_scores(Q[0], K[0], counts, 1)
    assert np.all(m == 1.0)  # no count scaling (attention.rs:194-218)
    A = O.masked_softmax_rows(S, mask)
    assert A[0, 2] == 0 and A[2, 0] == 0 and np.allclose(A.sum(axis=1), 1.0)
    assert np.allclose(A[0, :2], O.softmax_exact(np.array([1.0, 0.0])))
    # a window covering every pair equals Full attention with counts below 1
    assert np.allclose(O.attention_mean(ATT_ROWS, np.full(3, 0.5), 1, 1, window=5),
                       O.attention_mean(ATT_ROWS, np.full(3, 0.5), 1, 1))


@pytest.mark.parametrize("window", [0, 1, 2])
def test_sliding_attention_backward_vs_fd(window):
    rng = np.random.default_rng(window + 3)
    heads, dk, dm, n = 2, 2, 3, 6
    rows = rng.normal(size=(n, heads * (2 * dk + dm)))
    counts = rng.integers(1, 3, size=n).astype(np.float64)
    g = rng.normal(size=dm)
    an = O.attention_mean_backward(rows, counts, heads, dk, g, window)
    eps = 1e-6
    for i in range(n):
        for j in range(rows.shape[1]):
            rp, rm = rows.copy(), rows.copy()
            rp[i, j] += eps
            rm[i, j] -= eps
            fd = (O.attention_mean(rp, counts, heads, dk, window) @ g
                  - O.attention_mean(rm, counts, heads, dk, window) @ g) / (2 * eps)
            assert abs(fd - an[i, j]) < 1e-7


def test_sliding_items_keep_the_listed_order():
    f_indptr = np.array([0, 4], dtype=np.uint64)
    f_ids = np.array([5, 2, 5, 7], dtype=np.uint32)
    assert O.attention_node_items(f_indptr, f_ids, 0, sliding=True) == ([5, 2, 5, 7], [4.0, 2.0, 4.0, 2.0])
    assert O.attention_node_items(f_indptr, f_ids, 0) == ([5, 2, 7], [4.0, 2.0, 2.0])


def test_sliding_attention_example_grads_fd():
    table, f_indptr, f_ids, v, nbrs, negs = _toy_example(5, V=12, D=2 * (2 * 2 + 3))
    r = O.attention_example_grads(table, f_indptr, f_ids, v, nbrs, negs, 2.0, 2, 2, "exact", window=1)
    assert r.contributes
    eps = 1e-6
    loss = lambda t: O.attention_example_grads(t, f_indptr, f_ids, v, nbrs, negs, 2.0, 2, 2, "exact", window=1).loss
    for fr in range(table.shape[0]):
        for j in range(table.shape[1]):
            tp, tm = table.copy(), table.copy()
            tp[fr, j] += eps
            tm[fr, j] -= eps
            g = r.grads[fr][j] if fr in r.grads else 0.0
            assert abs((loss(tp) - loss(tm)) / (2 * eps) - g) < 1e-6


# --------------------------------------------------------------------------------------
# Random attention (attention.rs:221-250)
# --------------------------------------------------------------------------------------
def test_random_attention_full_masks_equal_scaled_rows():
    rng = np.random.default_rng(12)
    heads, dk, dm, n = 2, 3, 4, 5
    rows = rng.normal(size=(n, heads * (2 * dk + dm)))
    low = np.full(n, 0.5)  # counts below 1: no count scaling under Full attention
    full = [(np.ones((n, n), bool), 1.0)] * heads
    assert np.allclose(O.attention_mean(rows, low, heads, dk, rand=full), O.attention_mean(rows, low, heads, dk))
    # the item scale multiplies q, k and v: the same as scaling every row
    sc = [(np.ones((n, n), bool), 1.7)] * heads
    assert np.allclose(O.attention_mean(rows, low, heads, dk, rand=sc), O.attention_mean(1.7 * rows, low, heads, dk))
    # counts do not scale the scores
    assert np.allclose(O.attention_mean(rows, np.full(n, 3.0), heads, dk, rand=full),
                       O.attention_mean(rows, low, heads, dk))


def test_random_attention_mask_kat():
    # one head, d_k 1: row 0 attends to {0}, row 1 to {0, 2}, row 2 to {1}
    mask = np.array([[1, 0, 0], [1, 0, 1], [0, 1, 0]], dtype=bool)
    Q, K, V, dm = O.attention_split(ATT_ROWS, heads=1, d_k=1)
    out = O.attention_mean(ATT_ROWS, np.ones(3), 1, 1, rand=[(mask, 1.0)])
    a1 = O.softmax_exact(np.array([0.0 * -1.0, 0.0 * 1.0]))  # row 1: q_1 . k_0, q_1 . k_2
    A = np.array([[1.0, 0.0, 0.0], [a1[0], 0.0, a1[1]], [0.0, 1.0, 0.0]])
    assert np.allclose(out, (A.sum(axis=0) @ V[0]) / 3)


@pytest.mark.parametrize("seed", [1, 2])
def test_random_attention_backward_vs_fd(seed):
    rng = np.random.default_rng(seed)
    heads, dk, dm, n = 2, 2, 3, 6
    rows = rng.normal(size=(n, heads * (2 * dk + dm)))
    counts = rng.integers(1, 3, size=n).astype(np.float64)
    rand = []
    for h in range(heads):
        m = rng.random((n, n)) < 0.5
        m[np.arange(n), rng.integers(0, n, n)] = True  # every row has a column
        rand.append((m, float(rng.uniform(0.5, 2.0))))
    g = rng.normal(size=dm)
    an = O.attention_mean_backward(rows, counts, heads, dk, g, rand=rand)
    eps = 1e-6
    for i in range(n):
        for j in range(rows.shape[1]):
            rp, rm = rows.copy(), rows.copy()
            rp[i, j] += eps
            rm[i, j] -= eps
            fd = (O.attention_mean(rp, counts, heads, dk, rand=rand) @ g
                  - O.attention_mean(rm, counts, heads, dk, rand=rand) @ g) / (2 * eps)
            assert abs(fd - an[i, j]) < 1e-7
