"""FP64 CPU oracle for one Embedding Propagation (EP) training step.

TEST INFRASTRUCTURE ONLY.  Nothing in the shipped package imports this module;
only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` use it, and only as the checker.

It restates, in float64 numpy with an *explicit-samples* interface, the
algorithm of the Rust reference (paths relative to the reference checkout):

* graph construction          src/graph.rs:70-142, 414-422, 630-691
* ``Sample::sample``           src/algos/utils.rs:193-216
* feature collection / mean    src/algos/ep/model.rs:257-301, 342-416, 448-463
* margin loss                  src/algos/ep/loss.rs:59-73, 241-243
* loss weighting               src/algos/ep/mod.rs:190-200
* gradient extraction          src/algos/ep/mod.rs:329-378
* batch merge                  src/algos/ep/mod.rs:223-233
* cosine LR schedule           src/algos/grad_utils/scheduler.rs:25-54, mod.rs:132-137
* sparse lazy Adam             src/algos/grad_utils/optimizer.rs:43-81
* random init                  src/algos/ep/mod.rs:381-392
* multi-head attention         src/algos/ep/attention.rs:100-296, model.rs:306-335, 419-446

Random streams (rand 0.7 XorShift, ``choose_multiple``, reservoir sampling)
are *not* restated: callers pass the sampled neighbour and negative lists
explicitly, so the oracle and the GPU path see identical samples.

Pinning (what checks each function; see tests/test_oracle.py):
  csr_from_edges / cdf_convert   reference KATs graph.rs:707-776 (+ dedup order)
  euclidean                      reference KAT loss.rs:250-255 (sqrt 8)
  construct_mean                 closed forms on hand-made inputs, invariants
  margin_loss_adjoints           central finite differences, closed form
  loss_adjoints                  il2norm KAT (loss.rs:258-264); central finite differences and
                                 torch fp64 autograd of every loss kind; closed forms
  example_grads(first_leaf)      brute force with per-construction autograd leaves (torch fp64)
  example_grads(exact)           central finite differences of the loss w.r.t. the table
  adam_update                    closed form of the first update (m=v=0)
  cos_decay_lr                   closed form at step 1, warm-up end, warm-up+1, T
  attention_forward              reference KATs attention.rs:316-425
  attention_backward             central finite differences, torch fp64 autograd
  attention_example_grads        central finite differences of the example loss w.r.t. the
                                 table (exact); torch fp64 autograd with one leaf per
                                 (construction, feature) (first_leaf); item rules checked
                                 on duplicate features and single-feature nodes
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

# --------------------------------------------------------------------------------------
# Graph construction (src/graph.rs)
# --------------------------------------------------------------------------------------


def csr_from_edges(edges: Sequence[Tuple[int, int, float]], deduplicate: bool):
    """CSR::construct_from_edges (src/graph.rs:70-109) with dedup (120-142).

    Returns (rows int64[N+1], columns int64[E], weights float64[E]).  With
    ``deduplicate`` the edge list is first sorted by (from, to) and repeated
    edges are merged by summing weights; without it, columns keep insertion
    order within a row.
    """
    edges = [(int(f), int(t), float(w)) for f, t, w in edges]
    if deduplicate and edges:
        edges.sort(key=lambda e: (e[0], e[1]))
        merged = [list(edges[0])]
        for f, t, w in edges[1:]:
            if f == merged[-1][0] and t == merged[-1][1]:
                merged[-1][2] += w
            else:
                merged.append([f, t, w])
        edges = [tuple(e) for e in merged]
    max_node = max((max(f, t) for f, t, _ in edges), default=0)
    rows = np.zeros(max_node + 2, dtype=np.int64)
    for f, _, _ in edges:
        rows[f + 1] += 1
    rows = np.cumsum(rows)
    counts = np.zeros(max_node + 1, dtype=np.int64)
    cols = np.zeros(len(edges), dtype=np.int64)
    wts = np.zeros(len(edges), dtype=np.float64)
    for f, t, w in edges:
        idx = rows[f] + counts[f]
        cols[idx] = t
        wts[idx] = w
        counts[f] += 1
    return rows, cols, wts


def cdf_convert(weights: np.ndarray) -> np.ndarray:
    """convert_edges_to_cdf (src/graph.rs:664-691) for one row of weights."""
    w = np.asarray(weights, dtype=np.float64).copy()
    if w.size == 0:
        return w
    denom = w.sum()
    if denom == 0.0:
        w[:] = 1.0
        denom = float(w.size)
    w = np.minimum(np.cumsum(w) / denom, 1.0)
    w[-1] = 1.0
    return w


def cumcsr_convert(rows: np.ndarray, weights: np.ndarray) -> np.ndarray:
    """CumCSR::convert (src/graph.rs:414-422): per-row CDF of the weights."""
    out = np.asarray(weights, dtype=np.float64).copy()
    for i in range(len(rows) - 1):
        s, e = rows[i], rows[i + 1]
        if s < e:
            out[s:e] = cdf_convert(out[s:e])
    return out


def cdf_to_p(cdf: np.ndarray) -> np.ndarray:
    """CDFtoP (src/graph.rs:630-661)."""
    cdf = np.asarray(cdf, dtype=np.float64)
    if cdf.size == 0:
        return cdf
    return np.concatenate([cdf[:1], np.diff(cdf)])


# --------------------------------------------------------------------------------------
# Sampling specification (src/algos/utils.rs:166-216)
# --------------------------------------------------------------------------------------


def sample_spec(kind: str, k: float, n: int) -> Tuple[int, float]:
    """Deterministic part of ``Sample::sample`` for All and Fixed(k).

    Returns (number to draw, unbiasing scale).  ``Fixed(k)``: (min(k, n), n/r).
    ``All``: (n, 1).  (Probability(p) draws a Binomial and is not restated.)
    """
    if kind == "all":
        return n, 1.0
    if kind == "fixed":
        r = min(int(k), n)
        return r, (float(n) / float(r)) if r > 0 else float("nan")
    raise ValueError(kind)


def probability_sample(p: float, n: int, k: int) -> Tuple[int, float]:
    """``Sample::Probability(p).sample(n, at_least_one=true)`` given the Binomial draw k
    (utils.rs:203-210), in f32 as the reference: (k, 1/p), or (1, (1 - p^n)/p) when k = 0."""
    p32 = np.float32(p)
    if k == 0:
        return 1, float((np.float32(1.0) - np.float32(p32 ** np.float32(n))) / p32)
    return k, float(np.float32(1.0) / p32)


# --------------------------------------------------------------------------------------
# Constructions (src/algos/ep/model.rs)
# --------------------------------------------------------------------------------------


@dataclass
class Construction:
    """One node construction: ``feature -> accumulated count`` in first-seen order.

    Mirrors the ``NodeCounts`` map built by ``collect_embeddings_from_node``
    (model.rs:257-278): the first occurrence of a feature creates its leaf with
    the occurrence weight; later occurrences add their weight to the count.
    """

    counts: Dict[int, float] = field(default_factory=dict)

    def add(self, feat: int, weight: float):
        if feat in self.counts:
            self.counts[feat] += weight
        else:
            self.counts[feat] = weight

    @property
    def total(self) -> float:
        return float(sum(self.counts.values()))


def collect_node(con: Construction, feats: Sequence[int], weight: float, feat_scale: float = 1.0):
    """collect_embeddings_from_node (model.rs:257-278) with ``max_features`` = All.

    ``weight += scalar`` (model.rs:268) is additive: each occurrence adds
    ``weight + feat_scale`` to its feature's count (feat_scale = 1 for All).
    """
    w = weight + feat_scale
    for f in feats:
        con.add(int(f), w)


def node_features(f_indptr, f_ids, node) -> np.ndarray:
    return np.asarray(f_ids[int(f_indptr[node]):int(f_indptr[node + 1])], dtype=np.int64)


def _selected(f_indptr, f_ids, u, sel):
    """The node's features taken by max_features (sel = (positions, scale)) or all of them."""
    feats = node_features(f_indptr, f_ids, u)
    if sel is None:
        return feats, 1.0
    pos, scale = sel
    return feats[np.asarray(sorted(pos), dtype=np.int64)], float(scale)


def build_anchor_construction(f_indptr, f_ids, v, sel=None) -> Construction:
    """h(v): construct_node_embedding (model.rs:283-301) called with weight 1 (mod.rs:300-301).
    ``sel`` = (positions, scale): the features max_features took and its scale."""
    c = Construction()
    feats, sc = _selected(f_indptr, f_ids, v, sel)
    collect_node(c, feats, 1.0, sc)
    return c


def build_neighbour_construction(f_indptr, f_ids, nbrs: Sequence[int], scale_weight, sels=None) -> Construction:
    """~h(v): reconstruct_node_embedding (model.rs:342-386) -> construct_from_multiple_nodes (388-416).

    ``nbrs`` is the explicit sampled neighbour list.  ``scale_weight`` is each
    neighbour's weight (model.rs:357-361): one number for every neighbour (the
    unbiasing scale of ``Sample::sample``, deg/S when subsampled, 1 otherwise), or
    one weight per neighbour (weighted averaging, see ``neighbour_weights``).  ``sels[j]``
    = (positions, scale) are the features max_features took from neighbour j.
    """
    c = Construction()
    ws = [float(scale_weight)] * len(nbrs) if np.isscalar(scale_weight) else [float(x) for x in scale_weight]
    for j, (u, w) in enumerate(zip(nbrs, ws)):
        feats, sc = _selected(f_indptr, f_ids, u, None if sels is None else sels[j])
        collect_node(c, feats, w, sc)
    return c


def neighbour_weights(cdf_row: np.ndarray, picked: Sequence[int], n_sample: int, weighted_averaging: bool):
    """Weights of the picked edges of one row (model.rs:355-361), in f32 as the reference.

    ``Sample::sample`` gives (r, deg/r) (deg/deg = 1 when every neighbour is taken); with
    weighted averaging neighbour j weighs scale * p_j * deg, p_j = CDFtoP(cdf)[j]
    (graph.rs:630-661), else scale.  ``picked`` are positions in the row.
    """
    deg = len(cdf_row)
    r = min(n_sample, deg)
    scale = np.float32(deg) / np.float32(r)
    if not weighted_averaging:
        return [float(scale)] * len(picked)
    p = cdf_to_p(np.asarray(cdf_row, dtype=np.float32)).astype(np.float32)
    return [float(scale * p[j] * np.float32(deg)) for j in picked]


def construct_mean(table: np.ndarray, con: Construction) -> np.ndarray:
    """mean_embeddings (model.rs:448-463): sum_f c_f E_f / sum_f c_f."""
    acc = np.zeros(table.shape[1], dtype=np.float64)
    n = 0.0
    for f, c in con.counts.items():
        acc += c * table[f].astype(np.float64)
        n += c
    return acc / n


# --------------------------------------------------------------------------------------
# Margin loss (src/algos/ep/loss.rs:59-73, 241-243)
# --------------------------------------------------------------------------------------


def euclidean(x: np.ndarray, y: np.ndarray) -> float:
    """euclidean_distance (loss.rs:241-243): sqrt(sum (x-y)^2)."""
    d = np.asarray(x, dtype=np.float64) - np.asarray(y, dtype=np.float64)
    return float(np.sqrt(np.sum(d * d)))


def margin_loss_adjoints(a: np.ndarray, p: np.ndarray, negs: Sequence[np.ndarray], gamma: float):
    """Margin loss and its adjoints.

    a = h(v), p = ~h(v), n_i = h(u_i).  D0 = |p-a|, D_i = |p-n_i|,
    l_i = gamma + D0 - D_i, A = {i: l_i > 0}; L = mean_{i in A} l_i (0 if A empty).
    Returns (L, active mask, g_a, g_p, [g_n_i]) with g = dL/dx; inactive
    negatives get None (they are not in the loss graph).  Division by a zero
    distance yields NaN exactly as the reference's tape does.
    """
    a = np.asarray(a, dtype=np.float64)
    p = np.asarray(p, dtype=np.float64)
    D0 = euclidean(p, a)
    Ds = [euclidean(p, n) for n in negs]
    ls = [gamma + D0 - d for d in Ds]
    active = np.array([l > 0.0 for l in ls], dtype=bool)
    nA = int(active.sum())
    if nA == 0:
        return 0.0, active, None, None, [None] * len(negs)
    L = float(sum(l for l, act in zip(ls, active) if act) / nA)
    with np.errstate(divide="ignore", invalid="ignore"):
        g_a = (a - p) / D0
        g_p = (p - a) / D0
        g_n = []
        for n, d, act in zip(negs, Ds, active):
            if act:
                gn = (p - np.asarray(n, dtype=np.float64)) / d / nA
                g_p = g_p - gn
                g_n.append(gn)
            else:
                g_n.append(None)
    return L, active, g_a, g_p, g_n


def _unit(x: np.ndarray):
    n = float(np.sqrt(np.sum(x * x)))
    return x / n, n


def _unit_adjoint(x: np.ndarray, g_unit: np.ndarray) -> np.ndarray:
    """Adjoint through x -> x/|x| (il2norm, loss.rs:228-234): (g - u (u.g)) / |x|."""
    u, n = _unit(x)
    return (g_unit - u * float(u @ g_unit)) / n


def loss_adjoints(kind: str, a: np.ndarray, p: np.ndarray, negs: Sequence[np.ndarray], params: Sequence[float]):
    """Loss::compute (loss.rs:56-151) and its adjoints, float64.

    a = h(v), p = ~h(v) (thv), n_i = h(u_i).  Returns (L, reached, g_a, g_p, [g_n_i]):
    ``reached`` marks the constructions inside the loss graph (the others have no
    gradient, like an unreached autodiff leaf); a gradient is None when not reached.
      margin / ppr (gamma)     mean over active i of gamma + |p-a| - |p-n_i|
      starspace (gamma)        mean over active i of gamma - (p^.a^ - a^.n_i^)   (x^ = x/|x|)
      contrastive (pm, nm)     mean over the positive terms of max(0, pm - p^.a^) and
                               max(0, a^.n_i^ - nm)
      rank (tau)               -ln s_pos when s_pos < tau, s = softmax(n_1.a, .., n_K.a, a.p)
                               (exact exp; the reference's exp_approx is a series kept in
                               simple_grad)
      rankspace (gamma)        starspace(gamma) + rank(gamma)
    """
    a = np.asarray(a, dtype=np.float64)
    p = np.asarray(p, dtype=np.float64)
    negs = [np.asarray(n, dtype=np.float64) for n in negs]
    K = len(negs)
    if kind in ("margin", "ppr"):
        L, active, g_a, g_p, g_n = margin_loss_adjoints(a, p, negs, params[0])
        reached = [L > 0, L > 0] + [bool(x) and L > 0 for x in active]
        return L, reached, g_a, g_p, g_n
    ga = np.zeros_like(a)
    gp = np.zeros_like(p)
    gn = [np.zeros_like(n) for n in negs]
    reached = [False] * (K + 2)
    L = 0.0
    if kind in ("starspace", "rankspace"):
        gamma = params[0]
        au, _ = _unit(a)
        pu, _ = _unit(p)
        nus = [_unit(n)[0] for n in negs]
        rd = float(pu @ au)
        ls = [gamma - (rd - float(au @ nu)) for nu in nus]
        act = [l > 0 for l in ls]
        nA = sum(act)
        if nA:
            L += sum(l for l, x in zip(ls, act) if x) / nA
            g_au = sum((nus[i] - pu) for i in range(K) if act[i]) / nA
            g_pu = -au * (sum(act) / nA)
            ga += _unit_adjoint(a, g_au)
            gp += _unit_adjoint(p, g_pu)
            reached[0] = reached[1] = True
            for i in range(K):
                if act[i]:
                    gn[i] += _unit_adjoint(negs[i], au / nA)
                    reached[2 + i] = True
    if kind == "contrastive":
        pm, nm = params[0], params[1]
        au, _ = _unit(a)
        pu, _ = _unit(p)
        nus = [_unit(n)[0] for n in negs]
        pos = pm - float(pu @ au)
        negt = [float(au @ nu) - nm for nu in nus]
        terms = [t for t in negt if t > 0] + ([pos] if pos > 0 else [])
        if terms:
            m = len(terms)
            L = sum(terms) / m
            g_au = np.zeros_like(a)
            if pos > 0:
                g_au -= pu / m
                gp += _unit_adjoint(p, -au / m)
                reached[1] = True
            for i in range(K):
                if negt[i] > 0:
                    g_au += nus[i] / m
                    gn[i] += _unit_adjoint(negs[i], au / m)
                    reached[2 + i] = True
            ga += _unit_adjoint(a, g_au)
            reached[0] = True
    if kind in ("rank", "rankspace"):
        tau = params[0]
        ds = np.array([float(n @ a) for n in negs] + [float(a @ p)])
        z = np.exp(ds - ds.max())
        sm = z / z.sum()
        if sm[-1] < tau:
            L += -math.log(sm[-1])
            ga += sum(sm[i] * negs[i] for i in range(K)) + (sm[-1] - 1.0) * p
            gp += (sm[-1] - 1.0) * a
            for i in range(K):
                gn[i] += sm[i] * a
            reached = [True] * (K + 2)
    if not (L > 0):
        reached = [False] * (K + 2)
    return (L, reached, ga if reached[0] else None, gp if reached[1] else None,
            [g if r else None for g, r in zip(gn, reached[2:])])


def loss_weight_divisor(kind: str, degree: int, weight: float = 0.0) -> float:
    """LossWeighting (mod.rs:190-200): loss is divided by this value."""
    if kind == "none":
        return 1.0
    if kind == "log":
        return math.log(1.0 + degree)
    if kind == "exponential":
        return float(degree) ** weight
    raise ValueError(kind)


# --------------------------------------------------------------------------------------
# One example and one batch (src/algos/ep/mod.rs:174-233, 289-378)
# --------------------------------------------------------------------------------------


@dataclass
class ExampleResult:
    loss: float                      # weighted loss (0 when no gradient)
    contributes: bool                # loss > 0 and not NaN (mod.rs:206-214)
    grads: Dict[int, np.ndarray]     # feature -> gradient (after first_leaf/exact rule)
    vectors: List[np.ndarray]        # [a, p, n_1..n_K]
    adjoints: List[Optional[np.ndarray]]


def example_grads(table, f_indptr, f_ids, anchor, nbrs, nbr_scale, negs, gamma,
                  grad_mode="first_leaf", loss_div=1.0, loss=None, feat_sel=None) -> ExampleResult:
    """Forward + adjoint + gradient extraction for one anchor with explicit samples.

    Constructions are ordered a=h(v), p=~h(v), n_1..n_K (mod.rs:300-320).
    A construction is *eligible* for gradient extraction when it is part of
    the loss graph (a and p always, negatives only when active) and its
    adjoint is finite (mod.rs:370-375 drops NaN/Inf feature gradients).  The
    gradient of feature f through construction x is (c_f/C_x) dL/dx.

    grad_mode:
      "first_leaf"  the reference rule (mod.rs:347-352, 366-367): the first
                    eligible construction containing f, in the order above.
      "exact"       the true gradient: sum over every eligible construction.

    feat_sel (max_features sampling): one (positions, scale) per node in the order
    v, nbrs, negs; None takes every feature.
    """
    fs = [None] * (1 + len(nbrs) + len(negs)) if feat_sel is None else list(feat_sel)
    cons = [build_anchor_construction(f_indptr, f_ids, anchor, fs[0]),
            build_neighbour_construction(f_indptr, f_ids, nbrs, nbr_scale, fs[1:1 + len(nbrs)])]
    for k, u in enumerate(negs):
        cons.append(build_anchor_construction(f_indptr, f_ids, u, fs[1 + len(nbrs) + k]))
    vecs = [construct_mean(table, c) for c in cons]
    kind, params = loss if loss is not None else ("margin", (gamma,))
    L, _, g_a, g_p, g_n = loss_adjoints(kind, vecs[0], vecs[1], vecs[2:], params)
    L = L / loss_div
    adjs = [g_a, g_p] + list(g_n)
    adjs = [None if g is None else g / loss_div for g in adjs]
    if not (L > 0.0) or math.isnan(L):
        return ExampleResult(L if not math.isnan(L) else float("nan"), False, {}, vecs, adjs)
    grads: Dict[int, np.ndarray] = {}
    for con, adj in zip(cons, adjs):
        if adj is None or not np.all(np.isfinite(adj)):
            continue
        C = con.total
        for f, c in con.counts.items():
            g = (c / C) * adj
            if not np.all(np.isfinite(g)):
                continue
            if grad_mode == "first_leaf":
                if f in grads:
                    continue
                grads[f] = g
            elif grad_mode == "exact":
                grads[f] = grads[f] + g if f in grads else g
            else:
                raise ValueError(grad_mode)
    return ExampleResult(L, True, grads, vecs, adjs)


def merge_batch(results: Sequence[ExampleResult], dims: int) -> Dict[int, np.ndarray]:
    """Batch merge (mod.rs:223-233): sum of per-example gradients per feature."""
    out: Dict[int, np.ndarray] = {}
    for r in results:
        if not r.contributes:
            continue
        for f, g in r.grads.items():
            out[f] = out[f] + g if f in out else g.copy()
    return out


# --------------------------------------------------------------------------------------
# Schedule and optimizer
# --------------------------------------------------------------------------------------


def cos_decay_lr(step: int, alpha: float, total_updates: int) -> float:
    """LRScheduler::cos_decay as built at mod.rs:132-137 (scheduler.rs:39-47).

    warm = floor(T/5), min = alpha/100; step <= warm: linear warm-up from min;
    otherwise min + (alpha-min)/2 (1 + cos(pi step/T)).  Steps are numbered from 1.
    """
    warm = int(total_updates / 5.0)
    lo = alpha / 100.0
    if step > warm:
        return lo + 0.5 * (alpha - lo) * (1.0 + math.cos(math.pi * step / total_updates))
    return lo + (step / warm) * (alpha - lo) if warm > 0 else float("nan")


def noise_sd(step: int, noise: float, total_updates: int) -> float:
    """Gradient-noise standard deviation at `step` (mod.rs:139-147 + scheduler.rs:49-51):
    for noise > 1e-9, exp_decay(min 0, noise, decay) with decay = exp((ln 1e-9 - ln noise) / T),
    i.e. noise * decay^step, which reaches 1e-9 at step T; otherwise the no-op scheduler (0).
    """
    if noise <= 1e-9:
        return 0.0
    decay = math.exp((math.log(1e-9) - math.log(noise)) / total_updates)
    return max(noise * decay ** step, 0.0)


ADAM_B1 = 0.9
ADAM_B2 = 0.999
ADAM_EPS = 1e-8


def adam_update(table, mom, var, grads: Dict[int, np.ndarray], lr: float, pass_: int):
    """AdamOptimizer::update (optimizer.rs:43-81), in place, float64.

    Bias-correction time step t = pass + 1 (pass is 1-based, optimizer.rs:50).
    A row whose gradient has any NaN is skipped (optimizer.rs:56).
    """
    t = float(pass_) + 1.0
    for f, g in grads.items():
        if np.any(np.isnan(g)):
            continue
        mom[f] = ADAM_B1 * mom[f] + (1.0 - ADAM_B1) * g
        var[f] = ADAM_B2 * var[f] + (1.0 - ADAM_B2) * g * g
        mh = mom[f] / (1.0 - ADAM_B1 ** t)
        vh = var[f] / (1.0 - ADAM_B2 ** t)
        table[f] = table[f] - lr * (mh / (np.sqrt(vh) + ADAM_EPS))


def _example(table, f_indptr, f_ids, v, nbrs, nbr_scale, negs, gamma, grad_mode, div, attention, loss=None,
             feat_sel=None):
    if attention is None:
        return example_grads(table, f_indptr, f_ids, v, nbrs, nbr_scale, negs, gamma, grad_mode, div, loss, feat_sel)
    heads, d_k = attention[:2]
    window = attention[2] if len(attention) > 2 else None  # (heads, d_k, window): Sliding attention
    rand_blocks = attention[3] if len(attention) > 3 else None  # Random attention: per-block masks
    return attention_example_grads(table, f_indptr, f_ids, v, nbrs, negs, gamma, heads, d_k, grad_mode, div, loss,
                                   window, rand_blocks)


def ep_batch_step(table, mom, var, f_indptr, f_ids, anchors, nbr_lists, nbr_scales, neg_lists,
                  gamma, lr, pass_, grad_mode="first_leaf", loss_divs=None, attention=None, loss=None,
                  feat_sels=None):
    """One batch: per-example forward/backward, merge, Adam (mod.rs:174-252).

    Updates table/mom/var in place (float64 arrays).  Returns (results, merged grads).
    The Adam update runs only when at least one example contributes (mod.rs:237).
    attention = (heads, d_k) selects the attention model (attention_example_grads);
    loss = (kind, params) another loss than margin(gamma) (loss_adjoints).
    """
    results = []
    for i, v in enumerate(anchors):
        div = 1.0 if loss_divs is None else loss_divs[i]
        att = attention[i] if isinstance(attention, list) else attention  # per example (Random attention)
        results.append(_example(table, f_indptr, f_ids, v, nbr_lists[i], nbr_scales[i], neg_lists[i], gamma,
                                grad_mode, div, att, loss, None if feat_sels is None else feat_sels[i]))
    merged = merge_batch(results, table.shape[1])
    if any(r.contributes for r in results):
        adam_update(table, mom, var, merged, lr, pass_)
    return results, merged


def ep_group_step(table, mom, var, f_indptr, f_ids, anchors, nbr_lists, nbr_scales, neg_lists,
                  gamma, lrs, batch_of, pass_, grad_mode="first_leaf", loss_divs=None, attention=None):
    """A group of concurrently running batches (mod.rs:174: par_iter over the batch
    chunks, Hogwild updates of the shared store), serialised: every batch of the group
    computes its gradients from the parameters as they were at the group's start, then
    the batches' updates are applied in batch order, batch k with lrs[k] (its own step's
    learning rate, mod.rs:234-248).  batch_of[i] is example i's batch (0 .. len(lrs)-1);
    an int B means consecutive batches of B examples.  Updates table/mom/var in place;
    returns the per-example results and the list of per-batch merged gradients.
    """
    if isinstance(batch_of, int):
        batch_of = [i // batch_of for i in range(len(anchors))]
    snap = table.copy()
    results = []
    for i, v in enumerate(anchors):
        div = 1.0 if loss_divs is None else loss_divs[i]
        results.append(_example(snap, f_indptr, f_ids, v, nbr_lists[i], nbr_scales[i], neg_lists[i], gamma,
                                grad_mode, div, attention))
    merged = []
    for k in range(len(lrs)):
        part = [r for r, b in zip(results, batch_of) if b == k]
        mk = merge_batch(part, table.shape[1])
        if any(r.contributes for r in part):
            adam_update(table, mom, var, mk, lrs[k], pass_)
        merged.append(mk)
    return results, merged


def batch_error(results: Sequence[ExampleResult]) -> float:
    """Per-batch train error (mod.rs:227-260): sum of contributing losses / batch size."""
    cnt = sum(1 for r in results if r.contributes)
    if cnt == 0:
        return 0.0
    return sum(r.loss for r in results if r.contributes) / len(results)


# --------------------------------------------------------------------------------------
# Initialization (mod.rs:381-392)
# --------------------------------------------------------------------------------------


def normalize_rows(raw_uniform01: np.ndarray) -> np.ndarray:
    """x = 2u - 1 then divide each row by its L2 norm (mod.rs:385-390)."""
    x = 2.0 * np.asarray(raw_uniform01, dtype=np.float64) - 1.0
    return x / np.linalg.norm(x, axis=1, keepdims=True)


# --------------------------------------------------------------------------------------
# Attention aggregator (src/algos/ep/attention.rs, model.rs:306-335, 419-446)
# --------------------------------------------------------------------------------------


def attention_split(rows: np.ndarray, heads: int, d_k: int):
    """Row layout [Q_1..Q_H | K_1..K_H | V_1..V_H] (attention.rs:38-66)."""
    d_model = (rows.shape[1] - 2 * heads * d_k) // heads
    Q = [rows[:, h * d_k:(h + 1) * d_k] for h in range(heads)]
    K = [rows[:, heads * d_k + h * d_k: heads * d_k + (h + 1) * d_k] for h in range(heads)]
    base = 2 * heads * d_k
    V = [rows[:, base + h * d_model: base + (h + 1) * d_model] for h in range(heads)]
    return Q, K, V, d_model


def softmax_exact(x: np.ndarray) -> np.ndarray:
    """softmax (attention.rs:283-296) with the exact exponential."""
    z = np.exp(x - np.max(x))
    return z / z.sum()


def attention_matrix_full(Q: np.ndarray, K: np.ndarray, counts: np.ndarray) -> np.ndarray:
    """compute_full_attention_matrix (attention.rs:171-191): S_ij = Q_i.K_j * (c_i c_j if >= 1)."""
    S = Q @ K.T
    num = np.outer(counts, counts)
    return np.where(num >= 1.0, S * num, S)


def attention_window_mask(n: int, window: Optional[int]) -> np.ndarray:
    """Entries (i, j) that exist in the attention matrix: all of them (Full), or
    |i - j| <= window_size (compute_sliding_attention_matrix, attention.rs:194-218)."""
    if window is None:
        return np.ones((n, n), dtype=bool)
    idx = np.arange(n)
    return np.abs(idx[:, None] - idx[None, :]) <= int(window)


def attention_scores(Q: np.ndarray, K: np.ndarray, counts: np.ndarray, window: Optional[int]):
    """Scores before the 1/sqrt(d_k): Full scales by the counts (attention.rs:171-191),
    Sliding does not (194-218).  Returns (S, mask, m) with m the score scale."""
    n = Q.shape[0]
    if window is None:
        num = np.outer(counts, counts)
        m = np.where(num >= 1.0, num, 1.0)
    else:
        m = np.ones((n, n))
    return (Q @ K.T) * m, attention_window_mask(n, window), m


def masked_softmax_rows(S: np.ndarray, mask: np.ndarray) -> np.ndarray:
    """compute_attention_softmax (attention.rs:253-280): each row's softmax over its
    existing entries; missing entries are 0."""
    A = np.zeros_like(S)
    for i in range(S.shape[0]):
        A[i, mask[i]] = softmax_exact(S[i, mask[i]])
    return A


def attention_mean(rows: np.ndarray, counts: np.ndarray, heads: int, d_k: int,
                   window: Optional[int] = None, rand=None) -> np.ndarray:
    """attention_mean (attention.rs:100-131), float64; Full attention, or Sliding with
    window_size ``window``.

    rows: n x W feature rows, counts: n multiplicities.  For n == 1 the value
    vector of head 0 is returned (attention.rs:113-115).  Otherwise, per head,
    A = softmax_row(S / sqrt(d_k)), out_h = (1/n) sum_i sum_j A_ij V_j, and
    the result is the mean over heads.

    rand: Random attention (compute_random_attention_matrix, attention.rs:221-250), one
    (mask, scale) per head: row i attends to the columns mask[i], every item's q, k and v
    are scaled by `scale` (so the scores by scale^2), no count scaling.
    """
    rows = np.asarray(rows, dtype=np.float64)
    Q, K, V, d_model = attention_split(rows, heads, d_k)
    n = rows.shape[0]
    if n == 1:
        return V[0][0].copy()
    out = np.zeros(d_model)
    for h in range(heads):
        S, mask, _ = attention_scores(Q[h], K[h], np.asarray(counts, dtype=np.float64), window)
        sc = 1.0
        if rand is not None:
            mask, sc = np.asarray(rand[h][0], dtype=bool), float(rand[h][1])
            S = (Q[h] @ K[h].T) * sc * sc
        A = masked_softmax_rows(S / math.sqrt(d_k), mask)
        out += (A.sum(axis=0) @ V[h]) * sc / n
    return out / heads


def attention_mean_backward(rows: np.ndarray, counts: np.ndarray, heads: int, d_k: int,
                            g_out: np.ndarray, window: Optional[int] = None, rand=None) -> np.ndarray:
    """Adjoint of ``attention_mean`` w.r.t. the feature rows (float64).

    With w_j = sum_i A_ij:  dV_j = w_j g / (n H);
    dA_ij = (V_j . g)/(n H);  dS_ij = A_ij (dA_ij - sum_k A_ik dA_ik);
    dT = dS * m / sqrt(d_k) with m_ij = c_i c_j if >= 1 else 1;
    dQ_i = sum_j dT_ij K_j;  dK_j = sum_i dT_ij Q_i.
    """
    rows = np.asarray(rows, dtype=np.float64)
    counts = np.asarray(counts, dtype=np.float64)
    Q, K, V, d_model = attention_split(rows, heads, d_k)
    n = rows.shape[0]
    grad = np.zeros_like(rows)
    base = 2 * heads * d_k
    if n == 1:
        grad[0, base:base + d_model] = g_out
        return grad
    for h in range(heads):
        S, mask, m = attention_scores(Q[h], K[h], counts, window)
        sc = 1.0
        if rand is not None:  # Random attention: q, k, v scaled by sc, no count scaling
            mask, sc = np.asarray(rand[h][0], dtype=bool), float(rand[h][1])
            m = np.full((n, n), sc * sc)
            S = (Q[h] @ K[h].T) * m
        A = masked_softmax_rows(S / math.sqrt(d_k), mask)
        w = A.sum(axis=0)
        scale = 1.0 / (n * heads)
        grad[:, base + h * d_model: base + (h + 1) * d_model] += np.outer(w, g_out) * scale * sc
        dA = np.tile((V[h] @ g_out) * scale * sc, (n, 1))
        dS = A * (dA - np.sum(A * dA, axis=1, keepdims=True))
        dT = dS * m / math.sqrt(d_k)
        grad[:, h * d_k:(h + 1) * d_k] += dT @ K[h]
        grad[:, heads * d_k + h * d_k: heads * d_k + (h + 1) * d_k] += dT.T @ Q[h]
    return grad


def attention_node_items(f_indptr, f_ids, node, sliding: bool = False):
    """Items of h(node) under the attention model.

    attention_construct_node_embedding (model.rs:306-335) collects the node's features
    with weight 1 (each occurrence adds 1 + 1, model.rs:267-268) into a map and runs
    attention over the map's entries: one item per distinct feature, count = 2 x its
    multiplicity.  Items are in first-seen order; Full attention is permutation
    invariant, so the map's iteration order does not change the result.  Sliding
    attention keeps the feature order (model.rs:320-329): one item per listed
    occurrence (a repeated feature is repeated), each reading its feature's entry.
    """
    con = Construction()
    feats_listed = node_features(f_indptr, f_ids, node)
    collect_node(con, feats_listed, 1.0)
    feats = [int(f) for f in feats_listed] if sliding else list(con.counts.keys())
    return feats, [con.counts[f] for f in feats]


def attention_neighbour_items(f_indptr, f_ids, node):
    """Items of one neighbour's attention inside ~h(v) (attention_multiple, model.rs:419-446).

    The neighbour's feature list as listed (a repeated feature is repeated), each
    occurrence with the feature's multiplicity within that neighbour as its count.
    """
    feats = [int(f) for f in node_features(f_indptr, f_ids, node)]
    mult: Dict[int, int] = {}
    for f in feats:
        mult[f] = mult.get(f, 0) + 1
    return feats, [float(mult[f]) for f in feats]


def attention_inference_items(feats):
    """Items of AttentionAggregator::construct (aggregator.rs:137-153): every listed
    feature with count 1."""
    return [int(f) for f in feats], [1.0] * len(feats)


def _scatter_rows(out: Dict[int, np.ndarray], feats, rows: np.ndarray):
    for f, r in zip(feats, rows):
        out[f] = out[f] + r if f in out else r.copy()


def attention_example_grads(table, f_indptr, f_ids, anchor, nbrs, negs, gamma, heads, d_k,
                            grad_mode="first_leaf", loss_div=1.0, loss=None, window=None,
                            rand_blocks=None) -> ExampleResult:
    """Forward + adjoint + gradient extraction for one anchor under the attention model.

    AttentionFeatureModel (model.rs:158-249):
      a   = h(v)   attention over the anchor's distinct features (attention_node_items)
      p   = ~h(v)  mean over the sampled neighbours of each neighbour's attention
                   (attention_neighbour_items; construct_from_multiple_nodes with mha,
                   model.rs:388-416, then mean_embeddings with count 1 per neighbour)
      n_k = h(u_k) as a
    Margin loss and weighting as the mean model.  Each construction has one autodiff
    leaf per distinct feature (the NodeCounts map); inside ~h(v) every occurrence of a
    feature in every neighbour reads the same leaf, so the leaf's gradient sums them.
    Leaf gradients are full table rows [Q | K | V] of width heads * (2 d_k + d_model).
    The first_leaf / exact rules and the eligibility of constructions are those of
    example_grads.  window: Sliding attention with that window_size (None: Full).
    rand_blocks: Random attention, per block (0 = a, 1.. the neighbours, then the
    negatives) the per-head (mask, scale) of attention_mean.
    """
    table = np.asarray(table)
    W = table.shape[1]
    sl = window is not None
    node_sets = [attention_node_items(f_indptr, f_ids, anchor, sl)]
    nb_sets = [attention_neighbour_items(f_indptr, f_ids, u) for u in nbrs]
    neg_sets = [attention_node_items(f_indptr, f_ids, u, sl) for u in negs]

    rb = (lambda e: None) if rand_blocks is None else (lambda e: rand_blocks[e])

    def fwd(items, e):
        feats, counts = items
        return attention_mean(table[np.asarray(feats, dtype=np.int64)], np.asarray(counts), heads, d_k, window, rb(e))

    a = fwd(node_sets[0], 0)
    p = sum(fwd(it, 1 + j) for j, it in enumerate(nb_sets)) / float(len(nb_sets))
    nv = [fwd(it, 1 + len(nb_sets) + k) for k, it in enumerate(neg_sets)]
    vecs = [a, p] + nv
    kind, params = loss if loss is not None else ("margin", (gamma,))
    L, _, g_a, g_p, g_n = loss_adjoints(kind, a, p, nv, params)
    L = L / loss_div
    adjs = [g_a, g_p] + list(g_n)
    adjs = [None if g is None else g / loss_div for g in adjs]
    if not (L > 0.0) or math.isnan(L):
        return ExampleResult(L if not math.isnan(L) else float("nan"), False, {}, vecs, adjs)

    def bwd(items, g, e):
        feats, counts = items
        rows = table[np.asarray(feats, dtype=np.int64)]
        return attention_mean_backward(rows, np.asarray(counts), heads, d_k, g, window, rb(e))

    # leaf gradients per construction, in construction order a, p, n_1..n_K
    leaves: List[Optional[Dict[int, np.ndarray]]] = []
    for ci, adj in enumerate(adjs):
        if adj is None or not np.all(np.isfinite(adj)):
            leaves.append(None)
            continue
        out: Dict[int, np.ndarray] = {}
        if ci == 1:
            for j, it in enumerate(nb_sets):
                _scatter_rows(out, it[0], bwd(it, adj / float(len(nb_sets)), 1 + j))
        else:
            it = node_sets[0] if ci == 0 else neg_sets[ci - 2]
            _scatter_rows(out, it[0], bwd(it, adj, 0 if ci == 0 else 1 + len(nb_sets) + ci - 2))
        leaves.append(out)
    grads: Dict[int, np.ndarray] = {}
    for out in leaves:
        if out is None:
            continue
        for f, g in out.items():
            if g.shape != (W,) or not np.all(np.isfinite(g)):
                continue
            if grad_mode == "first_leaf":
                if f not in grads:
                    grads[f] = g
            elif grad_mode == "exact":
                grads[f] = grads[f] + g if f in grads else g
            else:
                raise ValueError(grad_mode)
    return ExampleResult(L, True, grads, vecs, adjs)


def attention_embed(table, feats, heads, d_k, window=None) -> np.ndarray:
    """AttentionAggregator::construct (aggregator.rs:137-153): attention over the listed
    features with count 1 each (float64); Full, or Sliding with ``window``."""
    items = attention_inference_items(feats)
    return attention_mean(np.asarray(table)[np.asarray(items[0], dtype=np.int64)], np.asarray(items[1]), heads, d_k,
                          window)


# --------------------------------------------------------------------------------------
# Inference aggregate and holdout AUC (src/algos/aggregator.rs:34-47)
# --------------------------------------------------------------------------------------


def averaged_node_embeddings(table, f_indptr, f_ids) -> np.ndarray:
    """AvgAggregator::construct (aggregator.rs:34-47): plain mean of the node's feature rows.

    Nodes without features keep a zero row (lib.rs:2116-2118 only writes when
    the node has features).
    """
    n = len(f_indptr) - 1
    out = np.zeros((n, table.shape[1]), dtype=np.float64)
    for i in range(n):
        fs = f_ids[f_indptr[i]:f_indptr[i + 1]]
        if len(fs):
            out[i] = table[np.asarray(fs, dtype=np.int64)].astype(np.float64).mean(axis=0)
    return out


def weighted_node_embeddings(table, f_indptr, f_ids, p_w, alpha) -> np.ndarray:
    """WeightedAggregator::construct (aggregator.rs:98-116), float64: sum_f w_f E_f / sum_f w_f
    with w_f = alpha / (alpha + p_w[f]) over the listed features; featureless nodes zero."""
    n = len(f_indptr) - 1
    out = np.zeros((n, table.shape[1]), dtype=np.float64)
    p_w = np.asarray(p_w, dtype=np.float64)
    for i in range(n):
        fs = np.asarray(f_ids[f_indptr[i]:f_indptr[i + 1]], dtype=np.int64)
        if len(fs):
            w = alpha / (alpha + p_w[fs])
            out[i] = (w[:, None] * table[fs].astype(np.float64)).sum(axis=0) / w.sum()
    return out


def roc_auc(pos_scores: np.ndarray, neg_scores: np.ndarray) -> float:
    """Mann-Whitney ROC-AUC with average ranks for ties."""
    pos = np.asarray(pos_scores, dtype=np.float64)
    neg = np.asarray(neg_scores, dtype=np.float64)
    allv = np.concatenate([pos, neg])
    order = np.argsort(allv, kind="mergesort")
    ranks = np.empty(len(allv), dtype=np.float64)
    sv = allv[order]
    i = 0
    while i < len(sv):
        j = i
        while j + 1 < len(sv) and sv[j + 1] == sv[i]:
            j += 1
        ranks[order[i:j + 1]] = 0.5 * (i + j) + 1.0
        i = j + 1
    rp = ranks[:len(pos)].sum()
    return float((rp - len(pos) * (len(pos) + 1) / 2.0) / (len(pos) * len(neg)))


# --------------------------------------------------------------------------------------
# NeighborhoodAligner (src/algos/alignment.rs:25-75)
# --------------------------------------------------------------------------------------
def neighborhood_align(indptr, indices, cdf, emb, alpha: Optional[float], max_neighbors: Optional[int]):
    """NeighborhoodAligner::align for every node, FP64 (alignment.rs:25-75).

    new_v = (1 - a) * sum_{j < m} p_j E[u_j] + a * E[v], m = min(deg, max_neighbors)
    (the first m edges; p_j = CDFtoP of the row, not renormalised after the cut),
    a = alpha, or min(1 / ln(deg + 1), 1) with the full degree when alpha is None
    (alignment.rs:62-71).  The reference indexes weights[0] of an edgeless node; here
    such a node gets a * E[v].
    """
    from scipy import sparse
    indptr = np.asarray(indptr, dtype=np.int64)
    n = len(indptr) - 1
    emb = np.asarray(emb, dtype=np.float64)
    deg = np.diff(indptr)
    rows, cols, vals = [], [], []
    for v in range(n):
        s, e = int(indptr[v]), int(indptr[v + 1])
        m = e - s if max_neighbors is None else min(e - s, int(max_neighbors))
        p = cdf_to_p(np.asarray(cdf[s:e], dtype=np.float32).astype(np.float64))[:m]
        rows.extend([v] * m)
        cols.extend(int(u) for u in indices[s:s + m])
        vals.extend(p.tolist())
    A = sparse.csr_matrix((vals, (rows, cols)), shape=(n, n))
    if alpha is None:
        with np.errstate(divide="ignore"):
            a = np.minimum(1.0 / np.log(deg.astype(np.float64) + 1.0), 1.0)
    else:
        a = np.full(n, float(alpha))
    return (A @ emb) * (1.0 - a)[:, None] + emb * a[:, None]
