"""GPU sampler options through the C ABI, checked against numpy simulations of the
reference's sampling rules: random-walk hard negatives (RandomWalkHardSampler,
node_sampler.rs:83-140)."""
import numpy as np
import pytest

from cloverleaf_amd import _native as N
from tests.problems import Problem

pytestmark = pytest.mark.gpu


def path_problem(n_path=7, n_iso=3, D=8):
    """An undirected path 0-1-...-(n_path-1) plus n_iso isolated nodes; one feature per node."""
    n = n_path + n_iso
    nbrs = [[] for _ in range(n)]
    for a in range(n_path - 1):
        nbrs[a].append(a + 1)
        nbrs[a + 1].append(a)
    indptr = np.concatenate([[0], np.cumsum([len(x) for x in nbrs])]).astype(np.uint64)
    indices = np.array([b for x in nbrs for b in sorted(x)], dtype=np.uint32)
    f_indptr = np.arange(n + 1, dtype=np.uint64)
    f_ids = np.arange(n, dtype=np.uint32)
    return Problem(indptr, indices, f_indptr, f_ids, n, D, np.ones((n, D), np.float32))


def ctx_for(prob, K, B, hard=0, S=4, seed=20222022):
    cfg = N.default_config()
    cfg.dims, cfg.batch_size, cfg.negatives = prob.D, B, K
    cfg.hard_negatives = hard
    cfg.max_nodes_kind, cfg.max_nodes = N.EP_SAMPLE_FIXED, S
    cfg.valid_pct = 0.0
    cfg.seed = seed
    ctx = N.Context(cfg)
    ctx.set_graph(prob.indptr, prob.indices)
    ctx.set_features(prob.f_indptr, prob.f_ids, prob.V)
    ctx.set_table(prob.table)
    return ctx


def simulate_hard_negatives(indptr, indices, anchor, K, H, pool, M, rng, restart=0.25):
    """M negative lists for one anchor by the reference's rules, vectorised over lists."""
    deg = np.diff(indptr.astype(np.int64))
    adj = set(indices[int(indptr[anchor]):int(indptr[anchor + 1])].tolist())
    cands = np.full((M, 2 * H), -1, dtype=np.int64)
    for j in range(2 * H):
        node = np.full(M, anchor, dtype=np.int64)
        live = np.ones(M, dtype=bool)
        for i in range(1, 11):
            live &= deg[node] > 0
            pick = indptr[node].astype(np.int64) + (rng.random(M) * np.maximum(deg[node], 1)).astype(np.int64)
            node = np.where(live, indices[np.minimum(pick, len(indices) - 1)].astype(np.int64), node)
            if i > 1:
                live &= ~((rng.random(M) < restart) & (node != anchor))
        bad = (node == anchor) | np.isin(node, list(adj))
        cands[:, j] = np.where(bad, -1, node)
    out = np.empty((M, K), dtype=np.int64)
    for r in range(M):  # take the first H distinct endpoints in walk order
        taken = []
        for x in cands[r]:
            if x >= 0 and x not in taken:
                taken.append(int(x))
            if len(taken) >= H:
                break
        out[r, :len(taken)] = taken
        out[r, len(taken):] = pool[rng.integers(0, len(pool), K - len(taken))]
    return out


@pytest.mark.parametrize("anchor,H", [(1, 2), (3, 3), (8, 2)])
def test_hard_negatives_match_walk_rules(anchor, H):
    prob = path_problem()
    K, B = 4, 4096
    ctx = ctx_for(prob, K, B, hard=H)
    anchors = np.full(B, anchor, dtype=np.uint32)
    _, _, neg = ctx.sample_batch(anchors, 1)
    ctx.close()
    assert neg.shape == (B, K) and neg.max() < prob.n
    # the walk part: never the anchor or one of its neighbours, and distinct
    adj = set(prob.indices[int(prob.indptr[anchor]):int(prob.indptr[anchor + 1])].tolist()) | {anchor}
    rng = np.random.default_rng(anchor)
    sim = simulate_hard_negatives(prob.indptr, prob.indices, anchor, K, min(H, K), np.arange(prob.n), 40000, rng)
    # frequency of each node in each slot, GPU against the simulation (chi-square)
    chi2, dof = 0.0, 0
    for k in range(K):
        got = np.bincount(neg[:, k], minlength=prob.n).astype(np.float64)
        exp = np.bincount(sim[:, k], minlength=prob.n) / len(sim) * B
        keep = exp > 5
        chi2 += float((((got[keep] - exp[keep]) ** 2) / exp[keep]).sum())
        dof += int(keep.sum()) - 1
        assert got[~keep].sum() <= max(20.0, 4 * exp[~keep].sum())
    print(f"[sampling] hard negatives anchor={anchor} H={H}: chi2 {chi2:.1f} dof {dof}", flush=True)
    assert chi2 < dof + 6 * np.sqrt(2 * dof)
    # lists whose first slot is a walk endpoint keep the rules
    if anchor != 8:
        first_hard = ~np.isin(neg[:, 0], list(adj))
        assert first_hard.mean() > 0.8


def test_hard_negatives_zero_is_uniform_and_reproducible():
    prob = path_problem()
    ctx = ctx_for(prob, 4, 2048, hard=0)
    anchors = np.full(2048, 1, dtype=np.uint32)
    _, _, n0 = ctx.sample_batch(anchors, 1)
    _, _, n1 = ctx.sample_batch(anchors, 1)
    ctx.close()
    assert np.array_equal(n0, n1)
    got = np.bincount(n0.ravel(), minlength=prob.n)
    exp = n0.size / prob.n
    assert float(((got - exp) ** 2 / exp).sum()) < 40  # 9 dof
