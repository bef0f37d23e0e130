"""GPU NeighborhoodAligner (k_align through ep_aligner_*) against the FP64 oracle
(oracle.ep_oracle.neighborhood_align, alignment.rs:25-75).  Bar: 1e-5 normwise relative."""
import numpy as np
import pytest

from cloverleaf_amd import _native as N
from oracle import ep_oracle as O

pytestmark = pytest.mark.gpu


def random_graph(seed, n, avg_deg, hubs=2, hub_deg=300, isolated=5):
    rng = np.random.default_rng(seed)
    rows = []
    for v in range(n):
        if v >= n - isolated:
            rows.append(np.zeros(0, dtype=np.int64))
            continue
        d = hub_deg if v < hubs else int(rng.integers(1, 2 * avg_deg))
        rows.append(rng.integers(0, n, size=d))
    indptr = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.uint64)
    indices = np.concatenate(rows).astype(np.uint32)
    w = rng.uniform(0.1, 2.0, size=len(indices)).astype(np.float32)
    cdf = np.empty_like(w)
    for v in range(n):
        s, e = int(indptr[v]), int(indptr[v + 1])
        if e > s:
            c = np.cumsum(w[s:e], dtype=np.float32)
            cdf[s:e] = c / c[-1]
    return indptr, indices, cdf


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30))


@pytest.mark.parametrize("dims", [8, 32, 100, 128, 256])
@pytest.mark.parametrize("alpha,max_nb", [(0.3, None), (None, None), (0.0, 7), (None, 64)])
def test_align_matches_oracle(dims, alpha, max_nb):
    n = 3000
    indptr, indices, cdf = random_graph(dims, n, 12)
    emb = np.random.default_rng(dims + 1).normal(size=(n, dims)).astype(np.float32)
    print(f"[align] dims={dims} alpha={alpha} max_nb={max_nb}", flush=True)
    with N.Aligner(indptr, indices, cdf, emb) as al:
        got = al.run(alpha, max_nb)
        part = al.run(alpha, max_nb, first=1234, count=517)
    want = O.neighborhood_align(indptr, indices, cdf, emb, alpha, max_nb)
    assert rel(got, want) <= 1e-5
    assert np.array_equal(part, got[1234:1234 + 517])
    # nodes without edges keep alpha * E[v] (adaptive alpha is 1 there)
    a = 1.0 if alpha is None else alpha
    assert np.allclose(got[-5:], a * emb[-5:], rtol=1e-6)


def test_align_facade_and_to_disk(tmp_path):
    import cloverleaf_amd as C
    n = 500
    indptr, indices, cdf = random_graph(3, n, 6)
    g = C.Graph(indptr, indices, cdf, C.Vocab.anonymous(n))
    emb = C.NodeEmbeddings(C.Vocab.anonymous(n), np.random.default_rng(5).normal(size=(n, 16)).astype(np.float32))
    al = C.NeighborhoodAligner(alpha=0.5, max_neighbors=20)
    out = al.align(emb, g)
    want = O.neighborhood_align(indptr, indices, cdf, emb.embeddings, 0.5, 20)
    assert rel(out.embeddings, want) <= 1e-5
    path = str(tmp_path / "aligned.tsv.gz")
    al.align_to_disk(path, emb, g, chunk_size=37)
    back = C.NodeEmbeddings.load(path)
    assert back.len() == n
    for i in (0, 1, 36, 37, 499):
        assert np.array_equal(np.asarray(back.get_embedding(("node", str(i))), dtype=np.float32), out.embeddings[i])
    with pytest.raises(ValueError):
        al.align(C.NodeEmbeddings(C.Vocab.anonymous(n - 1), np.zeros((n - 1, 16), np.float32)), g)
