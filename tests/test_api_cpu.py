"""Host-side pieces of the facade that need no GPU: unigram probabilities and the
aggregator file formats (src/lib.rs:1855-2000, aggregator.rs:50-79)."""
import numpy as np

import cloverleaf_amd as C
from cloverleaf_amd.api import unigram_probability


def _feature_set():
    g = C.Graph.from_csr(np.array([0, 1, 2, 3], dtype=np.uint64), np.array([1, 2, 0], dtype=np.uint32))
    fs = C.FeatureSet.new_from_graph(g)
    fs.set_features(("node", "0"), [("w", "a"), ("w", "b")])
    fs.set_features(("node", "1"), [("w", "b"), ("w", "c")])
    fs.set_features(("node", "2"), [("w", "b")])
    return fs


def test_unigram_probability_known_answer():
    # counts a:1, b:3, c:1 over 5 occurrences (UnigramProbability::new)
    p = unigram_probability(_feature_set())
    assert p.dtype == np.float32
    assert np.allclose(p, [0.2, 0.6, 0.2])


def test_aggregator_save_load_round_trip(tmp_path):
    fs = _feature_set()
    for agg in (C.FeatureAggregator.Averaged(), C.FeatureAggregator.Attention(4, 32),
                C.FeatureAggregator.Weighted(1e-3, fs)):
        path = tmp_path / "agg.txt"
        agg.save(str(path))
        back = C.FeatureAggregator.load(str(path))
        assert back.kind == agg.kind
        assert (back.num_heads, back.d_k) == (agg.num_heads, agg.d_k)
        if agg.kind == "weighted":
            assert back.alpha == agg.alpha
            assert back.vocab.names == agg.vocab.names
            assert np.allclose(back.unigrams, agg.unigrams)
    text = (tmp_path / "agg.txt").read_text().splitlines()
    assert text[:2] == ["Weighted", "0.001"] and text[2] == "w\ta\t0.20000000298023224"


def test_graph_save_load_round_trip(tmp_path):
    # io.rs:270-323 / lib.rs:414-430: TSV edges with probabilities; Undirected adds reverses
    path = tmp_path / "edges.tsv"
    path.write_text("q\ta\ti\tx\t1\nq\ta\ti\ty\t3\nq\tb\ti\tx\t2\n")
    g = C.Graph.load(str(path), C.EdgeType.Directed)
    assert g.nodes() == 4 and g.edges() == 3
    names, w = g.get_edges(("q", "a"), normalized=True)
    assert names == [("i", "x"), ("i", "y")] and np.allclose(w, [0.25, 0.75])
    u = C.Graph.load(str(path), C.EdgeType.Undirected)
    assert u.edges() == 6
    assert u.get_edges(("i", "x"))[0] == [("q", "a"), ("q", "b")]
    # unweighted loading ignores the weight column
    names, w = C.Graph.load(str(path), C.EdgeType.Directed, weighted=False).get_edges(("q", "a"), normalized=True)
    assert np.allclose(w, [0.5, 0.5])
    for out in (tmp_path / "g.tsv", tmp_path / "g.tsv.gz"):
        g.save(str(out))
        back = C.Graph.load(str(out), C.EdgeType.Directed)
        assert back.vocab_.names[:2] == g.vocab_.names[:2]
        assert np.array_equal(back.indptr[:3], g.indptr[:3])
        assert np.allclose(back.get_edges(("q", "a"), normalized=True)[1], [0.25, 0.75])
    assert (tmp_path / "g.tsv").read_text().splitlines()[0] == "q\ta\ti\tx\t0.25"


def test_graph_load_rejects_malformed(tmp_path):
    path = tmp_path / "bad.tsv"
    path.write_text("q\ta\ti\tx\n")
    try:
        C.Graph.load(str(path), C.EdgeType.Directed)
    except ValueError as e:
        assert "Expected 5 fields" in str(e)
    else:
        raise AssertionError("malformed line accepted")


def test_embeddings_save_load_round_trip(tmp_path):
    v = C.api.Vocab()
    for i in range(4):
        v.get_or_insert("feat" if i % 2 else "node", str(i))
    emb = np.array([[0.1, -2.5, 1e-7], [1.0, 0.0, 3.25], [123456.75, -1e-30, 0.5], [2.0, 2.0, 2.0]], np.float32)
    ne = C.NodeEmbeddings(v, emb, C.Distance.Cosine)
    for name in ("e.tsv", "e.tsv.gz"):
        ne.save(str(tmp_path / name))
        back = C.NodeEmbeddings.load(str(tmp_path / name))
        assert back.vocab_.names == v.names
        assert np.array_equal(back.embeddings, emb)  # shortest round-trip digits are exact
    assert (tmp_path / "e.tsv").read_text().splitlines()[0] == "node\t0\t[0.1,-2.5,1e-7]"  # ryu's layout
    only = C.NodeEmbeddings.load(str(tmp_path / "e.tsv"), filter_type="feat")
    assert only.vocab_.names == [("feat", "1"), ("feat", "3")]
    assert np.array_equal(only.embeddings, emb[[1, 3]])
    dup = tmp_path / "dup.tsv"
    dup.write_text("a\t1\t[1,2]\na\t1\t[3,4]\n")
    try:
        C.NodeEmbeddings.load(str(dup))
    except KeyError:
        pass
    else:
        raise AssertionError("duplicate node accepted")
