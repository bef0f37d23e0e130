"""The native graph file loader (ep_graph_load_tsv, GraphReader::load io.rs:270-323) against
the reference's rules written with the facade's GraphBuilder (CSR::construct_from_edges,
graph.rs:70-142; CumCSR::convert 414-422): node order, columns, CDFs, deduplication,
undirected edges, skipped rows, gzip and the reference's KAT graph."""
import gzip

import numpy as np
import pytest

import cloverleaf_amd as C
from cloverleaf_amd import _native as N


def write_edges(path, rows, gz=False):
    text = "".join("\t".join(map(str, r)) + "\n" for r in rows)
    if gz:
        with gzip.open(path, "wt") as fh:
            fh.write(text)
    else:
        with open(path, "w") as fh:
            fh.write(text)


def builder_graph(rows, undirected, dedup, weighted=True):
    b = C.GraphBuilder()
    for ft, f, tt, t, w in rows:
        b.add_edge((ft, f), (tt, t), float(np.float32(w)) if weighted else 1.0,
                   C.EdgeType.Undirected if undirected else C.EdgeType.Directed)
    return b.build_graph(deduplicate=dedup)


@pytest.mark.parametrize("undirected", [False, True])
@pytest.mark.parametrize("dedup", [False, True])
@pytest.mark.parametrize("gz", [False, True])
def test_native_loader_matches_builder(tmp_path, undirected, dedup, gz):
    rng = np.random.default_rng(int(undirected) * 4 + int(dedup) * 2 + int(gz))
    rows = [("q" if a % 3 else "i", f"n{a}", "i", f"n{b}", f"{w:.4f}")
            for a, b, w in zip(rng.integers(0, 40, 300), rng.integers(0, 40, 300), rng.uniform(0, 3, 300))]
    path = tmp_path / ("g.tsv.gz" if gz else "g.tsv")
    write_edges(path, rows, gz)
    g = C.Graph.load(str(path), C.EdgeType.Undirected if undirected else C.EdgeType.Directed, deduplicate=dedup)
    ref = builder_graph(rows, undirected, dedup)
    assert list(g.vocab()) == list(ref.vocab())
    assert np.array_equal(g.indptr, ref.indptr) and np.array_equal(g.indices, ref.indices)
    assert np.array_equal(g.weights, ref.weights)  # the same f32 arithmetic


def test_native_loader_skip_rows_unweighted_and_errors(tmp_path):
    path = tmp_path / "g.tsv"
    path.write_text("from_type\tfrom\tto_type\tto\tweight\nu\ta\tu\tb\t2\nu\tb\tu\tc\tnot-a-number\n")
    g = C.Graph.load(str(path), C.EdgeType.Directed, skip_rows=1, weighted=False)
    assert g.nodes() == 3 and g.edges() == 2
    with pytest.raises(ValueError, match="Malformed graph file"):
        C.Graph.load(str(path), C.EdgeType.Directed, skip_rows=1)
    bad = tmp_path / "bad.tsv"
    bad.write_text("u\ta\tu\tb\n")
    with pytest.raises(ValueError, match="Expected 5 fields"):
        C.Graph.load(str(bad), C.EdgeType.Directed)
    with pytest.raises(ValueError, match="cannot open"):
        C.Graph.load(str(tmp_path / "missing.tsv"), C.EdgeType.Directed)


def test_native_loader_float_forms_and_crlf(tmp_path):
    # str::parse::<f32>: sign, exponent, inf; CRLF line ends (BufRead::lines)
    path = tmp_path / "g.tsv"
    path.write_bytes(b"u\ta\tu\tb\t+1.5\r\nu\ta\tu\tc\t2e-1\r\nu\ta\tu\td\t0.3\n")
    g = C.Graph.load(str(path), C.EdgeType.Directed)
    names, w = g.get_edges(("u", "a"), normalized=True)
    assert [n for _, n in names] == ["b", "c", "d"]
    tot = np.float32(1.5) + np.float32(0.2) + np.float32(0.3)
    assert np.allclose(w, [1.5 / tot, 0.2 / tot, 0.3 / tot], rtol=1e-6)


def test_reference_csr_kat_through_the_file_loader(tmp_path):
    # graph.rs:707-776: edges (0,1,1), (1,1,3), (1,2,2), (2,0,2.5), (1,0,10) with dedup=false
    edges = [(0, 1, 1.0), (1, 1, 3.0), (1, 2, 2.0), (2, 0, 2.5), (1, 0, 10.0)]
    path = tmp_path / "kat.tsv"
    write_edges(path, [("n", a, "n", b, w) for a, b, w in edges])
    g = C.Graph.load(str(path), C.EdgeType.Directed)
    assert [n for _, n in g.vocab()] == ["0", "1", "2"]
    assert g.indptr.tolist() == [0, 1, 4, 5]
    assert g.indices.tolist() == [1, 1, 2, 0, 0]
    assert np.allclose(g.weights, [1.0, 3 / 15, 5 / 15, 1.0, 1.0], rtol=1e-6)
