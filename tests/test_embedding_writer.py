"""The native embedding file writer (NodeEmbeddings.save, EmbeddingWriter io.rs:24-78).

Values are printed as ryu::Buffer::format prints an f32 (io.rs:16, 69-76): the shortest
digits that read back to the same f32 (the closest when several are as short), laid out
as fixed notation for decimal magnitudes 1e-6 up to below 1e13 (integers with ".0") and
as d[.ddd]e<exp> outside that range.  Checks:

* known answers at each boundary of the layout (1e-6 / 1e-7, 1e12 / 1e13, integers,
  zero and its sign, NaN and infinities);
* the digits against numpy's independent shortest-digit algorithm (Dragon4, unique mode)
  on random f32 bit patterns: the printed decimal equals numpy's exactly and reads back
  to the same f32;
* whole files: write, then the native loader reads back the same names and the same bits,
  plain and gzip.
"""
import gzip
from decimal import Decimal

import numpy as np
import pytest

from cloverleaf_amd import _native as N
from cloverleaf_amd.api import NodeEmbeddings, Vocab

KNOWN = [
    (1e-7, "1e-7"), (1e-6, "0.000001"), (1.5e-5, "0.000015"), (1e-5, "0.00001"),
    (-1.234567e-6, "-0.000001234567"), (0.1, "0.1"), (0.3, "0.3"), (2.0 / 3.0, "0.6666667"),
    (1.0, "1.0"), (-2.5, "-2.5"), (100.0, "100.0"), (123456.7, "123456.7"),
    (16777216.0, "16777216.0"), (1e12, "1000000000000.0"), (9.999999e12, "9999999000000.0"),
    (1e13, "1e13"), (3.4028235e38, "3.4028235e38"), (1.17549435e-38, "1.1754944e-38"),
    (1e-45, "1e-45"), (0.0, "0.0"), (-0.0, "-0.0"),
    (float("nan"), "NaN"), (float("inf"), "inf"), (float("-inf"), "-inf"),
]


@pytest.mark.parametrize("x,text", KNOWN)
def test_known_answers(x, text):
    assert N.format_f32(x) == text


def test_digits_match_an_independent_shortest_algorithm():
    rng = np.random.default_rng(11)
    bits = rng.integers(0, 2 ** 32, size=20000, dtype=np.uint64).astype(np.uint32)
    vals = bits.view(np.float32)
    vals = vals[np.isfinite(vals)]
    # plus values near the layout boundaries
    extra = np.array([1e-6, 9.9999e-7, 1.0000001e-6, 9.999999e12, 1.0000001e13, 0.5, 1.5, 65504.0],
                     dtype=np.float32)
    for x in np.concatenate([vals, extra]):
        ours = N.format_f32(float(x))
        ref = np.format_float_scientific(x, unique=True, trim="-")
        assert Decimal(ours) == Decimal(ref), (x, ours, ref)
        assert np.float32(float(ours)) == x or (x == 0 and float(ours) == 0)


@pytest.mark.parametrize("suffix", [".tsv", ".tsv.gz"])
def test_write_then_load_round_trip(tmp_path, suffix):
    rng = np.random.default_rng(5)
    table = (rng.standard_normal((40, 7)) * 10.0 ** rng.integers(-9, 15, size=(40, 7))).astype(np.float32)
    table[0, :3] = [0.0, -0.0, 1e-45]
    names = [("user" if i % 2 else "item", f"n{i}") for i in range(40)]
    path = str(tmp_path / ("emb" + suffix))
    N.write_embeddings_tsv(path, names, table)
    got_names, got = N.load_embeddings_tsv(path)
    assert got_names == names
    assert np.array_equal(got.view(np.uint32), table.view(np.uint32))
    opener = gzip.open if suffix.endswith(".gz") else open
    with opener(path, "rt") as fh:
        first = fh.readline().rstrip("\n")
    assert first.startswith("item\tn0\t[0.0,-0.0,1e-45,")


def test_facade_save_uses_the_writer(tmp_path):
    names = [("node", "a"), ("node", "b")]
    emb = NodeEmbeddings(Vocab(names), np.array([[1e-7, 2.0], [1e13, -0.5]], dtype=np.float32))
    path = str(tmp_path / "e.tsv")
    emb.save(path)
    with open(path) as fh:
        assert fh.read() == "node\ta\t[1e-7,2.0]\nnode\tb\t[1e13,-0.5]\n"


def test_errors(tmp_path):
    with pytest.raises(ValueError, match="comp_level"):
        N.write_embeddings_tsv(str(tmp_path / "x.gz"), [("t", "n")], np.zeros((1, 2), np.float32), comp_level=12)
    with pytest.raises(ValueError, match="cannot open"):
        N.write_embeddings_tsv(str(tmp_path / "missing" / "x.tsv"), [("t", "n")], np.zeros((1, 2), np.float32))
