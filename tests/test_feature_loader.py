"""The native feature file loader (ep_features_load_tsv) behind FeatureSet.load_into
(src/lib.rs:1579-1606, FeatureStore::set_features src/feature_store.rs:36-48,
FeatureNamespace::parse lib.rs:1391-1403): known answers worked out from those lines, random
files against a restatement of them, Rust's whitespace set, gzip, errors and the facade."""
import gzip

import numpy as np
import pytest

import cloverleaf_amd as C
from cloverleaf_amd import _native as N

# char::is_whitespace (the White_Space property): what str::split_whitespace splits on
RUST_WS = set("\t\n\x0b\x0c\r \x85\xa0\u1680\u2028\u2029\u202f\u205f\u3000") | {chr(c) for c in range(0x2000, 0x200B)}


def rust_split_whitespace(s):
    out, cur = [], []
    for ch in s:
        if ch in RUST_WS:
            if cur:
                out.append("".join(cur))
                cur = []
        else:
            cur.append(ch)
    if cur:
        out.append("".join(cur))
    return out


def restated_load_into(lines, nodes, feats, kind, value):
    """lib.rs:1588-1605 line by line: exactly three tab fields; tokens through the namespace
    (("feat", token) when it gives none); features enter the vocabulary only for known nodes;
    set_features replaces the node's list."""
    node_id = {k: i for i, k in enumerate(nodes)}
    vocab = list(feats)
    fid = {k: i for i, k in enumerate(vocab)}
    lists = {}
    for i, line in enumerate(lines):
        pieces = line.split("\t")
        if len(pieces) != 3:
            raise ValueError(f"Line {i + 1}:Malformed feature line!")
        v = node_id.get((pieces[0], pieces[1]))
        if v is None:
            continue
        ids = []
        for tok in rust_split_whitespace(pieces[2]):
            if kind == N.EP_FNS_STATIC:
                key = (value, tok)
            elif kind == N.EP_FNS_NODE_TYPE:
                key = (pieces[0], tok)
            else:
                j = tok.find(value)
                key = ("feat", tok) if j < 0 else (tok[:j], tok[j + 1:])
            if key not in fid:
                fid[key] = len(vocab)
                vocab.append(key)
            ids.append(fid[key])
        lists[v] = ids
    return lists, vocab[len(feats):]


def as_lists(nodes, ip, ids):
    return {int(v): ids[int(ip[k]):int(ip[k + 1])].tolist() for k, v in enumerate(nodes)}


def write(path, text, gz=False):
    if gz:
        with gzip.open(path, "wb") as fh:
            fh.write(text.encode("utf-8"))
    else:
        path.write_bytes(text.encode("utf-8"))


def test_known_answers(tmp_path):
    p = tmp_path / "f.tsv"
    # line 3 names a node outside the vocabulary: ignored with its tokens; line 4 replaces a's list
    write(p, "n\ta\tx y  z\nn\tb\tq:1 r\nm\tc\tunused\nn\ta\ty w v\n")
    nodes = [("n", "a"), ("n", "b")]
    got = N.load_features_tsv(str(p), nodes, [("feat", "y")], N.EP_FNS_STATIC, "feat")
    assert as_lists(*got[:3]) == {0: [0, 5, 6], 1: [3, 4]}
    assert got[3] == [("feat", "x"), ("feat", "z"), ("feat", "q:1"), ("feat", "r"), ("feat", "w"), ("feat", "v")]
    got = N.load_features_tsv(str(p), nodes, [], N.EP_FNS_PREFIX, ":")
    assert as_lists(*got[:3]) == {0: [1, 5, 6], 1: [3, 4]}
    assert got[3][3] == ("q", "1") and got[3][4] == ("feat", "r")
    got = N.load_features_tsv(str(p), nodes, [], N.EP_FNS_NODE_TYPE, "")
    assert got[3][0] == ("n", "x")


def test_rust_whitespace_set(tmp_path):
    # U+3000, U+00A0 and \x0b separate tokens (char::is_whitespace); \x1c, which Python's
    # str.split also splits on, does not; a line's lone \r stays and separates tokens
    p = tmp_path / "f.tsv"
    write(p, "n\ta\tp\u3000q\xa0r\x0bs\x1ct\ru\n")
    got = N.load_features_tsv(str(p), [("n", "a")], [], N.EP_FNS_STATIC, "feat")
    assert [nm for _, nm in got[3]] == ["p", "q", "r", "s\x1ct", "u"]
    assert "s\x1ct".split() != ["s\x1ct"]  # the case where the two rules differ


@pytest.mark.parametrize("kind,value", [(N.EP_FNS_STATIC, "feat"), (N.EP_FNS_NODE_TYPE, ""), (N.EP_FNS_PREFIX, ":"),
                                        (N.EP_FNS_PREFIX, "_")])
@pytest.mark.parametrize("gz", [False, True])
def test_random_files_match_restatement(tmp_path, kind, value, gz):
    rng = np.random.default_rng(kind * 10 + len(value) + int(gz))
    nodes = [("u" if i % 2 else "i", f"n{i}") for i in range(60)]
    known = [("feat", f"w{i}") for i in range(5)]
    seps = [" ", "  ", "\u3000", "\xa0", "\x0b", "\u2009", "\u0085"]
    lines = []
    for _ in range(400):
        a = int(rng.integers(0, 70))  # some nodes outside the vocabulary
        t = "u" if a % 2 else "i"
        toks = [f"{rng.choice(['a', 'b', 'c'])}{rng.choice([':', '_', ''])}{int(rng.integers(0, 30))}"
                for _ in range(int(rng.integers(0, 6)))]
        body = ""
        for tok in toks:
            body += str(rng.choice(seps)) + tok
        lines.append(f"{t}\tn{a}\t{body}")
    p = tmp_path / ("f.tsv.gz" if gz else "f.tsv")
    write(p, "\n".join(lines) + "\n", gz)
    nodes_set, ip, ids, new = N.load_features_tsv(str(p), nodes, known, kind, value)
    ref_lists, ref_new = restated_load_into(lines, nodes, known, kind, value)
    assert as_lists(nodes_set, ip, ids) == ref_lists
    assert new == ref_new
    assert np.all(np.diff(nodes_set.astype(np.int64)) > 0)


def test_errors(tmp_path):
    p = tmp_path / "f.tsv"
    write(p, "n\ta\tx\n\nn\ta\ty\n")
    with pytest.raises(ValueError, match="Line 2:Malformed feature line"):
        N.load_features_tsv(str(p), [("n", "a")], [], N.EP_FNS_STATIC, "feat")
    write(p, "n\ta\tx\tz\n")
    with pytest.raises(ValueError, match="Line 1:Malformed"):
        N.load_features_tsv(str(p), [("n", "a")], [], N.EP_FNS_STATIC, "feat")
    with pytest.raises(ValueError, match="cannot open"):
        N.load_features_tsv(str(tmp_path / "missing.tsv"), [("n", "a")], [], N.EP_FNS_STATIC, "feat")


def small_graph():
    b = C.GraphBuilder()
    for a, z in [(0, 1), (1, 2), (2, 3), (3, 0)]:
        b.add_edge(("n", str(a)), ("n", str(z)), 1.0, C.EdgeType.Undirected)
    return b.build_graph()


def test_facade_load_into(tmp_path):
    g = small_graph()
    p = tmp_path / "f.tsv"
    # every node: the set keeps the loader's CSR
    write(p, "n\t0\ta b\nn\t1\tb\nn\t2\tc a\nn\t3\td\n")
    fs = C.FeatureSet.new_from_graph(g)
    fs.load_into(str(p))
    assert fs._csr is not None
    assert fs.get_features(("n", "2")) == [("feat", "c"), ("feat", "a")]
    ip, ids = fs.to_csr()
    assert ip.tolist() == [0, 2, 3, 5, 6] and ids.tolist() == [0, 1, 1, 2, 0, 3]
    # some nodes, after set_features: lists replaced, others kept, ids continue the vocabulary
    fs = C.FeatureSet.new_from_graph(g)
    fs.set_features(("n", "3"), [("feat", "z")])
    write(p, "n\t1\tz y\nx\t9\tignored\n")
    fs.load_into(str(p), C.FeatureNamespace.node_type())
    assert fs.get_features(("n", "1")) == [("n", "z"), ("n", "y")]
    assert fs.get_features(("n", "3")) == [("feat", "z")]
    assert fs.get_features(("n", "0")) == []
    assert fs.num_features() == 3
    fs.fill_missing_nodes()
    assert fs.get_features(("n", "0")) == [("node", "0")]


# --------------------------------------------------------------------------------------
# The native embedding file loader (ep_embeddings_load_tsv) behind NodeEmbeddings.load
# (src/lib.rs:2868-2894, EmbeddingReader::load io.rs:135-200, line_to_embedding 255-268,
# is_node_type 226-231, RecordReader's skip 99)
# --------------------------------------------------------------------------------------
def test_embeddings_known_answers(tmp_path):
    p = tmp_path / "e.tsv"
    # values trimmed of Unicode whitespace; Rust float forms (sign, exponent, inf, nan)
    write(p, "u\ta\t[1.5, -2e-3,　3]\ni\tb\t[+0.25,inf,-1E2]\nu\tc\t[NaN, 7 ,8.]\n")
    names, t = N.load_embeddings_tsv(str(p))
    assert names == [("u", "a"), ("i", "b"), ("u", "c")]
    assert t.dtype == np.float32 and t.shape == (3, 3)
    assert t[0].tolist() == [1.5, np.float32(-2e-3), 3.0]
    assert t[1].tolist() == [0.25, np.inf, -100.0]
    assert np.isnan(t[2, 0]) and t[2, 1:].tolist() == [7.0, 8.0]
    # skip_rows counts lines before the type filter; lines without a tab never match a filter
    write(p, "u\tskip\t[bad\nno-tab-here\nu\ta\t[1,2]\ni\tb\t[3,4]\nu\tc\t[5,6]\n")
    names, t = N.load_embeddings_tsv(str(p), ["u"], skip_rows=1)
    assert names == [("u", "a"), ("u", "c")] and t.tolist() == [[1, 2], [5, 6]]


def test_embeddings_rounding_is_correct(tmp_path):
    # a whole-string f32 parse rounds once (fast_float).  Just above 1 + 2^-24, the midpoint of
    # 1 and its f32 successor: through float64 it becomes the midpoint and ties to even (1.0);
    # rounded once it goes up.  Also the largest finite value, a subnormal, a tie to even.
    vals = ["1.000000059604644775390625000001", "3.4028235e38", "1e-45", "16777217"]
    p = tmp_path / "e.tsv"
    write(p, "u\ta\t[" + ",".join(vals) + "]\n")
    _, t = N.load_embeddings_tsv(str(p))
    assert np.float32(float(vals[0])) == np.float32(1.0)  # the double-rounded value
    assert t[0, 0] == np.nextafter(np.float32(1.0), np.float32(2.0))
    assert t[0, 1] == np.finfo(np.float32).max
    assert t[0, 2] == np.float32(2.0 ** -149)
    assert t[0, 3] == np.float32(16777216.0)


def test_embeddings_errors(tmp_path):
    p = tmp_path / "e.tsv"
    write(p, "u\ta\t[1,2]\nu\ta\t[3,4]\n")
    with pytest.raises(KeyError, match="found duplicate node at 1"):
        N.load_embeddings_tsv(str(p))
    write(p, "u\ta\t[1,2]\nu\tb\t[3]\n")
    with pytest.raises(ValueError, match="different sizes"):
        N.load_embeddings_tsv(str(p))
    for bad in ["u\ta\t[1,x]", "u\ta\t[]", "u\ta\t[1,,2]", "u\ta\t[1]\textra", "u\ta", "u\ta\t["]:
        write(p, bad + "\n")
        with pytest.raises(ValueError, match="Error parsing line"):
            N.load_embeddings_tsv(str(p))


@pytest.mark.parametrize("gz", [False, True])
def test_embeddings_save_load_round_trip(tmp_path, gz):
    rng = np.random.default_rng(int(gz))
    vocab = C.api.Vocab([("u" if i % 3 else "i", f"n{i}") for i in range(50)])
    table = rng.standard_normal((50, 7)).astype(np.float32)
    table[3, 2] = np.float32(1e-7)
    emb = C.NodeEmbeddings(vocab, table, C.Distance.Cosine)
    p = tmp_path / ("e.tsv.gz" if gz else "e.tsv")
    emb.save(str(p))
    back = C.NodeEmbeddings.load(str(p))
    assert list(back.vocab_.names) == vocab.names
    assert np.array_equal(back.embeddings, table)
    only_i = C.NodeEmbeddings.load(str(p), filter_type="i")
    assert [t for t, _ in only_i.vocab_.names] == ["i"] * 17
