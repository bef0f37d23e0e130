"""Python facade with the reference's EP API surface (src/lib.rs PyO3 classes).

Class and argument names, defaults and side effects follow the reference:

* ``EmbeddingPropagator`` / ``learn_features``   src/lib.rs:1119-1377
* ``EPLoss``, ``LossWeighting``                  src/lib.rs:939-1110
* ``Graph``, ``GraphBuilder``, ``EdgeType``      src/lib.rs:292-502, 836-937
* ``FeatureSet``                                 src/lib.rs:1432-1691, src/feature_store.rs
* ``NodeEmbeddings``, ``Distance``               src/lib.rs:2535-2894
* ``FeatureAggregator``, ``NodeEmbedder``        src/lib.rs:1766-2207 (Averaged, Weighted, Attention)
* ``NeighborhoodAligner``                        src/lib.rs:3040-3171, src/algos/alignment.rs

The training step itself runs in libcloverleaf_ep.so (HIP kernels on the GPU);
this module only converts names to integer CSR arrays and back.
"""
from __future__ import annotations

import math
import sys
from typing import Dict, Iterable, List, Optional, Sequence, Tuple

import numpy as np

from . import _native as N

FQNode = Tuple[str, str]
SEED = 20222022  # src/lib.rs:90


def _f32_debug(x) -> str:
    """An f32 the way Rust's `{:?}` prints it: shortest round-trip digits, at least one
    fractional digit."""
    return np.format_float_positional(np.float32(x), unique=True, trim="0")


# --------------------------------------------------------------------------------------
# Vocabulary (src/vocab.rs): (node_type, name) <-> dense id, in insertion order
# --------------------------------------------------------------------------------------
class Vocab:
    def __init__(self, names: Optional[Iterable[FQNode]] = None):
        self.names: List[FQNode] = []
        self.index: Dict[FQNode, int] = {}
        for nm in names or ():
            self.get_or_insert(*nm)

    def get_or_insert(self, node_type: str, name: str) -> int:
        key = (str(node_type), str(name))
        i = self.index.get(key)
        if i is None:
            i = len(self.names)
            self.names.append(key)
            self.index[key] = i
        return i

    def extend_new(self, names: Sequence[FQNode]):
        """Appends names known to be absent, ids len(self), len(self) + 1, ..."""
        base = len(self.names)
        self.names.extend(names)
        self.index.update(zip(names, range(base, base + len(names))))

    def get(self, node: FQNode) -> Optional[int]:
        return self.index.get((str(node[0]), str(node[1])))

    def name(self, i: int) -> FQNode:
        return self.names[i]

    def __len__(self):
        return len(self.names)

    @classmethod
    def anonymous(cls, n: int, node_type: str = "node") -> "Vocab":
        v = cls()
        v.names = [(node_type, str(i)) for i in range(n)]
        v.index = {k: i for i, k in enumerate(v.names)}
        return v


def _node_id(vocab: Vocab, node: FQNode) -> int:
    i = vocab.get(node)
    if i is None:
        raise KeyError(f"Node {node} does not exist!")
    return i


class EdgeType:
    Directed = "Directed"
    Undirected = "Undirected"


class Distance:
    Cosine = "Cosine"
    Euclidean = "Euclidean"
    Dot = "Dot"
    ALT = "ALT"
    Hamming = "Hamming"
    Jaccard = "Jaccard"


# --------------------------------------------------------------------------------------
# Graph
# --------------------------------------------------------------------------------------
class Graph:
    """CSR graph with CDF edge weights (CumCSR, src/graph.rs:399-550)."""

    def __init__(self, indptr: np.ndarray, indices: np.ndarray, weights_cdf: Optional[np.ndarray], vocab: Vocab):
        self.indptr = np.ascontiguousarray(indptr, dtype=np.uint64)
        self.indices = np.ascontiguousarray(indices, dtype=np.uint32)
        self.weights = None if weights_cdf is None else np.ascontiguousarray(weights_cdf, dtype=np.float32)
        self.vocab_ = vocab

    @classmethod
    def from_csr(cls, indptr, indices, weights=None, vocab: Optional[Vocab] = None) -> "Graph":
        """Builds a Graph from integer CSR arrays (node i named ("node", str(i)) by default).

        ``weights`` are raw edge weights; they are stored as per-row CDFs like CumCSR::convert.
        """
        n = len(indptr) - 1
        cdf = None
        if weights is not None:
            cdf = _rows_to_cdf(np.asarray(indptr, dtype=np.int64), np.asarray(weights, dtype=np.float32))
        return cls(indptr, indices, cdf, vocab or Vocab.anonymous(n))

    def nodes(self) -> int:
        return len(self.indptr) - 1

    def edges(self) -> int:
        return int(self.indptr[-1])

    def contains_node(self, name: FQNode) -> bool:
        return self.vocab_.get(name) is not None

    def degree_array(self) -> np.ndarray:
        return np.diff(self.indptr.astype(np.int64))

    def edge_cdf(self) -> np.ndarray:
        """Per-row CDF of the edge weights (CumCSR), uniform rows for an unweighted graph."""
        if self.weights is not None:
            return self.weights
        deg = self.degree_array()
        start = np.repeat(self.indptr[:-1].astype(np.int64), deg)
        pos = (np.arange(len(self.indices), dtype=np.int64) - start + 1).astype(np.float32)
        return np.minimum(pos / np.repeat(deg, deg).astype(np.float32), np.float32(1.0))

    def get_edges(self, node: FQNode, normalized: Optional[bool] = None):
        i = _node_id(self.vocab_, node)
        s, e = int(self.indptr[i]), int(self.indptr[i + 1])
        names = [self.vocab_.name(int(j)) for j in self.indices[s:e]]
        w = self.weights[s:e] if self.weights is not None else _uniform_cdf(e - s)
        if normalized:
            w = np.concatenate([w[:1], np.diff(w)]) if len(w) else w
        return names, [float(x) for x in w]

    def vocab(self):
        return iter(self.vocab_.names)

    def __len__(self):
        return self.nodes()

    def __repr__(self):
        return f"Graph<Nodes={self.nodes()}, Edges={self.edges()}>"

    def save(self, path: str, comp_level: Optional[int] = None):
        """Graph.save (src/lib.rs:414-430): from_type<TAB>from<TAB>to_type<TAB>to<TAB>weight per
        edge, the weight the edge's probability (CDFtoP of the row); `.gz` paths are gzipped."""
        with _open_text_write(path, comp_level) as fh:
            for i in range(self.nodes()):
                s, e = int(self.indptr[i]), int(self.indptr[i + 1])
                if s == e:
                    continue
                cdf = self.weights[s:e] if self.weights is not None else _uniform_cdf(e - s)
                p = np.empty(e - s, dtype=np.float32)
                p[0] = cdf[0]
                p[1:] = cdf[1:] - cdf[:-1]
                ft, fn = self.vocab_.name(i)
                for j, w in zip(self.indices[s:e], p):
                    tt, tn = self.vocab_.name(int(j))
                    fh.write(f"{ft}\t{fn}\t{tt}\t{tn}\t{_f32_display(w)}\n")

    @staticmethod
    def load(path: str, edge_type: str, chunk_size: Optional[int] = None, skip_rows: Optional[int] = None,
             weighted: Optional[bool] = None, deduplicate: Optional[bool] = None) -> "Graph":
        """Graph.load (src/lib.rs:472-500, GraphReader::load io.rs:270-323) in the native
        loader: five tab-separated fields per line, from_type, from, to_type, to, weight
        (ignored unless weighted, default true); Undirected adds the reverse edge;
        deduplicate defaults to false; gzip for `.gz` paths."""
        indptr, indices, cdf, names = N.load_graph_tsv(
            path, edge_type == EdgeType.Undirected, True if weighted is None else bool(weighted),
            False if deduplicate is None else bool(deduplicate), int(skip_rows or 0))
        return Graph(indptr, indices, cdf, Vocab(names))


def _uniform_cdf(k: int) -> np.ndarray:
    if k == 0:
        return np.zeros(0, dtype=np.float32)
    w = np.minimum(np.cumsum(np.ones(k, dtype=np.float32)) / np.float32(k), np.float32(1.0))
    w[-1] = 1.0
    return w


def _rows_to_cdf(indptr: np.ndarray, weights: np.ndarray) -> np.ndarray:
    """convert_edges_to_cdf per row (src/graph.rs:664-691), float32 like the reference."""
    out = weights.astype(np.float32).copy()
    for i in range(len(indptr) - 1):
        s, e = int(indptr[i]), int(indptr[i + 1])
        if s == e:
            continue
        w = out[s:e]
        denom = np.cumsum(w, dtype=np.float32)[-1]  # sequential f32 sum, as iter().sum()
        if denom == 0:
            w[:] = 1.0
            denom = np.float32(e - s)
        acc = np.cumsum(w, dtype=np.float32)
        w[:] = np.minimum(acc / denom, np.float32(1.0))
        w[-1] = 1.0
    return out


class GraphBuilder:
    """GraphBuilder (src/lib.rs:850-937)."""

    def __init__(self):
        self.vocab = Vocab()
        self.edges_: List[Tuple[int, int, float]] = []

    def __repr__(self):
        return f"GraphBuilder<Nodes={len(self.vocab)}, Edges={len(self.edges_)}>"

    def add_edge(self, from_node: FQNode, to_node: FQNode, weight: float, node_type: str):
        f = self.vocab.get_or_insert(*from_node)
        t = self.vocab.get_or_insert(*to_node)
        self.edges_.append((f, t, float(weight)))
        if node_type == EdgeType.Undirected:
            self.edges_.append((t, f, float(weight)))

    def build_graph(self, deduplicate: Optional[bool] = None) -> Optional[Graph]:
        """CSR::construct_from_edges (src/graph.rs:70-142) + CumCSR::convert (414-422)."""
        if not self.edges_:
            return None
        vocab, edges = self.vocab, self.edges_
        self.vocab, self.edges_ = Vocab(), []
        dedup = True if deduplicate is None else bool(deduplicate)
        e = np.asarray(edges, dtype=np.float64)
        src, dst, w = e[:, 0].astype(np.int64), e[:, 1].astype(np.int64), e[:, 2].astype(np.float32)
        n = max(int(max(src.max(), dst.max())) + 1, len(vocab))
        if dedup:
            order = np.lexsort((dst, src))
            src, dst, w = src[order], dst[order], w[order]
            new = np.ones(len(src), dtype=bool)
            new[1:] = (src[1:] != src[:-1]) | (dst[1:] != dst[:-1])
            grp = np.cumsum(new) - 1
            ws = np.zeros(int(grp[-1]) + 1, dtype=np.float32)
            for gi, wi in zip(grp, w):  # sequential float32 sums, like the reference
                ws[gi] = np.float32(ws[gi] + wi)
            src, dst, w = src[new], dst[new], ws
        else:
            order = np.argsort(src, kind="stable")
            src, dst, w = src[order], dst[order], w[order]
        indptr = np.zeros(n + 1, dtype=np.uint64)
        np.add.at(indptr, src + 1, 1)
        indptr = np.cumsum(indptr).astype(np.uint64)
        cdf = _rows_to_cdf(indptr.astype(np.int64), w)
        return Graph(indptr, dst.astype(np.uint32), cdf, vocab)


# --------------------------------------------------------------------------------------
# FeatureSet (src/feature_store.rs, src/lib.rs:1432-1691)
# --------------------------------------------------------------------------------------
class FeatureNamespace:
    def __init__(self, kind: str, value: Optional[str] = None):
        self.kind, self.value = kind, value

    @staticmethod
    def single(ns: str):
        return FeatureNamespace("static", ns)

    @staticmethod
    def node_type():
        return FeatureNamespace("node_type")

    @staticmethod
    def prefix(delim: str):
        return FeatureNamespace("prefix", delim)

    def parse(self, node_type: str, token: str) -> Optional[FQNode]:
        if self.kind == "static":
            return (self.value, token)
        if self.kind == "node_type":
            return (node_type, token)
        idx = token.find(self.value)
        return None if idx < 0 else (token[:idx], token[idx + 1:])


class FeatureSet:
    def __init__(self, vocab: Vocab, n_nodes: int):
        self.vocab_ = vocab
        self.features: List[List[int]] = [[] for _ in range(n_nodes)]
        self.feature_vocab = Vocab()
        self._csr = None  # (f_indptr, f_ids) fast path for large synthetic sets

    @staticmethod
    def new_from_graph(graph: Graph) -> "FeatureSet":
        return FeatureSet(graph.vocab_, graph.nodes())

    @classmethod
    def from_csr(cls, graph: Graph, f_indptr: np.ndarray, f_ids: np.ndarray, n_vocab: int) -> "FeatureSet":
        """Integer feature lists (feature k named ("feat", str(k))) for large synthetic inputs."""
        fs = cls(graph.vocab_, 0)
        fs._csr = (np.ascontiguousarray(f_indptr, dtype=np.uint64), np.ascontiguousarray(f_ids, dtype=np.uint32))
        fs._csr_nodes = graph.nodes()
        fs.feature_vocab = Vocab.anonymous(int(n_vocab), "feat")
        return fs

    def _materialize(self):
        if self._csr is not None:
            ip, ids = self._csr
            self.features = [ids[int(ip[i]):int(ip[i + 1])].astype(np.int64).tolist() for i in range(self._csr_nodes)]
            self._csr = None

    def set_features(self, node: FQNode, features: Sequence[FQNode]):
        self._materialize()
        i = _node_id(self.vocab_, node)
        self.features[i] = [self.feature_vocab.get_or_insert(t, n) for t, n in features]

    def get_features(self, node: FQNode) -> List[FQNode]:
        i = _node_id(self.vocab_, node)
        if self._csr is not None:
            ip, ids = self._csr
            return [self.feature_vocab.name(int(f)) for f in ids[int(ip[i]):int(ip[i + 1])]]
        return [self.feature_vocab.name(f) for f in self.features[i]]

    def load_into(self, path: str, f_ns: Optional[FeatureNamespace] = None):
        """FeatureSet.load_into (src/lib.rs:1579-1606): node_type<TAB>name<TAB>f1 f2 ... lines,
        parsed by the library's native loader (ep_features_load_tsv).  Nodes missing from
        the vocabulary are skipped; a later line of a node replaces its features."""
        ns = f_ns or FeatureNamespace.single("feat")
        kind = {"static": N.EP_FNS_STATIC, "node_type": N.EP_FNS_NODE_TYPE, "prefix": N.EP_FNS_PREFIX}[ns.kind]
        nodes, ip, ids, new = N.load_features_tsv(path, self.vocab_.names, self.feature_vocab.names, kind,
                                                  ns.value or "")
        self.feature_vocab.extend_new(new)  # ids len(feature_vocab), len + 1, ... in the loader's order
        n = len(self.vocab_)
        if len(nodes) == n and self._csr is None and not any(self.features):
            # every node set by the file: keep its CSR
            self._csr = (np.ascontiguousarray(ip, dtype=np.uint64), np.ascontiguousarray(ids, dtype=np.uint32))
            self._csr_nodes = n
            self.features = []
            return
        self._materialize()
        for k, v in enumerate(nodes.tolist()):
            self.features[v] = ids[int(ip[k]):int(ip[k + 1])].astype(np.int64).tolist()

    def fill_missing_nodes(self):
        """feature_store.rs:83-89: featureless nodes get the anonymous feature ("node", str(i))."""
        if self._csr is not None:
            ip, ids = self._csr
            empty = np.flatnonzero(np.diff(ip.astype(np.int64)) == 0)
            if len(empty) == 0:
                return
            self._materialize()
        for i in range(len(self.features)):
            if not self.features[i]:
                self.features[i] = [self.feature_vocab.get_or_insert("node", str(i))]

    def to_csr(self) -> Tuple[np.ndarray, np.ndarray]:
        if self._csr is not None:
            return self._csr
        counts = np.fromiter((len(f) for f in self.features), dtype=np.uint64, count=len(self.features))
        ip = np.zeros(len(self.features) + 1, dtype=np.uint64)
        np.cumsum(counts, out=ip[1:])
        ids = np.fromiter((x for f in self.features for x in f), dtype=np.uint32, count=int(ip[-1]))
        return ip, ids

    def nodes(self) -> int:
        return len(self.vocab_)

    def num_features(self) -> int:
        return len(self.feature_vocab)

    def vocab(self):
        return iter(self.vocab_.names)

    def __getitem__(self, idx: int) -> FQNode:
        n = self.num_features()
        if idx < 0:
            idx += n
        if idx >= n:
            raise IndexError("Feature Index larger than FeatureSet!")
        return self.feature_vocab.name(idx)

    def __repr__(self):
        return f"FeatureSet<Nodes={len(self.vocab_)},UniqueFeatures={self.num_features()}>"


def _open_text(path: str):
    if path.endswith(".gz"):
        import gzip
        return gzip.open(path, "rt")
    return open(path, "r")


def _open_text_write(path: str, comp_level: Optional[int] = None):
    """open_file_for_writing (io.rs:219-231): gzip for `.gz` paths (level 1 by default)."""
    if path.endswith(".gz"):
        import gzip
        return gzip.open(path, "wt", compresslevel=1 if comp_level is None else int(comp_level))
    return open(path, "w")


def _f32_display(x) -> str:
    """An f32 the way Rust's `{}` prints it: the shortest digits that read back to the same
    value, positional."""
    return np.format_float_positional(np.float32(x), unique=True, trim="-")


# --------------------------------------------------------------------------------------
# NodeEmbeddings (src/lib.rs:2535-2894)
# --------------------------------------------------------------------------------------
class NodeEmbeddings:
    def __init__(self, vocab: Vocab, embeddings: np.ndarray, distance: str = Distance.Cosine):
        self.vocab_ = vocab
        self.embeddings = np.ascontiguousarray(embeddings, dtype=np.float32)
        self.distance = distance

    def dims(self) -> int:
        return int(self.embeddings.shape[1]) if self.embeddings.ndim == 2 else 0

    def len(self) -> int:
        return int(self.embeddings.shape[0])

    def __len__(self):
        return self.len()

    def contains(self, node: FQNode) -> bool:
        return self.vocab_.get(node) is not None

    def get_embedding(self, node: FQNode) -> List[float]:
        return self.embeddings[_node_id(self.vocab_, node)].tolist()

    def set_embedding(self, node: FQNode, embedding: Sequence[float]):
        self.embeddings[_node_id(self.vocab_, node)] = np.asarray(embedding, dtype=np.float32)

    def vocab(self):
        return iter(self.vocab_.names)

    def get_distance(self) -> str:
        return self.distance

    def __repr__(self):
        return f"NodeEmbeddings<Nodes={self.len()}, Dims={self.dims()}>"

    def save(self, path: str, comp_level: Optional[int] = None):
        """NodeEmbeddings.save (src/lib.rs:2828-2840, EmbeddingWriter io.rs:24-78): one line per
        node, node_type<TAB>name<TAB>[v1,v2,...], each value as ryu prints an f32; `.gz`
        paths are gzipped (level comp_level, default 1).  Written by the library's native
        writer (ep_embeddings_write_tsv)."""
        N.write_embeddings_tsv(path, [self.vocab_.name(i) for i in range(self.len())], self.embeddings, comp_level)

    @staticmethod
    def load(path: str, distance: Optional[str] = None, filter_type=None, chunk_size: Optional[int] = None,
             skip_rows: Optional[int] = None) -> "NodeEmbeddings":
        """NodeEmbeddings.load (src/lib.rs:2868-2894, EmbeddingReader::load io.rs:135-200):
        node_type<TAB>name<TAB>[v1,...] lines, optionally only the node types in filter_type
        (a string or a list); a repeated node or a different width is an error.  Parsed by
        the library's native loader (ep_embeddings_load_tsv)."""
        types = None
        if filter_type is not None:
            types = sorted({filter_type} if isinstance(filter_type, str) else set(filter_type))
        names, table = N.load_embeddings_tsv(path, types, skip_rows or 0)
        vocab = Vocab()
        vocab.extend_new(names)
        if not names:
            table = np.zeros((0, 0), dtype=np.float32)
        return NodeEmbeddings(vocab, table, distance or Distance.Cosine)


# --------------------------------------------------------------------------------------
# Losses and weighting (src/lib.rs:939-1110)
# --------------------------------------------------------------------------------------
class LossWeighting:
    def __init__(self, kind: int, weight: float = 0.0):
        self.kind, self.weight = kind, weight

    @staticmethod
    def Log():
        return LossWeighting(N.EP_LW_LOG)

    @staticmethod
    def Exponential(weight: float):
        return LossWeighting(N.EP_LW_EXPONENTIAL, float(weight))


class EPLoss:
    def __init__(self, kind: str, params: tuple, negatives: int):
        self.kind, self.params, self.negatives = kind, params, negatives

    @staticmethod
    def margin(gamma: float, negatives: Optional[int] = None):
        return EPLoss("margin", (float(gamma),), int(negatives) if negatives is not None else 1)

    @staticmethod
    def contrastive(positive_margin: float, negative_margin: float, negatives: int):
        return EPLoss("contrastive", (positive_margin, negative_margin), max(1, negatives))

    @staticmethod
    def starspace(gamma: float, negatives: int):
        return EPLoss("starspace", (gamma,), max(1, negatives))

    @staticmethod
    def rank(tau: float, negatives: int):
        return EPLoss("rank", (tau,), max(1, negatives))

    @staticmethod
    def rankspace(tau: float, negatives: int):
        return EPLoss("rankspace", (tau,), max(1, negatives))

    @staticmethod
    def ppr(gamma: float, negatives: int, restart_p: float):
        return EPLoss("ppr", (gamma, restart_p), max(1, negatives))

    def __repr__(self):
        return f"EPLoss<{self.kind}{self.params}, negatives={self.negatives}>"


def _parse_sample(x: Optional[float], what: str) -> Tuple[int, int, float]:
    """Sample::new (src/algos/utils.rs:173-184): (0,1) -> Probability(p), else Fixed(int).
    Returns (kind, k, p)."""
    if x is None:
        return N.EP_SAMPLE_ALL, 0, 0.0
    x = float(x)
    if x < 0:
        raise ValueError(f"{what} must be either None, or (0, N)")
    if 0 < x < 1:
        return N.EP_SAMPLE_PROBABILITY, 0, float(np.float32(x))
    k = int(x)
    if k == 0:
        raise ValueError(f"{what}: Fixed(0) samples nothing")
    return N.EP_SAMPLE_FIXED, k, 0.0


# --------------------------------------------------------------------------------------
# EmbeddingPropagator (src/lib.rs:1119-1377)
# --------------------------------------------------------------------------------------
class EmbeddingPropagator:
    """Embedding Propagation trainer; the step runs on the GPU.

    Reference arguments keep their names and defaults.  Extra keyword arguments
    select the MI355X execution: ``precision`` ("fp32" | "bf16" | "bf16-master"), ``grad_mode``
    ("first_leaf", the reference rule, | "exact"), ``scatter`` ("sort" | "atomic"),
    ``use_graph`` (replay steps from a captured hipGraph), ``concurrent_batches`` (1, 2, 4
    or 8 batches per step, reading the same parameters and updating in batch order, like
    the reference's concurrently running batches; 1 = strictly sequential), ``device``,
    ``fp16_adjoints`` (BF16 tables, mean model: the adjoint rows the update gathers are
    held in FP16 with a power-of-two scale per row).
    ``attention=d_k`` with ``attention_heads`` trains the attention aggregator (Full
    attention, rows of attention_heads * (2 d_k + dims) values, at most 64 features per
    node); ``context_window`` selects Sliding attention (window_size >= 1).  Weighted
    neighbour sampling and averaging use the graph's edge weights; as in the reference,
    the attention model ignores weighted averaging and its weighted sampling is uniform.
    """

    def __init__(self, alpha=None, loss=None, batch_size=None, dims=None, passes=None, seed=None, max_nodes=None,
                 weighted_neighbor_sampling=None, weighted_neighbor_averaging=None, max_features=None,
                 loss_weighting=None, valid_pct=None, hard_negatives=None, indicator=None, attention=None,
                 attention_heads=None, context_window=None, noise=None, *, precision="fp32",
                 grad_mode="first_leaf", scatter="sort", use_graph=False, concurrent_batches=1, device=0,
                 fp16_adjoints=False):
        loss = loss or EPLoss.margin(1.0, 1)
        mf_kind, mf, mf_p = _parse_sample(max_features, "max_features")
        mn_kind, mn, mn_p = _parse_sample(max_nodes, "max_nodes")
        if attention is not None and context_window is not None and int(context_window) < 1:
            raise ValueError("context_window must be >= 1 (window_size 0 is not supported by this build)")
        if attention is not None and context_window is not None:
            # lib.rs:1288-1300: Sliding attention ignores max_features; without a window,
            # max_features is Random attention's column sample (kept in the config)
            mf_kind, mf, mf_p = N.EP_SAMPLE_ALL, 0, 0.0
        cfg = N.default_config()
        cfg.alpha = 0.9 if alpha is None else float(alpha)
        cfg.batch_size = 50 if batch_size is None else int(batch_size)
        cfg.dims = 100 if dims is None else int(dims)
        cfg.passes = 100 if passes is None else int(passes)
        cfg.negatives = loss.negatives
        cfg.hard_negatives = 0 if hard_negatives is None else int(hard_negatives)
        cfg.weighted_neighbor_sampling = 1 if weighted_neighbor_sampling else 0
        cfg.weighted_neighbor_averaging = 1 if weighted_neighbor_averaging else 0
        cfg.loss_kind = N.LOSS_KINDS[loss.kind]
        cfg.gamma = loss.params[0]
        cfg.loss_param2 = float(loss.params[1]) if len(loss.params) > 1 else 0.0
        cfg.valid_pct = 0.1 if valid_pct is None else float(valid_pct)
        cfg.seed = SEED if seed is None else int(seed)
        cfg.max_nodes_kind, cfg.max_nodes, cfg.max_nodes_p = mn_kind, mn, mn_p
        cfg.max_features_kind, cfg.max_features, cfg.max_features_p = mf_kind, mf, mf_p
        if loss_weighting is not None:
            cfg.loss_weighting = loss_weighting.kind
            cfg.loss_weight_exp = loss_weighting.weight
        cfg.precision = {"fp32": N.EP_PREC_FP32, "bf16": N.EP_PREC_BF16,
                         "bf16-master": N.EP_PREC_BF16_MASTER}[precision]
        cfg.grad_mode = {"first_leaf": N.EP_GRAD_FIRST_LEAF, "exact": N.EP_GRAD_EXACT}[grad_mode]
        cfg.scatter = {"sort": N.EP_SCATTER_SORT, "atomic": N.EP_SCATTER_ATOMIC}[scatter]
        cfg.use_graph = 1 if use_graph else 0
        cfg.concurrent_batches = int(concurrent_batches)
        cfg.fp16_adjoints = 1 if fp16_adjoints else 0
        cfg.noise = 0.0 if noise is None else float(noise)
        cfg.device = int(device)
        if attention is not None:
            # lib.rs:1288-1299: AttentionFeatureModel with Full attention (max_features All)
            cfg.attention_dk = int(attention)
            cfg.attention_heads = 1 if attention_heads is None else int(attention_heads)
            cfg.context_window = 0 if context_window is None else int(context_window)
        self.cfg = cfg
        self.indicator = True if indicator is None else bool(indicator)
        self.history: List[dict] = []

    def __repr__(self):
        c = self.cfg
        k, g, p2, K = c.loss_kind, _f32_debug(c.gamma), _f32_debug(c.loss_param2), c.negatives
        # the Debug form of Loss (src/algos/ep/loss.rs:15-36)
        loss = {N.EP_LOSS_MARGIN: f"MarginLoss({g}, {K})", N.EP_LOSS_CONTRASTIVE: f"Contrastive({g}, {p2}, {K})",
                N.EP_LOSS_STARSPACE: f"StarSpace({g}, {K})", N.EP_LOSS_RANK: f"RankLoss({g}, {K})",
                N.EP_LOSS_RANKSPACE: f"RankSpace({g}, {K})", N.EP_LOSS_PPR: f"PPR({g}, {K}, {p2})"}[k]
        return (f"EmbeddingPropagation {{ alpha: {_f32_debug(c.alpha)}, loss: {loss}, "
                f"batch_size: {c.batch_size}, d_model: {c.dims}, passes: {c.passes}, seed: {c.seed}, "
                f"valid_pct: {_f32_debug(c.valid_pct)} }}")

    def learn_features(self, graph: Graph, features: FeatureSet,
                       feature_embeddings: Optional[NodeEmbeddings] = None) -> NodeEmbeddings:
        """lib.rs:1327-1370.  Side effects kept: ``features`` gets fill_missing_nodes;
        a warm-start ``feature_embeddings`` is moved out (left as an empty store)."""
        features.fill_missing_nodes()
        f_indptr, f_ids = features.to_csr()
        n_vocab = features.num_features()
        warm = None
        if feature_embeddings is not None:
            warm = feature_embeddings.embeddings
            feature_embeddings.embeddings = np.zeros((len(feature_embeddings.vocab_), 0), dtype=np.float32)
            width = N.row_width(self.cfg)
            if warm.shape != (n_vocab, width):
                raise ValueError(f"warm-start table has shape {warm.shape}, expected {(n_vocab, width)}")
        with N.Context(self.cfg) as ctx:
            ctx.set_graph(graph.indptr, graph.indices)
            if self.cfg.weighted_neighbor_averaging:
                ctx.set_edge_weights(graph.edge_cdf())
            ctx.set_features(f_indptr, f_ids, n_vocab)
            ctx.set_table(warm)
            self.history = []

            def progress(p, step, tr, va, lr):
                self.history.append(dict(pass_=p, step=step, train=tr, valid=va, lr=lr))
                if self.indicator:
                    print(f"Pass {p}/{self.cfg.passes}, Train: {tr:.5f}, Valid: {va:.5f}, LR: {lr:.5f}",
                          file=sys.stderr)

            ctx.learn(progress)
            table = ctx.get_table()
        return NodeEmbeddings(features.feature_vocab, table, Distance.Cosine)


# --------------------------------------------------------------------------------------
# Inference aggregate (src/lib.rs:1766-2207, aggregator.rs:34-47)
# --------------------------------------------------------------------------------------
def unigram_probability(fs: "FeatureSet") -> np.ndarray:
    """UnigramProbability::new (aggregator.rs:55-72): count of each feature id over every
    node's feature list / total, in f32."""
    ip, ids = fs.to_csr()
    counts = np.bincount(ids.astype(np.int64), minlength=fs.num_features())
    total = int(counts.sum())
    return counts.astype(np.float32) / np.float32(max(total, 1))


class FeatureAggregator:
    def __init__(self, kind: str, num_heads: int = 0, d_k: int = 0, window: Optional[int] = None,
                 alpha: float = 0.0, vocab: Optional[Vocab] = None, unigrams: Optional[np.ndarray] = None):
        self.kind = kind
        self.num_heads, self.d_k, self.window = num_heads, d_k, window
        self.alpha, self.vocab, self.unigrams = alpha, vocab, unigrams

    @staticmethod
    def Averaged():
        return FeatureAggregator("averaged")

    @staticmethod
    def Weighted(alpha: float, fs: "FeatureSet"):
        """lib.rs:1855-1861: SIF weights alpha / (alpha + p(f)) from fs's unigram counts."""
        return FeatureAggregator("weighted", alpha=float(alpha), vocab=fs.feature_vocab, unigrams=unigram_probability(fs))

    def save(self, path: str):
        """lib.rs:1876-1911: the aggregator as text (Weighted keeps its unigram table)."""
        with open(path, "w") as fh:
            if self.kind == "averaged":
                fh.write("Averaged\n")
            elif self.kind == "attention":
                fh.write(f"Attention\n{self.num_heads}\n{self.d_k}\n{self.window or 0}\n")
            else:
                fh.write(f"Weighted\n{self.alpha}\n")
                for i, p in enumerate(self.unigrams):
                    if i >= len(self.vocab):
                        break
                    t, n = self.vocab.name(i)
                    fh.write(f"{t}\t{n}\t{float(p)}\n")

    @staticmethod
    def load(path: str) -> "FeatureAggregator":
        """lib.rs:1925-2000."""
        with open(path) as fh:
            kind = fh.readline().rstrip("\n")
            if kind == "Averaged":
                return FeatureAggregator.Averaged()
            if kind == "Attention":
                h, dk, win = (int(fh.readline().strip()) for _ in range(3))
                return FeatureAggregator.Attention(h, dk, None if win == 0 else win)
            if kind == "Weighted":
                alpha = float(fh.readline().strip())
                vocab, p_w = Vocab(), []
                for line in fh:
                    pieces = line.rstrip("\n").split("\t")
                    if len(pieces) != 3:
                        raise ValueError("Malformed feature line! Need node_type<TAB>name<TAB>weight ...")
                    i = vocab.get_or_insert(pieces[0], pieces[1])
                    if i < len(p_w):
                        raise ValueError(f"Duplicate feature found:{pieces[0]} {pieces[1]}")
                    p_w.append(float(pieces[2]))
                return FeatureAggregator("weighted", alpha=alpha, vocab=vocab, unigrams=np.asarray(p_w, np.float32))
        raise ValueError(f"unknown aggregator type {kind!r}")

    @staticmethod
    def Attention(num_heads: int, d_k: int, window: Optional[int] = None):
        """lib.rs:1835-1838; window: Sliding attention with that window_size."""
        if window is not None and int(window) < 1:
            raise ValueError("window must be >= 1 (window_size 0 is not supported by this build)")
        return FeatureAggregator("attention", int(num_heads), int(d_k), None if window is None else int(window))

    def __repr__(self):
        if self.kind == "attention":
            w = "None" if self.window is None else f"Some({self.window})"
            return f"FeatureAggregator<Attention<num_heads={self.num_heads},d_k={self.d_k},window={w}>"
        if self.kind == "weighted":
            return f"FeatureAggregator<Weighted<alpha={self.alpha}>>"
        return "FeatureAggregator<Averaged>"


class NodeEmbedder:
    def __init__(self, feat_agg: FeatureAggregator):
        if feat_agg.kind not in ("averaged", "attention", "weighted"):
            raise ValueError(f"unknown aggregator {feat_agg.kind}")
        self.feat_agg = feat_agg

    def embed_feature_set(self, feat_set: FeatureSet, feature_embeddings: NodeEmbeddings,
                          device: int = 0) -> NodeEmbeddings:
        """lib.rs:2092-2124: node embedding from its (known) feature rows, on the GPU: their
        mean (Averaged), SIF-weighted mean (Weighted) or Full attention over them with
        count 1 each (Attention)."""
        ip, ids = feat_set.to_csr()
        # translate FeatureSet feature ids into the embedding vocabulary, dropping unknown ones
        fv, ev = feat_set.feature_vocab, feature_embeddings.vocab_
        if fv is ev or fv.names == ev.names[:len(fv.names)]:
            t_ip, t_ids = ip, ids
        else:
            lut = np.array([ev.get(nm) if ev.get(nm) is not None else -1 for nm in fv.names], dtype=np.int64)
            mapped = lut[ids.astype(np.int64)]
            keep = mapped >= 0
            counts = np.add.reduceat(keep.astype(np.int64), ip[:-1].astype(np.int64)) if len(ids) else \
                np.zeros(len(ip) - 1, dtype=np.int64)
            counts[np.diff(ip.astype(np.int64)) == 0] = 0
            t_ip = np.zeros(len(ip), dtype=np.uint64)
            np.cumsum(counts, out=t_ip[1:])
            t_ids = mapped[keep].astype(np.uint32)
        table = feature_embeddings.embeddings
        cfg = N.default_config()
        cfg.device = int(device)
        fa = self.feat_agg
        if fa.kind == "attention":
            # lib.rs:2041-2046: d_model = (width - 2 H d_k) / H
            cfg.attention_heads, cfg.attention_dk = fa.num_heads, fa.d_k
            cfg.context_window = fa.window or 0
            cfg.dims = (table.shape[1] - 2 * fa.num_heads * fa.d_k) // fa.num_heads
            if N.row_width(cfg) != table.shape[1]:
                raise ValueError(f"embedding width {table.shape[1]} does not fit {fa.num_heads} heads of d_k={fa.d_k}")
        else:
            cfg.dims = table.shape[1]
        with N.Context(cfg) as ctx:
            # a one-node placeholder feature list: only the table is needed
            ctx.set_features(np.array([0, 1], dtype=np.uint64), np.array([0], dtype=np.uint32), table.shape[0])
            ctx.set_table(table)
            if fa.kind == "attention":
                out = ctx.embed_attention(t_ip, t_ids)
            elif fa.kind == "weighted":
                # the unigram table is indexed by the translated (embedding) feature ids, as
                # WeightedAggregator::construct does (aggregator.rs:103-108)
                out = ctx.embed_weighted(t_ip, t_ids, fa.unigrams, fa.alpha)
            else:
                out = ctx.embed_averaged(t_ip, t_ids)
        return NodeEmbeddings(feat_set.vocab_, out, Distance.Cosine)


# --------------------------------------------------------------------------------------
# NeighborhoodAligner (src/lib.rs:3040-3171, src/algos/alignment.rs)
# --------------------------------------------------------------------------------------
class NeighborhoodAligner:
    """Shifts node embeddings toward the weighted centroid of their neighbourhood
    (alignment.rs:25-75), on the GPU:

        new_v = (1 - a) * sum_j p_j E[u_j] + a * E[v]

    over the first ``max_neighbors`` edges of v (all by default; the edge probabilities
    p_j are not renormalised), with a = ``alpha``, or min(1 / ln(deg + 1), 1) when alpha
    is None.  Row i of the embeddings belongs to graph node i.  A node without edges
    keeps a * E[v] (the reference indexes a missing first weight there).
    """

    def __init__(self, alpha: Optional[float] = None, max_neighbors: Optional[int] = None, *, device: int = 0):
        self.alpha = None if alpha is None else float(np.float32(alpha))
        self.max_neighbors = None if max_neighbors is None else int(max_neighbors)
        self.device = device

    def __repr__(self):
        a = "None" if self.alpha is None else f"Some({_f32_debug(self.alpha)})"
        m = "None" if self.max_neighbors is None else f"Some({self.max_neighbors})"
        return f"NeighborhoodAligner<Alpha={a}, MaxNeighbors={m}>"

    def _aligner(self, embeddings: NodeEmbeddings, graph: Graph) -> N.Aligner:
        if embeddings.len() != graph.nodes():
            raise ValueError("embeddings must have one row per graph node (the graph's vocabulary)")
        return N.Aligner(graph.indptr, graph.indices, graph.edge_cdf(), embeddings.embeddings, self.device)

    def align(self, embeddings: NodeEmbeddings, graph: Graph) -> NodeEmbeddings:
        """NeighborhoodAligner.align (lib.rs:3095-3113): a new NodeEmbeddings."""
        with self._aligner(embeddings, graph) as al:
            out = al.run(self.alpha, self.max_neighbors)
        return NodeEmbeddings(embeddings.vocab_, out, embeddings.distance)

    def align_to_disk(self, path: str, embeddings: NodeEmbeddings, graph: Graph, chunk_size: Optional[int] = None,
                      comp_level: Optional[int] = None):
        """NeighborhoodAligner.align_to_disk (lib.rs:3140-3168): aligns chunk_size nodes at a
        time (default 10000) and streams them out in the NodeEmbeddings text format."""
        cs = 10_000 if chunk_size is None else max(1, int(chunk_size))
        with self._aligner(embeddings, graph) as al, _open_text_write(path, comp_level) as fh:
            for first in range(0, graph.nodes(), cs):
                count = min(cs, graph.nodes() - first)
                block = al.run(self.alpha, self.max_neighbors, first, count)
                for k in range(count):
                    t, n = embeddings.vocab_.name(first + k)
                    fh.write(f"{t}\t{n}\t[{','.join(str(x) for x in block[k])}]\n")
