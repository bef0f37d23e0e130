"""ctypes binding of the C ABI in include/cloverleaf_ep.h (libcloverleaf_ep.so).

The library is built in-tree by ``cloverleaf_amd.build_native`` (or
``__graft_entry__.build()``).  There is no fallback: if the library is missing
or fails to load, every entry point raises.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, Structure, c_char_p, c_double, c_float, c_int, c_int32, c_uint16, c_uint32, c_uint64, c_void_p
from pathlib import Path
from typing import Optional

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libcloverleaf_ep.so"

EP_SAMPLE_ALL, EP_SAMPLE_FIXED, EP_SAMPLE_PROBABILITY = 0, 1, 2
EP_LW_NONE, EP_LW_LOG, EP_LW_EXPONENTIAL = 0, 1, 2
EP_PREC_FP32, EP_PREC_BF16, EP_PREC_BF16_MASTER = 0, 1, 2
EP_GRAD_FIRST_LEAF, EP_GRAD_EXACT = 0, 1
EP_SCATTER_SORT, EP_SCATTER_ATOMIC = 0, 1
EP_LOSS_MARGIN, EP_LOSS_CONTRASTIVE, EP_LOSS_STARSPACE, EP_LOSS_RANK, EP_LOSS_RANKSPACE, EP_LOSS_PPR = range(6)
LOSS_KINDS = {"margin": EP_LOSS_MARGIN, "contrastive": EP_LOSS_CONTRASTIVE, "starspace": EP_LOSS_STARSPACE,
              "rank": EP_LOSS_RANK, "rankspace": EP_LOSS_RANKSPACE, "ppr": EP_LOSS_PPR}

ERROR_NAMES = {-1: "EP_EINVAL", -2: "EP_ESTATE", -3: "EP_EHIP", -4: "EP_ENOMEM", -5: "EP_ECAPACITY",
               -6: "EP_EDUPLICATE"}
EP_EDUPLICATE = -6


class EPConfig(Structure):
    _fields_ = [
        ("dims", c_uint32),
        ("batch_size", c_uint32),
        ("passes", c_uint32),
        ("negatives", c_uint32),
        ("alpha", c_float),
        ("gamma", c_float),
        ("valid_pct", c_float),
        ("max_nodes_kind", c_int32),
        ("max_nodes", c_uint32),
        ("max_features_kind", c_int32),
        ("max_features", c_uint32),
        ("loss_weighting", c_int32),
        ("loss_weight_exp", c_float),
        ("hard_negatives", c_uint32),
        ("noise", c_float),
        ("weighted_neighbor_sampling", c_int32),
        ("weighted_neighbor_averaging", c_int32),
        ("precision", c_int32),
        ("grad_mode", c_int32),
        ("scatter", c_int32),
        ("use_graph", c_int32),
        ("seed", c_uint64),
        ("device", c_int32),
        ("seg_vpl", c_int32),
        ("seg_unroll", c_int32),
        ("concurrent_batches", c_uint32),
        ("attention_dk", c_uint32),
        ("attention_heads", c_uint32),
        ("context_window", c_uint32),
        ("loss_kind", c_int32),
        ("loss_param2", c_float),
        ("max_features_p", c_float),
        ("max_nodes_p", c_float),
        ("fp16_adjoints", c_int32),
        ("forward_variant", c_int32),
    ]


def list_cap(cfg: "EPConfig") -> int:
    """Positive-list capacity per anchor: max_nodes, or K random-walk endpoints (PPR)."""
    return int(cfg.negatives) if cfg.loss_kind == EP_LOSS_PPR else int(cfg.max_nodes)


def row_width(cfg: "EPConfig") -> int:
    """Table row width: d_model (mean model) or H (2 d_k + d_model) (attention)."""
    if cfg.attention_dk:
        return int(cfg.attention_heads) * (2 * int(cfg.attention_dk) + int(cfg.dims))
    return int(cfg.dims)


class EPBenchResult(Structure):
    _fields_ = [
        ("ms_total", c_double),
        ("ms_forward_avg", c_double),
        ("ms_scatter_avg", c_double),
        ("examples", c_uint64),
        ("rows", c_uint64),
        ("nodes", c_uint64),
        ("unique_rows", c_uint64),
        ("steps", c_uint32),
        ("forward_launches", c_uint32),
        ("seg_vpl", c_int32),
        ("seg_unroll", c_int32),
        ("ms_phase", c_double * 8),
        ("short_rows", c_uint64),
        ("last_long_rows", c_uint32),
        ("last_pieces", c_uint32),
        ("att_items_sq", c_uint64),
        ("att_items", c_uint64),
        ("fwd_x", c_int32),
        ("reserved", c_int32),
    ]


PROGRESS_FN = ctypes.CFUNCTYPE(None, c_void_p, c_uint32, c_uint64, c_float, c_float, c_float)

_u32p = POINTER(c_uint32)
_u64p = POINTER(c_uint64)
_f32p = POINTER(c_float)
_u16p = POINTER(c_uint16)

# name -> (restype, argtypes)
SIGNATURES = {
    "ep_config_default": (c_int, [POINTER(EPConfig)]),
    "ep_create": (c_int, [POINTER(EPConfig), POINTER(c_void_p)]),
    "ep_destroy": (None, [c_void_p]),
    "ep_set_graph": (c_int, [c_void_p, c_uint64, _u64p, _u32p]),
    "ep_set_edge_weights": (c_int, [c_void_p, _f32p]),
    "ep_aligner_create": (c_int, [c_int, c_uint64, _u64p, _u32p, _f32p, c_uint32, _f32p, POINTER(c_void_p)]),
    "ep_aligner_run": (c_int, [c_void_p, c_float, c_uint32, c_uint64, c_uint64, _f32p]),
    "ep_aligner_destroy": (None, [c_void_p]),
    "ep_aligner_last_kernel_ms": (c_float, [c_void_p]),
    "ep_graph_load_tsv": (c_int, [c_char_p, c_int, c_int, c_int, c_uint64, POINTER(c_void_p)]),
    "ep_graph_file_sizes": (c_int, [c_void_p, _u64p, _u64p, _u64p]),
    "ep_graph_file_copy": (c_int, [c_void_p, _u64p, _u32p, _f32p, c_char_p]),
    "ep_graph_file_free": (None, [c_void_p]),
    "ep_graph_file_error": (c_char_p, []),
    "ep_features_load_tsv": (c_int, [c_char_p, c_char_p, _u64p, c_uint64, c_char_p, _u64p, c_uint64, c_int, c_char_p,
                                     POINTER(c_void_p)]),
    "ep_feature_file_sizes": (c_int, [c_void_p, _u64p, _u64p, _u64p, _u64p]),
    "ep_feature_file_copy": (c_int, [c_void_p, _u32p, _u64p, _u32p, c_char_p, _u64p]),
    "ep_feature_file_free": (None, [c_void_p]),
    "ep_embeddings_load_tsv": (c_int, [c_char_p, c_char_p, _u64p, c_uint64, c_int, c_uint64, POINTER(c_void_p)]),
    "ep_embedding_file_sizes": (c_int, [c_void_p, _u64p, _u64p, _u64p]),
    "ep_embedding_file_copy": (c_int, [c_void_p, _f32p, c_char_p, _u64p]),
    "ep_embedding_file_free": (None, [c_void_p]),
    "ep_embeddings_write_tsv": (c_int, [c_char_p, c_int, c_uint64, c_uint64, _f32p, c_char_p, _u64p]),
    "ep_format_f32": (c_int, [c_float, c_char_p]),
    "ep_sample_weights": (c_int, [c_void_p, c_uint32, _f32p]),
    "ep_set_explicit_weights": (c_int, [c_void_p, c_uint32, _f32p]),
    "ep_set_features": (c_int, [c_void_p, c_uint64, c_uint32, _u64p, _u32p]),
    "ep_set_table": (c_int, [c_void_p, _f32p]),
    "ep_get_table": (c_int, [c_void_p, _f32p]),
    "ep_get_table_bf16": (c_int, [c_void_p, _u16p]),
    "ep_get_adam": (c_int, [c_void_p, _f32p, _f32p]),
    "ep_set_split": (c_int, [c_void_p, _u32p, c_uint64, _u32p, c_uint64]),
    "ep_learn": (c_int, [c_void_p, PROGRESS_FN, c_void_p]),
    "ep_sample_batch": (c_int, [c_void_p, _u32p, c_uint32, c_uint32, _u32p, _u32p, _u32p]),
    "ep_step_explicit": (c_int, [c_void_p, _u32p, c_uint32, _u32p, _u32p, _u32p, c_uint32, c_float, c_int,
                                 _f32p, _f32p, _u32p, _u32p, _f32p]),
    "ep_step_explicit_noalloc": (c_int, [c_void_p, _u32p, c_uint32, _u32p, _u32p, _u32p, c_uint32, c_float, _f32p]),
    "ep_bench_steps": (c_int, [c_void_p, c_uint32, c_uint32, c_int, POINTER(EPBenchResult)]),
    "ep_embed_averaged": (c_int, [c_void_p, c_uint64, _u64p, _u32p, _f32p]),
    "ep_embed_attention": (c_int, [c_void_p, c_uint64, _u64p, _u32p, _f32p]),
    "ep_embed_weighted": (c_int, [c_void_p, c_uint64, _u64p, _u32p, _f32p, c_uint32, c_float, _f32p]),
    "ep_philox4x32": (c_int, [_u32p, _u32p, _u32p]),
    "ep_last_error": (c_char_p, []),
}

_lib = None


class NativeError(RuntimeError):
    """A nonzero return code of the C ABI."""


def load(path: os.PathLike | None = None) -> ctypes.CDLL:
    """Loads libcloverleaf_ep.so (once) and declares every exported signature."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path is not None else LIB_PATH
    if not p.exists():
        raise RuntimeError(
            f"{p} is missing: build it with `python -m cloverleaf_amd.build_native` "
            "(or __graft_entry__.build()); there is no CPU fallback")
    lib = ctypes.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def check(rc: int, what: str = ""):
    if rc != 0:
        msg = load().ep_last_error().decode(errors="replace")
        raise NativeError(f"{what or 'ep call'} failed ({ERROR_NAMES.get(rc, rc)}): {msg}")


def ptr(a: np.ndarray, ctype):
    """Pointer to a C-contiguous numpy array (None passes through)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return a.ctypes.data_as(POINTER(ctype))


def default_config() -> EPConfig:
    cfg = EPConfig()
    check(load().ep_config_default(ctypes.byref(cfg)), "ep_config_default")
    return cfg


class Aligner:
    """The graph and node embeddings resident on the device for NeighborhoodAligner runs."""

    def __init__(self, indptr, indices, cdf, emb, device: int = 0):
        indptr = np.ascontiguousarray(indptr, dtype=np.uint64)
        indices = np.ascontiguousarray(indices, dtype=np.uint32)
        cdf = np.ascontiguousarray(cdf, dtype=np.float32)
        emb = np.ascontiguousarray(emb, dtype=np.float32)
        self.n, self.dims = emb.shape
        if len(indptr) != self.n + 1:
            raise ValueError("embeddings must have one row per graph node")
        h = c_void_p()
        check(load().ep_aligner_create(int(device), self.n, ptr(indptr, c_uint64), ptr(indices, c_uint32),
                                       ptr(cdf, c_float), self.dims, ptr(emb, c_float), ctypes.byref(h)),
              "ep_aligner_create")
        self.handle = h

    def run(self, alpha: Optional[float], max_neighbors: Optional[int], first: int = 0,
            count: Optional[int] = None) -> np.ndarray:
        count = self.n - first if count is None else count
        out = np.empty((count, self.dims), dtype=np.float32)
        a = float("nan") if alpha is None else float(alpha)
        check(load().ep_aligner_run(self.handle, a, int(max_neighbors or 0), int(first), int(count),
                                    ptr(out, c_float)), "ep_aligner_run")
        return out

    def last_kernel_ms(self) -> float:
        return float(load().ep_aligner_last_kernel_ms(self.handle))

    def close(self):
        if getattr(self, "handle", None):
            load().ep_aligner_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    """Owning handle of an ``ep_ctx`` (device memory + stream)."""

    def __init__(self, cfg: EPConfig):
        lib = load()
        h = c_void_p()
        check(lib.ep_create(ctypes.byref(cfg), ctypes.byref(h)), "ep_create")
        self._h = h
        self.cfg = cfg
        self.n_vocab = 0
        self.n_nodes = 0

    @property
    def handle(self):
        if self._h is None:
            raise RuntimeError("context is closed")
        return self._h

    def close(self):
        if getattr(self, "_h", None) is not None:
            load().ep_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- uploads ---------------------------------------------------------------
    def set_graph(self, indptr: np.ndarray, indices: np.ndarray):
        indptr = np.ascontiguousarray(indptr, dtype=np.uint64)
        indices = np.ascontiguousarray(indices, dtype=np.uint32)
        n = len(indptr) - 1
        check(load().ep_set_graph(self.handle, n, ptr(indptr, c_uint64), ptr(indices, c_uint32)), "ep_set_graph")
        self.n_nodes = n
        self.deg_max = int(np.diff(indptr.astype(np.int64)).max(initial=0))

    def list_width(self) -> int:
        """Width of the neighbour lists of sample_batch / step_explicit: max_nodes (Fixed),
        K (PPR), the largest degree (max_nodes Probability)."""
        if self.cfg.max_nodes_kind == EP_SAMPLE_PROBABILITY and self.cfg.loss_kind != EP_LOSS_PPR:
            return max(self.deg_max, 1)
        return list_cap(self.cfg)

    def set_edge_weights(self, cdf: np.ndarray):
        """Per-edge CDF of each row (CumCSR weights), one value per CSR entry."""
        cdf = np.ascontiguousarray(cdf, dtype=np.float32)
        check(load().ep_set_edge_weights(self.handle, ptr(cdf, c_float)), "ep_set_edge_weights")

    def sample_weights(self, n: int) -> np.ndarray:
        """Neighbour weights w_j of the last sample_batch (n x max_nodes)."""
        w = np.zeros((n, self.list_width()), dtype=np.float32)
        check(load().ep_sample_weights(self.handle, n, ptr(w, c_float)), "ep_sample_weights")
        return w

    def set_explicit_weights(self, w: np.ndarray):
        """Neighbour weights (n x max_nodes) used by the next step_explicit calls."""
        w = np.ascontiguousarray(w, dtype=np.float32)
        check(load().ep_set_explicit_weights(self.handle, w.shape[0], ptr(w, c_float)), "ep_set_explicit_weights")

    def set_features(self, f_indptr: np.ndarray, f_ids: np.ndarray, n_vocab: int):
        f_indptr = np.ascontiguousarray(f_indptr, dtype=np.uint64)
        f_ids = np.ascontiguousarray(f_ids, dtype=np.uint32)
        n = len(f_indptr) - 1
        check(load().ep_set_features(self.handle, n, int(n_vocab), ptr(f_indptr, c_uint64), ptr(f_ids, c_uint32)),
              "ep_set_features")
        self.n_vocab = int(n_vocab)

    def set_table(self, rows: np.ndarray | None):
        if rows is None:
            check(load().ep_set_table(self.handle, None), "ep_set_table")
            return
        rows = np.ascontiguousarray(rows, dtype=np.float32)
        assert rows.shape == (self.n_vocab, row_width(self.cfg)), rows.shape
        check(load().ep_set_table(self.handle, ptr(rows, c_float)), "ep_set_table")

    def get_table(self) -> np.ndarray:
        out = np.empty((self.n_vocab, row_width(self.cfg)), dtype=np.float32)
        check(load().ep_get_table(self.handle, ptr(out, c_float)), "ep_get_table")
        return out

    def get_table_bf16(self) -> np.ndarray:
        out = np.empty((self.n_vocab, row_width(self.cfg)), dtype=np.uint16)
        check(load().ep_get_table_bf16(self.handle, ptr(out, c_uint16)), "ep_get_table_bf16")
        return out

    def get_adam(self):
        m = np.empty((self.n_vocab, row_width(self.cfg)), dtype=np.float32)
        v = np.empty_like(m)
        check(load().ep_get_adam(self.handle, ptr(m, c_float), ptr(v, c_float)), "ep_get_adam")
        return m, v

    def set_split(self, train: np.ndarray, valid: np.ndarray):
        train = np.ascontiguousarray(train, dtype=np.uint32)
        valid = np.ascontiguousarray(valid, dtype=np.uint32)
        check(load().ep_set_split(self.handle, ptr(train, c_uint32), len(train), ptr(valid, c_uint32), len(valid)),
              "ep_set_split")

    # -- training ----------------------------------------------------------------
    def learn(self, progress=None):
        if progress is None:
            cb = PROGRESS_FN(0)
        else:
            def _cb(_user, p, step, tr, va, lr):
                progress(int(p), int(step), float(tr), float(va), float(lr))
            cb = PROGRESS_FN(_cb)
        check(load().ep_learn(self.handle, cb, None), "ep_learn")

    def sample_batch(self, anchors: np.ndarray, pass_: int):
        anchors = np.ascontiguousarray(anchors, dtype=np.uint32)
        n = len(anchors)
        S, K = self.list_width(), self.cfg.negatives
        cnt = np.zeros(n, dtype=np.uint32)
        nbr = np.zeros((n, S), dtype=np.uint32)
        neg = np.zeros((n, K), dtype=np.uint32)
        check(load().ep_sample_batch(self.handle, ptr(anchors, c_uint32), n, int(pass_), ptr(cnt, c_uint32),
                                     ptr(nbr, c_uint32), ptr(neg, c_uint32)), "ep_sample_batch")
        return cnt, nbr, neg

    def step_explicit(self, anchors, nbr_cnt, nbr, neg, pass_: int, lr: float, apply_update: bool = True,
                      want_cons: bool = False):
        """One step on explicit samples.  Returns dict(loss, cons, feat, grad)."""
        anchors = np.ascontiguousarray(anchors, dtype=np.uint32)
        n = len(anchors)
        S, K, D, W = self.list_width(), self.cfg.negatives, self.cfg.dims, row_width(self.cfg)
        nbr_cnt = np.ascontiguousarray(nbr_cnt, dtype=np.uint32)
        nbr = np.ascontiguousarray(nbr, dtype=np.uint32).reshape(n, S)
        neg = np.ascontiguousarray(neg, dtype=np.uint32).reshape(n, K)
        loss = np.zeros(n, dtype=np.float32)
        cons = np.zeros((n, K + 2, D), dtype=np.float32) if want_cons else None
        nu = c_uint32(0)
        cap = self.n_vocab * max(1, int(self.cfg.concurrent_batches))  # one record per (row, batch)
        feat = np.zeros(cap, dtype=np.uint32)
        grad = np.zeros((cap, W), dtype=np.float32)
        check(load().ep_step_explicit(self.handle, ptr(anchors, c_uint32), n, ptr(nbr_cnt, c_uint32),
                                      ptr(nbr, c_uint32), ptr(neg, c_uint32), int(pass_), float(lr),
                                      1 if apply_update else 0, ptr(loss, c_float), ptr(cons, c_float),
                                      ctypes.byref(nu), ptr(feat, c_uint32), ptr(grad, c_float)),
              "ep_step_explicit")
        k = nu.value
        return {"loss": loss, "cons": cons, "feat": feat[:k].copy(), "grad": grad[:k].copy()}

    def step_explicit_noalloc(self, anchors, nbr_cnt, nbr, neg, pass_: int, lr: float) -> np.ndarray:
        """step_explicit with the update and only the loss output, through the entry point
        whose step arrays live in the library's per-GPU scratch (no device allocation in
        the call).  Returns the per-example loss."""
        anchors = np.ascontiguousarray(anchors, dtype=np.uint32)
        n = len(anchors)
        S, K = self.list_width(), self.cfg.negatives
        nbr_cnt = np.ascontiguousarray(nbr_cnt, dtype=np.uint32)
        nbr = np.ascontiguousarray(nbr, dtype=np.uint32).reshape(n, S)
        neg = np.ascontiguousarray(neg, dtype=np.uint32).reshape(n, K)
        loss = np.zeros(n, dtype=np.float32)
        check(load().ep_step_explicit_noalloc(self.handle, ptr(anchors, c_uint32), n, ptr(nbr_cnt, c_uint32),
                                              ptr(nbr, c_uint32), ptr(neg, c_uint32), int(pass_), float(lr),
                                              ptr(loss, c_float)), "ep_step_explicit_noalloc")
        return loss

    def bench_steps(self, warmup: int, steps: int, instrument: bool = True) -> EPBenchResult:
        out = EPBenchResult()
        check(load().ep_bench_steps(self.handle, int(warmup), int(steps), 1 if instrument else 0, ctypes.byref(out)),
              "ep_bench_steps")
        return out

    def embed_averaged(self, f_indptr: np.ndarray, f_ids: np.ndarray) -> np.ndarray:
        f_indptr = np.ascontiguousarray(f_indptr, dtype=np.uint64)
        f_ids = np.ascontiguousarray(f_ids, dtype=np.uint32)
        n = len(f_indptr) - 1
        out = np.empty((n, row_width(self.cfg)), dtype=np.float32)
        check(load().ep_embed_averaged(self.handle, n, ptr(f_indptr, c_uint64), ptr(f_ids, c_uint32),
                                       ptr(out, c_float)), "ep_embed_averaged")
        return out

    def embed_attention(self, f_indptr: np.ndarray, f_ids: np.ndarray) -> np.ndarray:
        f_indptr = np.ascontiguousarray(f_indptr, dtype=np.uint64)
        f_ids = np.ascontiguousarray(f_ids, dtype=np.uint32)
        n = len(f_indptr) - 1
        out = np.empty((n, self.cfg.dims), dtype=np.float32)
        check(load().ep_embed_attention(self.handle, n, ptr(f_indptr, c_uint64), ptr(f_ids, c_uint32),
                                        ptr(out, c_float)), "ep_embed_attention")
        return out

    def embed_weighted(self, f_indptr: np.ndarray, f_ids: np.ndarray, p_w: np.ndarray, alpha: float) -> np.ndarray:
        f_indptr = np.ascontiguousarray(f_indptr, dtype=np.uint64)
        f_ids = np.ascontiguousarray(f_ids, dtype=np.uint32)
        p_w = np.ascontiguousarray(p_w, dtype=np.float32)
        n = len(f_indptr) - 1
        out = np.empty((n, row_width(self.cfg)), dtype=np.float32)
        check(load().ep_embed_weighted(self.handle, n, ptr(f_indptr, c_uint64), ptr(f_ids, c_uint32), ptr(p_w, c_float),
                                       len(p_w), float(alpha), ptr(out, c_float)), "ep_embed_weighted")
        return out


def load_graph_tsv(path: str, undirected: bool, weighted: bool = True, deduplicate: bool = False,
                   skip_rows: int = 0):
    """The native graph file loader (GraphReader::load).  Returns (indptr u64, indices u32,
    per-row CDF f32, [(node_type, name)] in id order)."""
    lib = load()
    h = c_void_p()
    rc = lib.ep_graph_load_tsv(os.fsencode(path), int(bool(undirected)), int(bool(weighted)),
                               int(bool(deduplicate)), int(skip_rows), ctypes.byref(h))
    if rc != 0:
        raise ValueError(lib.ep_graph_file_error().decode(errors="replace"))
    try:
        n, e, nb = c_uint64(), c_uint64(), c_uint64()
        check(lib.ep_graph_file_sizes(h, ctypes.byref(n), ctypes.byref(e), ctypes.byref(nb)), "ep_graph_file_sizes")
        indptr = np.zeros(n.value + 1, dtype=np.uint64)
        indices = np.zeros(max(e.value, 1), dtype=np.uint32)
        cdf = np.zeros(max(e.value, 1), dtype=np.float32)
        names = ctypes.create_string_buffer(max(nb.value, 1))
        check(lib.ep_graph_file_copy(h, ptr(indptr, c_uint64), ptr(indices, c_uint32), ptr(cdf, c_float), names),
              "ep_graph_file_copy")
        text = names.raw[:nb.value].decode("utf-8", errors="surrogateescape")
        pairs = [tuple(line.split("\t", 1)) for line in text.split("\n")[:-1]]
        return indptr, indices[:e.value], cdf[:e.value], pairs
    finally:
        lib.ep_graph_file_free(h)


EP_FNS_STATIC, EP_FNS_NODE_TYPE, EP_FNS_PREFIX = 0, 1, 2


def _name_table(names):
    """[(type, name)] -> (blob bytes, offsets u64[2n + 1]) of the loaders' name tables."""
    parts = []
    for t, nm in names:
        parts.append(str(t).encode("utf-8", errors="surrogateescape"))
        parts.append(str(nm).encode("utf-8", errors="surrogateescape"))
    off = np.zeros(len(parts) + 1, dtype=np.uint64)
    if parts:
        np.cumsum(np.fromiter((len(x) for x in parts), dtype=np.uint64, count=len(parts)), out=off[1:])
    return b"".join(parts), off


def _split_names(blob: bytes, off: np.ndarray, n: int):
    o = off.tolist()
    if blob.isascii():  # byte offsets are character offsets: decode once
        t = blob.decode("ascii")
        return [(t[o[2 * k]:o[2 * k + 1]], t[o[2 * k + 1]:o[2 * k + 2]]) for k in range(n)]
    return [(blob[o[2 * k]:o[2 * k + 1]].decode("utf-8", errors="surrogateescape"),
             blob[o[2 * k + 1]:o[2 * k + 2]].decode("utf-8", errors="surrogateescape")) for k in range(n)]


def load_features_tsv(path: str, node_names, feature_names, ns_kind: int, ns_value: str = ""):
    """The native feature file loader (FeatureSet.load_into).  Returns (set nodes u32,
    indptr u64, feature ids u32, [(type, name)] of the features first seen in the file,
    whose ids follow the len(feature_names) known ones)."""
    lib = load()
    nb, noff = _name_table(node_names)
    fb, foff = _name_table(feature_names)
    h = c_void_p()
    rc = lib.ep_features_load_tsv(os.fsencode(path), nb, ptr(noff, c_uint64), len(node_names), fb,
                                  ptr(foff, c_uint64), len(feature_names), int(ns_kind),
                                  str(ns_value).encode("utf-8", errors="surrogateescape"), ctypes.byref(h))
    if rc != 0:
        raise ValueError(lib.ep_graph_file_error().decode(errors="replace"))
    try:
        ns_, ni, nn, nbytes = c_uint64(), c_uint64(), c_uint64(), c_uint64()
        check(lib.ep_feature_file_sizes(h, ctypes.byref(ns_), ctypes.byref(ni), ctypes.byref(nn),
                                        ctypes.byref(nbytes)), "ep_feature_file_sizes")
        nodes = np.zeros(max(ns_.value, 1), dtype=np.uint32)
        indptr = np.zeros(ns_.value + 1, dtype=np.uint64)
        ids = np.zeros(max(ni.value, 1), dtype=np.uint32)
        blob = ctypes.create_string_buffer(max(nbytes.value, 1))
        off = np.zeros(2 * nn.value + 1, dtype=np.uint64)
        check(lib.ep_feature_file_copy(h, ptr(nodes, c_uint32), ptr(indptr, c_uint64), ptr(ids, c_uint32), blob,
                                       ptr(off, c_uint64)), "ep_feature_file_copy")
        return nodes[:ns_.value], indptr, ids[:ni.value], _split_names(blob.raw[:nbytes.value], off, nn.value)
    finally:
        lib.ep_feature_file_free(h)


def load_embeddings_tsv(path: str, filter_types=None, skip_rows: int = 0):
    """The native embedding file loader (NodeEmbeddings.load).  Returns ([(type, name)] in
    file order, FP32 table rows x dims).  A repeated node raises KeyError, other
    malformed input ValueError."""
    lib = load()
    types = [] if filter_types is None else [str(t).encode("utf-8", errors="surrogateescape") for t in filter_types]
    toff = np.zeros(len(types) + 1, dtype=np.uint64)
    if types:
        np.cumsum(np.fromiter((len(t) for t in types), dtype=np.uint64, count=len(types)), out=toff[1:])
    h = c_void_p()
    rc = lib.ep_embeddings_load_tsv(os.fsencode(path), b"".join(types), ptr(toff, c_uint64), len(types),
                                    int(filter_types is not None), int(skip_rows), ctypes.byref(h))
    if rc != 0:
        msg = lib.ep_graph_file_error().decode(errors="replace")
        raise (KeyError if rc == EP_EDUPLICATE else ValueError)(msg)
    try:
        n, d, nb = c_uint64(), c_uint64(), c_uint64()
        check(lib.ep_embedding_file_sizes(h, ctypes.byref(n), ctypes.byref(d), ctypes.byref(nb)),
              "ep_embedding_file_sizes")
        table = np.zeros((n.value, d.value), dtype=np.float32)
        blob = ctypes.create_string_buffer(max(nb.value, 1))
        off = np.zeros(2 * n.value + 1, dtype=np.uint64)
        check(lib.ep_embedding_file_copy(h, ptr(table, c_float) if table.size else None, blob, ptr(off, c_uint64)),
              "ep_embedding_file_copy")
        return _split_names(blob.raw[:nb.value], off, n.value), table
    finally:
        lib.ep_embedding_file_free(h)


def write_embeddings_tsv(path: str, names, table: np.ndarray, comp_level=None):
    """The native embedding file writer (NodeEmbeddings.save): names[i] = (type, name) of
    row i of the FP32 table."""
    lib = load()
    table = np.ascontiguousarray(table, dtype=np.float32)
    n = table.shape[0]
    d = table.shape[1] if table.ndim == 2 else 0
    parts = [f"{t}\t{nm}".encode("utf-8", errors="surrogateescape") for t, nm in names]
    if len(parts) != n:
        raise ValueError("one (type, name) per table row")
    off = np.zeros(n + 1, dtype=np.uint64)
    if n:
        np.cumsum(np.fromiter((len(x) for x in parts), dtype=np.uint64, count=n), out=off[1:])
    rc = lib.ep_embeddings_write_tsv(os.fsencode(path), -1 if comp_level is None else int(comp_level), n, d,
                                     ptr(table, c_float) if table.size else None, b"".join(parts),
                                     ptr(off, c_uint64))
    if rc != 0:
        raise ValueError(lib.ep_graph_file_error().decode(errors="replace"))


def format_f32(x: float) -> str:
    """One f32 as the embedding writer prints it (ryu's layout)."""
    buf = ctypes.create_string_buffer(32)
    n = load().ep_format_f32(float(np.float32(x)), buf)
    return buf.raw[:n].decode()


def philox4x32(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    check(load().ep_philox4x32(ptr(c, c_uint32), ptr(k, c_uint32), ptr(out, c_uint32)), "ep_philox4x32")
    return out


def exported_symbols():
    return list(SIGNATURES.keys())
