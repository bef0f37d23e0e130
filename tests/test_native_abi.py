"""CPU-only checks of the native boundary: the library loads, exports every symbol the
public header declares, validates configurations, and its host Philox matches the
published Random123 known answers.  No GPU compute is called."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from cloverleaf_amd import _native as N

ROOT = Path(__file__).resolve().parents[1]


def header_functions():
    text = (ROOT / "include" / "cloverleaf_ep.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ep_[a-z0-9_]+)\s*\(", text)) - {"ep_progress_fn"})


def test_library_exports_every_header_symbol(native):
    funcs = header_functions()
    assert len(funcs) >= 15
    for name in funcs:
        assert hasattr(native, name), name
    assert set(funcs) == set(N.SIGNATURES), set(funcs) ^ set(N.SIGNATURES)


def test_config_defaults_match_reference():
    # EmbeddingPropagator::new defaults (src/lib.rs:1262-1274, SEED lib.rs:90)
    c = N.default_config()
    assert (c.dims, c.batch_size, c.passes, c.negatives) == (100, 50, 100, 1)
    assert abs(c.alpha - 0.9) < 1e-7 and c.gamma == 1.0 and abs(c.valid_pct - 0.1) < 1e-7
    assert c.seed == 20222022
    assert c.max_nodes_kind == N.EP_SAMPLE_ALL and c.max_features_kind == N.EP_SAMPLE_ALL


@pytest.mark.parametrize("field,value", [("dims", 7), ("negatives", 0), ("batch_size", 0),
                                         ("noise", -0.5), ("noise", float("nan")), ("max_features_kind", 1),
                                         ("valid_pct", 1.5), ("concurrent_batches", 3), ("seg_vpl", 5),
                                         ("loss_kind", 6), ("loss_kind", -1), ("gamma", float("inf"))])
def test_invalid_config_rejected(field, value):
    c = N.default_config()
    setattr(c, field, value)
    h = ctypes.c_void_p()
    rc = N.load().ep_create(ctypes.byref(c), ctypes.byref(h))
    assert rc == -1 and not h.value
    assert N.load().ep_last_error()


@pytest.mark.parametrize("fields", [
    dict(attention_dk=5, attention_heads=4, dims=128),            # d_k not built
    dict(attention_dk=8, attention_heads=2, dims=12),             # d_model not a multiple of 8
    dict(attention_dk=8, attention_heads=0, dims=16),             # no heads
    dict(attention_dk=8, attention_heads=3, dims=16),             # row width 96 not built
    dict(attention_dk=8, attention_heads=2, dims=16, scatter=1),  # attention needs the sort scatter
])
def test_invalid_attention_config_rejected(fields):
    c = N.default_config()
    for k, v in fields.items():
        setattr(c, k, v)
    h = ctypes.c_void_p()
    rc = N.load().ep_create(ctypes.byref(c), ctypes.byref(h))
    assert rc == -1 and not h.value
    assert N.load().ep_last_error()


def test_ppr_restart_probability_checked():
    for r in (-0.1, 1.5, float("nan")):
        c = N.default_config()
        c.loss_kind, c.loss_param2 = N.EP_LOSS_PPR, r
        h = ctypes.c_void_p()
        assert N.load().ep_create(ctypes.byref(c), ctypes.byref(h)) == -1 and not h.value
        assert b"restart" in N.load().ep_last_error() or b"finite" in N.load().ep_last_error()


def test_attention_row_width():
    c = N.default_config()
    c.dims, c.attention_dk, c.attention_heads = 128, 32, 4
    assert N.row_width(c) == 4 * (2 * 32 + 128) == 768  # model.rs:241-243
    c.attention_dk = 0
    assert N.row_width(c) == 128


# Random123 kat_vectors for philox4x32 with 10 rounds
PHILOX_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,exp", PHILOX_KAT)
def test_philox_known_answers(ctr, key, exp):
    assert tuple(int(x) for x in N.philox4x32(ctr, key)) == exp


def test_null_arguments_rejected(native):
    assert native.ep_create(None, None) == -1
    assert native.ep_set_graph(None, 0, None, None) == -1


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        N.load(tmp_path / "nope.so")


def test_aligner_rejects_bad_input():
    ip = np.array([0, 1, 2], dtype=np.uint64)
    ix = np.array([1, 0], dtype=np.uint32)
    cdf = np.ones(2, dtype=np.float32)
    with pytest.raises(N.NativeError, match="width"):
        N.Aligner(ip, ix, cdf, np.zeros((2, 7), np.float32))
    with pytest.raises(N.NativeError, match="out of range"):
        N.Aligner(ip, np.array([1, 5], dtype=np.uint32), cdf, np.zeros((2, 8), np.float32))
    with pytest.raises(ValueError):
        N.Aligner(ip, ix, cdf, np.zeros((3, 8), np.float32))
