"""C-ABI checks that need no GPU: every entry point declared in
include/comer2_hip.h is exported by the built library, argument errors come
back as error codes, the host-side threshold helper matches the oracle's
restatement of JobDispatcher::CalculateScoreThreshold bit for bit, and the
phase-1 key encoding round-trips."""
import ctypes as C
import os
import re
import numpy as np
import pytest
from comer2_amd import _ffi, tables
from comer2_amd.host import decode_keys, encode_key
from gpu_common import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    inc = os.path.join(ROOT, "include")
    for f in os.listdir(inc):
        if f.endswith(".h"):
            txt = open(os.path.join(inc, f)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(cbp_[a-z0-9_]+)\s*\(", txt))
    return sorted(names)


def test_every_declared_symbol_is_exported():
    L = _ffi.load_hip()
    names = _declared()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_argument_errors_are_codes():
    L = _ffi.load_hip()
    L.cbp_destroy.argtypes = [C.c_void_p]
    L.cbp_last_error.restype = C.c_char_p
    L.cbp_last_error.argtypes = [C.c_void_p]
    assert L.cbp_destroy(None) == -1
    assert L.cbp_last_error(None) == b"null handle"
    L.cbp_create.argtypes = [C.c_int, C.c_size_t, C.c_int, C.c_void_p]
    assert L.cbp_create(0, 0, 0, None) == -1
    L.cbp_score_threshold.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_void_p, C.c_void_p]
    assert L.cbp_score_threshold(None, 300, 1000, None, None) == -1


def test_create_without_gpu_reports_nodev():
    import torch
    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is present")
    L = _ffi.load_hip()
    h = C.c_void_p()
    L.cbp_create.argtypes = [C.c_int, C.c_size_t, C.c_int, C.c_void_p]
    assert L.cbp_create(0, 0, 0, C.byref(h)) == -5
    assert not h.value


@pytest.mark.parametrize("evalue", [10.0, 1.0, 1e-3, 0.0])
def test_score_threshold_matches_oracle(evalue):
    L = _ffi.load_hip()
    L.cbp_score_threshold.argtypes = [C.POINTER(_ffi.Options), C.c_uint64, C.c_uint64,
                                      C.POINTER(C.c_float), C.POINTER(C.c_float)]
    opts = dict(tables.default_options(), evalue=evalue)
    o = _ffi.options_struct(opts)
    orc = oracle(dict(evalue=evalue))
    for qlen in (30, 150, 300, 1000, 5000):
        for dblen in (450_000, 50_000_000, 10**10):
            s, l = C.c_float(), C.c_float()
            assert L.cbp_score_threshold(C.byref(o), qlen, dblen, C.byref(s), C.byref(l)) == 0
            es, el = orc.score_threshold(qlen, dblen)
            assert np.float32(s.value).tobytes() == np.float32(es).tobytes()
            assert np.float32(l.value).tobytes() == np.float32(el).tobytes()


def test_key_encoding_round_trip():
    rng = np.random.default_rng(5)
    sc = rng.uniform(0.01, 300, 200).astype(np.float32)
    x = rng.integers(0, 65535, 200)
    y = rng.integers(0, 65535, 200)
    keys = np.array([encode_key(s, a, b) for s, a, b in zip(sc, x, y)], np.uint64)
    s2, x2, y2 = decode_keys(keys)
    assert np.array_equal(s2, sc) and np.array_equal(x2, x) and np.array_equal(y2, y)
    # larger score wins; on equal score the smaller x, then the smaller y
    assert encode_key(2.0, 5, 5) > encode_key(1.0, 0, 0)
    assert encode_key(2.0, 4, 9) > encode_key(2.0, 5, 0)
    assert encode_key(2.0, 5, 3) > encode_key(2.0, 5, 4)
    assert encode_key(0.0, 1, 1) == 0
