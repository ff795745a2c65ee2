"""The allocation-free phase-1 entry point (cbp_phase1_keys_noalloc) against
cbp_phase1_keys through the C ABI: two calls in a row, the second with a
smaller input (a shorter query against a smaller uploaded chunk), plus a
two-stripe query (the stripe hand-off slice of the scratch array) and the
size check (an input beyond the scratch array is refused before any launch).
"""
import ctypes as C
import numpy as np
import pytest
from comer2_amd._ffi import SynthDB
from gpu_common import processing

pytestmark = pytest.mark.gpu


def _noalloc(cbp, qidx, nprof):
    L = cbp.L
    L.cbp_phase1_keys_noalloc.argtypes = [C.c_void_p, C.c_int32, np.ctypeslib.ndpointer(np.uint64, flags="C")]
    keys = np.zeros(nprof, np.uint64)
    rc = L.cbp_phase1_keys_noalloc(cbp.h, qidx, keys)
    return rc, keys


def test_noalloc_keys_equal_phase1_keys():
    q = SynthDB(2302, 3, 0, 0, 150, 150)                 # three queries of 150
    q2 = SynthDB(2305, 1, 0, 0, 700, 700)                # two stripes of 320 rows (R = 5)
    big = SynthDB(2303, 3000, 200, 0.668, 30, 3000)
    small = SynthDB(2304, 500, 150, 0.668, 30, 800)
    cbp = processing()
    # first call: a query of 150 against 3,000 profiles
    cbp.TransferQueries(q)
    cbp.TransferCPMData(big)
    print(f"noalloc: call 1, {big.nprof} profiles, query 150", flush=True)
    rc, k1 = _noalloc(cbp, 0, big.nprof)
    assert rc == 0, cbp.L.cbp_last_error(cbp.h)
    assert np.array_equal(k1, cbp.phase1_keys(0))
    assert np.count_nonzero(k1) > big.nprof // 2
    # second call, smaller input: another query against 500 profiles
    cbp.TransferCPMData(small)
    print(f"noalloc: call 2, {small.nprof} profiles", flush=True)
    rc, k2 = _noalloc(cbp, 2, small.nprof)
    assert rc == 0, cbp.L.cbp_last_error(cbp.h)
    assert np.array_equal(k2, cbp.phase1_keys(2))
    # a two-stripe query uses the stripe hand-off slice
    cbp.TransferQueries(q2)
    print("noalloc: two-stripe query of 700", flush=True)
    rc, k3 = _noalloc(cbp, 0, small.nprof)
    assert rc == 0, cbp.L.cbp_last_error(cbp.h)
    assert np.array_equal(k3, cbp.phase1_keys(0))


def test_noalloc_refuses_input_beyond_scratch():
    # 20,000 profiles (about 4.5 M positions) with a two-stripe query: the
    # stripe hand-off alone needs 20 B per position, beyond 64 MiB
    q = SynthDB(2402, 1, 0, 0, 700, 700)
    d = SynthDB(2403, 20000, 200, 0.668, 30, 3000)
    cbp = processing()
    cbp.TransferQueries(q)
    cbp.TransferCPMData(d)
    rc, _ = _noalloc(cbp, 0, d.nprof)
    assert rc == -1
    assert b"scratch" in cbp.L.cbp_last_error(cbp.h)
