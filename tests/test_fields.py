"""The reference's per-field profile buffers through the C ABI, without a GPU.

A profile set is laid out the way the reference hands it to
CuBatchProcessing::TransferCPMData -- one packed host buffer, field after
field at 32-byte aligned offsets, begin/end pointers per field and the
offsets in szCpm2dvfields (cupro/PM2DVectorFields.h:41-81,
cupro/CuDeviceMemory.cu:1036-1104) -- by the tests' driver of the C++
drop-in class; cbp_fields_count / cbp_fields_pack must return the set's
field-major arrays exactly, for a chunk, for a query whose pointers are
advanced into its batch (PMDataSkipNPros) and with the optional CV, HDP and
address fields present or empty.
"""
import ctypes as C
import os

import numpy as np
import pytest

from comer2_amd import _ffi
from comer2_amd._ffi import SynthDB

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NF = 81
PP = C.c_void_p * NF


def _libs():
    L = _ffi.load_hip()
    T = C.CDLL(os.path.join(ROOT, "tests", "libadapter_check.so"))
    T.c2t_layout.restype = C.c_void_p
    T.c2t_layout.argtypes = [C.POINTER(_ffi.Profiles), C.c_int, C.POINTER(C.c_size_t), PP, PP]
    T.c2t_free.argtypes = [C.c_void_p]
    T.c2t_query_fields.argtypes = [C.POINTER(_ffi.Profiles), PP, C.c_int, PP]
    L.cbp_fields_count.argtypes = [PP, PP, C.POINTER(C.c_int32), C.POINTER(C.c_int64)]
    L.cbp_fields_pack.argtypes = [PP, PP] + [C.c_void_p] * 9
    return L, T


def _arrays(view):
    np_, ns = view.nprof, view.npos
    a = lambda p, n, t: np.ctypeslib.as_array(C.cast(p, C.POINTER(t)), (n,)).copy()
    return dict(bkg=a(view.bkg, 20 * np_, C.c_float), eno=a(view.eno, np_, C.c_float),
                len=a(view.len, np_, C.c_int32), dist=a(view.dist, np_, C.c_uint32),
                trgfrq=a(view.trgfrq, 20 * ns, C.c_float), ssprb=a(view.ssprb, 3 * ns, C.c_float),
                aa=a(view.aa, ns, C.c_char), sss=a(view.sss, ns, C.c_char),
                trans=a(view.trans, 7 * (ns + np_), C.c_float))


def _pack(L, beg, end):
    n, s = C.c_int32(), C.c_int64()
    rc = L.cbp_fields_count(beg, end, C.byref(n), C.byref(s))
    if rc:
        return rc, None
    np_, ns = n.value, s.value
    out = dict(bkg=np.zeros(20 * np_, np.float32), eno=np.zeros(np_, np.float32),
               len=np.zeros(np_, np.int32), dist=np.zeros(np_, np.uint32),
               trgfrq=np.zeros(20 * ns, np.float32), ssprb=np.zeros(3 * ns, np.float32),
               aa=np.zeros(ns, np.uint8), sss=np.zeros(ns, np.uint8),
               trans=np.zeros(7 * (ns + np_), np.float32))
    order = ["bkg", "eno", "len", "dist", "trgfrq", "ssprb", "aa", "sss", "trans"]
    rc = L.cbp_fields_pack(beg, end, *[out[k].ctypes.data_as(C.c_void_p) for k in order])
    return rc, out


def _same(got, want):
    for k, v in want.items():
        g = got[k]
        if v.dtype.kind == "S":
            v = v.view(np.uint8)
        assert np.array_equal(g.view(np.uint8) if g.dtype != v.dtype else g, v), k


@pytest.mark.parametrize("with_optional", [1, 0])
def test_chunk_fields_pack_to_profile_set(with_optional):
    L, T = _libs()
    d = SynthDB(61, 40, 120, 0.668, 30, 400)
    sz = (C.c_size_t * (NF + 1))()
    beg, end = PP(), PP()
    buf = T.c2t_layout(C.byref(d.view), with_optional, sz, beg, end)
    try:
        # 32-byte aligned field offsets, total at index pmv2DTotFlds
        offs = [sz[f] for f in range(NF + 1)]
        assert all(o % 32 == 0 for o in offs) and offs == sorted(offs)
        assert all(beg[f] == buf + offs[f] for f in range(NF))
        rc, got = _pack(L, beg, end)
        assert rc == 0
        _same(got, _arrays(d.view))
    finally:
        T.c2t_free(buf)


def test_query_fields_advanced_into_batch():
    """querypmbeg advanced to query q (PMDataSkipNPros): the one-profile set
    read from those pointers and the query length is query q, its distance
    re-based to 0."""
    L, T = _libs()
    qs = SynthDB(62, 5, 150, 0.5, 40, 300)
    sz = (C.c_size_t * (NF + 1))()
    beg, end = PP(), PP()
    buf = T.c2t_layout(C.byref(qs.view), 0, sz, beg, end)
    try:
        full = _arrays(qs.view)
        lens = qs.lengths()
        for q in range(qs.nprof):
            qpm = PP()
            T.c2t_query_fields(C.byref(qs.view), beg, q, qpm)
            n = int(lens[q])
            qend = PP()
            for f in range(NF):
                if f < 23:
                    qend[f] = qpm[f] + 4
                elif f < 30:
                    qend[f] = qpm[f] + 4 * (n + 1)
                elif f in (79, 80):
                    qend[f] = qpm[f] + n
                elif (50 <= f < 71) or (74 <= f <= 78):
                    qend[f] = qpm[f]              # empty optional field
                else:
                    qend[f] = qpm[f] + 4 * n
            rc, got = _pack(L, qpm, qend)
            assert rc == 0
            d0 = int(qs.dists()[q])
            assert got["len"][0] == n and got["dist"][0] == 0
            assert np.array_equal(got["bkg"], full["bkg"].reshape(20, -1)[:, q])
            assert np.array_equal(got["trgfrq"].reshape(20, n), full["trgfrq"].reshape(20, -1)[:, d0:d0 + n])
            ntr = qs.npos + qs.nprof
            assert np.array_equal(got["trans"].reshape(7, n + 1),
                                  full["trans"].reshape(7, ntr)[:, d0 + q:d0 + q + n + 1])
            assert bytes(got["aa"]) == bytes(full["aa"][d0:d0 + n])
    finally:
        T.c2t_free(buf)


def test_inconsistent_fields_are_rejected():
    L, T = _libs()
    d = SynthDB(63, 10, 100, 0.5, 30, 200)
    sz = (C.c_size_t * (NF + 1))()
    beg, end = PP(), PP()
    buf = T.c2t_layout(C.byref(d.view), 1, sz, beg, end)
    try:
        assert _pack(L, beg, end)[0] == 0
        keep = end[30]
        end[30] = end[30] - 4                   # a target-frequency field one entry short
        assert _pack(L, beg, end)[0] == -1
        end[30] = keep
        keep = end[72]
        end[72] = beg[72]                       # a required SS field left empty
        assert _pack(L, beg, end)[0] == -1
        end[72] = keep
        dist = np.ctypeslib.as_array(C.cast(beg[22], C.POINTER(C.c_uint32)), (d.nprof,))
        dist[3] += 1                            # distances that do not follow the lengths
        assert _pack(L, beg, end)[0] == -1
        dist[3] -= 1
        assert _pack(L, beg, end)[0] == 0
    finally:
        T.c2t_free(buf)
