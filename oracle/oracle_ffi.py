"""ctypes bindings to the oracle (oracle/liboracle.so).

Test infrastructure: imported only by tests/, the smoke check and the
benchmark's CPU-baseline leg, never by the product package.
"""
import ctypes as C
import os
import numpy as np
from comer2_amd._ffi import Profiles, Options

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "liboracle.so")


class Tables(C.Structure):
    _fields_ = [("ssss", (C.c_float * 30) * 30), ("nnw", C.c_float * 93),
                ("cvkeys", C.c_float * 128), ("cvvals", C.c_float * 128),
                ("cvcard", C.c_int32), ("cvshift", C.c_int32)]


class DpMax(C.Structure):
    _fields_ = [("maxsc", C.c_float), ("x", C.c_int32), ("y", C.c_int32)]


class Fin(C.Structure):
    _fields_ = [("score", C.c_float), ("psts", C.c_int32), ("begx", C.c_int32),
                ("begy", C.c_int32), ("endx", C.c_int32), ("endy", C.c_int32)]


class Aln(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("alnlen", "idts", "psts", "gaps", "begx", "begy", "endx", "endy")]


_lib = None


def lib():
    global _lib
    if _lib is None:
        L = C.CDLL(LIB)
        P = C.POINTER(Profiles)
        fp = np.ctypeslib.ndpointer(np.float32, flags="C")
        dp = np.ctypeslib.ndpointer(np.float64, flags="C")
        cp = np.ctypeslib.ndpointer(np.int8, flags="C")
        T = C.POINTER(Tables)
        O = C.POINTER(Options)
        L.or_score_f32.argtypes = [P, C.c_int, P, C.c_int, T, O, fp, fp]
        L.or_score_f64.argtypes = [P, C.c_int, P, C.c_int, T, O, dp, dp]
        L.or_colm1.argtypes = [P, C.c_int, fp, fp]
        L.or_viterbi_f32.argtypes = [P, C.c_int, P, C.c_int, fp, cp, fp]
        L.or_viterbi_f32.restype = DpMax
        L.or_viterbi_f64.argtypes = [P, C.c_int, P, C.c_int, dp]
        L.or_viterbi_f64.restype = C.c_double
        L.or_refemu_viterbi.argtypes = [P, C.c_int, P, C.c_int, fp, cp, fp]
        L.or_refemu_viterbi.restype = DpMax
        L.or_finalize.argtypes = [C.c_int, C.c_int, fp, cp, DpMax]
        L.or_finalize.restype = Fin
        L.or_mask_corners_f32.argtypes = [C.c_int, C.c_int, fp]
        L.or_aln_final.argtypes = [C.c_int, C.c_int, fp, cp, DpMax, C.c_char_p, C.c_char_p, C.c_void_p]
        L.or_aln_final.restype = Aln
        L.or_score_threshold.argtypes = [O, C.c_uint64, C.c_uint64,
                                         C.POINTER(C.c_float), C.POINTER(C.c_float)]
        _lib = L
    return _lib


def make_tables(ssss, nnw, cvmap=None):
    """cvmap: (keys, values, shift) of the context-vector map, or None"""
    t = Tables()
    arr = np.ctypeslib.as_array(t.ssss)
    arr[:] = ssss
    np.ctypeslib.as_array(t.nnw)[:] = nnw
    if cvmap is not None:
        keys, vals, shift = cvmap
        n = len(keys)
        np.ctypeslib.as_array(t.cvkeys)[:n] = keys
        np.ctypeslib.as_array(t.cvvals)[:n] = vals
        t.cvcard = n
        t.cvshift = shift
    return t


class Oracle:
    def __init__(self, ssss, nnw, opts, cvmap=None):
        self.L = lib()
        self.t = make_tables(ssss, nnw, cvmap)
        self.o = Options()
        for k, _ in Options._fields_:
            setattr(self.o, k, opts.get(k, 0.0) if k == "cvswgt" else opts[k])

    def score_f32(self, q, qp, d, dp):
        Lq, Ld = q.len[qp], d.len[dp]
        S = np.empty((Lq, Ld), np.float32); M = np.empty((Lq, Ld), np.float32)
        self.L.or_score_f32(C.byref(q), qp, C.byref(d), dp, C.byref(self.t), C.byref(self.o), S, M)
        return S, M

    def score_f64(self, q, qp, d, dp):
        Lq, Ld = q.len[qp], d.len[dp]
        S = np.empty((Lq, Ld), np.float64); M = np.empty((Lq, Ld), np.float64)
        self.L.or_score_f64(C.byref(q), qp, C.byref(d), dp, C.byref(self.t), C.byref(self.o), S, M)
        return S, M

    def viterbi_f32(self, q, qp, d, dp, S, want_btck=True, want_mm=False):
        Lq, Ld = q.len[qp], d.len[dp]
        bt = np.zeros((Lq, Ld), np.int8); mm = np.zeros((Lq, Ld), np.float32)
        r = self.L.or_viterbi_f32(C.byref(q), qp, C.byref(d), dp, np.ascontiguousarray(S), bt, mm)
        return r, bt, mm

    def refemu(self, q, qp, d, dp, S):
        Lq, Ld = q.len[qp], d.len[dp]
        bt = np.zeros((Lq, Ld), np.int8); mm = np.zeros((Lq, Ld), np.float32)
        r = self.L.or_refemu_viterbi(C.byref(q), qp, C.byref(d), dp, np.ascontiguousarray(S), bt, mm)
        return r, bt, mm

    def viterbi_f64(self, q, qp, d, dp, S):
        return self.L.or_viterbi_f64(C.byref(q), qp, C.byref(d), dp, np.ascontiguousarray(S))

    def finalize(self, S, bt, m):
        Lq, Ld = S.shape
        return self.L.or_finalize(Lq, Ld, np.ascontiguousarray(S), bt, m)

    def aln_final(self, S, bt, m, qaa, daa):
        """MAPALN=0 drawing of a pair from its phase-1 backtrace: (record, (query, middle, target))"""
        Lq, Ld = S.shape
        buf = np.zeros((3, Lq + Ld + 1), np.uint8)
        r = self.L.or_aln_final(Lq, Ld, np.ascontiguousarray(S), bt, m, qaa, daa, buf.ctypes.data)
        return r, tuple(bytes(buf[j]).split(b"\0")[0].decode() for j in range(3))

    def colm1(self, q, qp):
        Lq = q.len[qp]
        a = np.zeros((Lq, 5), np.float32); b = np.zeros((Lq, 5), np.float32)
        self.L.or_colm1(C.byref(q), qp, a, b)
        return a, b

    def score_threshold(self, qlen, dblen):
        s = C.c_float(); l = C.c_float()
        self.L.or_score_threshold(C.byref(self.o), qlen, dblen, C.byref(s), C.byref(l))
        return s.value, l.value


class Stats(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("Min", "Max", "E", "H", "Lmbd", "K",
                                          "LmbdEst", "KEst", "Evalue", "EpA")]


def _ss_sigs(L):
    fp = np.ctypeslib.ndpointer(np.float32, flags="C")
    i64 = np.ctypeslib.ndpointer(np.int64, flags="C")
    L.or_histogram.argtypes = [fp, C.c_int, C.c_int, C.c_char_p, C.c_char_p, C.c_int, i64]
    L.or_statistics.argtypes = [i64, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float,
                                C.c_float, C.POINTER(Options), C.POINTER(C.c_float)]
    L.or_statistics.restype = Stats
    L.or_length_adjustment.argtypes = [C.POINTER(Options), C.c_uint64, C.c_uint64, C.c_uint64,
                                       C.POINTER(C.c_uint64)]
    L.or_length_adjustment.restype = C.c_uint64
    for n in ("or_lbetainc",):
        getattr(L, n).argtypes = [C.c_double, C.c_double, C.c_double]
        getattr(L, n).restype = C.c_double
    for n in ("or_lgammainc_Q", "or_lgammainc_P"):
        getattr(L, n).argtypes = [C.c_double, C.c_double]
        getattr(L, n).restype = C.c_double
    L.or_set_flags.argtypes = [C.c_int]


def _oracle_histogram(self, S, qaa, daa, xuninf=1):
    Lq, Ld = S.shape
    cnt = np.zeros(641, np.int64)
    self.L.or_histogram(np.ascontiguousarray(S, np.float32), Lq, Ld, qaa, daa, xuninf, cnt)
    return cnt


def _oracle_statistics(self, counts, Lq, Ld, qeno, deno, score, psts, searchspace):
    nn = np.ascontiguousarray(np.ctypeslib.as_array(self.t.nnw), np.float32)
    return self.L.or_statistics(np.ascontiguousarray(counts, np.int64), Lq, Ld, qeno, deno, score,
                                psts, searchspace, C.byref(self.o),
                                nn.ctypes.data_as(C.POINTER(C.c_float)))


def _oracle_length_adjustment(self, m, n, N):
    d = C.c_uint64()
    ss = self.L.or_length_adjustment(C.byref(self.o), m, n, N, C.byref(d))
    return ss, d.value


Oracle.histogram = _oracle_histogram
Oracle.statistics = _oracle_statistics
Oracle.length_adjustment = _oracle_length_adjustment
_orig_lib = lib


def lib():  # noqa: F811
    L = _orig_lib()
    if not getattr(L, "_ss_ready", False):
        _ss_sigs(L)
        L._ss_ready = True
    return L


class MapRes(C.Structure):
    _fields_ = [("logZ", C.c_float), ("maxF", C.c_float), ("maxsc", C.c_float),
                ("x", C.c_int32), ("y", C.c_int32), ("alnlen", C.c_int32), ("idts", C.c_int32),
                ("psts", C.c_int32), ("gaps", C.c_int32), ("begx", C.c_int32), ("begy", C.c_int32)]


def _map_sigs(L):
    P = C.POINTER(Profiles)
    fp = np.ctypeslib.ndpointer(np.float32, flags="C")
    dp = np.ctypeslib.ndpointer(np.float64, flags="C")
    L.or_map_refemu.argtypes = [P, C.c_int, P, C.c_int, fp, fp, C.c_float, C.c_float, C.c_int,
                                C.c_void_p, C.c_void_p, C.c_void_p]
    L.or_map_refemu.restype = MapRes
    L.or_map_direct.argtypes = [P, C.c_int, P, C.c_int, fp, fp, C.c_float, C.c_float, C.c_int,
                                C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.or_map_direct.restype = MapRes
    L.or_map_f64.argtypes = [P, C.c_int, P, C.c_int, dp, dp, C.c_float, C.c_int, C.c_void_p]
    L.or_map_f64.restype = C.c_double
    L.or_map_dynparams.argtypes = [C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float)]


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _oracle_map(self, q, qp, d, dp, S, M, epa, direct=True, want=True):
    L = self.L
    if not getattr(L, "_map_ready", False):
        _map_sigs(L)
        L._map_ready = True
    Lq, Ld = q.len[qp], d.len[dp]
    F = np.zeros((Lq, Ld), np.float32) if want else None
    FB = np.zeros((Lq, Ld), np.float32) if want else None
    bt = np.zeros((Lq, Ld), np.int8) if want else None
    # the mod scores are in use with either term (GetModScoreMatrixInUse,
    # cupro/CuDeviceMemory.cuh:142-144)
    modin = 1 if (self.o.ssswgt > 0 or self.o.cvswgt > 0) else 0
    if direct:
        aln = np.zeros((3, Lq + Ld + 1), np.uint8) if want else None
        r = L.or_map_direct(C.byref(q), qp, C.byref(d), dp, np.ascontiguousarray(S, np.float32),
                            np.ascontiguousarray(M, np.float32), epa, self.o.minpp, modin,
                            _ptr(F), _ptr(FB), _ptr(bt), _ptr(aln))
        return r, F, FB, bt, aln
    r = L.or_map_refemu(C.byref(q), qp, C.byref(d), dp, np.ascontiguousarray(S, np.float32),
                        np.ascontiguousarray(M, np.float32), epa, self.o.minpp, modin,
                        _ptr(F), _ptr(FB), _ptr(bt))
    return r, F, FB, bt, None


def _oracle_map64(self, q, qp, d, dp, S64, M64, epa):
    L = self.L
    if not getattr(L, "_map_ready", False):
        _map_sigs(L)
        L._map_ready = True
    post = np.zeros((q.len[qp], d.len[dp]), np.float64)
    # the mod scores are in use with either term (GetModScoreMatrixInUse,
    # cupro/CuDeviceMemory.cuh:142-144)
    modin = 1 if (self.o.ssswgt > 0 or self.o.cvswgt > 0) else 0
    lz = L.or_map_f64(C.byref(q), qp, C.byref(d), dp, np.ascontiguousarray(S64), np.ascontiguousarray(M64),
                      epa, modin, _ptr(post))
    return lz, post


Oracle.map = _oracle_map
Oracle.map64 = _oracle_map64


class SearchOut(C.Structure):
    _fields_ = [("cells", C.c_int64), ("map_cells", C.c_int64), ("npassed", C.c_int32),
                ("nmap", C.c_int32), ("checksum", C.c_double)]


def _oracle_search(self, q, qp, d, pros, scorethld, logevthld, searchspace, nthreads):
    """Whole pipeline for query qp against the listed DB profiles (OpenMP)."""
    L = self.L
    if not getattr(L, "_search_ready", False):
        _ss_sigs(L)
        L._ss_ready = True
        L.or_search.argtypes = [C.POINTER(Profiles), C.c_int, C.POINTER(Profiles),
                                np.ctypeslib.ndpointer(np.int32, flags="C"), C.c_int32,
                                C.POINTER(Tables), C.POINTER(Options), C.c_float, C.c_float,
                                C.c_float, C.c_int]
        L.or_search.restype = SearchOut
        L._search_ready = True
    pros = np.ascontiguousarray(pros, np.int32)
    return L.or_search(C.byref(q), qp, C.byref(d), pros, pros.size, C.byref(self.t), C.byref(self.o),
                       scorethld, logevthld, searchspace, nthreads)


Oracle.search = _oracle_search


def _sdp_sigs(L):
    P = C.POINTER(Profiles)
    dp = np.ctypeslib.ndpointer(np.float64, flags="C")
    fp = np.ctypeslib.ndpointer(np.float32, flags="C")
    L.or_map_colm1.argtypes = [P, C.c_int, fp, fp]
    L.or_sdp_logz.argtypes = [P, C.c_int, P, C.c_int, dp, dp, fp, fp]
    L.or_sdp_logz.restype = C.c_double
    L.or_sdp_grad.argtypes = [P, C.c_int, P, C.c_int, dp, dp, fp, fp, dp, dp, C.c_void_p]
    L.or_sdp_grad.restype = C.c_double
    L._sdp_ready = True


def _oracle_map_colm1(self, q, qp):
    if not getattr(self.L, "_sdp_ready", False):
        _sdp_sigs(self.L)
    Lq = q.len[qp]
    a = np.zeros((Lq, 5), np.float32); b = np.zeros((Lq, 5), np.float32)
    self.L.or_map_colm1(C.byref(q), qp, a, b)
    return a, b


def _oracle_sdp_logz(self, q, qp, d, dp, S, theta, bnd=None):
    """FP64 soft-DP log Z (natural log) of S [Lq][Ld] and theta[5]."""
    if not getattr(self.L, "_sdp_ready", False):
        _sdp_sigs(self.L)
    ll, ld = bnd if bnd is not None else self.map_colm1(q, qp)
    return self.L.or_sdp_logz(C.byref(q), qp, C.byref(d), dp, np.ascontiguousarray(S, np.float64),
                              np.ascontiguousarray(theta, np.float64), ll, ld)


def _oracle_sdp_grad(self, q, qp, d, dp, S, theta, bnd=None):
    """(logZ, dS, dtheta, posterior) of the FP64 soft DP."""
    if not getattr(self.L, "_sdp_ready", False):
        _sdp_sigs(self.L)
    ll, ld = bnd if bnd is not None else self.map_colm1(q, qp)
    S = np.ascontiguousarray(S, np.float64)
    dS = np.zeros_like(S); dth = np.zeros(5); post = np.zeros_like(S)
    z = self.L.or_sdp_grad(C.byref(q), qp, C.byref(d), dp, S, np.ascontiguousarray(theta, np.float64),
                           ll, ld, dS, dth, post.ctypes.data_as(C.c_void_p))
    return z, dS, dth, post


Oracle.map_colm1 = _oracle_map_colm1
Oracle.sdp_logz = _oracle_sdp_logz
Oracle.sdp_grad = _oracle_sdp_grad
