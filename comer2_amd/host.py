"""Python mirror of the reference's CuBatchProcessing interface over the C ABI.

The method set follows cupro/CuBatchProcessing.cuh:50-112
(TransferCPMData, SetDbDetails, SetQueryLen, SetScoreThreshold,
SetLogEThreshold, ProcessScoreMatrix); every call goes to
``libcomer2_hip.so`` through the functions declared in include/comer2_hip.h.
There is no CPU fallback: constructing the object without the library or a
GPU raises.
"""
import ctypes as C
import os
import numpy as np

from . import _ffi
from ._ffi import Profiles, Options, Results

_lib = None


class CbpError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_ffi.LIB_HIP):
        raise CbpError("libcomer2_hip.so is not built (run comer2_amd.build)")
    L = _ffi.load_hip()
    H = C.c_void_p
    L.cbp_create.argtypes = [C.c_int, C.c_size_t, C.c_int, C.POINTER(H)]
    L.cbp_destroy.argtypes = [H]
    L.cbp_last_error.argtypes = [H]
    L.cbp_last_error.restype = C.c_char_p
    L.cbp_set_options.argtypes = [H, C.POINTER(Options)]
    L.cbp_upload_tables.argtypes = [H, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int32]
    L.cbp_upload_cvs2s.argtypes = [H, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int32, C.c_int32]
    L.cbp_upload_queries.argtypes = [H, C.POINTER(Profiles)]
    L.cbp_upload_chunk.argtypes = [H, C.POINTER(Profiles)]
    L.cbp_set_db_details.argtypes = [H, C.c_uint64, C.c_uint64]
    L.cbp_score_threshold.argtypes = [C.POINTER(Options), C.c_uint64, C.c_uint64,
                                      C.POINTER(C.c_float), C.POINTER(C.c_float)]
    L.cbp_phase1_keys.argtypes = [H, C.c_int32, np.ctypeslib.ndpointer(np.uint64, flags="C")]
    L.cbp_last_kernel_ms.argtypes = [H, C.POINTER(C.c_double), C.POINTER(C.c_int32)]
    L.cbp_set_tuning.argtypes = [H, C.c_int32, C.c_int32, C.c_int32]
    L.cbp_set_graph.argtypes = [H, C.c_int32]
    L.cbp_set_queue.argtypes = [H, C.c_int32]
    L.cbp_set_mfma.argtypes = [H, C.c_int32]
    L.cbp_set_map_rows.argtypes = [H, C.c_int32]
    L.cbp_set_map_pipe.argtypes = [H, C.c_int32]
    L.cbp_set_device_lists.argtypes = [H, C.c_int32]
    L.cbp_set_device_list_caps.argtypes = [H, C.c_int32, C.c_int32, C.c_int64, C.c_int64]
    L.cbp_last_fallbacks.argtypes = [H, C.POINTER(C.c_int32)]
    L.cbp_chunk_limits.argtypes = [H, C.c_uint64] + [C.POINTER(C.c_uint64)] * 4
    L.cbp_autotune.argtypes = [H, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
    L.cbp_process.argtypes = [H, C.c_int32, C.c_float, C.c_float, C.POINTER(Results)]
    L.cbp_process_batch.argtypes = [H, C.c_void_p, C.c_int32, C.POINTER(Results)]
    L.cbp_score_block.argtypes = [H, C.c_int32, C.c_int32, C.POINTER(C.c_float), C.POINTER(C.c_float)]
    L.cbp_search_db.argtypes = [H, C.POINTER(Profiles), C.c_uint64, C.c_void_p, C.c_int32, C.POINTER(Results)]
    L.cbp_last_search_chunks.argtypes = [H, C.POINTER(C.c_int32)]
    V = C.c_void_p
    L.cbp_sdp_prepare.argtypes = [H, C.POINTER(_ffi.SdpPair), C.c_int32, C.c_int32, C.POINTER(C.c_int64)]
    L.cbp_sdp_scores.argtypes = [H, V, V]
    L.cbp_sdp_forward.argtypes = [H, V, V, V, V, V]
    L.cbp_sdp_adjoint.argtypes = [H, V, V, V, V, V, V, V, V, V]
    _lib = L
    return L


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class CuBatchProcessing:
    """One device worker (the reference constructs one per DevCommThread)."""

    def __init__(self, device=0, mem_budget=0, area=0, options=None, ssss=None, nnw=None, cvmap=None):
        from . import tables
        self.L = lib()
        self.h = C.c_void_p()
        rc = self.L.cbp_create(device, mem_budget, area, C.byref(self.h))
        if rc != 0:
            raise CbpError("cbp_create failed (%d): no usable GPU or HIP error" % rc)
        self.opts = dict(tables.default_options(), **(options or {}))
        self._o = _ffi.options_struct(self.opts)
        self._check(self.L.cbp_set_options(self.h, C.byref(self._o)))
        ss = np.ascontiguousarray(tables.load_ssss() if ssss is None else ssss, np.float32)
        nw = np.ascontiguousarray(tables.load_nn_weights() if nnw is None else nnw, np.float32)
        self._check(self.L.cbp_upload_tables(self.h, _fptr(ss), _fptr(nw), nw.size))
        keys, vals, shift = tables.load_cvs2s() if cvmap is None else cvmap
        self._cvk = np.ascontiguousarray(keys, np.float32)
        self._cvv = np.ascontiguousarray(vals, np.float32)
        self._check(self.L.cbp_upload_cvs2s(self.h, _fptr(self._cvk), _fptr(self._cvv), self._cvk.size, int(shift)))
        self.scorethld = None
        self.logevthld = None
        self.dblen = 0
        self.nseqs = 0
        self.qlen = 0

    def _check(self, rc):
        if rc != 0:
            raise CbpError("comer2 C ABI error %d: %s" % (rc, self.L.cbp_last_error(self.h).decode()))

    def close(self):
        if self.h:
            self.L.cbp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- reference method set ---
    def TransferQueries(self, qset):
        self._queries = qset
        self._check(self.L.cbp_upload_queries(self.h, C.byref(qset.view)))

    def TransferCPMData(self, dbset):
        self._chunk = dbset
        self._check(self.L.cbp_upload_chunk(self.h, C.byref(dbset.view)))

    def SetDbDetails(self, dblen, nseqs):
        self.dblen, self.nseqs = int(dblen), int(nseqs)
        self._check(self.L.cbp_set_db_details(self.h, self.dblen, self.nseqs))

    def SetQueryLen(self, qlen):
        self.qlen = int(qlen)

    def SetScoreThreshold(self, thld):
        self.scorethld = float(thld)

    def SetLogEThreshold(self, thld):
        self.logevthld = float(thld)

    def score_threshold(self, qlen, dblen):
        s, l = C.c_float(), C.c_float()
        self._check(self.L.cbp_score_threshold(C.byref(self._o), qlen, dblen, C.byref(s), C.byref(l)))
        return s.value, l.value

    # --- stage entry points ---
    def phase1_keys(self, qidx):
        keys = np.zeros(self._chunk.nprof, np.uint64)
        self._check(self.L.cbp_phase1_keys(self.h, qidx, keys))
        return keys

    def ProcessScoreMatrix(self, qidx, scorethld=None, logevthld=None):
        """Full search of query qidx against the uploaded chunk; returns a dict of
        numpy arrays (one row per pair that passed phase 1, profile order)."""
        if scorethld is None or logevthld is None:
            qlen = self._queries.lengths()[qidx]
            s, l = self.score_threshold(qlen, self.dblen or self._chunk.npos)
            scorethld = s if scorethld is None else scorethld
            logevthld = l if logevthld is None else logevthld
        r = Results()
        self._check(self.L.cbp_process(self.h, qidx, scorethld, logevthld, C.byref(r)))
        return results_to_dict(r)

    def ProcessBatch(self, qidxs):
        """All listed queries against the uploaded chunk (per-query thresholds
        from the options' EVAL); returns one result dict per query."""
        q = np.ascontiguousarray(qidxs, np.int32)
        outs = (Results * max(1, q.size))()
        self._check(self.L.cbp_process_batch(self.h, q.ctypes.data_as(C.c_void_p), q.size, outs))
        return [results_to_dict(outs[i]) for i in range(q.size)]

    def SearchDB(self, dbset, qidxs, chunk_pos=0):
        """All listed queries against a whole DB in host memory, uploaded in
        chunks of at most chunk_pos positions (0: from the chunk limits) while
        the previous chunk is searched; per query the chunks' hit lists are
        merged (one result dict per query, rows in whole-DB profile order).
        The uploaded chunk, if any, is released."""
        q = np.ascontiguousarray(qidxs, np.int32)
        outs = (Results * max(1, q.size))()
        self._check(self.L.cbp_search_db(self.h, C.byref(dbset.view), int(chunk_pos),
                                         q.ctypes.data_as(C.c_void_p), q.size, outs))
        self._chunk = None
        return [results_to_dict(outs[i]) for i in range(q.size)]

    def last_search_chunks(self):
        n = C.c_int32()
        self._check(self.L.cbp_last_search_chunks(self.h, C.byref(n)))
        return n.value

    def score_block(self, qidx, pronr, want_mod=True):
        """FP32 score matrix (and SS component) of one pair, [Lq][Ldb]."""
        Lq = int(self._queries.lengths()[qidx])
        Ld = int(self._chunk.lengths()[pronr])
        S = np.empty((Lq, Ld), np.float32)
        M = np.empty((Lq, Ld), np.float32) if want_mod else None
        self._check(self.L.cbp_score_block(self.h, qidx, pronr, _fptr(S), _fptr(M) if want_mod else None))
        return S, M

    def last_kernel_ms(self):
        ms, n = C.c_double(), C.c_int32()
        self._check(self.L.cbp_last_kernel_ms(self.h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def set_tuning(self, seg_cols=4096, rmax=5, nw=8):
        self._check(self.L.cbp_set_tuning(self.h, seg_cols, rmax, nw))

    def set_device_lists(self, on):
        """stage lists built on the device (default) or on the host"""
        self._check(self.L.cbp_set_device_lists(self.h, 1 if on else 0))

    def set_device_list_caps(self, cand, passed, code_words, map_cells):
        self._check(self.L.cbp_set_device_list_caps(self.h, cand, passed, code_words, map_cells))

    def last_fallbacks(self):
        n = C.c_int32()
        self._check(self.L.cbp_last_fallbacks(self.h, C.byref(n)))
        return n.value

    def chunk_limits(self, qlen):
        """(max DB positions, max profiles, pass-2 positions, pass-2 profiles)
        a chunk may hold for a query of qlen positions."""
        v = [C.c_uint64() for _ in range(4)]
        self._check(self.L.cbp_chunk_limits(self.h, qlen, *[C.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def GetCurrentMaxDbPos(self, qlen):
        return self.chunk_limits(qlen)[0]

    def GetCurrentMaxNDbPros(self, qlen):
        return self.chunk_limits(qlen)[1]

    def GetCurrentMaxDbPosPass2(self, qlen):
        return self.chunk_limits(qlen)[2]

    def GetCurrentMaxNDbProsPass2(self, qlen):
        return self.chunk_limits(qlen)[3]

    def set_graph(self, on=True):
        """Replay the phase-1 stage from a captured HIP graph while the batch,
        chunk and tuning stay unchanged."""
        self._check(self.L.cbp_set_graph(self.h, 1 if on else 0))

    def set_queue(self, on=True):
        """Batches of two queries or more run phase 1 and the candidates'
        backtrack pass from the device work queue (on, the default) or one
        launch per query (off)."""
        self._check(self.L.cbp_set_queue(self.h, 1 if on else 0))

    def set_mfma(self, on=True):
        """The phase-1 DB sweep forms the 20-term score sums on the matrix
        cores (on) or the vector ALUs (off, the default); bit-identical."""
        self._check(self.L.cbp_set_mfma(self.h, 1 if on else 0))

    def set_map_rows(self, rmax):
        """MAP passes of one-query calls: the shorter pairs' group takes at
        most rmax rows per lane (1-5), or the long pairs' geometry (0)."""
        self._check(self.L.cbp_set_map_rows(self.h, int(rmax)))

    def set_map_pipe(self, on=True):
        """MAP forward and backward passes stream each workgroup's pairs
        through one band (on, the default) or run one workgroup per pair."""
        self._check(self.L.cbp_set_map_pipe(self.h, 1 if on else 0))

    def autotune(self, qidx=0):
        """Times phase 1 of query qidx over segment widths, waves per workgroup
        and rows per lane; keeps and returns the fastest (seg_cols, nw, rmax)."""
        sc, nw, rm = C.c_int32(), C.c_int32(), C.c_int32()
        self._check(self.L.cbp_autotune(self.h, qidx, C.byref(sc), C.byref(nw), C.byref(rm)))
        return sc.value, nw.value, rm.value


def results_to_dict(r):
    n = r.npassed
    if n > 0:
        a = np.ctypeslib.as_array(r.alndata, (n, _ffi.NFIELDS)).copy()
        order = np.ctypeslib.as_array(r.order, (n,)).copy()
    else:
        a = np.zeros((0, _ffi.NFIELDS), np.float32)
        order = np.zeros(0, np.int32)
    u = a.view(np.uint32)
    out = SearchResults(npassed=n, nposits=r.nposits, maxlen=r.maxlen, nmap=r.nmap,
               scorethld=r.scorethld, logevthld=r.logevthld, search_space=r.search_space,
               alndata=a, order=order,
               score=a[:, _ffi.F_SCORE], pronr=u[:, _ffi.F_ORGPRONO].astype(np.int64),
               dst=u[:, _ffi.F_PRONEWDST].astype(np.int64), psts=a[:, _ffi.F_PSTVS],
               beg=u[:, _ffi.F_BEGCOORDS], end=u[:, _ffi.F_ENDCOORDS],
               logE=a[:, _ffi.F_EVALUE],
               lmbd=a[:, _ffi.F_LMBD], K=a[:, _ffi.F_K], H=a[:, _ffi.F_H], E=a[:, _ffi.F_E],
               smin=a[:, _ffi.F_MIN], smax=a[:, _ffi.F_MAX],
               lmbdest=a[:, _ffi.F_LMBDEST], kest=a[:, _ffi.F_KEST],
               idnts=a[:, _ffi.F_IDNTS], ngaps=a[:, _ffi.F_NGAPS])
    out["epa"] = np.ctypeslib.as_array(r.epa, (n,)).copy() if n > 0 else np.zeros(0, np.float32)
    out["alnlen"] = u[:, _ffi.F_EPA].astype(np.int64)
    out["map_rows"] = np.zeros(0, np.int32)
    out["aln_raw"] = np.zeros((0, 3, 1), np.uint8)
    if r.nmap > 0 and r.alns:
        out["aln_raw"] = np.ctypeslib.as_array(C.cast(r.alns, C.POINTER(C.c_uint8)),
                                               (r.nmap, 3, r.alnstride)).copy()
        out["map_rows"] = np.ctypeslib.as_array(r.map_rows, (r.nmap,)).copy()
    return out


class SearchResults(dict):
    """Result dict of one query; 'alns' (query, middle, target strings per
    realigned pair) is decoded from the raw NUL-terminated rows on first use."""

    def __missing__(self, key):
        if key != "alns":
            raise KeyError(key)
        raw = self["aln_raw"]
        val = [tuple(bytes(raw[k, j]).split(b"\0")[0].decode() for j in range(3))
               for k in range(raw.shape[0])]
        self["alns"] = val
        return val


def decode_keys(keys):
    """(score float32, x, y) arrays from phase-1 keys."""
    keys = np.asarray(keys, np.uint64)
    sc = (keys >> np.uint64(32)).astype(np.uint32).view(np.float32)
    lo = np.uint32(0xFFFFFFFF) - (keys & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    x = (lo >> np.uint32(16)).astype(np.int32)
    y = (lo & np.uint32(0xFFFF)).astype(np.int32)
    z = keys == 0
    x[z] = 0
    y[z] = 0
    return sc, x, y


def encode_key(score, x, y):
    hi = int(np.float32(score).view(np.uint32))
    lo = 0xFFFFFFFF - ((int(x) << 16) | int(y))
    return (hi << 32) | lo if score > 0 else 0
