"""A query's search results as the reference's output files: plain text
(cbp_format_plain) and JSON (cbp_format_json), and the search space behind
their summaries (cbp_search_space).

``format_plain(result, options, query=..., db=...)`` takes a result dict of
``CuBatchProcessing.ProcessScoreMatrix`` / ``ProcessBatch`` / ``SearchDB``
and returns the text of the reference's per-query output file
(cualn/AlnWriterPlain.cpp:19-121); ``format_json`` takes the same
arguments and returns the JSON file (cualn/AlnWriterJSON.cpp:20-143).
"""
import ctypes as C
import numpy as np
from . import _ffi
from ._ffi import LIB_HIP, Options, Results

_lib = None


class PlainInfo(C.Structure):
    _fields_ = [("progname", C.c_char_p), ("progversion", C.c_char_p), ("references", C.POINTER(C.c_char_p)),
                ("query_desc", C.c_char_p), ("query_len", C.c_int32), ("query_eno", C.c_float),
                ("query_sss", C.c_char_p),
                ("db_names", C.POINTER(C.c_char_p)), ("n_db_names", C.c_int32),
                ("db_size", C.c_uint64), ("db_entries", C.c_uint64),
                ("pro_desc", C.POINTER(C.c_char_p)), ("pro_len", C.POINTER(C.c_int32)),
                ("pro_eno", C.POINTER(C.c_float)), ("pro_dist", C.POINTER(C.c_uint32)),
                ("pro_sss", C.c_char_p), ("deltalen", C.c_int32),
                ("nohits", C.c_int32), ("noalns", C.c_int32),
                ("dsclen", C.c_int32), ("dscwidth", C.c_int32), ("alnwidth", C.c_int32), ("show", C.c_int32)]


def _L():
    global _lib
    if _lib is None:
        L = _ffi.load_hip()
        for f in (L.cbp_format_plain, L.cbp_format_json):
            f.argtypes = [C.POINTER(Options), C.POINTER(PlainInfo), C.POINTER(Results),
                          C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]
        L.cbp_search_space.argtypes = [C.POINTER(Options), C.c_uint64, C.c_uint64, C.c_uint64,
                                       C.POINTER(C.c_uint64), C.POINTER(C.c_int32)]
        _lib = L
    return _lib


def search_space(options, qlen, dblen, nseqs):
    """(effective search space, length adjustment) of the length adjustment"""
    o = _ffi.options_struct(options)
    ss, dl = C.c_uint64(), C.c_int32()
    if _L().cbp_search_space(C.byref(o), qlen, dblen, nseqs, C.byref(ss), C.byref(dl)) != 0:
        raise ValueError("cbp_search_space")
    return ss.value, dl.value


def _strarr(items):
    arr = (C.c_char_p * (len(items) + 1))()
    for i, s in enumerate(items):
        arr[i] = s.encode() if isinstance(s, str) else s
    arr[len(items)] = None
    return arr


def format_plain(result, options, *args, **kw):
    """the plain-text output of one query (see the module doc)"""
    return _format("cbp_format_plain", result, options, *args, **kw)


def format_json(result, options, *args, **kw):
    """the JSON output of one query (see the module doc)"""
    return _format("cbp_format_json", result, options, *args, **kw)


def _format(fn, result, options, query_desc, query_len, query_eno, db_names, db_size, db_entries,
            pro_desc, pro_len, pro_eno, query_sss=None, pro_sss=None, pro_dist=None,
            deltalen=None, nseqs=None, progname=None, progversion=None, references=None,
            nohits=700, noalns=700, dsclen=1024, dscwidth=80, alnwidth=60, show=1):
    """pro_* are indexed by the records' profile index; deltalen defaults to
    the length adjustment of (query_len, db_size, nseqs or db_entries)"""
    n = int(result["npassed"])
    nm = int(result["nmap"])
    a = np.ascontiguousarray(result["alndata"], np.float32).reshape(n, _ffi.NFIELDS) if n else np.zeros((1, _ffi.NFIELDS), np.float32)
    order = np.ascontiguousarray(result["order"], np.int32) if n else np.zeros(1, np.int32)
    rows = np.ascontiguousarray(result["map_rows"], np.int32) if nm else np.zeros(1, np.int32)
    raw = np.ascontiguousarray(result["aln_raw"], np.uint8) if nm else np.zeros((1, 3, 1), np.uint8)
    r = Results()
    r.npassed, r.nmap = n, nm
    r.alndata = a.ctypes.data_as(C.POINTER(C.c_float))
    r.order = order.ctypes.data_as(C.POINTER(C.c_int32))
    r.alns = raw.ctypes.data_as(C.POINTER(C.c_char))
    r.alnstride = raw.shape[2]
    r.map_rows = rows.ctypes.data_as(C.POINTER(C.c_int32))
    r.logevthld = float(result["logevthld"])
    r.search_space = float(result["search_space"])
    if deltalen is None:
        _, deltalen = search_space(options, query_len, db_size, nseqs or db_entries)
    info = PlainInfo()
    keep = []
    info.progname = progname.encode() if progname else None
    info.progversion = progversion.encode() if progversion else None
    if references:
        refs = _strarr(references); keep.append(refs); info.references = refs
    info.query_desc = query_desc.encode() if query_desc is not None else None
    info.query_len, info.query_eno = int(query_len), float(query_eno)
    info.query_sss = bytes(query_sss) if query_sss is not None else None
    names = _strarr(db_names); keep.append(names)
    info.db_names, info.n_db_names = names, len(db_names)
    info.db_size, info.db_entries = int(db_size), int(db_entries)
    descs = _strarr(pro_desc); keep.append(descs)
    ln = np.ascontiguousarray(pro_len, np.int32); en = np.ascontiguousarray(pro_eno, np.float32)
    keep += [ln, en]
    info.pro_desc = descs
    info.pro_len = ln.ctypes.data_as(C.POINTER(C.c_int32))
    info.pro_eno = en.ctypes.data_as(C.POINTER(C.c_float))
    if pro_sss is not None:
        ds = np.ascontiguousarray(pro_dist, np.uint32); keep.append(ds)
        info.pro_dist = ds.ctypes.data_as(C.POINTER(C.c_uint32))
        info.pro_sss = bytes(pro_sss)
    info.deltalen = int(deltalen)
    info.nohits, info.noalns = nohits, noalns
    info.dsclen, info.dscwidth, info.alnwidth, info.show = dsclen, dscwidth, alnwidth, show
    o = _ffi.options_struct(options)
    f = getattr(_L(), fn)
    size = C.c_size_t()
    if f(C.byref(o), C.byref(info), C.byref(r), None, 0, C.byref(size)) != 0:
        raise ValueError(fn + ": invalid arguments")
    buf = C.create_string_buffer(size.value + 1)
    if f(C.byref(o), C.byref(info), C.byref(r), buf, size.value + 1, C.byref(size)) != 0:
        raise ValueError(fn + ": invalid arguments")
    return buf.raw[:size.value].decode("latin-1")
