"""The C++ drop-in comer2hip::HipBatchProcessing on the GPU.

The tests' driver (tests/adapter_check.cpp) lays the query batch and the DB
out in the reference's packed per-field buffers and runs DevCommThread's
loop: TransferCPMData once, then per query SetQueryLen / SetScoreThreshold /
SetLogEThreshold / ProcessScoreMatrix with the query's field pointers
advanced into the batch, the DB given as a chunk alone or as a cached part
(bdb1) plus a chunk.  Every query's results must equal those of the C ABI
called directly on the same profile sets, bit for bit, and the chunk
counter must be released once per ProcessScoreMatrix.
"""
import ctypes as C
import os

import numpy as np
import pytest

from comer2_amd import _ffi, tables
from comer2_amd._ffi import SynthDB
from comer2_amd.host import results_to_dict
from gpu_common import processing

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SINK = C.CFUNCTYPE(None, C.c_int, C.POINTER(_ffi.Results), C.c_int)


def _lib():
    import torch  # noqa: F401  (one HIP runtime in the process)
    T = C.CDLL(os.path.join(ROOT, "tests", "libadapter_check.so"))
    P = C.POINTER(_ffi.Profiles)
    fp = C.POINTER(C.c_float)
    T.c2t_run.argtypes = [P, P, C.c_int, C.POINTER(_ffi.Options), fp, fp, C.c_void_p, C.c_int,
                          C.c_void_p, C.c_void_p, SINK, C.POINTER(C.c_int)]
    return T


def _same(a, b):
    return (a["npassed"] == b["npassed"] and a["nmap"] == b["nmap"]
            and np.array_equal(a["alndata"].view(np.uint32), b["alndata"].view(np.uint32))
            and np.array_equal(a["order"], b["order"]) and np.array_equal(a["map_rows"], b["map_rows"])
            and np.array_equal(a["aln_raw"], b["aln_raw"]) and a["search_space"] == b["search_space"])


@pytest.mark.parametrize("nsplit", [0, 37])
def test_drop_in_class_equals_c_abi(nsplit):
    T = _lib()
    q = SynthDB(71, 5, 200, 0.6, 60, 700)
    d = SynthDB(72, 110, 180, 0.668, 30, 900)
    d.plant(q, 73, 30, 3.0, 60.0)
    opts = dict(tables.default_options(), max_map_pairs=20)
    order = np.array([3, 0, 4, 1, 2, 0], np.int32)        # a query searched twice
    cbp = processing(dict(max_map_pairs=20))
    cbp.TransferQueries(q)
    cbp.TransferCPMData(d)
    cbp.SetDbDetails(d.npos, d.nprof)
    st = np.zeros(order.size, np.float32)
    lt = np.zeros(order.size, np.float32)
    want = []
    for i, qp in enumerate(order):
        st[i], lt[i] = cbp.score_threshold(int(q.lengths()[qp]), d.npos)
        want.append(cbp.ProcessScoreMatrix(int(qp), float(st[i]), float(lt[i])))
    cbp.close()
    got = []

    def sink(qn, res, ndesc):
        got.append((qn, results_to_dict(res.contents), ndesc))

    cb = SINK(sink)
    o = _ffi.options_struct(opts)
    ss = np.ascontiguousarray(tables.load_ssss(), np.float32)
    nn = np.ascontiguousarray(tables.load_nn_weights(), np.float32)
    decs = C.c_int()
    print(f"drop-in class: nsplit={nsplit}", flush=True)
    rc = T.c2t_run(C.byref(q.view), C.byref(d.view), nsplit, C.byref(o),
                   ss.ctypes.data_as(C.POINTER(C.c_float)), nn.ctypes.data_as(C.POINTER(C.c_float)),
                   order.ctypes.data_as(C.c_void_p), order.size,
                   st.ctypes.data_as(C.c_void_p), lt.ctypes.data_as(C.c_void_p), cb, C.byref(decs))
    assert rc == 0
    assert decs.value == order.size
    assert [g[0] for g in got] == order.tolist()
    assert all(g[2] == d.nprof - nsplit for g in got)
    for (qn, r, _), w in zip(got, want):
        assert _same(r, w), qn
    assert sum(w["nmap"] for w in want) > 0
