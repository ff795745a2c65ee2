"""The phase-1 sweep with its 20-term score sums on the matrix cores
(v_mfma_f32_4x4x1_16b_f32 chained over the terms, cbp_set_mfma): keys
bit-exact against the FP32 oracle for one-query launches at 8 and 4 waves
per workgroup, every rows-per-lane count, one and several stripes and the
edge lengths; whole searches (one query, batches through the work queue)
bit-identical to the vector-ALU sweep."""
import math

import numpy as np
import pytest
from comer2_amd._ffi import SynthDB
from gpu_common import expected_keys, oracle, processing

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("rmax,seg_cols,nw", [(5, 4096, 8), (4, 2000, 8), (3, 64, 4), (2, 300, 8), (1, 500, 4)])
def test_mfma_phase1_keys_bitexact(rmax, seg_cols, nw):
    q = SynthDB(11, 3, 150, 0.668, 30, 400)
    d = SynthDB(12, 60, 180, 0.668, 30, 800)
    orc = oracle()
    cbp = processing()
    cbp.set_tuning(seg_cols, rmax, nw)
    cbp.set_mfma(True)
    cbp.TransferQueries(q)
    cbp.TransferCPMData(d)
    for qp in list(range(q.nprof)) + [0]:
        print(f"mfma phase1: rmax={rmax} seg_cols={seg_cols} nw={nw} query {qp}", flush=True)
        got = cbp.phase1_keys(qp)
        exp = expected_keys(orc, q, qp, d)
        bad = np.nonzero(got != exp)[0]
        assert bad.size == 0, "query %d: %d/%d pairs differ, first %s got %x exp %x" % (
            qp, bad.size, d.nprof, bad[:5], int(got[bad[0]]), int(exp[bad[0]]))


EDGE_Q = [1, 2, 31, 32, 33, 63, 64, 65, 97, 319, 320, 321, 640, 641]
EDGE_D = [1, 2, 3, 31, 32, 33, 60, 61, 64, 65, 100, 177, 1, 2, 3, 4, 5, 6, 7, 8]


@pytest.mark.parametrize("rmax", [5, 2])
def test_mfma_phase1_keys_edge_lengths(rmax):
    q = SynthDB.from_lengths(111, EDGE_Q)
    d = SynthDB.from_lengths(112, EDGE_D)
    orc = oracle()
    cbp = processing()
    cbp.set_tuning(128, rmax, 4)
    cbp.set_mfma(True)
    cbp.TransferQueries(q)
    cbp.TransferCPMData(d)
    for qp in range(q.nprof):
        print(f"mfma phase1 edge: rmax={rmax} Lq={EDGE_Q[qp]}", flush=True)
        got = cbp.phase1_keys(qp)
        exp = expected_keys(orc, q, qp, d)
        bad = np.nonzero(got != exp)[0]
        assert bad.size == 0, "Lq %d: profiles %s got %s exp %s" % (
            EDGE_Q[qp], bad[:5], [hex(int(got[i])) for i in bad[:3]], [hex(int(exp[i])) for i in bad[:3]])


def _same(a, b):
    return (a["npassed"] == b["npassed"] and a["nmap"] == b["nmap"]
            and np.array_equal(a["alndata"].view(np.uint32), b["alndata"].view(np.uint32))
            and np.array_equal(a["order"], b["order"]) and np.array_equal(a["aln_raw"], b["aln_raw"]))


def test_mfma_searches_equal_vector_sweep():
    """one-query searches and a work-queue batch whose queries fall into
    every rows-per-lane bin, with the matrix-core sweep on and off"""
    q = SynthDB.from_lengths(501, [40, 100, 150, 230, 300, 330, 700, 1000])
    d = SynthDB(502, 400, 200, 0.668, 30, 3000)
    d.plant(q, 503, 40, 3.0, 60.0)
    thld, logev = 6.0, math.log(1e-3)
    res = {}
    for on in (False, True):
        cbp = processing()
        cbp.set_mfma(on)
        cbp.TransferQueries(q)
        cbp.TransferCPMData(d)
        cbp.SetDbDetails(d.npos, d.nprof)
        print(f"mfma searches: mfma={on}", flush=True)
        one = [cbp.ProcessScoreMatrix(i, thld, logev) for i in range(q.nprof)]
        batch = cbp.ProcessBatch(list(range(q.nprof)))
        res[on] = (one, batch)
        cbp.close()
    assert sum(r["nmap"] for r in res[False][0]) > 0
    for a, b in zip(res[False][0], res[True][0]):
        assert _same(a, b)
    for a, b in zip(res[False][1], res[True][1]):
        assert _same(a, b)
