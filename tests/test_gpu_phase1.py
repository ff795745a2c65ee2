"""Phase-1 fused score+Viterbi on the GPU vs the FP32 oracle (bit-exact keys)."""
import numpy as np
import pytest
from comer2_amd._ffi import SynthDB
from gpu_common import oracle, expected_keys, processing

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("rmax,seg_cols,nw", [(4, 4096, 8), (1, 300, 2), (3, 64, 4), (5, 2000, 1)])
def test_phase1_keys_bitexact(rmax, seg_cols, nw):
    q = SynthDB(11, 3, 150, 0.668, 30, 400)
    d = SynthDB(12, 60, 180, 0.668, 30, 800)
    orc = oracle()
    cbp = processing()
    cbp.set_tuning(seg_cols, rmax, nw)
    cbp.TransferQueries(q)
    cbp.TransferCPMData(d)
    # query 0 (400 rows) runs in several stripes; repeating it on the same
    # handle checks that the first stripe never reads the hand-off buffer the
    # previous launch left behind
    for qp in list(range(q.nprof)) + [0]:
        got = cbp.phase1_keys(qp)
        exp = expected_keys(orc, q, qp, d)
        bad = np.nonzero(got != exp)[0]
        assert bad.size == 0, "query %d: %d/%d pairs differ, first %s got %x exp %x" % (
            qp, bad.size, d.nprof, bad[:5], int(got[bad[0]]), int(exp[bad[0]]))


EDGE_Q = [1, 2, 31, 32, 33, 63, 64, 65, 97, 319, 320, 321, 640, 641]
EDGE_D = [1, 2, 3, 31, 32, 33, 60, 61, 64, 65, 100, 177, 1, 2, 3, 4, 5, 6, 7, 8]


@pytest.mark.parametrize("rmax", [5, 2])
def test_phase1_keys_edge_lengths(rmax):
    """Query and profile lengths at the 32-row block edges, the lane-row and
    stripe edges (64R rows), the corner-mask edge and down to 1, with runs of
    very short profiles inside one segment."""
    q = SynthDB.from_lengths(111, EDGE_Q)
    d = SynthDB.from_lengths(112, EDGE_D)
    orc = oracle()
    cbp = processing()
    cbp.set_tuning(128, rmax, 4)
    cbp.TransferQueries(q)
    cbp.TransferCPMData(d)
    for qp in range(q.nprof):
        print(f"phase1 edge: rmax={rmax} Lq={EDGE_Q[qp]}", flush=True)
        got = cbp.phase1_keys(qp)
        exp = expected_keys(orc, q, qp, d)
        bad = np.nonzero(got != exp)[0]
        assert bad.size == 0, "Lq %d: profiles %s (lengths %s) got %s exp %s" % (
            EDGE_Q[qp], bad[:5], [EDGE_D[i] for i in bad[:5]],
            [hex(int(got[i])) for i in bad[:3]], [hex(int(exp[i])) for i in bad[:3]])


@pytest.mark.parametrize("rmax", [5, 3])
def test_phase1_keys_ss_term_off(rmax):
    """ssswgt = 0: the kernel adds the SS term with weight 0 (an exact +-0),
    the oracle skips it; the keys must agree bit for bit."""
    q = SynthDB(21, 2, 150, 0.668, 30, 400)
    d = SynthDB(22, 50, 180, 0.668, 30, 700)
    opts = dict(ssswgt=0.0)
    orc = oracle(opts)
    cbp = processing(opts)
    cbp.set_tuning(1024, rmax, 4)
    cbp.TransferQueries(q)
    cbp.TransferCPMData(d)
    for qp in range(q.nprof):
        print(f"phase1 ss off: rmax={rmax} query {qp}", flush=True)
        got = cbp.phase1_keys(qp)
        exp = expected_keys(orc, q, qp, d)
        bad = np.nonzero(got != exp)[0]
        assert bad.size == 0, "query %d: %d/%d pairs differ" % (qp, bad.size, d.nprof)
