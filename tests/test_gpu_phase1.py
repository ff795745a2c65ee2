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
