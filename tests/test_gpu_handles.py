"""Handle-level behaviour on the GPU: the device memory budget (refusal with
CBP_ERR_MEM, recovery on the same handle, limits that do not drift as
queries and tunings are re-uploaded) and distinct handles driven from two
host threads at once (one handle per worker thread, as DevCommThread runs
one CuBatchProcessing per thread, cualn/DevCommThread.cpp:149-356)."""
import math
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest
from comer2_amd._ffi import SynthDB

pytestmark = pytest.mark.gpu


def _same(a, b):
    return (a["npassed"] == b["npassed"] and a["nmap"] == b["nmap"]
            and np.array_equal(a["alndata"].view(np.uint32), b["alndata"].view(np.uint32))
            and np.array_equal(a["order"], b["order"]) and np.array_equal(a["aln_raw"], b["aln_raw"]))


def _small_case():
    q = SynthDB(301, 3, 220, 0.4, 60, 400)
    d = SynthDB(302, 300, 180, 0.5, 30, 800)
    d.plant(q, 303, 12, 3.0, 60.0)
    return q, d


def test_memory_budget_refusal_and_recovery():
    from comer2_amd.host import CbpError, CuBatchProcessing
    q, d = _small_case()
    big = SynthDB(304, 16000, 200, 0.5, 30, 800)          # ~3.4 M positions, > 256 MiB of fields
    print(f"budget: big DB {big.npos} positions, small DB {d.npos}", flush=True)
    cbp = CuBatchProcessing(0, mem_budget=256 << 20)
    cbp.TransferQueries(q)
    with pytest.raises(CbpError, match="-3"):
        cbp.TransferCPMData(big)
    # the refused upload leaves the handle usable: the small DB searches as
    # on a handle without a budget
    cbp.TransferCPMData(d)
    cbp.SetDbDetails(d.npos, d.nprof)
    ref = CuBatchProcessing(0)
    ref.TransferQueries(q)
    ref.TransferCPMData(d)
    ref.SetDbDetails(d.npos, d.nprof)
    thld, logev = 4.0, math.log(1e-3)
    first = [cbp.ProcessScoreMatrix(i, thld, logev) for i in range(q.nprof)]
    for i in range(q.nprof):
        assert _same(first[i], ref.ProcessScoreMatrix(i, thld, logev)), i
    assert sum(r["nmap"] for r in first) > 0
    # re-uploading queries and re-tuning frees what they replace: once both
    # tunings have run, the chunk limits (budget minus live buffers) stay
    # where they are
    lim = None
    for k in range(6):
        cbp.set_tuning(seg_cols=(2048, 4096)[k % 2])
        cbp.TransferQueries(q)
        for i in range(q.nprof):
            assert _same(cbp.ProcessScoreMatrix(i, thld, logev), first[i]), (k, i)
        if k == 1:
            lim = cbp.chunk_limits(300)
        elif k > 1:
            assert cbp.chunk_limits(300) == lim, k
    cbp.close(); ref.close()


def test_two_handles_on_two_threads():
    """Two handles searching at once from two threads give the results of
    one handle searching alone."""
    from comer2_amd.host import CuBatchProcessing
    q, d = _small_case()
    thld, logev = 4.0, math.log(1e-3)

    def run(_):
        cbp = CuBatchProcessing(0)
        cbp.TransferQueries(q)
        cbp.TransferCPMData(d)
        cbp.SetDbDetails(d.npos, d.nprof)
        out = [cbp.ProcessScoreMatrix(i, thld, logev) for _ in range(3) for i in range(q.nprof)]
        out.append(cbp.ProcessBatch(list(range(q.nprof))))
        cbp.close()
        return out

    alone = run(0)
    print("two threads: one handle each", flush=True)
    with ThreadPoolExecutor(2) as ex:
        both = list(ex.map(run, range(2)))
    for out in both:
        for a, b in zip(out[:-1], alone[:-1]):
            assert _same(a, b)
        for a, b in zip(out[-1], alone[-1]):
            assert _same(a, b)
