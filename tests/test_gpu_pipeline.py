"""Phase-1 finalize and statistics on the GPU vs the oracle.

Finalize records (score after the correlation bonus, positive-score count,
beginning/end coordinates) must match the FP32 oracle bit for bit; the
statistics are FP64 on both sides and must agree to 1e-9 relative.
"""
import numpy as np
import pytest
from comer2_amd._ffi import SynthDB
from gpu_common import oracle, processing

pytestmark = pytest.mark.gpu


def _expected(orc, q, qp, d, thld, searchspace):
    exp = []
    qaa = bytes(q.view.aa[q.view.dist[qp]:q.view.dist[qp] + q.view.len[qp]])
    for dp in range(d.nprof):
        S, _ = orc.score_f32(q.view, qp, d.view, dp)
        m, bt, _ = orc.viterbi_f32(q.view, qp, d.view, dp, S)
        if not (m.maxsc > thld):
            continue
        f = orc.finalize(S, bt, m)
        if not (thld < f.score):
            continue
        daa = bytes(d.view.aa[d.view.dist[dp]:d.view.dist[dp] + d.view.len[dp]])
        cnt = orc.histogram(S, qaa, daa, 1)
        st = orc.statistics(cnt, q.view.len[qp], d.view.len[dp], q.view.eno[qp], d.view.eno[dp],
                            f.score, float(f.psts), np.float32(searchspace))
        exp.append((dp, f, st))
    return exp


@pytest.mark.parametrize("rmax", [4, 2])
def test_pipeline_phase1_finalize_and_stats(rmax):
    q = SynthDB(21, 2, 150, 0.668, 30, 400)
    d = SynthDB(22, 80, 180, 0.668, 30, 800)
    orc = oracle()
    cbp = processing()
    cbp.set_tuning(2048, rmax, 8)
    cbp.TransferQueries(q)
    cbp.TransferCPMData(d)
    cbp.SetDbDetails(d.npos, d.nprof)
    thld = 1.0
    for qp in range(q.nprof):
        res = cbp.ProcessScoreMatrix(qp, scorethld=thld, logevthld=np.log(10.0))
        ss, _ = orc.length_adjustment(q.view.len[qp], d.npos, d.nprof)
        assert res["search_space"] == ss
        exp = _expected(orc, q, qp, d, thld, ss)
        assert res["npassed"] == len(exp) > 0
        for k, (dp, f, st) in enumerate(exp):
            assert res["pronr"][k] == dp
            assert res["score"][k].view(np.uint32) == np.float32(f.score).view(np.uint32)
            if k not in set(res["map_rows"].tolist()):   # realigned rows carry MAP's values
                assert res["psts"][k] == f.psts
                assert res["beg"][k] == ((f.begx << 16) | f.begy)
                assert res["end"][k] == ((f.endx << 16) | f.endy)
            for name, val in (("lmbd", st.Lmbd), ("K", st.K), ("H", st.H), ("E", st.E),
                              ("logE", st.Evalue), ("epa", st.EpA), ("smin", st.Min),
                              ("smax", st.Max)):
                got = float(res[name][k])
                assert abs(got - val) <= 1e-6 * max(1.0, abs(val)), (name, dp, got, val)
        le = res["logE"]
        assert np.all(np.diff(le[res["order"]]) >= 0)


def test_score_block_bitexact():
    q = SynthDB(23, 2, 150, 0.668, 30, 400)
    d = SynthDB(24, 12, 180, 0.668, 30, 800)
    orc = oracle()
    cbp = processing()
    cbp.TransferQueries(q)
    cbp.TransferCPMData(d)
    for qp in range(q.nprof):
        for dp in range(d.nprof):
            S, M = cbp.score_block(qp, dp)
            eS, eM = orc.score_f32(q.view, qp, d.view, dp)
            assert np.array_equal(S.view(np.uint32), eS.view(np.uint32)), (qp, dp)
            assert np.array_equal(M.view(np.uint32), eM.view(np.uint32)), (qp, dp)


def test_batch_equals_one_by_one():
    q = SynthDB(25, 4, 150, 0.668, 60, 500)
    d = SynthDB(26, 120, 180, 0.668, 30, 800)
    d.plant(q, 27, 20, 3.0, 60.0)
    cbp = processing()
    cbp.TransferQueries(q)
    cbp.TransferCPMData(d)
    cbp.SetDbDetails(d.npos, d.nprof)
    batch = cbp.ProcessBatch([3, 0, 2, 1])
    for qp, rb in zip([3, 0, 2, 1], batch):
        r1 = cbp.ProcessScoreMatrix(qp)
        assert rb["npassed"] == r1["npassed"] and rb["nmap"] == r1["nmap"]
        assert np.array_equal(rb["alndata"].view(np.uint32), r1["alndata"].view(np.uint32))
        assert np.array_equal(rb["order"], r1["order"])
        assert rb["alns"] == r1["alns"]
        assert np.array_equal(rb["aln_raw"], r1["aln_raw"])      # NUL-padded to the stride
    assert sum(r["nmap"] for r in batch) > 0


def test_graph_replay_matches_eager():
    """Phase 1 captured in a HIP graph: every replay, with eager work and
    synchronisations between replays, gives the eager result bit for bit."""
    q = SynthDB(28, 3, 150, 0.668, 60, 500)
    d = SynthDB(29, 100, 180, 0.668, 30, 800)
    d.plant(q, 30, 16, 3.0, 60.0)
    cbp = processing()
    cbp.TransferQueries(q)
    cbp.TransferCPMData(d)
    cbp.SetDbDetails(d.npos, d.nprof)
    qs = [2, 0, 1]
    eager = cbp.ProcessBatch(qs)
    cbp.set_graph(True)
    for rep in range(4):
        got = cbp.ProcessBatch(qs)
        for e, g in zip(eager, got):
            assert g["npassed"] == e["npassed"] and g["nmap"] == e["nmap"], rep
            assert np.array_equal(g["alndata"].view(np.uint32), e["alndata"].view(np.uint32)), rep
            assert np.array_equal(g["order"], e["order"]), rep
            assert g["alns"] == e["alns"], rep
        # eager work between replays: a single query and a score block
        r1 = cbp.ProcessScoreMatrix(1)
        assert np.array_equal(r1["alndata"].view(np.uint32), eager[2]["alndata"].view(np.uint32))
        cbp.score_block(0, rep)
    # a tuning change invalidates the graph; the re-captured one still agrees
    cbp.set_tuning(2048, 5, 4)
    for rep in range(2):
        got = cbp.ProcessBatch(qs)
        for e, g in zip(eager, got):
            assert np.array_equal(g["alndata"].view(np.uint32), e["alndata"].view(np.uint32)), rep
    cbp.set_graph(False)


def test_autotune_keeps_results():
    q = SynthDB(31, 2, 150, 0.668, 60, 500)
    d = SynthDB(32, 200, 180, 0.668, 30, 800)
    cbp = processing()
    cbp.TransferQueries(q)
    cbp.TransferCPMData(d)
    cbp.SetDbDetails(d.npos, d.nprof)
    before = cbp.phase1_keys(0)
    sc, nw, rm = cbp.autotune(0)
    assert sc in (2048, 4096, 8192, 16384) and nw in (4, 8) and rm in (3, 4, 5)
    assert np.array_equal(cbp.phase1_keys(0), before)
    assert np.array_equal(cbp.phase1_keys(1), cbp.phase1_keys(1))


def test_chunk_limits():
    """Chunk limits shrink with the query length and grow with the budget;
    pass 2 holds 10% of the positions at 50 positions per profile."""
    from comer2_amd.host import CuBatchProcessing
    small = CuBatchProcessing(0, mem_budget=8 << 30)
    big = CuBatchProcessing(0, mem_budget=32 << 30)
    a, b, c = small.chunk_limits(300), small.chunk_limits(2000), big.chunk_limits(300)
    for lim in (a, b, c):
        assert lim[0] % 32 == 0 and lim[2] % 32 == 0
        assert lim[1] == lim[0] // 50 and lim[3] == lim[2] // 50
        assert lim[2] <= lim[0] // 10 and lim[2] > lim[0] // 10 - 32
    assert a[0] > b[0] > 0 and c[0] > 3 * a[0]
    free = CuBatchProcessing(0).chunk_limits(300)
    assert free[0] > 50_000_000          # the 50 M-position DB is one chunk on one GPU
    small.close(); big.close()


def _same(a, b):
    return (a["npassed"] == b["npassed"] and a["nmap"] == b["nmap"]
            and np.array_equal(a["alndata"].view(np.uint32), b["alndata"].view(np.uint32))
            and np.array_equal(a["order"], b["order"]) and np.array_equal(a["aln_raw"], b["aln_raw"]))


def test_work_queue_batch_equals_per_query_launches():
    """A batch whose queries fall into every rows-per-lane bin (1..5), one
    and several stripes, runs phase 1 and the backtrack pass from the device
    work queue; the hit lists equal those of per-query launches and of
    one-query runs bit for bit, also after the batch buffers grow."""
    lens = [40, 100, 150, 230, 300, 350, 500, 700, 1000]
    q = SynthDB.from_lengths(33, lens)
    d = SynthDB(34, 160, 180, 0.668, 30, 900)
    d.plant(q, 35, 40, 3.0, 60.0)
    cbp = processing()
    cbp.set_tuning(1024, 5, 8)           # many tiles per query
    cbp.TransferQueries(q)
    cbp.TransferCPMData(d)
    cbp.SetDbDetails(d.npos, d.nprof)
    ql = q.lengths()
    qs = [int(i) for i in np.argsort(-ql, kind="stable")][::2] + [int(i) for i in np.argsort(-ql, kind="stable")][1::2]
    small = cbp.ProcessBatch(qs[:3])                 # smaller batch first: buffers grow later
    queued = cbp.ProcessBatch(qs)
    cbp.set_queue(False)
    launched = cbp.ProcessBatch(qs)
    cbp.set_queue(True)
    for k, qp in enumerate(qs):
        one = cbp.ProcessScoreMatrix(qp)
        assert _same(queued[k], one), (k, qp, int(ql[qp]))
        assert _same(launched[k], one), (k, qp, int(ql[qp]))
        if k < 3:
            assert _same(small[k], one), (k, qp)
    assert sum(r["npassed"] for r in queued) > len(qs)
