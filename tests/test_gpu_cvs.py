"""The context-vector score term (CVSWGT > 0) on the GPU against the oracle.

Profiles carry synthetic context vectors (c2s_add_cv) and the planted
homologs their source's vectors plus noise; the map is the reference's
src/env/cvs2s as cbp_cvs2s_read smooths it.  Queries of 40, 100, 300 and 700
positions cover one row per lane, two rows per lane, and 3 and 6 stripes
(the CV term runs at most 2 rows per lane); a tuning of one row per lane
covers a 5-stripe query of one row per lane.

* score blocks (S and mod) bit-exact vs or_score_f32, SS term on and off,
  MAPALN 1 and 0 (priors truncated toward zero);
* phase-1 maxima of every profile bit-exact;
* the whole pipeline (one-query calls and a batch) equal to the oracle's
  hit lists, records and strings, MAPALN 1 (MAP) and 0 (phase-1 drawing),
  with a DB without context vectors (the uninformative vector), and with the
  SS term off (the MAP's dynamic score then takes the CV term's mod);
* a whole-DB search in three chunks equal to the one-chunk search.
"""
import numpy as np
import pytest
from comer2_amd._ffi import SynthDB
from gpu_common import oracle, processing, search_expect, assert_search_equal, expected_keys

pytestmark = pytest.mark.gpu

QLENS = [40, 100, 300, 700]


def _sets(db_cv=True):
    q = SynthDB.from_lengths(7001, np.array(QLENS, np.int32)).add_cv(7101)
    d = SynthDB(7002, 160, 200, 0.668, 30, 900)
    if db_cv:
        d.add_cv(7102)
    d.plant(q, 7003, 48, 3.0, 60.0)
    return q, d


def _cbp(q, d, opts):
    cbp = processing(opts)
    cbp.TransferQueries(q)
    cbp.TransferCPMData(d)
    cbp.SetDbDetails(d.npos, d.nprof)
    return cbp


@pytest.mark.parametrize("ssswgt,mapaln", [(0.12, 1), (0.0, 1), (0.12, 0)])
def test_cv_score_blocks_bitexact(ssswgt, mapaln):
    q, d = _sets()
    opts = dict(cvswgt=0.15, ssswgt=ssswgt, mapaln=mapaln)
    orc = oracle(opts)
    cbp = _cbp(q, d, opts)
    print(f"cv score blocks: ssswgt={ssswgt} mapaln={mapaln}", flush=True)
    for qp in range(len(QLENS)):
        for dp in (0, 7, d.nprof - 1):
            S, M = cbp.score_block(qp, dp)
            eS, eM = orc.score_f32(q.view, qp, d.view, dp)
            assert np.array_equal(S.view(np.uint32), eS.view(np.uint32)), (qp, dp)
            assert np.array_equal(M.view(np.uint32), eM.view(np.uint32)), (qp, dp)


def test_cv_phase1_keys_bitexact():
    q, d = _sets()
    opts = dict(cvswgt=0.15)
    orc = oracle(opts)
    cbp = _cbp(q, d, opts)
    for qp in range(len(QLENS)):
        print(f"cv phase 1: query {qp} (L={QLENS[qp]})", flush=True)
        keys = cbp.phase1_keys(qp)
        assert np.array_equal(keys, expected_keys(orc, q, qp, d)), qp
    # one row per lane: the 300-position query in 5 stripes
    cbp.set_tuning(4096, 1, 8)
    print("cv phase 1: one row per lane", flush=True)
    assert np.array_equal(cbp.phase1_keys(2), expected_keys(orc, q, 2, d))
    # the term changes the maxima
    plain = _cbp(q, d, dict(cvswgt=0.0))
    assert not np.array_equal(plain.phase1_keys(2), cbp.phase1_keys(2))


@pytest.mark.parametrize("mapaln,db_cv,ssswgt", [(1, True, 0.12), (0, True, 0.12), (1, False, 0.12), (1, True, 0.0)])
def test_cv_pipeline_matches_oracle(mapaln, db_cv, ssswgt):
    q, d = _sets(db_cv)
    opts = dict(cvswgt=0.15, mapaln=mapaln, max_map_pairs=12, ssswgt=ssswgt)
    orc = oracle(opts)
    cbp = _cbp(q, d, opts)
    thld, logev = 6.0, float(np.log(1e-2))
    drawn = 0
    one = []
    for qp in range(len(QLENS)):
        print(f"cv pipeline: mapaln={mapaln} db_cv={db_cv} ssswgt={ssswgt} query {qp}", flush=True)
        res = cbp.ProcessScoreMatrix(qp, thld, logev)
        exp = search_expect(orc, q, qp, d, thld, logev, d.npos, d.nprof, 12)
        assert_search_equal(res, exp, ("cv", mapaln, db_cv, qp))
        drawn += len(exp["sel"])
        one.append(res)
    assert drawn >= 8
    # a batch of the queries equals the one-query calls
    batch = cbp.ProcessBatch(list(range(len(QLENS))))
    for qp in range(len(QLENS)):
        r = cbp.ProcessScoreMatrix(qp)
        b = batch[qp]
        assert b["npassed"] == r["npassed"] and list(b["order"]) == list(r["order"]), qp
        assert np.array_equal(b["alndata"].view(np.uint32), r["alndata"].view(np.uint32)), qp
        assert list(b["alns"]) == list(r["alns"]), qp


def test_cv_search_db_in_chunks_equals_one_chunk():
    q, d = _sets()
    opts = dict(cvswgt=0.15, max_map_pairs=12)
    cbp = processing(opts)
    cbp.TransferQueries(q)
    qs = list(range(len(QLENS)))
    whole = cbp.SearchDB(d, qs, chunk_pos=d.npos)
    assert cbp.last_search_chunks() == 1
    parts = cbp.SearchDB(d, qs, chunk_pos=d.npos // 3 + 1)
    assert cbp.last_search_chunks() >= 3
    for qp in qs:
        a, b = whole[qp], parts[qp]
        assert a["npassed"] == b["npassed"] and list(a["order"]) == list(b["order"]), qp
        assert np.array_equal(a["alndata"].view(np.uint32), b["alndata"].view(np.uint32)), qp
        assert list(a["alns"]) == list(b["alns"]), qp
    assert sum(r["nmap"] for r in whole) >= 8
