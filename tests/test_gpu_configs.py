"""The benchmark configurations on the GPU against the oracle, at sizes the
oracle finishes in seconds.

* C1 sample: the plain pipeline (cbp_process) for a 300-position query
  against 300 DB profiles of the C1 length law (log-normal, median 200,
  sigma 0.668, clipped to 30-3000) with its three longest profiles at 3,000,
  2,950 and 2,800 positions and homologs of the query planted; the hit
  list, phase-1 and MAP records and alignment strings equal the oracle's.
* C2 sample: 64 queries with lengths uniform in 50-1000 (seed 3) in one
  cbp_process_batch call from the device work queue against a C2-law DB
  sample; every query's results equal the oracle's directly.
* SSEMODEL 0 and 1 and X_UNINF with 'X' residues in query and DB.
* C4 sample: long x long soft-DP pairs (2,000-5,000 per side) with sqrt(L)
  checkpointing (the automatic interval resolves to 64 steps) against all
  cells kept, the FP64 oracle and FP64 central differences.

Tolerances: hit lists, coordinates, records and strings exact; statistics
1e-6 relative (FP64 on both sides); soft DP log Z 1e-5 relative,
posteriors 1e-4 relative above 1e-2 (1e-6 absolute below) for the checkpointed
vs all-cells comparison and 2e-3 relative above 1e-2 (2e-5 absolute below)
vs the FP64 oracle on the long pairs (FP32 rounding of log-space states of up to ~200 nats, ulp 1.5e-5,
accumulated over up to 7,000 anti-diagonal steps: log Z is 1.2e-3 off in
absolute terms on a 2,000 x 5,000 pair, and every posterior inherits that as
relative error), dS 1e-4 absolute, dtheta 1e-3 relative, gradients vs
central differences 1e-3.
"""
import numpy as np
import pytest
from comer2_amd._ffi import SynthDB
from gpu_common import (oracle, processing, search_expect, assert_search_equal,
                        lognormal_lengths, set_residues)

pytestmark = pytest.mark.gpu


def _c1_sample():
    q = SynthDB(1002, 1, 300, 0.0, 300, 300)
    lens = np.concatenate([[3000, 2950, 2800], lognormal_lengths(2, 297, 200, 0.668, 30, 3000)])
    lens = np.sort(lens)[::-1].copy()
    d = SynthDB.from_lengths(2, lens)
    d.plant(q, 3002, 40, 3.0, 60.0)
    return q, d


def test_c1_sample_plain_pipeline_matches_oracle():
    q, d = _c1_sample()
    print(f"C1 sample: {d.nprof} profiles, {d.npos} positions, longest {d.lengths()[:3].tolist()}", flush=True)
    orc = oracle()
    cbp = processing(dict(max_map_pairs=2000))
    cbp.TransferQueries(q)
    cbp.TransferCPMData(d)
    cbp.SetDbDetails(d.npos, d.nprof)
    thld, logev = cbp.score_threshold(300, d.npos)
    res = cbp.ProcessScoreMatrix(0, thld, logev)
    exp = search_expect(orc, q, 0, d, thld, logev, d.npos, d.nprof, 2000)
    assert len(exp["sel"]) >= 10
    assert_search_equal(res, exp, ("C1",))
    # phase-1 maxima of every profile, the longest included
    from gpu_common import expected_keys
    assert np.array_equal(cbp.phase1_keys(0), expected_keys(orc, q, 0, d))


def test_c2_sample_batch_matches_oracle():
    nq = 64
    q = SynthDB(3, nq, 0, 0, 50, 1000, sort_desc=False)
    d = SynthDB.from_lengths(33, np.sort(lognormal_lengths(34, 110, 200, 0.668, 30, 3000))[::-1].copy())
    d.plant(q, 3003, 80, 3.0, 60.0)
    print(f"C2 sample: {nq} queries, {d.nprof} profiles, {d.npos} positions", flush=True)
    orc = oracle(dict(max_map_pairs=100))
    cbp = processing(dict(max_map_pairs=100))
    cbp.TransferQueries(q)
    cbp.TransferCPMData(d)
    cbp.SetDbDetails(d.npos, d.nprof)
    qs = list(range(nq))
    batch = cbp.ProcessBatch(qs)
    ql = q.lengths()
    hits = 0
    for qp, res in zip(qs, batch):
        thld, logev = cbp.score_threshold(int(ql[qp]), d.npos)
        exp = search_expect(orc, q, qp, d, thld, logev, d.npos, d.nprof, 100)
        assert_search_equal(res, exp, ("C2", qp, int(ql[qp])))
        hits += len(exp["sel"])
    assert hits >= 40


@pytest.mark.parametrize("model", [0, 1])
def test_ssemodel_variants_match_oracle(model):
    q = SynthDB(1402, 2, 250, 0.5, 100, 500)
    d = SynthDB(1403, 120, 200, 0.668, 30, 900)
    d.plant(q, 1404, 24, 3.0, 60.0)
    opts = dict(ssemodel=model, max_map_pairs=0)
    orc = oracle(opts)
    cbp = processing(opts)
    cbp.TransferQueries(q)
    cbp.TransferCPMData(d)
    cbp.SetDbDetails(50_000_000, 200_000)
    for qp in range(q.nprof):
        thld, logev = cbp.score_threshold(int(q.lengths()[qp]), 50_000_000)
        thld = min(thld, 8.0)
        print(f"ssemodel {model}: query {qp}", flush=True)
        res = cbp.ProcessScoreMatrix(qp, thld, logev)
        exp = search_expect(orc, q, qp, d, thld, logev, 50_000_000, 200_000, 0)
        assert len(exp["rows"]) > 3
        assert_search_equal(res, exp, ("ssemodel", model, qp))


@pytest.mark.parametrize("xuninf", [1, 0])
def test_x_residues_match_oracle(xuninf):
    """'X' residues in the query and the DB: with X_UNINF the histogram skips
    every cell whose query or DB residue is 'X'
    (CalcScoreProbsXuninfDynPlmProfile, cuss/CuBatchSS_dynplm.cu:80-182)."""
    q = SynthDB(1502, 2, 220, 0.5, 100, 400)
    d = SynthDB(1503, 90, 200, 0.668, 30, 700)
    d.plant(q, 1504, 20, 3.0, 60.0)
    rng = np.random.default_rng(1505)
    set_residues(q, rng.choice(q.npos, q.npos // 10, replace=False))
    set_residues(d, rng.choice(d.npos, d.npos // 8, replace=False))
    opts = dict(xuninf=xuninf, max_map_pairs=0)
    orc = oracle(opts)
    cbp = processing(opts)
    cbp.TransferQueries(q)
    cbp.TransferCPMData(d)
    cbp.SetDbDetails(d.npos, d.nprof)
    for qp in range(q.nprof):
        print(f"X residues: xuninf={xuninf} query {qp}", flush=True)
        res = cbp.ProcessScoreMatrix(qp, 6.0, float(np.log(1e-2)))
        exp = search_expect(orc, q, qp, d, 6.0, float(np.log(1e-2)), d.npos, d.nprof, 0, xuninf=xuninf)
        assert len(exp["rows"]) > 3
        assert_search_equal(res, exp, ("xuninf", xuninf, qp))
        assert any("X" in s[0] or "X" in s[2] for s in res["alns"])


def _post_close(got, want, rtol=1e-4):
    """posteriors: rtol relative above 1e-2, rtol * 1e-2 absolute below"""
    tol = rtol * np.maximum(want, 1e-2)
    bad = np.abs(got - want) > tol
    return not bad.any(), float(np.max(np.abs(got - want) / np.maximum(want, 1e-2)))


def test_c4_long_pairs_checkpointed():
    import torch
    from comer2_amd.softdp import SoftDP
    q = SynthDB.from_lengths(5, [2000, 3500])
    d = SynthDB.from_lengths(1005, [5000, 2200, 2100])
    cbp = processing()
    cbp.TransferQueries(q)
    cbp.TransferCPMData(d)
    pairs = [(0, 0), (1, 1), (0, 2)]
    print(f"C4 sample: pairs {pairs}", flush=True)
    orc = oracle()
    th = torch.tensor([0.15, -0.1, 0.1, -0.2, 0.05], dtype=torch.float32, device="cuda")
    out = {}
    for ckpt in (-1, 64, 0):
        sdp = SoftDP(cbp, pairs, ckpt=ckpt)
        S = sdp.scores()
        logZ = sdp.forward(S, th)
        dS, dth, post = sdp.adjoint(S, th, logZ, want_post=True)
        torch.cuda.synchronize()
        out[ckpt] = (S.cpu().numpy(), logZ.cpu().numpy(), dS.cpu().numpy(), dth.cpu().numpy(),
                     post.cpu().numpy(), sdp.ws_bytes)
        del sdp, S, dS, post
        torch.cuda.empty_cache()
    # the automatic interval is sqrt(5063) -> 64 steps: the same plan as K = 64
    assert out[-1][5] == out[64][5] < out[0][5]
    Sh, lz, dS, dth, post, _ = out[-1]
    for k in range(1, 5):
        assert np.array_equal(out[-1][k], out[64][k])
    # checkpointed vs all cells kept
    assert np.allclose(lz, out[0][1], rtol=1e-6, atol=0)
    assert np.max(np.abs(dS - out[0][2])) <= 1e-6
    ok, worst = _post_close(post, out[0][4])
    assert ok, ("checkpointed vs all cells", worst)
    assert np.allclose(dth, out[0][3], rtol=1e-5, atol=1e-5)
    # vs the FP64 oracle
    thd = th.cpu().numpy().astype(np.float64)
    off = 0
    for k, (qp, dp) in enumerate(pairs):
        L1, L2 = int(q.lengths()[qp]), int(d.lengths()[dp])
        n = L1 * L2
        Sk = Sh[off:off + n].reshape(L1, L2).astype(np.float64)
        z, eS, eth, epost = orc.sdp_grad(q.view, qp, d.view, dp, Sk, thd)
        ctx = (qp, dp, L1, L2)
        assert abs(lz[k] - z) <= 1e-5 * abs(z), ctx
        assert np.max(np.abs(dS[off:off + n] - eS.ravel())) < 1e-4, ctx
        ok, worst = _post_close(post[off:off + n], epost.ravel(), rtol=2e-3)
        print(f"C4 pair {ctx}: logZ {lz[k]:.4f} vs {z:.4f}, posterior worst rel {worst:.2e}", flush=True)
        assert ok, (ctx, worst)
        assert np.all(np.abs(dth[k] - eth) <= 1e-3 * np.maximum(1.0, np.abs(eth))), (ctx, dth[k], eth)
        if k == 2:
            # FP64 central differences (step 1e-3) on all five theta and four cells
            h = 1e-3
            f = lambda SS, tt: orc.sdp_logz(q.view, qp, d.view, dp, SS, tt)
            for j in range(5):
                tp, tm = thd.copy(), thd.copy()
                tp[j] += h; tm[j] -= h
                fd = (f(Sk, tp) - f(Sk, tm)) / (2 * h)
                assert abs(dth[k, j] - fd) <= 1e-3 * max(1.0, abs(fd)), (j, dth[k, j], fd)
            flat = np.argsort(epost.ravel())[::-1]
            rng = np.random.default_rng(9)
            cells = [int(flat[0]), int(flat[50]), int(rng.integers(0, n)), int(rng.integers(0, n))]
            for c in cells:
                i, j = divmod(c, L2)
                Sp, Sm = Sk.copy(), Sk.copy()
                Sp[i, j] += h; Sm[i, j] -= h
                fd = (f(Sp, thd) - f(Sm, thd)) / (2 * h)
                assert abs(dS[off + c] - fd) <= 1e-3, (i, j, dS[off + c], fd)
        off += n


def test_mapaln0_phase1_drawing_matches_oracle():
    """MAPALN=0: the selected pairs are drawn from their phase-1 backtrace
    (FinalizeDP_ALN, cudp/CuBatchDP_aln_final.cu:61-438) instead of being
    realigned, and the score shift is 0.0003 (CuBatchScoreMatrix_Init.cuh:14-15);
    one-query calls and one batch call against the oracle.  Query lengths up
    to 700 give two-stripe phase-1 codes."""
    q = SynthDB(1602, 3, 300, 0.5, 100, 700)
    d = SynthDB(1603, 120, 200, 0.668, 30, 1500)
    d.plant(q, 1604, 30, 3.0, 60.0)
    opts = dict(mapaln=0, max_map_pairs=0)
    orc = oracle(opts)
    cbp = processing(opts)
    cbp.TransferQueries(q)
    cbp.TransferCPMData(d)
    cbp.SetDbDetails(d.npos, d.nprof)
    ql = q.lengths()
    exps = []
    drawn = 0
    for qp in range(q.nprof):
        print(f"MAPALN=0: query {qp} len {int(ql[qp])}", flush=True)
        res = cbp.ProcessScoreMatrix(qp, 6.0, float(np.log(1e-2)))
        exp = search_expect(orc, q, qp, d, 6.0, float(np.log(1e-2)), d.npos, d.nprof, 0)
        assert_search_equal(res, exp, ("mapaln0", qp))
        drawn += len(exp["sel"])
        exps.append(exp)
    assert drawn >= 10
    batch = cbp.ProcessBatch(list(range(q.nprof)))
    for qp, res in enumerate(batch):
        thld, logev = cbp.score_threshold(int(ql[qp]), d.npos)
        exp = search_expect(orc, q, qp, d, thld, logev, d.npos, d.nprof, 0)
        assert_search_equal(res, exp, ("mapaln0 batch", qp))
