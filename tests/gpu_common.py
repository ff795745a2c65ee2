"""Shared helpers for GPU parity tests."""
import numpy as np
import pytest
from comer2_amd import tables
from comer2_amd._ffi import SynthDB
from oracle_ffi import Oracle


def oracle(opts=None):
    return Oracle(tables.load_ssss(), tables.load_nn_weights(), dict(tables.default_options(), **(opts or {})),
                  tables.load_cvs2s())


def expected_keys(orc, q, qp, d):
    from comer2_amd.host import encode_key
    keys = np.zeros(d.nprof, np.uint64)
    for dp in range(d.nprof):
        S, _ = orc.score_f32(q.view, qp, d.view, dp)
        m, _, _ = orc.viterbi_f32(q.view, qp, d.view, dp, S)
        keys[dp] = encode_key(m.maxsc, m.x, m.y)
    return keys


def processing(opts=None):
    from comer2_amd.host import CuBatchProcessing
    return CuBatchProcessing(0, 0, 0, options=opts)


def _pair_expect(orc, q, qp, d, dp, thld, ss, xuninf):
    """phase 1, finalize and statistics of one pair on the oracle; None when
    the pair does not pass phase 1 (restates the stage sequence of
    CuBatchProcessing::BatchProcessScoreMatrixDevice,
    cupro/CuBatchProcessing.cu:229-572)."""
    S, M = orc.score_f32(q.view, qp, d.view, dp)
    m, bt, _ = orc.viterbi_f32(q.view, qp, d.view, dp, S)
    if not (m.maxsc > thld):
        return None
    f = orc.finalize(S, bt, m)
    if not (thld < f.score):
        return None
    q0, d0 = q.view.dist[qp], d.view.dist[dp]
    qaa = bytes(q.view.aa[q0:q0 + q.view.len[qp]])
    daa = bytes(d.view.aa[d0:d0 + d.view.len[dp]])
    cnt = orc.histogram(S, qaa, daa, xuninf)
    st = orc.statistics(cnt, q.view.len[qp], d.view.len[dp], q.view.eno[qp], d.view.eno[dp],
                        f.score, float(f.psts), np.float32(ss))
    r = dict(dp=dp, fin=f, st=st)
    if not orc.o.mapaln:
        # MAPALN=0: the alignment drawn from the phase-1 backtrace (FinalizeDP_ALN)
        r["aln"] = orc.aln_final(S, bt, m, qaa, daa)
    return r


def search_expect(orc, q, qp, d, thld, logev, dblen, nseqs, cap, xuninf=1, threads=16):
    """Expected search results of query qp: passed pairs in profile order with
    their finalize records and statistics, the (log E, index) order, the MAP
    selection and, per realigned pair, the oracle's MAP record and strings."""
    from concurrent.futures import ThreadPoolExecutor
    ss, _ = orc.length_adjustment(q.view.len[qp], dblen, nseqs)
    with ThreadPoolExecutor(threads) as ex:
        rows = [r for r in ex.map(lambda dp: _pair_expect(orc, q, qp, d, dp, thld, ss, xuninf),
                                  range(d.nprof)) if r is not None]
    logE = np.array([np.float32(r["st"].Evalue) for r in rows], np.float32)
    order = np.argsort(logE, kind="stable")
    sel = [int(i) for i in order if logE[i] <= np.float32(logev)]
    if cap > 0:
        sel = sel[:cap]

    def one_map(i):
        r = rows[i]
        if "aln" in r:
            a, strs = r["aln"]
            return dict(alnlen=a.alnlen, idts=a.idts, psts=a.psts, gaps=a.gaps,
                        beg=(a.begx << 16) | a.begy, end=(a.endx << 16) | a.endy), strs
        S, M = orc.score_f32(q.view, qp, d.view, r["dp"])
        mr, _, _, _, aln = orc.map(q.view, qp, d.view, r["dp"], S, M, np.float32(r["st"].EpA))
        return (dict(alnlen=mr.alnlen, idts=mr.idts, psts=mr.psts, gaps=mr.gaps,
                     beg=(mr.begx << 16) | mr.begy, end=(mr.x << 16) | mr.y),
                tuple(bytes(aln[j]).split(b"\0")[0].decode() for j in range(3)))

    with ThreadPoolExecutor(threads) as ex:
        maps = list(ex.map(one_map, sel))
    return dict(rows=rows, order=order, sel=sel, maps=maps, ss=ss)


def assert_search_equal(res, exp, ctx, stat_rtol=1e-6):
    """GPU results equal the oracle's: the hit list (profiles, order, MAP
    selection), phase-1 records and MAP records bit for bit, strings exactly,
    statistics within stat_rtol relative (both FP64)."""
    rows = exp["rows"]
    assert res["search_space"] == exp["ss"], ctx
    assert res["npassed"] == len(rows), (ctx, res["npassed"], len(rows))
    assert [int(x) for x in res["pronr"]] == [r["dp"] for r in rows], ctx
    assert list(res["order"]) == [int(i) for i in exp["order"]], ctx
    assert list(res["map_rows"]) == exp["sel"], ctx
    sel = set(exp["sel"])
    for k, r in enumerate(rows):
        f, st = r["fin"], r["st"]
        c = ctx + (k, r["dp"])
        assert res["score"][k].view(np.uint32) == np.float32(f.score).view(np.uint32), c
        if k not in sel:
            assert res["psts"][k] == f.psts, c
            assert res["beg"][k] == ((f.begx << 16) | f.begy), c
            assert res["end"][k] == ((f.endx << 16) | f.endy), c
            assert res["alnlen"][k] == 0, c
        for name, val in (("lmbd", st.Lmbd), ("K", st.K), ("H", st.H), ("E", st.E),
                          ("logE", st.Evalue), ("epa", st.EpA), ("smin", st.Min), ("smax", st.Max),
                          ("lmbdest", st.LmbdEst), ("kest", st.KEst)):
            got = float(res[name][k])
            assert abs(got - val) <= stat_rtol * max(1.0, abs(val)), (c, name, got, val)
    for j, (i, (mr, strs)) in enumerate(zip(exp["sel"], exp["maps"])):
        c = ctx + ("drawn", i, rows[i]["dp"])
        assert res["alnlen"][i] == mr["alnlen"], c
        assert res["idnts"][i] == mr["idts"], c
        assert res["psts"][i] == mr["psts"], c
        assert res["ngaps"][i] == mr["gaps"], c
        assert res["beg"][i] == mr["beg"], c
        assert res["end"][i] == mr["end"], c
        assert res["alns"][j] == strs, c


def lognormal_lengths(seed, n, median, sigma, lo, hi):
    rng = np.random.default_rng(seed)
    return np.clip(np.rint(median * np.exp(sigma * rng.standard_normal(n))), lo, hi).astype(np.int32)


def set_residues(P, positions, letter=b"X"):
    """overwrite residue letters of a profile set in place (the generator
    emits none of 'X')."""
    import ctypes as C
    buf = C.cast(P.view.aa, C.POINTER(C.c_char))
    for i in positions:
        buf[int(i)] = letter
