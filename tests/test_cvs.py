"""The context-vector score term on the CPU: the map reader and the oracle.

* ``cbp_cvs2s_read`` on the reference's map (src/env/cvs2s, copied as
  comer2_amd/data/cvs2s): 93 keys -2.5..43.5 at step 1/2, shift 5 (the
  negative keys), every score replaced by the centred five-point mean of the
  file's scores (the first two and the last two -- the two '+' entries and
  the two numbers before them -- by the mean of the nearest full window),
  recomputed here in float64 from the file's numbers (CVS2Scores::
  SmoothScores, libpro/srcpro/CVS2Scores.cpp:366-466).
* a map with a missing key gets it inserted with the score of the key before.
* the model's constants (_TCTXVECT, libpro/srcpro/TCTXVECT.cpp:20-54):
  CTERM = (19/2 - 21/2)(ln 0.1 - ln 2.1) + ln 19! - 19 ln 2 in float64.
* the oracle's CV map score against the formula of CalcCVS2ScoreSMEMUnroll2x
  (cusco/CuBatchScoreMatrix_cvs2s.cuh:47-136) evaluated in float64, the map
  looked up at the float64 key (pairs whose key lies within 1e-3 of a bin
  edge are skipped: the FP32 key may fall on the other side); priors are
  truncated toward zero with MAPALN = 0.
* the oracle's score and mod with the term: S = S(CVSWGT=0) + t and
  mod = t - CVSWGT/2 with t = SS term + CVSWGT * c when SSSWGT > 0; with the
  SS term off, S gains CVSWGT * c and mod = (c - 1/2) CVSWGT where c != 0.
"""
import ctypes as C
import math
import os
import numpy as np
import pytest
from comer2_amd import tables
from comer2_amd._ffi import SynthDB
import oracle_ffi
from oracle_ffi import Oracle

MAP = os.path.join(os.path.dirname(tables.__file__), "data", "cvs2s")


def _raw_map():
    lines = [l for l in open(MAP).read().split("\n") if l.strip() and not l.lstrip().startswith("#")]
    i = [k for k, l in enumerate(lines) if l.startswith("Cardinality")][0]
    keys = [float(t) for t in lines[i + 1].split()]
    vals = [t for t in lines[i + 2].split()]
    return keys, vals


def test_reference_map_is_smoothed_by_centred_means():
    keys, vals, shift = tables.load_cvs2s()
    rk, rv = _raw_map()
    assert len(keys) == 93 and len(rk) == 93
    assert np.array_equal(keys, np.float32(rk))
    assert np.allclose(keys, -2.5 + 0.5 * np.arange(93))
    num = np.array([float(v) for v in rv if v != "+"])
    assert len(num) == 91 and rv[-2:] == ["+", "+"]
    want = np.empty(93)
    for n in range(93):
        c = min(max(n, 2), 88)              # centre of the nearest full window
        want[n] = num[c - 2:c + 3].mean()
    assert np.allclose(vals, want, rtol=0, atol=2e-6), np.abs(vals - want).max()
    assert shift == 5


def test_missing_key_takes_the_previous_score(tmp_path):
    p = tmp_path / "m"
    ks = [-1.0, -0.5, 0.0, 0.5, 1.5, 2.0, 2.5]          # 1.0 missing
    vs = [1.0, 2.0, 4.0, 8.0, 16.0, 32.0, 64.0]
    p.write_text("cvs2s= 1\n11\n## comment\nScale = 1.0\nDimensions = 19\nCardinality = 7\n"
                 + " ".join(map(str, ks)) + "\n" + " ".join(map(str, vs)) + "\n")
    keys, vals, shift = tables.load_cvs2s(str(p))
    assert list(keys) == [-1.0, -0.5, 0.0, 0.5, 1.0, 1.5, 2.0, 2.5]
    assert shift == 2
    m = [np.mean(vs[0:5]), np.mean(vs[1:6]), np.mean(vs[2:7])]
    want = [m[0], m[0], m[0], m[1], m[1], m[2], m[2], m[2]]
    assert np.allclose(vals, want, atol=1e-5)


def test_bad_map_is_refused(tmp_path):
    p = tmp_path / "m"
    p.write_text("cvs2s= 1\n11\nDimensions = 19\nCardinality = 3\n0 1 0.5\n1 2 3\n")
    with pytest.raises(ValueError):
        tables.load_cvs2s(str(p))


def _consts():
    L = oracle_ffi.lib()
    L.or_cv_constants.argtypes = [C.POINTER(C.c_float)] * 3
    k, nu, ct = C.c_float(), C.c_float(), C.c_float()
    L.or_cv_constants(C.byref(k), C.byref(nu), C.byref(ct))
    return k.value, nu.value, ct.value


def test_model_constants_closed_form():
    k, nu, ct = _consts()
    want = (9.5 - 10.5) * (math.log(0.1) - math.log(2.1)) + math.lgamma(20) - 19 * math.log(2)
    assert abs(k - 10.0) < 1e-6 and nu == -10.5
    assert abs(ct - want) < 2e-5, (ct, want)


def _cv_sets():
    q = SynthDB.from_lengths(7001, np.array([40, 25], np.int32)).add_cv(7101)
    d = SynthDB.from_lengths(7002, np.array([30, 20], np.int32)).add_cv(7102)
    return q, d


@pytest.mark.parametrize("mapaln", [1, 0])
def test_oracle_cv_score_matches_the_formula_in_float64(mapaln):
    keys, vals, shift = tables.load_cvs2s()
    L = oracle_ffi.lib()
    L.or_cvs_score.argtypes = [C.POINTER(oracle_ffi.Profiles), C.c_int64, C.POINTER(oracle_ffi.Profiles),
                               C.c_int64, C.POINTER(oracle_ffi.Tables), C.c_int]
    L.or_cvs_score.restype = C.c_float
    t = oracle_ffi.make_tables(tables.load_ssss(), tables.load_nn_weights(), (keys, vals, shift))
    q, d = _cv_sets()
    qv, dv = q.view, d.view
    cq = np.ctypeslib.as_array(qv.cv, (19, qv.npos))
    cd = np.ctypeslib.as_array(dv.cv, (19, dv.npos))
    pq = np.ctypeslib.as_array(qv.cvprior, (qv.npos,))
    pd = np.ctypeslib.as_array(dv.cvprior, (dv.npos,))
    nq = np.ctypeslib.as_array(qv.cvnorm2, (qv.npos,))
    nd = np.ctypeslib.as_array(dv.cvnorm2, (dv.npos,))
    kinv, nu, ct = 10.0, -10.5, (9.5 - 10.5) * (math.log(0.1) - math.log(2.1)) + math.lgamma(20) - 19 * math.log(2)
    checked = 0
    lo = hi = 0
    for i in range(qv.npos):
        for j in range(dv.npos):
            dot = float(np.dot(cq[:, i].astype(np.float64), cd[:, j].astype(np.float64))) + kinv
            det = (nq[i] + 1 + kinv) * (nd[j] + 1 + kinv) - dot * dot
            a, b = float(pq[i]), float(pd[j])
            if not mapaln:
                a, b = float(math.trunc(a)), float(math.trunc(b))
            x = ct + nu * math.log(det) - a - b
            if abs(2 * x - math.floor(2 * x) - 0.5) < 2e-3 or abs(x - keys[0]) < 1e-3 or abs(x - keys[-1]) < 1e-3:
                continue
            want = vals[0] if x <= keys[0] else (vals[-1] if x >= keys[-1] else vals[shift + int(round(2 * x))])
            got = L.or_cvs_score(C.byref(qv), i, C.byref(dv), j, C.byref(t), 0 if mapaln else 1)
            assert got == np.float32(want), (i, j, x, got, want)
            checked += 1
            lo += x <= keys[0]
            hi += x >= keys[-1]
    assert checked > 0.9 * qv.npos * dv.npos
    # the synthetic vectors span the map, not just its ends
    assert lo < checked // 4 and hi < checked // 4


@pytest.mark.parametrize("ssswgt,mapaln", [(0.12, 1), (0.0, 1), (0.12, 0)])
def test_oracle_score_adds_the_cv_term(ssswgt, mapaln):
    cvmap = tables.load_cvs2s()
    base = dict(tables.default_options(), ssswgt=ssswgt, mapaln=mapaln)
    # the log term alone, the SS term (as mod with CVSWGT = 0), both terms
    oL = Oracle(tables.load_ssss(), tables.load_nn_weights(), dict(base, ssswgt=0.0, cvswgt=0.0), cvmap)
    o0 = Oracle(tables.load_ssss(), tables.load_nn_weights(), dict(base, cvswgt=0.0), cvmap)
    o1 = Oracle(tables.load_ssss(), tables.load_nn_weights(), dict(base, cvswgt=0.15), cvmap)
    L = oracle_ffi.lib()
    L.or_cvs_score.argtypes = [C.POINTER(oracle_ffi.Profiles), C.c_int64, C.POINTER(oracle_ffi.Profiles),
                               C.c_int64, C.POINTER(oracle_ffi.Tables), C.c_int]
    L.or_cvs_score.restype = C.c_float
    q, d = _cv_sets()
    w = np.float32(0.15)
    SL, _ = oL.score_f32(q.view, 0, d.view, 1)
    S0, M0 = o0.score_f32(q.view, 0, d.view, 1)
    S1, M1 = o1.score_f32(q.view, 0, d.view, 1)
    q0, d0 = q.view.dist[0], d.view.dist[1]
    for y in range(S0.shape[0]):
        for x in range(S0.shape[1]):
            c = np.float32(L.or_cvs_score(C.byref(q.view), q0 + y, C.byref(d.view), d0 + x, C.byref(o1.t),
                                          0 if mapaln else 1))
            if ssswgt > 0:
                tt = np.float32(M0[y, x] + np.float32(c * w))
                assert S1[y, x] == np.float32(SL[y, x] + tt)
                assert M1[y, x] == np.float32(tt - np.float32(0.5) * w)
            elif c != 0:
                assert S1[y, x] == np.float32(SL[y, x] + np.float32(c * w))
                assert M1[y, x] == np.float32((c - np.float32(0.5)) * w)
            else:
                assert S1[y, x] == SL[y, x] and M1[y, x] == 0
