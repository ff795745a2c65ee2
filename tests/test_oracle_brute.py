"""Pins of the oracle that do not restate its own recurrences.

* Path enumeration.  On pairs of 1-4 rows and columns every path through the
  five-state model is listed explicitly, and the phase-1 Viterbi maximum
  (ExecDP_Btck_Unroll32x, cudp/CuBatchDP_init_btck.cu:303-360) must equal the
  largest path weight, and the soft-DP log Z (the MAP forward model of
  cumapdp/CuBatchMAPDP_fwd.cu:291-372 with exact log-sum-exp) the
  log-sum-exp of all path weights.  The move table below is written from the
  reference text (which neighbour each state reads, which query and DB
  transition each move adds, DPLocCacheTrnProbs at cudp/CuBatchDP_init.cuh:
  148-154 for the row of each transition); a path is a start followed by
  moves, so a wrong predecessor cell, a dropped term or a max taken in the
  wrong place fails here even though it would be shared by every DP-shaped
  restatement.
* Elementary functions.  c2_logf and c2_expf, shared by the kernels and the
  oracle, against the C library in double precision at their stated ulp
  bounds and range ends.
* SS18 E-value.  Invariants of CalculateEvalueSS18 / PredictEVDParametersSS18
  (cuss/CuBatchSS_SS18.cuh:118-355) that a plausible slip fails: monotone in
  the score and in the positive-substitution count, the Obs formula when mu
  <= 0, the variant-2 offset of log 1e-4 applied before the Brown
  combination, the asymptotic slope -1/(1.485656 scale), and the ENO net
  unused when its weight in mu and scale is zero (net blocks not swapped;
  layout cupro/CuBatchProcessing_com.cuh:44-94).
"""
import ctypes as C
import math

import numpy as np
import pytest

from comer2_amd import tables
from comer2_amd._ffi import SynthDB
from gpu_common import oracle
import oracle_ffi

MM, MI, IM, DN, ND = range(5)
TMM, TMI, TMD, TIM, TII, TDM, TDD = range(7)


def _trans(P, p):
    """[7][len+1] log transitions of profile p (row 0 = beginning)."""
    n = P.npos + P.nprof
    tr = np.ctypeslib.as_array(P.trans, (7 * n,)).reshape(7, n)
    b = P.dist[p] + p
    return tr[:, b:b + P.len[p] + 1].astype(np.float64)


def _t7(T, i):
    """transitions a DP position reads: MM, MD, IM, DM, DD at row i, MI and
    II at row i+1 (DPLocCacheTrnProbs)."""
    return {"MM": T[TMM, i], "MI": T[TMI, i + 1], "MD": T[TMD, i], "IM": T[TIM, i],
            "II": T[TII, i + 1], "DM": T[TDM, i], "DD": T[TDD, i]}


def _moves(qt, dt, th):
    """moves into a cell: (state, predecessor offset (dr, dc), predecessor
    state, added weight without the cell score); th = theta (soft DP) or
    zeros (Viterbi)."""
    return [
        # MM from the diagonal neighbour (the score is added separately)
        (MM, (1, 1), MI, qt["MM"] + dt["IM"] + th[4]),
        (MM, (1, 1), IM, qt["IM"] + dt["MM"] + th[4]),
        (MM, (1, 1), DN, qt["DM"] + dt["MM"] + th[4]),
        (MM, (1, 1), ND, qt["MM"] + dt["DM"] + th[4]),
        (MM, (1, 1), MM, qt["MM"] + dt["MM"]),
        # MI from the cell above
        (MI, (1, 0), MM, qt["MM"] + dt["MI"] + th[0]),
        (MI, (1, 0), MI, qt["MM"] + dt["II"] + th[1]),
        # IM from the cell on the left
        (IM, (0, 1), MM, qt["MI"] + dt["MM"] + th[0]),
        (IM, (0, 1), IM, qt["II"] + dt["MM"] + th[1]),
        # DN from above (query transitions only)
        (DN, (1, 0), MM, qt["MD"] + th[2]),
        (DN, (1, 0), DN, qt["DD"] + th[3]),
        # ND from the left (DB transitions only)
        (ND, (0, 1), MM, dt["MD"] + th[2]),
        (ND, (0, 1), ND, dt["DD"] + th[3]),
    ]


def _paths(S, qT, dT, th, row_m1, lleft, ldiag):
    """Every path weight ending in MM at each cell.  A path starts in MM at
    any cell with weight 0 (the local start), or in a boundary state: row -1
    holds row_m1 for every column, column -1 holds lleft[r] as the left
    neighbour of (r, 0) and ldiag[r] as its diagonal neighbour."""
    Lq, Ld = S.shape

    def ending(r, c, s):
        """all weights of paths ending in state s of cell (r, c) (r, c may be
        a boundary cell, reached only through the move that reads it)."""
        out = []
        if s == MM:
            out.append(0.0)                     # local start
        qt, dt = _t7(qT, r), _t7(dT, c)
        for (st, (dr, dc), ps, w) in _moves(qt, dt, th):
            if st != s:
                continue
            pr, pc = r - dr, c - dc
            add = w + (S[r, c] if s == MM else 0.0)
            if pr < 0 and pc < 0:               # diagonal of (0, 0)
                src = [ldiag[0][ps]]
            elif pr < 0:                        # row -1
                src = [row_m1[ps]]
            elif pc < 0:                        # column -1
                src = [(ldiag if (dr, dc) == (1, 1) else lleft)[r][ps]]
            else:
                src = ending(pr, pc, ps)
            out.extend(v + add for v in src)
        return out

    return {(r, c): ending(r, c, MM) for r in range(Lq) for c in range(Ld)}


def _corner_mask(S):
    """corner cells of CuBatchDP_init_btck.cuh:117-170 get score -9:
    edge e = min(60, rint(0.34 max(Lq, Ld)))."""
    Lq, Ld = S.shape
    e = min(60, int(np.rint(np.float32(0.34) * np.float32(max(Lq, Ld)))))
    T = S.copy()
    for y in range(Lq):
        for x in range(Ld):
            if (Lq <= y + e and x <= e + y - Lq) or (Ld - e + y <= x and y < e):
                T[y, x] = -9.0
    return T


SHAPES = [(1, 1), (1, 3), (2, 2), (2, 4), (3, 3), (4, 2), (3, 4)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_viterbi_is_max_over_paths(shape, seed):
    Lq, Ld = shape
    q = SynthDB.from_lengths(500 + seed, [Lq])
    d = SynthDB.from_lengths(600 + seed, [Ld])
    rng = np.random.default_rng(seed)
    S = rng.normal(0.5, 1.5, (Lq, Ld))
    orc = oracle()
    ll, ld = orc.colm1(q.view, 0)
    paths = _paths(_corner_mask(S), _trans(q.view, 0), _trans(d.view, 0), np.zeros(5),
                   np.zeros(5), ll.astype(np.float64), ld.astype(np.float64))
    brute = max(0.0, max(max(v) for v in paths.values()))
    got = orc.viterbi_f64(q.view, 0, d.view, 0, S)
    assert got == pytest.approx(brute, rel=1e-12, abs=1e-12)
    # the FP32 mirror (what the kernels reproduce bit for bit) agrees too
    m, _, _ = orc.viterbi_f32(q.view, 0, d.view, 0, S.astype(np.float32))
    assert m.maxsc == pytest.approx(brute, rel=1e-5, abs=1e-5)


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("seed", [0, 1])
def test_soft_logz_is_lse_over_paths(shape, seed):
    Lq, Ld = shape
    q = SynthDB.from_lengths(700 + seed, [Lq])
    d = SynthDB.from_lengths(800 + seed, [Ld])
    rng = np.random.default_rng(10 + seed)
    S = rng.normal(0.0, 1.0, (Lq, Ld))
    theta = rng.normal(0.0, 0.3, 5)
    orc = oracle()
    ll, ld = orc.map_colm1(q.view, 0)
    row_m1 = np.array([0.0, -99.0, -99.0, -99.0, -99.0])     # MAP forward's row -1
    paths = _paths(S, _trans(q.view, 0), _trans(d.view, 0), theta, row_m1,
                   ll.astype(np.float64), ld.astype(np.float64))
    w = np.concatenate([np.asarray(v) for v in paths.values()])
    m = w.max()
    brute = m + math.log(np.exp(w - m).sum())
    got = orc.sdp_logz(q.view, 0, d.view, 0, S, theta)
    assert got == pytest.approx(brute, rel=1e-12, abs=1e-12)


def test_path_enumeration_sees_a_planted_slip():
    """The enumeration is sharp enough to see a one-term slip: dropping the
    ND -> MM move changes the maximum for some pair."""
    orc = oracle()
    seen = False
    for seed in range(12):
        q = SynthDB.from_lengths(900 + seed, [3])
        d = SynthDB.from_lengths(950 + seed, [4])
        S = np.random.default_rng(seed).normal(0.5, 1.5, (3, 4))
        ll, ld = orc.colm1(q.view, 0)
        full = _paths(_corner_mask(S), _trans(q.view, 0), _trans(d.view, 0), np.zeros(5),
                      np.zeros(5), ll.astype(np.float64), ld.astype(np.float64))
        global _moves
        keep = _moves
        try:
            _moves = lambda qt, dt, th: [m for m in keep(qt, dt, th) if not (m[0] == MM and m[2] == ND)]
            cut = _paths(_corner_mask(S), _trans(q.view, 0), _trans(d.view, 0), np.zeros(5),
                         np.zeros(5), ll.astype(np.float64), ld.astype(np.float64))
        finally:
            _moves = keep
        a = max(0.0, max(max(v) for v in full.values()))
        b = max(0.0, max(max(v) for v in cut.values()))
        if a != b:
            seen = True
            assert orc.viterbi_f64(q.view, 0, d.view, 0, S) == pytest.approx(a, abs=1e-12)
            break
    assert seen


# ---------------------------------------------------------------------------
# elementary functions

def _ulp_err(got, x, fn):
    ref = fn(x.astype(np.float64))
    g = got.astype(np.float64)
    ulp = np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64)
    return np.abs(g - ref) / ulp


def _vf(name, x):
    L = oracle_ffi.lib()
    f = getattr(L, name)
    f.argtypes = [np.ctypeslib.ndpointer(np.float32, flags="C"),
                  np.ctypeslib.ndpointer(np.float32, flags="C"), C.c_int64]
    x = np.ascontiguousarray(x, np.float32)
    y = np.empty_like(x)
    f(x, y, x.size)
    return y


def test_logf_within_stated_ulp():
    rng = np.random.default_rng(3)
    x = np.concatenate([
        np.exp(rng.uniform(np.log(1e-35), np.log(1e35), 2_000_000)).astype(np.float32),
        rng.uniform(0.5, 2.0, 500_000).astype(np.float32),
        np.array([1e-35, 1.17549435e-38, 0.70710677, 0.7071068, 1.4142135, 1.4142137,
                  1.0, 2.0, 0.5, 3.4e38], np.float32)])
    x = x[(x >= np.float32(1.17549435e-38))]
    y = _vf("or_vlogf", x)
    err = _ulp_err(y, x, np.log)
    nz = x != np.float32(1.0)
    assert err[nz].max() <= 1.5, float(err.max())
    assert _vf("or_vlogf", np.array([1.0], np.float32))[0] == 0.0


def test_expf_within_stated_ulp_and_range():
    rng = np.random.default_rng(4)
    x = np.concatenate([rng.uniform(-87.0, 88.0, 2_000_000).astype(np.float32),
                        rng.uniform(-0.5, 0.5, 200_000).astype(np.float32),
                        np.array([-87.0, 88.0, 0.0, -16.0, 1e-8, -1e-8], np.float32)])
    y = _vf("or_vexpf", x)
    assert _ulp_err(y, x, np.exp).max() <= 1.02
    # documented range [-87.68, 88.37]: finite and accurate up to its ends
    ends = np.array([-87.68, -87.5, 88.3, 88.37], np.float32)
    ye = _vf("or_vexpf", ends)
    assert np.all(np.isfinite(ye))
    rel = np.abs(ye.astype(np.float64) - np.exp(ends.astype(np.float64))) / np.exp(ends.astype(np.float64))
    assert rel.max() < 1e-6


def test_lse1p_table_within_stated_error():
    """c2_lse1p, the MAP passes' log(1 + e^d): 1.2e-7 absolute over (-16, 0]
    against log1p(exp(d)) in double, interval edges and the clamped top
    interval included; exactly 0 at and below the cut-off -16; never
    decreasing by more than rounding over a dense grid."""
    rng = np.random.default_rng(5)
    edges = np.float32(-16.0) + np.float32(0.25) * np.arange(1, 65, dtype=np.float32)
    x = np.concatenate([rng.uniform(-16.0, 0.0, 2_000_000).astype(np.float32),
                        edges, np.nextafter(edges, np.float32(-20.0)),
                        np.array([0.0, -1e-8, -15.999999, np.nextafter(np.float32(-16.0), np.float32(0.0))],
                                 np.float32)])
    x = x[x > np.float32(-16.0)]
    y = _vf("or_vlse1p", x)
    err = np.abs(y.astype(np.float64) - np.log1p(np.exp(x.astype(np.float64))))
    assert err.max() <= 1.2e-7, float(err.max())
    below = np.array([-16.0, -16.5, -99.0, -1e30], np.float32)
    assert np.all(_vf("or_vlse1p", below) == 0.0)
    g = np.linspace(-16.0, 0.0, 400_001).astype(np.float32)[1:]
    yg = _vf("or_vlse1p", g).astype(np.float64)
    assert np.diff(yg).min() >= -1.2e-7


# ---------------------------------------------------------------------------
# SS18 E-value invariants

def _hist_case(orc):
    q = SynthDB(41, 1, 220, 0.0, 220, 220)
    d = SynthDB(42, 4, 250, 0.4, 150, 400)
    S, _ = orc.score_f32(q.view, 0, d.view, 1)
    qaa = bytes(q.view.aa[0:q.view.len[0]])
    daa = bytes(d.view.aa[d.view.dist[1]:d.view.dist[1] + d.view.len[1]])
    cnt = orc.histogram(S, qaa, daa, 1)
    return cnt, int(q.view.len[0]), int(d.view.len[1])


def _ev(orc, cnt, Lq, Ld, score, psts, qeno=7.0, deno=9.0, ss=1.0e9):
    return orc.statistics(cnt, Lq, Ld, qeno, deno, score, psts, ss).Evalue


def _oracle_with(nnw_mod=None, **opts):
    nnw = tables.load_nn_weights().copy()
    if nnw_mod is not None:
        nnw_mod(nnw)
    return oracle_ffi.Oracle(tables.load_ssss(), nnw, dict(tables.default_options(), **opts))


@pytest.mark.parametrize("model", [1, 2])
def test_ss18_monotone_in_score_and_positives(model):
    orc = _oracle_with(ssemodel=model)
    cnt, Lq, Ld = _hist_case(orc)
    assert orc.statistics(cnt, Lq, Ld, 7.0, 9.0, 30.0, 10.0, 1e9).Lmbd > 0
    ev = np.array([_ev(orc, cnt, Lq, Ld, s, 20.0) for s in np.linspace(5.0, 400.0, 80)])
    # non-increasing; strictly decreasing once below the cap at low scores
    # (log E = 77 + log(ss/(Lq Ld)) when the combined p-value reaches 1)
    assert np.all(np.diff(ev) <= 0)
    assert np.all(np.diff(ev[ev < ev[0]]) < 0) and ev[-1] < ev[0] - 10
    evp = [_ev(orc, cnt, Lq, Ld, 60.0, p) for p in range(0, 200, 5)]
    assert np.all(np.diff(evp) <= 0) and evp[-1] < evp[0]


def test_ss18_falls_back_to_obs_when_mu_not_positive():
    def neg_mu(w):
        for base in (75, 84):             # both variants: slopes 0, intercepts -1
            w[base + 0] = 0.0; w[base + 1] = -1.0
            w[base + 2] = 0.0; w[base + 3] = -1.0
    for model in (1, 2):
        orc = _oracle_with(neg_mu, ssemodel=model)
        cnt, Lq, Ld = _hist_case(orc)
        o = tables.default_options()
        for sc, ss in ((25.0, 1e9), (80.0, 3.3e7)):
            obs = math.log(o["gap_K"] * np.float32(ss)) - o["gap_lambda"] * sc
            got = _ev(orc, cnt, Lq, Ld, sc, 30.0, ss=ss)
            assert got == pytest.approx(obs, rel=1e-6, abs=1e-6)
    orc0 = _oracle_with(ssemodel=0)
    cnt, Lq, Ld = _hist_case(orc0)
    o = tables.default_options()
    assert _ev(orc0, cnt, Lq, Ld, 40.0, 5.0) == pytest.approx(
        math.log(o["gap_K"] * 1e9) - o["gap_lambda"] * 40.0, rel=1e-6)


def test_ss18_variant_offset_and_asymptotic_slope():
    """With both variants' adjustment blocks equal, variant 2 is variant 1 at
    a score shifted by 9.210340372 scale (the log 1e-4 offset sits in log E
    before the beta/gamma combination), the same shift at every score; and
    the slope of log E at large scores is -1/(1.485656 scale) for that
    scale."""
    def same_adj(w):
        w[75:84] = w[84:93]
    o1 = _oracle_with(same_adj, ssemodel=1)
    o2 = _oracle_with(same_adj, ssemodel=2)
    cnt, Lq, Ld = _hist_case(o1)

    def shift_at(sc):
        target = _ev(o2, cnt, Lq, Ld, sc, 15.0)
        lo, hi = 0.0, 200.0
        for _ in range(200):
            mid = 0.5 * (lo + hi)
            if _ev(o1, cnt, Lq, Ld, sc + mid, 15.0) > target:
                lo = mid
            else:
                hi = mid
        return 0.5 * (lo + hi)

    d1, d2 = shift_at(30.0), shift_at(70.0)
    assert d1 == pytest.approx(d2, rel=1e-5)
    scale = d1 / 9.210340372
    assert 0.1 <= scale <= 5.0
    s1, s2 = 4000.0, 4400.0
    slope = (_ev(o1, cnt, Lq, Ld, s2, 15.0) - _ev(o1, cnt, Lq, Ld, s1, 15.0)) / (s2 - s1)
    assert slope == pytest.approx(-1.0 / (1.485656 * scale), rel=2e-3)


def test_ss18_eno_net_unused_without_weight():
    """alpha = 0 and ibeta = 1 give mu = muR and scale = scaleR: the E-value
    then does not depend on the ENO net's weights (0..24) nor on the ENOs,
    and does depend on the two compositional nets (25..74)."""
    def only_r(w):
        w[84 + 7] = 0.0
        w[84 + 8] = 1.0
    base = _oracle_with(only_r, ssemodel=2)
    cnt, Lq, Ld = _hist_case(base)
    e0 = _ev(base, cnt, Lq, Ld, 45.0, 25.0)

    def perturbed(lo, hi):
        def f(w):
            only_r(w)
            w[lo:hi] += np.float32(0.37)
        return _ev(_oracle_with(f, ssemodel=2), cnt, Lq, Ld, 45.0, 25.0)

    assert perturbed(0, 25) == e0
    assert _ev(base, cnt, Lq, Ld, 45.0, 25.0, qeno=2.0, deno=13.0) == e0
    assert perturbed(25, 50) != e0
    assert perturbed(50, 75) != e0
    # with the shipped weights the ENOs do matter
    full = _oracle_with(ssemodel=2)
    assert _ev(full, cnt, Lq, Ld, 45.0, 25.0, qeno=2.0, deno=13.0) != _ev(full, cnt, Lq, Ld, 45.0, 25.0)
