/*
 * c2_score.h -- one score-matrix cell recomputed from the DB chunk and the
 * prepared query (device code shared by the finalize, histogram, MAP and
 * score-block kernels).
 *
 * S(y, x) = log(sum_a q_a(y) d_a(x) / (0.5 (bq_a + bd_a))) - shift
 *           + ssswgt (sum_j p'_j QT[10j + int(9.999 p'_j)](y) + (p_H+0.1)(p'_H+0.1) 0.1111111 - 0.05)
 * (CalcSM_Init_SMEMUnroll2x, cusco/CuBatchScoreMatrix_Init.cu:45-186, and
 * CalcSM_SSSS_SMEMUnroll2x, cusco/CuBatchScoreMatrix_ssss.cu:51-280), in the
 * operation order of the fused phase-1 kernel so that every stage sees the
 * same FP32 bits.  QT is the query row's SS table sum_s p_s T[10s+int(9.999 p_s)][.]
 * prepared on the host.
 */
#ifndef C2_SCORE_H
#define C2_SCORE_H
#include "c2_dev.h"
#include "c2_kernels.h"

/* the context-vector map score of a pair of positions
 * (CalcCVS2ScoreSMEMUnroll2x, cusco/CuBatchScoreMatrix_cvs2s.cuh:47-136, and
 * GetScoreE1S1Step2Boundary, cupro/SerializedCVS2Scores.cuh:377-410):
 *   d = q.d + 1/kappa, det = (|q|^2 + 1 + 1/kappa)(|d|^2 + 1 + 1/kappa) - d^2,
 *   x = CTERM + nu log det - prior_q - prior_d (priors truncated toward zero
 *   with MAPALN = 0), the map score at key x (the first or last score beyond
 *   the keys);
 * qa = |q|^2 + 1 + 1/kappa and da = |d|^2 + 1 + 1/kappa, rq, rd the priors
 * as used.  A non-positive det (inconsistent norms) takes log(FLT_MIN). */
__device__ __forceinline__ float c2_cvs_lookup(const CvsCtx& v, float x)
{
    if (x <= v.k1) return v.v1;
    if (v.kl <= x) return v.vl;
    return v.map[v.shift + (int)rintf(2.0f * x)];
}
__device__ __forceinline__ float c2_cvs_prior(const CvsCtx& v, float p) { return v.trunc ? truncf(p) : p; }
__device__ __forceinline__ float c2_cvs_norm(const CvsCtx& v, float n) { return (n + 1.0f) + v.kinv; }
__device__ __forceinline__ float c2_cvs_x(const CvsCtx& v, const float* q, float qa, float rq,
                                          const float* d, float da, float rd)
{
    float dot = q[0] * d[0];
#pragma unroll
    for (int i = 1; i < 19; i++) dot = fmaf(q[i], d[i], dot);
    dot = dot + v.kinv;
    const float det = fmaf(qa, da, -(dot * dot));
    float x = fmaf(v.nu, c2_logf_clamped(det), v.cterm);
    x = x - rq;
    return x - rd;
}
/* the term of query row y and chunk column col (map score, unweighted) */
__device__ __forceinline__ float c2_cvs_cell(const CvsCtx& v, int Lpad, int y, int64_t col, int64_t npos)
{
    float q[19], d[19];
#pragma unroll
    for (int i = 0; i < 19; i++) {
        q[i] = v.qcv[(size_t)i * Lpad + y];
        d[i] = v.cv[(size_t)i * npos + col];
    }
    const float qa = c2_cvs_norm(v, v.qcv[(size_t)20 * Lpad + y]);
    const float rq = c2_cvs_prior(v, v.qcv[(size_t)19 * Lpad + y]);
    const float da = c2_cvs_norm(v, v.cvn[col]);
    const float rd = c2_cvs_prior(v, v.cvp[col]);
    return c2_cvs_lookup(v, c2_cvs_x(v, q, qa, rq, d, da, rd));
}
/* adds the CV term to a cell's score: sc enters as log(...) - shift plus
 * nothing, ss as the weighted SS term (0 when off); on return sc is the
 * cell's score and ss its 'mod' value.  SS and CV terms together
 * (CalcSM_SSSS_CVS2S_SMEMUnroll2x, cusco/CuBatchScoreMatrix_ssss_cvs2s.cu:255-271):
 * t = ss + wgt c, S += t, mod = t - 0.5 wgt.  CV term alone
 * (CalcSM_CVS2S_SMEMUnroll2x, cusco/CuBatchScoreMatrix_cvs2s.cu:288-299):
 * for c != 0, S += wgt c and mod = (c - 0.5) wgt (the reference adds this to
 * a mod buffer nothing initialises; here it starts at 0) */
__device__ __forceinline__ void c2_cvs_combine(float wgt, float ssswgt, float c, float& sc, float& ss)
{
    if (ssswgt > 0.0f) {
        const float t = ss + c * wgt;
        sc = sc + t;
        ss = t - 0.5f * wgt;
    } else if (c != 0.0f) {
        sc = sc + c * wgt;
        ss = (c - 0.5f) * wgt;
    }
}

/* score of cell (y, col) of the query vs chunk profile p; *ssout receives the
 * SS and CV component (the reference's 'mod' matrix value for this cell) */
__device__ __forceinline__ float c2_score_cell(const ScoreCtx& c, int y, int64_t col, int p, float* ssout)
{
    float acc = 0.0f;
#pragma unroll
    for (int k = 0; k < 20; k++) {
        const float w = 1.0f / (0.5f * (c.qbkg[k] + c.bkg[(size_t)k * c.nprof + p]));
        const float dw = c.trgfrq[(size_t)k * c.npos + col] * w;
        const float qk = c.qv[(size_t)k * c.Lpad + y];
        acc = (k == 0) ? qk * dw : fmaf(qk, dw, acc);
    }
    float sc = c2_logf_clamped(acc) - c.shift, ss = 0.0f;
    if (c.ssswgt > 0.0f) {
        float pj[3]; int co[3];
#pragma unroll
        for (int j = 0; j < 3; j++) {
            pj[j] = c.ssprb[(size_t)j * c.npos + col];
            co[j] = 10 * j + (int)(pj[j] * 9.999f);
        }
        const float dh = (pj[2] + 0.1f) * 0.1111111f;
        ss = c.QT[(size_t)co[0] * c.Lpad + y] * pj[0];
        ss = fmaf(c.QT[(size_t)co[1] * c.Lpad + y], pj[1], ss);
        ss = fmaf(c.QT[(size_t)co[2] * c.Lpad + y], pj[2], ss);
        ss = fmaf(c.qh[y], dh, ss);
        ss = (ss - 0.05f) * c.ssswgt;
        if (c.cvs.wgt <= 0.0f) sc = sc + ss;
    }
    if (c.cvs.wgt > 0.0f)
        c2_cvs_combine(c.cvs.wgt, c.ssswgt, c2_cvs_cell(c.cvs, c.Lpad, y, col, c.npos), sc, ss);
    *ssout = ss;
    return sc;
}

/* the pair's 20 weights 1/(0.5 (bq_a + bd_a)), computed as c2_score_cell does */
__device__ __forceinline__ void c2_pair_weights(const ScoreCtx& c, int p, float* w)
{
#pragma unroll
    for (int k = 0; k < 20; k++) w[k] = 1.0f / (0.5f * (c.qbkg[k] + c.bkg[(size_t)k * c.nprof + p]));
}

/* c2_score_cell with the pair's weights precomputed (same FP32 result) */
__device__ __forceinline__ float c2_score_cell_w(const ScoreCtx& c, int y, int64_t col, const float* w, float* ssout)
{
    float acc = 0.0f;
#pragma unroll
    for (int k = 0; k < 20; k++) {
        const float dw = c.trgfrq[(size_t)k * c.npos + col] * w[k];
        const float qk = c.qv[(size_t)k * c.Lpad + y];
        acc = (k == 0) ? qk * dw : fmaf(qk, dw, acc);
    }
    float sc = c2_logf_clamped(acc) - c.shift, ss = 0.0f;
    if (c.ssswgt > 0.0f) {
        float pj[3]; int co[3];
#pragma unroll
        for (int j = 0; j < 3; j++) {
            pj[j] = c.ssprb[(size_t)j * c.npos + col];
            co[j] = 10 * j + (int)(pj[j] * 9.999f);
        }
        const float dh = (pj[2] + 0.1f) * 0.1111111f;
        ss = c.QT[(size_t)co[0] * c.Lpad + y] * pj[0];
        ss = fmaf(c.QT[(size_t)co[1] * c.Lpad + y], pj[1], ss);
        ss = fmaf(c.QT[(size_t)co[2] * c.Lpad + y], pj[2], ss);
        ss = fmaf(c.qh[y], dh, ss);
        ss = (ss - 0.05f) * c.ssswgt;
        if (c.cvs.wgt <= 0.0f) sc = sc + ss;
    }
    if (c.cvs.wgt > 0.0f)
        c2_cvs_combine(c.cvs.wgt, c.ssswgt, c2_cvs_cell(c.cvs, c.Lpad, y, col, c.npos), sc, ss);
    *ssout = ss;
    return sc;
}

#endif
