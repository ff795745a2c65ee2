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

/* score of cell (y, col) of the query vs chunk profile p; *ssout receives the
 * SS component (the reference's 'mod' matrix value for this cell) */
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
        sc = sc + ss;
    }
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
        sc = sc + ss;
    }
    *ssout = ss;
    return sc;
}

#endif
