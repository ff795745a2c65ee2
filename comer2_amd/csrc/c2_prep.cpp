/*
 * c2_prep.cpp -- per-query tables consumed by the phase-1 kernel.
 *
 *  - qv, qtr: target frequencies and the 7 DP transitions per row, as
 *    DPLocCacheTrnProbs caches them (cudp/CuBatchDP_init.cuh:190-201);
 *    padding rows get log0 = -32768 transitions, which pins their MM state
 *    to 0 (no score can be reached from them).
 *  - QT[g][r] = sum_i p_i(r) T[10 i + int(9.999 p_i(r))][g]: the query half
 *    of the SS score of CalcSSSScoreSMEMUnroll2x
 *    (cusco/CuBatchScoreMatrix_ssss.cuh:48-131), so a cell needs 3 lookups.
 *  - lleft/ldiag: the reference's column -1 boundary.  ExecDP_Btck_Unroll32x
 *    computes cells left of column 0 inside the x0=-32 block of each 32-row
 *    block row, with log0 DB transitions, score -9 and zero neighbours
 *    outside the block; (r,0) sees (r,-1) as its left neighbour except on
 *    rows r%32==31 (zeros), and (r-1,-1) as its diagonal except on rows
 *    r%32 in {0,31} (zeros) (cudp/CuBatchDP_init_btck.cu:203-240).
 */
#include <cmath>
#include <cstring>
#include <algorithm>
#include "c2_prep.h"
#include "c2_kernels.h"
#include "c2_fmath.h"

namespace {
enum { TMM = 0, TMI, TMD, TIM, TII, TDM, TDD };

inline float trn(const cbp_profiles* P, int p, int k, int i) {
    return P->trans[(size_t)k * (size_t)(P->npos + P->nprof) + P->dist[p] + p + i];
}
void trn7(const cbp_profiles* P, int p, int i, float t[7]) {
    t[0] = trn(P, p, TMM, i);     t[1] = trn(P, p, TMI, i + 1);
    t[2] = trn(P, p, TMD, i);     t[3] = trn(P, p, TIM, i);
    t[4] = trn(P, p, TII, i + 1); t[5] = trn(P, p, TDM, i);
    t[6] = trn(P, p, TDD, i);
}
inline float mx(float a, float b) { return a < b ? b : a; }

struct St { float s[5]; };   /* MM MI IM DN ND */

St cell(const St& dg, const St& up, const St& lf, const float* qt, const float* dt, float sc) {
    St n; float v1, v2;
    v1 = dg.s[1] + qt[0] + dt[3];
    v2 = dg.s[2] + qt[3] + dt[0]; v1 = mx(v1, v2);
    v2 = dg.s[3] + qt[5] + dt[0]; v1 = mx(v1, v2);
    v2 = dg.s[4] + qt[0] + dt[5]; v1 = mx(v1, v2);
    v2 = dg.s[0] + qt[0] + dt[0]; v1 = mx(v1, v2);
    n.s[0] = mx(v1 + sc, 0.0f);
    n.s[1] = mx(up.s[0] + qt[0] + dt[1], up.s[1] + qt[0] + dt[4]);
    n.s[2] = mx(lf.s[0] + qt[1] + dt[0], lf.s[2] + qt[4] + dt[0]);
    n.s[3] = mx(up.s[0] + qt[2], up.s[3] + qt[6]);
    n.s[4] = mx(lf.s[0] + dt[2], lf.s[4] + dt[6]);
    return n;
}
}  // namespace

void c2_colm1(const cbp_profiles* q, int qp, float* lleft, float* ldiag)
{
    const int Lq = q->len[qp];
    const St zero = {{0, 0, 0, 0, 0}};
    float dt[7];
    for (int k = 0; k < 7; k++) dt[k] = -32768.0f;
    std::vector<St> V(32 * 33);     /* [row in block][column + 32] */
    for (int b = 0; 32 * b < Lq; b++) {
        for (int rr = 0; rr < 32 && 32 * b + rr < Lq; rr++) {
            const int r = 32 * b + rr, first = 31 - rr;   /* first column index (+32) */
            float qt[7]; trn7(q, qp, r, qt);
            for (int cc = first; cc < 32; cc++) {
                const St& up = (rr > 0 && cc > first) ? V[(rr - 1) * 33 + cc] : zero;
                const St& dg = (rr > 0 && cc - 1 > first) ? V[(rr - 1) * 33 + cc - 1] : zero;
                const St& lf = (cc > first) ? V[rr * 33 + cc - 1] : zero;
                V[rr * 33 + cc] = cell(dg, up, lf, qt, dt, -9.0f);
            }
            for (int s = 0; s < 5; s++) {
                lleft[r * 5 + s] = (rr == 31) ? 0.0f : V[rr * 33 + 31].s[s];
                ldiag[r * 5 + s] = (rr == 0 || rr == 31) ? 0.0f : V[(rr - 1) * 33 + 31].s[s];
            }
        }
    }
}

namespace {
inline float mexp(float x) { return (-16.0f < x) ? c2_expf(x) : 0.0f; }
inline float mlog(float x) { return (3.8e-11f < x) ? c2_logf(x) : -99.0f; }
inline float lse2(float a, float b) { float m = mx(a, b); return mlog(mexp(a - m) + mexp(b - m)) + m; }
/* forward cell of ExecMAPDP_Fwd_Unroll32x (cumapdp/CuBatchMAPDP_fwd.cu:291-372) */
St fcell(const St& dg, const St& up, const St& lf, const float* qt, const float* dt, float sco) {
    St n; float v1, v2, v3, v4, v5, m = 0.0f;
    v1 = dg.s[1] + qt[0] + dt[3] + sco; m = mx(m, v1);
    v2 = dg.s[2] + qt[3] + dt[0] + sco; m = mx(m, v2);
    v3 = dg.s[3] + qt[5] + dt[0] + sco; m = mx(m, v3);
    v4 = dg.s[4] + qt[0] + dt[5] + sco; m = mx(m, v4);
    v5 = dg.s[0] + qt[0] + dt[0] + sco; m = mx(m, v5);
    v1 = mexp(v1 - m) + mexp(v2 - m) + mexp(v3 - m) + mexp(v4 - m) + mexp(v5 - m) + mexp(-m);
    n.s[0] = mlog(v1) + m;
    n.s[1] = lse2(up.s[0] + qt[0] + dt[1], up.s[1] + qt[0] + dt[4]);
    n.s[2] = lse2(lf.s[0] + qt[1] + dt[0], lf.s[2] + qt[4] + dt[0]);
    n.s[3] = lse2(up.s[0] + qt[2], up.s[3] + qt[6]);
    n.s[4] = lse2(lf.s[0] + dt[2], lf.s[4] + dt[6]);
    return n;
}
}  // namespace

/* The MAP forward pass has the phase-1 column -1 structure with the MAP
 * kernels' initial values (MM 0, other states log 0 = -99) and out-of-range
 * transitions and scores of -99 (cumapdp/CuBatchMAPDP_fwd.cu:210-268). */
void c2_map_colm1(const cbp_profiles* q, int qp, float* mleft, float* mdiag)
{
    const int Lq = q->len[qp];
    const St init = {{0.0f, -99.0f, -99.0f, -99.0f, -99.0f}};
    float dt[7];
    for (int k = 0; k < 7; k++) dt[k] = -99.0f;
    std::vector<St> V(32 * 33);
    for (int b = 0; 32 * b < Lq; b++)
        for (int rr = 0; rr < 32 && 32 * b + rr < Lq; rr++) {
            const int r = 32 * b + rr, first = 31 - rr;
            float qt[7]; trn7(q, qp, r, qt);
            for (int cc = first; cc < 32; cc++) {
                const St& up = (rr > 0 && cc > first) ? V[(rr - 1) * 33 + cc] : init;
                const St& dg = (rr > 0 && cc - 1 > first) ? V[(rr - 1) * 33 + cc - 1] : init;
                const St& lf = (cc > first) ? V[rr * 33 + cc - 1] : init;
                V[rr * 33 + cc] = fcell(dg, up, lf, qt, dt, -99.0f);
            }
            for (int s = 0; s < 5; s++) {
                mleft[r * 5 + s] = (rr == 31) ? init.s[s] : V[rr * 33 + 31].s[s];
                mdiag[r * 5 + s] = (rr == 0 || rr == 31) ? init.s[s] : V[(rr - 1) * 33 + 31].s[s];
            }
        }
}

void c2_geometry(int Lq, int rmax, int* R, int* nstripes)
{
    if (rmax < 1) rmax = 1;
    if (rmax > 5) rmax = 5;
    for (;;) {
        int ns = (Lq + 64 * rmax - 1) / (64 * rmax);
        if (ns < 1) ns = 1;
        int r = (Lq + 64 * ns - 1) / (64 * ns);
        if (r < 1) r = 1;
        /* a query of several stripes also stages its top boundary in the LDS
         * ring: fewer rows per lane rather than fewer waves per workgroup */
        if (ns > 1 && rmax > 1 && c2_phase1_lds_bytes(r, 8, ns) > 160 * 1024) { rmax--; continue; }
        *R = r; *nstripes = ns;
        return;
    }
}

void c2_prepare_query(const cbp_profiles* q, int qp, const float* ssss, int rmax, QueryPrep* o)
{
    const int Lq = q->len[qp];
    const size_t q0 = q->dist[qp];
    c2_geometry(Lq, rmax, &o->R, &o->nstripes);
    o->Lq = Lq;
    o->Lpad = o->nstripes * 64 * o->R;
    const int Lp = o->Lpad;
    o->qv.assign((size_t)20 * Lp, 0.0f);
    o->qtr.assign((size_t)7 * Lp, -32768.0f);
    o->qtrb.assign((size_t)7 * Lp, -99.0f);
    o->QT.assign((size_t)30 * Lp, 0.0f);
    o->qh.assign((size_t)Lp, 0.0f);
    o->lleft.assign((size_t)5 * Lp, 0.0f);
    o->ldiag.assign((size_t)5 * Lp, 0.0f);
    o->qbkg.resize(20);
    for (int a = 0; a < 20; a++) o->qbkg[a] = q->bkg[(size_t)a * q->nprof + qp];
    std::vector<float> ll((size_t)5 * Lq), ld((size_t)5 * Lq);
    c2_colm1(q, qp, ll.data(), ld.data());
    std::vector<float> ml((size_t)5 * Lq), md((size_t)5 * Lq);
    c2_map_colm1(q, qp, ml.data(), md.data());
    o->mleft.assign((size_t)5 * Lq, 0.0f);
    o->mdiag.assign((size_t)5 * Lq, 0.0f);
    for (int r = 0; r < Lq; r++)
        for (int s = 0; s < 5; s++) {
            o->mleft[(size_t)s * Lq + r] = ml[(size_t)r * 5 + s];
            o->mdiag[(size_t)s * Lq + r] = md[(size_t)r * 5 + s];
        }
    for (int r = 0; r < Lq; r++) {
        for (int a = 0; a < 20; a++) o->qv[(size_t)a * Lp + r] = q->trgfrq[(size_t)a * q->npos + q0 + r];
        float t[7]; trn7(q, qp, r, t);
        for (int k = 0; k < 7; k++) o->qtr[(size_t)k * Lp + r] = t[k];
        for (int k = 0; k < 7; k++) o->qtrb[(size_t)k * Lp + r] = trn(q, qp, k, r + 1);
        float p[3]; int ri[3];
        for (int i = 0; i < 3; i++) {
            p[i] = q->ssprb[(size_t)i * q->npos + q0 + r];
            ri[i] = i * 10 + (int)(p[i] * 9.999f);
        }
        for (int g = 0; g < 30; g++) {
            float v = p[0] * ssss[ri[0] * 30 + g];
            v = fmaf(p[1], ssss[ri[1] * 30 + g], v);
            v = fmaf(p[2], ssss[ri[2] * 30 + g], v);
            o->QT[(size_t)g * Lp + r] = v;
        }
        o->qh[r] = p[2] + 0.1f;
        for (int s = 0; s < 5; s++) {
            o->lleft[(size_t)s * Lp + r] = ll[(size_t)r * 5 + s];
            o->ldiag[(size_t)s * Lp + r] = ld[(size_t)r * 5 + s];
        }
    }
}
