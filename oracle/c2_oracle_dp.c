/*
 * c2_oracle_dp.c -- oracle for the score matrix and the phase-1 Viterbi DP.
 * TEST INFRASTRUCTURE ONLY (see c2_oracle.h).
 */
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include "c2_oracle.h"
#include "../comer2_amd/csrc/c2_fmath.h"

enum { TMM = 0, TMI, TMD, TIM, TII, TDM, TDD };
enum { SMM = 0, SMI, SIM, SDN, SND };

/* transition k at row i (row 0 = beginning) of profile p; the layout keeps
 * len+1 rows per profile (cupro/PMBatchProData.h:412) */
static inline float trn(const cbp_profiles* P, int p, int k, int i) {
    return P->trans[(size_t)k * (size_t)(P->npos + P->nprof) + P->dist[p] + p + i];
}

/* the 7 transitions used at a DP position, as cached by DPLocCacheTrnProbs
 * (cudp/CuBatchDP_init.cuh:190-201): MMp MIc MDp IMp IIc DMp DDp */
static void trn7(const cbp_profiles* P, int p, int i, float t[7]) {
    t[0] = trn(P, p, TMM, i);     t[1] = trn(P, p, TMI, i + 1);
    t[2] = trn(P, p, TMD, i);     t[3] = trn(P, p, TIM, i);
    t[4] = trn(P, p, TII, i + 1); t[5] = trn(P, p, TDM, i);
    t[6] = trn(P, p, TDD, i);
}
static void trn7_log0(float t[7]) { for (int k = 0; k < 7; k++) t[k] = -32768.0f; }

/* ---------------------------------------------------------------------- */
/* scores                                                                   */

/* SS table row/column index, GetSSSSTableIndex (cusco/CuBatchScoreMatrix_ssss.cuh:25-31) */
static inline int ssidx(int s, float p) { return s * 10 + (int)(p * 9.999f); }

/* FP32 restatement of CalcSM_Init_SMEMUnroll2x (cusco/CuBatchScoreMatrix_Init.cu:45-186)
 * followed by CalcSM_SSSS_SMEMUnroll2x (cusco/CuBatchScoreMatrix_ssss.cu:51-280), in the
 * fixed operation order of the gfx950 kernels:
 *   w_a = 1/(0.5(bq_a+bd_a)), d'_a = d_a w_a, acc = fma chain over a = 0..19,
 *   S = log(acc) - shift;  QT_r[g] = sum_i p_i T[idx_i][g] (fma chain),
 *   ss = sum_j p'_j QT_r[idx_j] + (p_H+0.1)((p'_H+0.1)*0.1111111), ss = (ss-0.05)*w */
void or_cv_constants(float* kinv, float* nu, float* cterm)
{
    /* _TCTXVECT: DIM 19, KAPPA0 0.1, NU0 = DIM, loKAPPA0 = 1/KAPPA0,
     * PowerNU0 = -0.5 (NU0 + 2), CTERM = (0.5 DIM + PowerNU0)(ln KAPPA0 -
     * ln(KAPPA0 + 2)) + sum_{i=1..DIM} (ln(NU0 + 1 - i) - ln 2), in float */
    const float kappa = 0.1f, dimf = 19.0f;
    float pw = -0.5f * (dimf + 2.0f);
    float ct = (0.5f * 19 + pw) * (logf(kappa) - logf(kappa + 2.0f));
    float acc = 0.0f;
    for (int i = 1; i <= 19; i++) acc += logf(dimf + 1.0f - (float)i) - 0.69314718055994530942f;
    *kinv = 1.0f / kappa; *nu = pw; *cterm = ct + acc;
}

static float cv_at(const cbp_profiles* s, int64_t pos, int i) {
    return s->cv ? s->cv[(size_t)i * s->npos + pos] : 0.17302947f;
}

float or_cvs_score(const cbp_profiles* q, int64_t qpos, const cbp_profiles* d, int64_t dpos,
                   const or_tables* t, int trunc)
{
    float kinv, nu, cterm;
    or_cv_constants(&kinv, &nu, &cterm);
    /* dot product of the two vectors plus the off-diagonal 1/kappa */
    float dot = cv_at(q, qpos, 0) * cv_at(d, dpos, 0);
    for (int i = 1; i < 19; i++) dot = fmaf(cv_at(q, qpos, i), cv_at(d, dpos, i), dot);
    dot = dot + kinv;
    const float nq = q->cv ? q->cvnorm2[qpos] : 0.0f, nd = d->cv ? d->cvnorm2[dpos] : 0.0f;
    float pq = q->cv ? q->cvprior[qpos] : 0.0f, pd = d->cv ? d->cvprior[dpos] : 0.0f;
    if (trunc) { pq = (float)(int)pq; pd = (float)(int)pd; }
    /* det(C_2 + H'H) of the shifted centring matrix */
    const float det = fmaf((nq + 1.0f) + kinv, (nd + 1.0f) + kinv, -(dot * dot));
    float x = fmaf(nu, c2_logf_clamped(det), cterm);
    x = x - pq;
    x = x - pd;
    const int card = t->cvcard;
    if (x <= t->cvkeys[0]) return t->cvvals[0];
    if (t->cvkeys[card - 1] <= x) return t->cvvals[card - 1];
    return t->cvvals[t->cvshift + (int)rintf(2.0f * x)];
}

void or_score_f32(const cbp_profiles* q, int qp, const cbp_profiles* d, int dp,
                  const or_tables* t, const cbp_options* o, float* S, float* mod)
{
    const int Lq = q->len[qp], Ld = d->len[dp];
    const size_t q0 = q->dist[qp], d0 = d->dist[dp];
    const float shift = o->mapaln ? 0.002f : 0.0003f;   /* CuBatchScoreMatrix_Init.cuh:14-15 */
    float w[20];
    for (int a = 0; a < 20; a++)
        w[a] = 1.0f / (0.5f * (q->bkg[(size_t)a * q->nprof + qp] + d->bkg[(size_t)a * d->nprof + dp]));
    float* dw = (float*)malloc(sizeof(float) * 20 * (size_t)Ld);
    float* QT = (float*)malloc(sizeof(float) * 30 * (size_t)Lq);
    float* qh = (float*)malloc(sizeof(float) * (size_t)Lq);
    int*   cj = (int*)malloc(sizeof(int) * 3 * (size_t)Ld);
    float* pj = (float*)malloc(sizeof(float) * 3 * (size_t)Ld);
    float* dh = (float*)malloc(sizeof(float) * (size_t)Ld);
    for (int c = 0; c < Ld; c++) {
        for (int a = 0; a < 20; a++)
            dw[(size_t)c * 20 + a] = d->trgfrq[(size_t)a * d->npos + d0 + c] * w[a];
        for (int j = 0; j < 3; j++) {
            float p = d->ssprb[(size_t)j * d->npos + d0 + c];
            pj[c * 3 + j] = p; cj[c * 3 + j] = ssidx(j, p);
        }
        dh[c] = (pj[c * 3 + 2] + 0.1f) * 0.1111111f;
    }
    for (int r = 0; r < Lq; r++) {
        float p[3]; int ri[3];
        for (int i = 0; i < 3; i++) { p[i] = q->ssprb[(size_t)i * q->npos + q0 + r]; ri[i] = ssidx(i, p[i]); }
        for (int g = 0; g < 30; g++) {
            float v = p[0] * t->ssss[ri[0]][g];
            v = fmaf(p[1], t->ssss[ri[1]][g], v);
            v = fmaf(p[2], t->ssss[ri[2]][g], v);
            QT[(size_t)r * 30 + g] = v;
        }
        qh[r] = p[2] + 0.1f;
    }
    for (int r = 0; r < Lq; r++) {
        float qv[20];
        for (int a = 0; a < 20; a++) qv[a] = q->trgfrq[(size_t)a * q->npos + q0 + r];
        for (int c = 0; c < Ld; c++) {
            const float* dv = dw + (size_t)c * 20;
            float acc = qv[0] * dv[0];
            for (int a = 1; a < 20; a++) acc = fmaf(qv[a], dv[a], acc);
            float s = c2_logf_clamped(acc) - shift;
            float ss = 0.0f;
            if (o->ssswgt > 0.0f) {
                const float* qt = QT + (size_t)r * 30;
                ss = qt[cj[c * 3 + 0]] * pj[c * 3 + 0];
                ss = fmaf(qt[cj[c * 3 + 1]], pj[c * 3 + 1], ss);
                ss = fmaf(qt[cj[c * 3 + 2]], pj[c * 3 + 2], ss);
                ss = fmaf(qh[r], dh[c], ss);
                ss = (ss - 0.05f) * o->ssswgt;
                if (!(o->cvswgt > 0.0f)) s = s + ss;
            }
            if (o->cvswgt > 0.0f) {
                /* CalcSM_SSSS_CVS2S_SMEMUnroll2x (cusco/CuBatchScoreMatrix_ssss_cvs2s.cu:255-271)
                 * and CalcSM_CVS2S_SMEMUnroll2x (cusco/CuBatchScoreMatrix_cvs2s.cu:288-299) */
                const float cm = or_cvs_score(q, (int64_t)q0 + r, d, (int64_t)d0 + c, t, !o->mapaln);
                if (o->ssswgt > 0.0f) {
                    const float tot = ss + cm * o->cvswgt;
                    s = s + tot;
                    ss = tot - 0.5f * o->cvswgt;
                } else if (cm != 0.0f) {
                    s = s + cm * o->cvswgt;
                    ss = (cm - 0.5f) * o->cvswgt;
                }
            }
            S[(size_t)r * Ld + c] = s;
            if (mod) mod[(size_t)r * Ld + c] = ss;
        }
    }
    free(dw); free(QT); free(qh); free(cj); free(pj); free(dh);
}

/* FP64 restatement of the same two kernels with the reference's formulas:
 *   S = log(sum_a q_a d_a / (0.5(bq_a+bd_a))) - shift
 *     + ssswgt (sum_ij p_i p'_j T[idx_i][idx_j] + (p_H+0.1)(p'_H+0.1) 0.1111111 - 0.05) */
void or_score_f64(const cbp_profiles* q, int qp, const cbp_profiles* d, int dp,
                  const or_tables* t, const cbp_options* o, double* S, double* mod)
{
    const int Lq = q->len[qp], Ld = d->len[dp];
    const size_t q0 = q->dist[qp], d0 = d->dist[dp];
    const double shift = o->mapaln ? (double)0.002f : (double)0.0003f;
    for (int r = 0; r < Lq; r++) {
        for (int c = 0; c < Ld; c++) {
            double acc = 0.0;
            for (int a = 0; a < 20; a++) {
                double bq = q->bkg[(size_t)a * q->nprof + qp], bd = d->bkg[(size_t)a * d->nprof + dp];
                acc += (double)q->trgfrq[(size_t)a * q->npos + q0 + r] *
                       (double)d->trgfrq[(size_t)a * d->npos + d0 + c] / (0.5 * (bq + bd));
            }
            double s = log(acc) - shift, ss = 0.0;
            if (o->ssswgt > 0.0f) {
                for (int i = 0; i < 3; i++) {
                    float pi = q->ssprb[(size_t)i * q->npos + q0 + r];
                    for (int j = 0; j < 3; j++) {
                        float pj = d->ssprb[(size_t)j * d->npos + d0 + c];
                        ss += (double)pi * (double)pj * (double)t->ssss[ssidx(i, pi)][ssidx(j, pj)];
                    }
                }
                ss += ((double)q->ssprb[(size_t)2 * q->npos + q0 + r] + 0.1) *
                      ((double)d->ssprb[(size_t)2 * d->npos + d0 + c] + 0.1) * (double)0.1111111f;
                ss = (ss - (double)0.05f) * (double)o->ssswgt;
                s += ss;
            }
            if (o->cvswgt > 0.0f) {
                /* the map lookup is a step function of its key: the FP32
                 * key of the kernels is used here too */
                const double cm = (double)or_cvs_score(q, (int64_t)q0 + r, d, (int64_t)d0 + c, t, !o->mapaln);
                if (o->ssswgt > 0.0f) {
                    s += cm * (double)o->cvswgt;
                    ss = ss + cm * (double)o->cvswgt - 0.5 * (double)o->cvswgt;
                } else if (cm != 0.0) {
                    s += cm * (double)o->cvswgt;
                    ss = (cm - 0.5) * (double)o->cvswgt;
                }
            }
            S[(size_t)r * Ld + c] = s;
            if (mod) mod[(size_t)r * Ld + c] = ss;
        }
    }
}

/* ---------------------------------------------------------------------- */
/* corners (CuBatchDP_init_btck.cuh:117-170)                               */

static int corner_edge(int Lq, int Ld) {
    int m = Lq > Ld ? Lq : Ld;
    int e = (int)rintf(0.34f * (float)m);           /* CUDP_CORNERLENPRC */
    return e > 60 ? 60 : e;                           /* CUDP_MAXCORNERLEN */
}
static int in_corner(int Lq, int Ld, int e, int y, int x) {
    return (Lq <= y + e && x <= e + y - Lq) || (Ld - e + y <= x && y < e);
}
void or_mask_corners_f32(int Lq, int Ld, float* S) {
    int e = corner_edge(Lq, Ld);
    for (int y = 0; y < Lq; y++)
        for (int x = 0; x < Ld; x++)
            if (in_corner(Lq, Ld, e, y, x)) S[(size_t)y * Ld + x] = -9.0f;
}
void or_mask_corners_f64(int Lq, int Ld, double* S) {
    int e = corner_edge(Lq, Ld);
    for (int y = 0; y < Lq; y++)
        for (int x = 0; x < Ld; x++)
            if (in_corner(Lq, Ld, e, y, x)) S[(size_t)y * Ld + x] = -9.0;
}

/* ---------------------------------------------------------------------- */
/* one DP cell, ExecDP_Btck_Unroll32x inner loop (cudp/CuBatchDP_init_btck.cu:291-360) */

typedef struct { float v[5]; } st5;

static inline float fmx(float a, float b) { return a < b ? b : a; }   /* myhdmax */

static st5 cell_f32(const st5* dg, const st5* up, const st5* lf,
                    const float* qt, const float* dt, float s, char* btck)
{
    st5 n; float v1, v2, mx; char b = OR_STOP;
    /* MM (diagonal) */
    v1 = dg->v[SMI] + qt[0] + dt[3];
    v2 = dg->v[SIM] + qt[3] + dt[0]; v1 = fmx(v1, v2);
    v2 = dg->v[SDN] + qt[5] + dt[0]; v1 = fmx(v1, v2);
    v2 = dg->v[SND] + qt[0] + dt[5]; v1 = fmx(v1, v2);
    v2 = dg->v[SMM] + qt[0] + dt[0]; v1 = fmx(v1, v2);
    v1 = fmx(v1 + s, 0.0f);
    n.v[SMM] = v1; mx = v1; if (mx != 0.0f) b = OR_DIAG;
    /* MI (up) */
    v1 = up->v[SMM] + qt[0] + dt[1];
    v2 = up->v[SMI] + qt[0] + dt[4]; v1 = fmx(v1, v2);
    n.v[SMI] = v1; if (mx < v1) { mx = v1; b = OR_UP; }
    /* IM (left) */
    v1 = lf->v[SMM] + qt[1] + dt[0];
    v2 = lf->v[SIM] + qt[4] + dt[0]; v1 = fmx(v1, v2);
    n.v[SIM] = v1; if (mx < v1) { mx = v1; b = OR_LEFT; }
    /* DN (up) */
    v1 = up->v[SMM] + qt[2];
    v2 = up->v[SDN] + qt[6]; v1 = fmx(v1, v2);
    n.v[SDN] = v1; if (mx < v1) { mx = v1; b = OR_UP; }
    /* ND (left) */
    v1 = lf->v[SMM] + dt[2];
    v2 = lf->v[SND] + dt[6]; v1 = fmx(v1, v2);
    n.v[SND] = v1; if (mx < v1) { mx = v1; b = OR_LEFT; }
    if (btck) *btck = b;
    return n;
}

/* Column -1 boundary.  In the reference the cells left of column 0 are
 * computed inside the x0=-32 oblique block of every 32-row block-row, with
 * out-of-range transitions log0=-32768 (DPLocInitTrnProbs), score -9 and
 * zero-initialised neighbours outside the block; the cell (r,0) then sees
 * (r,-1) as its left neighbour except on rows r%32==31, whose column 0 lies
 * in the next block and reads zeros (x-1<0 branch, CuBatchDP_init_btck.cu:
 * 203-224); its diagonal neighbour (r-1,-1) is seen as zeros on rows
 * r%32==0 (bottom buffer, :228-240) and r%32==31.  This restates that. */
static int g_flags = 0;
void or_set_flags(int flags) { g_flags = flags; }

/* the FP32 elementary functions the kernels and this restatement share
 * (c2_fmath.h), exposed element-wise so that tests can measure them against
 * the C library in double precision */
void or_vlogf(const float* x, float* y, int64_t n) { for (int64_t i = 0; i < n; i++) y[i] = c2_logf(x[i]); }
void or_vexpf(const float* x, float* y, int64_t n) { for (int64_t i = 0; i < n; i++) y[i] = c2_expf(x[i]); }

void or_colm1(const cbp_profiles* q, int qp, float (*lleft)[5], float (*ldiag)[5])
{
    const int noquirk = g_flags & 1;
    const int Lq = q->len[qp];
    st5 zero = {{0, 0, 0, 0, 0}};
    float dt[7]; trn7_log0(dt);
    st5 V[32][32];  /* V[rr][cc]: row 32k+rr, column cc-32 */
    for (int k = 0; k * 32 < Lq; k++) {
        for (int rr = 0; rr < 32 && 32 * k + rr < Lq; rr++) {
            int r = 32 * k + rr, t = 31 - rr;
            float qt[7]; trn7(q, qp, r, qt);
            for (int cc = t; cc < 32; cc++) {          /* column cc-32 in [-32+t, -1] */
                const st5* up = (rr > 0 && cc >= t + 1) ? &V[rr - 1][cc] : &zero;
                const st5* dg = (rr > 0 && cc - 1 >= t + 1) ? &V[rr - 1][cc - 1] : &zero;
                const st5* lf = (cc > t) ? &V[rr][cc - 1] : &zero;
                V[rr][cc] = cell_f32(dg, up, lf, qt, dt, -9.0f, NULL);
            }
            for (int s = 0; s < 5; s++) {
                lleft[r][s] = (rr == 31 && !noquirk) ? 0.0f : V[rr][31].v[s];
                ldiag[r][s] = (rr == 0 || (rr == 31 && !noquirk)) ? 0.0f : V[rr - 1][31].v[s];
            }
        }
    }
}

or_dpmax or_viterbi_f32(const cbp_profiles* q, int qp, const cbp_profiles* d, int dp,
                        const float* Sin, char* btck, float* mm)
{
    const int Lq = q->len[qp], Ld = d->len[dp];
    or_dpmax res = {0.0f, 0, 0};
    float (*lleft)[5] = malloc(sizeof(float) * 5 * Lq);
    float (*ldiag)[5] = malloc(sizeof(float) * 5 * Lq);
    or_colm1(q, qp, lleft, ldiag);
    float* S = (float*)malloc(sizeof(float) * (size_t)Lq * Ld);
    memcpy(S, Sin, sizeof(float) * (size_t)Lq * Ld);
    or_mask_corners_f32(Lq, Ld, S);
    st5* prev = (st5*)calloc((size_t)Ld, sizeof(st5));   /* row -1: zeros */
    st5* cur = (st5*)malloc(sizeof(st5) * (size_t)Ld);
    float* dts = (float*)malloc(sizeof(float) * 7 * (size_t)Ld);
    for (int c = 0; c < Ld; c++) trn7(d, dp, c, dts + 7 * c);
    int bx = 0, by = 0; float best = 0.0f;
    for (int r = 0; r < Lq; r++) {
        float qt[7]; trn7(q, qp, r, qt);
        st5 L, D;
        memcpy(L.v, lleft[r], sizeof L.v); memcpy(D.v, ldiag[r], sizeof D.v);
        for (int c = 0; c < Ld; c++) {
            const st5* dg = (c == 0) ? &D : &prev[c - 1];
            const st5* lf = (c == 0) ? &L : &cur[c - 1];
            char b;
            cur[c] = cell_f32(dg, &prev[c], lf, qt, dts + 7 * c, S[(size_t)r * Ld + c], &b);
            float v = cur[c].v[SMM];
            if (btck) btck[(size_t)r * Ld + c] = b;
            if (mm) mm[(size_t)r * Ld + c] = v;
            /* canonical tie rule: larger score, then smaller x, then smaller y */
            if (v > 0.0f && (v > best || (v == best && (c < bx || (c == bx && r < by))))) {
                best = v; bx = c; by = r;
            }
        }
        st5* tmp = prev; prev = cur; cur = tmp;
    }
    res.maxsc = best; res.x = bx; res.y = by;
    free(lleft); free(ldiag); free(S); free(prev); free(cur); free(dts);
    return res;
}

/* FP64 version (values only), same recurrences and boundaries */
typedef struct { double v[5]; } st5d;
static inline double dmx(double a, double b) { return a < b ? b : a; }
double or_viterbi_f64(const cbp_profiles* q, int qp, const cbp_profiles* d, int dp,
                      const double* Sin)
{
    const int Lq = q->len[qp], Ld = d->len[dp];
    float (*lleft)[5] = malloc(sizeof(float) * 5 * Lq);
    float (*ldiag)[5] = malloc(sizeof(float) * 5 * Lq);
    or_colm1(q, qp, lleft, ldiag);
    double* S = (double*)malloc(sizeof(double) * (size_t)Lq * Ld);
    memcpy(S, Sin, sizeof(double) * (size_t)Lq * Ld);
    or_mask_corners_f64(Lq, Ld, S);
    st5d* prev = (st5d*)calloc((size_t)Ld, sizeof(st5d));
    st5d* cur = (st5d*)malloc(sizeof(st5d) * (size_t)Ld);
    double best = 0.0;
    for (int r = 0; r < Lq; r++) {
        float qtf[7]; trn7(q, qp, r, qtf);
        double qt[7]; for (int k = 0; k < 7; k++) qt[k] = qtf[k];
        st5d L, D;
        for (int s = 0; s < 5; s++) { L.v[s] = lleft[r][s]; D.v[s] = ldiag[r][s]; }
        for (int c = 0; c < Ld; c++) {
            float dtf[7]; trn7(d, dp, c, dtf);
            double dt[7]; for (int k = 0; k < 7; k++) dt[k] = dtf[k];
            const st5d* dg = (c == 0) ? &D : &prev[c - 1];
            const st5d* up = &prev[c];
            const st5d* lf = (c == 0) ? &L : &cur[c - 1];
            st5d n; double v1, v2;
            v1 = dg->v[SMI] + qt[0] + dt[3];
            v2 = dg->v[SIM] + qt[3] + dt[0]; v1 = dmx(v1, v2);
            v2 = dg->v[SDN] + qt[5] + dt[0]; v1 = dmx(v1, v2);
            v2 = dg->v[SND] + qt[0] + dt[5]; v1 = dmx(v1, v2);
            v2 = dg->v[SMM] + qt[0] + dt[0]; v1 = dmx(v1, v2);
            n.v[SMM] = dmx(v1 + S[(size_t)r * Ld + c], 0.0);
            n.v[SMI] = dmx(up->v[SMM] + qt[0] + dt[1], up->v[SMI] + qt[0] + dt[4]);
            n.v[SIM] = dmx(lf->v[SMM] + qt[1] + dt[0], lf->v[SIM] + qt[4] + dt[0]);
            n.v[SDN] = dmx(up->v[SMM] + qt[2], up->v[SDN] + qt[6]);
            n.v[SND] = dmx(lf->v[SMM] + dt[2], lf->v[SND] + dt[6]);
            cur[c] = n;
            if (n.v[SMM] > best) best = n.v[SMM];
        }
        st5d* tmp = prev; prev = cur; cur = tmp;
    }
    free(lleft); free(ldiag); free(S); free(prev); free(cur);
    return best;
}

/* ---------------------------------------------------------------------- */
/* Literal emulation of the reference's phase-1 schedule for one profile:
 * the host loop over block diagonals (cudp/CuBatchDP.cu:208-222,260-312),
 * ExecDP_Btck_Unroll32x blocks of 32 threads (cudp/CuBatchDP_init_btck.cu:
 * 65-494) with their diagonal/bottom/max buffers, and FinalizeDPPro's
 * maximum reduction over the last 64 DiagM entries (CuBatchDP_final.cu:
 * 96-147).  Used only to pin or_viterbi_f32 / or_colm1.                   */

or_dpmax or_refemu_viterbi(const cbp_profiles* q, int qp, const cbp_profiles* d, int dp,
                           const float* Sraw, char* btck, float* mm)
{
    const int Lq = q->len[qp], Ld = d->len[dp], W = 32;
    st5* dg1 = (st5*)calloc(Ld, sizeof(st5));
    st5* dg2 = (st5*)calloc(Ld, sizeof(st5));
    st5* bot = (st5*)calloc(Ld, sizeof(st5));
    float* dgM = (float*)calloc(Ld, sizeof(float));
    uint32_t* crd = (uint32_t*)calloc(Ld, sizeof(uint32_t));
    const int nblkdiags = ((Ld + Lq) + W - 1) / W + (Lq - 1) / W;
    const int lastdiag = 2 * ((Lq - 1) / W);
    int maxblk = (Lq + W - 1) / W;
    { int bpw = 2 * W; int m2 = ((Ld + bpw) + bpw - 1) / bpw; if (m2 < maxblk) maxblk = m2; }
    const int edge = corner_edge(Lq, Ld);
    for (int dd = 0; dd < nblkdiags; dd++) {
        for (int s = 0; s < maxblk; s++) {
            int x0, y;
            if (dd <= lastdiag) { x0 = (2 * s - !(dd & 1)) * W; y = ((dd >> 1) + 1 - s) * W - 1; }
            else { x0 = (2 * s + (dd - lastdiag - 1)) * W; y = ((lastdiag >> 1) + 1 - s) * W - 1; }
            /* GetMaqxNoIterations */
            int ilim;
            { int xx = Ld - x0; if (xx <= 0) ilim = 0; else { int yy = y - (Lq - 1); if (0 < yy) xx -= yy;
              ilim = xx <= 0 ? 0 : (W <= xx ? W : xx); } }
            if (y < 0 || Lq <= y + 1 - W || Ld <= x0 || ilim < 1) continue;
            float qt[32][7], dtc[64][7], sc[32][32];
            st5 A[33], B[33], botc[32];
            float maxsc[32]; uint32_t maxcrd[32];
            char bt[32][32];
            for (int t = 0; t < W; t++) {
                int qpos = y - t;
                if (0 <= qpos && qpos < Lq) trn7(q, qp, qpos, qt[t]); else trn7_log0(qt[t]);
                int x = x0 + t;
                if (0 <= x && x < Ld) trn7(d, dp, x, dtc[t]); else trn7_log0(dtc[t]);
                if (0 <= x + W && x + W < Ld) trn7(d, dp, x + W, dtc[t + W]); else trn7_log0(dtc[t + W]);
                for (int i = 0; i < W; i++) {
                    float v = -9.0f;
                    int yy = y - i, xx = x + i;
                    if (!in_corner(Lq, Ld, edge, yy, xx) && 0 <= yy && yy < Lq && 0 <= xx && xx < Ld)
                        v = Sraw[(size_t)yy * Ld + xx];
                    sc[i][t] = v;
                }
            }
            /* diagonal caches: A = diag1Cache (shift 0), B = diag2Cache (shift 1) */
            memset(A, 0, sizeof A); memset(B, 0, sizeof B);
            for (int t = 0; t < W; t++) {
                int x = x0 + t;
                if (0 <= x - 1 && x - 1 < Ld) {
                    A[t] = dg1[x - 1]; B[t + 1] = dg2[x - 1];
                    maxsc[t] = dgM[x - 1]; maxcrd[t] = crd[x - 1];
                } else { maxsc[t] = 0.0f; maxcrd[t] = 0; }
                if (W <= y && 0 <= x + W - 1 && x + W - 1 < Ld) botc[t] = bot[x + W - 1];
                else memset(&botc[t], 0, sizeof(st5));
            }
            st5* p1 = A; st5* p2 = B;
            for (int i = 0; i < ilim; i++) {
                p1[W] = botc[i];
                st5 nw[32];
                for (int t = 0; t < W; t++) {
                    char b;
                    nw[t] = cell_f32(&p2[t + 1], &p1[t + 1], &p1[t], qt[t], dtc[t + i], sc[t][i], &b);
                    bt[t][i] = b;
                    float v = nw[t].v[SMM];
                    if (maxsc[t] < v) { maxsc[t] = v; maxcrd[t] = ((uint32_t)(x0 + t + i) << 16) | (uint32_t)(y - t); }
                    int yy = y - t, xx = x0 + t + i;
                    if (mm && 0 <= yy && yy < Lq && 0 <= xx && xx < Ld) mm[(size_t)yy * Ld + xx] = v;
                }
                for (int t = 0; t < W; t++) p2[t] = nw[t];
                botc[i] = nw[0];
                st5* tmp = p1; p1 = p2; p2 = tmp;
            }
            for (int t = 0; t < W; t++) {
                int x = x0 + t;
                if (0 <= x + W - 1 && x + W - 1 < Ld) { dg1[x + W - 1] = p1[t]; dg2[x + W - 1] = p2[t + 1]; }
                int ndx = W - 1, xtl = x0 + W - 1;
                if (xtl >= Ld) ndx = 0; else if (xtl + ndx >= Ld) ndx = Ld - xtl - 1;
                if (0 <= x + ndx && x + ndx < Ld) {
                    if (W <= y) { if (dgM[x + ndx] < maxsc[t]) { dgM[x + ndx] = maxsc[t]; crd[x + ndx] = maxcrd[t]; } }
                    else { dgM[x + ndx] = maxsc[t]; crd[x + ndx] = maxcrd[t]; }
                }
                if (0 <= x && x < Ld) bot[x] = botc[t];
                for (int i = 0; i < W; i++) {
                    int yy = y - i, xx = x + i;
                    if (btck && 0 <= yy && yy < Lq && 0 <= xx && xx < Ld) btck[(size_t)yy * Ld + xx] = bt[i][t];
                }
            }
        }
    }
    /* FinalizeDPPro reduction */
    float mc[64]; int mi[64];
    for (int t = 0; t < W; t++) {
        mc[t] = (t < Ld) ? dgM[Ld - 1 - t] : 0.0f;
        mc[t + W] = (t + W < Ld) ? dgM[Ld - 1 - t - W] : 0.0f;
    }
    for (int t = 0; t < W; t++) {   /* dpfinmaxndxinit */
        float v2 = mc[t + W];
        if (mc[t] < v2) { mc[t] = v2; mi[t] = t + W; } else mi[t] = t;
    }
    for (int off = W >> 1; off >= 1; off >>= 1) {   /* dpfinmaxndx */
        float v2[32]; int i2[32];
        for (int t = 0; t < W; t++) { v2[t] = mc[t + off]; i2[t] = mi[t + off]; }
        for (int t = 0; t < W; t++) if (mc[t] < v2[t]) { mc[t] = v2[t]; mi[t] = i2[t]; }
    }
    or_dpmax res;
    res.maxsc = mc[0];
    uint32_t xx = crd[Ld - 1 - mi[0]];
    res.x = (int)(xx >> 16); res.y = (int)(xx & 0xffff);
    free(dg1); free(dg2); free(bot); free(dgM); free(crd);
    return res;
}

/* ---------------------------------------------------------------------- */
/* FinalizeDPPro backtrace and correlation bonus (CuBatchDP_final.cu:176-248) */

static float butterfly32(const float c[32]) {   /* __shfl_xor_sync 16,8,4,2,1 */
    float v[32], n[32];
    memcpy(v, c, sizeof v);
    for (int off = 16; off >= 1; off >>= 1) {
        for (int t = 0; t < 32; t++) n[t] = v[t] + v[t ^ off];
        memcpy(v, n, sizeof v);
    }
    return v[0];
}

or_fin or_finalize(int Lq, int Ld, const float* S, const char* btck, or_dpmax m)
{
    (void)Lq;
    or_fin f;
    int x = m.x, y = m.y, psts = 0;
    char b = OR_DIAG;
    float cache[36];
    memset(cache, 0, sizeof cache);
    float sumcorr = 0.0f;
    while (b != OR_STOP) {
        for (int t = 0; t < 4; t++) cache[t] = cache[t + 32];
        int ndx;
        for (ndx = 4; ndx < 36;) {
            if (x < 0 || y < 0) { b = OR_STOP; break; }
            size_t qpos = (size_t)y * Ld + x;
            b = btck[qpos];
            if (b == OR_STOP) break;
            if (b == OR_UP) { y--; continue; }
            else if (b == OR_LEFT) { x--; continue; }
            cache[ndx] = S[qpos];
            if (cache[ndx] > 0.0f) psts++;
            x--; y--; ndx++;
        }
        float corr[32];
        for (int t = 0; t < 32; t++)
            corr[t] = (ndx <= t + 4) ? 0.0f
                    : cache[t + 4] * (cache[t] + cache[t + 1] + cache[t + 2] + cache[t + 3]);
        sumcorr += butterfly32(corr);
    }
    sumcorr *= 0.5f;
    sumcorr = sumcorr < 0.0f ? -(float)sqrt((double)-sumcorr) : (float)sqrt((double)sumcorr);
    f.score = m.maxsc + sumcorr;
    f.psts = psts;
    f.begx = x + 1; f.begy = y + 1;
    f.endx = m.x; f.endy = m.y;
    return f;
}

/* FinalizeDP_ALN (cudp/CuBatchDP_aln_final.cu:206-345): the walk of
 * :246-306 (STOP at the matrix edge or a STOP code; UP y--, LEFT x--, DIAG
 * both), the middle symbol of :299-305, the reversed write of :314-328 and
 * the beginning coordinates (x+1, y+1) of :336 */
or_aln or_aln_final(int Lq, int Ld, const float* S, const char* btck, or_dpmax m,
                    const char* qaa, const char* daa, char* aln)
{
    or_aln r;
    memset(&r, 0, sizeof r);
    int x = (m.maxsc > 0.0f) ? m.x : 0, y = (m.maxsc > 0.0f) ? m.y : 0;
    r.endx = x; r.endy = y;
    const int stride = Lq + Ld + 1;
    char* buf = (char*)malloc(3 * (size_t)stride);
    int n = 0;
    for (;;) {
        if (x < 0 || y < 0) break;
        const size_t qpos = (size_t)y * Ld + x;
        const char b = btck[qpos];
        if (b == OR_STOP) break;
        char qa, ta, mid;
        if (b == OR_UP) { r.gaps++; qa = qaa[y]; ta = '-'; mid = ' '; y--; }
        else if (b == OR_LEFT) { r.gaps++; qa = '-'; ta = daa[x]; mid = ' '; x--; }
        else {
            qa = qaa[y]; ta = daa[x];
            if (qa == ta) { mid = qa; r.idts++; }
            else if (S[qpos] > 0.0f) { mid = '+'; r.psts++; }
            else mid = ' ';
            x--; y--;
        }
        buf[n] = qa; buf[stride + n] = mid; buf[2 * (size_t)stride + n] = ta; n++;
    }
    r.alnlen = n;
    r.begx = x + 1; r.begy = y + 1;
    if (aln) {
        for (int k = 0; k < n; k++) {
            aln[k] = buf[n - 1 - k];
            aln[stride + k] = buf[stride + n - 1 - k];
            aln[2 * (size_t)stride + k] = buf[2 * (size_t)stride + n - 1 - k];
        }
        aln[n] = aln[stride + n] = aln[2 * (size_t)stride + n] = 0;
    }
    free(buf);
    return r;
}

/* JobDispatcher::CalculateScoreThreshold (cualn/JobDispatcher.cpp:809-860) */
void or_score_threshold(const cbp_options* o, uint64_t qlen, uint64_t dblen,
                        float* scorethld, float* logevthld)
{
    uint64_t sspace = qlen * dblen;
    float ev = o->evalue, lmb = o->gap_lambda, K = o->gap_K;
    if (ev <= 0.0f) { *scorethld = 999.0f; *logevthld = -99.0f; return; }
    *logevthld = logf(ev);
    if (ev < 1.0f) ev *= 10.0f; else ev *= 3.0f;
    ev /= K * (float)sspace;
    if (ev <= 0.0f) { *scorethld = 999.0f; return; }
    ev = logf(ev);
    float sc = -ev / lmb;
    if (sc < 0.0f) sc = 0.0f;
    *scorethld = sc;
}
