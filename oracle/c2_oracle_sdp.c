/*
 * c2_oracle_sdp.c -- FP64 oracle of the differentiable soft DP.
 * TEST INFRASTRUCTURE ONLY (see c2_oracle.h).
 *
 * The soft DP has no reference counterpart.  Its model is the reference's
 * MAP forward model (ExecMAPDP_Fwd_Unroll32x, cumapdp/CuBatchMAPDP_fwd.cu:291-372):
 * five states MM MI IM DN ND per cell, MM = log(1 + sum of the five
 * diagonal predecessor terms, each carrying the cell score), gap states
 * log-sum-exp of an open and an extend term, row -1 = (MM 0, others -99),
 * column -1 = the MAP pass's boundary vectors, logZ = log sum_ij exp(MM(i,j)).
 * Here exp/log are exact (no -16 / 3.8e-11 cut-offs), and five global
 * log-offsets theta are added to the transition classes:
 *   theta[0] insert open   (MM->MI, MM->IM)
 *   theta[1] insert extend (MI->MI, IM->IM)
 *   theta[2] delete open   (MM->DN, MM->ND)
 *   theta[3] delete extend (DN->DN, ND->ND)
 *   theta[4] gap close     (MI, IM, DN, ND -> MM)
 * The gradients come from the backward (outside) recursion:
 *   dlogZ/dS(i,j)     = exp(B_MM(i,j) - logZ) * (exp(F_MM(i,j)) - 1)
 *   dlogZ/dtheta[k]   = sum over class-k transitions of their posterior counts
 *   posterior MM(i,j) = exp(F_MM(i,j) + B_MM(i,j) - logZ).
 */
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include "c2_oracle.h"

enum { TMM = 0, TMI, TMD, TIM, TII, TDM, TDD };
enum { SMM = 0, SMI, SIM, SDN, SND };
#define NEG (-1.0e300)

static double trn(const cbp_profiles* P, int p, int k, int i) {
    return (double)P->trans[(size_t)k * (size_t)(P->npos + P->nprof) + P->dist[p] + p + i];
}
/* forward caching of the MAP pass: MI and II at i+1, the rest at i */
static void trnf(const cbp_profiles* P, int p, int i, double t[7]) {
    t[0] = trn(P, p, TMM, i); t[1] = trn(P, p, TMI, i + 1); t[2] = trn(P, p, TMD, i);
    t[3] = trn(P, p, TIM, i); t[4] = trn(P, p, TII, i + 1); t[5] = trn(P, p, TDM, i);
    t[6] = trn(P, p, TDD, i);
}
static double lse(const double* a, int n) {
    double m = NEG;
    for (int k = 0; k < n; k++) if (a[k] > m) m = a[k];
    if (m <= NEG) return NEG;
    double s = 0.0;
    for (int k = 0; k < n; k++) s += exp(a[k] - m);
    return m + log(s);
}

typedef struct { double v[5]; } sd5;

/* the 14 forward terms of cell (i,j): MM gets a[0..5] (a[5] = start, 0),
 * MI b[0..1], IM c[0..1], DN d[0..1], ND e[0..1] */
static void terms(const sd5* dg, const sd5* up, const sd5* lf, const double* qt, const double* dt,
                  double s, const double* th, double a[6], double b[2], double c[2], double d[2], double e[2])
{
    a[0] = dg->v[SMI] + qt[0] + dt[3] + th[4] + s;
    a[1] = dg->v[SIM] + qt[3] + dt[0] + th[4] + s;
    a[2] = dg->v[SDN] + qt[5] + dt[0] + th[4] + s;
    a[3] = dg->v[SND] + qt[0] + dt[5] + th[4] + s;
    a[4] = dg->v[SMM] + qt[0] + dt[0] + s;
    a[5] = 0.0;
    b[0] = up->v[SMM] + qt[0] + dt[1] + th[0]; b[1] = up->v[SMI] + qt[0] + dt[4] + th[1];
    c[0] = lf->v[SMM] + qt[1] + dt[0] + th[0]; c[1] = lf->v[SIM] + qt[4] + dt[0] + th[1];
    d[0] = up->v[SMM] + qt[2] + th[2];         d[1] = up->v[SDN] + qt[6] + th[3];
    e[0] = lf->v[SMM] + dt[2] + th[2];         e[1] = lf->v[SND] + dt[6] + th[3];
}

/* forward over all cells; F [Lq*Ld] of sd5; returns logZ */
static double forward(const cbp_profiles* q, int qp, const cbp_profiles* d, int dp,
                      const double* S, const double* th, const float* ll, const float* ld, sd5* F)
{
    const int Lq = q->len[qp], Ld = d->len[dp];
    sd5 init = {{0.0, -99.0, -99.0, -99.0, -99.0}};
    double zm = NEG, zs = 0.0;
    for (int r = 0; r < Lq; r++) {
        double qt[7]; trnf(q, qp, r, qt);
        sd5 lft, dgn;
        for (int s = 0; s < 5; s++) { lft.v[s] = ll[r * 5 + s]; dgn.v[s] = ld[r * 5 + s]; }
        for (int c = 0; c < Ld; c++) {
            double dt[7]; trnf(d, dp, c, dt);
            const sd5* up = (r > 0) ? &F[(size_t)(r - 1) * Ld + c] : &init;
            const sd5* dg = (c == 0) ? &dgn : ((r > 0) ? &F[(size_t)(r - 1) * Ld + c - 1] : &init);
            const sd5* lf = (c == 0) ? &lft : &F[(size_t)r * Ld + c - 1];
            double a[6], b[2], cc[2], dd[2], e[2];
            terms(dg, up, lf, qt, dt, S[(size_t)r * Ld + c], th, a, b, cc, dd, e);
            sd5 n;
            n.v[SMM] = lse(a, 6); n.v[SMI] = lse(b, 2); n.v[SIM] = lse(cc, 2);
            n.v[SDN] = lse(dd, 2); n.v[SND] = lse(e, 2);
            F[(size_t)r * Ld + c] = n;
            const double f = n.v[SMM];
            if (f > zm) { zs = zs * exp(zm - f) + 1.0; zm = f; } else zs += exp(f - zm);
        }
    }
    return zm + log(zs);
}

double or_sdp_logz(const cbp_profiles* q, int qp, const cbp_profiles* d, int dp,
                   const double* S, const double* theta, const float* ll, const float* ld)
{
    const size_t N = (size_t)q->len[qp] * d->len[dp];
    sd5* F = (sd5*)malloc(sizeof(sd5) * N);
    const double z = forward(q, qp, d, dp, S, theta, ll, ld, F);
    free(F);
    return z;
}

double or_sdp_grad(const cbp_profiles* q, int qp, const cbp_profiles* d, int dp,
                   const double* S, const double* theta, const float* ll, const float* ld,
                   double* dS, double* dtheta, double* post)
{
    const int Lq = q->len[qp], Ld = d->len[dp];
    const size_t N = (size_t)Lq * Ld;
    sd5* F = (sd5*)malloc(sizeof(sd5) * N);
    sd5* B = (sd5*)malloc(sizeof(sd5) * N);
    const double Z = forward(q, qp, d, dp, S, theta, ll, ld, F);
    sd5 init = {{0.0, -99.0, -99.0, -99.0, -99.0}};
    for (int k = 0; k < 5; k++) dtheta[k] = 0.0;
    /* outside recursion: B_X(i,j) = log sum over the forward terms that read
     * X(i,j) of exp(term - X(i,j) + B_dest); the end term of MM is 0 */
    for (int r = Lq - 1; r >= 0; r--) {
        for (int c = Ld - 1; c >= 0; c--) {
            double acc[5][6]; int na[5] = {0, 0, 0, 0, 0};
            acc[SMM][na[SMM]++] = 0.0;
            const sd5* Fc = &F[(size_t)r * Ld + c];
            /* successor (r+1, c+1): reads this cell as its diagonal */
            if (r + 1 < Lq && c + 1 < Ld) {
                double qt[7], dt[7]; trnf(q, qp, r + 1, qt); trnf(d, dp, c + 1, dt);
                const sd5* Bd = &B[(size_t)(r + 1) * Ld + c + 1];
                const double s = S[(size_t)(r + 1) * Ld + c + 1];
                acc[SMI][na[SMI]++] = qt[0] + dt[3] + theta[4] + s + Bd->v[SMM];
                acc[SIM][na[SIM]++] = qt[3] + dt[0] + theta[4] + s + Bd->v[SMM];
                acc[SDN][na[SDN]++] = qt[5] + dt[0] + theta[4] + s + Bd->v[SMM];
                acc[SND][na[SND]++] = qt[0] + dt[5] + theta[4] + s + Bd->v[SMM];
                acc[SMM][na[SMM]++] = qt[0] + dt[0] + s + Bd->v[SMM];
            }
            /* successor (r+1, c): reads this cell as up */
            if (r + 1 < Lq) {
                double qt[7], dt[7]; trnf(q, qp, r + 1, qt); trnf(d, dp, c, dt);
                const sd5* Bn = &B[(size_t)(r + 1) * Ld + c];
                acc[SMM][na[SMM]++] = qt[0] + dt[1] + theta[0] + Bn->v[SMI];
                acc[SMI][na[SMI]++] = qt[0] + dt[4] + theta[1] + Bn->v[SMI];
                acc[SMM][na[SMM]++] = qt[2] + theta[2] + Bn->v[SDN];
                acc[SDN][na[SDN]++] = qt[6] + theta[3] + Bn->v[SDN];
            }
            /* successor (r, c+1): reads this cell as left */
            if (c + 1 < Ld) {
                double qt[7], dt[7]; trnf(q, qp, r, qt); trnf(d, dp, c + 1, dt);
                const sd5* Br = &B[(size_t)r * Ld + c + 1];
                acc[SMM][na[SMM]++] = qt[1] + dt[0] + theta[0] + Br->v[SIM];
                acc[SIM][na[SIM]++] = qt[4] + dt[0] + theta[1] + Br->v[SIM];
                acc[SMM][na[SMM]++] = dt[2] + theta[2] + Br->v[SND];
                acc[SND][na[SND]++] = dt[6] + theta[3] + Br->v[SND];
            }
            sd5 b;
            for (int s = 0; s < 5; s++) b.v[s] = lse(acc[s], na[s]);
            B[(size_t)r * Ld + c] = b;
            const double fm = Fc->v[SMM];
            dS[(size_t)r * Ld + c] = exp(b.v[SMM] - Z) * expm1(fm);
            if (post) post[(size_t)r * Ld + c] = exp(fm + b.v[SMM] - Z);
            (void)init;
        }
    }
    /* class counts from the forward terms of every cell: exp(term + B_dest - Z) */
    for (int r = 0; r < Lq; r++) {
        double qt[7]; trnf(q, qp, r, qt);
        sd5 lft, dgn;
        for (int s = 0; s < 5; s++) { lft.v[s] = ll[r * 5 + s]; dgn.v[s] = ld[r * 5 + s]; }
        for (int c = 0; c < Ld; c++) {
            double dt[7]; trnf(d, dp, c, dt);
            const sd5* up = (r > 0) ? &F[(size_t)(r - 1) * Ld + c] : &init;
            const sd5* dg = (c == 0) ? &dgn : ((r > 0) ? &F[(size_t)(r - 1) * Ld + c - 1] : &init);
            const sd5* lf = (c == 0) ? &lft : &F[(size_t)r * Ld + c - 1];
            double a[6], b[2], cc[2], dd[2], e[2];
            terms(dg, up, lf, qt, dt, S[(size_t)r * Ld + c], theta, a, b, cc, dd, e);
            const sd5* Bc = &B[(size_t)r * Ld + c];
            for (int k = 0; k < 4; k++) dtheta[4] += exp(a[k] + Bc->v[SMM] - Z);
            dtheta[0] += exp(b[0] + Bc->v[SMI] - Z) + exp(cc[0] + Bc->v[SIM] - Z);
            dtheta[1] += exp(b[1] + Bc->v[SMI] - Z) + exp(cc[1] + Bc->v[SIM] - Z);
            dtheta[2] += exp(dd[0] + Bc->v[SDN] - Z) + exp(e[0] + Bc->v[SND] - Z);
            dtheta[3] += exp(dd[1] + Bc->v[SDN] - Z) + exp(e[1] + Bc->v[SND] - Z);
        }
    }
    free(F); free(B);
    return Z;
}
