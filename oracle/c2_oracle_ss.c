/*
 * c2_oracle_ss.c -- FP64 oracle for alignment statistics and E-values.
 * TEST INFRASTRUCTURE ONLY (see c2_oracle.h).
 *
 * The reference computes these in FP32 with CUDA intrinsics; the north-star
 * allows FP64 for this reduction, so the restatement follows the reference's
 * algorithm step by step in double precision.  Two reference behaviours are
 * fixed deliberately:
 *   - histogram counts are exact integers (the reference adds 1.0f with
 *     atomicAdd, exact below 2^24 cells per pair);
 *   - CalcProbOfAlnLengthj reads neighbour cells while other warps write
 *     theirs without a barrier (cuss/CuBatchSS_K.cuh:262-283); this oracle
 *     uses the synchronous reading (all reads see the previous length).
 */
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include "c2_oracle.h"

#define SMIN (-10)
#define NUNS 21
#define SCL 32
#define SSMIN (-320)
#define NSC 641
#define LOG_SP_MIN (-8.7336545e+01)
#define LOG_SP_MAX (8.8722839e+01)
#define LN1K 6.90775527898213681510242167860

void or_histogram(const float* S, int Lq, int Ldb, const char* qaa, const char* daa,
                  int xuninf, int64_t* counts)
{
    memset(counts, 0, sizeof(int64_t) * NSC);
    for (int y = 0; y < Lq; y++) {
        if (xuninf && qaa[y] == 'X') continue;
        for (int x = 0; x < Ldb; x++) {
            if (xuninf && daa[x] == 'X') continue;
            float s = S[(size_t)y * Ldb + x];
            if (s < -10.0f) s = -10.0f; else if (s > 10.0f) s = 10.0f;   /* VerifyScore */
            int i = (int)rintf(s * 32.0f);
            counts[i - SSMIN]++;
        }
    }
}

/* interval of scaled-score bins owned by unscaled score u (CuBatchSS_dynplm.cu:207-222) */
static void interval(int y, int* l, int* r) {
    int u = y + SMIN;
    int left = u * SCL - (SCL >> 1) - SSMIN;
    int rght = u * SCL + (SCL >> 1) - SSMIN;
    if (u <= 0) { left++; if (left < 0) left = 0; }
    if (0 <= u) { rght--; if (NSC - 1 < rght) rght = NSC - 1; }
    *l = left; *r = rght;
}

/* FindLmbdRootByNRandBisection (cuss/CuBatchSS_lambda.cuh:131-226) */
static void lambda_f(double x, double P[NUNS][SCL], const int* L, double mn, double mx,
                     double* f, double* df)
{
    double lf = 0.0, ldf = 0.0;
    for (int y = 0; y < NUNS; y++)
        for (int t = 0; t < SCL; t++) {
            double sc = L[y] + t + SSMIN;
            if (mn <= sc && sc <= mx) {
                sc -= mx;
                double v = x * sc;
                if (LOG_SP_MIN < v) { double e = exp(v) * P[y][t]; lf += e; ldf += e * sc; }
            }
        }
    double z = x * -mx;
    if (LOG_SP_MIN < z) { z = exp(z); lf -= z; ldf += z * mx; }
    *f = lf; if (df) *df = ldf;
}

static double find_lambda(double P[NUNS][SCL], const int* L, double mn, double mx)
{
    const double x1 = 1.0e-5, x2 = 10.0, acc = 1.0e-4;
    double f, df, fl, dx, dxold, xh, xl, rts;
    lambda_f(x1, P, L, mn, mx, &fl, NULL);
    lambda_f(x2, P, L, mn, mx, &f, NULL);
    if ((fl > 0.0 && f > 0.0) || (fl < 0.0 && f < 0.0)) return -1.0;
    if (fl < 0.0) { xl = x1; xh = x2; } else { xh = x1; xl = x2; }
    rts = 0.5 * (x1 + x2);
    dxold = fabs(x2 - x1);
    dx = dxold;
    lambda_f(rts, P, L, mn, mx, &f, &df);
    int n;
    for (n = 0; n < 32; n++) {
        int stop = 0;
        if ((((rts - xh) * df - f) * ((rts - xl) * df - f) > 0.0) || (fabs(2.0 * f) > fabs(dxold * df))) {
            dxold = dx; dx = 0.5 * (xh - xl); rts = xl + dx;
            if (xl == rts) stop = 1;
        } else {
            dxold = dx; dx = f / df;
            double tmp = rts; rts -= dx;
            if (tmp == rts) stop = 1;
        }
        if (fabs(dx) < acc) stop = 1;
        if (stop) break;
        lambda_f(rts, P, L, mn, mx, &f, &df);
        if (f < 0.0) xl = rts; else xh = rts;
    }
    if (n == 32) return -1.0;
    return rts;
}

/* CalcProbOfAlnLengthj on the 21x32 thread layout (cuss/CuBatchSS_K.cuh:240-321),
 * synchronous reads; C holds the previous length's values and is overwritten */
static double prob_len_j(double C[NUNS][SCL], const double* pu, const int* L,
                         int minu, int maxu, int range, int minsc, int maxsc, double lambda)
{
    double nw[NUNS][SCL], pj = 0.0;
    for (int y = 0; y < NUNS; y++)
        for (int t = 0; t < SCL; t++) {
            double pji = 0.0;
            int sc = L[y] + t + SSMIN;
            if (minsc <= sc && sc <= maxsc && -32 <= sc && sc <= 32) {
                double etols = (sc < 0) ? lambda * (double)sc : 0.0;
                if (LOG_SP_MIN < etols) {
                    int lo = (maxsc - sc < range) ? maxu - (maxsc - sc) : minu;
                    int hi = (sc - minsc < range) ? minu + (sc - minsc) : maxu;
                    for (int i = lo; i <= hi; i++) {
                        int yy = y, xx = t - i;
                        if (SCL <= xx) { xx -= SCL; if (yy + 1 < NUNS) yy++; else continue; }
                        else if (xx < 0) { xx += SCL; if (0 < yy) yy--; else continue; }
                        pji += C[yy][xx] * pu[i - SMIN];
                    }
                    pj += pji * exp(etols);
                }
            }
            nw[y][t] = pji;
        }
    memcpy(C, nw, sizeof nw);
    return pj;
}

static void nn_e(const double in[4], const float* w, double* out) {
    double a[4];
    for (int h = 0; h < 4; h++) {
        double z = (double)w[h * 5] + in[0] * w[h * 5 + 1] + in[1] * w[h * 5 + 2] +
                   in[2] * w[h * 5 + 3] + in[3] * w[h * 5 + 4];
        a[h] = (2.0 / (1.0 + exp(-z)) - 1.0) * w[20 + h + 1];
    }
    double o = (double)w[20] + a[0] + a[1] + a[2] + a[3];
    *out = (2.0 / (1.0 + exp(-o)) - 1.0) * 30.0;
}

static double clampd(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }

or_stats or_statistics(const int64_t* counts, int Lq, int Ldb, float qeno, float deno,
                       float score, float psts, float searchspace,
                       const cbp_options* o, const float* nnw)
{
    or_stats r;
    memset(&r, 0, sizeof r);
    int L[NUNS], Rr[NUNS];
    double P[NUNS][SCL], pu[NUNS], tot = 0.0;
    for (int y = 0; y < NUNS; y++) {
        interval(y, &L[y], &Rr[y]);
        double s = 0.0;
        for (int t = 0; t < SCL; t++) {
            P[y][t] = (t + L[y] <= Rr[y]) ? (double)counts[L[y] + t] : 0.0;
            s += P[y][t];
        }
        pu[y] = s;
        tot += s;
    }
    for (int y = 0; y < NUNS; y++) {
        pu[y] /= tot;
        for (int t = 0; t < SCL; t++) P[y][t] /= tot;
    }
    /* expected score and min/max of scaled scores (CuBatchSS_expected.cuh, _minmax.cuh) */
    double E = 0.0, mn = 0.0, mx = 0.0;
    for (int y = 0; y < NUNS; y++)
        for (int t = 0; t < SCL; t++) {
            double sc = L[y] + t + SSMIN;
            E += sc * P[y][t];
            double v = P[y][t] != 0.0 ? sc : 0.0;
            if (v < mn) mn = v;
            if (v > mx) mx = v;
        }
    /* lambda (CalcLambdaProfile) */
    double lmb = (0.0 <= E || mx <= 0.0) ? -1.0 : find_lambda(P, L, mn, mx);
    /* relative entropy (CalcEntropyProfile, scaled version) */
    double H = -1.0;
    if (lmb > 0.0) {
        H = 0.0;
        for (int y = 0; y < NUNS; y++)
            for (int t = 0; t < SCL; t++) {
                double sc = L[y] + t + SSMIN;
                if (mn <= sc && sc <= mx) {
                    double v = lmb * sc;
                    if (LOG_SP_MIN < v && v < LOG_SP_MAX - 3.0) H += exp(v) * v * P[y][t];
                }
            }
    }
    /* rescale; P row of unscaled score 0 becomes the length-0 distribution */
    for (int t = 0; t < SCL; t++) P[0 - SMIN][t] = 1.0;
    if (lmb > 0.0) lmb *= SCL;
    E /= SCL;
    int minu = 0, maxu = 0;
    for (int i = 0; i < NUNS; i++) {
        int v = pu[i] != 0.0 ? i + SMIN : 0;
        if (v < minu) minu = v;
        if (v > maxu) maxu = v;
    }
    /* K (CalcKProfile, cuss/CuBatchSS_K.cuh:109-180) */
    double K;
    if (0.0 <= E || lmb <= 0.0 || H <= 0.0) K = -1.0;
    else if (minu == -1 && maxu == 1) {
        double a = pu[-1 - SMIN], b = pu[1 - SMIN];
        K = (a - b) * (a - b) / a;
    } else if (minu == -1 || maxu == 1) {
        if (maxu != 1) K = lmb / H * (1.0 - exp(-lmb)) * E * E;
        else K = H / lmb * (1.0 - exp(-lmb));
    } else {
        double sigma = 0.0, Pj = 1.0;
        int minsc = minu, maxsc = maxu, range = maxu - minu + 1;
        for (int j = 1; j <= 32 && 1.0e-4 < Pj; j++) {
            Pj = prob_len_j(P, pu, L, minu, maxu, range, minsc, maxsc, lmb);
            sigma += Pj / j;
            minsc += minu; maxsc += maxu;
        }
        K = lmb * exp(-2.0 * sigma) / (H * (1.0 - exp(-lmb)));
    }
    r.Min = minu; r.Max = maxu; r.E = E; r.H = H; r.Lmbd = lmb; r.K = K;

    /* significance (CalcSignificanceSS18DynPlmProfile / Obs) */
    const double lref = o->ung_lambda, kref = o->ung_K, lg = o->gap_lambda, kg = o->gap_K;
    const double sc = score, ss = searchspace;
    if (lmb > 0.0) r.LmbdEst = lmb * lg / lref;
    if (K > 0.0) r.KEst = K * kg / kref;
    const double l1l2 = (double)(float)((uint32_t)Lq * (uint32_t)Ldb);
    if (lmb <= 0.0 || K <= 0.0) r.EpA = log(kref * l1l2) - lref * sc;
    else r.EpA = log(K * l1l2) - lmb * sc;
    if (o->ssemodel == 0) {
        r.Evalue = log(kg * ss) - lg * sc;
        return r;
    }
    /* PredictEVDParametersSS18 (cuss/CuBatchSS_SS18.cuh:118-237) */
    double in[4], eMU, lMU, lSC;
    {
        double qe = qeno, de = deno; uint32_t lq = Lq, ld = Ldb;
        if (qe < de) { double t = qe; qe = de; de = t; uint32_t u = lq; lq = ld; ld = u; }
        else if (qe == de && lq < ld) { uint32_t u = lq; lq = ld; ld = u; }
        in[0] = 0.05 * qe; in[1] = log((double)lq) / LN1K;
        in[2] = 0.05 * de; in[3] = log((double)ld) / LN1K;
        nn_e(in, nnw, &eMU);
    }
    {
        double lm = rint(lmb * 10.0) / 10.0;
        if (lm <= 0.0) lm = 0.1;
        if (10.0 < lm) lm = 10.0;
        uint32_t lq = Lq, ld = Ldb;
        if (lq < ld) { uint32_t u = lq; lq = ld; ld = u; }
        in[0] = 0.2 * lm; in[1] = log((double)lq) / LN1K; in[2] = log((double)ld) / LN1K; in[3] = 0.0;
        const float* w2 = nnw + 25;
        double outs[2];
        for (int net = 0; net < 2; net++) {
            const float* w = w2 + net * 25;
            double a[4];
            for (int h = 0; h < 4; h++) {
                double z = (double)w[h * 5] + in[0] * w[h * 5 + 1] + in[1] * w[h * 5 + 2] +
                           in[2] * w[h * 5 + 3] + in[3] * w[h * 5 + 4];
                a[h] = (2.0 / (1.0 + exp(-z)) - 1.0) * w[20 + h + 1];
            }
            double oo = (double)w[20] + a[0] + a[1] + a[2] + a[3];
            outs[net] = (2.0 / (1.0 + exp(-oo)) - 1.0) * (net == 0 ? 30.0 : 5.0);
        }
        lMU = outs[0]; lSC = outs[1];
    }
    const float* adj = nnw + (o->ssemodel == 2 ? 84 : 75);
    if (eMU < 3.0) eMU = 3.0;
    if (lMU < 3.0) lMU = 3.0;
    if (lSC < 0.1) lSC = 0.1;
    lSC = adj[6] * (exp(lSC) - 1.0);
    lSC = clampd(lSC, 0.1, 5.0);
    double scaleE = clampd(adj[4] * eMU + adj[5], 0.1, 5.0);
    eMU = adj[0] * eMU + adj[1];
    lMU = adj[2] * lMU + adj[3];
    double alpha = adj[7], ibeta = adj[8];
    if (0.0 < lmb) {
        double rc = 1.0 / lmb, m10 = rc < 10.0 ? rc : 10.0;
        alpha -= alpha * 0.1 * m10;
        ibeta += (1.0 - ibeta) * 0.1 * m10;
    }
    double mu = alpha * eMU + (1.0 - alpha) * lMU;
    double scale = (1.0 - ibeta) * scaleE + ibeta * lSC;
    /* CalculateEvalueSS18 (cuss/CuBatchSS_SS18.cuh:258-355) */
    if (mu <= 0.0 || scale <= 0.0) { r.Evalue = log(kg * ss) - lg * sc; return r; }
    double logKmn = mu / scale;
    if (K > kref && 0.0 < K && K < 1.0 && 0.0 < kref && kref < 1.0) logKmn += log(K / kref);
    double logexp = logKmn - sc / scale;
    if (o->ssemodel == 2) logexp += -9.210340372;
    double logp = logexp;
    if (-4.605170186 <= logexp) {
        if (LOG_SP_MAX - 3.0 < logexp) logp = 0.0;
        else {
            double p = exp(logexp);
            logp = (LOG_SP_MAX - 3.0 < p) ? 0.0 : log(1.0 - exp(-p));
        }
    }
    double l12 = l1l2 < 625.0 ? 625.0 : l1l2;
    double pn = rint(100000.0 * (double)psts / pow(l12, 1.5));
    double lpp = or_lbetainc(pn + 1.0, 0.1, 0.757213);
    logp = or_lgammainc_Q(1.346206, (-logp - lpp) / 1.485656);
    logexp = logp;
    if (-4.605170186 <= logp) {
        double p = exp(logp);
        logexp = 77.0;
        if (p < 1.0) logexp = log(-log(1.0 - p));
    }
    logexp += log(ss / l1l2);
    r.Evalue = logexp;
    return r;
}

/* ---- Cephes-derived log incomplete beta / gamma, FP64 ---- */
#define MACHEP 1.11022302462515654042e-16
#define BIG 4.503599627370496e15
#define BIGINV 2.22044604925031308085e-16

static double incbcf(double a, double b, double x) {
    double xk, pk, pkm1, pkm2, qk, qkm1, qkm2, k1, k2, k3, k4, k5, k6, k7, k8, r, t, ans;
    k1 = a; k2 = a + b; k3 = a; k4 = a + 1.0; k5 = 1.0; k6 = b - 1.0; k7 = k4; k8 = a + 2.0;
    pkm2 = 0.0; qkm2 = 1.0; pkm1 = 1.0; qkm1 = 1.0; ans = 1.0; r = 1.0;
    for (int n = 0; n < 100; n++) {   /* iteration cap of cusf_lbetainccff/dff */
        xk = -(x * k1 * k2) / (k3 * k4);
        pk = pkm1 + pkm2 * xk; qk = qkm1 + qkm2 * xk;
        pkm2 = pkm1; pkm1 = pk; qkm2 = qkm1; qkm1 = qk;
        xk = (x * k5 * k6) / (k7 * k8);
        pk = pkm1 + pkm2 * xk; qk = qkm1 + qkm2 * xk;
        pkm2 = pkm1; pkm1 = pk; qkm2 = qkm1; qkm1 = qk;
        if (qk != 0.0) r = pk / qk;
        if (r != 0.0) { t = fabs((ans - r) / r); ans = r; } else t = 1.0;
        if (t < 3.0 * MACHEP) break;
        k1 += 1.0; k2 += 1.0; k3 += 2.0; k4 += 2.0; k5 += 1.0; k6 -= 1.0; k7 += 2.0; k8 += 2.0;
        if (fabs(qk) + fabs(pk) > BIG) { pkm2 *= BIGINV; pkm1 *= BIGINV; qkm2 *= BIGINV; qkm1 *= BIGINV; }
        if (fabs(qk) < BIGINV || fabs(pk) < BIGINV) { pkm2 *= BIG; pkm1 *= BIG; qkm2 *= BIG; qkm1 *= BIG; }
    }
    return ans;
}
static double incbd(double a, double b, double x) {
    double xk, pk, pkm1, pkm2, qk, qkm1, qkm2, k1, k2, k3, k4, k5, k6, k7, k8, r, t, ans, z;
    k1 = a; k2 = b - 1.0; k3 = a; k4 = a + 1.0; k5 = 1.0; k6 = a + b; k7 = a + 1.0; k8 = a + 2.0;
    pkm2 = 0.0; qkm2 = 1.0; pkm1 = 1.0; qkm1 = 1.0; z = x / (1.0 - x); ans = 1.0; r = 1.0;
    for (int n = 0; n < 100; n++) {   /* iteration cap of cusf_lbetainccff/dff */
        xk = -(z * k1 * k2) / (k3 * k4);
        pk = pkm1 + pkm2 * xk; qk = qkm1 + qkm2 * xk;
        pkm2 = pkm1; pkm1 = pk; qkm2 = qkm1; qkm1 = qk;
        xk = (z * k5 * k6) / (k7 * k8);
        pk = pkm1 + pkm2 * xk; qk = qkm1 + qkm2 * xk;
        pkm2 = pkm1; pkm1 = pk; qkm2 = qkm1; qkm1 = qk;
        if (qk != 0.0) r = pk / qk;
        if (r != 0.0) { t = fabs((ans - r) / r); ans = r; } else t = 1.0;
        if (t < 3.0 * MACHEP) break;
        k1 += 1.0; k2 -= 1.0; k3 += 2.0; k4 += 2.0; k5 += 1.0; k6 += 1.0; k7 += 2.0; k8 += 2.0;
        if (fabs(qk) + fabs(pk) > BIG) { pkm2 *= BIGINV; pkm1 *= BIGINV; qkm2 *= BIGINV; qkm1 *= BIGINV; }
        if (fabs(qk) < BIGINV || fabs(pk) < BIGINV) { pkm2 *= BIG; pkm1 *= BIG; qkm2 *= BIG; qkm1 *= BIG; }
    }
    return ans;
}
static double lpseries(double a, double b, double x) {
    double y = a * log(x) + (b - 1.0) * log(1.0 - x) - log(a);
    y -= lgamma(a) + lgamma(b);
    y += lgamma(a + b);
    double t = x / (1.0 - x), s = 0.0, u = 1.0;
    do {
        b -= 1.0;
        if (b == 0.0) break;
        a += 1.0;
        u *= t * b / a;
        s += u;
    } while (fabs(u) > MACHEP);
    return y + log(1.0 + s);
}
static double lbeta_helper(double a, double b, double x) {
    double ans, t = 1.0 - x;
    if (b > 10.0 && fabs(b * x / a) < 0.3) return lpseries(a, b, x);
    ans = x * (a + b - 2.0) / (a - 1.0);
    if (ans < 1.0) { ans = incbcf(a, b, x); t = b * log(t); }
    else { ans = incbd(a, b, x); t = (b - 1.0) * log(t); }
    t += a * log(x) + lgamma(a + b) - lgamma(a) - lgamma(b);
    t += log(ans / a);
    return t;
}
/* cusf_lbetaincf (cusf/CuSF_betainc.cuh:96-126) */
double or_lbetainc(double aa, double bb, double xx) {
    if (xx <= 0.0 || xx >= 1.0) {
        if (xx == 0.0) return -999.0;
        if (xx == 1.0) return 0.0;
        return 999.0;
    }
    if (aa <= 0.0) return 999.0;
    if (aa <= 1.0) {
        double ans = lbeta_helper(aa + 1.0, bb, xx);
        ans = ans > LOG_SP_MIN ? exp(ans) : 0.0;
        double t = aa * log(xx) + bb * log(1.0 - xx) + lgamma(aa + bb) - lgamma(aa + 1.0) - lgamma(bb);
        if (t > LOG_SP_MIN) ans += exp(t);
        return ans ? log(ans) : -999.0;
    }
    return lbeta_helper(aa, bb, xx);
}
/* cusf_lgammainc_Qf / _Pf (cusf/CuSF_gammainc.cuh:102-200) */
double or_lgammainc_P(double a, double x);
double or_lgammainc_Q(double a, double x) {
    if (x <= 0.0 || a <= 0.0) return 0.0;
    if (x < 1.0 || x < a) {
        double ans = or_lgammainc_P(a, x);
        ans = ans > LOG_SP_MIN ? exp(ans) : 0.0;
        return ans < 1.0 ? log(1.0 - ans) : -999.0;
    }
    double ax = a * log(x) - x - lgamma(a);
    double y = 1.0 - a, z = x + y + 1.0, c = 0.0;
    double pkm2 = 1.0, qkm2 = x, pkm1 = x + 1.0, qkm1 = z * x;
    double ans = pkm1 / qkm1, t, r, pk, qk, yc;
    do {
        c += 1.0; y += 1.0; z += 2.0; yc = y * c;
        pk = pkm1 * z - pkm2 * yc; qk = qkm1 * z - qkm2 * yc;
        if (qk != 0.0) { r = pk / qk; t = fabs((ans - r) / r); ans = r; } else t = 1.0;
        pkm2 = pkm1; pkm1 = pk; qkm2 = qkm1; qkm1 = qk;
        if (fabs(pk) > BIG) { pkm2 *= BIGINV; pkm1 *= BIGINV; qkm2 *= BIGINV; qkm1 *= BIGINV; }
    } while (t > MACHEP);
    return log(ans) + ax;
}
double or_lgammainc_P(double a, double x) {
    if (x <= 0.0 || a <= 0.0) return -999.0;
    if (x > 1.0 && x > a) {
        double ans = or_lgammainc_Q(a, x);
        ans = ans > LOG_SP_MIN ? exp(ans) : 0.0;
        return ans < 1.0 ? log(1.0 - ans) : -999.0;
    }
    double ax = a * log(x) - x - lgamma(a);
    double r = a, c = 1.0, ans = 1.0;
    do { r += 1.0; c *= x / r; ans += c; } while (c / ans > MACHEP);
    return log(ans) + ax - log(a);
}

/* AbstractScoreMatrix::ComputeLengthAdjustment, float arithmetic as written */
uint64_t or_length_adjustment(const cbp_options* o, uint64_t m, uint64_t n, uint64_t N,
                              uint64_t* delta)
{
    const float lambda = o->gap_lambda, K = o->gap_K, alpha = o->gap_alpha, beta = o->gap_beta;
    uint64_t ss = m * n;
    *delta = 0;
    if (lambda <= 0.0f || K <= 0.0f || N < 1) return ss;
    const int maxit = 20;                                  /* LENGTH_ADJUSTMENT_MAXIT */
    const float logK = logf(K), alratio = alpha / lambda;
    const float max_mn = (float)((m > n) ? m : n);
    float length = 0.0f, min_length = 0.0f, max_length = 0.0f, nxt_length = 0.0f;
    float sspace = 0.0f, space2 = 0.0f;
    int converged = 0;
    float a = (float)N, b = (float)(n + m * N), c = (float)(m * n) - max_mn / K;
    if (c < 0.0f) return ss;
    max_length = 2.0f * c / (b + sqrtf(b * b - 4.0f * a * c));
    for (int j = 0; j < maxit; j++) {
        length = nxt_length;
        sspace = ((float)m - length) * ((float)n - (float)N * length);
        nxt_length = alratio * (logK + logf(sspace)) + beta;
        if (length <= nxt_length) {
            min_length = length;
            if (nxt_length - length <= 1.0f) { converged = 1; break; }
            if (min_length == max_length) break;
        } else max_length = length;
        if (nxt_length < min_length || max_length < nxt_length)
            nxt_length = (!j) ? max_length : (min_length + max_length) * 0.5f;
    }
    if (converged) {
        length = ceilf(min_length);
        if (length <= max_length) {
            space2 = ((float)m - length) * ((float)n - (float)N * length);
            nxt_length = alratio * (logK + logf(space2)) + beta;
            if (length <= nxt_length) { min_length = length; sspace = space2; }
        }
    }
    *delta = (uint64_t)min_length;
    return (uint64_t)sspace;
}
