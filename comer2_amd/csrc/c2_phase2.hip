/*
 * c2_phase2.hip -- phase-1 finalize (backtrace + correlation bonus) and the
 * alignment statistics for gfx950.
 *
 * k_finalize replaces FinalizeDP/FinalizeDPPro (cudp/CuBatchDP_final.cu:61-396):
 * one wave per candidate pair walks the packed backtrack codes written by
 * the BTCK phase-1 launch from the maximum cell, scores the DIAG cells of
 * each 32-cell window together and accumulates the correlation sum in the
 * reference's order (32-score windows reduced by an xor butterfly, windows
 * added in sequence), then adds sign*sqrt(|0.5 sum|).  The passed list is compacted on the host in profile
 * order (the reference appends under a spin lock in block-scheduling order).
 *
 * k_hist replaces InitScoreProbsDynPlmProfile + CalcScoreProbsXuninfDynPlmProfile
 * (cuss/CuBatchSS_dynplm.cu:61-182): one workgroup per passed pair recomputes
 * the pair's score matrix and bins rint(32*clamp(S,-10,10)) into an LDS
 * integer histogram (exact counts instead of float atomics).
 *
 * k_stats replaces CalcStatParamsDynPlmProfile and the SS18/Obs significance
 * kernels (cuss/CuBatchSS_dynplm.cu:193-674, cuss/CuBatchSS_{lambda,entropy,K,
 * SS18,NN}.cuh, cusf/*): one wave per pair, FP64 throughout; the score
 * distribution lives in LDS, the sums over its 672 entries are spread over
 * the lanes and combined by a fixed xor butterfly, so the Newton/bisection
 * decisions are wave-uniform and the result does not depend on scheduling.
 */
#include "c2_draw.h"

namespace {

__device__ __forceinline__ float score_at(const ScoreCtx& c, int y, int64_t col, const float* w)
{
    float ss;
    return c2_score_cell_w(c, y, col, w, &ss);
}

__device__ void finalize_one(const ScoreCtx& c, const int32_t* cand, const unsigned long long* keys,
                             const uint32_t* btck, const int64_t* btck_off, FinRec* out, uint32_t* win, int i)
{
    const int lane = threadIdx.x;
    const int p = cand[i];
    const unsigned long long key = keys[p];
    const float maxsc = __uint_as_float((uint32_t)(key >> 32));
    const uint32_t lo = 0xFFFFFFFFu - (uint32_t)(key & 0xFFFFFFFFull);
    const int xe = (int)(lo >> 16), ye = (int)(lo & 0xFFFF);
    const int64_t col0 = c.dist[p];
    P1Codes codes;
    codes.bt = btck + btck_off[i]; codes.nsteps = c.len[p] + 63; codes.R = c.R; codes.win = win;
    int x = xe, y = ye, psts = 0;
    float pw[20];
    c2_pair_weights(c, p, pw);
    float carry[4] = {0.0f, 0.0f, 0.0f, 0.0f};   /* the previous window's last four scores */
    float sumcorr = 0.0f;
    if (key != 0ull) {
        int b = 3;
        while (b != 0) {
            int nd = 0, myy = 0, myx = 0;
            while (nd < 32) {
                if (x < 0 || y < 0) { b = 0; break; }
                b = codes.at(y, x, lane);
                if (b == 0) break;
                if (b == 2) { y--; continue; }
                if (b == 1) { x--; continue; }
                if (nd == lane) { myy = y; myx = x; }
                x--; y--; nd++;
            }
            const float sv = (lane < nd) ? score_at(c, myy, col0 + myx, pw) : 0.0f;
            psts += __popcll(__ballot(lane < nd && sv > 0.0f));
            /* the four scores before this lane's: carried ones for the first lanes */
            float prev[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const float up = __shfl_up(sv, 4 - k);
                const int src = lane - (4 - k);
                const int j = lane + k;               /* carried slot when src < 0 */
                const float cj = (j == 0) ? carry[0] : (j == 1) ? carry[1] : (j == 2) ? carry[2] : carry[3];
                prev[k] = (src < 0) ? cj : up;
            }
            float v = (lane < 32 && lane < nd) ? sv * (((prev[0] + prev[1]) + prev[2]) + prev[3]) : 0.0f;
            for (int off = 16; off >= 1; off >>= 1) v = v + __shfl_xor(v, off, 32);
            sumcorr += __shfl(v, 0);
#pragma unroll
            for (int k = 0; k < 4; k++) carry[k] = __shfl(sv, 28 + k);
        }
    }
    sumcorr *= 0.5f;
    sumcorr = sumcorr < 0.0f ? -__fsqrt_rn(-sumcorr) : __fsqrt_rn(sumcorr);
    if (lane == 0) {
        FinRec r;
        r.score = maxsc + sumcorr;
        r.maxsc = maxsc;
        r.psts = psts;
        r.beg = ((uint32_t)(x + 1) << 16) | (uint32_t)(y + 1);
        r.end = ((uint32_t)xe << 16) | (uint32_t)ye;
        r.pad = 0;
        out[i] = r;
    }
}

/* one wave per candidate: every lane follows the backtrace through the LDS
 * window of the codes (P1Codes, c2_draw.h); lane j keeps the j-th DIAG cell
 * of the current 32-cell window, the lanes score the window's cells
 * together, and the window's correlation terms
 *   v[t] = s[t+4] * (((s[t] + s[t+1]) + s[t+2]) + s[t+3])   (s[0..3] carried over)
 * are summed by the same xor butterfly the reference uses
 * (CuBatchDP_final.cu:200-292), one term per lane; windows are added in
 * path order, so the sum has the one-thread loop's bits */
__global__ __launch_bounds__(64) void k_finalize(ScoreCtx c, const int32_t* cand, int32_t ncand,
                                                 const unsigned long long* keys, const uint32_t* btck,
                                                 const int64_t* btck_off, FinRec* out, const int32_t* ncand_dev)
{
    __shared__ __align__(16) uint32_t win[C2_DRAW_WIN / 4];
    const int n = ncand_dev ? __builtin_amdgcn_readfirstlane(*ncand_dev) : ncand;
    for (int i = blockIdx.x; i < n; i += gridDim.x)
        finalize_one(c, cand, keys, btck, btck_off, out, win, i);
}
/* FinalizeDP_ALN (cudp/CuBatchDP_aln_final.cu:61-438): alignments of the
 * MAPALN=0 path drawn from the phase-1 backtrace, one wave per pair;
 * end coordinates are the phase-1 maximum's */
__global__ __launch_bounds__(64) void k_aln_final(ScoreCtx c, AlnArgs a, int n)
{
    __shared__ __align__(16) uint32_t win[C2_DRAW_WIN / 4];
    const int nn = a.dcnt ? __builtin_amdgcn_readfirstlane(a.dcnt[DL_NSEL]) : n;
    for (int i = blockIdx.x; i < nn; i += gridDim.x) {
    const int p = a.pro[i];
    const unsigned long long key = a.keys[p];
    const uint32_t lo = 0xFFFFFFFFu - (uint32_t)(key & 0xFFFFFFFFull);
    const int xe = key ? (int)(lo >> 16) : 0, ye = key ? (int)(lo & 0xFFFF) : 0;
    P1Codes codes;
    codes.bt = a.btck + a.btck_off[a.cand[i]]; codes.nsteps = c.len[p] + 63; codes.R = c.R; codes.win = win;
    const int stride = (int)a.alnstride;
    int32_t* o = a.res + (size_t)i * 8;
    c2_walk_draw(c, p, xe, ye, codes, a.alns + (size_t)i * 3 * stride, a.alntmp + (size_t)i * 3 * stride,
                 stride, o);
    if (threadIdx.x == 0) o[5] = (int32_t)(((uint32_t)xe << 16) | (uint32_t)ye);
    }
}

constexpr int NSC = 641;

/* query rows staged per tile: 20 target frequencies + p_H + 0.1 as 21
 * words per row (broadcast reads), and the 30-entry SS table per row */
constexpr int HTY = 128;
constexpr int HQW = 21;

template <bool CV>
__global__ __launch_bounds__(256) void k_hist(ScoreCtx c, const int32_t* pro, int xuninf, int32_t* counts,
                                              int32_t npro, const int32_t* npro_dev)
{
    constexpr int NWH = 4;                  /* waves of the 256-thread workgroup, one histogram each */
    __shared__ int h[NWH][NSC];
    __shared__ float sq[HTY * HQW];
    __shared__ float sqt[30 * HTY];
    __shared__ unsigned char sx[HTY];
    const int npairs = npro_dev ? __builtin_amdgcn_readfirstlane(*npro_dev) : npro;
    for (int pi = blockIdx.x; pi < npairs; pi += gridDim.x) {
    __syncthreads();
    for (int i = threadIdx.x; i < NWH * NSC; i += blockDim.x) (&h[0][0])[i] = 0;
    int* hw = h[threadIdx.x >> 6];
    const int p = pro[pi];
    const int Ldb = c.len[p];
    const int64_t col0 = c.dist[p];
    float w[20];
    c2_pair_weights(c, p, w);
    /* a profile shorter than the workgroup is covered several times over:
     * thread (phase ph, column xc) takes the rows r = ph (mod nph) */
    const int Lw = Ldb < (int)blockDim.x ? Ldb : (int)blockDim.x;
    const int nph = (int)blockDim.x / Lw;
    const int ph = (int)threadIdx.x / Lw, xc = (int)threadIdx.x - ph * Lw;
    for (int x0 = 0; x0 < Ldb; x0 += Lw) {
        const int x = x0 + xc;
        const int64_t col = col0 + (x < Ldb ? x : 0);
        const bool xok = ph < nph && x < Ldb && !(xuninf && c.aa[col] == 'X');
        float dw[20], pj[3];
        int co[3];
#pragma unroll
        for (int k = 0; k < 20; k++) dw[k] = c.trgfrq[(size_t)k * c.npos + col] * w[k];
#pragma unroll
        for (int j = 0; j < 3; j++) {
            pj[j] = c.ssprb[(size_t)j * c.npos + col];
            co[j] = 10 * j + (int)(pj[j] * 9.999f);
        }
        const float dh = (pj[2] + 0.1f) * 0.1111111f;
        /* the column's context vector, |d|^2 + 1 + 1/kappa and prior as used */
        float dcv[CV ? 19 : 1], dca = 0.0f, dcr = 0.0f;
        if constexpr (CV) {
#pragma unroll
            for (int i = 0; i < 19; i++) dcv[i] = c.cvs.cv[(size_t)i * c.npos + col];
            dca = c2_cvs_norm(c.cvs, c.cvs.cvn[col]);
            dcr = c2_cvs_prior(c.cvs, c.cvs.cvp[col]);
        }
        for (int y0 = 0; y0 < c.Lq; y0 += HTY) {
            const int ny = min(HTY, c.Lq - y0);
            __syncthreads();
            for (int i = threadIdx.x; i < ny * HQW; i += blockDim.x) {
                const int r = i / HQW, k = i - r * HQW;
                sq[i] = (k < 20) ? c.qv[(size_t)k * c.Lpad + y0 + r] : c.qh[y0 + r];
            }
            for (int i = threadIdx.x; i < 30 * ny; i += blockDim.x) {
                const int j = i / ny, r = i - j * ny;
                sqt[j * HTY + r] = c.QT[(size_t)j * c.Lpad + y0 + r];
            }
            for (int i = threadIdx.x; i < ny; i += blockDim.x)
                sx[i] = (xuninf && c.qaa[y0 + i] == 'X') ? 1 : 0;
            __syncthreads();
            if (!xok) continue;
            for (int r = ph; r < ny; r += nph) {
                if (sx[r]) continue;
                const float* qr = sq + r * HQW;
                float acc = qr[0] * dw[0];
#pragma unroll
                for (int k = 1; k < 20; k++) acc = fmaf(qr[k], dw[k], acc);
                float sc = c2_logf_clamped(acc) - c.shift, ss = 0.0f;
                if (c.ssswgt > 0.0f) {
                    ss = sqt[co[0] * HTY + r] * pj[0];
                    ss = fmaf(sqt[co[1] * HTY + r], pj[1], ss);
                    ss = fmaf(sqt[co[2] * HTY + r], pj[2], ss);
                    ss = fmaf(qr[20], dh, ss);
                    ss = (ss - 0.05f) * c.ssswgt;
                    if (!CV) sc = sc + ss;
                }
                if constexpr (CV) {
                    /* the row's values are the same for the whole wave */
                    const int yy = y0 + r;
                    float qc[19];
#pragma unroll
                    for (int i = 0; i < 19; i++) qc[i] = c.cvs.qcv[(size_t)i * c.Lpad + yy];
                    const float qca = c2_cvs_norm(c.cvs, c.cvs.qcv[(size_t)20 * c.Lpad + yy]);
                    const float qcr = c2_cvs_prior(c.cvs, c.cvs.qcv[(size_t)19 * c.Lpad + yy]);
                    const float cm = c2_cvs_lookup(c.cvs, c2_cvs_x(c.cvs, qc, qca, qcr, dcv, dca, dcr));
                    c2_cvs_combine(c.cvs.wgt, c.ssswgt, cm, sc, ss);
                }
                if (sc < -10.0f) sc = -10.0f; else if (sc > 10.0f) sc = 10.0f;
                const int bin = (int)rintf(sc * 32.0f) + 320;
                atomicAdd(&hw[bin], 1);
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < NSC; i += blockDim.x)
        counts[(size_t)pi * NSC + i] = h[0][i] + h[1][i] + h[2][i] + h[3][i];
    }
}

/* ---------------- FP64 statistics (one wave per pair) ---------------- */

constexpr int SMIN = -10, NUNS = 21, SCL = 32, SSMIN = -320;
constexpr double LOG_SP_MIN = -8.7336545e+01, LOG_SP_MAX = 8.8722839e+01;
constexpr double LN1K = 6.90775527898213681510242167860;
constexpr double MACHEP = 1.11022302462515654042e-16;
constexpr double BIG = 4.503599627370496e15, BIGINV = 2.22044604925031308085e-16;

__device__ void interval(int y, int* l, int* r) {
    const int u = y + SMIN;
    int left = u * SCL - (SCL >> 1) - SSMIN, rght = u * SCL + (SCL >> 1) - SSMIN;
    if (u <= 0) { left++; if (left < 0) left = 0; }
    if (0 <= u) { rght--; if (NSC - 1 < rght) rght = NSC - 1; }
    *l = left; *r = rght;
}

__device__ double incbcf(double a, double b, double x) {
    double xk, pk, pkm1, pkm2, qk, qkm1, qkm2, k1, k2, k3, k4, k5, k6, k7, k8, r, t, ans;
    k1 = a; k2 = a + b; k3 = a; k4 = a + 1.0; k5 = 1.0; k6 = b - 1.0; k7 = k4; k8 = a + 2.0;
    pkm2 = 0.0; qkm2 = 1.0; pkm1 = 1.0; qkm1 = 1.0; ans = 1.0; r = 1.0;
    for (int n = 0; n < 100; n++) {
        xk = -(x * k1 * k2) / (k3 * k4);
        pk = pkm1 + pkm2 * xk; qk = qkm1 + qkm2 * xk; pkm2 = pkm1; pkm1 = pk; qkm2 = qkm1; qkm1 = qk;
        xk = (x * k5 * k6) / (k7 * k8);
        pk = pkm1 + pkm2 * xk; qk = qkm1 + qkm2 * xk; pkm2 = pkm1; pkm1 = pk; qkm2 = qkm1; qkm1 = qk;
        if (qk != 0.0) r = pk / qk;
        if (r != 0.0) { t = fabs((ans - r) / r); ans = r; } else t = 1.0;
        if (t < 3.0 * MACHEP) break;
        k1 += 1.0; k2 += 1.0; k3 += 2.0; k4 += 2.0; k5 += 1.0; k6 -= 1.0; k7 += 2.0; k8 += 2.0;
        if (fabs(qk) + fabs(pk) > BIG) { pkm2 *= BIGINV; pkm1 *= BIGINV; qkm2 *= BIGINV; qkm1 *= BIGINV; }
        if (fabs(qk) < BIGINV || fabs(pk) < BIGINV) { pkm2 *= BIG; pkm1 *= BIG; qkm2 *= BIG; qkm1 *= BIG; }
    }
    return ans;
}
__device__ double incbd(double a, double b, double x) {
    double xk, pk, pkm1, pkm2, qk, qkm1, qkm2, k1, k2, k3, k4, k5, k6, k7, k8, r, t, ans, z;
    k1 = a; k2 = b - 1.0; k3 = a; k4 = a + 1.0; k5 = 1.0; k6 = a + b; k7 = a + 1.0; k8 = a + 2.0;
    pkm2 = 0.0; qkm2 = 1.0; pkm1 = 1.0; qkm1 = 1.0; z = x / (1.0 - x); ans = 1.0; r = 1.0;
    for (int n = 0; n < 100; n++) {
        xk = -(z * k1 * k2) / (k3 * k4);
        pk = pkm1 + pkm2 * xk; qk = qkm1 + qkm2 * xk; pkm2 = pkm1; pkm1 = pk; qkm2 = qkm1; qkm1 = qk;
        xk = (z * k5 * k6) / (k7 * k8);
        pk = pkm1 + pkm2 * xk; qk = qkm1 + qkm2 * xk; pkm2 = pkm1; pkm1 = pk; qkm2 = qkm1; qkm1 = qk;
        if (qk != 0.0) r = pk / qk;
        if (r != 0.0) { t = fabs((ans - r) / r); ans = r; } else t = 1.0;
        if (t < 3.0 * MACHEP) break;
        k1 += 1.0; k2 -= 1.0; k3 += 2.0; k4 += 2.0; k5 += 1.0; k6 += 1.0; k7 += 2.0; k8 += 2.0;
        if (fabs(qk) + fabs(pk) > BIG) { pkm2 *= BIGINV; pkm1 *= BIGINV; qkm2 *= BIGINV; qkm1 *= BIGINV; }
        if (fabs(qk) < BIGINV || fabs(pk) < BIGINV) { pkm2 *= BIG; pkm1 *= BIG; qkm2 *= BIG; qkm1 *= BIG; }
    }
    return ans;
}
__device__ double lpseries(double a, double b, double x) {
    double y = a * log(x) + (b - 1.0) * log(1.0 - x) - log(a);
    y -= lgamma(a) + lgamma(b);
    y += lgamma(a + b);
    double t = x / (1.0 - x), s = 0.0, u = 1.0;
    do { b -= 1.0; if (b == 0.0) break; a += 1.0; u *= t * b / a; s += u; } while (fabs(u) > MACHEP);
    return y + log(1.0 + s);
}
__device__ double lbeta_helper(double a, double b, double x) {
    double ans, t = 1.0 - x;
    if (b > 10.0 && fabs(b * x / a) < 0.3) return lpseries(a, b, x);
    ans = x * (a + b - 2.0) / (a - 1.0);
    if (ans < 1.0) { ans = incbcf(a, b, x); t = b * log(t); }
    else { ans = incbd(a, b, x); t = (b - 1.0) * log(t); }
    t += a * log(x) + lgamma(a + b) - lgamma(a) - lgamma(b);
    t += log(ans / a);
    return t;
}
__device__ double lbetainc(double aa, double bb, double xx) {
    if (xx <= 0.0 || xx >= 1.0) { if (xx == 0.0) return -999.0; if (xx == 1.0) return 0.0; return 999.0; }
    if (aa <= 0.0) return 999.0;
    if (aa <= 1.0) {
        double ans = lbeta_helper(aa + 1.0, bb, xx);
        ans = ans > LOG_SP_MIN ? exp(ans) : 0.0;
        const double t = aa * log(xx) + bb * log(1.0 - xx) + lgamma(aa + bb) - lgamma(aa + 1.0) - lgamma(bb);
        if (t > LOG_SP_MIN) ans += exp(t);
        return ans != 0.0 ? log(ans) : -999.0;
    }
    return lbeta_helper(aa, bb, xx);
}
__device__ double lgammainc_P(double a, double x);
__device__ double lgammainc_Q(double a, double x) {
    if (x <= 0.0 || a <= 0.0) return 0.0;
    if (x < 1.0 || x < a) {
        double ans = lgammainc_P(a, x);
        ans = ans > LOG_SP_MIN ? exp(ans) : 0.0;
        return ans < 1.0 ? log(1.0 - ans) : -999.0;
    }
    const double ax = a * log(x) - x - lgamma(a);
    double y = 1.0 - a, z = x + y + 1.0, c = 0.0, pkm2 = 1.0, qkm2 = x, pkm1 = x + 1.0, qkm1 = z * x;
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
__device__ double lgammainc_P(double a, double x) {
    if (x <= 0.0 || a <= 0.0) return -999.0;
    if (x > 1.0 && x > a) {
        double ans = lgammainc_Q(a, x);
        ans = ans > LOG_SP_MIN ? exp(ans) : 0.0;
        return ans < 1.0 ? log(1.0 - ans) : -999.0;
    }
    const double ax = a * log(x) - x - lgamma(a);
    double r = a, c = 1.0, ans = 1.0;
    do { r += 1.0; c *= x / r; ans += c; } while (c / ans > MACHEP);
    return log(ans) + ax - log(a);
}

__device__ double clampd(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }

__device__ void nnet(const double in[4], const float* w, double factor, double* out) {
    double a[4];
    for (int h = 0; h < 4; h++) {
        const double z = (double)w[h * 5] + in[0] * w[h * 5 + 1] + in[1] * w[h * 5 + 2] +
                         in[2] * w[h * 5 + 3] + in[3] * w[h * 5 + 4];
        a[h] = (2.0 / (1.0 + exp(-z)) - 1.0) * w[20 + h + 1];
    }
    const double o = (double)w[20] + a[0] + a[1] + a[2] + a[3];
    *out = (2.0 / (1.0 + exp(-o)) - 1.0) * factor;
}

/* wave-wide sums / extrema with a fixed xor-butterfly order (every lane gets the result) */
__device__ __forceinline__ double wsum(double v)
{
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
    return v;
}
__device__ __forceinline__ double wmin(double v)
{
    for (int off = 32; off >= 1; off >>= 1) { const double o = __shfl_xor(v, off); v = o < v ? o : v; }
    return v;
}
__device__ __forceinline__ double wmax(double v)
{
    for (int off = 32; off >= 1; off >>= 1) { const double o = __shfl_xor(v, off); v = o > v ? o : v; }
    return v;
}

/* sum_s p(s) e^{x s} - 1 on the scaled scores (GetFdFlmbdequation,
 * cuss/CuBatchSS_lambda.cuh:79-170), entries spread over the lanes */
__device__ void lambda_fw(double x, const double* P, const int* L, double mn, double mx, int lane,
                          double* f, double* df)
{
    double lf = 0.0, ldf = 0.0;
    for (int j = lane; j < NUNS * SCL; j += 64) {
        const int y = j / SCL, t = j - y * SCL;
        double sc = L[y] + t + SSMIN;
        if (mn <= sc && sc <= mx) {
            sc -= mx;
            const double v = x * sc;
            if (LOG_SP_MIN < v) { const double e = exp(v) * P[j]; lf += e; ldf += e * sc; }
        }
    }
    lf = wsum(lf); ldf = wsum(ldf);
    double z = x * -mx;
    if (LOG_SP_MIN < z) { z = exp(z); lf -= z; ldf += z * mx; }
    *f = lf; if (df) *df = ldf;
}

/* root of the lambda equation by Newton-Raphson safeguarded with bisection
 * (FindLmbdRootByNRandBisection, cuss/CuBatchSS_lambda.cuh:199-398) */
__device__ double find_lambda_w(const double* P, const int* L, double mn, double mx, int lane)
{
    const double x1 = 1.0e-5, x2 = 10.0, acc = 1.0e-4;
    double f, df, fl, dx, dxold, xh, xl, rts;
    lambda_fw(x1, P, L, mn, mx, lane, &fl, &df);
    lambda_fw(x2, P, L, mn, mx, lane, &f, &df);
    if ((fl > 0.0 && f > 0.0) || (fl < 0.0 && f < 0.0)) return -1.0;
    if (fl < 0.0) { xl = x1; xh = x2; } else { xh = x1; xl = x2; }
    rts = 0.5 * (x1 + x2); dxold = fabs(x2 - x1); dx = dxold;
    lambda_fw(rts, P, L, mn, mx, lane, &f, &df);
    int it;
    for (it = 0; it < 32; it++) {
        bool stop = false;
        if ((((rts - xh) * df - f) * ((rts - xl) * df - f) > 0.0) || (fabs(2.0 * f) > fabs(dxold * df))) {
            dxold = dx; dx = 0.5 * (xh - xl); rts = xl + dx;
            if (xl == rts) stop = true;
        } else {
            dxold = dx; dx = f / df;
            const double tmp = rts; rts -= dx;
            if (tmp == rts) stop = true;
        }
        if (fabs(dx) < acc) stop = true;
        if (stop) break;
        lambda_fw(rts, P, L, mn, mx, lane, &f, &df);
        if (f < 0.0) xl = rts; else xh = rts;
    }
    return it == 32 ? -1.0 : rts;
}

/* probability of alignment length j by convolution (CalcProbOfAlnLengthj,
 * cuss/CuBatchSS_K.cuh:214-321); C (= P) advanced in place through NW */
__device__ double prob_len_jw(double* C, double* NW, const double* pu, const int* L, int minu, int maxu,
                              int range, int minsc, int maxsc, double lambda, int lane)
{
    double pj = 0.0;
    for (int j = lane; j < NUNS * SCL; j += 64) {
        const int y = j / SCL, t = j - y * SCL;
        double pji = 0.0;
        const int sc = L[y] + t + SSMIN;
        if (minsc <= sc && sc <= maxsc && -32 <= sc && sc <= 32) {
            const double etols = (sc < 0) ? lambda * (double)sc : 0.0;
            if (LOG_SP_MIN < etols) {
                const int lo = (maxsc - sc < range) ? maxu - (maxsc - sc) : minu;
                const int hi = (sc - minsc < range) ? minu + (sc - minsc) : maxu;
                for (int i = lo; i <= hi; i++) {
                    int yy = y, xx = t - i;
                    if (SCL <= xx) { xx -= SCL; if (yy + 1 < NUNS) yy++; else continue; }
                    else if (xx < 0) { xx += SCL; if (0 < yy) yy--; else continue; }
                    pji += C[yy * SCL + xx] * pu[i - SMIN];
                }
                pj += pji * exp(etols);
            }
        }
        NW[j] = pji;
    }
    pj = wsum(pj);
    wave_lds_sync();
    for (int j = lane; j < NUNS * SCL; j += 64) C[j] = NW[j];
    wave_lds_sync();
    return pj;
}

/* Karlin-Altschul statistics and E-values of one passed pair per wave
 * (CalcStatParamsDynPlmProfile + CalcSignificanceSS18/ObsDynPlmProfile,
 * cuss/CuBatchSS_dynplm.cu:193-674), FP64 */
__device__ void stats_one(const StatParams& sp, const float* nnw, const int32_t* counts, const int32_t* crow,
                          const StatIn* in, double* out, int i, double* P, double* NW, double* pu, int* L, int* Rr)
{
    const int lane = threadIdx.x & 63;
    const int32_t* cnt = counts + (size_t)(crow ? crow[i] : i) * NSC;
    if (lane < NUNS) {
        int l, r;
        interval(lane, &l, &r);
        L[lane] = l; Rr[lane] = r;
        double s = 0.0;
        for (int t = 0; t < SCL; t++) {
            P[lane * SCL + t] = (t + l <= r) ? (double)cnt[l + t] : 0.0;
            s += P[lane * SCL + t];
        }
        pu[lane] = s;
    }
    wave_lds_sync();
    double tot = 0.0;
    for (int y = 0; y < NUNS; y++) tot += pu[y];
    wave_lds_sync();
    for (int j = lane; j < NUNS * SCL; j += 64) P[j] /= tot;
    if (lane < NUNS) pu[lane] /= tot;
    wave_lds_sync();
    double E = 0.0, mn = 0.0, mx = 0.0;
    for (int j = lane; j < NUNS * SCL; j += 64) {
        const int y = j / SCL, t = j - y * SCL;
        const double sc = L[y] + t + SSMIN;
        E += sc * P[j];
        const double v = P[j] != 0.0 ? sc : 0.0;
        if (v < mn) mn = v;
        if (v > mx) mx = v;
    }
    E = wsum(E); mn = wmin(mn); mx = wmax(mx);
    double lmb = (0.0 <= E || mx <= 0.0) ? -1.0 : find_lambda_w(P, L, mn, mx, lane);
    double H = -1.0;
    if (lmb > 0.0) {
        H = 0.0;
        for (int j = lane; j < NUNS * SCL; j += 64) {
            const int y = j / SCL, t = j - y * SCL;
            const double sc = L[y] + t + SSMIN;
            if (mn <= sc && sc <= mx) {
                const double v = lmb * sc;
                if (LOG_SP_MIN < v && v < LOG_SP_MAX - 3.0) H += exp(v) * v * P[j];
            }
        }
        H = wsum(H);
    }
    wave_lds_sync();
    if (lane < SCL) P[(0 - SMIN) * SCL + lane] = 1.0;
    wave_lds_sync();
    if (lmb > 0.0) lmb *= SCL;
    E /= SCL;
    int minu = 0, maxu = 0;
    for (int k = 0; k < NUNS; k++) {
        const int v = pu[k] != 0.0 ? k + SMIN : 0;
        if (v < minu) minu = v;
        if (v > maxu) maxu = v;
    }
    double K;
    if (0.0 <= E || lmb <= 0.0 || H <= 0.0) K = -1.0;
    else if (minu == -1 && maxu == 1) { const double a = pu[-1 - SMIN], b = pu[1 - SMIN]; K = (a - b) * (a - b) / a; }
    else if (minu == -1 || maxu == 1) {
        if (maxu != 1) K = lmb / H * (1.0 - exp(-lmb)) * E * E;
        else K = H / lmb * (1.0 - exp(-lmb));
    } else {
        double sigma = 0.0, Pj = 1.0;
        int minsc = minu, maxsc = maxu;
        const int range = maxu - minu + 1;
        for (int j = 1; j <= 32 && 1.0e-4 < Pj; j++) {
            Pj = prob_len_jw(P, NW, pu, L, minu, maxu, range, minsc, maxsc, lmb, lane);
            sigma += Pj / j;
            minsc += minu; maxsc += maxu;
        }
        K = lmb * exp(-2.0 * sigma) / (H * (1.0 - exp(-lmb)));
    }
    if (lane != 0) return;
    double* o = out + (size_t)i * 10;
    o[0] = minu; o[1] = maxu; o[2] = E; o[3] = H; o[4] = lmb; o[5] = K;
    const double lref = sp.ung_lambda, kref = sp.ung_K, lg = sp.gap_lambda, kg = sp.gap_K;
    const double sc = in[i].score, ss = sp.searchspace;
    const int Ldb = in[i].Ldb;
    o[6] = lmb > 0.0 ? lmb * lg / lref : 0.0;
    o[7] = K > 0.0 ? K * kg / kref : 0.0;
    const double l1l2 = (double)(float)((uint32_t)sp.Lq * (uint32_t)Ldb);
    o[9] = (lmb <= 0.0 || K <= 0.0) ? log(kref * l1l2) - lref * sc : log(K * l1l2) - lmb * sc;
    if (sp.ssemodel == 0) { o[8] = log(kg * ss) - lg * sc; return; }
    double inp[4], eMU, lMU, lSC;
    {
        double qe = sp.qeno, de = in[i].deno;
        uint32_t lq = sp.Lq, ld = Ldb;
        if (qe < de) { const double t = qe; qe = de; de = t; const uint32_t u = lq; lq = ld; ld = u; }
        else if (qe == de && lq < ld) { const uint32_t u = lq; lq = ld; ld = u; }
        inp[0] = 0.05 * qe; inp[1] = log((double)lq) / LN1K; inp[2] = 0.05 * de; inp[3] = log((double)ld) / LN1K;
        nnet(inp, nnw, 30.0, &eMU);
    }
    {
        double lm = rint(lmb * 10.0) / 10.0;
        if (lm <= 0.0) lm = 0.1;
        if (10.0 < lm) lm = 10.0;
        uint32_t lq = sp.Lq, ld = Ldb;
        if (lq < ld) { const uint32_t u = lq; lq = ld; ld = u; }
        inp[0] = 0.2 * lm; inp[1] = log((double)lq) / LN1K; inp[2] = log((double)ld) / LN1K; inp[3] = 0.0;
        nnet(inp, nnw + 25, 30.0, &lMU);
        nnet(inp, nnw + 50, 5.0, &lSC);
    }
    const float* adj = nnw + (sp.ssemodel == 2 ? 84 : 75);
    if (eMU < 3.0) eMU = 3.0;
    if (lMU < 3.0) lMU = 3.0;
    if (lSC < 0.1) lSC = 0.1;
    lSC = clampd(adj[6] * (exp(lSC) - 1.0), 0.1, 5.0);
    const double scaleE = clampd(adj[4] * eMU + adj[5], 0.1, 5.0);
    eMU = adj[0] * eMU + adj[1];
    lMU = adj[2] * lMU + adj[3];
    double alpha = adj[7], ibeta = adj[8];
    if (0.0 < lmb) {
        const double rc = 1.0 / lmb, m10 = rc < 10.0 ? rc : 10.0;
        alpha -= alpha * 0.1 * m10;
        ibeta += (1.0 - ibeta) * 0.1 * m10;
    }
    const double mu = alpha * eMU + (1.0 - alpha) * lMU;
    const double scale = (1.0 - ibeta) * scaleE + ibeta * lSC;
    if (mu <= 0.0 || scale <= 0.0) { o[8] = log(kg * ss) - lg * sc; return; }
    double logKmn = mu / scale;
    if (K > kref && 0.0 < K && K < 1.0 && 0.0 < kref && kref < 1.0) logKmn += log(K / kref);
    double logexp = logKmn - sc / scale;
    if (sp.ssemodel == 2) logexp += -9.210340372;
    double logp = logexp;
    if (-4.605170186 <= logexp) {
        if (LOG_SP_MAX - 3.0 < logexp) logp = 0.0;
        else {
            const double pp = exp(logexp);
            logp = (LOG_SP_MAX - 3.0 < pp) ? 0.0 : log(1.0 - exp(-pp));
        }
    }
    const double l12 = l1l2 < 625.0 ? 625.0 : l1l2;
    const double pn = rint(100000.0 * (double)in[i].psts / pow(l12, 1.5));
    const double lpp = lbetainc(pn + 1.0, 0.1, 0.757213);
    logp = lgammainc_Q(1.346206, (-logp - lpp) / 1.485656);
    logexp = logp;
    if (-4.605170186 <= logp) {
        const double pp = exp(logp);
        logexp = 77.0;
        if (pp < 1.0) logexp = log(-log(1.0 - pp));
    }
    o[8] = logexp + log(ss / l1l2);
}

__global__ __launch_bounds__(256) void k_stats(StatParams sp, const float* nnw, const int32_t* counts,
                                               const int32_t* crow, const StatIn* in, int32_t n, double* out,
                                               const int32_t* n_dev)
{
    __shared__ double sP[4][NUNS * SCL], sNW[4][NUNS * SCL], spu[4][NUNS];
    __shared__ int sL[4][NUNS], sR[4][NUNS];
    const int wv = threadIdx.x >> 6;
    const int nn = n_dev ? __builtin_amdgcn_readfirstlane(*n_dev) : n;
    const int i = blockIdx.x * 4 + wv;      /* the grid covers the capacity */
    if (i < nn) stats_one(sp, nnw, counts, crow, in, out, i, sP[wv], sNW[wv], spu[wv], sL[wv], sR[wv]);
}

/* S and its SS component for one pair, row-major [Lq][Ldb] */
__global__ __launch_bounds__(256) void k_score_block(ScoreCtx c, int p, float* S, float* mod)
{
    const int Ld = c.len[p];
    const int64_t n = (int64_t)c.Lq * Ld;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int y = (int)(i / Ld), x = (int)(i - (int64_t)y * Ld);
        float ss;
        S[i] = c2_score_cell(c, y, (int64_t)c.dist[p] + x, p, &ss);
        if (mod) mod[i] = ss;
    }
}

}  // namespace

/* a device count: the host count is a capacity; the grid loops */
static int grid_of(int n, const void* dev, int cap) { return dev && n > cap ? cap : n; }

hipError_t c2_launch_finalize(const ScoreCtx& c, const int32_t* cand, int32_t ncand,
                              const unsigned long long* keys, const uint32_t* btck,
                              const int64_t* btck_off, FinRec* out, hipStream_t stream,
                              const int32_t* ncand_dev)
{
    (void)hipGetLastError();   /* the status read after the launches is theirs alone */
    if (ncand <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_finalize, dim3(grid_of(ncand, ncand_dev, 8192)), dim3(64), 0, stream, c, cand, ncand,
                       keys, btck, btck_off, out, ncand_dev);
    return hipGetLastError();
}

hipError_t c2_launch_aln_final(const ScoreCtx& c, const AlnArgs& a, int n, hipStream_t stream)
{
    (void)hipGetLastError();   /* the status read after the launches is theirs alone */
    if (n <= 0) return hipSuccess;
    if (!a.pro || !a.cand || !a.keys || !a.btck || !a.btck_off || !a.alns || !a.alntmp || !a.res)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_aln_final, dim3(grid_of(n, a.dcnt, 8192)), dim3(64), 0, stream, c, a, n);
    return hipGetLastError();
}

hipError_t c2_launch_hist(const ScoreCtx& c, const int32_t* pro, int32_t npro, int xuninf,
                          int32_t* counts, hipStream_t stream, const int32_t* npro_dev)
{
    (void)hipGetLastError();   /* the status read after the launches is theirs alone */
    if (npro <= 0) return hipSuccess;
    if (c.cvs.wgt > 0.0f)
        hipLaunchKernelGGL(k_hist<true>, dim3(grid_of(npro, npro_dev, 4096)), dim3(256), 0, stream, c, pro, xuninf,
                           counts, npro, npro_dev);
    else
        hipLaunchKernelGGL(k_hist<false>, dim3(grid_of(npro, npro_dev, 4096)), dim3(256), 0, stream, c, pro, xuninf,
                           counts, npro, npro_dev);
    return hipGetLastError();
}

hipError_t c2_launch_stats(const StatParams& sp, const float* nnw, const int32_t* counts, const int32_t* crow,
                           const StatIn* in, int32_t n, double* out, hipStream_t stream, const int32_t* n_dev)
{
    (void)hipGetLastError();   /* the status read after the launches is theirs alone */
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_stats, dim3((n + 3) / 4), dim3(256), 0, stream, sp, nnw, counts,
                       crow, in, n, out, n_dev);
    return hipGetLastError();
}

hipError_t c2_launch_score_block(const ScoreCtx& c, int p, int Ld, float* S, float* mod, hipStream_t stream)
{
    (void)hipGetLastError();   /* the status read after the launches is theirs alone */
    const int64_t n = (int64_t)c.Lq * Ld;
    if (n <= 0) return hipSuccess;
    int64_t g = (n + 255) / 256;
    if (g > 65536) g = 65536;
    hipLaunchKernelGGL(k_score_block, dim3((unsigned)g), dim3(256), 0, stream, c, p, S, mod);
    return hipGetLastError();
}
