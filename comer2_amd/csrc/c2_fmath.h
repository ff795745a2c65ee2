/*
 * c2_fmath.h -- deterministic FP32 elementary functions shared by the HIP
 * kernels and the host code.
 *
 * Every function here is written only with IEEE-754 basic operations
 * (add, sub, mul, correctly rounded fmaf, float<->int conversion and bit
 * manipulation), so a gfx950 kernel built with -ffp-contract=off and a host
 * build with gcc -ffp-contract=off return bit-identical results.  That
 * property is what lets the Viterbi backtrack decisions on the GPU agree bit
 * for bit with a CPU restatement of the same arithmetic.
 *
 * The polynomial coefficients are least-squares fits on Chebyshev nodes
 * (fitted for this project, not taken from any math library):
 *   c2_logf : log1p(f) = f + f^2 P7(f), f in [sqrt(1/2)-1, sqrt(2)-1];
 *             <= 1.5 ulp against log() in double over (1e-35, 1e35).
 *   c2_expf : exp(r) = 1 + r + r^2 Q4(r), |r| <= ln2/2, Cody-Waite reduction;
 *             <= 1.02 ulp against exp() in double over [-87, 88]; defined
 *             for x in [-87.68, 88.37]: 2^k is built in the exponent field,
 *             so k = rint(x log2 e) must lie in [-126, 127] (x >= 88.38
 *             gives +inf although exp(x) is finite there).  Every caller in
 *             this library passes x <= 0.
 * The reference uses CUDA fast intrinsics (__logf, __expf) and Schraudolph's
 * approximation for these operations (cusco/CuBatchScoreMatrix_Init.cu:129-177,
 * cucom/myfastapproxmath.cuh:12-40); neither is reproduced here.
 */
#ifndef C2_FMATH_H
#define C2_FMATH_H

#include <stdint.h>

#if defined(__HIPCC__)
#  include <hip/hip_runtime.h>
#  define C2_HD __host__ __device__ __forceinline__
#else
#  include <math.h>
#  include <string.h>
#  define C2_HD static inline
#endif

C2_HD uint32_t c2_f2u(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __float_as_uint(x);
#else
    uint32_t u; memcpy(&u, &x, sizeof u); return u;
#endif
}
C2_HD float c2_u2f(uint32_t u) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __uint_as_float(u);
#else
    float x; memcpy(&x, &u, sizeof x); return x;
#endif
}

/* smallest normal float; arguments of c2_logf are clamped to it */
#define C2_FLT_MIN_NORMAL 1.17549435e-38f

/* natural logarithm for x >= FLT_MIN (callers clamp; see c2_logf_clamped) */
#define C2_LOG_P7 0.08516032248735428f
#define C2_LOG_P6 (-0.14337590336799622f)
#define C2_LOG_P5 0.15015284717082977f
#define C2_LOG_P4 (-0.16570599377155304f)
#define C2_LOG_P3 0.19951635599136353f
#define C2_LOG_P2 (-0.25001534819602966f)
#define C2_LOG_P1 0.33334264159202576f
#define C2_LOG_P0 (-0.4999999403953552f)
#define C2_LN2_HI 0.693145751953125f
#define C2_LN2_LO 1.428606765330187e-06f
C2_HD float c2_logf(float x) {
    uint32_t ix = c2_f2u(x) - 0x3f3504f3u;          /* 0x3f3504f3 = sqrt(1/2) */
    int32_t k = ((int32_t)ix) >> 23;
    uint32_t iy = (ix & 0x007fffffu) + 0x3f3504f3u;
    float f = c2_u2f(iy) - 1.0f;                     /* exact */
    float p = C2_LOG_P7;
    p = fmaf(p, f, C2_LOG_P6);
    p = fmaf(p, f, C2_LOG_P5);
    p = fmaf(p, f, C2_LOG_P4);
    p = fmaf(p, f, C2_LOG_P3);
    p = fmaf(p, f, C2_LOG_P2);
    p = fmaf(p, f, C2_LOG_P1);
    p = fmaf(p, f, C2_LOG_P0);
    float r = fmaf(f * f, p, f);
    float kf = (float)k;
    return fmaf(kf, C2_LN2_HI, fmaf(kf, C2_LN2_LO, r));
}

/* arguments are sums of non-negative products (never NaN): fmaxf gives the
 * same value as a compare-and-select and is one instruction on gfx950 */
C2_HD float c2_logf_clamped(float x) {
    return c2_logf(fmaxf(x, C2_FLT_MIN_NORMAL));
}

#if defined(__HIPCC__)
/* c2_logf_clamped of N values, the same operations step by step across the
 * values, so that the N independent chains interleave in one wave */
template <int N>
__device__ __forceinline__ void c2_logf_clamped_n(const float* x, float* out) {
    float f[N], p[N], kf[N];
#pragma unroll
    for (int i = 0; i < N; i++) {
        uint32_t ix = c2_f2u(fmaxf(x[i], C2_FLT_MIN_NORMAL)) - 0x3f3504f3u;
        kf[i] = (float)(((int32_t)ix) >> 23);
        f[i] = c2_u2f((ix & 0x007fffffu) + 0x3f3504f3u) - 1.0f;
        p[i] = C2_LOG_P7;
    }
#pragma unroll
    for (int i = 0; i < N; i++) p[i] = fmaf(p[i], f[i], C2_LOG_P6);
#pragma unroll
    for (int i = 0; i < N; i++) p[i] = fmaf(p[i], f[i], C2_LOG_P5);
#pragma unroll
    for (int i = 0; i < N; i++) p[i] = fmaf(p[i], f[i], C2_LOG_P4);
#pragma unroll
    for (int i = 0; i < N; i++) p[i] = fmaf(p[i], f[i], C2_LOG_P3);
#pragma unroll
    for (int i = 0; i < N; i++) p[i] = fmaf(p[i], f[i], C2_LOG_P2);
#pragma unroll
    for (int i = 0; i < N; i++) p[i] = fmaf(p[i], f[i], C2_LOG_P1);
#pragma unroll
    for (int i = 0; i < N; i++) p[i] = fmaf(p[i], f[i], C2_LOG_P0);
#pragma unroll
    for (int i = 0; i < N; i++) {
        const float r = fmaf(f[i] * f[i], p[i], f[i]);
        out[i] = fmaf(kf[i], C2_LN2_HI, fmaf(kf[i], C2_LN2_LO, r));
    }
}
#endif

/* exponential for x in [-87.68, 88.37] (see the header); callers clamp or
 * threshold */
C2_HD float c2_expf(float x) {
    float kf = rintf(x * 1.4426950408889634f);
    float r = fmaf(kf, -0.693145751953125f, x);
    r = fmaf(kf, -1.428606765330187e-06f, r);
    float p = 0.0013946504332125187f;
    p = fmaf(p, r, 0.008369187824428082f);
    p = fmaf(p, r, 0.04166626185178757f);
    p = fmaf(p, r, 0.16666515171527863f);
    p = fmaf(p, r, 0.5f);
    float e = fmaf(r * r, p, r) + 1.0f;
    int32_t k = (int32_t)kf;
    return e * c2_u2f((uint32_t)(k + 127) << 23);
}

/* max without NaN handling; the DP never produces NaN or -0.0, so this is the
 * same value as the reference's  a<b ? b : a  (cucom/cucommon.h:61) */
C2_HD float c2_maxf(float a, float b) { return fmaxf(a, b); }

#endif /* C2_FMATH_H */
