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

/* log(1 + e^d) for d <= 0, the correction term of a two-term log-sum-exp
 * (log(e^a + e^b) = max + log(1 + e^(min - max))): 0 for d <= -16 (the
 * reference's cut-off exp(x) = 0 for x <= -16), otherwise a degree-4
 * polynomial in u = d - (-16 + i/4) on the interval i = floor(4 (d + 16))
 * (clamped to 63); the coefficients are least-squares fits for this project
 * (comer2_amd/fit_lse_table.py), 8 floats per interval (c0..c4, padding).
 * Largest error against log1p(exp(d)) in double: 1.13e-7 absolute, against
 * about 1.2e-7 for log(1 + exp(d)) through c2_logf and c2_expf (whose
 * rounding of 1 + e^d also loses most of e^d's digits for d near -16). */
#define C2_LSE1P_TABLE { \
    0x1.e355ba0000000p-24f, 0x1.e355480000000p-24f, 0x1.e371a20000000p-25f, 0x1.3fce8e0000000p-26f, 0x1.6d696a0000000p-28f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.364ea00000000p-23f, 0x1.364e560000000p-23f, 0x1.36608a0000000p-24f, 0x1.9aa3dc0000000p-26f, 0x1.d532a60000000p-28f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.8e71360000000p-23f, 0x1.8e70d60000000p-23f, 0x1.8e88340000000p-24f, 0x1.07a2d80000000p-25f, 0x1.2d3b1a0000000p-27f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.ff9c240000000p-23f, 0x1.ff9baa0000000p-23f, 0x1.ffb9a80000000p-24f, 0x1.5283f00000000p-25f, 0x1.82c9b40000000p-27f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.4875c80000000p-22f, 0x1.4875780000000p-22f, 0x1.4888b80000000p-23f, 0x1.b2a9880000000p-25f, 0x1.f0a52c0000000p-27f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.a5c0460000000p-22f, 0x1.a5bfde0000000p-22f, 0x1.a5d8940000000p-23f, 0x1.170f000000000p-24f, 0x1.3eda300000000p-26f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.0ec5000000000p-21f, 0x1.0ec4bc0000000p-21f, 0x1.0ed4980000000p-22f, 0x1.6651740000000p-24f, 0x1.9969fe0000000p-26f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.5bacc80000000p-21f, 0x1.5bac700000000p-21f, 0x1.5bc0c80000000p-22f, 0x1.cc16e00000000p-24f, 0x1.06d9420000000p-25f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.be6c640000000p-21f, 0x1.be6bf00000000p-21f, 0x1.be860c0000000p-22f, 0x1.27620c0000000p-23f, 0x1.5180fa0000000p-25f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.1e9c060000000p-20f, 0x1.1e9bb80000000p-20f, 0x1.1eac780000000p-21f, 0x1.7b47680000000p-23f, 0x1.b15cce0000000p-25f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.7003880000000p-20f, 0x1.7003200000000p-20f, 0x1.70189a0000000p-21f, 0x1.e700d40000000p-23f, 0x1.1639360000000p-24f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.d889fa0000000p-20f, 0x1.d889700000000p-20f, 0x1.d8a4f40000000p-21f, 0x1.38a96a0000000p-22f, 0x1.653ea80000000p-24f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.2f603e0000000p-19f, 0x1.2f5fe00000000p-19f, 0x1.2f71820000000p-20f, 0x1.9177040000000p-22f, 0x1.cab5920000000p-24f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.858ad40000000p-19f, 0x1.858a540000000p-19f, 0x1.85a0e60000000p-20f, 0x1.01beb40000000p-21f, 0x1.267ed60000000p-23f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.f42e980000000p-19f, 0x1.f42de60000000p-19f, 0x1.f44ac60000000p-20f, 0x1.4af3320000000p-21f, 0x1.7a23080000000p-23f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.411f7e0000000p-18f, 0x1.411f020000000p-18f, 0x1.4131760000000p-19f, 0x1.a8f2480000000p-21f, 0x1.e5885a0000000p-23f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.9c54720000000p-18f, 0x1.9c53be0000000p-18f, 0x1.9c6b4a0000000p-19f, 0x1.10d1ce0000000p-20f, 0x1.37b6da0000000p-22f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.08b8840000000p-17f, 0x1.08b8020000000p-17f, 0x1.08c7020000000p-18f, 0x1.5e4e0c0000000p-20f, 0x1.903e220000000p-22f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.53e8660000000p-17f, 0x1.53e7a60000000p-17f, 0x1.53fab80000000p-18f, 0x1.c1cbe20000000p-20f, 0x1.00f4a60000000p-21f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.b4731c0000000p-17f, 0x1.b471fc0000000p-17f, 0x1.b48a240000000p-18f, 0x1.20c5940000000p-19f, 0x1.49edd20000000p-21f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.1834aa0000000p-16f, 0x1.1833ce0000000p-16f, 0x1.18430c0000000p-17f, 0x1.72c8f40000000p-19f, 0x1.a79f7c0000000p-21f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.67ca580000000p-16f, 0x1.67c9060000000p-16f, 0x1.67dc280000000p-17f, 0x1.dc16a80000000p-19f, 0x1.0ff5a80000000p-20f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.cdfa860000000p-16f, 0x1.cdf87a0000000p-16f, 0x1.ce10520000000p-17f, 0x1.31a5ca0000000p-18f, 0x1.5d2f080000000p-20f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.2898500000000p-15f, 0x1.2896b40000000p-15f, 0x1.28a56a0000000p-16f, 0x1.8872860000000p-18f, 0x1.c054340000000p-20f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.7cd5660000000p-15f, 0x1.7cd2d60000000p-15f, 0x1.7ce4be0000000p-16f, 0x1.f7e4840000000p-18f, 0x1.1fce7c0000000p-19f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.e8ff300000000p-15f, 0x1.e8fb180000000p-15f, 0x1.e910780000000p-16f, 0x1.437d3e0000000p-17f, 0x1.7181f80000000p-19f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.39f0880000000p-14f, 0x1.39ed3e0000000p-14f, 0x1.39f9a00000000p-15f, 0x1.9f57720000000p-17f, 0x1.da62c80000000p-19f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.931a240000000p-14f, 0x1.9314d20000000p-14f, 0x1.9322860000000p-15f, 0x1.0aa1ea0000000p-16f, 0x1.3080c00000000p-18f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.02cb0c0000000p-13f, 0x1.02c6ba0000000p-13f, 0x1.02cdb40000000p-14f, 0x1.5653940000000p-16f, 0x1.86e47c0000000p-18f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.4c4a8a0000000p-13f, 0x1.4c43800000000p-13f, 0x1.4c49780000000p-14f, 0x1.b77ef00000000p-16f, 0x1.f5c18c0000000p-18f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.aaa9240000000p-13f, 0x1.aa9da40000000p-13f, 0x1.aaa0600000000p-14f, 0x1.1a1cc80000000p-15f, 0x1.4200840000000p-17f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.11e9e60000000p-12f, 0x1.11e07e0000000p-12f, 0x1.11de300000000p-13f, 0x1.6a28700000000p-15f, 0x1.9d3ddc0000000p-17f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.5fb2f60000000p-12f, 0x1.5fa38c0000000p-12f, 0x1.5f99e20000000p-13f, 0x1.d0e2b60000000p-15f, 0x1.0920a40000000p-16f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.c391ae0000000p-12f, 0x1.c378620000000p-12f, 0x1.c360ec0000000p-13f, 0x1.2a5a060000000p-14f, 0x1.5422bc0000000p-16f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.21e5360000000p-11f, 0x1.21d06e0000000p-11f, 0x1.21b8460000000p-12f, 0x1.7ee8940000000p-14f, 0x1.b441de0000000p-16f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.74342a0000000p-11f, 0x1.7412040000000p-11f, 0x1.73e4000000000p-12f, 0x1.eb5d0a0000000p-14f, 0x1.17aed00000000p-15f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.dddef60000000p-11f, 0x1.dda6cc0000000p-11f, 0x1.dd53000000000p-12f, 0x1.3b36b80000000p-13f, 0x1.6676280000000p-15f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.32c2700000000p-10f, 0x1.32943c0000000p-10f, 0x1.324a160000000p-11f, 0x1.9455920000000p-13f, 0x1.cb30b20000000p-15f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.89d2540000000p-10f, 0x1.89864a0000000p-10f, 0x1.8905920000000p-11f, 0x1.03406e0000000p-12f, 0x1.25ea300000000p-14f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.f991b40000000p-10f, 0x1.f914880000000p-10f, 0x1.f838180000000p-11f, 0x1.4c55220000000p-12f, 0x1.77edd40000000p-14f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.447e360000000p-9f, 0x1.44172e0000000p-9f, 0x1.435c460000000p-10f, 0x1.a9d0340000000p-12f, 0x1.e04bc60000000p-14f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.a082d00000000p-9f, 0x1.9fd9360000000p-9f, 0x1.9e9e940000000p-10f, 0x1.10a1160000000p-11f, 0x1.3261c20000000p-13f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.0b48ec0000000p-8f, 0x1.0abd5a0000000p-8f, 0x1.09b6060000000p-9f, 0x1.5cd55a0000000p-11f, 0x1.862aca0000000p-13f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.5700900000000p-8f, 0x1.561ae40000000p-8f, 0x1.5464140000000p-9f, 0x1.bde48a0000000p-11f, 0x1.efb1660000000p-13f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.b818da0000000p-8f, 0x1.b69f0e0000000p-8f, 0x1.b3c6740000000p-9f, 0x1.1c9d4c0000000p-10f, 0x1.39ec500000000p-12f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.1a47880000000p-7f, 0x1.1910f20000000p-7f, 0x1.16b5f80000000p-8f, 0x1.6abe8a0000000p-10f, 0x1.8c0ede0000000p-12f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.6a03340000000p-7f, 0x1.6804e20000000p-7f, 0x1.6421ac0000000p-8f, 0x1.cd59800000000p-10f, 0x1.f127a60000000p-12f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.d01bb00000000p-7f, 0x1.ccd5ee0000000p-7f, 0x1.c66fca0000000p-8f, 0x1.2496ce0000000p-9f, 0x1.35fc860000000p-11f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.295e500000000p-6f, 0x1.26af700000000p-6f, 0x1.216f000000000p-7f, 0x1.71d7500000000p-9f, 0x1.7f48c40000000p-11f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.7cda8c0000000p-6f, 0x1.7875da0000000p-6f, 0x1.6fddbe0000000p-7f, 0x1.d16cfe0000000p-9f, 0x1.d4ad1e0000000p-11f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.e76e4c0000000p-6f, 0x1.e0402a0000000p-6f, 0x1.d23c3a0000000p-7f, 0x1.23342a0000000p-8f, 0x1.1a66fc0000000p-10f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.37a28a0000000p-5f, 0x1.31c8080000000p-5f, 0x1.2666540000000p-6f, 0x1.69c4ae0000000p-8f, 0x1.4dd58c0000000p-10f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.8e07100000000p-5f, 0x1.8483260000000p-5f, 0x1.721dfa0000000p-6f, 0x1.bd4c2c0000000p-8f, 0x1.8097800000000p-10f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.fba3740000000p-5f, 0x1.ec3ac00000000p-5f, 0x1.cead280000000p-6f, 0x1.0ed83c0000000p-7f, 0x1.abc0360000000p-10f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.4321e20000000p-4f, 0x1.36b70c0000000p-4f, 0x1.1f262e0000000p-5f, 0x1.448fb00000000p-7f, 0x1.c4bdd00000000p-10f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.9a72320000000p-4f, 0x1.868d320000000p-4f, 0x1.614e080000000p-5f, 0x1.7d90a60000000p-7f, 0x1.bd16ac0000000p-10f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.03f2d60000000p-3f, 0x1.e8416e0000000p-4f, 0x1.ae07460000000p-5f, 0x1.b5cb6a0000000p-7f, 0x1.835b480000000p-10f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.48239a0000000p-3f, 0x1.2f33780000000p-3f, 0x1.0249ec0000000p-4f, 0x1.e6f74a0000000p-7f, 0x1.062c180000000p-10f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.9c7e900000000p-3f, 0x1.759ba80000000p-3f, 0x1.316a6a0000000p-4f, 0x1.0446ba0000000p-6f, 0x1.cf892e0000000p-13f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.01f9b20000000p-2f, 0x1.c817340000000p-3f, 0x1.6278b20000000p-4f, 0x1.083f340000000p-6f, -0x1.be6fae0000000p-11f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.40c7ac0000000p-2f, 0x1.13657c0000000p-2f, 0x1.929bbe0000000p-4f, 0x1.f4ee9c0000000p-7f, -0x1.16ee460000000p-9f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.8c27ea0000000p-2f, 0x1.48857a0000000p-2f, 0x1.be32ee0000000p-4f, 0x1.aef4bc0000000p-7f, -0x1.c1477e0000000p-9f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.e574700000000p-2f, 0x1.829a180000000p-2f, 0x1.e140300000000p-4f, 0x1.3db66a0000000p-7f, -0x1.27420a0000000p-8f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
    0x1.26e1880000000p-1f, 0x1.c054d40000000p-2f, 0x1.f8114e0000000p-4f, 0x1.5135700000000p-8f, -0x1.4f83aa0000000p-8f, 0x0p+0f, 0x0p+0f, 0x0p+0f, \
}

/* interval and offset of d (d > -16) in the table */
C2_HD int c2_lse1p_index(float d, float* u) {
    int i = (int)((d + 16.0f) * 4.0f);
    i = i < 63 ? i : 63;
    *u = d - fmaf(0.25f, (float)i, -16.0f);
    return i;
}
/* the polynomial of one interval at offset u */
C2_HD float c2_lse1p_poly(float u, float c0, float c1, float c2, float c3, float c4) {
    float p = fmaf(c4, u, c3);
    p = fmaf(p, u, c2);
    p = fmaf(p, u, c1);
    return fmaf(p, u, c0);
}
/* log(1 + e^d) from a table tab[64][8] (host memory or a device copy) */
C2_HD float c2_lse1p(float d, const float* tab) {
    if (!(-16.0f < d)) return 0.0f;
    float u;
    const float* q = tab + 8 * c2_lse1p_index(d, &u);
    return c2_lse1p_poly(u, q[0], q[1], q[2], q[3], q[4]);
}

/* max without NaN handling; the DP never produces NaN or -0.0, so this is the
 * same value as the reference's  a<b ? b : a  (cucom/cucommon.h:61) */
C2_HD float c2_maxf(float a, float b) { return fmaxf(a, b); }

#endif /* C2_FMATH_H */
