/*
 * c2_dev.h -- device-side building blocks shared by the gfx950 kernels.
 */
#ifndef C2_DEV_H
#define C2_DEV_H
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "c2_fmath.h"

/* raise a kernel's dynamic LDS limit once per process; the block-scope
 * static is initialised exactly once even when several host threads (one
 * handle each) launch the kernel for the first time together */
#define C2_LDS_LIMIT_ONCE(fn, bytes)                                                              \
    do {                                                                                          \
        static const bool c2_once_ = ((void)hipFuncSetAttribute((const void*)(fn),                \
                                                                hipFuncAttributeMaxDynamicSharedMemorySize, \
                                                                (int)(bytes)),                    \
                                      true);                                                      \
        (void)c2_once_;                                                                           \
    } while (0)

/* the five DP states of the reference (cudp/CuBatchDP_com.h:77-84) */
struct S5 { float mm, mi, im, dn, nd; };

__device__ __forceinline__ S5 s5_zero() { S5 z; z.mm = z.mi = z.im = z.dn = z.nd = 0.0f; return z; }

/* Transition slots as cached by DPLocCacheTrnProbs (cudp/CuBatchDP_init.cuh:190-201):
 * 0 MMp, 1 MIc, 2 MDp, 3 IMp, 4 IIc, 5 DMp, 6 DDp. */

/* One local-Viterbi cell, the recurrences of ExecDP_Btck_Unroll32x
 * (cudp/CuBatchDP_init_btck.cu:303-360).  Terms that share the same DB
 * transition are maximised before that transition is added; because FP32
 * rounding is monotonic, max(fl(a+d), fl(b+d)) == fl(max(a,b)+d), so the
 * result is bit-identical to the reference's term-by-term order.
 * dp_mm is the MM state alone, the only state that reads the diagonal. */
__device__ __forceinline__ float dp_mm(const S5& dg, const float* qt, const float* dt, float s)
{
    float m0 = fmaxf(fmaxf(dg.im + qt[3], dg.dn + qt[5]), dg.mm + qt[0]) + dt[0];
    float v1 = fmaxf(fmaxf((dg.mi + qt[0]) + dt[3], (dg.nd + qt[0]) + dt[5]), m0);
    return fmaxf(v1 + s, 0.0f);
}

/* the whole cell */
__device__ __forceinline__ S5 dp_cell(const S5& dg, const S5& up, const S5& lf,
                                      const float* qt, const float* dt, float s)
{
    S5 n;
    n.mm = dp_mm(dg, qt, dt, s);
    n.mi = fmaxf((up.mm + qt[0]) + dt[1], (up.mi + qt[0]) + dt[4]);
    n.im = fmaxf(lf.mm + qt[1], lf.im + qt[4]) + dt[0];
    n.dn = fmaxf(up.mm + qt[2], up.dn + qt[6]);
    n.nd = fmaxf(lf.mm + dt[2], lf.nd + dt[6]);
    return n;
}

/* backtrack code of a cell from its five states (btck logic of the same
 * kernel): STOP 0, LEFT 1, UP 2, DIAG 3; strict < as myhdmaxassgn */
__device__ __forceinline__ uint32_t dp_btck(const S5& n)
{
    float mx = n.mm; uint32_t b = (n.mm != 0.0f) ? 3u : 0u;
    if (mx < n.mi) { mx = n.mi; b = 2u; }
    if (mx < n.im) { mx = n.im; b = 1u; }
    if (mx < n.dn) { mx = n.dn; b = 2u; }
    if (mx < n.nd) { b = 1u; }
    return b;
}

/* whole-wave shift by one lane towards higher lanes (DPP wave_shr:1);
 * lane 0 receives 'old' */
__device__ __forceinline__ float wave_shr1(float v, float old)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(
        __builtin_bit_cast(int, old), __builtin_bit_cast(int, v), 0x138, 0xf, 0xf, false));
}

/* wave_shr1 with 0 into lane 0 (DPP bound_ctrl: the missing source reads
 * as zero, so no register has to hold the old value) */
__device__ __forceinline__ float wave_shr1_z(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(
        0, __builtin_bit_cast(int, v), 0x138, 0xf, 0xf, true));
}

/* whole-wave shift by one lane towards lower lanes (DPP wave_shl:1);
 * lane 63 receives 'old' */
__device__ __forceinline__ float wave_shl1(float v, float old)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(
        __builtin_bit_cast(int, old), __builtin_bit_cast(int, v), 0x130, 0xf, 0xf, false));
}

__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

#endif
