/*
 * c2_phase1.hip -- fused score matrix + phase-1 local Viterbi for gfx950.
 *
 * Replaces CalcSM_Init_SMEMUnroll2x + CalcSM_SSSS_SMEMUnroll2x
 * (cusco/CuBatchScoreMatrix_Init.cu:45-186, cusco/CuBatchScoreMatrix_ssss.cu:51-280),
 * the host loop of ExecDP_Btck_Unroll32x launches (cudp/CuBatchDP.cu:208-312,
 * kernel cudp/CuBatchDP_init_btck.cu:65-494) and the maximum reduction of
 * FinalizeDPPro (cudp/CuBatchDP_final.cu:96-147).  The score matrix is never
 * written to HBM.
 *
 * Work decomposition (one kernel launch per query / stripe set):
 *   - a wave owns 64*R query rows (R contiguous rows per lane) and sweeps a
 *     segment of whole DB profiles laid end to end; at step t lane l
 *     processes column c_begin + t - l for its R rows (anti-diagonal
 *     wavefront), so the cell above a lane's first row arrives from lane l-1
 *     through one DPP wave_shr:1 per state per step;
 *   - DB column data (20 weighted target frequencies, 7 transitions, 3 SS
 *     probabilities, profile-local x, length, profile id) is staged in a
 *     96-column LDS ring per wave, 32 columns at a time;
 *   - query data (target frequencies, transitions) lives in registers, the
 *     per-row 30-entry SS table QT and the column -1 boundary vectors in LDS
 *     shared by the workgroup;
 *   - a profile start is handled in a wave-uniform rare branch: flush the
 *     previous profile's maximum with a 64-bit atomicMax of a key that
 *     encodes (score, -x, -y), reset the row states to the reference's
 *     column -1 values;
 *   - queries longer than 64*R rows run in stripes; lane 63 hands its last
 *     row to the next stripe through a per-column HBM buffer.
 * BTCK=true additionally stores 2-bit backtrack codes packed per (step, lane)
 * for the pairs that pass to the finalize stage.
 */
#include <type_traits>
#include "c2_dev.h"
#include "c2_kernels.h"

namespace {

constexpr int RING = 96;      /* columns in the per-wave LDS ring */
constexpr int BLK = 32;       /* columns loaded per refill */
/* ring field offsets (each field is RING words) */
constexpr int F_DW = 0, F_DT = 20, F_PJ = 27, F_X = 30, F_LEN = 31, F_PID = 32, F_TOP = 33;
constexpr int NF_BASE = 33, NF_TOP = 38;

/* tops: the ring carries the stripe's top boundary; rtop: read it from the
 * hand-off buffer (stripes after the first) instead of the zero row -1 */
template <int R>
__device__ __forceinline__ void load_block(const P1Args& a, float* ring, int lane, int t0,
                                           int64_t c_begin, int64_t c_end, bool tops, bool rtop)
{
    if (lane >= BLK) return;
    const int slot = (t0 + lane) % RING;
    const int64_t c = c_begin + t0 + lane;
    if (c < c_end) {
        const int p = a.addrpro[c];
        const int x = (int)(c - (int64_t)a.dist[p]);
        for (int k = 0; k < 20; k++) {
            float w = 1.0f / (0.5f * (a.qbkg[k] + a.bkg[(size_t)k * a.nprof + p]));
            ring[(F_DW + k) * RING + slot] = a.trgfrq[(size_t)k * a.npos + c] * w;
        }
        const size_t ntr = (size_t)(a.npos + a.nprof);
        const size_t tb = (size_t)c + p;     /* row x of profile p: dist + p + x */
        ring[(F_DT + 0) * RING + slot] = a.trans[0 * ntr + tb];
        ring[(F_DT + 1) * RING + slot] = a.trans[1 * ntr + tb + 1];
        ring[(F_DT + 2) * RING + slot] = a.trans[2 * ntr + tb];
        ring[(F_DT + 3) * RING + slot] = a.trans[3 * ntr + tb];
        ring[(F_DT + 4) * RING + slot] = a.trans[4 * ntr + tb + 1];
        ring[(F_DT + 5) * RING + slot] = a.trans[5 * ntr + tb];
        ring[(F_DT + 6) * RING + slot] = a.trans[6 * ntr + tb];
        for (int j = 0; j < 3; j++)
            ring[(F_PJ + j) * RING + slot] = a.ssprb[(size_t)j * a.npos + c];
        ring[F_X * RING + slot] = __int_as_float(x);
        ring[F_LEN * RING + slot] = __int_as_float(a.len[p]);
        ring[F_PID * RING + slot] = __int_as_float(p);
        if (tops)
            for (int s = 0; s < 5; s++)
                ring[(F_TOP + s) * RING + slot] = rtop ? a.bot[(size_t)s * a.npos + c] : 0.0f;
    } else {
        for (int k = 0; k < 20; k++) ring[(F_DW + k) * RING + slot] = 0.0f;
        for (int k = 0; k < 7; k++) ring[(F_DT + k) * RING + slot] = -32768.0f;
        for (int j = 0; j < 3; j++) ring[(F_PJ + j) * RING + slot] = 0.0f;
        ring[F_X * RING + slot] = __int_as_float(-1);
        ring[F_LEN * RING + slot] = __int_as_float(1);
        ring[F_PID * RING + slot] = __int_as_float(-1);
        if (tops)
            for (int s = 0; s < 5; s++) ring[(F_TOP + s) * RING + slot] = 0.0f;
    }
}

__device__ __forceinline__ unsigned long long mkkey(float best, int x, int y)
{
    uint32_t lo = 0xFFFFFFFFu - (((uint32_t)x << 16) | (uint32_t)y);
    return ((unsigned long long)__float_as_uint(best) << 32) | lo;
}

template <int R, int NW, bool BTCK>
__global__ __launch_bounds__(64 * NW, 1) void k_phase1(P1Args a)
{
    extern __shared__ float lds[];
    const int SL = 64 * R;                      /* rows per stripe */
    float* sQT = lds;                           /* [30][SL] */
    float* sLL = sQT + 30 * SL;                 /* [5][SL] */
    float* sLD = sLL + 5 * SL;                  /* [5][SL] */
    const bool tops = a.nstripes > 1;
    const int nf = tops ? NF_TOP : NF_BASE;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    float* ring = sLD + 5 * SL + (size_t)wv * nf * RING;
    const int seg = blockIdx.x * NW + wv;

    for (int s = 0; s < a.nstripes; s++) {
        const int rbase = s * SL;
        __syncthreads();
        for (int i = threadIdx.x; i < 30 * SL; i += 64 * NW)
            sQT[i] = a.QT[(size_t)(i / SL) * a.Lpad + rbase + (i % SL)];
        for (int i = threadIdx.x; i < 5 * SL; i += 64 * NW) {
            sLL[i] = a.lleft[(size_t)(i / SL) * a.Lpad + rbase + (i % SL)];
            sLD[i] = a.ldiag[(size_t)(i / SL) * a.Lpad + rbase + (i % SL)];
        }
        __syncthreads();
        if (seg >= a.nseg) continue;

        const int pb = a.seg_pb[seg], pe = a.seg_pe[seg];
        const int64_t c_begin = a.dist[pb];
        const int64_t c_end = (int64_t)a.dist[pe - 1] + a.len[pe - 1];
        const int ncols = (int)(c_end - c_begin);
        const int nsteps = ncols + 63;
        const int y0 = rbase + R * lane;

        float q[R][20], qt[R][7], qh[R];
#pragma unroll
        for (int k = 0; k < R; k++) {
#pragma unroll
            for (int j = 0; j < 20; j++) q[k][j] = a.qv[(size_t)j * a.Lpad + y0 + k];
#pragma unroll
            for (int j = 0; j < 7; j++) qt[k][j] = a.qtr[(size_t)j * a.Lpad + y0 + k];
            qh[k] = a.qh[y0 + k];
        }
        int kf = -1;   /* row k>=1 with r%32==31 (its column-0 diagonal is seen as zeros) */
#pragma unroll
        for (int k = 1; k < R; k++)
            if (((y0 + k) & 31) == 31 && y0 + k < a.Lq) kf = k;

        /* ring: invalid columns everywhere, then the first block */
        for (int i = lane; i < RING; i += 64) {
            for (int k = 0; k < 20; k++) ring[(F_DW + k) * RING + i] = 0.0f;
            for (int k = 0; k < 7; k++) ring[(F_DT + k) * RING + i] = -32768.0f;
            for (int j = 0; j < 3; j++) ring[(F_PJ + j) * RING + i] = 0.0f;
            ring[F_X * RING + i] = __int_as_float(-1);
            ring[F_LEN * RING + i] = __int_as_float(1);
            ring[F_PID * RING + i] = __int_as_float(-1);
            if (tops) for (int j = 0; j < 5; j++) ring[(F_TOP + j) * RING + i] = 0.0f;
        }
        wave_lds_sync();
        load_block<R>(a, ring, lane, 0, c_begin, c_end, tops, s > 0);
        wave_lds_sync();

        S5 H[R];
#pragma unroll
        for (int k = 0; k < R; k++) H[k] = s5_zero();
        S5 U = s5_zero(), D = s5_zero();
        if (tops && lane == 0) {
            U.mm = ring[(F_TOP + 0) * RING]; U.mi = ring[(F_TOP + 1) * RING];
            U.im = ring[(F_TOP + 2) * RING]; U.dn = ring[(F_TOP + 3) * RING];
            U.nd = ring[(F_TOP + 4) * RING];
        }
        float best = 0.0f;
        int bpos = 0, curpid = -1, pstart = 0, cA = 0, cW = 0;
        const bool wbot = tops && (s + 1 < a.nstripes);
        uint32_t* btw = nullptr;
        if (BTCK) btw = a.btck + a.btck_off[seg] + (size_t)s * nsteps * 64;
        const bool ssterm = a.ssswgt > 0.0f;

        /* the R cells of one step; FIX: some lane is at a profile's column 0,
         * where row kf takes a zero diagonal (the reference's 32-row block
         * boundary) */
        auto cells = [&](auto fixc, int t, bool first, const float* dw, const float* dt, const float* pj,
                         int xloc) {
            constexpr bool FIX = decltype(fixc)::value;
            int co[3];
#pragma unroll
            for (int j = 0; j < 3; j++) co[j] = (10 * j + (int)(pj[j] * 9.999f)) * SL + R * lane;
            const float dh = (pj[2] + 0.1f) * 0.1111111f;
            const int u0 = (xloc - y0) - cA - 1;
            uint32_t bw = 0;
            S5 saved = D;
#pragma unroll
            for (int k = 0; k < R; k++) {
                /* score: log(sum_a q_a d_a w_a) - shift + SS term */
                float acc = q[k][0] * dw[0];
#pragma unroll
                for (int j = 1; j < 20; j++) acc = fmaf(q[k][j], dw[j], acc);
                float sc = c2_logf_clamped(acc) - a.shift;
                if (ssterm) {
                    float ss = sQT[co[0] + k] * pj[0];
                    ss = fmaf(sQT[co[1] + k], pj[1], ss);
                    ss = fmaf(sQT[co[2] + k], pj[2], ss);
                    ss = fmaf(qh[k], dh, ss);
                    ss = (ss - 0.05f) * a.ssswgt;
                    sc = sc + ss;
                }
                if ((uint32_t)(u0 - k) >= (uint32_t)cW) sc = -9.0f;   /* distant corners */
                S5 dg = saved;
                if (FIX && k > 0) {
                    const bool z = first && kf == k;
                    dg.mm = z ? 0.0f : dg.mm; dg.mi = z ? 0.0f : dg.mi; dg.im = z ? 0.0f : dg.im;
                    dg.dn = z ? 0.0f : dg.dn; dg.nd = z ? 0.0f : dg.nd;
                }
                const S5 up = (k == 0) ? U : H[k - 1];
                saved = H[k];
                const S5 n = dp_cell(dg, up, H[k], qt[k], dt, sc);
                H[k] = n;
                if (n.mm > best) { best = n.mm; bpos = (t << 3) | k; }
                if (BTCK) bw |= dp_btck(n) << (2 * k);
            }
            if (BTCK) btw[(size_t)t * 64 + lane] = bw;
        };

        /* one step t: ring refill, profile starts, cells, stripe hand-off,
         * row above for step t+1 */
        for (int t = 0; t < nsteps; t++) {
            if (((t + 1) % BLK) == 0) {
                load_block<R>(a, ring, lane, t + 1, c_begin, c_end, tops, s > 0);
                wave_lds_sync();
            }
            const int slot = (t - lane + RING * 4) % RING;
            float dw[20], dt[7], pj[3];
#pragma unroll
            for (int k = 0; k < 20; k++) dw[k] = ring[(F_DW + k) * RING + slot];
#pragma unroll
            for (int k = 0; k < 7; k++) dt[k] = ring[(F_DT + k) * RING + slot];
#pragma unroll
            for (int j = 0; j < 3; j++) pj[j] = ring[(F_PJ + j) * RING + slot];
            const int xloc = __float_as_int(ring[F_X * RING + slot]);

            /* profile start: flush, reset to the column -1 boundary */
            const bool first = (xloc == 0);
            if (__builtin_amdgcn_read_exec() & __ballot(first)) {
                if (first) {
                    if (curpid >= 0 && best > 0.0f) {
                        const int tb = bpos >> 3, kb = bpos & 7;
                        atomicMax(a.keys + curpid, mkkey(best, (tb - lane) - pstart, y0 + kb));
                    }
                    const int ldb = __float_as_int(ring[F_LEN * RING + slot]);
                    curpid = __float_as_int(ring[F_PID * RING + slot]);
                    pstart = t - lane;
                    best = 0.0f; bpos = 0;
                    const int mx = a.Lq > ldb ? a.Lq : ldb;
                    int e = (int)rintf(0.34f * (float)mx);
                    if (e > 60) e = 60;
                    cA = e - a.Lq;                 /* x - y <= A: bottom-left corner */
                    cW = (ldb - e) - cA - 1;       /* x - y >= B: top-right corner  */
#pragma unroll
                    for (int k = 0; k < R; k++) {
                        const int sl = R * lane + k;
                        H[k].mm = sLL[0 * SL + sl]; H[k].mi = sLL[1 * SL + sl];
                        H[k].im = sLL[2 * SL + sl]; H[k].dn = sLL[3 * SL + sl];
                        H[k].nd = sLL[4 * SL + sl];
                    }
                    const int sl0 = R * lane;
                    D.mm = sLD[0 * SL + sl0]; D.mi = sLD[1 * SL + sl0]; D.im = sLD[2 * SL + sl0];
                    D.dn = sLD[3 * SL + sl0]; D.nd = sLD[4 * SL + sl0];
                }
            }
            cells(std::true_type{}, t, first, dw, dt, pj, xloc);
            if (wbot && lane == 63) {
                const int64_t c = c_begin + t - 63;
                if (c >= c_begin && c < c_end) {
                    a.bot[0 * (size_t)a.npos + c] = H[R - 1].mm; a.bot[1 * (size_t)a.npos + c] = H[R - 1].mi;
                    a.bot[2 * (size_t)a.npos + c] = H[R - 1].im; a.bot[3 * (size_t)a.npos + c] = H[R - 1].dn;
                    a.bot[4 * (size_t)a.npos + c] = H[R - 1].nd;
                }
            }
            /* hand the last row to the next lane; lane 0 takes the stripe's top boundary */
            S5 top = s5_zero();
            if (tops) {
                const int ns = (t + 1) % RING;
                top.mm = ring[(F_TOP + 0) * RING + ns]; top.mi = ring[(F_TOP + 1) * RING + ns];
                top.im = ring[(F_TOP + 2) * RING + ns]; top.dn = ring[(F_TOP + 3) * RING + ns];
                top.nd = ring[(F_TOP + 4) * RING + ns];
            }
            D = U;
            U.mm = wave_shr1(H[R - 1].mm, top.mm);
            U.mi = wave_shr1(H[R - 1].mi, top.mi);
            U.im = wave_shr1(H[R - 1].im, top.im);
            U.dn = wave_shr1(H[R - 1].dn, top.dn);
            U.nd = wave_shr1(H[R - 1].nd, top.nd);
        }
        if (curpid >= 0 && best > 0.0f) {
            const int tb = bpos >> 3, kb = bpos & 7;
            atomicMax(a.keys + curpid, mkkey(best, (tb - lane) - pstart, y0 + kb));
        }
        if (tops) __threadfence();
    }
}

__global__ void k_zero_u64(unsigned long long* p, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

}  // namespace

/* clears the phase-1 maxima with a kernel (safe inside a captured graph) */
hipError_t c2_launch_zero_u64(unsigned long long* p, int64_t n, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_zero_u64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, p, n);
    return hipGetLastError();
}

size_t c2_phase1_lds_bytes(int R, int NW, int nstripes)
{
    const int nf = nstripes > 1 ? NF_TOP : NF_BASE;
    return sizeof(float) * ((size_t)40 * 64 * R + (size_t)NW * nf * RING);
}

#define C2_P1_CASE(RR, NWW, BT)                                                         \
    case RR:                                                                              \
        hipLaunchKernelGGL((k_phase1<RR, NWW, BT>), dim3(grid), dim3(64 * NWW), lds, stream, a); \
        break;

template <int NWW, bool BT>
static hipError_t launch_nw(const P1Args& a, int R, int grid, size_t lds, hipStream_t stream)
{
    static bool attr_set[6] = {false, false, false, false, false, false};
    switch (R) {
        case 1: if (!attr_set[1]) { (void)hipFuncSetAttribute((const void*)k_phase1<1, NWW, BT>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024); attr_set[1] = true; } break;
        case 2: if (!attr_set[2]) { (void)hipFuncSetAttribute((const void*)k_phase1<2, NWW, BT>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024); attr_set[2] = true; } break;
        case 3: if (!attr_set[3]) { (void)hipFuncSetAttribute((const void*)k_phase1<3, NWW, BT>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024); attr_set[3] = true; } break;
        case 4: if (!attr_set[4]) { (void)hipFuncSetAttribute((const void*)k_phase1<4, NWW, BT>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024); attr_set[4] = true; } break;
        case 5: if (!attr_set[5]) { (void)hipFuncSetAttribute((const void*)k_phase1<5, NWW, BT>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024); attr_set[5] = true; } break;
        default: return hipErrorInvalidValue;
    }
    switch (R) {
        C2_P1_CASE(1, NWW, BT)
        C2_P1_CASE(2, NWW, BT)
        C2_P1_CASE(3, NWW, BT)
        C2_P1_CASE(4, NWW, BT)
        C2_P1_CASE(5, NWW, BT)
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t c2_launch_phase1(const P1Args& a, int R, int NW, bool btck, hipStream_t stream)
{
    if (a.nseg <= 0) return hipSuccess;
    const int grid = (a.nseg + NW - 1) / NW;
    const size_t lds = c2_phase1_lds_bytes(R, NW, a.nstripes);
    if (btck) {
        switch (NW) {
            case 8: return launch_nw<8, true>(a, R, grid, lds, stream);
            case 4: return launch_nw<4, true>(a, R, grid, lds, stream);
            case 2: return launch_nw<2, true>(a, R, grid, lds, stream);
            case 1: return launch_nw<1, true>(a, R, grid, lds, stream);
            default: return hipErrorInvalidValue;
        }
    }
    switch (NW) {
        case 8: return launch_nw<8, false>(a, R, grid, lds, stream);
        case 4: return launch_nw<4, false>(a, R, grid, lds, stream);
        case 2: return launch_nw<2, false>(a, R, grid, lds, stream);
        case 1: return launch_nw<1, false>(a, R, grid, lds, stream);
        default: return hipErrorInvalidValue;
    }
}
