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
 *     96-column LDS ring per wave, 32 columns at a time, and read by each
 *     lane one step ahead of its use;
 *   - the R rows of a step are scored stage by stage across the rows (R
 *     independent FMA and logarithm chains interleaved in one wave), then
 *     the Viterbi recurrences run row by row;
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
#include "c2_dev.h"
#include "c2_kernels.h"

namespace {

constexpr int RING = 96;      /* columns in the per-wave LDS ring */
constexpr int BLK = 32;       /* columns loaded per refill */
/* the ring is slot-major: column slot s holds its fields at words
 * [s*NS, s*NS + NS), so a lane reads all of its column's fields at immediate
 * offsets from one address; NS is odd, so the 64 lanes (64 consecutive slots)
 * hit 64 distinct banks.  Field offsets within a slot: */
constexpr int F_DW = 0, F_DT = 20, F_PJ = 27, F_X = 30, F_LEN = 31, F_PID = 32, F_BIN = 33, F_TOP = 34;
constexpr int NS_BASE = 35, NS_TOP = 39;   /* slot stride without / with the stripe-top fields */
/* with the context-vector term: the column's 19 entries, its prior as used
 * and |d|^2 + 1 + 1/kappa after the stripe-top fields */
constexpr int F_CV = 39, F_CVP = 58, F_CVA = 59, NS_CV = 61;
constexpr int CV_MAP = 128;                /* LDS words for the CV map */

/* tops: the ring carries the stripe's top boundary; rtop: read it from the
 * hand-off buffer (stripes after the first) instead of the zero row -1 */
/* the profile of the column lane loads in the refill of block t0 (loaded
 * half a block ahead, so the refill's loads do not wait on it) */
template <int BK, class PA>
__device__ __forceinline__ int block_profile(PA& a, int lane, int t0, int64_t c_begin, int64_t c_end)
{
    const int64_t c = c_begin + t0 + lane;
    return (lane < BK && c < c_end) ? a.addrpro[c] : -1;
}

template <int R, bool CV, int BK, class PA>
__device__ __forceinline__ void load_block(PA& a, float* ring, int ns, int lane, int t0,
                                           int64_t c_begin, int64_t c_end, bool tops, bool rtop, int p,
                                           const float* bot, int64_t bld, int64_t bc0)
{
    if (lane >= BK) return;
    float* rc = ring + ((t0 + lane) % RING) * ns;
    const int64_t c = c_begin + t0 + lane;
    if (c < c_end) {
        const int x = (int)(c - (int64_t)a.dist[p]);
        /* the field stride is made opaque here so that the 20 per-field base
         * addresses are formed per refill instead of being held in registers
         * across the step loop */
        int64_t npos = a.npos;
        asm volatile("" : "+s"(npos));
        const float* tf = a.trgfrq + c;
        if constexpr (R >= 5) {
            /* the widest kernel keeps its registers for the rows: the weights
             * are formed here (the same FP32 operations as the table) */
            for (int k = 0; k < 20; k++) {
                const float w = 1.0f / (0.5f * (a.qbkg[k] + a.bkg[(size_t)k * a.nprof + p]));
                rc[F_DW + k] = a.trgfrq[(size_t)k * a.npos + c] * w;
            }
        } else {
            const float* wp = a.wts + (size_t)p * 20;
            for (int k = 0; k < 20; k++)
                rc[F_DW + k] = tf[(size_t)k * npos] * wp[k];
        }
        size_t ntr = (size_t)(a.npos + a.nprof);
        asm volatile("" : "+s"(ntr));
        const float* tr = a.trans + (size_t)c + p;     /* row x of profile p: dist + p + x */
        rc[F_DT + 0] = tr[0 * ntr];
        rc[F_DT + 1] = tr[1 * ntr + 1];
        rc[F_DT + 2] = tr[2 * ntr];
        rc[F_DT + 3] = tr[3 * ntr];
        rc[F_DT + 4] = tr[4 * ntr + 1];
        rc[F_DT + 5] = tr[5 * ntr];
        rc[F_DT + 6] = tr[6 * ntr];
        const float* sp = a.ssprb + c;
        for (int j = 0; j < 3; j++)
            rc[F_PJ + j] = sp[(size_t)j * npos];
        rc[F_X] = __int_as_float(x);
        rc[F_LEN] = __int_as_float(a.len[p]);
        rc[F_PID] = __int_as_float(p);
        /* the column's SS-table bins int(9.999 p'_j), 4 bits each */
        const uint32_t bins = (uint32_t)(int)(sp[0] * 9.999f) | ((uint32_t)(int)(sp[(size_t)npos] * 9.999f) << 4) |
                              ((uint32_t)(int)(sp[2 * (size_t)npos] * 9.999f) << 8);
        rc[F_BIN] = __uint_as_float(bins);
        if (tops) {
            const float* bt = bot + (c - bc0);
            for (int s = 0; s < 5; s++)
                rc[F_TOP + s] = rtop ? bt[(size_t)s * bld] : 0.0f;
        }
        if constexpr (CV) {
            const float* cv = a.cvs.cv + c;
            for (int i = 0; i < 19; i++) rc[F_CV + i] = cv[(size_t)i * npos];
            const float pr = a.cvs.cvp[c];
            rc[F_CVP] = a.cvs.trunc ? truncf(pr) : pr;
            rc[F_CVA] = (a.cvs.cvn[c] + 1.0f) + a.cvs.kinv;
        }
    } else {
        for (int k = 0; k < 20; k++) rc[F_DW + k] = 0.0f;
        for (int k = 0; k < 7; k++) rc[F_DT + k] = -32768.0f;
        for (int j = 0; j < 3; j++) rc[F_PJ + j] = 0.0f;
        rc[F_X] = __int_as_float(-1);
        rc[F_LEN] = __int_as_float(1);
        rc[F_PID] = __int_as_float(-1);
        rc[F_BIN] = 0.0f;
        if (tops)
            for (int s = 0; s < 5; s++) rc[F_TOP + s] = 0.0f;
        if constexpr (CV) {
            for (int i = 0; i < 19; i++) rc[F_CV + i] = 0.0f;
            rc[F_CVP] = 0.0f;
            rc[F_CVA] = 1.0f;
        }
    }
}

__device__ __forceinline__ unsigned long long mkkey(float best, int x, int y)
{
    uint32_t lo = 0xFFFFFFFFu - (((uint32_t)x << 16) | (uint32_t)y);
    return ((unsigned long long)__float_as_uint(best) << 32) | lo;
}

/* P1Args read through the constant address space (loads may be scalar) */
typedef const __attribute__((address_space(4))) P1Args CP1Args;

/* the sweep of one segment (all stripes of the query) by wave wv of the
 * workgroup; every wave of the workgroup calls it with the same query, and
 * waves without a segment (seg >= nseg) still take part in the barriers.
 * botw != nullptr: the stripe hand-off goes to the wave's own scratch rows
 * of bot_ld columns (indexed from the segment's first column) instead of the
 * query's per-column buffer a.bot */
template <int R, int NW, bool BTCK, bool CV, bool MF, class PA>
__device__ __forceinline__ void p1_sweep(PA& a, int seg, float* lds, float* botw, int64_t bot_ld)
{
    static_assert(!(MF && CV), "the matrix-core contraction has no context-vector form");
    /* refill: BK columns, LEAD steps ahead of the column that starts the
     * block.  The matrix-core form reads columns up to 3 ahead of lane 0 at
     * every fourth step, so it refills 16 columns 4 steps ahead (the ring
     * still holds every column a lane reads: a refill at the end of step t
     * overwrites columns t-92..t-77, last read at step t-14) */
    constexpr int BK = MF ? 16 : BLK;
    constexpr int LEAD = MF ? 4 : 2;
    const int SL = 64 * R;                      /* rows per stripe */
    float* sQT = lds;                           /* [30][SL] */
    float* sLL = sQT + 30 * SL;                 /* [5][SL] */
    float* sLD = sLL + 5 * SL;                  /* [5][SL] */
    const bool tops = a.nstripes > 1;
    const int ns = CV ? NS_CV : (tops ? NS_TOP : NS_BASE);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    float* ring = sLD + 5 * SL + (size_t)wv * ns * RING;
    /* the CV map after the rings */
    float* sMap = sLD + 5 * SL + (size_t)NW * ns * RING;
    if constexpr (CV) {
        for (int i = threadIdx.x; i < CV_MAP; i += 64 * NW)
            sMap[i] = i < a.cvs.card ? a.cvs.map[i] : 0.0f;
    }

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
        const int64_t c_full = (int64_t)a.dist[pe - 1] + a.len[pe - 1];
        int64_t c_end = c_full;
        if (BTCK) {
            /* one candidate profile: the backtrace starts at the maximum's
             * column x (the key phase 1 left), so columns past x are not swept;
             * the codes keep the whole profile's stride */
            const uint32_t xy = 0xFFFFFFFFu - (uint32_t)a.keys[pb];
            const int64_t cut = c_begin + (int64_t)(xy >> 16) + 1;
            if (cut < c_end) c_end = cut;
        }
        const int ncols = (int)(c_end - c_begin);
        const int nsteps = ncols + 63;
        const int bstride = (int)(c_full - c_begin) + 63;   /* backtrack-code steps per stripe */
        const int y0 = rbase + R * lane;
        float* bot = a.bot;
        int64_t bld = a.npos, bc0 = 0;
        if (botw) { bot = botw; bld = bot_ld; bc0 = c_begin; }

        float q[R][20], qt[R][7], qh[R];
#pragma unroll
        for (int k = 0; k < R; k++) {
#pragma unroll
            for (int j = 0; j < 20; j++) q[k][j] = a.qv[(size_t)j * a.Lpad + y0 + k];
#pragma unroll
            for (int j = 0; j < 7; j++) qt[k][j] = a.qtr[(size_t)j * a.Lpad + y0 + k];
            qh[k] = a.qh[y0 + k];
        }
        /* the rows' context vectors, |q|^2 + 1 + 1/kappa and priors as used */
        float qc[CV ? R : 1][19], qa[CV ? R : 1], rq[CV ? R : 1];
        if constexpr (CV) {
#pragma unroll
            for (int k = 0; k < R; k++) {
#pragma unroll
                for (int i = 0; i < 19; i++) qc[k][i] = a.cvs.qcv[(size_t)i * a.Lpad + y0 + k];
                const float pr = a.cvs.qcv[(size_t)19 * a.Lpad + y0 + k];
                rq[k] = a.cvs.trunc ? truncf(pr) : pr;
                qa[k] = (a.cvs.qcv[(size_t)20 * a.Lpad + y0 + k] + 1.0f) + a.cvs.kinv;
            }
        }
        int kf = -1;   /* row k>=1 with r%32==31 (its column-0 diagonal is seen as zeros) */
#pragma unroll
        for (int k = 1; k < R; k++)
            if (((y0 + k) & 31) == 31 && y0 + k < a.Lq) kf = k;

        /* ring: invalid columns everywhere, then the first block */
        for (int i = lane; i < RING; i += 64) {
            float* rc = ring + i * ns;
            for (int k = 0; k < 20; k++) rc[F_DW + k] = 0.0f;
            for (int k = 0; k < 7; k++) rc[F_DT + k] = -32768.0f;
            for (int j = 0; j < 3; j++) rc[F_PJ + j] = 0.0f;
            rc[F_X] = __int_as_float(-1);
            rc[F_LEN] = __int_as_float(1);
            rc[F_PID] = __int_as_float(-1);
            rc[F_BIN] = 0.0f;
            if (tops) for (int j = 0; j < 5; j++) rc[F_TOP + j] = 0.0f;
            if constexpr (CV) {
                for (int j = 0; j < 19; j++) rc[F_CV + j] = 0.0f;
                rc[F_CVP] = 0.0f;
                rc[F_CVA] = 1.0f;
            }
        }
        wave_lds_sync();
        int pn = block_profile<BK>(a, lane, 0, c_begin, c_end);
        load_block<R, CV, BK>(a, ring, ns, lane, 0, c_begin, c_end, tops, s > 0, pn, bot, bld, bc0);
        wave_lds_sync();

        S5 H[R];
#pragma unroll
        for (int k = 0; k < R; k++) H[k] = s5_zero();
        S5 U = s5_zero(), D = s5_zero();
        if (tops && lane == 0) {
            U.mm = ring[F_TOP + 0]; U.mi = ring[F_TOP + 1];
            U.im = ring[F_TOP + 2]; U.dn = ring[F_TOP + 3];
            U.nd = ring[F_TOP + 4];
        }
        float best = 0.0f;
        int bpos = 0, curpid = -1, pstart = 0, cA = 0, cW = 0;
        const bool wbot = tops && (s + 1 < a.nstripes);
        uint32_t* btw = nullptr;
        if (BTCK) btw = a.btck + a.btck_off[seg] + (size_t)s * bstride * 64;

        /* word offset of the lane's column slot, ((t - lane) mod RING) * ns,
         * advanced by one slot per step */
        const int soff_end = (RING - 1) * ns;
        int soff = ((RING * 4 - lane) % RING) * ns;
        /* a step's column fields (with the profile length and id a profile
         * start needs) are read at the end of the step before, after that
         * step's refill: the ring then holds the 64 columns the lanes read next
         * and the 32 columns of the block after them */
        float dw[20], dt[7], pj[3];
        float dc[CV ? 19 : 1], dpr = 0.0f, dna = 0.0f;
        int xloc, nlen, npid;
        uint32_t bb;
        auto read_col = [&]() {
            /* one LDS address for all of the column's fields: opaque, so that
             * the paired reads take small offsets from it instead of one
             * address per pair */
            const __attribute__((address_space(3))) float* rs =
                (const __attribute__((address_space(3))) float*)(ring + soff);
            asm volatile("" : "+v"(rs));
            if constexpr (!MF) {
#pragma unroll
                for (int k = 0; k < 20; k++) dw[k] = rs[F_DW + k];
            }
#pragma unroll
            for (int k = 0; k < 7; k++) dt[k] = rs[F_DT + k];
#pragma unroll
            for (int j = 0; j < 3; j++) pj[j] = rs[F_PJ + j];
            xloc = __float_as_int(rs[F_X]);
            nlen = __float_as_int(rs[F_LEN]);
            npid = __float_as_int(rs[F_PID]);
            bb = __float_as_uint(rs[F_BIN]);
            if constexpr (CV) {
#pragma unroll
                for (int i = 0; i < 19; i++) dc[i] = rs[F_CV + i];
                dpr = rs[F_CVP];
                dna = rs[F_CVA];
            }
        };
        /* matrix-core form: at every fourth step t = 4s each group of four
         * lanes (g = lane/4, j = lane%4) forms, on the matrix cores, the
         * 20-term sums of its rows against four columns with
         * v_mfma_f32_4x4x1_16b_f32 chained over the 20 terms (one product
         * added per instruction: the same roundings as the fmaf chain).  Lane
         * 4g+i supplies column t-4g+i (operand A), lane 4g+j its row's
         * frequency (operand B); register i of lane 4g+j then holds row j x
         * column t-4g+i.  Lane 4g+j meets column t+u-4g-j at step t+u, so it
         * reads register u-j of this tile (u >= j) or 4+u-j of the previous
         * one; the logarithm is taken on the tile, four cells per row. */
        float tcur[MF ? R : 1][4], tprv[MF ? R : 1][4];
        const int jl = lane & 3;
        if constexpr (MF) {
            const float z = c2_logf_clamped(0.0f);   /* what an empty (invalid) column scores */
#pragma unroll
            for (int k = 0; k < R; k++)
#pragma unroll
                for (int i = 0; i < 4; i++) tcur[k][i] = z;
        }
        auto tile = [&](int t) {
            typedef float f4 __attribute__((ext_vector_type(4)));
#pragma unroll
            for (int k = 0; k < R; k++)
#pragma unroll
                for (int i = 1; i < 4; i++) tprv[k][i] = tcur[k][i];
            const int ca = t - (lane & ~3) + jl;             /* >= t - 60 */
            const int slot = (ca + 4 * RING) % RING;
            const __attribute__((address_space(3))) float* rs =
                (const __attribute__((address_space(3))) float*)(ring + slot * ns + F_DW);
            f4 acc[R];
#pragma unroll
            for (int k = 0; k < R; k++) acc[k] = f4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int kk = 0; kk < 20; kk++) {
                const float av = rs[kk];
#pragma unroll
                for (int k = 0; k < R; k++) acc[k] = __builtin_amdgcn_mfma_f32_4x4x1f32(av, q[k][kk], acc[k], 0, 0, 0);
            }
            /* the logarithms one column of the tile at a time (R chains) */
#pragma unroll
            for (int i = 0; i < 4; i++) {
                float raw[R], lg[R];
#pragma unroll
                for (int k = 0; k < R; k++) raw[k] = acc[k][i];
                c2_logf_clamped_n<R>(raw, lg);
#pragma unroll
                for (int k = 0; k < R; k++) tcur[k][i] = lg[k];
            }
        };
        /* the step's R scores from the tiles: step t = 4s+u, lane 4g+j reads
         * register u-j of the current tile or 4+u-j of the previous one */
        auto select = [&](auto UC, float* sc) __attribute__((always_inline)) {
            constexpr int u = decltype(UC)::value;
#pragma unroll
            for (int k = 0; k < R; k++) {
                const float x0 = tcur[k][u];
                const float x1 = u >= 1 ? tcur[k][u >= 1 ? u - 1 : 0] : tprv[k][3];
                const float x2 = u >= 2 ? tcur[k][u >= 2 ? u - 2 : 0] : tprv[k][u + 2 < 4 ? u + 2 : 3];
                const float x3 = u >= 3 ? tcur[k][0] : tprv[k][u + 1 < 4 ? u + 1 : 3];
                sc[k] = jl == 0 ? x0 : (jl == 1 ? x1 : (jl == 2 ? x2 : x3));
            }
        };
        read_col();
        for (int t = 0; t < nsteps; t++) {
            /* profile start: flush, reset to the column -1 boundary */
            const bool first = (xloc == 0);
            uint32_t fixbits = 0;
            if (__builtin_amdgcn_read_exec() & __ballot(first)) {
                if (first) {
                    if (curpid >= 0 && best > 0.0f) {
                        const int tb = bpos >> 3, kb = bpos & 7;
                        atomicMax(a.keys + curpid, mkkey(best, (tb - lane) - pstart, y0 + kb));
                    }
                    const int ldb = nlen;
                    curpid = npid;
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
#pragma unroll
                for (int k = 1; k < R; k++)
                    if (__ballot(first && kf == k)) fixbits |= 1u << k;
            }

            /* per-column SS table offsets: QT[10j + int(9.999 p'_j)] */
            int co[3];
            {
#pragma unroll
                for (int j = 0; j < 3; j++) co[j] = (10 * j + (int)((bb >> (4 * j)) & 15u)) * SL + R * lane;
            }
            const float dh = (pj[2] + 0.1f) * 0.1111111f;
            const int u0 = (xloc - y0) - cA - 1;

            /* the rows of one step: their scores first, each stage across all
             * rows (independent chains interleave), then the DP row by row; the
             * SS term is added with weight 0 when it is off (an exact +-0) */
            float sc[R];
            {
                if constexpr (MF) {
                    /* t is wave-uniform: a scalar branch picks the selection */
                    switch (t & 3) {
                        case 0: tile(t); select(std::integral_constant<int, 0>{}, sc); break;
                        case 1: select(std::integral_constant<int, 1>{}, sc); break;
                        case 2: select(std::integral_constant<int, 2>{}, sc); break;
                        default: select(std::integral_constant<int, 3>{}, sc); break;
                    }
                } else {
                    float acc[R];
#pragma unroll
                    for (int k = 0; k < R; k++) acc[k] = q[k][0] * dw[0];
#pragma unroll
                    for (int j = 1; j < 20; j++)
#pragma unroll
                        for (int k = 0; k < R; k++) acc[k] = fmaf(q[k][j], dw[j], acc[k]);
                    c2_logf_clamped_n<R>(acc, sc);
                }
                const float ssw = a.ssswgt > 0.0f ? a.ssswgt : 0.0f;
                float t0[R], t1[R], t2[R];
#pragma unroll
                for (int k = 0; k < R; k++) {
                    t0[k] = sQT[co[0] + k]; t1[k] = sQT[co[1] + k]; t2[k] = sQT[co[2] + k];
                }
#pragma unroll
                for (int k = 0; k < R; k++) {
                    float ss = t0[k] * pj[0];
                    ss = fmaf(t1[k], pj[1], ss);
                    ss = fmaf(t2[k], pj[2], ss);
                    ss = fmaf(qh[k], dh, ss);
                    ss = (ss - 0.05f) * ssw;
                    if constexpr (CV) {
                        /* the CV map score (c2_cvs_x + the map lookup of the
                         * score kernels), then the terms combined as there */
                        float dot = qc[k][0] * dc[0];
#pragma unroll
                        for (int i = 1; i < 19; i++) dot = fmaf(qc[k][i], dc[i], dot);
                        dot = dot + a.cvs.kinv;
                        const float det = fmaf(qa[k], dna, -(dot * dot));
                        float x = fmaf(a.cvs.nu, c2_logf_clamped(det), a.cvs.cterm);
                        x = x - rq[k];
                        x = x - dpr;
                        const float cm = x <= a.cvs.k1 ? a.cvs.v1
                                       : (a.cvs.kl <= x ? a.cvs.vl : sMap[a.cvs.shift + (int)rintf(2.0f * x)]);
                        const float wv = a.cvs.wgt;
                        sc[k] = sc[k] - a.shift;
                        if (a.ssswgt > 0.0f) sc[k] = sc[k] + (ss + cm * wv);
                        else if (cm != 0.0f) sc[k] = sc[k] + cm * wv;
                    } else {
                        sc[k] = (sc[k] - a.shift) + ss;
                    }
                    if ((uint32_t)(u0 - k) >= (uint32_t)cW) sc[k] = -9.0f;   /* distant corners */
                }
            }
            /* the stripe top of the next step's column, read ahead of the rows */
            S5 top;
            if (tops) {
                const float* rt = ring + ((t + 1) % RING) * ns;
                top.mm = rt[F_TOP + 0]; top.mi = rt[F_TOP + 1];
                top.im = rt[F_TOP + 2]; top.dn = rt[F_TOP + 3];
                top.nd = rt[F_TOP + 4];
            }
            S5 saved = D;
            uint32_t bw = 0;
            auto row = [&](const int k, const bool fix) {
                const S5 dg = saved;
                const S5 up = (k == 0) ? U : H[k - 1];
                saved = H[k];
                S5 n = dp_cell(dg, up, H[k], qt[k], dt, sc[k]);
                /* row kf at a profile's first column: its diagonal reads zeros */
                if (fix && k > 0 && (fixbits & (1u << k)) && first && kf == k)
                    n.mm = dp_mm(s5_zero(), qt[k], dt, sc[k]);
                H[k] = n;
                if (BTCK) bw |= dp_btck(n) << (2 * k);
            };
            if (fixbits == 0) {
#pragma unroll
                for (int k = 0; k < R; k++) row(k, false);
            } else {
#pragma unroll
                for (int k = 0; k < R; k++) row(k, true);
            }
            /* running maximum: the step's largest cell, the first row on ties
             * (the order of a row-by-row strict-greater update) */
            {
                float m = H[0].mm;
#pragma unroll
                for (int k = 1; k < R; k++) m = fmaxf(m, H[k].mm);
                if (m > best) {
                    int kb = R - 1;
#pragma unroll
                    for (int k = R - 2; k >= 0; k--) if (H[k].mm == m) kb = k;
                    best = m;
                    bpos = (t << 3) | kb;
                }
            }
            if (BTCK) btw[(size_t)t * 64 + lane] = bw;
            if (wbot && t >= 63 && t - 63 < ncols) {
                if (lane == 63) {
                    float* bt = bot + ((c_begin + t - 63) - bc0);
                    bt[0 * (size_t)bld] = H[R - 1].mm; bt[1 * (size_t)bld] = H[R - 1].mi;
                    bt[2 * (size_t)bld] = H[R - 1].im; bt[3 * (size_t)bld] = H[R - 1].dn;
                    bt[4 * (size_t)bld] = H[R - 1].nd;
                }
            }
            /* hand the last row to the next lane; lane 0 takes the stripe's
             * top boundary (zeros in the first stripe) */
            D = U;
            if (tops) {
                U.mm = wave_shr1(H[R - 1].mm, top.mm);
                U.mi = wave_shr1(H[R - 1].mi, top.mi);
                U.im = wave_shr1(H[R - 1].im, top.im);
                U.dn = wave_shr1(H[R - 1].dn, top.dn);
                U.nd = wave_shr1(H[R - 1].nd, top.nd);
            } else {
                U.mm = wave_shr1_z(H[R - 1].mm);
                U.mi = wave_shr1_z(H[R - 1].mi);
                U.im = wave_shr1_z(H[R - 1].im);
                U.dn = wave_shr1_z(H[R - 1].dn);
                U.nd = wave_shr1_z(H[R - 1].nd);
            }
            soff = (soff == soff_end) ? 0 : soff + ns;
            /* the block after next: its first column is the top boundary the
             * next step reads ahead of its rows */
            if (((t + LEAD) % BK) == 0) {
                load_block<R, CV, BK>(a, ring, ns, lane, t + LEAD, c_begin, c_end, tops, s > 0,
                                      block_profile<BK>(a, lane, t + LEAD, c_begin, c_end), bot, bld, bc0);
                wave_lds_sync();
            }
            read_col();
        }
        if (curpid >= 0 && best > 0.0f) {
            const int tb = bpos >> 3, kb = bpos & 7;
            atomicMax(a.keys + curpid, mkkey(best, (tb - lane) - pstart, y0 + kb));
        }
        if (tops) __threadfence();
    }
}

template <int R, int NW, bool BTCK, bool CV = false, bool MF = false>
__global__ __launch_bounds__(64 * NW, 1) void k_phase1(P1Args a)
{
    extern __shared__ float lds[];
    p1_sweep<R, NW, BTCK, CV, MF, const P1Args>(a, blockIdx.x * NW + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)),
                                                lds, nullptr, 0);
}

/* persistent form for a batch of queries: workgroups pull tiles (query,
 * group of NW segments) from a device work queue until it is empty.  The
 * tiles of one launch share the rows per lane R (one length bin) and come in
 * the order the host sorted them (largest first); every wave hands its
 * stripes over through its own scratch rows (5 x bot_ld floats per wave). */
template <int R, int NW, bool BTCK, bool MF = false>
__global__ __launch_bounds__(64 * NW, 1) void k_phase1q(P1QueueArgs qa)
{
    extern __shared__ float lds[];
    __shared__ int tile_s;
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    float* botw = qa.botw ? qa.botw + ((size_t)blockIdx.x * NW + wv) * 5 * qa.bot_ld : nullptr;
    const int ntiles = qa.ntiles_dev ? __builtin_amdgcn_readfirstlane(*qa.ntiles_dev) : qa.ntiles;
    for (;;) {
        if (threadIdx.x == 0) tile_s = (int)atomicAdd(qa.head, 1u);
        __syncthreads();
        const int ti = __builtin_amdgcn_readfirstlane(tile_s);   /* uniform: the tile's arguments load as scalars */
        __syncthreads();
        if (ti >= ntiles) break;
        const int tq = __builtin_amdgcn_readfirstlane(qa.tiles[ti].q);
        const int tg = __builtin_amdgcn_readfirstlane(qa.tiles[ti].g);
        /* the arguments are read-only for the whole launch: read through the
         * constant address space, as kernel arguments are */
        CP1Args* ap = (CP1Args*)(qa.args + tq);
        p1_sweep<R, NW, BTCK, false, MF, CP1Args>(*ap, tg * NW + wv, lds, botw, qa.bot_ld);
    }
}

__global__ void k_zero_u64(unsigned long long* p, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

/* wts[p][k] = 1 / (0.5 (qbkg_k + bkg_k(p))), the per-pair weights of the
 * score's 20-term sum (the same FP32 operations as every other user) */
__global__ void k_weights(const float* qbkg, const float* bkg, int nprof, float* wts)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nprof) return;
#pragma unroll
    for (int k = 0; k < 20; k++)
        wts[(size_t)p * 20 + k] = 1.0f / (0.5f * (qbkg[k] + bkg[(size_t)k * nprof + p]));
}

__global__ void k_zero_u32(uint32_t* p, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

__global__ void k_fill_u32(uint32_t* p, int64_t n, uint32_t v)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

}  // namespace

/* clears the phase-1 maxima with a kernel (safe inside a captured graph) */
hipError_t c2_launch_zero_u64(unsigned long long* p, int64_t n, hipStream_t stream)
{
    (void)hipGetLastError();   /* the status read after the launches is theirs alone */
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_zero_u64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, p, n);
    return hipGetLastError();
}

hipError_t c2_launch_weights(const float* qbkg, const float* bkg, int nprof, float* wts, hipStream_t stream)
{
    (void)hipGetLastError();   /* the status read after the launches is theirs alone */
    if (nprof <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_weights, dim3((unsigned)((nprof + 255) / 256)), dim3(256), 0, stream, qbkg, bkg, nprof, wts);
    return hipGetLastError();
}

/* clears n 32-bit words with a kernel (safe inside a captured graph) */
hipError_t c2_launch_zero_u32(uint32_t* p, int64_t n, hipStream_t stream)
{
    (void)hipGetLastError();   /* the status read after the launches is theirs alone */
    if (n <= 0) return hipSuccess;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_zero_u32, dim3((unsigned)blocks), dim3(256), 0, stream, p, n);
    return hipGetLastError();
}

hipError_t c2_launch_fill_u32(uint32_t* p, int64_t n, uint32_t v, hipStream_t stream)
{
    (void)hipGetLastError();   /* the status read after the launches is theirs alone */
    if (n <= 0) return hipSuccess;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)blocks), dim3(256), 0, stream, p, n, v);
    return hipGetLastError();
}

size_t c2_phase1_lds_bytes(int R, int NW, int nstripes)
{
    const int ns = nstripes > 1 ? NS_TOP : NS_BASE;
    return sizeof(float) * ((size_t)40 * 64 * R + (size_t)NW * ns * RING);
}

/* the context-vector variant: rows per lane at most 2 (their 21 CV values
 * each stay in registers), at most 4 waves per workgroup (a column slot
 * carries 61 words) */
static size_t phase1_cv_lds_bytes(int R, int NW)
{
    return sizeof(float) * ((size_t)40 * 64 * R + (size_t)NW * NS_CV * RING + CV_MAP);
}

template <int RR, int NWW, bool BT>
static hipError_t launch_cv(const P1Args& a, int grid, size_t lds, hipStream_t stream)
{
    (void)hipGetLastError();   /* the status read after the launches is theirs alone */
    C2_LDS_LIMIT_ONCE((k_phase1<RR, NWW, BT, true>), 160 * 1024);
    hipLaunchKernelGGL((k_phase1<RR, NWW, BT, true>), dim3(grid), dim3(64 * NWW), lds, stream, a);
    return hipGetLastError();
}

#define C2_P1_CASE(RR, NWW, BT, M)                                                         \
    case RR:                                                                                 \
        hipLaunchKernelGGL((k_phase1<RR, NWW, BT, false, M>), dim3(grid), dim3(64 * NWW), lds, stream, a); \
        break;

template <int NWW, bool BT, bool M = false>
static hipError_t launch_nw(const P1Args& a, int R, int grid, size_t lds, hipStream_t stream)
{
    switch (R) {
        case 1: C2_LDS_LIMIT_ONCE((k_phase1<1, NWW, BT, false, M>), 160 * 1024); break;
        case 2: C2_LDS_LIMIT_ONCE((k_phase1<2, NWW, BT, false, M>), 160 * 1024); break;
        case 3: C2_LDS_LIMIT_ONCE((k_phase1<3, NWW, BT, false, M>), 160 * 1024); break;
        case 4: C2_LDS_LIMIT_ONCE((k_phase1<4, NWW, BT, false, M>), 160 * 1024); break;
        case 5: C2_LDS_LIMIT_ONCE((k_phase1<5, NWW, BT, false, M>), 160 * 1024); break;
        default: return hipErrorInvalidValue;
    }
    switch (R) {
        C2_P1_CASE(1, NWW, BT, M)
        C2_P1_CASE(2, NWW, BT, M)
        C2_P1_CASE(3, NWW, BT, M)
        C2_P1_CASE(4, NWW, BT, M)
        C2_P1_CASE(5, NWW, BT, M)
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t c2_launch_phase1(const P1Args& a, int R, int NW, bool btck, hipStream_t stream)
{
    if (a.nseg <= 0) return hipSuccess;
    if (a.cvs.wgt > 0.0f) {
        if (R < 1 || R > 2 || a.cvs.card < 1 || a.cvs.card > CV_MAP) return hipErrorInvalidValue;
        const int nw = btck ? 2 : 4;
        const int grid = (a.nseg + nw - 1) / nw;
        const size_t lds = phase1_cv_lds_bytes(R, nw);
        if (btck) return R == 1 ? launch_cv<1, 2, true>(a, grid, lds, stream) : launch_cv<2, 2, true>(a, grid, lds, stream);
        return R == 1 ? launch_cv<1, 4, false>(a, grid, lds, stream) : launch_cv<2, 4, false>(a, grid, lds, stream);
    }
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
    /* the DB sweep with the 20-term sums on the matrix cores */
    if (a.mfma) {
        switch (NW) {
            case 8: return launch_nw<8, false, true>(a, R, grid, lds, stream);
            case 4: return launch_nw<4, false, true>(a, R, grid, lds, stream);
            default: break;
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

static constexpr size_t Q_MAX_DYN_LDS = 160 * 1024 - 256;

/* the queue kernels built: the DB sweep with 8 or 4 waves per workgroup, the
 * candidates' backtrack pass with 2 */
static const void* phase1q_fn(int R, int NW, bool btck, bool mf)
{
#define C2_P1Q(RR)                                                                   \
    if (R == RR) {                                                                   \
        if (btck) return NW == 2 ? (const void*)k_phase1q<RR, 2, true> : nullptr;    \
        if (mf && NW == 8) return (const void*)k_phase1q<RR, 8, false, true>;        \
        if (mf && NW == 4) return (const void*)k_phase1q<RR, 4, false, true>;        \
        if (NW == 8) return (const void*)k_phase1q<RR, 8, false>;                    \
        if (NW == 4) return (const void*)k_phase1q<RR, 4, false>;                    \
        return nullptr;                                                              \
    }
    C2_P1Q(1) C2_P1Q(2) C2_P1Q(3) C2_P1Q(4) C2_P1Q(5)
#undef C2_P1Q
    return nullptr;
}

int c2_phase1q_grid(int R, int NW, bool btck, size_t lds, bool mf)
{
    const void* fn = phase1q_fn(R, NW, btck, mf);
    if (!fn || lds > Q_MAX_DYN_LDS) return 0;
    /* the kernel's own pull slot is static LDS: the dynamic limit stays below 160 KiB */
    if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)Q_MAX_DYN_LDS) != hipSuccess) return 0;
    int dev = 0, ncu = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu < 1)
        return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, 64 * NW, lds) != hipSuccess || per < 1) {
        /* the LDS bound (160 KiB per CU): extra workgroups only find the queue empty */
        per = (int)((160 * 1024) / (lds + 256));
        if (per < 1) per = 1;
    }
    return per * ncu;
}

hipError_t c2_launch_phase1q(const P1QueueArgs& q, int R, int NW, bool btck, int grid, size_t lds, hipStream_t stream,
                             bool mf)
{
    if (q.ntiles <= 0) return hipSuccess;
    const void* fn = phase1q_fn(R, NW, btck, mf);
    if (!fn || grid < 1) return hipErrorInvalidValue;
    P1QueueArgs qa = q;
    void* args[] = {&qa};
    return hipLaunchKernel(fn, dim3((unsigned)grid), dim3(64 * NW), args, lds, stream);
}
