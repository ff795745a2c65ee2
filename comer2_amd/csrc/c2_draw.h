/*
 * c2_draw.h -- backtrace and alignment drawing, one wave per pair.
 *
 * Shared by the MAP drawing (FinalizeMAPDP_MAP, cumapdp/CuBatchMAPDP_map_final.cu:168-357)
 * and the phase-1 drawing of the MAPALN=0 path (FinalizeDP_ALN,
 * cudp/CuBatchDP_aln_final.cu:206-345).  Both walk the backtrack codes from
 * the maximum cell (STOP 0, LEFT 1, UP 2, DIAG 3) until STOP or the matrix
 * edge and write the query, middle and target strings: the middle symbol is
 * the residue where the two residues are identical, '+' where the pair's
 * score is positive, ' ' otherwise.
 *
 * Every lane of the wave follows the same path (uniform control flow) and
 * reads the codes from a window staged in LDS by the code source (Codes):
 * the codes of one stripe over a run of consecutive anti-diagonal steps are
 * one contiguous range of the lane-step layouts, read by the wave with
 * 16-byte loads.  A backtrace only moves up and left, so its step index
 * (column + lane) never grows inside a stripe, and the window is reloaded
 * ending at the current step when the path leaves it.  Lane j keeps move j
 * of every 64; after each 64 moves the lanes draw their moves together
 * (residues, middle symbol, counts) into the pair's scratch rows in path
 * order, and the three strings are finally written start first.
 */
#ifndef C2_DRAW_H
#define C2_DRAW_H
#include "c2_score.h"

constexpr int C2_DRAW_WIN = 16384;     /* LDS bytes of a code window */

/* codes of the MAP DP: one byte per cell, lane-step layout
 * [stripe][step][row in lane][lane] */
template <int R>
struct MapCodes {
    const char* bt; int nsteps; char* win; int ws = -1, wt0 = 0;
    static constexpr int WS = C2_DRAW_WIN / (64 * R);
    __device__ int at(int y, int x, int lane) {
        const int s = y / (64 * R), rr = y - s * 64 * R, ln = rr / R, k = rr - ln * R;
        const int t = x + ln;
        if (s != ws || t < wt0 || t >= wt0 + WS) {
            ws = s;
            wt0 = t - (WS - 1); if (wt0 < 0) wt0 = 0;
            const int wt1 = (wt0 + WS < nsteps) ? wt0 + WS : nsteps;
            const char* src = bt + ((int64_t)s * nsteps + wt0) * (R * 64);
            const int nbytes = (wt1 - wt0) * (R * 64);
            for (int i = lane * 16; i < nbytes; i += 64 * 16) *(uint4*)(win + i) = *(const uint4*)(src + i);
            wave_lds_sync();
        }
        return win[((t - wt0) * R + k) * 64 + ln];
    }
};

/* codes of the phase-1 backtrack pass: one 32-bit word per (stripe, step,
 * lane) holding the 2-bit codes of the lane's R rows */
struct P1Codes {
    const uint32_t* bt; int nsteps, R; uint32_t* win; int ws = -1, wt0 = 0;
    static constexpr int WS = C2_DRAW_WIN / (64 * 4);
    __device__ int at(int y, int x, int lane) {
        const int s = y / (64 * R), rr = y - s * 64 * R, ln = rr / R, k = rr - ln * R;
        const int t = x + ln;
        if (s != ws || t < wt0 || t >= wt0 + WS) {
            ws = s;
            wt0 = t - (WS - 1); if (wt0 < 0) wt0 = 0;
            const int wt1 = (wt0 + WS < nsteps) ? wt0 + WS : nsteps;
            const uint32_t* src = bt + ((int64_t)s * nsteps + wt0) * 64;
            const int nwords = (wt1 - wt0) * 64;
            for (int i = lane * 4; i < nwords; i += 64 * 4) *(uint4*)(win + i) = *(const uint4*)(src + i);
            wave_lds_sync();
        }
        return (int)((win[(t - wt0) * 64 + ln] >> (2 * k)) & 3u);
    }
};

/* walks from (x, y) and draws the pair's strings into out (3 rows of
 * stride chars, NUL-terminated) through tmp (same shape); o receives
 * alignment length, identities, positives, gaps and the beginning
 * coordinates (x+1)<<16 | (y+1) in slots 0..4 */
template <class Codes>
__device__ void c2_walk_draw(const ScoreCtx& c, int p, int x, int y, Codes& codes,
                             char* out, char* tmp, int stride, int32_t* o)
{
    const int lane = threadIdx.x & 63;
    const int64_t col0 = c.dist[p];
    float w[20];
    c2_pair_weights(c, p, w);
    int n = 0, idts = 0, psts = 0, gaps = 0;
    int myb = 0, myy = 0, myx = 0;
    /* lanes 0..cnt-1 draw moves base..base+cnt-1 */
    auto draw = [&](int base, int cnt) {
        if (lane < cnt) {
            char qa, ta, mid;
            if (myb == 2) { gaps++; qa = c.qaa[myy]; ta = '-'; mid = ' '; }
            else if (myb == 1) { gaps++; qa = '-'; ta = c.aa[col0 + myx]; mid = ' '; }
            else {
                qa = c.qaa[myy]; ta = c.aa[col0 + myx];
                float ss;
                if (qa == ta) { mid = qa; idts++; }
                else if (c2_score_cell_w(c, myy, col0 + myx, w, &ss) > 0.0f) { mid = '+'; psts++; }
                else mid = ' ';
            }
            const int i = base + lane;
            tmp[i] = qa; tmp[stride + i] = mid; tmp[2 * (size_t)stride + i] = ta;
        }
    };
    while (x >= 0 && y >= 0) {
        const int b = codes.at(y, x, lane);
        if (b == 0) break;
        if ((n & 63) == lane) { myb = b; myy = y; myx = x; }
        n++;
        if (b == 2) y--;
        else if (b == 1) x--;
        else { x--; y--; }
        if ((n & 63) == 0) draw(n - 64, 64);
    }
    if (n & 63) draw(n & ~63, n & 63);
    __threadfence_block();
    __syncthreads();
    for (int k = lane; k < n; k += 64) {
        out[k] = tmp[n - 1 - k];
        out[stride + k] = tmp[stride + n - 1 - k];
        out[2 * (size_t)stride + k] = tmp[2 * (size_t)stride + n - 1 - k];
    }
    for (int off = 32; off >= 1; off >>= 1) {
        idts += __shfl_xor(idts, off); psts += __shfl_xor(psts, off); gaps += __shfl_xor(gaps, off);
    }
    if (lane == 0) {
        out[n] = out[stride + n] = out[2 * (size_t)stride + n] = 0;
        o[0] = n; o[1] = idts; o[2] = psts; o[3] = gaps;
        o[4] = (int32_t)(((uint32_t)(x + 1) << 16) | (uint32_t)(y + 1));
    }
}

#endif
