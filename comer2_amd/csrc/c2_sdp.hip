/*
 * c2_sdp.hip -- differentiable soft DP for gfx950: log-sum-exp forward and
 * its hand-written adjoint.  No reference counterpart.
 *
 * Model: the MAP forward model of the reference (ExecMAPDP_Fwd_Unroll32x,
 * cumapdp/CuBatchMAPDP_fwd.cu:291-372) with exact log-sum-exp: per cell the
 * states MM MI IM DN ND,
 *   MM = log(1 + e^{MI'+t+c+s} + e^{IM'+t+c+s} + e^{DN'+t+c+s} + e^{ND'+t+c+s} + e^{MM'+t+s})
 *   MI = lse(MM^ + t + o_I, MI^ + t + e_I)      IM = lse(MM< + t + o_I, IM< + t + e_I)
 *   DN = lse(MM^ + t + o_D, DN^ + t + e_D)      ND = lse(MM< + t + o_D, ND< + t + e_D)
 * (' diagonal, ^ up, < left; t the profile transitions; theta = (o_I, e_I,
 * o_D, e_D, c) five global log-offsets), row -1 = (MM 0, others -99), column
 * -1 = the MAP pass's boundary vectors, log Z = log sum_ij e^{MM(i,j)}.
 * Adjoint: the outside recursion B over the same terms in reverse;
 *   dlogZ/dS(i,j)   = e^{B_MM(i,j) - Z} (e^{F_MM(i,j)} - 1)
 *   dlogZ/dtheta_k  = sum of e^{F_src + term + B_dst - Z} over class-k terms
 *   posterior MM    = e^{F_MM + B_MM - Z}.
 * Arithmetic is base 2 inside (inputs scaled by log2 e once), with the
 * hardware exp2/log2.
 *
 * Decomposition: persistent workgroups take pairs with a uniform stride
 * over the longest-first order; a pair is cut into stripes of 64*R query
 * rows, one wave per stripe, lane l owning rows R*l..R*l+R-1 and processing
 * column t-l at step t, the row above arriving from lane l-1 by DPP
 * wave_shr:1 (forward) and the row below from lane l+1 by wave_shl:1
 * (adjoint).  The waves of a workgroup take a band of up to 8 stripes that
 * advance in lockstep 64 steps apart (one barrier per step), passing their
 * boundary row through LDS; bands pass it through per-pair HBM buffers.
 * The forward either stores all five states of every cell (K = 0) or only
 * the wave state every K steps (K divides 64); the adjoint then recomputes
 * each K-step segment from its checkpoint into per-wave scratch and sweeps
 * it backwards (sqrt(L) checkpointing).
 */
#include "c2_dev.h"
#include "c2_kernels.h"

namespace {

constexpr float L2E = 1.4426950408889634f;
constexpr float LN2 = 0.6931471805599453f;
constexpr float NEG = -1.0e30f;

__device__ __forceinline__ float ex2(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ __forceinline__ float lg2(float x) { return __builtin_amdgcn_logf(x); }
__device__ __forceinline__ float lse2(float a, float b)
{
    const float m = fmaxf(a, b);
    return m + lg2(1.0f + ex2(fminf(a, b) - m));
}

/* log-sum-exp of two terms as (max, e^{a-max}, e^{b-max}) with one exp2 */
__device__ __forceinline__ float lse_parts(float a, float b, float& fa, float& fb)
{
    const float m = fmaxf(a, b);
    const float e = ex2(fminf(a, b) - m);
    const bool ab = a >= b;
    fa = ab ? 1.0f : e;
    fb = ab ? e : 1.0f;
    return m;
}

struct Th { float oi, ei, od, ed, cl; };

__device__ __forceinline__ S5 s5_fill(float mm, float rest)
{
    S5 s; s.mm = mm; s.mi = s.im = s.dn = s.nd = rest; return s;
}

/* transitions of query/profile row i in the MAP forward caching (MI, II at i+1), base 2 */
__device__ __forceinline__ void ld_trn(const float* T, int64_t ntr, int64_t base, int i, float t[7])
{
    t[0] = T[0 * ntr + base + i] * L2E;
    t[1] = T[1 * ntr + base + i + 1] * L2E;
    t[2] = T[2 * ntr + base + i] * L2E;
    t[3] = T[3 * ntr + base + i] * L2E;
    t[4] = T[4 * ntr + base + i + 1] * L2E;
    t[5] = T[5 * ntr + base + i] * L2E;
    t[6] = T[6 * ntr + base + i] * L2E;
}

/* Row-major S, dS and posterior blocks are read and written four columns at
 * a time: at anti-diagonal step t lane l holds column t-l of each of its R
 * rows, so the 4 steps of an aligned group (t = 4m..4m+3) cover 4
 * consecutive columns of every row.  A per-wave LDS buffer [R][4][64] holds
 * the group (slot t&3); one 16-byte access per row replaces four scattered
 * 4-byte accesses, and only the lanes at a row's two ends fall back to
 * single words. */
typedef float f4u __attribute__((ext_vector_type(4), aligned(4)));

/* row r = rbase + R*lane + k of a row-major [L1][L2] block starts at
 * (rbase + k)*L2 (wave-uniform) + lane*R*L2 (lo, the lane's part), so the
 * R row addresses share one per-lane register pair */
template <int R>
__device__ __forceinline__ void grp_load(const float* M, int L1, int L2, int rbase, int64_t lo,
                                         int c0, float* buf, int lane)
{
    const bool full = c0 >= 0 && c0 + 3 < L2;
#pragma unroll 1
    for (int k = 0; k < R; k++) {
        if (rbase + R * lane + k >= L1) continue;
        const float* row = M + (int64_t)(rbase + k) * L2 + lo;
        float v0 = 0.0f, v1 = 0.0f, v2 = 0.0f, v3 = 0.0f;
        if (full) {
            const f4u v = *(const f4u*)(row + c0);
            v0 = v.x; v1 = v.y; v2 = v.z; v3 = v.w;
        } else {
            if (c0 >= 0 && c0 < L2) v0 = row[c0];
            if (c0 + 1 >= 0 && c0 + 1 < L2) v1 = row[c0 + 1];
            if (c0 + 2 >= 0 && c0 + 2 < L2) v2 = row[c0 + 2];
            if (c0 + 3 >= 0 && c0 + 3 < L2) v3 = row[c0 + 3];
        }
        float* b = buf + k * 256 + lane;
        b[0] = v0; b[64] = v1; b[128] = v2; b[192] = v3;
    }
}

template <int R>
__device__ __forceinline__ void grp_store(float* M, int L1, int L2, int rbase, int64_t lo,
                                          int c0, const float* buf, int lane)
{
    const bool full = c0 >= 0 && c0 + 3 < L2;
#pragma unroll 1
    for (int k = 0; k < R; k++) {
        if (rbase + R * lane + k >= L1) continue;
        float* row = M + (int64_t)(rbase + k) * L2 + lo;
        const float* b = buf + k * 256 + lane;
        if (full) {
            f4u v;
            v.x = b[0]; v.y = b[64]; v.z = b[128]; v.w = b[192];
            *(f4u*)(row + c0) = v;
        } else {
            if (c0 >= 0 && c0 < L2) row[c0] = b[0];
            if (c0 + 1 >= 0 && c0 + 1 < L2) row[c0 + 1] = b[64];
            if (c0 + 2 >= 0 && c0 + 2 < L2) row[c0 + 2] = b[128];
            if (c0 + 3 >= 0 && c0 + 3 < L2) row[c0 + 3] = b[192];
        }
    }
}

/* LDS offset of the current step's word in a row's group (slot t & 3) */
__device__ __forceinline__ int k_grp_slot(int t, int lane) { return (t & 3) * 64 + lane; }

/* floats of one wave's group buffer */
template <int R>
__device__ __forceinline__ constexpr int grp_floats() { return R * 4 * 64; }

__device__ __forceinline__ S5 fcell(const S5& dg, const S5& up, const S5& lf, const float* qt,
                                    const float* dt, float s, const Th& th)
{
    const float a0 = dg.mi + qt[0] + dt[3] + th.cl + s;
    const float a1 = dg.im + qt[3] + dt[0] + th.cl + s;
    const float a2 = dg.dn + qt[5] + dt[0] + th.cl + s;
    const float a3 = dg.nd + qt[0] + dt[5] + th.cl + s;
    const float a4 = dg.mm + qt[0] + dt[0] + s;
    const float m = fmaxf(fmaxf(fmaxf(a0, a1), fmaxf(a2, a3)), fmaxf(a4, 0.0f));
    const float sum = ex2(a0 - m) + ex2(a1 - m) + ex2(a2 - m) + ex2(a3 - m) + ex2(a4 - m) + ex2(-m);
    S5 n;
    n.mm = m + lg2(sum);
    n.mi = lse2(up.mm + qt[0] + dt[1] + th.oi, up.mi + qt[0] + dt[4] + th.ei);
    n.im = lse2(lf.mm + qt[1] + dt[0] + th.oi, lf.im + qt[4] + dt[0] + th.ei);
    n.dn = lse2(up.mm + qt[2] + th.od, up.dn + qt[6] + th.ed);
    n.nd = lse2(lf.mm + dt[2] + th.od, lf.nd + dt[6] + th.ed);
    return n;
}

__device__ __forceinline__ Th load_theta(const float* theta)
{
    Th t;
    t.oi = theta[0] * L2E; t.ei = theta[1] * L2E; t.od = theta[2] * L2E;
    t.ed = theta[3] * L2E; t.cl = theta[4] * L2E;
    return t;
}

/* wave state of the forward sweep of one stripe */
template <int R>
struct FwdState {
    S5 H[R];      /* row states at the lane's current column */
    S5 U, D;      /* up / diagonal neighbour of the lane's first row */
};

/* per-stripe constants of a pair */
template <int R>
struct Stripe {
    int rbase;
    int r[R];
    bool valid[R];
    float qt[R][7];
};

template <int R>
__device__ __forceinline__ void load_stripe(const SdpArgs& a, const SdpPair& P, int s, int lane, Stripe<R>& st)
{
    st.rbase = s * 64 * R;
    const int64_t qb = (int64_t)a.qdist[P.q] + P.q;
#pragma unroll
    for (int k = 0; k < R; k++) {
        st.r[k] = st.rbase + R * lane + k;
        st.valid[k] = st.r[k] < P.L1;
        if (st.valid[k]) ld_trn(a.qtr, a.qntr, qb, st.r[k], st.qt[k]);
        else for (int j = 0; j < 7; j++) st.qt[k][j] = 0.0f;
    }
}

/* top boundary of stripe s at column c (row rbase-1 as seen by the stripe) */
__device__ __forceinline__ S5 top_at(const SdpArgs& a, const SdpPair& P, const float* fh, int s, int c)
{
    const float lp0 = -99.0f * L2E;
    if (s == 0 || c < 0 || c >= P.L2) return s5_fill(0.0f, lp0);
    const float* h = fh + (size_t)(s - 1) * 5 * P.L2;
    S5 v;
    v.mm = h[0 * P.L2 + c]; v.mi = h[1 * P.L2 + c]; v.im = h[2 * P.L2 + c];
    v.dn = h[3 * P.L2 + c]; v.nd = h[4 * P.L2 + c];
    return v;
}

/* the cells of forward step t of stripe s (active lanes only), Z
 * accumulated; lane 63 writes its last row to the HBM hand-off (fhw, when a
 * stripe follows) and to an LDS slot (lds, when the next stripe runs in the
 * same band) */
template <int R>
__device__ __forceinline__ void fwd_cells(const SdpArgs& a, const SdpPair& P, const Stripe<R>& sp,
                                          FwdState<R>& w, int s, int ns, int t, int lane, const Th& th,
                                          const float* sg, float* fhw, float* lds,
                                          float& zm, float& zs, bool accz)
{
    const int c = t - lane;
    const bool act = (c >= 0) && (c < P.L2);
    if (act) {
        float dt[7];
        ld_trn(a.dtr, a.dntr, (int64_t)a.ddist[P.p] + P.p, c, dt);
        S5 prevk = w.D;   /* diagonal of row k: old state of row k-1 (lane's D for k = 0) */
#pragma unroll
        for (int k = 0; k < R; k++) {
            S5 lf = w.H[k], dg = prevk;
            if (c == 0) {
                const float* bl = a.qll + ((int64_t)a.qdist[P.q] + (sp.valid[k] ? sp.r[k] : 0)) * 5;
                const float* bd = a.qld + ((int64_t)a.qdist[P.q] + (sp.valid[k] ? sp.r[k] : 0)) * 5;
                lf.mm = bl[0] * L2E; lf.mi = bl[1] * L2E; lf.im = bl[2] * L2E; lf.dn = bl[3] * L2E; lf.nd = bl[4] * L2E;
                dg.mm = bd[0] * L2E; dg.mi = bd[1] * L2E; dg.im = bd[2] * L2E; dg.dn = bd[3] * L2E; dg.nd = bd[4] * L2E;
            }
            const S5 up = (k == 0) ? w.U : w.H[k - 1];
            prevk = w.H[k];
            const float sc = sp.valid[k] ? sg[k * 256 + (t & 3) * 64 + lane] * L2E : 0.0f;
            const S5 n = fcell(dg, up, lf, sp.qt[k], dt, sc, th);
            w.H[k] = n;
            if (accz && sp.valid[k]) {
                const float f = n.mm;
                if (f > zm) { zs = zs * ex2(zm - f) + 1.0f; zm = f; }
                else zs += ex2(f - zm);
            }
        }
        if (lane == 63) {
            const S5& e = w.H[R - 1];
            if (s + 1 < ns && fhw) {
                float* h = fhw + (size_t)s * 5 * P.L2;
                h[0 * P.L2 + c] = e.mm; h[1 * P.L2 + c] = e.mi; h[2 * P.L2 + c] = e.im;
                h[3 * P.L2 + c] = e.dn; h[4 * P.L2 + c] = e.nd;
            }
            if (lds) { lds[0] = e.mm; lds[1] = e.mi; lds[2] = e.im; lds[3] = e.dn; lds[4] = e.nd; }
        }
    }
}

/* the row above for the next step: lane l takes lane l-1's last row, lane 0
 * the stripe's top boundary tp */
template <int R>
__device__ __forceinline__ void fwd_shift(FwdState<R>& w, const S5& tp)
{
    w.D = w.U;
    w.U.mm = wave_shr1(w.H[R - 1].mm, tp.mm);
    w.U.mi = wave_shr1(w.H[R - 1].mi, tp.mi);
    w.U.im = wave_shr1(w.H[R - 1].im, tp.im);
    w.U.dn = wave_shr1(w.H[R - 1].dn, tp.dn);
    w.U.nd = wave_shr1(w.H[R - 1].nd, tp.nd);
}

/* one whole forward step with the top boundary from the HBM hand-off (the
 * adjoint's recomputation, after the forward pass wrote every stripe's) */
template <int R>
__device__ __forceinline__ void fwd_step(const SdpArgs& a, const SdpPair& P, const Stripe<R>& sp,
                                         FwdState<R>& w, int s, int ns, int t, int lane, const Th& th,
                                         const float* S, const float* fh, float* sg)
{
    float zm = 0.0f, zs = 0.0f;
    if ((t & 3) == 0) grp_load<R>(S, P.L1, P.L2, sp.rbase, (int64_t)lane * R * P.L2, t - lane, sg, lane);
    fwd_cells<R>(a, P, sp, w, s, ns, t, lane, th, sg, nullptr, nullptr, zm, zs, false);
    fwd_shift<R>(w, top_at(a, P, fh, s, t + 1));
}

template <int R>
__device__ __forceinline__ void fwd_init(const SdpArgs& a, const SdpPair& P, const float* fh, int s, int lane,
                                         FwdState<R>& w)
{
#pragma unroll
    for (int k = 0; k < R; k++) w.H[k] = s5_fill(NEG, NEG);
    w.D = s5_fill(NEG, NEG);
    w.U = (lane == 0) ? top_at(a, P, fh, s, 0) : s5_fill(NEG, NEG);
}

/* store / load of the five states of the lane's R cells of one step */
template <int R>
__device__ __forceinline__ void put_cells(float* base, int lane, const S5* H)
{
#pragma unroll
    for (int k = 0; k < R; k++) {
        base[(0 * R + k) * 64 + lane] = H[k].mm; base[(1 * R + k) * 64 + lane] = H[k].mi;
        base[(2 * R + k) * 64 + lane] = H[k].im; base[(3 * R + k) * 64 + lane] = H[k].dn;
        base[(4 * R + k) * 64 + lane] = H[k].nd;
    }
}
template <int R>
__device__ __forceinline__ void get_cell(const float* base, int lane, int k, S5& v)
{
    v.mm = base[(0 * R + k) * 64 + lane]; v.mi = base[(1 * R + k) * 64 + lane];
    v.im = base[(2 * R + k) * 64 + lane]; v.dn = base[(3 * R + k) * 64 + lane];
    v.nd = base[(4 * R + k) * 64 + lane];
}
/* checkpoint of the wave state: H[R], U, D (5 floats each) */
template <int R>
__device__ __forceinline__ void put_ckpt(float* base, int lane, const FwdState<R>& w)
{
    put_cells<R>(base, lane, w.H);
    float* e = base + 5 * R * 64;
    e[0 * 64 + lane] = w.U.mm; e[1 * 64 + lane] = w.U.mi; e[2 * 64 + lane] = w.U.im;
    e[3 * 64 + lane] = w.U.dn; e[4 * 64 + lane] = w.U.nd;
    e[5 * 64 + lane] = w.D.mm; e[6 * 64 + lane] = w.D.mi; e[7 * 64 + lane] = w.D.im;
    e[8 * 64 + lane] = w.D.dn; e[9 * 64 + lane] = w.D.nd;
}
template <int R>
__device__ __forceinline__ void get_ckpt(const float* base, int lane, FwdState<R>& w)
{
#pragma unroll
    for (int k = 0; k < R; k++) get_cell<R>(base, lane, k, w.H[k]);
    const float* e = base + 5 * R * 64;
    w.U.mm = e[0 * 64 + lane]; w.U.mi = e[1 * 64 + lane]; w.U.im = e[2 * 64 + lane];
    w.U.dn = e[3 * 64 + lane]; w.U.nd = e[4 * 64 + lane];
    w.D.mm = e[5 * 64 + lane]; w.D.mi = e[6 * 64 + lane]; w.D.im = e[7 * 64 + lane];
    w.D.dn = e[8 * 64 + lane]; w.D.nd = e[9 * 64 + lane];
}

template <int R>
__device__ __forceinline__ int64_t ckpt_floats() { return (int64_t)(5 * R + 10) * 64; }

/* forward: one workgroup per pair (persistent over the longest-first
 * order), one wave per stripe of a band of NW stripes advancing in lockstep
 * 64 steps apart; wave w-1's last row crosses to wave w in a double-buffered
 * LDS slot, every stripe's last row also goes to HBM (band boundaries and
 * the adjoint's recomputation read it there) */
template <int R>
__global__ __launch_bounds__(512) void k_sdp_fwd(SdpArgs a)
{
    __shared__ float hx[2][8][5];
    __shared__ float wz[8][2];
    extern __shared__ float sdp_grp[];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, NW = blockDim.x >> 6;
    float* sg = sdp_grp + wv * grp_floats<R>();       /* this wave's S group */
    const Th th = load_theta(a.theta);
    const float lp0 = -99.0f * L2E;
    for (int slot = blockIdx.x; slot < a.npairs; slot += gridDim.x) {
        const int pi = a.order[slot];
        const SdpPair P = a.pairs[pi];
        const int ns = P.ns, nsteps = P.L2 + 63;
        const float* S = a.S + P.soff;
        float* store = a.ws + P.foff;
        float* fh = a.ws + P.hoff;
        float zm = NEG, zs = 0.0f;
        const int nb = (ns + NW - 1) / NW;
        for (int b = 0; b < nb; b++) {
            const int s = b * NW + wv;
            const bool have = s < ns;
            const int nwb = min(NW, ns - b * NW);
            Stripe<R> sp;
            load_stripe<R>(a, P, s, lane, sp);
            FwdState<R> w;
#pragma unroll
            for (int k = 0; k < R; k++) w.H[k] = s5_fill(NEG, NEG);
            w.U = s5_fill(NEG, NEG); w.D = s5_fill(NEG, NEG);
            const int TT = nsteps + 1 + 64 * (nwb - 1);
            for (int T = 0; T < TT; T++) {
                const int t = T - 1 - 64 * wv;       /* this wave's step; -1 only takes the row above */
                if (have && t >= 0 && t < nsteps) {
                    if (a.K > 0 && t > 0 && (t % a.K) == 0)
                        put_ckpt<R>(store + ((int64_t)s * P.ncp + t / a.K) * ckpt_floats<R>(), lane, w);
                    if ((t & 3) == 0) grp_load<R>(S, P.L1, P.L2, sp.rbase, (int64_t)lane * R * P.L2, t - lane, sg, lane);
                    fwd_cells<R>(a, P, sp, w, s, ns, t, lane, th, sg, fh,
                                 (wv + 1 < nwb) ? hx[T & 1][wv] : nullptr, zm, zs, true);
                    if (a.K == 0 && (t - lane) >= 0 && (t - lane) < P.L2)
                        put_cells<R>(store + ((int64_t)s * nsteps + t) * 5 * R * 64, lane, w.H);
                }
                __syncthreads();
                if (have && t >= -1 && t + 1 < nsteps) {
                    S5 tp = s5_fill(0.0f, lp0);
                    const int c = t + 1;
                    if (lane == 0 && s > 0 && c < P.L2) {
                        if (wv > 0) {
                            const float* o = hx[T & 1][wv - 1];
                            tp.mm = o[0]; tp.mi = o[1]; tp.im = o[2]; tp.dn = o[3]; tp.nd = o[4];
                        } else {
                            tp = top_at(a, P, fh, s, c);
                        }
                    }
                    fwd_shift<R>(w, tp);
                }
            }
            __syncthreads();
        }
        /* log Z: deterministic xor-butterfly of the lanes' (max, scaled sum),
         * then the waves in order */
        for (int off = 32; off >= 1; off >>= 1) {
            const float om = __shfl_xor(zm, off), os = __shfl_xor(zs, off);
            const float m = fmaxf(zm, om);
            zs = zs * ex2(zm - m) + os * ex2(om - m);
            zm = m;
        }
        if (lane == 0) { wz[wv][0] = zm; wz[wv][1] = zs; }
        __syncthreads();
        if (threadIdx.x == 0) {
            float m = wz[0][0], t = wz[0][1];
            for (int k = 1; k < NW; k++) {
                const float om = wz[k][0], os = wz[k][1];
                const float mm = fmaxf(m, om);
                t = t * ex2(m - mm) + os * ex2(om - mm);
                m = mm;
            }
            a.logZ[pi] = (m + lg2(t)) * LN2;
        }
        __syncthreads();
    }
}

/* outside-recursion values a lane hands to the row above: B_MM + S, B_MI, B_DN */
struct Up3 { float bs, mi, dn; };

/* adjoint: the forward's workgroup and band layout, bands bottom-up; inside
 * a band wave w runs 64 steps behind wave w+1 in the reverse sweep, which
 * hands its first row (B_MM+S, B_MI, B_DN) up through a double-buffered LDS
 * slot.  The sweep goes in phases of KK steps (KK = K, a divisor of 64, or
 * 64 when every cell is stored): with checkpoints, each wave first
 * recomputes its phase's forward states from the checkpoint into its own
 * scratch, then the band takes KK reverse steps in lockstep. */
template <int R>
__global__ __launch_bounds__(512) void k_sdp_adj(SdpArgs a)
{
    __shared__ float hb[2][8][3];
    __shared__ double wacc[8][5];
    extern __shared__ float sdp_grp[];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, NW = blockDim.x >> 6;
    float* sg = sdp_grp + (3 * wv + 0) * grp_floats<R>();   /* S group */
    float* dg = sdp_grp + (3 * wv + 1) * grp_floats<R>();   /* dS group */
    float* pg = sdp_grp + (3 * wv + 2) * grp_floats<R>();   /* posterior group */
    const int gw = blockIdx.x * NW + wv;
    const Th th = load_theta(a.theta);
    float* scr = a.ws + a.scr_off + (int64_t)gw * a.K * 5 * R * 64;
    const int KK = a.K > 0 ? a.K : 64;
    const int lag = 64 / KK;
    for (int slot = blockIdx.x; slot < a.npairs; slot += gridDim.x) {
        const int pi = a.order[slot];
        const SdpPair P = a.pairs[pi];
        const int ns = P.ns, nsteps = P.L2 + 63, L2 = P.L2;
        const float* S = a.S + P.soff;
        const float* store = a.ws + P.foff;
        float* fh = a.ws + P.hoff;
        float* bh = fh + (size_t)ns * 5 * L2;      /* [3][L2]: B_MM+S, B_MI, B_DN of a band's first row */
        float* dS = a.dS + P.soff;
        float* post = a.post ? a.post + P.soff : nullptr;
        const float Z = a.logZ[pi] * L2E;
        const float g = a.gZ ? a.gZ[pi] : 1.0f;
        const int64_t db = (int64_t)a.ddist[P.p] + P.p;
        const int64_t qb = (int64_t)a.qdist[P.q] + P.q;
        double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0, acc3 = 0.0, acc4 = 0.0;
        const int nb = (ns + NW - 1) / NW;
        const int nsegp = nsteps / KK + 1;         /* phases over steps 0..nsteps */
        for (int b = nb - 1; b >= 0; b--) {
            const int s = b * NW + wv;
            const bool have = s < ns;
            const int nwb = min(NW, ns - b * NW);
            Stripe<R> sp;
            load_stripe<R>(a, P, s, lane, sp);
            const int64_t lo = (int64_t)lane * R * L2;      /* the lane's part of its row offsets */
            float qn[7];                             /* transitions of the row below the lane's last row */
            const int rn = sp.rbase + R * (lane + 1);
            if (rn < P.L1) ld_trn(a.qtr, a.qntr, qb, rn, qn);
            else for (int j = 0; j < 7; j++) qn[j] = 0.0f;
            const bool below = (s + 1 < ns);
            S5 cur[R];
#pragma unroll
            for (int k = 0; k < R; k++) cur[k] = s5_fill(NEG, NEG);
            Up3 sh, shp;
            sh.bs = sh.mi = sh.dn = NEG; shp = sh;
            float dx[7];                             /* transitions of column c+1 */
            for (int j = 0; j < 7; j++) dx[j] = 0.0f;
            const int nph = nsegp + lag * (nwb - 1);
            for (int ph = 0; ph < nph; ph++) {
                const int gseg = nsegp - 1 - ph + lag * (nwb - 1 - wv);   /* this wave's phase */
                const bool segok = have && gseg >= 0 && gseg < nsegp;
                const int t0 = gseg * KK;
                if (a.K > 0 && segok && t0 < nsteps) {
                    /* recompute the phase's forward states from its checkpoint */
                    const int t1 = min(nsteps, t0 + KK);
                    FwdState<R> w;
                    if (gseg == 0) {
#pragma unroll
                        for (int k = 0; k < R; k++) w.H[k] = s5_fill(NEG, NEG);
                        w.D = s5_fill(NEG, NEG);
                        w.U = (lane == 0) ? top_at(a, P, fh, s, 0) : s5_fill(NEG, NEG);
                    } else {
                        get_ckpt<R>(store + ((int64_t)s * P.ncp + gseg) * ckpt_floats<R>(), lane, w);
                    }
                    for (int t = t0; t < t1; t++) {
                        fwd_step<R>(a, P, sp, w, s, ns, t, lane, th, S, fh, sg);
                        if ((t - lane) >= 0 && (t - lane) < L2)
                            put_cells<R>(scr + (int64_t)(t - t0) * 5 * R * 64, lane, w.H);
                    }
                    __threadfence_block();
                }
                /* this phase's parameter-gradient sums in FP32 (at most KK*R terms per
                 * lane), added to the pair's FP64 sums after the phase */
                float f0 = 0.0f, f1 = 0.0f, f2 = 0.0f, f3 = 0.0f, f4 = 0.0f;
                for (int j = 0; j < KK; j++) {
                    const int T = ph * KK + j;
                    const int t = t0 + KK - 1 - j;
                    if (segok && t < nsteps) {
                        const int c = t - lane;
                        const bool act = (c >= 0) && (c < L2);
                        /* the reverse sweep enters a group of 4 steps at its top step */
                        if ((t & 3) == 3 || t == nsteps - 1)
                            grp_load<R>(S, P.L1, L2, sp.rbase, lo, (t & ~3) - lane, sg, lane);
                        const int gs = k_grp_slot(t, lane);
                        float dc[7];
                        if (act) ld_trn(a.dtr, a.dntr, db, c, dc);
                        else for (int jj = 0; jj < 7; jj++) dc[jj] = 0.0f;
                        const float* fb = (a.K > 0) ? scr + (int64_t)(t - t0) * 5 * R * 64
                                                    : store + ((int64_t)s * nsteps + t) * 5 * R * 64;
                        if (act) {
                            S5 nxt = s5_fill(NEG, NEG);   /* new values of row k+1 (current column) */
                            float oldbs = shp.bs;         /* B_MM+S of row k+1 at column c+1 */
#pragma unroll
                            for (int k = R - 1; k >= 0; k--) {
                                const S5 rt = cur[k];                               /* (r, c+1) */
                                const float dgbs = oldbs;                           /* (r+1, c+1) */
                                oldbs = rt.mm;
                                const float dnmi = (k < R - 1) ? nxt.mi : sh.mi;    /* (r+1, c) */
                                const float dndn = (k < R - 1) ? nxt.dn : sh.dn;
                                const float* qs = sp.qt[k];
                                const float* qr = (k < R - 1) ? sp.qt[k + 1] : qn;  /* row r+1 */
                                S5 nb = s5_fill(NEG, NEG);
                                if (sp.valid[k]) {
                                    /* MM: end, MM->MM, MM->MI, MM->IM, MM->DN, MM->ND */
                                    const float m1 = dgbs + qr[0] + dx[0];
                                    const float m2 = dnmi + qr[0] + dc[1] + th.oi;
                                    const float m3 = rt.im + qs[1] + dx[0] + th.oi;
                                    const float m4 = dndn + qr[2] + th.od;
                                    const float m5 = rt.nd + dx[2] + th.od;
                                    const float mm = fmaxf(fmaxf(fmaxf(m1, m2), fmaxf(m3, m4)), fmaxf(m5, 0.0f));
                                    const float e2 = ex2(m2 - mm), e3 = ex2(m3 - mm), e4 = ex2(m4 - mm), e5 = ex2(m5 - mm);
                                    const float sum6 = ex2(-mm) + ex2(m1 - mm) + e2 + e3 + e4 + e5;
                                    const float bmm = mm + lg2(sum6);
                                    /* gap states: two terms each, the larger one is e^0 = 1 */
                                    float fi1, fi2, fj1, fj2, fd1, fd2, fn1, fn2;
                                    /* MI: close, extend */
                                    const float i1 = dgbs + qr[0] + dx[3] + th.cl, i2 = dnmi + qr[0] + dc[4] + th.ei;
                                    const float mi = lse_parts(i1, i2, fi1, fi2);
                                    const float bmi = mi + lg2(fi1 + fi2);
                                    /* IM */
                                    const float j1 = dgbs + qr[3] + dx[0] + th.cl, j2 = rt.im + qs[4] + dx[0] + th.ei;
                                    const float mj = lse_parts(j1, j2, fj1, fj2);
                                    const float bim = mj + lg2(fj1 + fj2);
                                    /* DN */
                                    const float d1 = dgbs + qr[5] + dx[0] + th.cl, d2 = dndn + qr[6] + th.ed;
                                    const float md = lse_parts(d1, d2, fd1, fd2);
                                    const float bdn = md + lg2(fd1 + fd2);
                                    /* ND */
                                    const float n1 = dgbs + qr[0] + dx[5] + th.cl, n2 = rt.nd + dx[6] + th.ed;
                                    const float mn = lse_parts(n1, n2, fn1, fn2);
                                    const float bnd = mn + lg2(fn1 + fn2);
                                    S5 f;
                                    get_cell<R>(fb, lane, k, f);
                                    const float wz = ex2(bmm - Z);
                                    /* e^{F_MM + B_MM - Z} = e^{F_MM + mm - Z} * sum6 */
                                    const float wm = ex2(f.mm + mm - Z), wi = ex2(f.mi + mi - Z), wj = ex2(f.im + mj - Z);
                                    const float pmm = wm * sum6;
                                    dg[k * 256 + gs] = (pmm - wz) * g;
                                    if (post) pg[k * 256 + gs] = pmm;
                                    const float wd = ex2(f.dn + md - Z), wn = ex2(f.nd + mn - Z);
                                    f0 = fmaf(wm, e2 + e3, f0);
                                    f2 = fmaf(wm, e4 + e5, f2);
                                    f1 = fmaf(wi, fi2, fmaf(wj, fj2, f1));
                                    f3 = fmaf(wd, fd2, fmaf(wn, fn2, f3));
                                    f4 = fmaf(wi, fi1, fmaf(wj, fj1, fmaf(wd, fd1, fmaf(wn, fn1, f4))));
                                    const float sc = sg[k * 256 + gs] * L2E;
                                    if (c == 0 || sp.r[k] == 0) {
                                        /* transitions whose source is the row -1 / column -1
                                         * boundary: counted here at their destination */
                                        const float lp0 = -99.0f * L2E;
                                        const S5 up0 = s5_fill(0.0f, lp0);
                                        S5 lf0 = up0, dg0 = up0;
                                        if (c == 0) {
                                            const int64_t bi = ((int64_t)a.qdist[P.q] + sp.r[k]) * 5;
                                            lf0.mm = a.qll[bi] * L2E; lf0.mi = a.qll[bi + 1] * L2E; lf0.im = a.qll[bi + 2] * L2E;
                                            lf0.dn = a.qll[bi + 3] * L2E; lf0.nd = a.qll[bi + 4] * L2E;
                                            dg0.mm = a.qld[bi] * L2E; dg0.mi = a.qld[bi + 1] * L2E; dg0.im = a.qld[bi + 2] * L2E;
                                            dg0.dn = a.qld[bi + 3] * L2E; dg0.nd = a.qld[bi + 4] * L2E;
                                            f0 += ex2(lf0.mm + qs[1] + dc[0] + th.oi + bim - Z);
                                            f1 += ex2(lf0.im + qs[4] + dc[0] + th.ei + bim - Z);
                                            f2 += ex2(lf0.mm + dc[2] + th.od + bnd - Z);
                                            f3 += ex2(lf0.nd + dc[6] + th.ed + bnd - Z);
                                        }
                                        if (sp.r[k] == 0) {
                                            f0 += ex2(up0.mm + qs[0] + dc[1] + th.oi + bmi - Z);
                                            f1 += ex2(up0.mi + qs[0] + dc[4] + th.ei + bmi - Z);
                                            f2 += ex2(up0.mm + qs[2] + th.od + bdn - Z);
                                            f3 += ex2(up0.dn + qs[6] + th.ed + bdn - Z);
                                        }
                                        f4 += (ex2(dg0.mi + qs[0] + dc[3] + th.cl + sc + bmm - Z) +
                                                         ex2(dg0.im + qs[3] + dc[0] + th.cl + sc + bmm - Z) +
                                                         ex2(dg0.dn + qs[5] + dc[0] + th.cl + sc + bmm - Z) +
                                                         ex2(dg0.nd + qs[0] + dc[5] + th.cl + sc + bmm - Z));
                                    }
                                    nb.mm = bmm + sc;
                                    nb.mi = bmi; nb.im = bim; nb.dn = bdn; nb.nd = bnd;
                                }
                                cur[k] = nb;
                                nxt = nb;
                            }
                            if (lane == 0) {
                                if (wv > 0) {
                                    float* o = hb[T & 1][wv];
                                    o[0] = cur[0].mm; o[1] = cur[0].mi; o[2] = cur[0].dn;
                                } else if (s > 0) {
                                    bh[0 * L2 + c] = cur[0].mm; bh[1 * L2 + c] = cur[0].mi; bh[2 * L2 + c] = cur[0].dn;
                                }
                            }
                        } else {
#pragma unroll
                            for (int k = 0; k < R; k++) cur[k] = s5_fill(NEG, NEG);
                        }
                        for (int jj = 0; jj < 7; jj++) dx[jj] = dc[jj];
                        /* the group's last step: its four columns go out */
                        if ((t & 3) == 0) {
                            grp_store<R>(dS, P.L1, L2, sp.rbase, lo, t - lane, dg, lane);
                            if (post) grp_store<R>(post, P.L1, L2, sp.rbase, lo, t - lane, pg, lane);
                        }
                    }
                    __syncthreads();
                    if (segok && t >= 1 && t <= nsteps) {
                        /* the row below for step t-1: lane l+1's first row, or for
                         * lane 63 the stripe below at column t-1-63 */
                        Up3 ob; ob.bs = ob.mi = ob.dn = NEG;
                        const int cb = t - 1 - 63;
                        if (lane == 63 && below && cb >= 0 && cb < L2) {
                            if (wv + 1 < nwb) {
                                const float* o = hb[T & 1][wv + 1];
                                ob.bs = o[0]; ob.mi = o[1]; ob.dn = o[2];
                            } else {
                                ob.bs = bh[0 * L2 + cb]; ob.mi = bh[1 * L2 + cb]; ob.dn = bh[2 * L2 + cb];
                            }
                        }
                        shp = sh;
                        sh.bs = wave_shl1(cur[0].mm, ob.bs);
                        sh.mi = wave_shl1(cur[0].mi, ob.mi);
                        sh.dn = wave_shl1(cur[0].dn, ob.dn);
                    }
                }
                acc0 += (double)f0; acc1 += (double)f1; acc2 += (double)f2;
                acc3 += (double)f3; acc4 += (double)f4;
            }
            __syncthreads();
        }
        /* dtheta: fixed-order xor-butterfly in double, then the waves in order */
        for (int off = 32; off >= 1; off >>= 1) {
            acc0 += __shfl_xor(acc0, off); acc1 += __shfl_xor(acc1, off); acc2 += __shfl_xor(acc2, off);
            acc3 += __shfl_xor(acc3, off); acc4 += __shfl_xor(acc4, off);
        }
        if (lane == 0) {
            wacc[wv][0] = acc0; wacc[wv][1] = acc1; wacc[wv][2] = acc2; wacc[wv][3] = acc3; wacc[wv][4] = acc4;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double t0 = 0.0, t1 = 0.0, t2 = 0.0, t3 = 0.0, t4 = 0.0;
            for (int k = 0; k < NW; k++) {
                t0 += wacc[k][0]; t1 += wacc[k][1]; t2 += wacc[k][2]; t3 += wacc[k][3]; t4 += wacc[k][4];
            }
            float* d = a.dth + (size_t)pi * 5;
            d[0] = (float)(t0 * g); d[1] = (float)(t1 * g); d[2] = (float)(t2 * g);
            d[3] = (float)(t3 * g); d[4] = (float)(t4 * g);
        }
        __syncthreads();
    }
}

/* score matrices of the prepared pairs from the raw profile fields:
 * S = log(sum_a q_a d_a / (0.5 (bq_a + bd_a))) - shift
 *     + ssswgt (sum_st p_s p'_t T[10s+int(9.999 p_s)][10t+int(9.999 p'_t)]
 *               + (p_H+0.1)(p'_H+0.1) 0.1111111 - 0.05)
 * (CalcSM_Init_SMEMUnroll2x + CalcSM_SSSS_SMEMUnroll2x); one workgroup per pair */
__global__ __launch_bounds__(256) void k_sdp_scores(SdpScoreArgs a)
{
    __shared__ float T[900];
    __shared__ float w[20];
    for (int i = threadIdx.x; i < 900; i += blockDim.x) T[i] = a.ssss[i];
    const SdpPair P = a.pairs[blockIdx.x];
    if (threadIdx.x < 20)
        w[threadIdx.x] = 1.0f / (0.5f * (a.qbkg[(size_t)threadIdx.x * a.qnprof + P.q] +
                                         a.dbkg[(size_t)threadIdx.x * a.dnprof + P.p]));
    __syncthreads();
    const int64_t q0 = a.qdist[P.q], d0 = a.ddist[P.p];
    const int64_t n = (int64_t)P.L1 * P.L2;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
        const int r = (int)(i / P.L2), c = (int)(i - (int64_t)r * P.L2);
        float acc = 0.0f;
        for (int k = 0; k < 20; k++)
            acc = fmaf(a.qtrg[(size_t)k * a.qnpos + q0 + r], a.dtrg[(size_t)k * a.dnpos + d0 + c] * w[k], acc);
        float sc = c2_logf_clamped(acc) - a.shift;
        if (a.ssswgt > 0.0f) {
            float ps[3], pt[3]; int is[3], it[3];
            for (int j = 0; j < 3; j++) {
                ps[j] = a.qss[(size_t)j * a.qnpos + q0 + r]; is[j] = 10 * j + (int)(ps[j] * 9.999f);
                pt[j] = a.dss[(size_t)j * a.dnpos + d0 + c]; it[j] = 10 * j + (int)(pt[j] * 9.999f);
            }
            float ss = 0.0f;
            for (int x = 0; x < 3; x++)
                for (int y = 0; y < 3; y++) ss = fmaf(ps[x] * pt[y], T[is[x] * 30 + it[y]], ss);
            ss = fmaf(ps[2] + 0.1f, (pt[2] + 0.1f) * 0.1111111f, ss);
            sc += (ss - 0.05f) * a.ssswgt;
        }
        a.S[P.soff + i] = sc;
    }
}

}  // namespace

int c2_sdp_rows_per_lane() { return 5; }

int64_t c2_sdp_ckpt_floats(int R) { return (int64_t)(5 * R + 10) * 64; }

hipError_t c2_launch_sdp(const SdpArgs& a, bool adjoint, int groups, int nw, hipStream_t stream)
{
    (void)hipGetLastError();   /* the status read after the launches is theirs alone */
    if (a.npairs <= 0) return hipSuccess;
    if (groups < 1 || nw < 1 || nw > 8) return hipErrorInvalidValue;
    if (a.K < 0 || (a.K > 0 && 64 % a.K != 0)) return hipErrorInvalidValue;
    const size_t grp = sizeof(float) * 4 * 64 * 5;      /* one wave's [R=5][4][64] group buffer */
    static bool attr = false;
    if (!attr) {
        /* up to 8 waves x 3 group buffers (120 KiB) of dynamic LDS */
        hipError_t e = hipFuncSetAttribute((const void*)k_sdp_adj<5>, hipFuncAttributeMaxDynamicSharedMemorySize, 24 * (int)grp);
        if (e == hipSuccess)
            e = hipFuncSetAttribute((const void*)k_sdp_fwd<5>, hipFuncAttributeMaxDynamicSharedMemorySize, 8 * (int)grp);
        if (e != hipSuccess) return e;
        attr = true;
    }
    if (adjoint) hipLaunchKernelGGL((k_sdp_adj<5>), dim3(groups), dim3(64 * nw), 3 * nw * grp, stream, a);
    else hipLaunchKernelGGL((k_sdp_fwd<5>), dim3(groups), dim3(64 * nw), nw * grp, stream, a);
    return hipGetLastError();
}

hipError_t c2_launch_sdp_scores(const SdpScoreArgs& a, int npairs, hipStream_t stream)
{
    (void)hipGetLastError();   /* the status read after the launches is theirs alone */
    if (npairs <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_sdp_scores, dim3(npairs), dim3(256), 0, stream, a);
    return hipGetLastError();
}
