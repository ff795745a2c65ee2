/*
 * c2_noalloc.hip -- the phase-1 maxima without any allocation: every device
 * buffer the call writes besides its inputs (the uploaded chunk and query
 * set) and its output (the caller's host array) lives in one file-scope
 * device array, `scratch`, of C2_NOALLOC_BYTES, which each GPU the library
 * runs on holds until the process ends.
 *
 * Layout per call (256-byte aligned slices): the 64-bit maxima [nprof], the
 * score weights [nprof][20], the stripe hand-off [5][npos] (queries of more
 * than one stripe), and the query's phase-1 tables (target frequencies,
 * transitions, SS tables, p_H, column -1 boundary, background), uploaded
 * from a host preparation on every call.
 *
 * The array's device address is looked up once per device, with the device
 * current, when a handle is created (c2_noalloc_init from cbp_create); host
 * code uses only that address.  Calls on one GPU share the array, so calls
 * that overlap on one GPU (two threads, two handles) race.
 */
#include <mutex>
#include <cstring>
#include "c2_lib.h"

namespace {

constexpr size_t C2_NOALLOC_BYTES = (size_t)64 << 20;
constexpr int C2_NOALLOC_MAXDEV = 64;

static __device__ __align__(256) unsigned char scratch[C2_NOALLOC_BYTES];

std::mutex g_mu;
void* g_ptr[C2_NOALLOC_MAXDEV] = {};
hipError_t g_err[C2_NOALLOC_MAXDEV] = {};
bool g_done[C2_NOALLOC_MAXDEV] = {};

size_t up256(size_t n) { return (n + 255) & ~(size_t)255; }

}  // namespace

/* looks up the scratch array's address on the current device (once per
 * device); returns the lookup's status */
hipError_t c2_noalloc_init(int device)
{
    if (device < 0 || device >= C2_NOALLOC_MAXDEV) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_done[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_err[device] = e;
        g_ptr[device] = (e == hipSuccess) ? p : nullptr;
        g_done[device] = true;
    }
    return g_err[device];
}

extern "C" int cbp_phase1_keys_noalloc(cbp_handle* h, int32_t qidx, uint64_t* keys)
{
    C2ApiScope c2_scope_;
    if (!h || !keys) return CBP_ERR_ARG;
    if (!h->have_opt || !h->have_tables || !h->have_queries || !h->have_chunk)
        return c2_fail(h, CBP_ERR_STATE, "options, tables, queries and chunk must be uploaded first");
    if (qidx < 0 || qidx >= h->queries.v.nprof) return c2_fail(h, CBP_ERR_ARG, "query index out of range");
    /* the scratch layout has no room for the query's context vectors */
    if (h->opt.cvswgt > 0.0f)
        return c2_fail(h, CBP_ERR_STATE, "the allocation-free entry point has no context-vector term (CVSWGT > 0)");
    if (hipSetDevice(h->device) != hipSuccess) return c2_fail(h, CBP_ERR_HIP, "hipSetDevice");
    unsigned char* base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (h->device >= C2_NOALLOC_MAXDEV || !g_done[h->device] || !g_ptr[h->device])
            return c2_fail(h, CBP_ERR_HIP, "the scratch array's device address is unavailable");
        base = (unsigned char*)g_ptr[h->device];
    }
    QueryPrep hp;
    c2_prepare_query(&h->queries.v, qidx, h->ssss.data(), h->rmax, &hp);
    const DevSet& d = h->dchunk;
    const size_t np = (size_t)d.nprof, ns = (size_t)d.npos;
    /* slices of the scratch array */
    size_t off = 0;
    auto take = [&](size_t n) { const size_t o = off; off += up256(n); return o; };
    const size_t o_keys = take(8 * np), o_wts = take(4 * 20 * np);
    const size_t o_bot = hp.nstripes > 1 ? take(4 * 5 * ns) : 0;
    const size_t o_qv = take(4 * hp.qv.size()), o_qtr = take(4 * hp.qtr.size()), o_QT = take(4 * hp.QT.size());
    const size_t o_qh = take(4 * hp.qh.size()), o_ll = take(4 * hp.lleft.size()), o_ld = take(4 * hp.ldiag.size());
    const size_t o_qb = take(4 * hp.qbkg.size());
    if (off > C2_NOALLOC_BYTES)
        return c2_fail(h, CBP_ERR_ARG, "query and chunk need " + std::to_string(off) +
                       " bytes of scratch; the allocation-free entry point holds " +
                       std::to_string(C2_NOALLOC_BYTES));
    hipStream_t st = h->stream;
    struct { size_t o; const std::vector<float>* v; } tabs[] = {
        {o_qv, &hp.qv}, {o_qtr, &hp.qtr}, {o_QT, &hp.QT}, {o_qh, &hp.qh},
        {o_ll, &hp.lleft}, {o_ld, &hp.ldiag}, {o_qb, &hp.qbkg}};
    for (auto& t : tabs)
        if (!t.v->empty() && hipMemcpyAsync(base + t.o, t.v->data(), 4 * t.v->size(), hipMemcpyHostToDevice, st) != hipSuccess)
            return c2_fail(h, CBP_ERR_HIP, "query tables upload");
    P1Args a;
    std::memset(&a, 0, sizeof a);
    a.trgfrq = d.trgfrq.as<float>(); a.ssprb = d.ssprb.as<float>();
    a.trans = d.trans.as<float>(); a.bkg = d.bkg.as<float>();
    a.len = d.len.as<int32_t>(); a.dist = d.dist.as<uint32_t>(); a.addrpro = d.addrpro.as<int32_t>();
    a.npos = d.npos; a.nprof = d.nprof;
    a.qv = (const float*)(base + o_qv); a.qtr = (const float*)(base + o_qtr); a.QT = (const float*)(base + o_QT);
    a.qh = (const float*)(base + o_qh); a.lleft = (const float*)(base + o_ll); a.ldiag = (const float*)(base + o_ld);
    a.qbkg = (const float*)(base + o_qb);
    float* wts = (float*)(base + o_wts);
    a.wts = wts;
    a.Lq = hp.Lq; a.Lpad = hp.Lpad; a.nstripes = hp.nstripes;
    a.shift = h->opt.mapaln ? 0.002f : 0.0003f;
    a.ssswgt = h->opt.ssswgt;
    a.seg_pb = h->dseg_b.as<int32_t>(); a.seg_pe = h->dseg_e.as<int32_t>(); a.nseg = h->nseg;
    a.keys = (unsigned long long*)(base + o_keys);
    a.bot = hp.nstripes > 1 ? (float*)(base + o_bot) : nullptr;
    int nw = h->nw;
    while (nw > 1 && c2_phase1_lds_bytes(hp.R, nw, hp.nstripes) > 160 * 1024) nw >>= 1;
    /* the maxima start at zero (a kernel, not a memset) */
    if (c2_launch_zero_u64(a.keys, (int64_t)np, st) != hipSuccess ||
        c2_launch_weights(a.qbkg, a.bkg, (int)np, wts, st) != hipSuccess ||
        c2_launch_phase1(a, hp.R, nw, false, st) != hipSuccess)
        return c2_fail(h, CBP_ERR_HIP, "phase-1 launch");
    if (hipMemcpyAsync(keys, a.keys, 8 * np, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return c2_fail(h, CBP_ERR_HIP, "maxima copy");
    return CBP_OK;
}
