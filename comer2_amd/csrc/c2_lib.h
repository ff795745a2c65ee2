/*
 * c2_lib.h -- internal state behind the C ABI.
 */
#ifndef C2_LIB_H
#define C2_LIB_H
#include <hip/hip_runtime.h>
#include <string>
#include <vector>
#include <memory>
#include <map>
#include <new>
#include "../../include/comer2_hip.h"
#include "c2_prep.h"
#include "c2_kernels.h"

/* every ABI call starts and ends with the thread's HIP error status
 * cleared.  The launch checks read the thread's last error, so a status left
 * pending by the caller's own HIP work, or by a call that succeeded (seen
 * when the process held two HIP runtimes: this library opened before
 * PyTorch's), must not be taken for a launch failure; and a status the
 * library left behind would be read by the caller's next check */
struct C2ApiScope {
    C2ApiScope() { (void)hipGetLastError(); }
    ~C2ApiScope() { (void)hipGetLastError(); }
    C2ApiScope(const C2ApiScope&) = delete;
    C2ApiScope& operator=(const C2ApiScope&) = delete;
};

/* device buffer; c2_dalloc charges its size to the handle's account
 * (acct), and the release returns it, so the memory budget sees only live
 * buffers */
struct DBuf {
    void* p = nullptr;
    size_t n = 0;
    size_t* acct = nullptr;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    DBuf(DBuf&& o) noexcept : p(o.p), n(o.n), acct(o.acct) { o.p = nullptr; o.n = 0; o.acct = nullptr; }
    DBuf& operator=(DBuf&& o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; acct = o.acct; o.p = nullptr; o.n = 0; o.acct = nullptr; }
        return *this;
    }
    void release() {
        if (p) { (void)hipFree(p); if (acct) *acct -= n; }
        p = nullptr; n = 0; acct = nullptr;
    }
    ~DBuf() { release(); }
    template <class T> T* as() const { return (T*)p; }
};

/* host vectors in pinned memory: the stage buffers that cross to and from
 * the device, so that their copies are asynchronous and the queries of a
 * batch overlap on their streams (a copy to or from pageable memory waits
 * for the stream); allocation failure throws std::bad_alloc, caught at the
 * C ABI */
template <class T>
struct PinAlloc {
    using value_type = T;
    PinAlloc() = default;
    template <class U> PinAlloc(const PinAlloc<U>&) {}
    T* allocate(size_t n) {
        void* p = nullptr;
        if (hipHostMalloc(&p, n * sizeof(T), hipHostMallocDefault) != hipSuccess) throw std::bad_alloc();
        return static_cast<T*>(p);
    }
    void deallocate(T* p, size_t) { (void)hipHostFree(p); }
    template <class U> bool operator==(const PinAlloc<U>&) const { return true; }
    template <class U> bool operator!=(const PinAlloc<U>&) const { return false; }
};
template <class T> using pvec = std::vector<T, PinAlloc<T>>;

/* owned host copy of a profile set */
struct HostSet {
    cbp_profiles v{};
    std::vector<float> bkg, eno, trgfrq, ssprb, trans;
    std::vector<float> cv, cvprior, cvnorm2;     /* empty when the set has no context vectors */
    std::vector<int32_t> len;
    std::vector<uint32_t> dist;
    std::vector<char> aa, sss;
    void copy_from(const cbp_profiles* s);
};

struct DevSet {
    int32_t nprof = 0;
    int64_t npos = 0;
    DBuf bkg, eno, len, dist, trgfrq, ssprb, aa, sss, trans, addrpro;
    /* context vectors [19][npos], priors and squared norms [npos]; has_cv
     * false: none uploaded (filled with the uninformative vector when the
     * CV term is on) */
    DBuf cv, cvp, cvn;
    bool has_cv = false;
};

struct DevQuery {
    QueryPrep hp;
    DBuf qv, qtr, qtrb, QT, qh, lleft, ldiag, qbkg, qaa, mleft, mdiag;
    DBuf qcv;            /* [21][Lpad] context vector, prior, squared norm (CV term on) */
};

/* device and host state of one query's pass through the stage sequence;
 * results stay valid until the next cbp_process* call */
struct QWork {
    int qidx = -1;
    DevQuery* dq = nullptr;
    float scorethld = 0.0f, logevthld = 0.0f;
    hipStream_t st = nullptr;
    int slot = 0;                       /* stream slot (stripe hand-off buffer) */
    DBuf keys;                          /* [nprof] phase-1 maxima (one-query runs) */
    unsigned long long* hkeys = nullptr; /* pinned host copy of keys (one-query runs) */
    size_t hkeys_n = 0;
    pvec<int32_t> cand, pe;
    pvec<int64_t> off;
    DBuf wts;                           /* [nprof][20] score weights of the query (one-query runs) */
    /* the buffers the stages use: the query's own above, or its slice of a
     * batch's work-queue buffers */
    unsigned long long* dk = nullptr;   /* device maxima [nprof] */
    unsigned long long* hk = nullptr;   /* pinned host copy [nprof] */
    float* dw = nullptr;                /* device score weights [nprof][20] */
    DBuf dcand, dpe, doff, btck, fin;
    pvec<FinRec> hfin;
    pvec<int32_t> pass;
    std::vector<int> passfin;
    pvec<StatIn> si;
    DBuf pass_pro, counts, statin, statout;
    /* the score histograms of all candidates were launched on a second stream
     * together with the backtrack pass (a query on its own); counts rows are
     * then candidate indices and the stats stage reads them through passfin */
    bool hist_ahead = false;
    pvec<double> hstat;
    std::vector<int> sel;
    pvec<MapPair> mp;
    pvec<int32_t> morder;
    pvec<int4> msitems;                 /* dynamic-score work items */
    int64_t alnstride = 0;
    DBuf mpairs, dmorder, dmsitems, mF, mB, mSd, mhand, mbt, mstat, malns, malntmp, mres;
    pvec<int32_t> apro, acand;          /* MAPALN=0: drawn pairs' profiles and candidate indices */
    DBuf dapro, dacand;
    pvec<int32_t> hmres;
    int64_t nposits = 0;
    int32_t maxlen = 0;
    uint64_t sspace = 0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    /* results */
    std::vector<float> aln, epa;
    std::vector<int32_t> order;
    char* strs = nullptr;               /* pinned host copy of the alignment strings */
    size_t strs_cap = 0;
    std::vector<int> map;
    std::vector<float> p1save;          /* [nmap][3] phase-1 PSTVS, BEG, END of the drawn rows */
    ~QWork() {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (hkeys) (void)hipHostFree(hkeys);
        if (strs) (void)hipHostFree(strs);
    }
};

/* merged result of one query of a whole-DB search in chunks */
struct StreamOut {
    std::vector<float> aln, epa;
    std::vector<int32_t> order, map_rows;
    std::vector<char> strs;
};

/* device-built stage lists of one query (c2_dpipe.cpp): the lists and
 * records that outlive the query's chain, and its counters in pinned host
 * memory */
struct DlQuery {
    DBuf cnt, aln, epa, order, sel, res, p1save, malns;
    DBuf dargs, tiles, ntiles;          /* the backtrack pass's queue: arguments, tiles, tile count */
    int32_t* hcnt = nullptr;            /* pinned copy of cnt */
    ~DlQuery() { if (hcnt) (void)hipHostFree(hcnt); }
};
/* per stream slot: the pools a query's chain uses while it runs (the
 * slot's queries run one after another on its stream) */
struct DlSlot {
    DBuf tilecnt, tileoff, cand, pe, off, btck, botw, fin, pass_pro, passfin, statin, counts, statout;
    DBuf mpairs, morder, sitems, apro, acand, mF, mB, mSd, mbt, mhand, mstat, malntmp;
};
/* a query's results from the device lists, as the caller sees them */
struct DlResult {
    std::vector<float> aln, epa;
    std::vector<int32_t> order, map;
    std::vector<float> p1save;
    std::vector<char> strs;
};

constexpr int C2_NSTREAMS = 4;
constexpr size_t C2_QCACHE_MAX = 128;
/* one-query calls run the candidates' histograms ahead of the finalize when
 * there are at most this many candidates (641 counts of 4 B each) */
constexpr int C2_HIST_AHEAD_MAX = 16384;

struct cbp_handle {
    int device = 0;
    size_t budget = 0, used = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
    cbp_options opt{};
    bool have_opt = false;
    std::vector<float> ssss;     /* 30x30 */
    std::vector<float> nnw;      /* 93 */
    bool have_tables = false;
    HostSet queries;
    bool have_queries = false;
    /* device query preparations, least recently used evicted beyond
     * C2_QCACHE_MAX entries (never one of the running call's queries) */
    struct QEnt { std::unique_ptr<DevQuery> dq; uint64_t tick = 0; };
    std::map<int, QEnt> qcache;
    uint64_t qtick = 0;          /* bumped once per call that prepares queries */
    HostSet chunk;
    DevSet dchunk;
    bool have_chunk = false;
    std::vector<int32_t> seg_p;
    DBuf dseg_b, dseg_e;
    DBuf dnnw;           /* SS18 parameters on the device */
    /* context-vector map (cbp_upload_cvs2s) and the model's constants */
    std::vector<float> cvkeys, cvvals;
    int32_t cvshift = 0;
    DBuf dcvmap;
    bool have_cvs = false;
    /* per-call scratch */
    DBuf cand, cand_off, btck, fin, pass_pro, counts, statin, statout;
    DBuf mpairs, mF, mSd, mhand, mbt, mstat, malns, malntmp, mres;
    int32_t nseg = 0;
    DBuf keys;           /* [nprof] u64 */
    DBuf bot;            /* stripe hand-off */
    DBuf wts;            /* [nprof][20] score weights of the query being searched */
    uint64_t dblen = 0, nseqs = 0;
    int seg_cols = 4096; /* tuning: target columns per segment */
    int rmax = 5;        /* tuning: max rows per lane */
    int nw = 8;          /* tuning: waves per workgroup */
    double last_ms = 0.0;
    int32_t last_nlaunch = 0;
    /* staged pipeline: per-query work, extra streams (default priority) and
     * one stripe hand-off buffer per stream */
    std::vector<std::unique_ptr<QWork>> work;
    hipStream_t xstream[C2_NSTREAMS - 1] = {nullptr, nullptr, nullptr};
    DBuf sbot[C2_NSTREAMS];
    /* phase-1 stage of a batch captured in a HIP graph and replayed while the
     * batch (queries, chunk, launch geometry) is unchanged; a few batches are
     * kept, keyed by (generation, geometry, query list) */
    struct GraphEnt { hipGraphExec_t exec = nullptr; int seen = 0; };
    bool use_graph = false;
    std::map<std::vector<int32_t>, GraphEnt> graphs;
    /* device work queue of a batch (two queries or more): persistent
     * phase-1 workgroups pull (query, segment group) tiles binned by rows per
     * lane; the candidates' backtrack pass runs the same way */
    struct QBin { int R = 1, NW = 8; bool tops = false; int64_t t0 = 0, nt = 0; int grid = 0; size_t lds = 0; int64_t bot0 = 0; double cost = 0.0; };
    struct QPlan {
        std::vector<P1Args> args;
        std::vector<P1Tile> tiles;
        std::vector<QBin> bins;
        int64_t bot_ld = 0, bot_floats = 0;
        DBuf dargs, dtiles, dhead, dbot;
    };
    bool use_queue = true;
    /* the DB sweep forms the 20-term score sums on the matrix cores (the
     * results are bit-identical either way) */
    bool use_mfma = false;
    /* MAP passes of a one-query call: the pairs no longer than half the
     * longest take at most this many rows per lane (0: the long pairs'
     * geometry for every pair) */
    int map_rmax2 = 0;
    /* MAP forward and backward passes stream a workgroup's pairs through one
     * band (queries whose stripes fit 8 waves) */
    bool map_pipe = true;
    QPlan qscan, qcand;
    DBuf qkeys, qwts;                   /* [nq][nprof] maxima, [nq][nprof][20] weights */
    unsigned long long* hqkeys = nullptr; size_t hqkeys_n = 0;
    int32_t seg_maxcols = 0;            /* columns of the widest segment */
    hipEvent_t gfork = nullptr, gjoin[C2_NSTREAMS - 1] = {nullptr, nullptr, nullptr};
    hipEvent_t gt0 = nullptr, gt1 = nullptr;
    /* fork/join of the MAP backward pass beside the forward pass (one-query
     * calls: the long-pair group and the rest) */
    hipEvent_t mfork[2] = {nullptr, nullptr}, mjoin[2] = {nullptr, nullptr};
    uint64_t chunk_gen = 0;             /* bumped by every chunk upload / tuning change */
    /* device-built stage lists (default on): per query and per stream slot
     * buffers, capacities, results; the host-decided stages remain the
     * fallback of a query that exceeds a capacity */
    bool use_dlists = true;
    std::vector<std::unique_ptr<DlQuery>> dlq;
    DlSlot dls[C2_NSTREAMS];
    std::vector<DlResult> dlres;
    int32_t dl_capC = 65536, dl_capP = 16384;
    int64_t dl_capW = (int64_t)1 << 27, dl_capCells = (int64_t)1 << 28;
    int32_t chunk_maxlen = 0;           /* longest profile of the chunk */
    int32_t last_dl_fallbacks = 0;      /* queries of the last call that fell back */
    /* whole-DB search in chunks: two device chunk sets, pinned staging
     * buffers and the copy stream of the copy thread, merged results */
    DevSet dsets[2];
    char* stage[2] = {nullptr, nullptr};
    size_t stage_n[2] = {0, 0};
    hipEvent_t up_ev[2] = {nullptr, nullptr};
    hipStream_t ustream = nullptr;
    std::vector<StreamOut> smerged;
    int32_t last_nchunks = 0;
    /* soft DP: query set on the device and the prepared pair plan */
    DBuf sdq_trans, sdq_dist, sdq_ll, sdq_ld, sdq_trg, sdq_ss, sdq_bkg, dssss;
    bool sdq_ready = false;
    std::vector<SdpPair> sdp_pairs;
    std::vector<int32_t> sdp_order;
    DBuf sdp_dpairs, sdp_dorder;
    int32_t sdp_K = 0, sdp_nw = 1, sdp_groups = 1;
    int64_t sdp_ws_floats = 0, sdp_scr_off = 0;
    bool sdp_ready = false;
};

/* internal helpers shared by the library's translation units */
int c2_fail(cbp_handle* h, int code, const std::string& msg);
int c2_dalloc(cbp_handle* h, DBuf& b, size_t n);
template <class T>
inline int c2_dupload(cbp_handle* h, DBuf& b, const T* src, size_t count, hipStream_t st) {
    int rc = c2_dalloc(h, b, sizeof(T) * count);
    if (rc) return rc;
    if (count) {
        const hipError_t e = hipMemcpyAsync(b.p, src, sizeof(T) * count, hipMemcpyHostToDevice, st);
        if (e != hipSuccess) return c2_fail(h, CBP_ERR_HIP, std::string("upload: ") + hipGetErrorString(e));
    }
    return CBP_OK;
}
DevQuery* c2_get_query(cbp_handle* h, int qidx, int* rc);
int c2_append_query(cbp_handle* h, const cbp_profiles* one, int32_t* qidx);
void c2_drop_graphs(cbp_handle* h);
/* device address of the allocation-free entry point's scratch array on the
 * current device, looked up once per device */
hipError_t c2_noalloc_init(int device);
int c2_fill_p1args(cbp_handle* h, DevQuery* dq, P1Args* a);
/* the CV term's arguments (wgt 0 when the term is off) */
void c2_fill_cvs(cbp_handle* h, DevQuery* dq, CvsCtx* v);
/* with the CV term on: the chunk's context vectors (the uninformative
 * vector when none were uploaded) */
int c2_ensure_chunk_cv(cbp_handle* h);
void c2_cv_constants(float* kappa_inv, float* nu_pow, float* cterm);
/* the uninformative context vector entry of the profile reader */
constexpr float C2_CV_UNINF = 0.17302947f;
void c2_fill_scorectx(cbp_handle* h, DevQuery* dq, ScoreCtx* c);
uint64_t c2_length_adjustment(const cbp_options& o, uint64_t m, uint64_t n, uint64_t N, uint64_t* delta);
int c2_build_segments(cbp_handle* h, const cbp_profiles* s);
int c2_run_phase1(cbp_handle* h, int qidx);
int c2_run_queries(cbp_handle* h, const int32_t* qidx, int32_t nq, const float* scorethld,
                   const float* logevthld, cbp_results* outs);
/* the same sequence with the stage lists built on the device */
int c2_run_queries_dev(cbp_handle* h, const int32_t* qidx, int32_t nq, const float* scorethld,
                       const float* logevthld, cbp_results* outs);
/* stage helpers shared by both sequences (c2_pipe.cpp) */
hipStream_t c2_slot_stream(cbp_handle* h, int slot);
int c2_ensure_fork_events(cbp_handle* h);
int c2_phase1_prepare(cbp_handle* h, QWork& w);
int c2_phase1_launch(cbp_handle* h, QWork& w, bool timed, bool copy);
int c2_queue_scan_prepare(cbp_handle* h, int nq);
int c2_queue_scan_launch(cbp_handle* h, int nq, bool timed, bool copy);

#endif
