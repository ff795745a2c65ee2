/*
 * c2_lib.cpp -- implementation of the C ABI (include/comer2_hip.h).
 */
#include <cstring>
#include <cmath>
#include <cstdio>
#include <algorithm>
#include "c2_lib.h"
#include "c2_fmath.h"

#define HIPCHK(call)                                                             \
    do {                                                                         \
        hipError_t e_ = (call);                                                  \
        if (e_ != hipSuccess) {                                                  \
            h->err = std::string(#call) + ": " + hipGetErrorString(e_);          \
            return CBP_ERR_HIP;                                                  \
        }                                                                        \
    } while (0)

int c2_fail(cbp_handle* h, int code, const std::string& msg) {
    if (h) h->err = msg;
    return code;
}
static int fail(cbp_handle* h, int code, const std::string& msg) { return c2_fail(h, code, msg); }

void HostSet::copy_from(const cbp_profiles* s) {
    const size_t np = (size_t)s->nprof, ns = (size_t)s->npos;
    bkg.assign(s->bkg, s->bkg + 20 * np);
    eno.assign(s->eno, s->eno + np);
    len.assign(s->len, s->len + np);
    dist.assign(s->dist, s->dist + np);
    trgfrq.assign(s->trgfrq, s->trgfrq + 20 * ns);
    ssprb.assign(s->ssprb, s->ssprb + 3 * ns);
    aa.assign(s->aa, s->aa + ns);
    sss.assign(s->sss, s->sss + ns);
    trans.assign(s->trans, s->trans + 7 * (ns + np));
    if (s->cv) {
        cv.assign(s->cv, s->cv + (size_t)CBP_NCVEL * ns);
        cvprior.assign(s->cvprior, s->cvprior + ns);
        cvnorm2.assign(s->cvnorm2, s->cvnorm2 + ns);
    } else {
        cv.clear(); cvprior.clear(); cvnorm2.clear();
    }
    v.nprof = s->nprof; v.npos = s->npos;
    v.bkg = bkg.data(); v.eno = eno.data(); v.len = len.data(); v.dist = dist.data();
    v.trgfrq = trgfrq.data(); v.ssprb = ssprb.data(); v.aa = aa.data(); v.sss = sss.data();
    v.trans = trans.data();
    v.cv = s->cv ? cv.data() : nullptr;
    v.cvprior = s->cv ? cvprior.data() : nullptr;
    v.cvnorm2 = s->cv ? cvnorm2.data() : nullptr;
}

int c2_dalloc(cbp_handle* h, DBuf& b, size_t n) {
    if (b.p && b.n >= n) return CBP_OK;
    b.release();
    if (n == 0) n = 8;
    if (h->budget && h->used + n > h->budget)
        return fail(h, CBP_ERR_MEM, "device memory budget exceeded");
    HIPCHK(hipMalloc(&b.p, n));
    /* a status left pending by a successful allocation (seen with two HIP
     * runtimes in the process) must not reach the next launch check */
    (void)hipGetLastError();
    b.n = n; b.acct = &h->used; h->used += n;
    return CBP_OK;
}

static int dalloc(cbp_handle* h, DBuf& b, size_t n) { return c2_dalloc(h, b, n); }

template <class T>
static int dupload(cbp_handle* h, DBuf& b, const T* src, size_t count) {
    return c2_dupload(h, b, src, count, h->stream);
}

static int check_set(const cbp_profiles* s) {
    if (!s || s->nprof < 1 || s->npos < 1 || !s->bkg || !s->eno || !s->len || !s->dist ||
        !s->trgfrq || !s->ssprb || !s->aa || !s->sss || !s->trans)
        return 0;
    /* the three context-vector fields come together or not at all */
    if ((s->cv == nullptr) != (s->cvprior == nullptr) || (s->cv == nullptr) != (s->cvnorm2 == nullptr))
        return 0;
    int64_t acc = 0;
    for (int32_t p = 0; p < s->nprof; p++) {
        if (s->len[p] < 1 || s->len[p] > 65535 || (int64_t)s->dist[p] != acc) return 0;
        acc += s->len[p];
    }
    if (acc != s->npos) return 0;
    /* SS probabilities index the 30-entry SS tables: they must lie in [0,1] */
    for (size_t i = 0; i < (size_t)3 * s->npos; i++)
        if (!(s->ssprb[i] >= 0.0f && s->ssprb[i] <= 1.0f)) return 0;
    return 1;
}

extern "C" {

int cbp_create(int device, size_t mem_budget, int area, cbp_handle** out)
{
    C2ApiScope c2_scope_;
    (void)area;
    if (!out) return CBP_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return CBP_ERR_NODEV;
    if (device < 0 || device >= ndev) return CBP_ERR_ARG;
    cbp_handle* h = new cbp_handle();
    h->device = device;
    h->budget = mem_budget;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&h->stream) != hipSuccess ||
        hipEventCreate(&h->ev0) != hipSuccess || hipEventCreate(&h->ev1) != hipSuccess) {
        delete h;
        return CBP_ERR_HIP;
    }
    /* the allocation-free entry point's scratch array: its address on this
     * device (a failed lookup is reported by that entry point) */
    (void)c2_noalloc_init(device);
    *out = h;
    return CBP_OK;
}

int cbp_destroy(cbp_handle* h)
{
    C2ApiScope c2_scope_;
    if (!h) return CBP_ERR_ARG;
    (void)hipSetDevice(h->device);
    (void)hipStreamSynchronize(h->stream);
    for (auto& g : h->graphs) if (g.second.exec) (void)hipGraphExecDestroy(g.second.exec);
    h->graphs.clear();
    if (h->gfork) (void)hipEventDestroy(h->gfork);
    for (int s = 0; s < C2_NSTREAMS - 1; s++) if (h->gjoin[s]) (void)hipEventDestroy(h->gjoin[s]);
    for (int s = 0; s < 2; s++) {
        if (h->mfork[s]) (void)hipEventDestroy(h->mfork[s]);
        if (h->mjoin[s]) (void)hipEventDestroy(h->mjoin[s]);
    }
    if (h->gt0) (void)hipEventDestroy(h->gt0);
    if (h->gt1) (void)hipEventDestroy(h->gt1);
    if (h->hqkeys) (void)hipHostFree(h->hqkeys);
    h->work.clear();
    h->qcache.clear();
    h->dchunk = DevSet();
    for (int s = 0; s < 2; s++) {
        h->dsets[s] = DevSet();
        if (h->stage[s]) (void)hipHostFree(h->stage[s]);
        if (h->up_ev[s]) (void)hipEventDestroy(h->up_ev[s]);
    }
    if (h->ustream) (void)hipStreamDestroy(h->ustream);
    for (int s = 0; s < C2_NSTREAMS - 1; s++)
        if (h->xstream[s]) (void)hipStreamDestroy(h->xstream[s]);
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    /* the teardown's own failures are not reported; clear the thread's last
     * error so that another handle's next launch check does not read it */
    (void)hipGetLastError();
    return CBP_OK;
}

const char* cbp_last_error(const cbp_handle* h)
{
    return h ? h->err.c_str() : "null handle";
}

int cbp_set_options(cbp_handle* h, const cbp_options* opt)
{
    C2ApiScope c2_scope_;
    if (!h || !opt) return CBP_ERR_ARG;
    if (!(opt->ssswgt >= 0.0f) || !(opt->cvswgt >= 0.0f) || !(opt->gap_lambda > 0.0f) || !(opt->gap_K > 0.0f))
        return fail(h, CBP_ERR_ARG, "invalid options");
    /* the CV term changes the queries' geometry (at most 2 rows per lane)
     * and tables: their preparations and the graphs holding them retire */
    if ((opt->cvswgt > 0.0f) != (h->opt.cvswgt > 0.0f) && !h->qcache.empty()) {
        (void)hipStreamSynchronize(h->stream);
        h->qcache.clear();
        c2_drop_graphs(h);
    }
    h->opt = *opt;
    h->have_opt = true;
    h->chunk_gen++;
    return CBP_OK;
}

int cbp_upload_tables(cbp_handle* h, const float* ssss900, const float* nnw, int32_t nnw_count)
{
    C2ApiScope c2_scope_;
    if (!h || !ssss900) return CBP_ERR_ARG;
    h->ssss.assign(ssss900, ssss900 + 900);
    if (nnw) {
        if (nnw_count != 93) return fail(h, CBP_ERR_ARG, "expected 93 SS18 parameters");
        h->nnw.assign(nnw, nnw + 93);
        int rc = dupload(h, h->dnnw, h->nnw.data(), 93);
        if (rc) return rc;
        HIPCHK(hipStreamSynchronize(h->stream));
    }
    h->have_tables = true;
    h->qcache.clear();
    h->chunk_gen++;
    return CBP_OK;
}

int cbp_upload_queries(cbp_handle* h, const cbp_profiles* q)
{
    C2ApiScope c2_scope_;
    if (!h) return CBP_ERR_ARG;
    if (!check_set(q)) return fail(h, CBP_ERR_ARG, "invalid query profile set");
    h->queries.copy_from(q);
    h->chunk_gen++;
    h->have_queries = true;
    h->qcache.clear();
    h->sdq_ready = false;
    h->sdp_ready = false;
    return CBP_OK;
}

}  /* extern "C" */

/* appends one profile to the query set; the preparations of the queries
 * already there stay cached (their indices and data are unchanged) */
int c2_append_query(cbp_handle* h, const cbp_profiles* one, int32_t* qidx)
{
    if (!check_set(one) || one->nprof != 1) return fail(h, CBP_ERR_ARG, "invalid query profile");
    if (!h->have_queries) {
        const int rc = cbp_upload_queries(h, one);
        if (rc == CBP_OK) *qidx = 0;
        return rc;
    }
    const cbp_profiles& o = h->queries.v;
    const size_t np = (size_t)o.nprof, ns = (size_t)o.npos, L = (size_t)one->npos;
    const size_t np2 = np + 1, ns2 = ns + L;
    std::vector<float> bkg(20 * np2), eno(np2), trg(20 * ns2), ss(3 * ns2), tr(7 * (ns2 + np2));
    std::vector<int32_t> len(np2);
    std::vector<uint32_t> dist(np2);
    std::vector<char> aa(ns2), sss(ns2);
    for (size_t a = 0; a < 20; a++) {
        std::copy(o.bkg + a * np, o.bkg + (a + 1) * np, bkg.begin() + a * np2);
        bkg[a * np2 + np] = one->bkg[a];
        std::copy(o.trgfrq + a * ns, o.trgfrq + (a + 1) * ns, trg.begin() + a * ns2);
        std::copy(one->trgfrq + a * L, one->trgfrq + (a + 1) * L, trg.begin() + a * ns2 + ns);
    }
    for (size_t k = 0; k < 3; k++) {
        std::copy(o.ssprb + k * ns, o.ssprb + (k + 1) * ns, ss.begin() + k * ns2);
        std::copy(one->ssprb + k * L, one->ssprb + (k + 1) * L, ss.begin() + k * ns2 + ns);
    }
    for (size_t k = 0; k < 7; k++) {
        std::copy(o.trans + k * (ns + np), o.trans + (k + 1) * (ns + np), tr.begin() + k * (ns2 + np2));
        std::copy(one->trans + k * (L + 1), one->trans + (k + 1) * (L + 1), tr.begin() + k * (ns2 + np2) + ns + np);
    }
    std::copy(o.eno, o.eno + np, eno.begin()); eno[np] = one->eno[0];
    std::copy(o.len, o.len + np, len.begin()); len[np] = one->len[0];
    std::copy(o.dist, o.dist + np, dist.begin()); dist[np] = (uint32_t)ns;
    std::copy(o.aa, o.aa + ns, aa.begin()); std::copy(one->aa, one->aa + L, aa.begin() + ns);
    std::copy(o.sss, o.sss + ns, sss.begin()); std::copy(one->sss, one->sss + L, sss.begin() + ns);
    cbp_profiles v{};
    v.nprof = (int32_t)np2; v.npos = (int64_t)ns2;
    v.bkg = bkg.data(); v.eno = eno.data(); v.len = len.data(); v.dist = dist.data();
    v.trgfrq = trg.data(); v.ssprb = ss.data(); v.aa = aa.data(); v.sss = sss.data(); v.trans = tr.data();
    /* context vectors: kept when either part has them, the other part's
     * positions taking the uninformative vector */
    std::vector<float> cv, cvp, cvn;
    if (o.cv || one->cv) {
        cv.assign((size_t)CBP_NCVEL * ns2, C2_CV_UNINF);
        cvp.assign(ns2, 0.0f);
        cvn.assign(ns2, 0.0f);
        for (size_t i = 0; i < (size_t)CBP_NCVEL; i++) {
            if (o.cv) std::copy(o.cv + i * ns, o.cv + (i + 1) * ns, cv.begin() + i * ns2);
            if (one->cv) std::copy(one->cv + i * L, one->cv + (i + 1) * L, cv.begin() + i * ns2 + ns);
        }
        if (o.cv) { std::copy(o.cvprior, o.cvprior + ns, cvp.begin()); std::copy(o.cvnorm2, o.cvnorm2 + ns, cvn.begin()); }
        if (one->cv) {
            std::copy(one->cvprior, one->cvprior + L, cvp.begin() + ns);
            std::copy(one->cvnorm2, one->cvnorm2 + L, cvn.begin() + ns);
        }
        v.cv = cv.data(); v.cvprior = cvp.data(); v.cvnorm2 = cvn.data();
    }
    h->queries.copy_from(&v);
    h->chunk_gen++;
    h->sdq_ready = false;
    h->sdp_ready = false;
    *qidx = (int32_t)np;
    return CBP_OK;
}

/* segments of whole profiles of about seg_cols columns each (the unit of
 * work of one phase-1 wave) */
int c2_build_segments(cbp_handle* h, const cbp_profiles* s)
{
    h->seg_p.clear();
    h->seg_p.push_back(0);
    int64_t acc = 0;
    for (int32_t p = 0; p < s->nprof; p++) {
        acc += s->len[p];
        if (acc >= h->seg_cols) { h->seg_p.push_back(p + 1); acc = 0; }
    }
    if (h->seg_p.back() != s->nprof) h->seg_p.push_back(s->nprof);
    h->nseg = (int32_t)h->seg_p.size() - 1;
    h->seg_maxcols = 0;
    h->chunk_maxlen = 0;
    for (int32_t p = 0; p < s->nprof; p++)
        if (s->len[p] > h->chunk_maxlen) h->chunk_maxlen = s->len[p];
    for (int32_t i = 0; i < h->nseg; i++) {
        const int32_t pl = h->seg_p[i + 1] - 1;
        const int64_t cols = (int64_t)s->dist[pl] + s->len[pl] - (int64_t)s->dist[h->seg_p[i]];
        if (cols > h->seg_maxcols) h->seg_maxcols = (int32_t)cols;
    }
    int rc;
    if ((rc = dupload(h, h->dseg_b, h->seg_p.data(), h->nseg))) return rc;
    if ((rc = dupload(h, h->dseg_e, h->seg_p.data() + 1, h->nseg))) return rc;
    HIPCHK(hipStreamSynchronize(h->stream));
    return CBP_OK;
}

extern "C" {

int cbp_upload_chunk(cbp_handle* h, const cbp_profiles* db)
{
    C2ApiScope c2_scope_;
    if (!h) return CBP_ERR_ARG;
    if (!check_set(db)) return fail(h, CBP_ERR_ARG, "invalid database profile set");
    (void)hipSetDevice(h->device);
    const cbp_profiles& s = *db;
    const size_t np = s.nprof, ns = s.npos;
    DevSet& d = h->dchunk;
    d.nprof = s.nprof; d.npos = s.npos;
    int rc;
    if ((rc = dupload(h, d.bkg, s.bkg, 20 * np))) return rc;
    if ((rc = dupload(h, d.eno, s.eno, np))) return rc;
    if ((rc = dupload(h, d.len, s.len, np))) return rc;
    if ((rc = dupload(h, d.dist, s.dist, np))) return rc;
    if ((rc = dupload(h, d.trgfrq, s.trgfrq, 20 * ns))) return rc;
    if ((rc = dupload(h, d.ssprb, s.ssprb, 3 * ns))) return rc;
    if ((rc = dupload(h, d.aa, s.aa, ns))) return rc;
    if ((rc = dupload(h, d.sss, s.sss, ns))) return rc;
    if ((rc = dupload(h, d.trans, s.trans, 7 * (ns + np)))) return rc;
    d.has_cv = s.cv != nullptr;
    if (d.has_cv) {
        if ((rc = dupload(h, d.cv, s.cv, (size_t)CBP_NCVEL * ns))) return rc;
        if ((rc = dupload(h, d.cvp, s.cvprior, ns))) return rc;
        if ((rc = dupload(h, d.cvn, s.cvnorm2, ns))) return rc;
    } else {
        d.cv.release(); d.cvp.release(); d.cvn.release();
    }
    /* profile index of every position (the reference's pmv2DAddrPro field) */
    std::vector<int32_t> ap(ns);
    for (int32_t p = 0; p < s.nprof; p++)
        std::fill(ap.begin() + s.dist[p], ap.begin() + s.dist[p] + s.len[p], p);
    if ((rc = dupload(h, d.addrpro, ap.data(), ns))) return rc;
    if ((rc = c2_build_segments(h, &s))) return rc;
    if ((rc = dalloc(h, h->keys, sizeof(unsigned long long) * np))) return rc;
    h->chunk.copy_from(db);
    h->chunk_gen++;
    HIPCHK(hipStreamSynchronize(h->stream));
    h->have_chunk = true;
    h->sdp_ready = false;
    return CBP_OK;
}

int cbp_set_db_details(cbp_handle* h, uint64_t dblen, uint64_t nseqs)
{
    C2ApiScope c2_scope_;
    if (!h) return CBP_ERR_ARG;
    h->dblen = dblen; h->nseqs = nseqs;
    return CBP_OK;
}

int cbp_score_threshold(const cbp_options* o, uint64_t qlen, uint64_t dblen,
                        float* scorethld, float* logevthld)
{
    C2ApiScope c2_scope_;
    if (!o || !scorethld || !logevthld) return CBP_ERR_ARG;
    const uint64_t sspace = qlen * dblen;
    float ev = o->evalue;
    if (sspace < 1 || o->gap_K <= 0.0f || o->gap_lambda <= 0.0f) return CBP_ERR_ARG;
    if (ev <= 0.0f) { *scorethld = 999.0f; *logevthld = -99.0f; return CBP_OK; }
    *logevthld = logf(ev);
    if (ev < 1.0f) ev *= 10.0f; else ev *= 3.0f;
    ev /= o->gap_K * (float)sspace;
    if (ev <= 0.0f) { *scorethld = 999.0f; return CBP_OK; }
    ev = logf(ev);
    float sc = -ev / o->gap_lambda;
    *scorethld = sc < 0.0f ? 0.0f : sc;
    return CBP_OK;
}

}  /* extern "C" */

/* device copy of a query's preparation (cached per query index) */
DevQuery* c2_get_query(cbp_handle* h, int qidx, int* rc)
{
    auto it = h->qcache.find(qidx);
    if (it != h->qcache.end()) { it->second.tick = h->qtick; return it->second.dq.get(); }
    /* evict the least recently used preparation of an earlier call; graphs
     * hold the evicted buffers' addresses, so they retire with a new
     * generation */
    while (h->qcache.size() >= C2_QCACHE_MAX) {
        auto lru = h->qcache.end();
        for (auto e = h->qcache.begin(); e != h->qcache.end(); ++e)
            if (e->second.tick < h->qtick && (lru == h->qcache.end() || e->second.tick < lru->second.tick)) lru = e;
        if (lru == h->qcache.end()) break;      /* all belong to the running call */
        (void)hipStreamSynchronize(h->stream);
        h->qcache.erase(lru);
        c2_drop_graphs(h);
        h->chunk_gen++;
    }
    auto dq = std::make_unique<DevQuery>();
    /* the CV term keeps each lane's rows' 21 context-vector values in
     * registers: at most 2 rows per lane */
    const bool cvon = h->opt.cvswgt > 0.0f;
    c2_prepare_query(&h->queries.v, qidx, h->ssss.data(), cvon && h->rmax > 2 ? 2 : h->rmax, &dq->hp);
    const QueryPrep& p = dq->hp;
    if (cvon) {
        /* [21][Lpad]: the 19 entries, the log prior, the squared norm; rows
         * past the query and a query without context vectors take the
         * uninformative vector */
        const cbp_profiles& qs = h->queries.v;
        const size_t Lp = (size_t)p.Lpad, q0 = qs.dist[qidx];
        std::vector<float> qc(21 * Lp, 0.0f);
        for (size_t i = 0; i < (size_t)CBP_NCVEL; i++)
            for (size_t y = 0; y < Lp; y++)
                qc[i * Lp + y] = (qs.cv && y < (size_t)p.Lq) ? qs.cv[i * (size_t)qs.npos + q0 + y] : C2_CV_UNINF;
        if (qs.cv)
            for (size_t y = 0; y < (size_t)p.Lq; y++) {
                qc[19 * Lp + y] = qs.cvprior[q0 + y];
                qc[20 * Lp + y] = qs.cvnorm2[q0 + y];
            }
        if ((*rc = dupload(h, dq->qcv, qc.data(), qc.size()))) return nullptr;
    }
    if ((*rc = dupload(h, dq->qv, p.qv.data(), p.qv.size()))) return nullptr;
    if ((*rc = dupload(h, dq->qtr, p.qtr.data(), p.qtr.size()))) return nullptr;
    if ((*rc = dupload(h, dq->qtrb, p.qtrb.data(), p.qtrb.size()))) return nullptr;
    if ((*rc = dupload(h, dq->mleft, p.mleft.data(), p.mleft.size()))) return nullptr;
    if ((*rc = dupload(h, dq->mdiag, p.mdiag.data(), p.mdiag.size()))) return nullptr;
    if ((*rc = dupload(h, dq->QT, p.QT.data(), p.QT.size()))) return nullptr;
    if ((*rc = dupload(h, dq->qh, p.qh.data(), p.qh.size()))) return nullptr;
    if ((*rc = dupload(h, dq->lleft, p.lleft.data(), p.lleft.size()))) return nullptr;
    if ((*rc = dupload(h, dq->ldiag, p.ldiag.data(), p.ldiag.size()))) return nullptr;
    if ((*rc = dupload(h, dq->qbkg, p.qbkg.data(), p.qbkg.size()))) return nullptr;
    {
        std::vector<char> qa((size_t)p.Lpad, 'X');
        const size_t q0 = h->queries.v.dist[qidx];
        std::memcpy(qa.data(), h->queries.v.aa + q0, (size_t)p.Lq);
        if ((*rc = dupload(h, dq->qaa, qa.data(), qa.size()))) return nullptr;
    }
    DevQuery* r = dq.get();
    cbp_handle::QEnt& e = h->qcache[qidx];
    e.dq = std::move(dq);
    e.tick = h->qtick;
    *rc = CBP_OK;
    return r;
}

int c2_fill_p1args(cbp_handle* h, DevQuery* dq, P1Args* a)
{
    DevSet& d = h->dchunk;
    std::memset(a, 0, sizeof *a);
    a->trgfrq = d.trgfrq.as<float>(); a->ssprb = d.ssprb.as<float>();
    a->trans = d.trans.as<float>(); a->bkg = d.bkg.as<float>();
    a->len = d.len.as<int32_t>(); a->dist = d.dist.as<uint32_t>();
    a->addrpro = d.addrpro.as<int32_t>();
    a->npos = d.npos; a->nprof = d.nprof;
    a->qv = dq->qv.as<float>(); a->qtr = dq->qtr.as<float>(); a->QT = dq->QT.as<float>();
    a->qh = dq->qh.as<float>(); a->lleft = dq->lleft.as<float>(); a->ldiag = dq->ldiag.as<float>();
    a->qbkg = dq->qbkg.as<float>();
    a->Lq = dq->hp.Lq; a->Lpad = dq->hp.Lpad; a->nstripes = dq->hp.nstripes;
    a->shift = h->opt.mapaln ? 0.002f : 0.0003f;
    a->ssswgt = h->opt.ssswgt;
    c2_fill_cvs(h, dq, &a->cvs);
    a->seg_pb = h->dseg_b.as<int32_t>(); a->seg_pe = h->dseg_e.as<int32_t>(); a->nseg = h->nseg;
    a->mfma = h->use_mfma;
    return CBP_OK;
}

void c2_fill_cvs(cbp_handle* h, DevQuery* dq, CvsCtx* v)
{
    std::memset(v, 0, sizeof *v);
    if (!(h->opt.cvswgt > 0.0f) || !h->have_cvs || !dq) return;
    const DevSet& d = h->dchunk;
    v->cv = d.cv.as<float>(); v->cvp = d.cvp.as<float>(); v->cvn = d.cvn.as<float>();
    v->qcv = dq->qcv.as<float>();
    v->map = h->dcvmap.as<float>();
    v->card = (int32_t)h->cvvals.size();
    v->shift = h->cvshift;
    v->trunc = h->opt.mapaln ? 0 : 1;
    v->wgt = h->opt.cvswgt;
    c2_cv_constants(&v->kinv, &v->nu, &v->cterm);
    v->k1 = h->cvkeys.front(); v->v1 = h->cvvals.front();
    v->kl = h->cvkeys.back(); v->vl = h->cvvals.back();
}

int c2_ensure_chunk_cv(cbp_handle* h)
{
    if (!(h->opt.cvswgt > 0.0f)) return CBP_OK;
    if (!h->have_cvs) return fail(h, CBP_ERR_STATE, "CVSWGT > 0 needs the context-vector map (cbp_upload_cvs2s)");
    DevSet& d = h->dchunk;
    if (d.has_cv) return CBP_OK;
    const size_t ns = (size_t)d.npos;
    int rc;
    if ((rc = dalloc(h, d.cv, sizeof(float) * CBP_NCVEL * ns))) return rc;
    if ((rc = dalloc(h, d.cvp, sizeof(float) * ns))) return rc;
    if ((rc = dalloc(h, d.cvn, sizeof(float) * ns))) return rc;
    uint32_t u;
    std::memcpy(&u, &C2_CV_UNINF, 4);
    HIPCHK(c2_launch_fill_u32(d.cv.as<uint32_t>(), (int64_t)(CBP_NCVEL * ns), u, h->stream));
    HIPCHK(c2_launch_fill_u32(d.cvp.as<uint32_t>(), (int64_t)ns, 0u, h->stream));
    HIPCHK(c2_launch_fill_u32(d.cvn.as<uint32_t>(), (int64_t)ns, 0u, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    d.has_cv = true;
    return CBP_OK;
}

extern "C" int cbp_upload_cvs2s(cbp_handle* h, const float* keys, const float* values, int32_t card,
                                int32_t shift)
{
    C2ApiScope c2_scope_;
    if (!h || !keys || !values) return CBP_ERR_ARG;
    /* the phase-1 kernels keep the map in 128 words of LDS */
    if (card < 1 || card > 128 || shift < 0 || shift >= card)
        return fail(h, CBP_ERR_ARG, "invalid context-vector map (1..128 entries, shift within it)");
    for (int32_t i = 1; i < card; i++)
        if (!(keys[i] > keys[i - 1])) return fail(h, CBP_ERR_ARG, "context-vector map keys are not increasing");
    h->cvkeys.assign(keys, keys + card);
    h->cvvals.assign(values, values + card);
    h->cvshift = shift;
    int rc;
    if ((rc = dupload(h, h->dcvmap, h->cvvals.data(), (size_t)card))) return rc;
    HIPCHK(hipStreamSynchronize(h->stream));
    h->have_cvs = true;
    h->chunk_gen++;
    return CBP_OK;
}

/* phase-1 launch for one query into h->keys; records kernel time */
int c2_run_phase1(cbp_handle* h, int qidx)
{
    (void)hipGetLastError();   /* an earlier call's unreported status is not this call's */
    if (!h->have_opt || !h->have_tables || !h->have_queries || !h->have_chunk)
        return fail(h, CBP_ERR_STATE, "options, tables, queries and chunk must be uploaded first");
    if (qidx < 0 || qidx >= h->queries.v.nprof) return fail(h, CBP_ERR_ARG, "query index out of range");
    (void)hipSetDevice(h->device);
    int rc;
    h->qtick++;
    DevQuery* dq = c2_get_query(h, qidx, &rc);
    if (!dq) return rc;
    if ((rc = c2_ensure_chunk_cv(h))) return rc;
    P1Args a;
    if ((rc = c2_fill_p1args(h, dq, &a))) return rc;
    a.keys = h->keys.as<unsigned long long>();
    if (dq->hp.nstripes > 1) {
        if ((rc = dalloc(h, h->bot, sizeof(float) * 5 * (size_t)h->dchunk.npos))) return rc;
        a.bot = h->bot.as<float>();
    }
    if ((rc = dalloc(h, h->wts, sizeof(float) * 20 * (size_t)h->dchunk.nprof))) return rc;
    a.wts = h->wts.as<float>();
    HIPCHK(c2_launch_weights(dq->qbkg.as<float>(), h->dchunk.bkg.as<float>(), h->dchunk.nprof, h->wts.as<float>(), h->stream));
    int nw = h->nw;
    while (nw > 1 && c2_phase1_lds_bytes(dq->hp.R, nw, dq->hp.nstripes) > 160 * 1024) nw >>= 1;
    HIPCHK(hipMemsetAsync(h->keys.p, 0, sizeof(unsigned long long) * h->dchunk.nprof, h->stream));
    HIPCHK(hipEventRecord(h->ev0, h->stream));
    HIPCHK(c2_launch_phase1(a, dq->hp.R, nw, false, h->stream));
    HIPCHK(hipEventRecord(h->ev1, h->stream));
    HIPCHK(hipEventSynchronize(h->ev1));
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    h->last_ms = ms; h->last_nlaunch = 1;
    HIPCHK(hipGetLastError());   /* nothing of this call's left unreported */
    return CBP_OK;
}

extern "C" {

int cbp_phase1_keys(cbp_handle* h, int32_t qidx, uint64_t* keys)
{
    C2ApiScope c2_scope_;
    if (!h || !keys) return CBP_ERR_ARG;
    int rc = c2_run_phase1(h, qidx);
    if (rc) return rc;
    HIPCHK(hipMemcpy(keys, h->keys.p, sizeof(uint64_t) * h->dchunk.nprof, hipMemcpyDeviceToHost));
    return CBP_OK;
}

int cbp_last_kernel_ms(const cbp_handle* h, double* ms, int32_t* nlaunch)
{
    C2ApiScope c2_scope_;
    if (!h || !ms) return CBP_ERR_ARG;
    *ms = h->last_ms;
    if (nlaunch) *nlaunch = h->last_nlaunch;
    return CBP_OK;
}

/* tuning knobs (autotuner): target segment columns, max rows per lane, waves per group */
int cbp_set_tuning(cbp_handle* h, int32_t seg_cols, int32_t rmax, int32_t nw)
{
    C2ApiScope c2_scope_;
    if (!h || seg_cols < 64 || rmax < 1 || rmax > 5 || !(nw == 1 || nw == 2 || nw == 4 || nw == 8))
        return CBP_ERR_ARG;
    const bool reseg = (seg_cols != h->seg_cols);
    h->seg_cols = seg_cols; h->rmax = rmax; h->nw = nw;
    h->qcache.clear();
    h->chunk_gen++;
    if (reseg && h->have_chunk) {
        const int rc = c2_build_segments(h, &h->chunk.v);
        if (rc) return rc;
    }
    return CBP_OK;
}

}  /* extern "C" */

/* ----------------------------------------------------------------------- */
/* full pipeline                                                            */

/* AbstractScoreMatrix::ComputeLengthAdjustment (libpro/srcsco/AbstractScoreMatrix.cpp:
 * 247-390): Altschul-Gertz iteration in float, as the reference computes it */
uint64_t c2_length_adjustment(const cbp_options& o, uint64_t m, uint64_t n, uint64_t N, uint64_t* delta)
{
    const float lambda = o.gap_lambda, K = o.gap_K, alpha = o.gap_alpha, beta = o.gap_beta;
    uint64_t ss = m * n;
    *delta = 0;
    if (lambda <= 0.0f || K <= 0.0f || N < 1) return ss;
    const float logK = logf(K), alratio = alpha / lambda;
    const float max_mn = (float)((m > n) ? m : n);
    float length = 0.0f, min_length = 0.0f, max_length = 0.0f, nxt_length = 0.0f;
    float sspace = 0.0f, space2 = 0.0f;
    bool converged = false;
    const float a = (float)N, b = (float)(n + m * N), c = (float)(m * n) - max_mn / K;
    if (c < 0.0f) return ss;
    max_length = 2.0f * c / (b + sqrtf(b * b - 4.0f * a * c));
    for (int j = 0; j < 20; j++) {
        length = nxt_length;
        sspace = ((float)m - length) * ((float)n - (float)N * length);
        nxt_length = alratio * (logK + logf(sspace)) + beta;
        if (length <= nxt_length) {
            min_length = length;
            if (nxt_length - length <= 1.0f) { converged = true; break; }
            if (min_length == max_length) break;
        } else max_length = length;
        if (nxt_length < min_length || max_length < nxt_length)
            nxt_length = (!j) ? max_length : (min_length + max_length) * 0.5f;
    }
    if (converged) {
        length = ceilf(min_length);
        if (length <= max_length) {
            space2 = ((float)m - length) * ((float)n - (float)N * length);
            nxt_length = alratio * (logK + logf(space2)) + beta;
            if (length <= nxt_length) { min_length = length; sspace = space2; }
        }
    }
    *delta = (uint64_t)min_length;
    return (uint64_t)sspace;
}

extern "C" int cbp_search_space(const cbp_options* opt, uint64_t qlen, uint64_t dblen, uint64_t nseqs,
                                uint64_t* sspace, int32_t* deltalen)
{
    C2ApiScope c2_scope_;
    if (!opt || !sspace || !deltalen) return CBP_ERR_ARG;
    uint64_t d = 0;
    *sspace = c2_length_adjustment(*opt, qlen, dblen, nseqs, &d);
    *deltalen = (int32_t)d;
    return CBP_OK;
}

void c2_fill_scorectx(cbp_handle* h, DevQuery* dq, ScoreCtx* c)
{
    DevSet& d = h->dchunk;
    c->trgfrq = d.trgfrq.as<float>(); c->ssprb = d.ssprb.as<float>(); c->bkg = d.bkg.as<float>();
    c->len = d.len.as<int32_t>(); c->dist = d.dist.as<uint32_t>(); c->aa = d.aa.as<char>();
    c->npos = d.npos; c->nprof = d.nprof;
    c->qv = dq->qv.as<float>(); c->QT = dq->QT.as<float>(); c->qh = dq->qh.as<float>();
    c->qbkg = dq->qbkg.as<float>(); c->qaa = dq->qaa.as<char>();
    c->Lq = dq->hp.Lq; c->Lpad = dq->hp.Lpad; c->R = dq->hp.R; c->nstripes = dq->hp.nstripes;
    c->shift = h->opt.mapaln ? 0.002f : 0.0003f;
    c->ssswgt = h->opt.ssswgt;
    c2_fill_cvs(h, dq, &c->cvs);
}


static uint32_t f2u(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
static float u2f(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

extern "C" int cbp_score_block(cbp_handle* h, int32_t qidx, int32_t pronr, float* S, float* mod)
{
    C2ApiScope c2_scope_;
    if (!h || !S) return CBP_ERR_ARG;
    if (!h->have_opt || !h->have_tables || !h->have_queries || !h->have_chunk)
        return fail(h, CBP_ERR_STATE, "options, tables, queries and chunk must be uploaded first");
    if (qidx < 0 || qidx >= h->queries.v.nprof) return fail(h, CBP_ERR_ARG, "query index out of range");
    if (pronr < 0 || pronr >= h->dchunk.nprof) return fail(h, CBP_ERR_ARG, "profile index out of range");
    (void)hipSetDevice(h->device);
    int rc;
    h->qtick++;
    DevQuery* dq = c2_get_query(h, qidx, &rc);
    if (!dq) return rc;
    if ((rc = c2_ensure_chunk_cv(h))) return rc;
    const int Ld = h->chunk.v.len[pronr];
    const size_t n = (size_t)dq->hp.Lq * Ld;
    DBuf dS, dM;
    if ((rc = dalloc(h, dS, sizeof(float) * n))) return rc;
    if (mod && (rc = dalloc(h, dM, sizeof(float) * n))) return rc;
    ScoreCtx sc;
    c2_fill_scorectx(h, dq, &sc);
    hipError_t e = c2_launch_score_block(sc, pronr, Ld, dS.as<float>(), mod ? dM.as<float>() : nullptr, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(S, dS.p, sizeof(float) * n, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess && mod) e = hipMemcpyAsync(mod, dM.p, sizeof(float) * n, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return fail(h, CBP_ERR_HIP, std::string("cbp_score_block: ") + hipGetErrorString(e));
    return CBP_OK;
}


/* ---- differentiable soft DP ---- */

static int sdp_upload_queries(cbp_handle* h)
{
    if (h->sdq_ready) return CBP_OK;
    const cbp_profiles& q = h->queries.v;
    int rc;
    if ((rc = dupload(h, h->sdq_trans, q.trans, (size_t)7 * (q.npos + q.nprof)))) return rc;
    if ((rc = dupload(h, h->sdq_dist, q.dist, (size_t)q.nprof))) return rc;
    std::vector<float> ll((size_t)q.npos * 5), ld((size_t)q.npos * 5);
    for (int32_t p = 0; p < q.nprof; p++)
        c2_map_colm1(&q, p, ll.data() + (size_t)q.dist[p] * 5, ld.data() + (size_t)q.dist[p] * 5);
    if ((rc = dupload(h, h->sdq_ll, ll.data(), ll.size()))) return rc;
    if ((rc = dupload(h, h->sdq_ld, ld.data(), ld.size()))) return rc;
    if ((rc = dupload(h, h->sdq_trg, q.trgfrq, (size_t)20 * q.npos))) return rc;
    if ((rc = dupload(h, h->sdq_ss, q.ssprb, (size_t)3 * q.npos))) return rc;
    if ((rc = dupload(h, h->sdq_bkg, q.bkg, (size_t)20 * q.nprof))) return rc;
    if ((rc = dupload(h, h->dssss, h->ssss.data(), h->ssss.size()))) return rc;
    HIPCHK(hipStreamSynchronize(h->stream));
    h->sdq_ready = true;
    return CBP_OK;
}

extern "C" int cbp_sdp_prepare(cbp_handle* h, const cbp_sdp_pair* pairs, int32_t npairs, int32_t ckpt,
                               int64_t* ws_bytes)
{
    C2ApiScope c2_scope_;
    if (!h || !ws_bytes || npairs < 0 || (npairs > 0 && !pairs)) return CBP_ERR_ARG;
    if (!h->have_queries || !h->have_chunk)
        return fail(h, CBP_ERR_STATE, "queries and chunk must be uploaded first");
    (void)hipSetDevice(h->device);
    h->sdp_ready = false;
    int rc;
    if ((rc = sdp_upload_queries(h))) return rc;
    const int R = c2_sdp_rows_per_lane();
    const cbp_profiles& q = h->queries.v;
    const cbp_profiles& d = h->chunk.v;
    int maxsteps = 1;
    for (int32_t i = 0; i < npairs; i++) {
        if (pairs[i].qidx < 0 || pairs[i].qidx >= q.nprof || pairs[i].pronr < 0 || pairs[i].pronr >= d.nprof ||
            pairs[i].soff < 0)
            return fail(h, CBP_ERR_ARG, "soft DP: pair index or offset out of range");
        maxsteps = std::max(maxsteps, d.len[pairs[i].pronr] + 63);
    }
    /* checkpoint interval: sqrt of the longest step count when automatic,
     * rounded down to 16, 32 or 64 (the adjoint's phases are 64 steps apart
     * between stripes of a band, so the interval divides 64) */
    int K = ckpt;
    if (K < 0) K = (int)std::ceil(std::sqrt((double)maxsteps));
    if (K > 0) K = K >= 64 ? 64 : (K >= 32 ? 32 : 16);
    const auto align = [](int64_t x) { return (x + 63) & ~(int64_t)63; };
    h->sdp_pairs.resize(npairs);
    int64_t off = 0;
    for (int32_t i = 0; i < npairs; i++) {
        SdpPair& P = h->sdp_pairs[i];
        P.q = pairs[i].qidx; P.p = pairs[i].pronr;
        P.L1 = q.len[P.q]; P.L2 = d.len[P.p];
        P.ns = (P.L1 + 64 * R - 1) / (64 * R);
        const int nsteps = P.L2 + 63;
        P.ncp = K > 0 ? (nsteps - 1) / K + 1 : 0;
        P.soff = pairs[i].soff;
        P.foff = off;
        off += align(K > 0 ? (int64_t)P.ns * P.ncp * c2_sdp_ckpt_floats(R)
                           : (int64_t)P.ns * nsteps * 5 * R * 64);
        P.hoff = off;
        off += align((int64_t)P.ns * 5 * P.L2 + 3 * (int64_t)P.L2);
    }
    /* longest first (cost = stripes x steps), ties by pair index */
    h->sdp_order.resize(npairs);
    for (int32_t i = 0; i < npairs; i++) h->sdp_order[i] = i;
    std::stable_sort(h->sdp_order.begin(), h->sdp_order.end(), [&](int a, int b) {
        const SdpPair& A = h->sdp_pairs[a]; const SdpPair& B = h->sdp_pairs[b];
        return (int64_t)A.ns * (A.L2 + 63) > (int64_t)B.ns * (B.L2 + 63);
    });
    /* workgroups of up to 8 waves (one stripe each), about 4096 waves in
     * all: a stripe per wave only when the pairs alone do not fill the chip */
    int maxns = 1;
    for (const SdpPair& P : h->sdp_pairs) maxns = std::max(maxns, P.ns);
    const int want = npairs > 0 ? (4096 + npairs - 1) / npairs : 1;
    const int nw = std::max(1, std::min(std::min(maxns, 8), want));
    const int groups = std::max(1, std::min(npairs, 4096 / nw));
    h->sdp_nw = nw;
    h->sdp_groups = groups;
    h->sdp_scr_off = off;
    if (K > 0) off += align((int64_t)groups * nw * K * 5 * R * 64);
    h->sdp_K = K;
    h->sdp_ws_floats = off;
    if (npairs > 0) {
        if ((rc = dupload(h, h->sdp_dpairs, h->sdp_pairs.data(), h->sdp_pairs.size()))) return rc;
        if ((rc = dupload(h, h->sdp_dorder, h->sdp_order.data(), h->sdp_order.size()))) return rc;
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    *ws_bytes = off * (int64_t)sizeof(float);
    h->sdp_ready = true;
    return CBP_OK;
}

static void sdp_fill(cbp_handle* h, SdpArgs* a, const float* S, const float* theta, void* ws)
{
    std::memset(a, 0, sizeof *a);
    a->pairs = h->sdp_dpairs.as<SdpPair>(); a->order = h->sdp_dorder.as<int32_t>();
    a->npairs = (int32_t)h->sdp_pairs.size();
    const cbp_profiles& q = h->queries.v;
    a->qtr = h->sdq_trans.as<float>(); a->qntr = q.npos + q.nprof; a->qdist = h->sdq_dist.as<uint32_t>();
    a->qll = h->sdq_ll.as<float>(); a->qld = h->sdq_ld.as<float>();
    a->dtr = h->dchunk.trans.as<float>(); a->dntr = h->dchunk.npos + h->dchunk.nprof;
    a->ddist = h->dchunk.dist.as<uint32_t>();
    a->S = S; a->theta = theta;
    a->ws = (float*)ws; a->scr_off = h->sdp_scr_off; a->K = h->sdp_K;
}

extern "C" int cbp_sdp_forward(cbp_handle* h, const float* S, const float* theta, float* logZ,
                               void* ws, void* stream)
{
    C2ApiScope c2_scope_;
    if (!h || !S || !theta || !logZ || !ws) return CBP_ERR_ARG;
    if (!h->sdp_ready) return fail(h, CBP_ERR_STATE, "soft DP: cbp_sdp_prepare first");
    (void)hipSetDevice(h->device);
    SdpArgs a;
    sdp_fill(h, &a, S, theta, ws);
    a.logZ = logZ;
    HIPCHK(c2_launch_sdp(a, false, h->sdp_groups, h->sdp_nw, stream ? (hipStream_t)stream : h->stream));
    return CBP_OK;
}

extern "C" int cbp_sdp_adjoint(cbp_handle* h, const float* S, const float* theta, const float* logZ,
                               const float* gZ, float* dS, float* dtheta, float* post, void* ws, void* stream)
{
    C2ApiScope c2_scope_;
    if (!h || !S || !theta || !logZ || !dS || !dtheta || !ws) return CBP_ERR_ARG;
    if (!h->sdp_ready) return fail(h, CBP_ERR_STATE, "soft DP: cbp_sdp_prepare first");
    (void)hipSetDevice(h->device);
    SdpArgs a;
    sdp_fill(h, &a, S, theta, ws);
    a.logZ = (float*)logZ; a.gZ = gZ; a.dS = dS; a.dth = dtheta; a.post = post;
    HIPCHK(c2_launch_sdp(a, true, h->sdp_groups, h->sdp_nw, stream ? (hipStream_t)stream : h->stream));
    return CBP_OK;
}

extern "C" int cbp_sdp_scores(cbp_handle* h, float* S, void* stream)
{
    C2ApiScope c2_scope_;
    if (!h || !S) return CBP_ERR_ARG;
    if (!h->sdp_ready || !h->have_tables || !h->have_opt)
        return fail(h, CBP_ERR_STATE, "soft DP: tables, options and cbp_sdp_prepare first");
    (void)hipSetDevice(h->device);
    const cbp_profiles& q = h->queries.v;
    SdpScoreArgs a;
    a.pairs = h->sdp_dpairs.as<SdpPair>();
    a.qtrg = h->sdq_trg.as<float>(); a.qss = h->sdq_ss.as<float>(); a.qbkg = h->sdq_bkg.as<float>();
    a.qdist = h->sdq_dist.as<uint32_t>(); a.qnpos = q.npos; a.qnprof = q.nprof;
    a.dtrg = h->dchunk.trgfrq.as<float>(); a.dss = h->dchunk.ssprb.as<float>(); a.dbkg = h->dchunk.bkg.as<float>();
    a.ddist = h->dchunk.dist.as<uint32_t>(); a.dnpos = h->dchunk.npos; a.dnprof = h->dchunk.nprof;
    a.ssss = h->dssss.as<float>();
    a.shift = h->opt.mapaln ? 0.002f : 0.0003f;
    a.ssswgt = h->opt.ssswgt;
    a.S = S;
    HIPCHK(c2_launch_sdp_scores(a, (int)h->sdp_pairs.size(), stream ? (hipStream_t)stream : h->stream));
    return CBP_OK;
}

/* autotuner: times the phase-1 launch of query qidx over a small grid of
 * (segment columns, waves per workgroup) and keeps the fastest */
extern "C" int cbp_autotune(cbp_handle* h, int32_t qidx, int32_t* seg_cols, int32_t* nw, int32_t* rmax)
{
    C2ApiScope c2_scope_;
    if (!h) return CBP_ERR_ARG;
    if (!h->have_opt || !h->have_tables || !h->have_queries || !h->have_chunk)
        return fail(h, CBP_ERR_STATE, "options, tables, queries and chunk must be uploaded first");
    /* tile width (DB columns per wave segment), waves per workgroup and the
     * anti-diagonal depth (query rows per lane, at most rmax) */
    static const int segs[] = {2048, 4096, 8192, 16384};
    static const int nws[] = {4, 8};
    static const int rms[] = {5, 4, 3};
    /* the current tuning is the baseline; another geometry replaces it only
     * when faster by more than 1% (geometries that are the same launch for
     * this query, or within timing noise, keep the baseline, which the other
     * query lengths of a batch were tuned for) */
    auto time_it = [&](int sc, int rm, int nwv, double* t) {
        int rc = cbp_set_tuning(h, sc, rm, nwv);
        if (rc) return rc;
        *t = 1e30;
        for (int rep = 0; rep < 2; rep++) {
            if ((rc = c2_run_phase1(h, qidx))) return rc;
            if (h->last_ms < *t) *t = h->last_ms;
        }
        return CBP_OK;
    };
    int bs = h->seg_cols, bn = h->nw, br = h->rmax;
    double best = 1e30;
    int rc = time_it(bs, br, bn, &best);
    if (rc) return rc;
    for (int ri = 0; ri < 3; ri++)
        for (int si = 0; si < 4; si++)
            for (int ni = 0; ni < 2; ni++) {
                double t = 1e30;
                if ((rc = time_it(segs[si], rms[ri], nws[ni], &t))) return rc;
                if (t < 0.99 * best) { best = t; bs = segs[si]; bn = nws[ni]; br = rms[ri]; }
            }
    if ((rc = cbp_set_tuning(h, bs, br, bn))) return rc;
    if (seg_cols) *seg_cols = bs;
    if (nw) *nw = bn;
    if (rmax) *rmax = br;
    return CBP_OK;
}

/* chunk limits for a query length (CuDeviceMemory::CalcMaxDbDataChunkSizeHelper,
 * cupro/CuDeviceMemory.cu:334-... with GetTotalMemoryReqs,
 * CuDeviceMemory.cuh:716-760): the most DB positions (and profiles, at the
 * reference's expected profile length of 50) whose chunk data and phase-1
 * buffers, plus the pass-2 buffers for 10% of those positions (the
 * reference's DEV_PASS2MEMP default), fit the memory budget (or the device's
 * free memory when the budget is 0).  This build never materialises a score
 * matrix, so the per-position cost is the profile data itself, not
 * Lq x 4 bytes of scores. */
extern "C" int cbp_chunk_limits(cbp_handle* h, uint64_t qlen, uint64_t* maxpos, uint64_t* maxpros,
                                uint64_t* maxpos2, uint64_t* maxpros2)
{
    C2ApiScope c2_scope_;
    if (!h || !maxpos || !maxpros || !maxpos2 || !maxpros2 || qlen < 1 || qlen > 65535) return CBP_ERR_ARG;
    (void)hipSetDevice(h->device);
    size_t avail = 0;
    if (h->budget) {
        avail = h->budget > h->used ? h->budget - h->used : 0;
    } else {
        size_t fr = 0, tot = 0;
        HIPCHK(hipMemGetInfo(&fr, &tot));
        avail = fr;
    }
    const double expct_len = 50.0, pass2 = 0.10;
    int R = 1, ns = 1;
    c2_geometry((int)qlen, h->rmax, &R, &ns);
    int mR = 1, mns = 1, mnw = 1;
    c2_map_geometry((int)qlen, &mR, &mns, &mnw);
    /* phase 1, per DB position: 20 target frequencies, 3 SS probabilities,
     * 7 transitions, residue, SS state, profile index; the stripe hand-off
     * (5 floats) for multi-stripe queries; per profile: 20 background
     * probabilities, ENO, length, offset, the extra transition row, the
     * 64-bit maximum and the 20 score weights */
    const double per_pos = 4.0 * (20 + 3 + 7) + 1 + 1 + 4 + (ns > 1 ? 20.0 : 0.0);
    const double per_pro = 4.0 * (20 + 1 + 1 + 1 + 7 + 2 + 20);
    const double p1 = per_pos + per_pro / expct_len;
    /* pass 2, per DB position of a realigned pair: S', F, B (4 B each) and
     * MAP backtrack (1 B) over 64*R*stripes rows; the candidates' 2-bit
     * phase-1 backtrack codes (4 B per step and lane per stripe); hand-off
     * rows */
    const double rows_m = 64.0 * mR * mns;
    const double p2 = 13.0 * rows_m + 4.0 * 64.0 * ns + 4.0 * (5 * mns + 4);
    double n = (double)avail / (p1 + pass2 * p2);
    uint64_t mp = (uint64_t)n;
    mp &= ~(uint64_t)31;                          /* positions padded to 32, as the reference */
    *maxpos = mp;
    *maxpros = (uint64_t)((double)mp / expct_len);
    *maxpos2 = ((uint64_t)((double)mp * pass2)) & ~(uint64_t)31;
    *maxpros2 = (uint64_t)((double)*maxpos2 / expct_len);
    return CBP_OK;
}
