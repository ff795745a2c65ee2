/*
 * c2_stream.cpp -- whole-database search in chunks: the DB is cut into
 * chunks of whole profiles, each chunk is uploaded by a copy thread from
 * pinned staging buffers while the previous chunk is searched, and every
 * query's per-chunk hit lists are merged into one result over the whole DB.
 *
 * The reference's chunk loop: JobDispatcher reads the DB chunk by chunk
 * (cualn/JobDispatcher.cpp:1103-1134) through TdDataReader's read-ahead
 * (cupro/TdDataReader.cpp:575-692), DevCommThread runs every query against
 * each chunk (cualn/DevCommThread.cpp:296-356), and AlnWriter::MergeResults
 * (cualn/AlnWriter.cpp:218-267) merges the parts of a query's results by
 * log E with std::merge.  Here the parts of a query are its chunks in DB
 * order, each sorted by (log E, profile index), so a stable merge in chunk
 * order gives the (log E, whole-DB profile index) order of a one-chunk
 * search.  The cap on drawn pairs applies per chunk while a chunk is
 * searched (a pair of the whole DB's top N is in its chunk's top N); the
 * merge then keeps the drawings of the whole DB's first N selected pairs
 * and gives the other rows back their phase-1 fields, so the merged result
 * equals the search of the whole DB as one chunk.
 *
 * Staging: two pinned buffers and two device chunk sets.  The copy thread
 * packs chunk k's slices of the field-major DB arrays into staging buffer
 * k%2 and copies it to device set k%2 on its own stream; it waits until the
 * search of chunk k-2 (the set's previous user) has finished.  The search
 * thread waits for chunk k's copies, points the handle's chunk at set k%2
 * and runs the batch stage sequence.
 */
#include <cstring>
#include <algorithm>
#include <thread>
#include <mutex>
#include <condition_variable>
#include "c2_lib.h"

#define SCHK(call)                                                               \
    do {                                                                         \
        hipError_t e_ = (call);                                                  \
        if (e_ != hipSuccess)                                                    \
            return c2_fail(h, CBP_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

namespace {

struct Chunk { int32_t p0, p1; int64_t pos0, npos; };

/* byte offsets of the fields of a chunk in a staging buffer (256-aligned) */
struct StageLayout {
    size_t bkg, eno, len, dist, trg, ss, aa, sss, trans, addr, cv = 0, cvp = 0, cvn = 0, total;
    bool has_cv;
    StageLayout(int64_t np, int64_t ns, bool with_cv) : has_cv(with_cv) {
        size_t o = 0;
        auto take = [&](size_t n) { const size_t r = o; o = (o + n + 255) & ~(size_t)255; return r; };
        bkg = take(4 * 20 * np); eno = take(4 * np); len = take(4 * np); dist = take(4 * np);
        trg = take(4 * 20 * ns); ss = take(4 * 3 * ns); aa = take(ns); sss = take(ns);
        trans = take(4 * 7 * (ns + np)); addr = take(4 * ns);
        if (with_cv) { cv = take(4 * (size_t)CBP_NCVEL * ns); cvp = take(4 * ns); cvn = take(4 * ns); }
        total = o;
    }
};

/* device pointers of one chunk set, read by the copy thread */
struct SetPtrs {
    char* bkg; char* eno; char* len; char* dist; char* trg; char* ss; char* aa; char* sss; char* trans; char* addr;
    char* cv; char* cvp; char* cvn;
};

SetPtrs set_ptrs(DevSet& d)
{
    return SetPtrs{d.bkg.as<char>(), d.eno.as<char>(), d.len.as<char>(), d.dist.as<char>(), d.trgfrq.as<char>(),
                   d.ssprb.as<char>(), d.aa.as<char>(), d.sss.as<char>(), d.trans.as<char>(), d.addrpro.as<char>(),
                   d.cv.as<char>(), d.cvp.as<char>(), d.cvn.as<char>()};
}

int alloc_set(cbp_handle* h, DevSet& d, int64_t np, int64_t ns, bool with_cv)
{
    int rc;
    if ((rc = c2_dalloc(h, d.bkg, 4 * 20 * (size_t)np))) return rc;
    if ((rc = c2_dalloc(h, d.eno, 4 * (size_t)np))) return rc;
    if ((rc = c2_dalloc(h, d.len, 4 * (size_t)np))) return rc;
    if ((rc = c2_dalloc(h, d.dist, 4 * (size_t)np))) return rc;
    if ((rc = c2_dalloc(h, d.trgfrq, 4 * 20 * (size_t)ns))) return rc;
    if ((rc = c2_dalloc(h, d.ssprb, 4 * 3 * (size_t)ns))) return rc;
    if ((rc = c2_dalloc(h, d.aa, (size_t)ns))) return rc;
    if ((rc = c2_dalloc(h, d.sss, (size_t)ns))) return rc;
    if ((rc = c2_dalloc(h, d.trans, 4 * 7 * (size_t)(ns + np)))) return rc;
    if ((rc = c2_dalloc(h, d.addrpro, 4 * (size_t)ns))) return rc;
    d.has_cv = with_cv;
    if (with_cv) {
        if ((rc = c2_dalloc(h, d.cv, 4 * (size_t)CBP_NCVEL * ns))) return rc;
        if ((rc = c2_dalloc(h, d.cvp, 4 * (size_t)ns))) return rc;
        if ((rc = c2_dalloc(h, d.cvn, 4 * (size_t)ns))) return rc;
    }
    return CBP_OK;
}

/* chunk c's slices of the whole-DB arrays into a staging buffer; distances
 * re-based to the chunk's first position */
void pack_chunk(const cbp_profiles& db, const Chunk& c, char* st, const StageLayout& L)
{
    const int64_t NP = db.nprof, NS = db.npos, np = c.p1 - c.p0, ns = c.npos;
    for (int a = 0; a < 20; a++) {
        std::memcpy(st + L.bkg + 4 * a * np, db.bkg + a * NP + c.p0, 4 * np);
        std::memcpy(st + L.trg + 4 * a * ns, db.trgfrq + a * NS + c.pos0, 4 * ns);
    }
    std::memcpy(st + L.eno, db.eno + c.p0, 4 * np);
    std::memcpy(st + L.len, db.len + c.p0, 4 * np);
    uint32_t* dist = (uint32_t*)(st + L.dist);
    int32_t* addr = (int32_t*)(st + L.addr);
    for (int64_t p = 0; p < np; p++) {
        const uint32_t d0 = (uint32_t)(db.dist[c.p0 + p] - c.pos0);
        dist[p] = d0;
        std::fill(addr + d0, addr + d0 + db.len[c.p0 + p], (int32_t)p);
    }
    for (int j = 0; j < 3; j++) std::memcpy(st + L.ss + 4 * j * ns, db.ssprb + j * NS + c.pos0, 4 * ns);
    std::memcpy(st + L.aa, db.aa + c.pos0, ns);
    std::memcpy(st + L.sss, db.sss + c.pos0, ns);
    for (int k = 0; k < 7; k++)
        std::memcpy(st + L.trans + 4 * k * (ns + np), db.trans + k * (NS + NP) + c.pos0 + c.p0, 4 * (ns + np));
    if (L.has_cv) {
        for (int i = 0; i < CBP_NCVEL; i++) std::memcpy(st + L.cv + 4 * i * ns, db.cv + i * NS + c.pos0, 4 * ns);
        std::memcpy(st + L.cvp, db.cvprior + c.pos0, 4 * ns);
        std::memcpy(st + L.cvn, db.cvnorm2 + c.pos0, 4 * ns);
    }
}

hipError_t copy_chunk(const SetPtrs& d, const char* st, const StageLayout& L, int64_t np, int64_t ns, hipStream_t s)
{
    struct { char* dst; size_t off, n; } parts[] = {
        {d.bkg, L.bkg, 4 * 20 * (size_t)np}, {d.eno, L.eno, 4 * (size_t)np}, {d.len, L.len, 4 * (size_t)np},
        {d.dist, L.dist, 4 * (size_t)np}, {d.trg, L.trg, 4 * 20 * (size_t)ns}, {d.ss, L.ss, 4 * 3 * (size_t)ns},
        {d.aa, L.aa, (size_t)ns}, {d.sss, L.sss, (size_t)ns}, {d.trans, L.trans, 4 * 7 * (size_t)(ns + np)},
        {d.addr, L.addr, 4 * (size_t)ns},
        {d.cv, L.cv, L.has_cv ? 4 * (size_t)CBP_NCVEL * ns : 0}, {d.cvp, L.cvp, L.has_cv ? 4 * (size_t)ns : 0},
        {d.cvn, L.cvn, L.has_cv ? 4 * (size_t)ns : 0}};
    for (auto& p : parts) {
        if (p.n == 0) continue;
        const hipError_t e = hipMemcpyAsync(p.dst, st + p.off, p.n, hipMemcpyHostToDevice, s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

/* one chunk's results of one query, in chunk-local rows */
struct Part {
    int32_t p0 = 0;
    int32_t npassed = 0, nmap = 0;
    int64_t nposits = 0;
    int32_t maxlen = 0;
    std::vector<float> aln, epa, p1save;
    std::vector<int32_t> order, map_rows;
    std::vector<char> strs;
    int64_t stride = 0;
    float scorethld = 0.0f, logevthld = 0.0f;
    double sspace = 0.0;
};

void keep_part(const cbp_results& r, const QWork& w, int32_t p0, Part& P)
{
    P.p0 = p0;
    P.npassed = r.npassed; P.nmap = r.nmap; P.nposits = r.nposits; P.maxlen = r.maxlen;
    P.aln.assign(r.alndata, r.alndata + (size_t)r.npassed * CBP_NFIELDS);
    P.epa.assign(r.epa, r.epa + r.npassed);
    P.order.assign(r.order, r.order + r.npassed);
    P.map_rows.assign(r.map_rows, r.map_rows + r.nmap);
    P.p1save = w.p1save;
    P.stride = r.alnstride;
    P.strs.assign(r.alns, r.alns + (size_t)r.nmap * 3 * r.alnstride);
    P.scorethld = r.scorethld; P.logevthld = r.logevthld; P.sspace = r.search_space;
    /* profile numbers of the whole DB */
    for (int32_t k = 0; k < r.npassed; k++) {
        uint32_t u;
        std::memcpy(&u, &P.aln[(size_t)k * CBP_NFIELDS + CBP_F_ORGPRONO], 4);
        u += (uint32_t)p0;
        std::memcpy(&P.aln[(size_t)k * CBP_NFIELDS + CBP_F_ORGPRONO], &u, 4);
    }
}

float logE_of(const std::vector<float>& aln, int row) { return aln[(size_t)row * CBP_NFIELDS + CBP_F_EVALUE]; }

/* one query's parts merged into M and the caller's view out */
int merge_parts(cbp_handle* h, const cbp_profiles& db, std::vector<Part>& parts, int32_t qlen, StreamOut& M,
                cbp_results* out)
{
    /* rows of the whole DB in profile order: the parts one after another */
    std::vector<int32_t> roff(parts.size() + 1, 0);
    for (size_t i = 0; i < parts.size(); i++) roff[i + 1] = roff[i] + parts[i].npassed;
    const int32_t n = roff.back();
    M.aln.clear(); M.epa.clear();
    M.aln.reserve((size_t)n * CBP_NFIELDS); M.epa.reserve(n);
    for (const Part& P : parts) {
        M.aln.insert(M.aln.end(), P.aln.begin(), P.aln.end());
        M.epa.insert(M.epa.end(), P.epa.begin(), P.epa.end());
    }
    /* AlnWriter::MergeResults: std::merge of the sorted parts, in part order */
    M.order.clear();
    std::vector<int32_t> tmp;
    for (size_t i = 0; i < parts.size(); i++) {
        std::vector<int32_t> add(parts[i].order.size());
        for (size_t k = 0; k < add.size(); k++) add[k] = parts[i].order[k] + roff[i];
        tmp.resize(M.order.size() + add.size());
        std::merge(M.order.begin(), M.order.end(), add.begin(), add.end(), tmp.begin(),
                   [&](int32_t a, int32_t b) { return logE_of(M.aln, a) < logE_of(M.aln, b); });
        M.order.swap(tmp);
    }
    /* the drawn pairs of the whole DB: the parts' drawn rows in merged order
     * up to the cap (a row the whole DB selects is drawn in its part) */
    std::vector<int32_t> part_of(n), drawn_at(n, -1);
    for (size_t i = 0; i < parts.size(); i++) {
        for (int32_t k = 0; k < parts[i].npassed; k++) part_of[roff[i] + k] = (int32_t)i;
        for (int32_t j = 0; j < parts[i].nmap; j++) drawn_at[roff[i] + parts[i].map_rows[j]] = j;
    }
    const float logev = parts.empty() ? 0.0f : parts[0].logevthld;
    const int32_t cap = h->opt.max_map_pairs;
    M.map_rows.clear();
    for (int32_t g : M.order) {
        if (drawn_at[g] < 0 || !(logE_of(M.aln, g) <= logev)) continue;
        if (cap > 0 && (int32_t)M.map_rows.size() >= cap) break;
        M.map_rows.push_back(g);
    }
    std::vector<char> keep(n, 0);
    for (int32_t g : M.map_rows) keep[g] = 1;
    int maxld = 0;
    for (int32_t g = 0; g < n; g++) {
        if (drawn_at[g] < 0) continue;
        const Part& P = parts[part_of[g]];
        float* r = &M.aln[(size_t)g * CBP_NFIELDS];
        if (keep[g]) {
            uint32_t pro;
            std::memcpy(&pro, &r[CBP_F_ORGPRONO], 4);
            maxld = std::max(maxld, (int)db.len[pro]);
            continue;
        }
        /* drawn in its chunk only: the row as a one-chunk search leaves it */
        const int j = drawn_at[g];
        const uint32_t z = 0u;
        r[CBP_F_PSTVS] = P.p1save[3 * j];
        r[CBP_F_BEGCOORDS] = P.p1save[3 * j + 1];
        r[CBP_F_ENDCOORDS] = P.p1save[3 * j + 2];
        r[CBP_F_IDNTS] = 0.0f; r[CBP_F_NGAPS] = 0.0f;
        std::memcpy(&r[CBP_F_EPA], &z, 4);
    }
    /* positions of the passed profiles before each row (PRONEWDST) */
    int64_t nposits = 0;
    int32_t maxlen = 0;
    for (int32_t g = 0; g < n; g++) {
        float* r = &M.aln[(size_t)g * CBP_NFIELDS];
        uint32_t pro;
        std::memcpy(&pro, &r[CBP_F_ORGPRONO], 4);
        const uint32_t u = (uint32_t)nposits;
        std::memcpy(&r[CBP_F_PRONEWDST], &u, 4);
        nposits += db.len[pro];
        maxlen = std::max(maxlen, (int32_t)db.len[pro]);
    }
    const int nm = (int)M.map_rows.size();
    const int64_t stride = nm ? (int64_t)qlen + maxld + 1 : 0;
    M.strs.assign((size_t)nm * 3 * stride, 0);
    for (int k = 0; k < nm; k++) {
        const int32_t g = M.map_rows[k];
        const Part& P = parts[part_of[g]];
        const int j = drawn_at[g];
        for (int row = 0; row < 3; row++) {
            const char* src = &P.strs[((size_t)j * 3 + row) * P.stride];
            const size_t len = strnlen(src, (size_t)P.stride);
            std::memcpy(&M.strs[((size_t)k * 3 + row) * stride], src, std::min<size_t>(len, (size_t)stride - 1));
        }
    }
    std::memset(out, 0, sizeof *out);
    out->npassed = n;
    out->nposits = nposits;
    out->maxlen = maxlen;
    out->nmap = nm;
    out->alndata = n ? M.aln.data() : nullptr;
    out->order = n ? M.order.data() : nullptr;
    out->alns = nm ? M.strs.data() : nullptr;
    out->alnstride = stride;
    out->map_rows = nm ? M.map_rows.data() : nullptr;
    out->epa = n ? M.epa.data() : nullptr;
    if (!parts.empty()) {
        out->scorethld = parts[0].scorethld;
        out->logevthld = parts[0].logevthld;
        out->search_space = parts[0].sspace;
    }
    return CBP_OK;
}

int check_db(const cbp_profiles* s)
{
    if (!s || s->nprof < 1 || s->npos < 1 || !s->bkg || !s->eno || !s->len || !s->dist ||
        !s->trgfrq || !s->ssprb || !s->aa || !s->sss || !s->trans)
        return 0;
    if ((s->cv == nullptr) != (s->cvprior == nullptr) || (s->cv == nullptr) != (s->cvnorm2 == nullptr))
        return 0;
    int64_t acc = 0;
    for (int32_t p = 0; p < s->nprof; p++) {
        if (s->len[p] < 1 || s->len[p] > 65535 || (int64_t)s->dist[p] != acc) return 0;
        acc += s->len[p];
    }
    if (acc != s->npos) return 0;
    for (size_t i = 0; i < (size_t)3 * s->npos; i++)
        if (!(s->ssprb[i] >= 0.0f && s->ssprb[i] <= 1.0f)) return 0;
    return 1;
}

/* state shared by the search thread and the copy thread */
struct Shared {
    std::mutex mu;
    std::condition_variable cv;
    int issued = -1;        /* last chunk whose copies were issued */
    int done = -1;          /* last chunk whose search finished */
    bool stop = false;
    hipError_t err = hipSuccess;
};

}  // namespace

int c2_search_db(cbp_handle* h, const cbp_profiles* db, uint64_t chunk_pos, const int32_t* qidx, int32_t nq,
                 cbp_results* outs)
{
    if (!h->have_opt || !h->have_tables || !h->have_queries)
        return c2_fail(h, CBP_ERR_STATE, "options, tables and queries must be uploaded first");
    if (!check_db(db)) return c2_fail(h, CBP_ERR_ARG, "invalid database profile set");
    (void)hipSetDevice(h->device);
    int32_t qmax = 1;
    for (int32_t i = 0; i < nq; i++) {
        if (qidx[i] < 0 || qidx[i] >= h->queries.v.nprof) return c2_fail(h, CBP_ERR_ARG, "query index out of range");
        qmax = std::max(qmax, h->queries.v.len[qidx[i]]);
    }
    if (chunk_pos == 0) {
        uint64_t a, b, c, d;
        const int rc = cbp_chunk_limits(h, (uint64_t)qmax, &a, &b, &c, &d);
        if (rc) return rc;
        /* two chunk sets are resident: half the limit each */
        chunk_pos = a / 2;
    }
    /* chunks of whole profiles of at most chunk_pos positions */
    std::vector<Chunk> ch;
    {
        Chunk c{0, 0, 0, 0};
        for (int32_t p = 0; p < db->nprof; p++) {
            const int64_t L = db->len[p];
            if ((uint64_t)L > chunk_pos) return c2_fail(h, CBP_ERR_ARG, "a profile is longer than the chunk size");
            if ((uint64_t)(c.npos + L) > chunk_pos) {
                c.p1 = p; ch.push_back(c);
                c = Chunk{p, p, (int64_t)db->dist[p], 0};
            }
            c.npos += L;
        }
        c.p1 = db->nprof; ch.push_back(c);
    }
    int64_t mnp = 0, mns = 0;
    for (const Chunk& c : ch) { mnp = std::max<int64_t>(mnp, c.p1 - c.p0); mns = std::max(mns, c.npos); }
    int rc;
    /* two device chunk sets and two pinned staging buffers of the largest chunk */
    const bool with_cv = db->cv != nullptr;
    if ((rc = alloc_set(h, h->dsets[0], mnp, mns, with_cv))) return rc;
    if ((rc = alloc_set(h, h->dsets[1], mnp, mns, with_cv))) return rc;
    if ((rc = c2_dalloc(h, h->keys, sizeof(unsigned long long) * (size_t)mnp))) return rc;
    const StageLayout Lmax(mnp, mns, with_cv);
    for (int s = 0; s < 2; s++) {
        if (h->stage_n[s] < Lmax.total) {
            if (h->stage[s]) { (void)hipHostFree(h->stage[s]); h->stage[s] = nullptr; h->stage_n[s] = 0; }
            SCHK(hipHostMalloc((void**)&h->stage[s], Lmax.total, hipHostMallocDefault));
            h->stage_n[s] = Lmax.total;
        }
        if (!h->up_ev[s]) SCHK(hipEventCreateWithFlags(&h->up_ev[s], hipEventDisableTiming));
    }
    if (!h->ustream) SCHK(hipStreamCreate(&h->ustream));
    const SetPtrs sp[2] = {set_ptrs(h->dsets[0]), set_ptrs(h->dsets[1])};
    /* whole-DB size for thresholds and statistics */
    const uint64_t dblen = h->dblen ? h->dblen : (uint64_t)db->npos;
    const uint64_t nseqs = h->nseqs ? h->nseqs : (uint64_t)db->nprof;
    const uint64_t sdblen = h->dblen, snseqs = h->nseqs;
    h->dblen = dblen; h->nseqs = nseqs;
    std::vector<float> st(nq), lt(nq);
    for (int32_t i = 0; i < nq; i++)
        if ((rc = cbp_score_threshold(&h->opt, (uint64_t)h->queries.v.len[qidx[i]], dblen, &st[i], &lt[i])))
            return c2_fail(h, rc, "score threshold");
    /* the handle's own chunk (if any) is not used by this search */
    h->have_chunk = false;
    h->chunk_gen++;
    Shared sh;
    const int device = h->device;
    hipStream_t us = h->ustream;
    char* stage[2] = {h->stage[0], h->stage[1]};
    hipEvent_t upev[2] = {h->up_ev[0], h->up_ev[1]};
    std::thread copier([&, device, us]() {
        (void)hipSetDevice(device);
        for (size_t k = 0; k < ch.size(); k++) {
            {
                std::unique_lock<std::mutex> lk(sh.mu);
                sh.cv.wait(lk, [&] { return sh.stop || sh.done >= (int)k - 2; });
                if (sh.stop) return;
            }
            const int s = (int)(k & 1);
            const Chunk& c = ch[k];
            const StageLayout L(c.p1 - c.p0, c.npos, with_cv);
            /* the staging buffer's previous copies (chunk k-2) have completed:
             * the search of chunk k-2 waited for them */
            pack_chunk(*db, c, stage[s], L);
            hipError_t e = copy_chunk(sp[s], stage[s], L, c.p1 - c.p0, c.npos, us);
            if (e == hipSuccess) e = hipEventRecord(upev[s], us);
            {
                std::lock_guard<std::mutex> lk(sh.mu);
                if (e != hipSuccess) { sh.err = e; sh.stop = true; }
                sh.issued = (int)k;
            }
            sh.cv.notify_all();
            if (e != hipSuccess) return;
        }
    });
    /* per query, per chunk results */
    std::vector<std::vector<Part>> parts(nq, std::vector<Part>(ch.size()));
    std::vector<cbp_results> couts(nq);
    std::vector<int32_t> rebased;
    int frc = CBP_OK;
    for (size_t k = 0; k < ch.size() && frc == CBP_OK; k++) {
        {
            std::unique_lock<std::mutex> lk(sh.mu);
            sh.cv.wait(lk, [&] { return sh.stop || sh.issued >= (int)k; });
            if (sh.err != hipSuccess) { frc = c2_fail(h, CBP_ERR_HIP, std::string("chunk upload: ") + hipGetErrorString(sh.err)); break; }
        }
        const int s = (int)(k & 1);
        if (hipEventSynchronize(upev[s]) != hipSuccess) { frc = c2_fail(h, CBP_ERR_HIP, "chunk upload wait"); break; }
        const Chunk& c = ch[k];
        /* the handle's chunk is set s: its device arrays and a host view of
         * the chunk's slices of the caller's arrays */
        std::swap(h->dchunk, h->dsets[s]);
        h->dchunk.nprof = c.p1 - c.p0; h->dchunk.npos = c.npos;
        rebased.resize(c.p1 - c.p0);
        for (int32_t p = c.p0; p < c.p1; p++) rebased[p - c.p0] = (int32_t)(db->dist[p] - c.pos0);
        cbp_profiles v{};
        v.nprof = c.p1 - c.p0; v.npos = c.npos;
        v.bkg = nullptr; v.eno = db->eno + c.p0; v.len = db->len + c.p0; v.dist = (const uint32_t*)rebased.data();
        v.trgfrq = nullptr; v.ssprb = nullptr; v.aa = db->aa + c.pos0; v.sss = db->sss + c.pos0; v.trans = nullptr;
        h->chunk.v = v;
        h->chunk_gen++;
        frc = c2_build_segments(h, &h->chunk.v);
        if (frc == CBP_OK) {
            h->have_chunk = true;
            frc = c2_run_queries(h, qidx, nq, st.data(), lt.data(), couts.data());
            h->have_chunk = false;
        }
        if (frc == CBP_OK)
            for (int32_t i = 0; i < nq; i++) keep_part(couts[i], *h->work[i], c.p0, parts[i][k]);
        std::swap(h->dchunk, h->dsets[s]);
        {
            std::lock_guard<std::mutex> lk(sh.mu);
            sh.done = (int)k;
            if (frc != CBP_OK) sh.stop = true;
        }
        sh.cv.notify_all();
    }
    {
        std::lock_guard<std::mutex> lk(sh.mu);
        sh.stop = true;
    }
    sh.cv.notify_all();
    copier.join();
    (void)hipStreamSynchronize(us);
    /* the host view pointed into the caller's arrays */
    h->chunk.v = cbp_profiles{};
    h->dchunk.nprof = 0; h->dchunk.npos = 0;
    h->chunk_gen++;
    h->dblen = sdblen; h->nseqs = snseqs;
    if (frc != CBP_OK) return frc;
    h->smerged.resize(nq);
    for (int32_t i = 0; i < nq; i++)
        if ((rc = merge_parts(h, *db, parts[i], h->queries.v.len[qidx[i]], h->smerged[i], &outs[i]))) return rc;
    h->last_nchunks = (int32_t)ch.size();
    return CBP_OK;
}

extern "C" int cbp_search_db(cbp_handle* h, const cbp_profiles* db, uint64_t chunk_pos, const int32_t* qidx,
                             int32_t nq, cbp_results* outs)
{
    C2ApiScope c2_scope_;
    if (!h || !db || !qidx || !outs || nq < 1) return CBP_ERR_ARG;
    try {
        return c2_search_db(h, db, chunk_pos, qidx, nq, outs);
    } catch (const std::bad_alloc&) {
        return c2_fail(h, CBP_ERR_MEM, "host memory allocation failed");
    }
}

extern "C" int cbp_last_search_chunks(const cbp_handle* h, int32_t* nchunks)
{
    C2ApiScope c2_scope_;
    if (!h || !nchunks) return CBP_ERR_ARG;
    *nchunks = h->last_nchunks;
    return CBP_OK;
}
