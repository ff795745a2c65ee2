/*
 * adapter_check.cpp -- test driver for the C++ drop-in
 * comer2hip::HipBatchProcessing (include/hip_batch_processing.h).
 *
 * TEST INFRASTRUCTURE: built into tests/libadapter_check.so and loaded by
 * the tests only.  It lays profile sets out the way the reference hands them
 * to CuBatchProcessing -- one packed host buffer, one field after another
 * at 32-byte aligned offsets (szCpm2dvfields holds the offsets and, at
 * index pmv2DTotFlds, the total; cupro/CuDeviceMemory.cu:1036-1104,
 * cupro/PM2DVectorFields.h:41-81) -- and drives the class as DevCommThread
 * does (cualn/DevCommThread.cpp:296-356): TransferCPMData once per chunk,
 * then per query SetQueryLen, SetScoreThreshold, SetLogEThreshold and
 * ProcessScoreMatrix with the query's field pointers advanced into the
 * query batch (PMBatchProData::PMDataSkipNPros).
 */
#include <cstdlib>
#include <cstring>
#include <vector>
#include "../include/hip_batch_processing.h"

namespace {

constexpr int NF = CBP_PM_NFIELDS;
enum { F_BKG = 0, F_ENO = 20, F_LEN = 21, F_DIST = 22, F_TRN = 23, F_TRG = 30, F_CV = 50,
       F_SS = 71, F_HDPPRB = 74, F_ADDRSS = 78, F_AA = 79, F_SSSTATE = 80 };

bool optional_field(int f) { return (f >= F_CV && f < F_SS) || (f >= F_HDPPRB && f <= F_ADDRSS); }

/* bytes of field f for np profiles and ns positions */
size_t field_bytes(int f, size_t np, size_t ns)
{
    if (f < F_TRN) return 4 * np;
    if (f < F_TRG) return 4 * (ns + np);
    if (f == F_AA || f == F_SSSTATE) return ns;
    return 4 * ns;
}

/* profile set s packed field after field; optional fields (CV, HDP,
 * addresses) present and zero-filled when with_optional, else empty */
char* layout(const cbp_profiles* s, int with_optional, size_t* sz, char** beg, char** end)
{
    const size_t np = (size_t)s->nprof, ns = (size_t)s->npos;
    size_t off = 0;
    size_t fb[NF];
    for (int f = 0; f < NF; f++) {
        fb[f] = (optional_field(f) && !with_optional) ? 0 : field_bytes(f, np, ns);
        sz[f] = off;
        off += (fb[f] + 31) & ~(size_t)31;
    }
    sz[NF] = off;
    char* buf = (char*)std::calloc(off ? off : 1, 1);
    if (!buf) return nullptr;
    for (int f = 0; f < NF; f++) { beg[f] = buf + sz[f]; end[f] = beg[f] + fb[f]; }
    for (int a = 0; a < 20; a++) std::memcpy(beg[F_BKG + a], s->bkg + a * np, 4 * np);
    std::memcpy(beg[F_ENO], s->eno, 4 * np);
    std::memcpy(beg[F_LEN], s->len, 4 * np);
    std::memcpy(beg[F_DIST], s->dist, 4 * np);
    for (int k = 0; k < 7; k++) std::memcpy(beg[F_TRN + k], s->trans + k * (ns + np), 4 * (ns + np));
    for (int a = 0; a < 20; a++) std::memcpy(beg[F_TRG + a], s->trgfrq + a * ns, 4 * ns);
    for (int k = 0; k < 3; k++) std::memcpy(beg[F_SS + k], s->ssprb + k * ns, 4 * ns);
    std::memcpy(beg[F_AA], s->aa, ns);
    std::memcpy(beg[F_SSSTATE], s->sss, ns);
    /* the profile-address field of the reference: index of each position's profile */
    if (with_optional)
        for (size_t p = 0; p < np; p++)
            for (int i = 0; i < s->len[p]; i++)
                ((int32_t*)beg[76])[s->dist[p] + i] = (int32_t)p;
    return buf;
}

/* field pointers advanced past n profiles (PMBatchProData::PMDataSkipNPros) */
void skip_pros(char** pm, const cbp_profiles* s, int from, int n)
{
    size_t npos = 0;
    for (int p = from; p < from + n; p++) npos += (size_t)s->len[p];
    for (int f = 0; f < NF; f++) {
        if (f < F_TRN) pm[f] += 4 * (size_t)n;
        else if (f < F_TRG) pm[f] += 4 * (npos + n);
        else if (f == F_AA || f == F_SSSTATE) pm[f] += npos;
        else if (!optional_field(f)) pm[f] += 4 * npos;
    }
}

/* profiles [p0, p0+n) of s as a set of its own, distances from 0 */
struct Part {
    std::vector<float> bkg, eno, trg, ss, tr;
    std::vector<int32_t> len;
    std::vector<uint32_t> dist;
    std::vector<char> aa, sss;
    cbp_profiles v{};
    Part(const cbp_profiles* s, int p0, int n) {
        const size_t np = s->nprof, ns = s->npos;
        const size_t a0 = s->dist[p0];
        size_t L = 0;
        for (int p = p0; p < p0 + n; p++) L += (size_t)s->len[p];
        bkg.resize(20 * (size_t)n); trg.resize(20 * L); ss.resize(3 * L); tr.resize(7 * (L + n));
        for (int a = 0; a < 20; a++) {
            std::memcpy(&bkg[a * (size_t)n], s->bkg + a * np + p0, 4 * (size_t)n);
            std::memcpy(&trg[a * L], s->trgfrq + a * ns + a0, 4 * L);
        }
        for (int k = 0; k < 3; k++) std::memcpy(&ss[k * L], s->ssprb + k * ns + a0, 4 * L);
        for (int k = 0; k < 7; k++) std::memcpy(&tr[k * (L + n)], s->trans + k * (ns + np) + a0 + p0, 4 * (L + n));
        eno.assign(s->eno + p0, s->eno + p0 + n);
        len.assign(s->len + p0, s->len + p0 + n);
        for (int p = p0; p < p0 + n; p++) dist.push_back((uint32_t)(s->dist[p] - a0));
        aa.assign(s->aa + a0, s->aa + a0 + L);
        sss.assign(s->sss + a0, s->sss + a0 + L);
        v.nprof = n; v.npos = (int64_t)L;
        v.bkg = bkg.data(); v.eno = eno.data(); v.len = len.data(); v.dist = dist.data();
        v.trgfrq = trg.data(); v.ssprb = ss.data(); v.aa = aa.data(); v.sss = sss.data(); v.trans = tr.data();
    }
};

struct Counter {
    int value = 0;
    void dec() { value--; }
};

}  // namespace

extern "C" {

/* the reference's packed layout of s; sz[NF+1], beg[NF], end[NF]; free the
 * returned buffer with c2t_free */
char* c2t_layout(const cbp_profiles* s, int with_optional, size_t* sz, char** beg, char** end)
{
    return layout(s, with_optional, sz, beg, end);
}
void c2t_free(char* buf) { std::free(buf); }

/* query q of the batch: begin pointers advanced to it */
void c2t_query_fields(const cbp_profiles* qs, char** batch_beg, int q, char** out)
{
    for (int f = 0; f < NF; f++) out[f] = batch_beg[f];
    skip_pros(out, qs, 0, q);
}

typedef void (*c2t_sink)(int qrysernr, const cbp_results* res, int ndesc);

/* DevCommThread's loop over the queries of a batch against one DB given as
 * a cached part (its first nsplit profiles; 0 = none) and a chunk (the
 * rest); thresholds per query; the sink receives every query's results;
 * *decs = how many times the chunk counter was released.  Returns 0 or the
 * error code thrown. */
int c2t_run(const cbp_profiles* qs, const cbp_profiles* db, int nsplit, const cbp_options* opt,
            const float* ssss900, const float* nnw93, const int32_t* qorder, int nq,
            const float* scorethld, const float* logevthld, c2t_sink sink, int* decs)
{
    try {
        comer2hip::HipBatchProcessing cbpc(0, 0, 0, *opt, ssss900, nnw93);
        size_t qsz[NF + 1], csz[NF + 1], bsz[NF + 1];
        char *qbeg[NF], *qend[NF], *cbeg[NF], *cend[NF], *bbeg[NF], *bend[NF];
        char* qbuf = layout(qs, 0, qsz, qbeg, qend);
        Part bdb1(db, 0, nsplit > 0 ? nsplit : 1);
        Part bdbC(db, nsplit, db->nprof - nsplit);
        char* cbuf = layout(&bdbC.v, 1, csz, cbeg, cend);
        char* bbuf = nsplit > 0 ? layout(&bdb1.v, 1, bsz, bbeg, bend) : nullptr;
        std::vector<const char*> desc(db->nprof - nsplit, "synthetic");
        Counter cnt;
        int rc = 0;
        try {
            cbpc.SetResultsSink([&](int qn, const cbp_results& r, const char** d) {
                sink(qn, &r, d == desc.data() ? (int)desc.size() : -1);
            });
            cbpc.SetDbDetails((uint64_t)db->npos, (size_t)db->nprof);
            cbpc.TransferCPMData(cbeg, cend, csz);
            for (int i = 0; i < nq; i++) {
                const int q = qorder[i];
                char* qpm[NF];
                c2t_query_fields(qs, qbeg, q, qpm);
                const size_t qlen = (size_t)qs->len[q];
                cbpc.SetQueryLen(qlen);
                cbpc.SetScoreThreshold(scorethld[i]);
                cbpc.SetLogEThreshold(logevthld[i]);
                cbpc.ProcessScoreMatrix(q, qlen, qpm,
                                        nsplit > 0 ? bbeg : nullptr, nsplit > 0 ? bend : nullptr,
                                        desc.data(), cbeg, cend, csz, &cnt);
            }
        } catch (const comer2hip::Error& e) {
            rc = e.code();
        }
        std::free(qbuf); std::free(cbuf); std::free(bbuf);
        *decs = -cnt.value;
        return rc;
    } catch (const comer2hip::Error& e) {
        return e.code();
    }
}

}  /* extern "C" */
