/*
 * c2_fields.cpp -- the reference's per-field profile buffers as the C ABI's
 * profile sets (host only).
 *
 * The reference passes profile data as an array of pmv2DTotFlds = 81 field
 * pointers (cupro/PM2DVectorFields.h:41-81): for a DB chunk, begin/end
 * pointers into one packed host buffer (cupro/CuDeviceMemory.cu:1036-1104,
 * szCpm2dvfields holding the field offsets); for a query, the begin pointers
 * advanced to the query by PMDataSkipNPros (cualn/DevCommThread.cpp:315-321).
 * Per profile: 20 background probabilities, ENO (float), length (int),
 * distance = positions before the profile (unsigned).  Per profile and
 * position plus one beginning row: 7 log transitions.  Per position: 20
 * target frequencies, 19 context-vector entries, CV prior and squared norm,
 * 3 SS probabilities, HDP probability and index, three addresses (int), the
 * residue (char) and the SS state (char).  The fields this path reads are
 * gathered into the field-major arrays of cbp_profiles; the CV, HDP and
 * address fields may be absent (empty) or present; the CV fields are read
 * when present (the context-vector score term), the HDP and address fields
 * are not.
 */
#include <cstring>
#include <vector>
#include "c2_lib.h"

namespace {

enum {
    F_BKG = 0, F_ENO = 20, F_LEN = 21, F_DIST = 22, F_TRN = 23, F_TRG = 30, F_CV = 50,
    F_CVPRIOR = 69, F_CVNORM = 70, F_SS = 71, F_HDPPRB = 74, F_HDPIND = 75, F_ADDRPRO = 76,
    F_ADDRCV = 77, F_ADDRSS = 78, F_AA = 79, F_SSSTATE = 80, F_TOTAL = 81
};
static_assert(F_TOTAL == CBP_PM_NFIELDS, "field count");

size_t fbytes(const void* const* beg, const void* const* end, int f)
{
    if (!beg[f] || !end[f]) return 0;
    const char* b = static_cast<const char*>(beg[f]);
    const char* e = static_cast<const char*>(end[f]);
    return e >= b ? (size_t)(e - b) : (size_t)-1;
}

}  // namespace

extern "C" int cbp_fields_count(const void* const* beg, const void* const* end, int32_t* nprof, int64_t* npos)
{
    C2ApiScope c2_scope_;
    if (!beg || !end || !nprof || !npos) return CBP_ERR_ARG;
    const size_t np = fbytes(beg, end, F_LEN) / sizeof(int32_t);
    const size_t ns = fbytes(beg, end, F_AA);
    if (np < 1 || ns < 1 || np > 0x7fffffff) return CBP_ERR_ARG;
    for (int f = 0; f < F_TOTAL; f++) {
        const size_t b = fbytes(beg, end, f);
        if (b == (size_t)-1) return CBP_ERR_ARG;
        size_t want;
        if (f < F_TRN) want = 4 * np;                              /* per profile */
        else if (f < F_TRG) want = 4 * (ns + np);                  /* transitions */
        else if (f == F_AA || f == F_SSSTATE) want = ns;           /* chars */
        else want = 4 * ns;                                        /* per position */
        const bool optional = (f >= F_CV && f <= F_CVNORM) || (f >= F_HDPPRB && f <= F_ADDRSS);
        if (b != want && !(optional && b == 0)) return CBP_ERR_ARG;
    }
    *nprof = (int32_t)np;
    *npos = (int64_t)ns;
    return CBP_OK;
}

extern "C" int cbp_fields_pack(const void* const* beg, const void* const* end,
                               float* bkg, float* eno, int32_t* len, uint32_t* dist,
                               float* trgfrq, float* ssprb, char* aa, char* sss, float* trans)
{
    C2ApiScope c2_scope_;
    int32_t np32 = 0;
    int64_t ns64 = 0;
    int rc = cbp_fields_count(beg, end, &np32, &ns64);
    if (rc) return rc;
    if (!bkg || !eno || !len || !dist || !trgfrq || !ssprb || !aa || !sss || !trans) return CBP_ERR_ARG;
    const size_t np = (size_t)np32, ns = (size_t)ns64;
    for (int a = 0; a < 20; a++) std::memcpy(bkg + a * np, beg[F_BKG + a], 4 * np);
    std::memcpy(eno, beg[F_ENO], 4 * np);
    std::memcpy(len, beg[F_LEN], 4 * np);
    /* distances re-based to the first profile: a chunk carries them from 0,
     * a query advanced into its batch from the query's offset */
    const uint32_t* d = static_cast<const uint32_t*>(beg[F_DIST]);
    uint64_t acc = 0;
    for (size_t p = 0; p < np; p++) {
        dist[p] = d[p] - d[0];
        if (dist[p] != acc || len[p] < 1) return CBP_ERR_ARG;
        acc += (uint64_t)len[p];
    }
    if (acc != ns) return CBP_ERR_ARG;
    for (int k = 0; k < 7; k++) std::memcpy(trans + k * (ns + np), beg[F_TRN + k], 4 * (ns + np));
    for (int a = 0; a < 20; a++) std::memcpy(trgfrq + a * ns, beg[F_TRG + a], 4 * ns);
    for (int s = 0; s < 3; s++) std::memcpy(ssprb + s * ns, beg[F_SS + s], 4 * ns);
    std::memcpy(aa, beg[F_AA], ns);
    std::memcpy(sss, beg[F_SSSTATE], ns);
    return CBP_OK;
}

extern "C" int cbp_fields_pack_cv(const void* const* beg, const void* const* end,
                                  float* cv, float* cvprior, float* cvnorm2)
{
    C2ApiScope c2_scope_;
    int32_t np32 = 0;
    int64_t ns64 = 0;
    int rc = cbp_fields_count(beg, end, &np32, &ns64);
    if (rc) return rc;
    const size_t ns = (size_t)ns64;
    int present = 0;
    for (int f = F_CV; f <= F_CVNORM; f++) present += fbytes(beg, end, f) != 0;
    if (present == 0) return 0;
    if (present != F_CVNORM - F_CV + 1) return CBP_ERR_ARG;
    if (!cv || !cvprior || !cvnorm2) return CBP_ERR_ARG;
    for (int i = 0; i < CBP_NCVEL; i++) std::memcpy(cv + i * ns, beg[F_CV + i], 4 * ns);
    std::memcpy(cvprior, beg[F_CVPRIOR], 4 * ns);
    std::memcpy(cvnorm2, beg[F_CVNORM], 4 * ns);
    return 1;
}

/* field-major copy of a field set, owned */
struct FieldSet {
    std::vector<float> bkg, eno, trgfrq, ssprb, trans;
    std::vector<float> cv, cvprior, cvnorm2;
    std::vector<int32_t> len;
    std::vector<uint32_t> dist;
    std::vector<char> aa, sss;
    cbp_profiles v{};
    int load(const void* const* beg, const void* const* end) {
        int32_t np = 0; int64_t ns = 0;
        int rc = cbp_fields_count(beg, end, &np, &ns);
        if (rc) return rc;
        bkg.resize((size_t)20 * np); eno.resize(np); len.resize(np); dist.resize(np);
        trgfrq.resize((size_t)20 * ns); ssprb.resize((size_t)3 * ns); aa.resize(ns); sss.resize(ns);
        trans.resize((size_t)7 * (ns + np));
        rc = cbp_fields_pack(beg, end, bkg.data(), eno.data(), len.data(), dist.data(), trgfrq.data(),
                             ssprb.data(), aa.data(), sss.data(), trans.data());
        if (rc) return rc;
        cv.resize((size_t)CBP_NCVEL * ns); cvprior.resize(ns); cvnorm2.resize(ns);
        rc = cbp_fields_pack_cv(beg, end, cv.data(), cvprior.data(), cvnorm2.data());
        if (rc < 0) return rc;
        if (rc == 0) { cv.clear(); cvprior.clear(); cvnorm2.clear(); }
        v.nprof = np; v.npos = ns;
        v.bkg = bkg.data(); v.eno = eno.data(); v.len = len.data(); v.dist = dist.data();
        v.trgfrq = trgfrq.data(); v.ssprb = ssprb.data(); v.aa = aa.data(); v.sss = sss.data();
        v.trans = trans.data();
        v.cv = cv.empty() ? nullptr : cv.data();
        v.cvprior = cv.empty() ? nullptr : cvprior.data();
        v.cvnorm2 = cv.empty() ? nullptr : cvnorm2.data();
        return CBP_OK;
    }
};

extern "C" int cbp_upload_chunk_fields(cbp_handle* h, const void* const* beg, const void* const* end)
{
    C2ApiScope c2_scope_;
    if (!h) return CBP_ERR_ARG;
    try {
        FieldSet fs;
        const int rc = fs.load(beg, end);
        if (rc) return c2_fail(h, rc, "invalid profile field buffers");
        return cbp_upload_chunk(h, &fs.v);
    } catch (const std::bad_alloc&) {
        return c2_fail(h, CBP_ERR_MEM, "host memory allocation failed");
    }
}

/* one query given by its field begin pointers (the reference's querypmbeg,
 * advanced to the query) and its length */
extern "C" int cbp_add_query_fields(cbp_handle* h, const void* const* beg, uint64_t nqyposs, int32_t* qidx)
{
    C2ApiScope c2_scope_;
    if (!h || !beg || !qidx || nqyposs < 1 || nqyposs > 65535) return CBP_ERR_ARG;
    const char* end[CBP_PM_NFIELDS];
    for (int f = 0; f < CBP_PM_NFIELDS; f++) {
        const char* b = static_cast<const char*>(beg[f]);
        size_t n;
        if (f < F_TRN) n = 4;
        else if (f < F_TRG) n = 4 * (nqyposs + 1);
        else if (f == F_AA || f == F_SSSTATE) n = nqyposs;
        else n = 4 * nqyposs;
        end[f] = b ? b + n : nullptr;
    }
    try {
        FieldSet one;
        int rc = one.load(beg, reinterpret_cast<const void* const*>(end));
        if (rc) return c2_fail(h, rc, "invalid query field buffers");
        if ((uint64_t)one.len[0] != nqyposs) return c2_fail(h, CBP_ERR_ARG, "query length differs from nqyposs");
        return c2_append_query(h, &one.v, qidx);
    } catch (const std::bad_alloc&) {
        return c2_fail(h, CBP_ERR_MEM, "host memory allocation failed");
    }
}

/* a DB made of a cached part and a chunk, searched as one set, the cached
 * part's profiles first */
extern "C" int cbp_upload_db_fields(cbp_handle* h, const void* const* beg1, const void* const* end1,
                                    const void* const* beg2, const void* const* end2)
{
    C2ApiScope c2_scope_;
    if (!h) return CBP_ERR_ARG;
    if (!beg1 || !end1) return cbp_upload_chunk_fields(h, beg2, end2);
    if (!beg2 || !end2) return cbp_upload_chunk_fields(h, beg1, end1);
    try {
        FieldSet a, b;
        int rc = a.load(beg1, end1);
        if (!rc) rc = b.load(beg2, end2);
        if (rc) return c2_fail(h, rc, "invalid profile field buffers");
        const size_t na = a.len.size(), nb = b.len.size(), sa = a.aa.size(), sb = b.aa.size();
        const size_t np = na + nb, ns = sa + sb;
        FieldSet c;
        c.bkg.resize(20 * np); c.trgfrq.resize(20 * ns); c.ssprb.resize(3 * ns); c.trans.resize(7 * (ns + np));
        for (size_t k = 0; k < 20; k++) {
            std::copy(a.bkg.begin() + k * na, a.bkg.begin() + (k + 1) * na, c.bkg.begin() + k * np);
            std::copy(b.bkg.begin() + k * nb, b.bkg.begin() + (k + 1) * nb, c.bkg.begin() + k * np + na);
            std::copy(a.trgfrq.begin() + k * sa, a.trgfrq.begin() + (k + 1) * sa, c.trgfrq.begin() + k * ns);
            std::copy(b.trgfrq.begin() + k * sb, b.trgfrq.begin() + (k + 1) * sb, c.trgfrq.begin() + k * ns + sa);
        }
        for (size_t k = 0; k < 3; k++) {
            std::copy(a.ssprb.begin() + k * sa, a.ssprb.begin() + (k + 1) * sa, c.ssprb.begin() + k * ns);
            std::copy(b.ssprb.begin() + k * sb, b.ssprb.begin() + (k + 1) * sb, c.ssprb.begin() + k * ns + sa);
        }
        for (size_t k = 0; k < 7; k++) {
            std::copy(a.trans.begin() + k * (sa + na), a.trans.begin() + (k + 1) * (sa + na),
                      c.trans.begin() + k * (ns + np));
            std::copy(b.trans.begin() + k * (sb + nb), b.trans.begin() + (k + 1) * (sb + nb),
                      c.trans.begin() + k * (ns + np) + sa + na);
        }
        c.eno = a.eno; c.eno.insert(c.eno.end(), b.eno.begin(), b.eno.end());
        c.len = a.len; c.len.insert(c.len.end(), b.len.begin(), b.len.end());
        c.dist = a.dist;
        for (size_t p = 0; p < nb; p++) c.dist.push_back(b.dist[p] + (uint32_t)sa);
        c.aa = a.aa; c.aa.insert(c.aa.end(), b.aa.begin(), b.aa.end());
        c.sss = a.sss; c.sss.insert(c.sss.end(), b.sss.begin(), b.sss.end());
        c.v.nprof = (int32_t)np; c.v.npos = (int64_t)ns;
        c.v.bkg = c.bkg.data(); c.v.eno = c.eno.data(); c.v.len = c.len.data(); c.v.dist = c.dist.data();
        c.v.trgfrq = c.trgfrq.data(); c.v.ssprb = c.ssprb.data(); c.v.aa = c.aa.data(); c.v.sss = c.sss.data();
        c.v.trans = c.trans.data();
        if (!a.cv.empty() || !b.cv.empty()) {
            /* a part without context vectors takes the uninformative vector */
            c.cv.assign((size_t)CBP_NCVEL * ns, C2_CV_UNINF);
            c.cvprior.assign(ns, 0.0f);
            c.cvnorm2.assign(ns, 0.0f);
            for (size_t k = 0; k < (size_t)CBP_NCVEL; k++) {
                if (!a.cv.empty()) std::copy(a.cv.begin() + k * sa, a.cv.begin() + (k + 1) * sa, c.cv.begin() + k * ns);
                if (!b.cv.empty()) std::copy(b.cv.begin() + k * sb, b.cv.begin() + (k + 1) * sb, c.cv.begin() + k * ns + sa);
            }
            if (!a.cv.empty()) {
                std::copy(a.cvprior.begin(), a.cvprior.end(), c.cvprior.begin());
                std::copy(a.cvnorm2.begin(), a.cvnorm2.end(), c.cvnorm2.begin());
            }
            if (!b.cv.empty()) {
                std::copy(b.cvprior.begin(), b.cvprior.end(), c.cvprior.begin() + sa);
                std::copy(b.cvnorm2.begin(), b.cvnorm2.end(), c.cvnorm2.begin() + sa);
            }
            c.v.cv = c.cv.data(); c.v.cvprior = c.cvprior.data(); c.v.cvnorm2 = c.cvnorm2.data();
        }
        return cbp_upload_chunk(h, &c.v);
    } catch (const std::bad_alloc&) {
        return c2_fail(h, CBP_ERR_MEM, "host memory allocation failed");
    }
}
