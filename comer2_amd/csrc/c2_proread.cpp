/*
 * c2_proread.cpp -- COMER text profiles (format "COMER profile v2.0", the
 * condensed form makepro writes: TextWriteProfileCondensed,
 * libpro/srcpro/PMProfileModel.cpp:2295-2510) read into the field-major
 * profile sets of the C ABI, as the device path's reader fills its profile
 * buffers (TextReadProfileHeader / TextReadProfileData,
 * cupro/IOProfileModel.cu:46-1021).  Host only.
 *
 * Per profile the reader keeps what the search path reads: background
 * probabilities (NULL line), ENO (EFFNOS), length, the 7 used log
 * transitions (MM MI MD IM II DM DD of the file's 9, zero as -32768, IM..DD
 * scaled by 0.4) for the beginning row and every position, the residue, the
 * 20 target probabilities, the context vector (CV line; the uninformative
 * vector when the profile has none), the 3 SS probabilities and the SS
 * state (SS line; zeros and ' ' when absent); plus the description and file
 * name for the output.  HDP cluster data, observed counts and the profile's
 * statistical parameters are checked for presence and skipped, as there.
 */
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include "c2_lib.h"

struct cbp_proset {
    std::vector<float> bkg, eno, trgfrq, ssprb, trans, cv, cvprior, cvnorm2;
    std::vector<int32_t> len;
    std::vector<uint32_t> dist;
    std::vector<char> aa, sss;
    std::vector<std::string> desc, name;
    std::vector<char> has_cv;           /* per profile */
    cbp_profiles v{};
    std::string err;
    /* per-profile staging before the field-major layout is built */
    struct One {
        std::string desc, name;
        int32_t L = 0;
        float eno = 0.0f;
        float bkg[20];
        std::vector<float> trg, tr, cv, cvp, cvn, ss;   /* [L][20], [L+1][7], [L][19], [L], [L], [L][3] */
        std::vector<char> aa, sss;
        bool has_cv = false;
    };
    std::vector<One> pro;
    void build();
};

namespace {

constexpr float kUninf = 0.17302947f;   /* IOProfileModel.cu:704 */
constexpr int kMaxColumns = 50000;      /* MAXCOLUMNS, liblib/mylimits.h:50 */
constexpr int kNStates = 9;             /* P_MM .. P_DD (libpro/srcpro/TRANSPROBS.h:14-19) */
constexpr int kMaxPosteriors = 1000;

struct Reader {
    FILE* fp;
    std::string path;
    long lineno = 0;
    std::string err;
    explicit Reader(FILE* f, const char* p) : fp(f), path(p) {}
    /* the next line that is not blank or a '#' comment (skip_comments,
     * liblib/myfiler.cpp:131-185) */
    bool next(std::string& line) {
        char buf[8192];
        while (std::fgets(buf, sizeof buf, fp)) {
            std::string l(buf);
            while (!l.empty() && l.back() != '\n' && std::fgets(buf, sizeof buf, fp)) l += buf;
            lineno++;
            size_t i = 0;
            while (i < l.size() && (l[i] == ' ' || l[i] == '\t')) i++;
            if (i < l.size() && l[i] != '#' && l[i] != '\n' && l[i] != '\r') { line = l; return true; }
        }
        return false;
    }
    bool fail(const std::string& m) {
        if (err.empty()) err = path + ":" + std::to_string(lineno) + ": " + m;
        return false;
    }
};

/* read_integer / read_float / read_symbol (liblib/myfiler.cpp:324-512):
 * a token after blanks, ended by a blank, ';', a parenthesis or the line end
 * (read_float also ends at ','); p moves past it */
bool token(const char*& p, bool comma_ends, std::string& tok)
{
    while (*p == ' ' || *p == '\t') p++;
    const char* b = p;
    while (*p && *p != ' ' && *p != '\t' && *p != ';' && *p != '(' && *p != ')' && *p != '\n' && *p != '\r' &&
           !(comma_ends && *p == ','))
        p++;
    tok.assign(b, p);
    return p != b;
}
bool read_int(const char*& p, int* v)
{
    std::string t;
    if (!token(p, false, t)) return false;
    char* e = nullptr;
    errno = 0;
    const long x = std::strtol(t.c_str(), &e, 10);
    if (errno || *e || x < -2147483647L - 1 || x > 2147483647L) return false;
    *v = (int)x;
    return true;
}
bool read_flt(const char*& p, float* v)
{
    std::string t;
    if (!token(p, true, t)) return false;
    char* e = nullptr;
    errno = 0;
    *v = std::strtof(t.c_str(), &e);
    return errno == 0 && *e == 0;
}
bool read_sym(const char*& p, char* c)
{
    while (*p == ' ' || *p == '\t') p++;
    const char* b = p;
    while (*p && *p != ' ' && *p != '\t' && *p != '\n' && *p != '\r') p++;
    if (p - b != 1) return false;
    *c = *b;
    return true;
}
bool at_end(const char* p) { return *p == 0; }

/* the 7 used transitions of the 9 on a line (P_ID and P_DI skipped), as
 * logs: IM, II, DM, DD (after P_MD) scaled by 0.4, zero as -32768 */
bool read_trans(Reader& r, const std::string& line, int scale, float* out)
{
    const char* p = line.c_str();
    for (int t = 0, tt = 0; t < kNStates; t++) {
        int iv;
        if (at_end(p) || !read_int(p, &iv)) return r.fail("incomplete transition probability data");
        const float value = (float)iv / (float)scale;
        if (t == 5 || t == 7) continue;
        out[tt++] = value != 0.0f ? (2 < t ? logf(value) * 0.4f : logf(value)) : -32768.0f;
    }
    return true;
}

bool read_profile(Reader& r, cbp_proset::One& P, bool* eof)
{
    std::string line;
    const char* p;
    *eof = false;
    /* header (TextReadProfileHeader, IOProfileModel.cu:46-277) */
    if (!r.next(line)) { *eof = true; return false; }
    if (!(p = std::strstr(line.c_str(), "COMER profile v")) || !std::strstr(p + 15, "2.0"))
        return r.fail("wrong profile format (version line)");
    auto rest = [](const std::string& l, const char* pat, std::string& out) {
        const char* q = std::strstr(l.c_str(), pat);
        if (!q) return false;
        q += std::strlen(pat);
        while (*q == ' ' || *q == '\t') q++;
        out = q;
        while (!out.empty() && (out.back() == '\n' || out.back() == '\r')) out.pop_back();
        return true;
    };
    if (!r.next(line) || !rest(line, "DESC:", P.desc)) return r.fail("wrong profile format (DESC)");
    if (!r.next(line) || !rest(line, "FILE:", P.name)) return r.fail("wrong profile format (FILE)");
    if (!r.next(line)) return r.fail("wrong profile format (CMD)");
    int L = 0, scale = 0;
    if (!r.next(line) || !(p = std::strstr(line.c_str(), "LEN:"))) return r.fail("wrong profile format (LEN)");
    p += 4;
    if (!read_int(p, &L) || L < 1) return r.fail("invalid profile length");
    if (kMaxColumns < L) return r.fail("too large profile length");
    if (!r.next(line)) return r.fail("wrong profile format (NOS)");
    if (!r.next(line) || !(p = std::strstr(line.c_str(), "EFFNOS:"))) return r.fail("wrong profile format (EFFNOS)");
    p += 7;
    if (!read_flt(p, &P.eno) || !(P.eno > 0.0f)) return r.fail("invalid effective number of observations");
    if (!r.next(line) || !(p = std::strstr(line.c_str(), "SCALE:"))) return r.fail("wrong profile format (SCALE)");
    p += 6;
    if (!read_int(p, &scale) || scale < 1) return r.fail("invalid scale factor");
    if (!r.next(line)) return r.fail("wrong profile format (residues)");
    if (!r.next(line) || !(p = std::strstr(line.c_str(), "NULL"))) return r.fail("no background probabilities");
    p += 4;
    float consv = 0.0f;
    for (int a = 0; a < 20; a++) {
        int iv;
        if (at_end(p) || !read_int(p, &iv)) return r.fail("incomplete background probability data");
        if (iv < 0 || scale < iv) return r.fail("invalid background probability value");
        P.bkg[a] = (float)iv / (float)scale;
        consv += P.bkg[a];
    }
    if (consv < 1.0f - 1.0e-3f || consv > 1.0f + 1.0e-3f) return r.fail("invalid background probabilities");
    if (!r.next(line)) return r.fail("wrong profile format (POST)");
    P.L = L;
    P.trg.assign((size_t)L * 20, 0.0f);
    P.tr.assign((size_t)(L + 1) * 7, 0.0f);
    P.cv.assign((size_t)L * 19, kUninf);
    P.cvp.assign(L, 0.0f);
    P.cvn.assign(L, 0.0f);
    P.ss.assign((size_t)L * 3, 0.0f);
    P.aa.assign(L, 'X');
    P.sss.assign(L, ' ');
    /* data (TextReadProfileData, IOProfileModel.cu:287-1021): beginning
     * transitions, beginning MID observations, then the positions */
    if (!r.next(line)) return r.fail("wrong profile format (beginning transitions)");
    if (!read_trans(r, line, scale, &P.tr[0])) return false;
    if (!r.next(line)) return r.fail("wrong profile format (beginning observations)");
    bool lineread = false, ctset = false, cvset = false, ssset = false, ssp3 = true;
    int t0 = 0, tn = 2, vsize = 0;
    for (int m = 0; m < L; m++) {
        const std::string at = " at pos " + std::to_string(m);
        if (!lineread && !r.next(line)) return r.fail("wrong profile format" + at);
        lineread = false;
        p = line.c_str();
        int n;
        if (!read_int(p, &n)) return r.fail("wrong profile format" + at);
        if (n != m + 1) return r.fail("wrong numbering" + at);
        if (at_end(p)) return r.fail("no residue" + at);
        char res;
        if (!read_sym(p, &res)) return r.fail("invalid residue" + at);
        P.aa[m] = res;
        for (int a = 0; a < 20; a++) {
            int iv;
            if (at_end(p) || !read_int(p, &iv)) return r.fail("incomplete target probability data" + at);
            P.trg[(size_t)m * 20 + a] = (float)iv / (float)scale;
        }
        if (!r.next(line)) return r.fail("wrong profile format" + at);
        if (!read_trans(r, line, scale, &P.tr[(size_t)(m + 1) * 7])) return false;
        if (!r.next(line)) return r.fail("wrong profile format (MID observations)" + at);
        /* HDP1: background posterior and the number of posteriors; the
         * first posterior and cluster index lines when there are any */
        if (!r.next(line)) return r.fail("wrong profile format (HDP1)" + at);
        p = line.c_str();
        int bpp, nopp;
        if (!read_int(p, &bpp) || bpp < 0) return r.fail("invalid HDP1 background posterior" + at);
        if (at_end(p) || !read_int(p, &nopp) || nopp < 0) return r.fail("invalid number of HDP1 posteriors" + at);
        if (kMaxPosteriors < nopp) return r.fail("too large number of HDP1 posteriors" + at);
        if (nopp >= 1) {
            int iv;
            if (!r.next(line)) return r.fail("no HDP1 posteriors" + at);
            p = line.c_str();
            if (!read_int(p, &iv) || iv < 0) return r.fail("invalid HDP1 posterior" + at);
            if (!r.next(line)) return r.fail("no HDP1 cluster indices" + at);
            p = line.c_str();
            if (!read_int(p, &iv) || (iv != -1 && iv < 0)) return r.fail("invalid HDP1 cluster index" + at);
        }
        /* HDP context line (probed at the first position) */
        if (m < 1) ctset = true;
        if (ctset) {
            if (!r.next(line)) return r.fail("wrong profile format" + at);
            lineread = true;
            if (!std::strstr(line.c_str(), "CT:")) {
                if (m < 1) ctset = false;
                else return r.fail("no HDP ctx probabilities" + at);
            } else {
                lineread = false;
            }
        }
        /* context vector (probed at the first position) */
        if (m < 1) cvset = true;
        if (cvset) {
            if (!lineread && !r.next(line)) return r.fail("wrong profile format" + at);
            lineread = true;
            const char* q = std::strstr(line.c_str(), "CV:");
            if (!q) {
                if (m < 1) cvset = false;
                else return r.fail("no context vector data" + at);
            } else {
                lineread = false;
                p = q + 3;
                int pr, nm, vs;
                if (at_end(p) || !read_int(p, &pr)) return r.fail("no prior for context vector" + at);
                if (at_end(p) || !read_int(p, &nm)) return r.fail("no context vector norm" + at);
                if (nm < 0) return r.fail("negative context vector norm" + at);
                if (at_end(p) || !read_int(p, &vs)) return r.fail("no context vector size" + at);
                if (vs < 0) return r.fail("negative context vector size" + at);
                if (1000 < vs) return r.fail("too large context vector size" + at);
                if (!vsize && vs != 19) return r.fail("only context vectors of size 19 are supported" + at);
                if (vsize && vsize != vs && vsize != 19) return r.fail("inconsistent context vector size" + at);
                vsize = vs;
                P.cvp[m] = (float)pr / (float)scale;
                P.cvn[m] = (float)nm / (float)scale;
                for (int t = 0; t < vs; t++) {
                    int iv;
                    if (at_end(p) || !read_int(p, &iv)) return r.fail("incomplete context vector data" + at);
                    if (t < 19) P.cv[(size_t)m * 19 + t] = (float)iv / (float)scale;
                }
                if (vs < 19)
                    for (int t = vs; t < 19; t++) P.cv[(size_t)m * 19 + t] = 0.0f;
            }
        }
        /* SS state and probabilities (probed at the first position; one
         * probability instead of three is detected there) */
        if (m < 1) { ssset = true; ssp3 = true; t0 = 0; tn = 2; }
        if (ssset) {
            if (!lineread && !r.next(line)) return r.fail("wrong profile format" + at);
            lineread = true;
            const char* q = std::strstr(line.c_str(), "SS:");
            if (!q) {
                if (m < 1) ssset = false;
                else return r.fail("no SS state" + at);
            } else {
                lineread = false;
                p = q + 3;
                char sc;
                if (!read_sym(p, &sc)) return r.fail("invalid SS state" + at);
                const char up = (char)std::toupper((unsigned char)sc);
                const int code = up == 'C' ? 0 : up == 'E' ? 1 : up == 'H' ? 2 : -1;
                if (code < 0) return r.fail("unrecognized SS state" + at);
                P.sss[m] = " eh"[code];
                if (!ssp3) { t0 = code; tn = code; }
                int iv = 0;
                for (int t = t0; t <= tn; t++) {
                    if (at_end(p)) {
                        if (m < 1 && t == t0 + 1) {
                            ssp3 = false;
                            P.ss[(size_t)m * 3 + code] = (float)iv / (float)scale;
                            break;
                        }
                        return r.fail("no SS state probability" + at);
                    }
                    if (!read_int(p, &iv)) return r.fail("invalid SS state probability" + at);
                    while (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n') p++;
                    P.ss[(size_t)m * 3 + t] = (float)iv / (float)scale;
                }
            }
        }
    }
    P.has_cv = cvset;
    /* statistical parameters, then the footer */
    if (!r.next(line)) return r.fail("no statistical parameters");
    if (!std::strstr(line.c_str(), "Expected score per pos. is non-negative,")) {
        if (!r.next(line)) return r.fail("no statistical parameters");
        if (!r.next(line)) return r.fail("no reference parameters");
        if (!r.next(line)) return r.fail("no entropy");
    }
    if (!r.next(line) || line.compare(0, 1, "*") != 0) return r.fail("no ending");
    /* the SS probabilities index the 30-entry SS table */
    for (float x : P.ss)
        if (!(x >= 0.0f && x <= 1.0f)) return r.fail("SS probability outside [0, 1]");
    return true;
}

}  // namespace

/* the field-major arrays of the profiles read so far */
void cbp_proset::build()
{
    const size_t np = pro.size();
    size_t ns = 0;
    for (const One& P : pro) ns += (size_t)P.L;
    bool anycv = false;
    for (const One& P : pro) anycv |= P.has_cv;
    bkg.assign(20 * np, 0.0f); eno.assign(np, 0.0f); len.assign(np, 0); dist.assign(np, 0);
    trgfrq.assign(20 * ns, 0.0f); ssprb.assign(3 * ns, 0.0f); trans.assign(7 * (ns + np), 0.0f);
    aa.assign(ns, 0); sss.assign(ns, 0);
    cv.assign(anycv ? 19 * ns : 0, kUninf); cvprior.assign(anycv ? ns : 0, 0.0f); cvnorm2.assign(anycv ? ns : 0, 0.0f);
    desc.clear(); name.clear(); has_cv.clear();
    size_t pos = 0;
    for (size_t p = 0; p < np; p++) {
        const One& P = pro[p];
        const size_t L = (size_t)P.L;
        for (int a = 0; a < 20; a++) bkg[a * np + p] = P.bkg[a];
        eno[p] = P.eno; len[p] = P.L; dist[p] = (uint32_t)pos;
        for (size_t m = 0; m < L; m++) {
            for (int a = 0; a < 20; a++) trgfrq[a * ns + pos + m] = P.trg[m * 20 + a];
            for (int j = 0; j < 3; j++) ssprb[j * ns + pos + m] = P.ss[m * 3 + j];
            aa[pos + m] = P.aa[m]; sss[pos + m] = P.sss[m];
            if (anycv) {
                for (int i = 0; i < 19; i++) cv[i * ns + pos + m] = P.cv[m * 19 + i];
                cvprior[pos + m] = P.cvp[m]; cvnorm2[pos + m] = P.cvn[m];
            }
        }
        for (size_t m = 0; m <= L; m++)
            for (int k = 0; k < 7; k++) trans[k * (ns + np) + pos + p + m] = P.tr[m * 7 + k];
        desc.push_back(P.desc); name.push_back(P.name); has_cv.push_back(P.has_cv ? 1 : 0);
        pos += L;
    }
    v = cbp_profiles{};
    v.nprof = (int32_t)np; v.npos = (int64_t)ns;
    v.bkg = bkg.data(); v.eno = eno.data(); v.len = len.data(); v.dist = dist.data();
    v.trgfrq = trgfrq.data(); v.ssprb = ssprb.data(); v.aa = aa.data(); v.sss = sss.data();
    v.trans = trans.data();
    if (anycv) { v.cv = cv.data(); v.cvprior = cvprior.data(); v.cvnorm2 = cvnorm2.data(); }
}

extern "C" {

int cbp_proset_create(cbp_proset** out)
{
    if (!out) return CBP_ERR_ARG;
    try { *out = new cbp_proset(); } catch (const std::bad_alloc&) { return CBP_ERR_MEM; }
    return CBP_OK;
}

int cbp_proset_read_text(cbp_proset* s, const char* path)
{
    if (!s || !path) return CBP_ERR_ARG;
    if (s->v.nprof && s->pro.empty()) { s->err = "a set read from a binary database takes no more files"; return CBP_ERR_ARG; }
    FILE* fp = std::fopen(path, "r");
    if (!fp) { s->err = std::string("cannot open ") + path; return CBP_ERR_ARG; }
    int rc = CBP_OK;
    try {
        Reader r(fp, path);
        int nread = 0;
        for (;;) {
            cbp_proset::One P;
            bool eof = false;
            if (!read_profile(r, P, &eof)) {
                if (eof && nread > 0) break;
                s->err = eof ? std::string(path) + ": no profile" : r.err;
                rc = CBP_ERR_ARG;
                break;
            }
            s->pro.push_back(std::move(P));
            nread++;
        }
        s->build();
    } catch (const std::bad_alloc&) {
        s->err = "host memory allocation failed";
        rc = CBP_ERR_MEM;
    }
    std::fclose(fp);
    return rc;
}

/* a COMER binary profile database, format v2.2 (Db::MakeBin_v2_2,
 * libpro/srcpro/Db.cpp:333-383; CalculateAndWriteAddressTable_v2_2 and
 * BinaryWriteProfile_v2_2, libpro/srcpro/PMProfileModel.cpp:1468-1960; read
 * as TdDataReader::OpenAndReadPreamble does, cupro/TdDataReader.cpp:385-467):
 * "COMER profile bin database v2.2", the number of profiles and of
 * positions (8 bytes each), a table of 79 file addresses, then one section
 * per field over all profiles -- lengths, ENOs, 20 background
 * probabilities, the 7 transition logs (a beginning row and one row per
 * position per profile), 20 target probabilities, 19 context-vector
 * entries, their prior and squared norm, 3 SS probabilities, HDP
 * probability and index, residues, SS states, the end addresses of the
 * descriptions and the NUL-terminated descriptions.  The sections are the
 * field-major arrays of cbp_profiles, read whole into the set. */
int cbp_proset_read_bin(cbp_proset* s, const char* path)
{
    if (!s || !path) return CBP_ERR_ARG;
    if (!s->pro.empty() || s->v.nprof) { s->err = "a binary database is read into an empty set"; return CBP_ERR_ARG; }
    FILE* fp = std::fopen(path, "rb");
    if (!fp) { s->err = std::string("cannot open ") + path; return CBP_ERR_ARG; }
    int rc = CBP_OK;
    auto bad = [&](const std::string& m) { s->err = std::string(path) + ": " + m; rc = CBP_ERR_ARG; };
    try {
        do {
            if (fseeko(fp, 0, SEEK_END) != 0) { bad("cannot seek"); break; }
            const uint64_t fsize = (uint64_t)ftello(fp);
            static const char magic[] = "COMER profile bin database v2.2";
            const size_t lm = sizeof magic - 1;
            std::vector<char> head(lm + 16);
            std::vector<uint64_t> tab(79);
            if (fseeko(fp, 0, SEEK_SET) != 0 || std::fread(head.data(), 1, head.size(), fp) != head.size() ||
                std::fread(tab.data(), 8, 79, fp) != 79) { bad("truncated header"); break; }
            if (std::memcmp(head.data(), magic, lm) != 0) { bad("wrong database format or version"); break; }
            uint64_t np, ns;
            std::memcpy(&np, head.data() + lm, 8);
            std::memcpy(&ns, head.data() + lm + 8, 8);
            if (np < 1 || ns < 1 || np > 0x7fffffffULL || ns > (1ULL << 40)) { bad("invalid number of profiles or positions"); break; }
            /* the sections' sizes, in table order */
            std::vector<uint64_t> sz;
            for (int i = 0; i < 2 + 20; i++) sz.push_back(4 * np);
            for (int i = 0; i < 7; i++) sz.push_back(4 * (np + ns));
            for (int i = 0; i < 20 + 19 + 2 + 3 + 2; i++) sz.push_back(4 * ns);
            sz.push_back(ns); sz.push_back(ns); sz.push_back(8 * np);
            bool ok = true;
            for (int i = 0; i < 78 && ok; i++)
                if (tab[i] + sz[i] != tab[i + 1] || tab[i + 1] > fsize) ok = false;
            if (!ok || tab[0] < lm + 16 + 79 * 8) { bad("inconsistent address table"); break; }
            auto rd = [&](int sec, void* dst) {
                return fseeko(fp, (off_t)tab[sec], SEEK_SET) == 0 && std::fread(dst, 1, sz[sec], fp) == sz[sec];
            };
            s->len.resize(np); s->eno.resize(np); s->bkg.resize(20 * np);
            s->trans.resize(7 * (np + ns)); s->trgfrq.resize(20 * ns); s->cv.resize(19 * ns);
            s->cvprior.resize(ns); s->cvnorm2.resize(ns); s->ssprb.resize(3 * ns);
            s->aa.resize(ns); s->sss.resize(ns);
            std::vector<uint64_t> dend(np);
            ok = rd(0, s->len.data()) && rd(1, s->eno.data());
            for (int a = 0; a < 20 && ok; a++) ok = rd(2 + a, s->bkg.data() + a * np);
            for (int k = 0; k < 7 && ok; k++) ok = rd(22 + k, s->trans.data() + k * (np + ns));
            for (int a = 0; a < 20 && ok; a++) ok = rd(29 + a, s->trgfrq.data() + a * ns);
            for (int i = 0; i < 19 && ok; i++) ok = rd(49 + i, s->cv.data() + i * ns);
            ok = ok && rd(68, s->cvprior.data()) && rd(69, s->cvnorm2.data());
            for (int j = 0; j < 3 && ok; j++) ok = rd(70 + j, s->ssprb.data() + j * ns);
            ok = ok && rd(75, s->aa.data()) && rd(76, s->sss.data()) && rd(77, dend.data());
            if (!ok) { bad("truncated section"); break; }
            /* lengths and distances */
            s->dist.resize(np);
            uint64_t acc = 0;
            for (uint64_t p = 0; p < np && ok; p++) {
                if (s->len[p] < 1 || s->len[p] > 65535) ok = false;
                s->dist[p] = (uint32_t)acc;
                acc += (uint64_t)s->len[p];
            }
            if (!ok || acc != ns || ns > 0xffffffffULL) { bad("profile lengths disagree with the database size"); break; }
            for (float x : s->ssprb)
                if (!(x >= 0.0f && x <= 1.0f)) { ok = false; break; }
            if (!ok) { bad("SS probability outside [0, 1]"); break; }
            /* descriptions: [end of the previous, end) each, NUL included */
            if (dend[np - 1] > fsize || dend[np - 1] < tab[78]) { bad("invalid description addresses"); break; }
            std::vector<char> dtext(dend[np - 1] - tab[78]);
            if (fseeko(fp, (off_t)tab[78], SEEK_SET) != 0 ||
                std::fread(dtext.data(), 1, dtext.size(), fp) != dtext.size()) { bad("truncated descriptions"); break; }
            s->desc.clear(); s->name.clear(); s->has_cv.assign(np, 1);
            uint64_t b = tab[78];
            for (uint64_t p = 0; p < np && ok; p++) {
                if (dend[p] <= b || dend[p] > dend[np - 1]) { ok = false; break; }
                s->desc.emplace_back(dtext.data() + (b - tab[78]), strnlen(dtext.data() + (b - tab[78]), dend[p] - b));
                s->name.emplace_back();
                b = dend[p];
            }
            if (!ok) { bad("invalid description addresses"); break; }
            cbp_profiles& v = s->v;
            v = cbp_profiles{};
            v.nprof = (int32_t)np; v.npos = (int64_t)ns;
            v.bkg = s->bkg.data(); v.eno = s->eno.data(); v.len = s->len.data(); v.dist = s->dist.data();
            v.trgfrq = s->trgfrq.data(); v.ssprb = s->ssprb.data(); v.aa = s->aa.data(); v.sss = s->sss.data();
            v.trans = s->trans.data();
            v.cv = s->cv.data(); v.cvprior = s->cvprior.data(); v.cvnorm2 = s->cvnorm2.data();
        } while (0);
    } catch (const std::bad_alloc&) {
        s->err = "host memory allocation failed";
        rc = CBP_ERR_MEM;
    }
    std::fclose(fp);
    if (rc != CBP_OK) {
        const std::string e = s->err.empty() ? std::string(path) + ": read failed" : s->err;
        *s = cbp_proset();
        s->err = e;
    }
    return rc;
}

const cbp_profiles* cbp_proset_profiles(const cbp_proset* s) { return s ? &s->v : nullptr; }

const char* cbp_proset_desc(const cbp_proset* s, int32_t i)
{
    return (s && i >= 0 && (size_t)i < s->desc.size()) ? s->desc[i].c_str() : nullptr;
}

const char* cbp_proset_name(const cbp_proset* s, int32_t i)
{
    return (s && i >= 0 && (size_t)i < s->name.size()) ? s->name[i].c_str() : nullptr;
}

const char* cbp_proset_error(const cbp_proset* s) { return s ? s->err.c_str() : "null profile set"; }

void cbp_proset_free(cbp_proset* s) { delete s; }

}  /* extern "C" */
