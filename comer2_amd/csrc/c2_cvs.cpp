/*
 * c2_cvs.cpp -- the context-vector score term's host side: the cvs2s map
 * file (CVS2Scores::ReadScores / ReadScoresHelper / SmoothScores,
 * libpro/srcpro/CVS2Scores.cpp:54-466), its upload
 * (CuDeviceMemory::CacheCVS2ScoresDevice, cupro/CuDeviceMemory.cu:671-723)
 * and the constants of the multivariate t model of context vectors
 * (_TCTXVECT, libpro/srcpro/TCTXVECT.cpp:20-54).
 *
 * Only the first table of the map is read: the score kernels look up the
 * first table alone (GetScoreE1S1Step2Boundary,
 * cupro/SerializedCVS2Scores.cuh:377-410), whatever the number of levels of
 * effective observations the file declares.
 */
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include "c2_lib.h"

namespace {

constexpr float kPlus = 9999.0f;    /* a '+' entry (CVS2Scores plusval_) */
constexpr float kNA = -9999.0f;     /* an NA entry (naval_) */

/* next line that is neither blank nor a '#' comment (skip_comments,
 * liblib/myfiler.cpp:131-185); false at the end of the file */
bool next_line(FILE* fp, std::string& line)
{
    char buf[4096];
    while (std::fgets(buf, sizeof buf, fp)) {
        std::string l(buf);
        while (!l.empty() && l.back() != '\n' && std::fgets(buf, sizeof buf, fp)) l += buf;
        size_t i = 0;
        while (i < l.size() && (l[i] == ' ' || l[i] == '\t')) i++;
        if (i < l.size() && l[i] != '#' && l[i] != '\n' && l[i] != '\r') { line = l; return true; }
    }
    return false;
}

/* the next token as read_float reads it (strtof of the characters up to a
 * blank, ';', ',', a parenthesis or the line end); false when there is none
 * or it is not a number */
bool next_float(const char*& p, float* v)
{
    while (*p == ' ' || *p == '\t') p++;
    const char* b = p;
    while (*p && *p != ' ' && *p != '\t' && *p != ';' && *p != ',' && *p != '(' && *p != ')' &&
           *p != '\n' && *p != '\r') p++;
    if (p == b) return false;
    std::string tok(b, p);
    char* end = nullptr;
    errno = 0;
    *v = std::strtof(tok.c_str(), &end);
    return errno == 0 && *end == 0;
}

bool after(const std::string& line, const char* pat, const char*& p)
{
    const char* f = std::strstr(line.c_str(), pat);
    if (!f) return false;
    p = f + std::strlen(pat);
    return true;
}

/* SmoothScores (CVS2Scores.cpp:366-466) for step 2: a centred moving
 * average over five scores ('+' entries are skipped and take the average
 * of their neighbourhood), with the running sum kept as the reference keeps
 * it (subtract the value leaving the window, add the new one, divide by 5);
 * every score is replaced by its smoothed value; missing keys (gaps wider
 * than 1/2) are inserted with the last smoothed score; shift = number of
 * negative keys */
int smooth(std::vector<float>& keys, std::vector<float>& sco, int* shift, std::string& err)
{
    const float incr = 1.0f / 2.0f;
    const int nSWIN = 5, nSWINh = nSWIN >> 1;
    const int n_all = (int)sco.size();
    if (n_all < 1) return 0;
    std::vector<float> smth(n_all, 0.0f);
    float rwin[nSWIN] = {0, 0, 0, 0, 0};
    float rsum = 0.0f, avg = 0.0f;
    int nv = 0, rn = 0, nn = 0;
    for (int n = 0; n < n_all; n++) {
        const float sc = sco[n];
        if (sc <= kNA) { err = "NA score in the map"; return -1; }
        if (nv < nSWINh) nn++;
        const bool plus = kPlus <= sc;
        if (!plus) {
            if (nSWIN <= rn) rn = 0;
            rsum -= rwin[rn];
            rwin[rn++] = sc;
            rsum += sc;
            avg = rsum / (float)nSWIN;
            nv++;
        }
        if (nSWIN <= nv) {
            if (!plus) {
                if (nSWIN == nv)
                    for (int p = 0; p < nn; p++) smth[p] = avg;
                smth[nn] = avg;
                for (++nn; nn < n && kPlus <= sco[nn]; nn++) smth[nn] = avg;
            }
            if (n_all == n + 1)
                for (int p = nn; p < n_all; p++) smth[p] = avg;
        }
    }
    float sm = 0.0f;
    for (int n = n_all - 1; n >= 0; n--) { sm = smth[n]; sco[n] = sm; }
    float pkk = 0.0f;
    for (size_t n = 0; n < keys.size(); n++) {
        const float kk = keys[n];
        if (n && kk <= pkk) { err = "map keys are not sorted"; return -1; }
        if (n && incr < kk - pkk) {
            pkk += incr;
            keys.insert(keys.begin() + n, pkk);
            sco.insert(sco.begin() + n, sm);
            smth.insert(smth.begin() + n, sm);
            continue;
        }
        pkk = kk;
        sm = smth[n];
        if (kPlus <= sco[n]) sco[n] = sm;
    }
    *shift = 0;
    for (size_t n = 0; n < keys.size(); n++) {
        if (0.0f <= keys[n]) break;
        (*shift)++;
    }
    return 0;
}

int read_map(const char* path, std::vector<float>& keys, std::vector<float>& sco, int* shift,
             std::string& err)
{
    FILE* fp = std::fopen(path, "r");
    if (!fp) { err = std::string("cannot open ") + path; return -1; }
    std::string line;
    const char* p = nullptr;
    int rc = -1;
    float v;
    do {
        /* levels of effective observations (ReadLevels, :85-140) */
        if (!next_line(fp, line) || !after(line, "cvs2s=", p)) { err = "no 'cvs2s=' line"; break; }
        std::vector<float> lev;
        while (next_float(p, &v)) lev.push_back(v);
        if (lev.empty() || lev.size() > 10) { err = "invalid number of levels"; break; }
        /* the first table's header: its pair of levels (ReadScoreTables, :142-205) */
        char levcmb[64];
        std::snprintf(levcmb, sizeof levcmb, "%d%d", (int)rintf(lev[0]), (int)rintf(lev[0]));
        if (!next_line(fp, line) || std::strncmp(line.c_str(), levcmb, std::strlen(levcmb)) != 0) {
            err = "no table header for the first pair of levels"; break;
        }
        /* scale (optional), dimensions, cardinality (ReadScoresHelper, :207-364) */
        if (!next_line(fp, line)) { err = "truncated map"; break; }
        if (after(line, "Scale =", p)) {
            if (!next_float(p, &v) || v <= 0.0f || v > 1000.0f) { err = "invalid scale"; break; }
            if (!next_line(fp, line)) { err = "truncated map"; break; }
        }
        if (!after(line, "Dimensions =", p)) { err = "no dimensions"; break; }
        long dim = std::strtol(p, nullptr, 10);
        if (dim < 19 || dim > 1900) { err = "invalid dimensions"; break; }
        if (!next_line(fp, line) || !after(line, "Cardinality =", p)) { err = "no cardinality"; break; }
        long card = std::strtol(p, nullptr, 10);
        if (card <= 0 || card > 4000) { err = "invalid cardinality"; break; }
        std::string lk, ls;
        if (!next_line(fp, lk) || !next_line(fp, ls)) { err = "no keys or scores"; break; }
        const char* pk = lk.c_str();
        const char* ps = ls.c_str();
        keys.clear(); sco.clear();
        bool ok = true;
        for (long c = 0; c < card && ok; c++) {
            float k, s;
            if (!next_float(pk, &k)) { err = "short of keys"; ok = false; break; }
            while (*ps == ' ' || *ps == '\t') ps++;
            if (*ps == '+') { ps++; s = kPlus; }
            else if (!next_float(ps, &s)) { err = "short of scores"; ok = false; break; }
            keys.push_back(k); sco.push_back(s);
        }
        if (!ok) break;
        rc = smooth(keys, sco, shift, err);
    } while (0);
    std::fclose(fp);
    return rc;
}

}  // namespace

void c2_cv_constants(float* kappa_inv, float* nu_pow, float* cterm)
{
    /* _TCTXVECT (TCTXVECT.cpp:20-54): DIM = 19, KAPPA0 = 0.1, NU0 = DIM,
     * loKAPPA0 = 1/KAPPA0, PowerNU0 = -(NU0+2)/2,
     * CTERM = (DIM/2 + PowerNU0)(log KAPPA0 - log(KAPPA0+2))
     *         + sum_{i=1..DIM} (log(NU0+1-i) - log 2), all in float */
    const int dim = 19;
    const float kappa0 = 0.1f, nu0 = (float)dim;
    const float lok = 1.f / kappa0;
    const float pnu = -0.5f * (nu0 + 2.f);
    float ct = (0.5f * dim + pnu) * (logf(kappa0) - logf(kappa0 + 2.f));
    float sum = 0.0f;
    for (int i = 1; i <= dim; i++) sum += logf(nu0 + 1.0f - i) - 0.69314718055994530941723212146f;
    ct += sum;
    *kappa_inv = lok; *nu_pow = pnu; *cterm = ct;
}

extern "C" int cbp_cvs2s_read(const char* path, float* keys, float* values, int32_t cap,
                              int32_t* card, int32_t* shift)
{
    C2ApiScope c2_scope_;
    if (!path || !card || !shift) return CBP_ERR_ARG;
    std::vector<float> k, s;
    std::string err;
    int sh = 0;
    if (read_map(path, k, s, &sh, err)) {
        std::fprintf(stderr, "cbp_cvs2s_read: %s: %s\n", path, err.c_str());
        return CBP_ERR_ARG;
    }
    *card = (int32_t)k.size();
    *shift = sh;
    if (keys || values) {
        if (cap < (int32_t)k.size()) return CBP_ERR_ARG;
        if (keys) std::memcpy(keys, k.data(), sizeof(float) * k.size());
        if (values) std::memcpy(values, s.data(), sizeof(float) * s.size());
    }
    return CBP_OK;
}
