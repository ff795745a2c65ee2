/*
 * c2_synth.cpp -- seeded synthetic profile databases in the reference's SoA
 * layout (cupro/PM2DVectorFields.h:48-81; transitions carry len+1 rows per
 * profile, cupro/PMBatchProData.h:412).
 *
 * Per position: 20 target probabilities ~ Dirichlet(conc * bg) with bg the
 * Robinson background (libpro/srcpro/SUBSTABLE.cpp:404-406), residue =
 * argmax, 3 SS probabilities ~ Dirichlet(1,1,1), SS state = argmax (stored as
 * the reader stores it: ' ', 'e', 'h' for C, E, H), and the
 * 7 used transitions (MM, MI, MD, IM, II, DM, DD) drawn row-wise: M-row ~
 * Dirichlet(18,1,1) (E[MM] = 0.9), I-row (IM,II) and D-row (DM,DD) ~
 * Dirichlet(3,2).  Transitions are stored as logs, with IM, II, DM, DD scaled
 * by 0.4 and zero stored as -32768 (cupro/IOProfileModel.cu:363).
 * Per profile: ENO ~ U(1,14), background = bg, and one extra "beginning"
 * transition row.  Lengths ~ round(LogNormal(log(median), sigma)) clipped to
 * [minlen, maxlen] (or uniform on it); the database is sorted by length, longest first, as
 * makedb stores it (libpro/srcpro/Db.h:268-275).
 *
 * Context vectors (optional, c2s_add_cv): 19 entries ~ N(0, 0.5^2) per
 * position, squared norm = their sum of squares, log prior ~ -U(12, 16);
 * c2s_plant_cv gives the planted windows the source positions' vectors plus
 * N(0, 0.1^2) noise, so homologous windows score high in the CV term.
 *
 * Randomness: Philox4x32-10 keyed by (seed, profile serial), so every profile
 * is generated independently and the result does not depend on the number of
 * OpenMP threads.
 */
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <cmath>
#include <vector>
#include <algorithm>
#include <numeric>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../../include/comer2_hip.h"

namespace {

const float kRobinson[20] = {
    0.078050f, 0.051290f, 0.044870f, 0.053640f, 0.019250f, 0.042640f, 0.062950f,
    0.073770f, 0.021990f, 0.051420f, 0.090190f, 0.057440f, 0.022430f, 0.038560f,
    0.052030f, 0.071200f, 0.058410f, 0.013300f, 0.032160f, 0.064410f};
const char kAlpha[21] = "ARNDCQEGHILKMFPSTWYV";

struct Philox {
    uint32_t key0, key1, ctr[4];
    uint32_t out[4];
    int idx;
    Philox(uint64_t seed, uint64_t stream) : idx(4) {
        key0 = (uint32_t)seed; key1 = (uint32_t)(seed >> 32);
        ctr[0] = 0; ctr[1] = 0; ctr[2] = (uint32_t)stream; ctr[3] = (uint32_t)(stream >> 32);
    }
    void round_(uint32_t c[4], uint32_t k0, uint32_t k1) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        c[0] = n0; c[1] = (uint32_t)p1; c[2] = n2; c[3] = (uint32_t)p0;
    }
    void block() {
        uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
        uint32_t k0 = key0, k1 = key1;
        for (int r = 0; r < 10; r++) {
            round_(c, k0, k1);
            k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
        }
        memcpy(out, c, sizeof out);
        if (++ctr[0] == 0) ++ctr[1];
        idx = 0;
    }
    uint32_t next() { if (idx >= 4) block(); return out[idx++]; }
    /* uniform in (0,1) with 53 random bits */
    double uniform() {
        uint64_t a = next() >> 5, b = next() >> 6;
        return ((double)(a * 67108864ull + b) + 0.5) / 9007199254740992.0;
    }
    double normal() {  /* Box-Muller, one value per call */
        double u1 = uniform(), u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
    /* Gamma(a,1): Marsaglia & Tsang (2000), with the a<1 boost U^(1/a) */
    double gamma(double a) {
        double boost = 1.0;
        if (a < 1.0) { boost = std::pow(uniform(), 1.0 / a); a += 1.0; }
        double d = a - 1.0 / 3.0, c = 1.0 / std::sqrt(9.0 * d);
        for (;;) {
            double x, v;
            do { x = normal(); v = 1.0 + c * x; } while (v <= 0.0);
            v = v * v * v;
            double u = uniform();
            if (u < 1.0 - 0.0331 * x * x * x * x) return d * v * boost;
            if (std::log(u) < 0.5 * x * x + d * (1.0 - v + std::log(v))) return d * v * boost;
        }
    }
};

void dirichlet(Philox& g, const double* alpha, int n, double* out) {
    double s = 0.0;
    for (int i = 0; i < n; i++) { out[i] = g.gamma(alpha[i]); s += out[i]; }
    for (int i = 0; i < n; i++) out[i] /= s;
}

float logtr(double p, bool scaled) {
    if (!(p > 0.0)) return -32768.0f;
    float v = logf((float)p);
    return scaled ? v * 0.4f : v;
}

/* one transition row: MM MI MD IM II DM DD */
void gen_trans(Philox& g, float t[7]) {
    const double am[3] = {18.0, 1.0, 1.0}, ai[2] = {3.0, 2.0}, ad[2] = {3.0, 2.0};
    double m[3], i2[2], d2[2];
    dirichlet(g, am, 3, m); dirichlet(g, ai, 2, i2); dirichlet(g, ad, 2, d2);
    t[0] = logtr(m[0], false); t[1] = logtr(m[1], false); t[2] = logtr(m[2], false);
    t[3] = logtr(i2[0], true); t[4] = logtr(i2[1], true);
    t[5] = logtr(d2[0], true); t[6] = logtr(d2[1], true);
}

struct SynthDB {
    cbp_profiles view;
    std::vector<float> bkg, eno, trgfrq, ssprb, trans;
    std::vector<float> cv, cvprior, cvnorm2;
    std::vector<int32_t> len;
    std::vector<uint32_t> dist;
    std::vector<char> aa, sss;
    std::vector<int32_t> orgidx;  /* generation serial of each (sorted) profile */
};

}  // namespace

extern "C" {

/* lengths: LogNormal(log(median), sigma) rounded, clipped, or uniform on
 * [minlen, maxlen] when median <= 0; sorted descending when sort_desc != 0.
 * Returns an owned handle; free with c2s_free. */
static void* generate_from(uint64_t seed, int32_t nprof, const std::vector<int32_t>& lens,
                           int32_t sort_desc, double conc);

void* c2s_generate(uint64_t seed, int32_t nprof, double median, double sigma,
                   int32_t minlen, int32_t maxlen, int32_t sort_desc, double conc)
{
    std::vector<int32_t> lens(nprof);
    {
        Philox g(seed, 0xFFFFFFFFull);
        for (int32_t p = 0; p < nprof; p++) {
            long v;
            if (median > 0.0) v = std::lround(std::exp(std::log(median) + sigma * g.normal()));
            else v = minlen + (long)(g.uniform() * (double)(maxlen - minlen + 1));
            if (v < minlen) v = minlen;
            if (v > maxlen) v = maxlen;
            lens[p] = (int32_t)v;
        }
    }
    return generate_from(seed, nprof, lens, sort_desc, conc);
}

/* profiles of the given lengths (in the given order), the same per-position
 * distributions as c2s_generate */
void* c2s_generate_lengths(uint64_t seed, int32_t nprof, const int32_t* lengths, double conc)
{
    std::vector<int32_t> lens(lengths, lengths + nprof);
    return generate_from(seed, nprof, lens, 0, conc);
}

}  /* extern "C" */

static void* generate_from(uint64_t seed, int32_t nprof, const std::vector<int32_t>& lens,
                           int32_t sort_desc, double conc)
{
    SynthDB* db = new SynthDB();
    std::vector<int32_t> order(nprof);
    std::iota(order.begin(), order.end(), 0);
    if (sort_desc)
        std::stable_sort(order.begin(), order.end(),
                         [&](int32_t a, int32_t b) { return lens[a] > lens[b]; });
    db->orgidx = order;
    db->len.resize(nprof);
    db->dist.resize(nprof);
    int64_t npos = 0;
    for (int32_t p = 0; p < nprof; p++) {
        db->len[p] = lens[order[p]];
        db->dist[p] = (uint32_t)npos;
        npos += db->len[p];
    }
    db->bkg.resize((size_t)20 * nprof);
    db->eno.resize(nprof);
    db->trgfrq.resize((size_t)20 * npos);
    db->ssprb.resize((size_t)3 * npos);
    db->trans.resize((size_t)7 * (npos + nprof));
    db->aa.resize(npos);
    db->sss.resize(npos);
    const int64_t ntr = npos + nprof;
#pragma omp parallel for schedule(dynamic, 16)
    for (int32_t p = 0; p < nprof; p++) {
        Philox g(seed, (uint64_t)order[p] + 1);
        int32_t L = db->len[p];
        int64_t d0 = db->dist[p];
        for (int a = 0; a < 20; a++) db->bkg[(size_t)a * nprof + p] = kRobinson[a];
        db->eno[p] = (float)(1.0 + 13.0 * g.uniform());
        double alpha[20], q[20], ss[3];
        for (int a = 0; a < 20; a++) alpha[a] = conc * (double)kRobinson[a];
        const double a3[3] = {1.0, 1.0, 1.0};
        float t[7];
        gen_trans(g, t);   /* beginning row */
        for (int k = 0; k < 7; k++) db->trans[(size_t)k * ntr + d0 + p] = t[k];
        for (int32_t i = 0; i < L; i++) {
            int64_t pos = d0 + i;
            dirichlet(g, alpha, 20, q);
            int am = 0;
            for (int a = 0; a < 20; a++) {
                db->trgfrq[(size_t)a * npos + pos] = (float)q[a];
                if (q[a] > q[am]) am = a;
            }
            db->aa[pos] = kAlpha[am];
            dirichlet(g, a3, 3, ss);
            int sm = 0;
            for (int s = 0; s < 3; s++) {
                db->ssprb[(size_t)s * npos + pos] = (float)ss[s];
                if (ss[s] > ss[sm]) sm = s;
            }
            db->sss[pos] = " eh"[sm];
            gen_trans(g, t);
            for (int k = 0; k < 7; k++) db->trans[(size_t)k * ntr + pos + p + 1] = t[k];
        }
    }
    cbp_profiles& v = db->view;
    v.nprof = nprof; v.npos = npos;
    v.bkg = db->bkg.data(); v.eno = db->eno.data(); v.len = db->len.data();
    v.dist = db->dist.data(); v.trgfrq = db->trgfrq.data(); v.ssprb = db->ssprb.data();
    v.aa = db->aa.data(); v.sss = db->sss.data(); v.trans = db->trans.data();
    return db;
}

extern "C" {

const cbp_profiles* c2s_view(void* h) { return &((SynthDB*)h)->view; }
const int32_t* c2s_orgidx(void* h) { return ((SynthDB*)h)->orgidx.data(); }
void c2s_free(void* h) { delete (SynthDB*)h; }

/* Plants homologous segments: nplant DB profiles (chosen by the seed) receive,
 * over a window of w = U(0.5,1) * min(Ldb, Ls) positions, target
 * probabilities ~ Dirichlet(k * t_src + 0.5 bg) and SS probabilities
 * ~ Dirichlet(k * p_src + 0.2) copied from a window of a source profile
 * (source profile chosen round-robin), with k ~ U(kmin, kmax) per planted
 * profile; residues and SS states are re-derived (argmax).  Returns the
 * number planted; plant_src[i] / plant_pro[i] record (source, DB profile). */
int32_t c2s_plant(void* h, const cbp_profiles* src, uint64_t seed, int32_t nplant,
                  double kmin, double kmax, int32_t* plant_src, int32_t* plant_pro)
{
    SynthDB* db = (SynthDB*)h;
    const int32_t np = db->view.nprof;
    if (nplant > np) nplant = np;
    Philox g(seed, 0xABCDEF0123ull);
    std::vector<int32_t> idx(np);
    std::iota(idx.begin(), idx.end(), 0);
    for (int32_t i = 0; i < nplant; i++) {       /* partial Fisher-Yates */
        int32_t j = i + (int32_t)(g.uniform() * (np - i));
        if (j >= np) j = np - 1;
        std::swap(idx[i], idx[j]);
    }
    std::sort(idx.begin(), idx.begin() + nplant);
    const int64_t npos = db->view.npos;
#pragma omp parallel for schedule(dynamic, 8)
    for (int32_t i = 0; i < nplant; i++) {
        const int32_t p = idx[i];
        const int32_t sp = i % src->nprof;
        Philox r(seed, 0x100000000ull + (uint64_t)p);
        const int32_t Ld = db->len[p], Ls = src->len[sp];
        const int32_t m = Ld < Ls ? Ld : Ls;
        int32_t w = (int32_t)((0.5 + 0.5 * r.uniform()) * m);
        if (w < 1) w = 1;
        const int32_t so = (int32_t)(r.uniform() * (Ls - w + 1));
        const int32_t dof = (int32_t)(r.uniform() * (Ld - w + 1));
        const double k = kmin + (kmax - kmin) * r.uniform();
        double alpha[20], tq[20], a3[3], s3[3];
        for (int32_t j = 0; j < w; j++) {
            const int64_t spos = (int64_t)src->dist[sp] + so + j;
            const int64_t pos = (int64_t)db->dist[p] + dof + j;
            for (int a = 0; a < 20; a++)
                alpha[a] = k * (double)src->trgfrq[(size_t)a * src->npos + spos] + 0.5 * (double)kRobinson[a];
            dirichlet(r, alpha, 20, tq);
            int am = 0;
            for (int a = 0; a < 20; a++) {
                db->trgfrq[(size_t)a * npos + pos] = (float)tq[a];
                if (tq[a] > tq[am]) am = a;
            }
            db->aa[pos] = kAlpha[am];
            for (int s = 0; s < 3; s++) a3[s] = k * (double)src->ssprb[(size_t)s * src->npos + spos] + 0.2;
            dirichlet(r, a3, 3, s3);
            int sm = 0;
            for (int s = 0; s < 3; s++) {
                db->ssprb[(size_t)s * npos + pos] = (float)s3[s];
                if (s3[s] > s3[sm]) sm = s;
            }
            db->sss[pos] = " eh"[sm];
        }
        if (plant_src) plant_src[i] = sp;
        if (plant_pro) plant_pro[i] = p;
    }
    return nplant;
}

/* Adds context vectors to every position (see the header). */
void c2s_add_cv(void* h, uint64_t seed)
{
    SynthDB* db = (SynthDB*)h;
    const int32_t np = db->view.nprof;
    const int64_t npos = db->view.npos;
    db->cv.assign((size_t)19 * npos, 0.0f);
    db->cvprior.assign(npos, 0.0f);
    db->cvnorm2.assign(npos, 0.0f);
#pragma omp parallel for schedule(dynamic, 64)
    for (int32_t p = 0; p < np; p++) {
        Philox r(seed, 0x200000000ull + (uint64_t)p);
        for (int32_t j = 0; j < db->len[p]; j++) {
            const int64_t pos = (int64_t)db->dist[p] + j;
            float n2 = 0.0f;
            for (int i = 0; i < 19; i++) {
                const float e = (float)(0.5 * r.normal());
                db->cv[(size_t)i * npos + pos] = e;
                n2 += e * e;
            }
            db->cvnorm2[pos] = n2;
            db->cvprior[pos] = (float)(-12.0 - 4.0 * r.uniform());
        }
    }
    cbp_profiles& v = db->view;
    v.cv = db->cv.data(); v.cvprior = db->cvprior.data(); v.cvnorm2 = db->cvnorm2.data();
}

/* After c2s_plant with the same seed and source: the planted windows take
 * the source positions' context vectors plus N(0, 0.1^2) noise (norms
 * recomputed, priors copied).  Both sets need context vectors. */
int32_t c2s_plant_cv(void* h, const cbp_profiles* src, uint64_t seed, int32_t nplant)
{
    SynthDB* db = (SynthDB*)h;
    if (!db->view.cv || !src->cv) return -1;
    const int32_t np = db->view.nprof;
    if (nplant > np) nplant = np;
    /* the same choice of profiles and windows as c2s_plant */
    Philox g(seed, 0xABCDEF0123ull);
    std::vector<int32_t> idx(np);
    std::iota(idx.begin(), idx.end(), 0);
    for (int32_t i = 0; i < nplant; i++) {
        int32_t j = i + (int32_t)(g.uniform() * (np - i));
        if (j >= np) j = np - 1;
        std::swap(idx[i], idx[j]);
    }
    std::sort(idx.begin(), idx.begin() + nplant);
    const int64_t npos = db->view.npos;
#pragma omp parallel for schedule(dynamic, 8)
    for (int32_t i = 0; i < nplant; i++) {
        const int32_t p = idx[i];
        const int32_t sp = i % src->nprof;
        Philox r(seed, 0x100000000ull + (uint64_t)p);
        const int32_t Ld = db->len[p], Ls = src->len[sp];
        const int32_t m = Ld < Ls ? Ld : Ls;
        int32_t w = (int32_t)((0.5 + 0.5 * r.uniform()) * m);
        if (w < 1) w = 1;
        const int32_t so = (int32_t)(r.uniform() * (Ls - w + 1));
        const int32_t dof = (int32_t)(r.uniform() * (Ld - w + 1));
        Philox q(seed, 0x300000000ull + (uint64_t)p);
        for (int32_t j = 0; j < w; j++) {
            const int64_t spos = (int64_t)src->dist[sp] + so + j;
            const int64_t pos = (int64_t)db->dist[p] + dof + j;
            float n2 = 0.0f;
            for (int k = 0; k < 19; k++) {
                const float e = src->cv[(size_t)k * src->npos + spos] + (float)(0.1 * q.normal());
                db->cv[(size_t)k * npos + pos] = e;
                n2 += e * e;
            }
            db->cvnorm2[pos] = n2;
            db->cvprior[pos] = src->cvprior[spos];
        }
    }
    return nplant;
}

/* Copies profiles [p0, p0+n) of a set into a new owned set (dist rebased). */
void* c2s_subset(void* h, int32_t p0, int32_t n)
{
    const SynthDB* s = (const SynthDB*)h;
    const cbp_profiles& a = s->view;
    SynthDB* db = new SynthDB();
    int64_t b0 = a.dist[p0];
    int64_t npos = 0;
    db->len.resize(n); db->dist.resize(n); db->eno.resize(n);
    db->bkg.resize((size_t)20 * n); db->orgidx.resize(n);
    for (int32_t p = 0; p < n; p++) {
        db->len[p] = a.len[p0 + p]; db->dist[p] = (uint32_t)npos; npos += db->len[p];
        db->eno[p] = a.eno[p0 + p]; db->orgidx[p] = s->orgidx[p0 + p];
        for (int k = 0; k < 20; k++) db->bkg[(size_t)k * n + p] = a.bkg[(size_t)k * a.nprof + p0 + p];
    }
    db->trgfrq.resize((size_t)20 * npos); db->ssprb.resize((size_t)3 * npos);
    db->aa.resize(npos); db->sss.resize(npos);
    db->trans.resize((size_t)7 * (npos + n));
    for (int k = 0; k < 20; k++)
        memcpy(&db->trgfrq[(size_t)k * npos], &a.trgfrq[(size_t)k * a.npos + b0], npos * sizeof(float));
    for (int k = 0; k < 3; k++)
        memcpy(&db->ssprb[(size_t)k * npos], &a.ssprb[(size_t)k * a.npos + b0], npos * sizeof(float));
    memcpy(db->aa.data(), a.aa + b0, npos); memcpy(db->sss.data(), a.sss + b0, npos);
    for (int k = 0; k < 7; k++)
        memcpy(&db->trans[(size_t)k * (npos + n)], &a.trans[(size_t)k * (a.npos + a.nprof) + b0 + p0],
               (npos + n) * sizeof(float));
    cbp_profiles& v = db->view;
    v.nprof = n; v.npos = npos;
    v.bkg = db->bkg.data(); v.eno = db->eno.data(); v.len = db->len.data();
    v.dist = db->dist.data(); v.trgfrq = db->trgfrq.data(); v.ssprb = db->ssprb.data();
    v.aa = db->aa.data(); v.sss = db->sss.data(); v.trans = db->trans.data();
    if (a.cv) {
        db->cv.resize((size_t)19 * npos); db->cvprior.resize(npos); db->cvnorm2.resize(npos);
        for (int k = 0; k < 19; k++)
            memcpy(&db->cv[(size_t)k * npos], &a.cv[(size_t)k * a.npos + b0], npos * sizeof(float));
        memcpy(db->cvprior.data(), a.cvprior + b0, npos * sizeof(float));
        memcpy(db->cvnorm2.data(), a.cvnorm2 + b0, npos * sizeof(float));
        v.cv = db->cv.data(); v.cvprior = db->cvprior.data(); v.cvnorm2 = db->cvnorm2.data();
    }
    return db;
}

}  /* extern "C" */
