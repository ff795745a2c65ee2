/*
 * c2_format.cpp -- a query's search results as the reference's plain-text
 * output file (host only): the programme header and query description
 * (AlnWriter::WritePrognamePlain / WriteQueryDescriptionPlain,
 * cualn/AlnWriterPlain.cpp:123-226), the database lines and hit-list
 * heading (WriteSearchInformationPlain, :228-320), one annotation line per
 * hit (MakeAnnotationPlain, cupro/CuBatchProcessingFinalizerPlain.cpp:144-178),
 * the alignment records (description, scores, alignment in fragments of
 * ALNWIDTH columns, statistics footer; CompressResultsPlain,
 * FormatScoresPlain, FormatAlignmentPlain, FormatFooterPlain, :31-401) and
 * the closing summary (WriteSummaryPlain, cualn/AlnWriterPlain.cpp:322-398).
 * Descriptions wrap and truncate as FormatDescription does
 * (liblib/fmtdescription.h:34-90).
 *
 * The hits are the query's pairs with log E <= the log E threshold and an
 * alignment drawn (its length is the EPA field), in the results' (log E,
 * profile) order; NOHITS bounds the annotations and NOALNS the records.
 * The secondary-structure lines of a record ("struct") come from the
 * query's and the subject's SS states along the alignment, as the MAP
 * drawing writes them (cumapdp/CuBatchMAPDP_map_final.cu:249-300), when
 * SSSWGT > 0 and both carry SS states.
 */
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include "c2_lib.h"

namespace {

constexpr int kIndent = 13;       /* OUTPUTINDENT, liblib/mylimits.h:56 */
constexpr int kAnnotLen = 75;     /* ANNOTATIONLEN, liblib/mylimits.h:59 */
constexpr int kTmpBuffer = 1024;  /* AlnWriter szTmpBuffer */
constexpr int kWriterBuffer = 4096;
constexpr float kLogDpMax = 7.0978271289338397e+02f;   /* SLC_LOG_DP_MAX */
constexpr float kLn2 = 0.69314718055994530941723212146f;

void put(std::string& o, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
void put(std::string& o, const char* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    const int n = std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (n > 0) o.append(buf, (size_t)std::min(n, (int)sizeof buf - 1));
}

/* FormatDescription: blanks collapse to one space, a word that would reach
 * the width starts a new line indented by indent (only when width <
 * maxoutlen), output stops at maxoutlen and its last three characters
 * become "..." then */
void fmt_desc(std::string& o, const char* desc, int maxoutlen, int indent, int width, int& outpos, int& linepos,
              char addsep)
{
    while (*desc == ' ' || *desc == '\t' || *desc == '\n' || *desc == '\r') desc++;
    if (addsep) { o += addsep; outpos++; linepos++; }
    while (*desc && outpos < maxoutlen) {
        if (*desc == ' ' || *desc == '\t') {
            while (*desc == ' ' || *desc == '\t') desc++;
            if (width < maxoutlen) {
                int loc = linepos + 1;
                for (const char* p = desc; *p && *p != ' ' && *p != '\t' && loc < width; p++, loc++) {}
                if (width <= loc) {
                    o += '\n'; outpos++;
                    for (int i = 0; i < indent; i++, outpos++) o += ' ';
                    linepos = indent;
                } else {
                    o += ' '; outpos++; linepos++;
                }
            } else {
                o += ' '; outpos++; linepos++;
            }
            continue;
        }
        o += *desc++; outpos++; linepos++;
    }
    if (maxoutlen <= outpos && 3 < maxoutlen)
        for (size_t i = 1; i <= 3 && i <= o.size(); i++) o[o.size() - i] = '.';
}

uint32_t bits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }

/* the hits: the pairs under the log E threshold with an alignment drawn, in
 * the results' order; drawn_at maps a record to its row of strings */
struct Hits {
    std::vector<int> list;
    std::vector<int> drawn_at;
};

void collect_hits(const cbp_results* res, Hits& h)
{
    h.drawn_at.assign(res->npassed > 0 ? res->npassed : 0, -1);
    for (int j = 0; j < res->nmap; j++)
        if (res->map_rows[j] >= 0 && res->map_rows[j] < res->npassed) h.drawn_at[res->map_rows[j]] = j;
    for (int i = 0; i < res->npassed; i++) {
        const int k = res->order[i];
        const float* r = res->alndata + (size_t)k * CBP_NFIELDS;
        if (res->logevthld < r[CBP_F_EVALUE]) continue;
        if ((int32_t)bits(r[CBP_F_EPA]) < 1) continue;
        h.list.push_back(k);
    }
}

/* the query, middle and subject strings of record k; false if it has none
 * or they are shorter than its alignment length */
bool record_rows(const cbp_results* res, const Hits& h, int k, int alnlen, const char*& rq, const char*& rm,
                 const char*& rt)
{
    const int j = h.drawn_at[k];
    if (j < 0 || !res->alns) return false;
    rq = res->alns + ((size_t)j * 3 + 0) * res->alnstride;
    rm = res->alns + ((size_t)j * 3 + 1) * res->alnstride;
    rt = res->alns + ((size_t)j * 3 + 2) * res->alnstride;
    return (int)strnlen(rq, (size_t)res->alnstride) >= alnlen;
}

/* SS rows along an alignment: a gap takes ' ', a residue its position's
 * state (cumapdp/CuBatchMAPDP_map_final.cu:249-300) */
void ss_rows(const cbp_plain_info* in, uint32_t pro, unsigned qrybeg, unsigned trgbeg, const char* rq,
             const char* rt, int alnlen, std::string& qss, std::string& tss)
{
    const size_t t0 = (size_t)in->pro_dist[pro];
    size_t qy = qrybeg - 1, tx = trgbeg - 1;
    for (int c = 0; c < alnlen; c++) {
        qss += rq[c] == '-' ? ' ' : in->query_sss[qy];
        tss += rt[c] == '-' ? ' ' : in->pro_sss[t0 + tx];
        if (rq[c] != '-') qy++;
        if (rt[c] != '-') tx++;
    }
}

double evalue_of(float le) { return kLogDpMax < le ? 1.0e+308 : exp((double)le); }

/* bit score and P-value of a record (GetBitScore, GetPvalue) */
float bit_score(float sspace, float logeval) { return sspace <= 0.0f ? 0.0f : (logf(sspace) - logeval) / kLn2; }
double pvalue(double evalue) { return evalue < 0.01 ? evalue : 1.0 - exp(-evalue); }

/* FormatDescriptionJSON (liblib/fmtdescription.h:96-126): leading blanks
 * dropped, runs of blanks one space, '"', '\\' and '/' escaped by a
 * backslash; at maxoutlen the last three characters, and backslashes just
 * before them, become '.' */
void fmt_desc_json(std::string& o, const char* desc, int maxoutlen, int& outpos)
{
    while (*desc == ' ' || *desc == '\t' || *desc == '\n' || *desc == '\r') desc++;
    while (*desc && outpos < maxoutlen) {
        if (*desc == ' ' || *desc == '\t') {
            while (*desc == ' ' || *desc == '\t') desc++;
            o += ' '; outpos++;
            continue;
        }
        if (*desc == '"' || *desc == '\\' || *desc == '/') {
            o += '\\'; outpos++;
            if (maxoutlen <= outpos) break;
        }
        o += *desc++; outpos++;
    }
    if (maxoutlen <= outpos && 3 < maxoutlen) {
        /* the reference steps back over the buffer it wrote, outpos
         * characters of which may precede this call's text */
        size_t i = 1;
        for (; i <= 3 && i <= o.size(); i++) o[o.size() - i] = '.';
        for (; i <= (size_t)outpos && i <= o.size() && o[o.size() - i] == '\\'; i++) o[o.size() - i] = '.';
    }
}

int finish(const std::string& o, char* out, size_t cap, size_t* len)
{
    *len = o.size();
    if (out) {
        if (cap < o.size() + 1) return CBP_ERR_ARG;
        std::memcpy(out, o.data(), o.size());
        out[o.size()] = 0;
    }
    return CBP_OK;
}

}  // namespace

extern "C" int cbp_format_plain(const cbp_options* opt, const cbp_plain_info* in, const cbp_results* res,
                                char* out, size_t cap, size_t* len)
{
    if (!opt || !in || !res || !len || in->query_len < 1 || in->dsclen < 1 || in->dscwidth < 1 || in->alnwidth < 1)
        return CBP_ERR_ARG;
    if (res->npassed > 0 && (!res->alndata || !res->order)) return CBP_ERR_ARG;
    if (res->npassed > 0 && (!in->pro_desc || !in->pro_len || !in->pro_eno)) return CBP_ERR_ARG;
    std::string o;
    try {
        /* the programme and its references */
        {
            size_t size0 = 0;
            if (in->progname) put(o, "%s", in->progname);
            if (in->progversion) put(o, " %s", in->progversion);
            if (in->progname || in->progversion) o += "\n\n";
            size0 = o.size();
            const int maxsize = kWriterBuffer - (int)size0 - 4;
            if (maxsize >= 1) {
                int outpos = 0, linepos = 0;
                for (int i = 0; in->references && in->references[i] && outpos < maxsize; i++) {
                    fmt_desc(o, in->references[i], maxsize, 0, in->dscwidth, outpos, linepos, 0);
                    o += '\n'; outpos++;
                    linepos = 0;
                }
                o += '\n';
            }
        }
        /* the query */
        {
            int left = kWriterBuffer - (int)o.size();
            if (left > in->dsclen) left = in->dsclen;
            if (left >= 40) {
                const size_t b = o.size();
                put(o, "\n Query (%d positions):\n", in->query_len);
                const int maxsize = left - (int)(o.size() - b);
                if (maxsize >= 1 && in->query_desc) {
                    int outpos = 0, linepos = 0;
                    fmt_desc(o, in->query_desc, maxsize, 0, in->dscwidth, outpos, linepos, 0);
                }
            }
        }
        /* the hits: drawn pairs under the threshold, in the sorted order */
        Hits H;
        collect_hits(res, H);
        const std::vector<int>& hits = H.list;
        /* database lines and the heading of the hit list (limited to the
         * writer's 1 KB scratch, with "..." when it runs out) */
        {
            std::string s;
            const int maxsize = kTmpBuffer;
            bool stop = false;
            auto room = [&](int delta) {
                if (maxsize < (int)s.size() + delta) {
                    if ((int)s.size() + 5 < maxsize) s += "...\n\n";
                    return false;
                }
                return true;
            };
            if (room(30)) {
                s += "\n\n Database(s):\n";
                for (int i = 0; i < in->n_db_names && !stop; i++) {
                    const std::string nm = in->db_names[i] ? in->db_names[i] : "";
                    size_t sz = nm.size() < 100 ? nm.size() : 100;
                    if (!room((int)sz + 2)) { stop = true; break; }
                    s.append(nm, 0, sz);
                    if (sz < nm.size() && 3 < sz)
                        for (size_t k = 1; k <= 3; k++) s[s.size() - k] = '.';
                    s += '\n';
                }
                if (!stop && room(kIndent + 40)) {
                    s.append(kIndent, ' ');
                    put(s, "%zu profiles\n", (size_t)in->db_entries);
                    if (room(kIndent + 50)) {
                        s.append(kIndent, ' ');
                        put(s, "%zu total positions\n\n\n", (size_t)in->db_size);
                        if (room(kAnnotLen * 2)) {
                            if (!hits.empty()) {
                                const size_t b = s.size();
                                s += " Profiles found below the e-value threshold:";
                                while (s.size() - b < (size_t)kAnnotLen) s += ' ';
                                put(s, " %7s %7s\n\n", "Score", "E-value");
                            } else {
                                put(s, " No profiles found below an e-value threshold of %g.\n\n\n",
                                    exp((double)res->logevthld));
                            }
                        }
                    }
                }
            }
            o += s;
        }
        const float sspace = (float)res->search_space;
        if (!hits.empty()) {
            /* annotations */
            for (size_t i = 0; i < hits.size() && (in->nohits < 0 || i < (size_t)in->nohits); i++) {
                const int k = hits[i];
                const float* r = res->alndata + (size_t)k * CBP_NFIELDS;
                const uint32_t pro = bits(r[CBP_F_ORGPRONO]);
                const size_t b = o.size();
                int outpos = 0, linepos = 0;
                if (in->pro_desc[pro]) fmt_desc(o, in->pro_desc[pro], kAnnotLen, 0, kAnnotLen, outpos, linepos, 0);
                while (o.size() - b < (size_t)kAnnotLen) o += ' ';
                put(o, " %7.0f %7.1g\n", r[CBP_F_SCORE], evalue_of(r[CBP_F_EVALUE]));
            }
            o += "\n\n\n";
            /* alignment records */
            const bool printsss = opt->ssswgt > 0.0f && in->query_sss && in->pro_sss;
            const int width = kIndent < in->dscwidth ? in->dscwidth : 1;
            for (size_t i = 0; i < hits.size() && (in->noalns < 0 || i < (size_t)in->noalns); i++) {
                const int k = hits[i];
                const float* r = res->alndata + (size_t)k * CBP_NFIELDS;
                const uint32_t pro = bits(r[CBP_F_ORGPRONO]);
                const int alnlen = (int32_t)bits(r[CBP_F_EPA]);
                const float logeval = r[CBP_F_EVALUE];
                const double evalue = evalue_of(logeval);
                const float score = r[CBP_F_SCORE];
                const int dbprolen = in->pro_len[pro];
                {
                    int outpos = 0, linepos = 0;
                    if (in->pro_desc[pro]) fmt_desc(o, in->pro_desc[pro], in->dsclen, kIndent, width, outpos, linepos, '>');
                    o += '\n';
                }
                /* scores (FormatScoresPlain) */
                const int psts = (int)r[CBP_F_PSTVS], idts = (int)r[CBP_F_IDNTS], gaps = (int)r[CBP_F_NGAPS];
                put(o, "  Length/ENO: Query = %d/%.1f, Sbjct = %d/%.1f\n\n", in->query_len, in->query_eno, dbprolen,
                    in->pro_eno[pro]);
                put(o, " Score = %.2f (%.1f bits),  Expect = %.2g, P-value = %.2g\n", score, bit_score(sspace, logeval),
                    evalue, pvalue(evalue));
                if (idts > 0) put(o, " Identities = %d/%d (%d%%)", idts, alnlen, idts * 100 / alnlen);
                if (psts > 0) {
                    if (idts) o += ",";
                    put(o, " Positives = %d/%d (%d%%)", psts, alnlen, psts * 100 / alnlen);
                }
                if (gaps > 0) {
                    if (idts || psts) o += ",";
                    put(o, " Gaps = %d/%d (%d%%)", gaps, alnlen, gaps * 100 / alnlen);
                }
                o += "\n\n";
                /* the alignment (FormatAlignmentPlain) */
                const char *rq, *rm, *rt;
                if (!record_rows(res, H, k, alnlen, rq, rm, rt)) return CBP_ERR_ARG;
                const uint32_t bc = bits(r[CBP_F_BEGCOORDS]);
                unsigned qrybeg = (bc & 0xFFFFu) + 1, trgbeg = (bc >> 16) + 1;
                std::string qss, tss;
                if (printsss) ss_rows(in, pro, qrybeg, trgbeg, rq, rt, alnlen, qss, tss);
                const int aw = in->alnwidth;
                for (int f = 0; f < alnlen; f += aw) {
                    int nbytes = alnlen - f;
                    if (aw < nbytes) nbytes = aw;
                    if (printsss) { put(o, "%-13s", "struct"); o.append(qss, f, nbytes); o += '\n'; }
                    {
                        put(o, "Query: %5u ", qrybeg);
                        o.append(rq + f, nbytes);
                        const int fg = (int)std::count(rq + f, rq + f + nbytes, '-');
                        qrybeg += nbytes - fg;
                        put(o, " %-5d\n", nbytes <= fg ? (int)qrybeg : (int)qrybeg - 1);
                    }
                    { put(o, "%13c", ' '); o.append(rm + f, nbytes); o += '\n'; }
                    {
                        put(o, "Sbjct: %5u ", trgbeg);
                        o.append(rt + f, nbytes);
                        const int fg = (int)std::count(rt + f, rt + f + nbytes, '-');
                        trgbeg += nbytes - fg;
                        put(o, " %-5d\n", nbytes <= fg ? (int)trgbeg : (int)trgbeg - 1);
                    }
                    if (printsss) { put(o, "%-13s", "struct"); o.append(tss, f, nbytes); o += '\n'; }
                    o += '\n';
                }
                /* statistics footer (FormatFooterPlain) */
                if (in->show) {
                    const float K = r[CBP_F_K], lmbd = r[CBP_F_LMBD], Ke = r[CBP_F_KEST], le = r[CBP_F_LMBDEST];
                    const float expected = r[CBP_F_E], entropy = r[CBP_F_H];
                    char kb[64], lb[64];
                    put(o, "%25c  %-6s   %-6s\n", ' ', "K", "Lambda");
                    const char* kp = "n/a";
                    const char* lp = "n/a";
                    if (expected < 0.0f) {
                        if (K > 0.0f) { std::snprintf(kb, sizeof kb, "%6.4f", K); kp = kb; }
                        if (lmbd > 0.0f) { std::snprintf(lb, sizeof lb, "%6.4f", lmbd); lp = lb; }
                    }
                    put(o, "%-25s  %6s   %6s\n", "Computed  ungapped,", kp, lp);
                    kp = lp = "n/a";
                    if (expected < 0.0f) {
                        if (Ke > 0.0f) { std::snprintf(kb, sizeof kb, "%6.4f", Ke); kp = kb; }
                        if (le > 0.0f) { std::snprintf(lb, sizeof lb, "%6.4f", le); lp = lb; }
                    }
                    put(o, "%-25s  %6s   %6s\n", "Estimated gapped,", kp, lp);
                    put(o, "Entropy, %6.4f; Expected, %6.4f; Min/Max, %.0f/%-.0f\n\n", entropy, expected,
                        r[CBP_F_MIN], r[CBP_F_MAX]);
                }
                o += "\n\n";
            }
        }
        /* summary (WriteSummaryPlain) */
        {
            int effq = in->query_len;
            uint64_t effdb = in->db_size;
            if (in->deltalen < in->query_len) effq -= in->deltalen;
            if (in->db_entries * (uint64_t)in->deltalen < in->db_size) effdb -= in->db_entries * (uint64_t)in->deltalen;
            put(o, "%-20s  %-6s   %-6s\n", "Reference values of", "K", "Lambda");
            put(o, "%-20s  %6.4f   %6.4f\n", "Ungapped", opt->ung_K, opt->ung_lambda);
            put(o, "%-20s  %6.4f   %6.4f\n\n", "Gapped", opt->gap_K, opt->gap_lambda);
            put(o, "Length of query, %d\n", in->query_len);
            put(o, "Length of database, %zu\n", (size_t)in->db_size);
            put(o, "Effective length of query, %d\n", effq);
            put(o, "Effective length of database, %zu\n", (size_t)effdb);
            put(o, "Effective search space, %zu\n\n", (size_t)sspace);
        }
    } catch (const std::bad_alloc&) {
        return CBP_ERR_MEM;
    }
    return finish(o, out, cap, len);
}

/* The JSON output file of a query (AlnWriter::WriteResultsJSON,
 * cualn/AlnWriterJSON.cpp:20-143): the programme and references
 * (WritePrognameJSON, :151-197), the query (WriteQueryDescriptionJSON,
 * :210-265), the database and message (WriteSearchInformationJSON,
 * :280-360), the summary entries (MakeAnnotationJSON,
 * cupro/CuBatchProcessingFinalizerJSON.cpp:165-195), the hit records
 * (CompressResultsJSON, FormatScoresJSON, FormatAlignmentJSON,
 * FormatFooterJSON, :35-382) and the search statistics (WriteSummaryJSON,
 * cualn/AlnWriterJSON.cpp:375-427).  The records' separators follow the
 * writer: a record is followed by ",\n" unless it is the last of the hits
 * or of the first NOHITS (for the records as well as the summary entries,
 * so NOALNS < NOHITS leaves a trailing comma, as the reference does). */
extern "C" int cbp_format_json(const cbp_options* opt, const cbp_plain_info* in, const cbp_results* res,
                               char* out, size_t cap, size_t* len)
{
    if (!opt || !in || !res || !len || in->query_len < 1 || in->dsclen < 1) return CBP_ERR_ARG;
    if (res->npassed > 0 && (!res->alndata || !res->order)) return CBP_ERR_ARG;
    if (res->npassed > 0 && (!in->pro_desc || !in->pro_len || !in->pro_eno)) return CBP_ERR_ARG;
    std::string o;
    try {
        /* the programme and its references: written when 1 KB for them
         * fits the writer's buffer, at most 964 bytes of reference text */
        {
            put(o, "{\"comer_search\": {\n  \"program\": \"%s\",\n  \"version\": \"%s\",\n",
                in->progname ? in->progname : "", in->progversion ? in->progversion : "");
            const int maxsize = kWriterBuffer - (int)o.size() - 1024;
            if (maxsize >= 1) {
                static const char* const none[1] = {nullptr};
                const char* const* refs = in->references ? in->references : none;
                o += "  \"references\": [\n";
                int totlen = 0;
                bool bwrt = true;
                /* one line at least (empty when there are no references);
                 * the list ends at its first null entry, which is never read
                 * past */
                for (int i = 0; !i || (bwrt && refs[i - 1] && refs[i]); i++) {
                    const char* r = refs[i];
                    totlen += r ? (int)std::strlen(r) : 0;
                    bwrt = totlen < 1024 - 60;
                    put(o, "    \"%s\"%s\n", (bwrt && r) ? r : "", (bwrt && r && refs[i + 1]) ? "," : "");
                }
                o += "  ],\n";
            }
        }
        /* the query */
        {
            int left = kWriterBuffer - (int)o.size();
            if (left > in->dsclen) left = in->dsclen;
            if (left >= 60) {
                const size_t b = o.size();
                put(o, "  \"query\": {\n    \"length\": %d,\n", in->query_len);
                const int maxsize = left - (int)(o.size() - b) - 256;
                if (maxsize >= 1) {
                    o += "    \"name\": \"\",\n    \"description\": \"";
                    int outpos = 36;
                    if (in->query_desc) fmt_desc_json(o, in->query_desc, maxsize, outpos);
                    o += "\"\n  },\n";
                }
            }
        }
        Hits H;
        collect_hits(res, H);
        const std::vector<int>& hits = H.list;
        /* the database and the message, within the writer's 1 KB scratch */
        {
            int maxsize = kTmpBuffer;
            auto lfn = [&maxsize](int d) { maxsize -= d; return maxsize < 1; };
            if (!lfn(40 * 7)) {
                o += "  \"database\": {\n    \"names\": [\n";
                for (int i = 0; i < in->n_db_names; i++) {
                    const char* nm = in->db_names && in->db_names[i] ? in->db_names[i] : "";
                    size_t sz = std::strlen(nm);
                    if (99 < sz) sz = 99;
                    if (lfn((int)sz + 15)) { o += "      \"...\"\n"; break; }
                    o += "      \"";
                    int outpos = 0;
                    fmt_desc_json(o, nm, (int)sz + 1, outpos);
                    o += i + 1 < in->n_db_names ? "\",\n" : "\"\n";
                }
                put(o, "    ],\n    \"number_of_profiles\": %zu,\n    \"number_of_positions\": %zu\n  },\n",
                    (size_t)in->db_entries, (size_t)in->db_size);
                if (!lfn(90 * 4)) {
                    if (!hits.empty())
                        o += "  \"message\": \"Profiles found below the e-value threshold:\",\n";
                    else
                        put(o, "  \"message\": \"No profiles found below an e-value threshold of %g\",\n",
                            exp((double)res->logevthld));
                }
            }
        }
        const size_t nh = in->nohits < 0 ? hits.size() : std::min(hits.size(), (size_t)in->nohits);
        const size_t na = in->noalns < 0 ? hits.size() : std::min(hits.size(), (size_t)in->noalns);
        auto sep = [&](size_t i) { return (i + 1 < hits.size() && i + 1 < nh) ? ",\n" : "\n"; };
        /* the summary entries */
        o += "  \"search_summary\": [\n";
        for (size_t i = 0; i < nh; i++) {
            const float* r = res->alndata + (size_t)hits[i] * CBP_NFIELDS;
            const uint32_t pro = bits(r[CBP_F_ORGPRONO]);
            o += "    {\"summary_entry\": {\n      \"description\": \"";
            int outpos = 0;
            if (in->pro_desc[pro]) fmt_desc_json(o, in->pro_desc[pro], kAnnotLen, outpos);
            put(o, "\",\n      \"score\": %.0f,\n      \"evalue\": %.2g\n    }}", r[CBP_F_SCORE],
                evalue_of(r[CBP_F_EVALUE]));
            o += sep(i);
        }
        o += "  ],\n  \"search_hits\": [\n";
        /* the hit records */
        const float sspace = (float)res->search_space;
        const bool qsss = in->query_sss != nullptr;
        for (size_t i = 0; i < na; i++) {
            const int k = hits[i];
            const float* r = res->alndata + (size_t)k * CBP_NFIELDS;
            const uint32_t pro = bits(r[CBP_F_ORGPRONO]);
            const int alnlen = (int32_t)bits(r[CBP_F_EPA]);
            const float logeval = r[CBP_F_EVALUE];
            const double evalue = evalue_of(logeval);
            const float score = r[CBP_F_SCORE];
            o += "    {\"hit_record\": {\n      \"target_description\": \"";
            {
                int outpos = 0;
                if (in->pro_desc[pro]) fmt_desc_json(o, in->pro_desc[pro], in->dsclen, outpos);
            }
            put(o, "\",\n      \"query_length\": %d,\n      \"query_eno\": %.1f,\n      \"target_length\": %d,\n"
                   "      \"target_eno\": %.1f,\n      \"alignment\": {\n",
                in->query_len, in->query_eno, in->pro_len[pro], in->pro_eno[pro]);
            /* scores (FormatScoresJSON) */
            put(o, "        \"score\": %.2f,\n        \"bit_score\": %.1f,\n        \"evalue\": %.2g,\n"
                   "        \"pvalue\": %.2g,\n",
                score, bit_score(sspace, logeval), evalue, pvalue(evalue));
            put(o, "        \"n_identities\": %d,\n        \"n_positives\": %d,\n        \"n_gaps\": %d,\n"
                   "        \"aln_length\": %d,\n",
                (int)r[CBP_F_IDNTS], (int)r[CBP_F_PSTVS], (int)r[CBP_F_NGAPS], alnlen);
            /* the alignment (FormatAlignmentJSON) */
            const char *rq, *rm, *rt;
            if (!record_rows(res, H, k, alnlen, rq, rm, rt)) return CBP_ERR_ARG;
            const uint32_t bc = bits(r[CBP_F_BEGCOORDS]);
            const unsigned qrybeg = (bc & 0xFFFFu) + 1, trgbeg = (bc >> 16) + 1;
            const int qg = (int)std::count(rq, rq + alnlen, '-'), tg = (int)std::count(rt, rt + alnlen, '-');
            const unsigned qryend = qrybeg + alnlen - qg, trgend = trgbeg + alnlen - tg;
            /* the flags say whether both carry SS states; the states are
             * drawn along the alignment only under SSSWGT > 0
             * (cumapdp/CuBatchMAPDP.cu:440), so the strings are empty
             * otherwise */
            const bool printsss = qsss && in->pro_sss;
            const bool drawsss = printsss && opt->ssswgt > 0.0f;
            put(o, "        \"query_from\": %u,\n        \"query_to\": %u,\n        \"target_from\": %u,\n"
                   "        \"target_to\": %u,\n",
                qrybeg, alnlen <= qg ? qryend : qryend - 1, trgbeg, alnlen <= tg ? trgend : trgend - 1);
            put(o, "        \"query_secstr_available\": %d,\n        \"target_secstr_available\": %d,\n"
                   "        \"query_secstr\": \"",
                (int)printsss, (int)printsss);
            std::string qss, tss;
            if (drawsss) ss_rows(in, pro, qrybeg, trgbeg, rq, rt, alnlen, qss, tss);
            o += qss;
            o += "\",\n        \"target_secstr\": \"";
            o += tss;
            o += "\",\n        \"query_aln\": \"";
            o.append(rq, alnlen);
            o += "\",\n        \"target_aln\": \"";
            o.append(rt, alnlen);
            o += "\",\n        \"middle\": \"";
            o.append(rm, alnlen);
            o += "\",\n";
            /* statistics (FormatFooterJSON) */
            {
                const float K = r[CBP_F_K], lmbd = r[CBP_F_LMBD], Ke = r[CBP_F_KEST], le = r[CBP_F_LMBDEST];
                const float expected = r[CBP_F_E], entropy = r[CBP_F_H];
                char kb[64], lb[64];
                const char* kp = "NA";
                const char* lp = "NA";
                if (expected < 0.0f) {
                    if (K > 0.0f) { std::snprintf(kb, sizeof kb, "%6.4f", K); kp = kb; }
                    if (lmbd > 0.0f) { std::snprintf(lb, sizeof lb, "%6.4f", lmbd); lp = lb; }
                }
                put(o, "        \"computed_K_ungapped\": \"%s\",\n        \"computed_lambda_ungapped\": \"%s\",\n", kp,
                    lp);
                kp = lp = "NA";
                if (expected < 0.0f) {
                    if (Ke > 0.0f) { std::snprintf(kb, sizeof kb, "%6.4f", Ke); kp = kb; }
                    if (le > 0.0f) { std::snprintf(lb, sizeof lb, "%6.4f", le); lp = lb; }
                }
                put(o, "        \"estimated_K_gapped\": \"%s\",\n        \"estimated_lambda_gapped\": \"%s\",\n", kp,
                    lp);
                put(o, "        \"entropy\": %.4f,\n        \"expected_score\": %.4f,\n        \"min_score\": %.0f,\n"
                       "        \"max_score\": %.0f\n",
                    entropy, expected, r[CBP_F_MIN], r[CBP_F_MAX]);
            }
            o += "      }\n    }}";
            o += sep(i);
        }
        o += "  ],\n";
        /* search statistics (WriteSummaryJSON) */
        {
            int effq = in->query_len;
            uint64_t effdb = in->db_size;
            if (in->deltalen < in->query_len) effq -= in->deltalen;
            if (in->db_entries * (uint64_t)in->deltalen < in->db_size) effdb -= in->db_entries * (uint64_t)in->deltalen;
            put(o, "  \"search_statistics\": {\n    \"reference_K_ungapped\": %.4f,\n"
                   "    \"reference_lambda_ungapped\": %.4f,\n    \"reference_K_gapped\": %.4f,\n"
                   "    \"reference_lambda_gapped\": %.4f,\n",
                opt->ung_K, opt->ung_lambda, opt->gap_K, opt->gap_lambda);
            put(o, "    \"query_length\": %d,\n    \"database_size\": %zu,\n    \"effective_query_length\": %d,\n"
                   "    \"effective_database_size\": %zu,\n    \"effective_search_space\": %zu\n  }\n}}\n",
                in->query_len, (size_t)in->db_size, effq, (size_t)effdb, (size_t)sspace);
        }
    } catch (const std::bad_alloc&) {
        return CBP_ERR_MEM;
    }
    return finish(o, out, cap, len);
}
