/*
 * c2_oracle_search.c -- the whole per-pair search pipeline on the CPU, used
 * as the timed CPU baseline and as a coarse end-to-end checker.
 *
 * TEST INFRASTRUCTURE ONLY (see c2_oracle.h).
 *
 * For one query against a list of DB profiles, each pair runs the stage
 * sequence of CuBatchProcessing::BatchProcessScoreMatrixDevice
 * (cupro/CuBatchProcessing.cu:229-572) with this oracle's FP32 restatements:
 *   score matrix + SS term             (or_score_f32)
 *   phase-1 Viterbi with backtrack     (or_viterbi_f32)
 *   finalize if max > threshold        (or_finalize)
 *   histogram + statistics if passed   (or_histogram, or_statistics)
 *   MAP realignment if log E <= thld   (or_map_direct)
 * Pairs are independent, so profiles are spread over OpenMP threads with a
 * dynamic schedule; the result does not depend on the thread count.
 */
#include <stdlib.h>
#include <string.h>
#include <omp.h>
#include "c2_oracle.h"

or_search_out or_search(const cbp_profiles* q, int qp, const cbp_profiles* d,
                        const int32_t* pros, int32_t npros, const or_tables* t,
                        const cbp_options* o, float scorethld, float logevthld,
                        float searchspace, int nthreads)
{
    or_search_out out;
    memset(&out, 0, sizeof out);
    const int Lq = q->len[qp];
    const char* qaa = q->aa + q->dist[qp];
    int64_t cells = 0, mcells = 0;
    int32_t npassed = 0, nmap = 0;
    double csum = 0.0;
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads) \
    reduction(+ : cells, mcells, npassed, nmap, csum)
    for (int32_t i = 0; i < npros; i++) {
        const int dp = pros[i];
        const int Ld = d->len[dp];
        const size_t n = (size_t)Lq * Ld;
        float* S = (float*)malloc(sizeof(float) * n);
        float* M = (float*)malloc(sizeof(float) * n);
        char* bt = (char*)malloc(n);
        or_score_f32(q, qp, d, dp, t, o, S, M);
        const or_dpmax m = or_viterbi_f32(q, qp, d, dp, S, bt, NULL);
        cells += (int64_t)n;
        if (m.maxsc > scorethld) {
            const or_fin f = or_finalize(Lq, Ld, S, bt, m);
            if (scorethld < f.score) {
                int64_t cnt[641];
                npassed++;
                or_histogram(S, Lq, Ld, qaa, d->aa + d->dist[dp], o->xuninf, cnt);
                const or_stats st = or_statistics(cnt, Lq, Ld, q->eno[qp], d->eno[dp], f.score,
                                                  (float)f.psts, searchspace, o, t->nnw);
                csum += st.Evalue;
                if (o->mapaln && st.Evalue <= logevthld) {
                    const or_map r = or_map_direct(q, qp, d, dp, S, M, (float)st.EpA, o->minpp,
                                                   o->ssswgt > 0.0f || o->cvswgt > 0.0f, NULL, NULL, bt, NULL);
                    nmap++;
                    mcells += (int64_t)n;
                    csum += r.alnlen;
                }
            }
        }
        free(S); free(M); free(bt);
    }
    out.cells = cells;
    out.map_cells = mcells;
    out.npassed = npassed;
    out.nmap = nmap;
    out.checksum = csum;
    return out;
}
