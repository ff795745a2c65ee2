/*
 * c2_dpipe.cpp -- the per-(chunk, query) stage sequence of
 * CuBatchProcessing::BatchProcessScoreMatrixDevice
 * (cupro/CuBatchProcessing.cu:229-572) with every stage list built on the
 * device (c2_dlist.hip): phase 1 (the batch's work queue, or one launch for a
 * query on its own), then per query, on its stream slot, the candidate list,
 * the backtrack pass from a device-built tile queue, finalize, the pass
 * list, histograms, statistics, the (log E, profile) order, the selection of
 * the pairs to draw, the MAP plan and passes (or the phase-1 drawing of
 * MAPALN=0) and the records.  No host decision separates the stages, so the
 * whole sequence of a batch is one HIP graph when graphs are on.  The host
 * reads each query's counters once at the end and copies exactly the
 * results' sizes.
 *
 * Lists and pools have fixed capacities (DlCaps); a call in which any query
 * exceeds one runs again through the host-decided sequence (c2_pipe.cpp), so
 * the results never depend on the capacities.
 */
#include <cstring>
#include <algorithm>
#include "c2_lib.h"

#define DCHK(call)                                                               \
    do {                                                                         \
        hipError_t e_ = (call);                                                  \
        if (e_ != hipSuccess)                                                    \
            return c2_fail(h, CBP_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

namespace {

struct DlCaps {
    int32_t capC, capP, capM;
    int64_t capW, capItems, capCells, capHand, stride;
    int mR, mns, mnw;
};

DlCaps dl_caps(cbp_handle* h, const std::vector<int>& lq, const std::vector<int>& nstripes)
{
    DlCaps k{};
    const int np = h->dchunk.nprof, maxlen = h->chunk_maxlen > 0 ? h->chunk_maxlen : 1;
    int Lmax = 1, smax = 1;
    for (int L : lq) Lmax = std::max(Lmax, L);
    for (int s : nstripes) smax = std::max(smax, s);
    k.capC = std::max(1, std::min(h->dl_capC, np));
    k.capP = std::max(1, std::min(h->dl_capP, k.capC));
    k.capM = h->opt.max_map_pairs > 0 ? std::min(h->opt.max_map_pairs, k.capP) : k.capP;
    /* no more code words than every profile of the chunk passing would take
     * (stripes x (len + 63) x 64 each), so a small chunk holds a small pool */
    const int64_t allw = (int64_t)smax * ((int64_t)h->dchunk.npos + 63 * (int64_t)np) * 64;
    k.capW = std::max<int64_t>(64, std::min<int64_t>(h->dl_capW, allw));
    c2_map_geometry(Lmax, &k.mR, &k.mns, &k.mnw);
    const int64_t chunks = (maxlen + C2_MAP_SCORE_COLS - 1) / C2_MAP_SCORE_COLS;
    k.capItems = (int64_t)k.capM * 8 * chunks;
    int64_t cells = 0, hand = 0;
    for (int L : lq) {
        int R = 1, ns = 1, nw = 1;
        c2_map_geometry(L, &R, &ns, &nw);
        cells = std::max(cells, c2_map_cells(R, ns, maxlen));
        hand = std::max(hand, (c2_map_hand(ns, maxlen) + 63) & ~(int64_t)63);
        if (h->map_rmax2 > 0) {
            c2_map_geometry_short(L, h->map_rmax2, &R, &ns, &nw);
            cells = std::max(cells, c2_map_cells(R, ns, maxlen));
            hand = std::max(hand, (c2_map_hand(ns, maxlen) + 63) & ~(int64_t)63);
        }
    }
    k.capCells = std::min<int64_t>(h->dl_capCells, (int64_t)k.capM * cells);
    k.capHand = std::min<int64_t>((int64_t)1 << 26, (int64_t)k.capM * hand);
    k.stride = (int64_t)Lmax + maxlen + 1;
    return k;
}

int dl_alloc_query(cbp_handle* h, DlQuery& q, const DlCaps& k)
{
    int rc;
    if ((rc = c2_dalloc(h, q.cnt, sizeof(int32_t) * DL_NCNT))) return rc;
    if (!q.hcnt) DCHK(hipHostMalloc((void**)&q.hcnt, sizeof(int32_t) * DL_NCNT, hipHostMallocDefault));
    if ((rc = c2_dalloc(h, q.aln, sizeof(float) * CBP_NFIELDS * (size_t)k.capP))) return rc;
    if ((rc = c2_dalloc(h, q.epa, sizeof(float) * (size_t)k.capP))) return rc;
    if ((rc = c2_dalloc(h, q.order, sizeof(int32_t) * (size_t)k.capP))) return rc;
    if ((rc = c2_dalloc(h, q.sel, sizeof(int32_t) * (size_t)k.capM))) return rc;
    if ((rc = c2_dalloc(h, q.res, sizeof(int32_t) * 8 * (size_t)k.capM))) return rc;
    if ((rc = c2_dalloc(h, q.p1save, sizeof(float) * 3 * (size_t)k.capM))) return rc;
    if ((rc = c2_dalloc(h, q.malns, (((size_t)k.capM * 3 * k.stride + 3) & ~(size_t)3)))) return rc;
    if ((rc = c2_dalloc(h, q.dargs, sizeof(P1Args)))) return rc;
    if ((rc = c2_dalloc(h, q.tiles, sizeof(P1Tile) * (size_t)((k.capC + 1) / 2)))) return rc;
    if ((rc = c2_dalloc(h, q.ntiles, sizeof(int32_t)))) return rc;
    return CBP_OK;
}

/* ahead: histogram rows for every candidate (one-query calls) */
int dl_alloc_slot(cbp_handle* h, DlSlot& s, const DlCaps& k, int64_t botw_floats, bool ahead)
{
    const size_t np = (size_t)h->dchunk.nprof, ntile = (np + DL_TILE - 1) / DL_TILE;
    int rc;
    if ((rc = c2_dalloc(h, s.tilecnt, sizeof(int32_t) * ntile))) return rc;
    if ((rc = c2_dalloc(h, s.tileoff, sizeof(int32_t) * ntile))) return rc;
    if ((rc = c2_dalloc(h, s.cand, sizeof(int32_t) * (size_t)k.capC))) return rc;
    if ((rc = c2_dalloc(h, s.pe, sizeof(int32_t) * (size_t)k.capC))) return rc;
    if ((rc = c2_dalloc(h, s.off, sizeof(int64_t) * ((size_t)k.capC + 1)))) return rc;
    if ((rc = c2_dalloc(h, s.btck, sizeof(uint32_t) * (size_t)k.capW))) return rc;
    if (botw_floats > 0 && (rc = c2_dalloc(h, s.botw, sizeof(float) * (size_t)botw_floats))) return rc;
    if ((rc = c2_dalloc(h, s.fin, sizeof(FinRec) * (size_t)k.capC))) return rc;
    if ((rc = c2_dalloc(h, s.pass_pro, sizeof(int32_t) * (size_t)k.capP))) return rc;
    if ((rc = c2_dalloc(h, s.passfin, sizeof(int32_t) * (size_t)k.capP))) return rc;
    if ((rc = c2_dalloc(h, s.statin, sizeof(StatIn) * (size_t)k.capP))) return rc;
    if ((rc = c2_dalloc(h, s.counts, sizeof(int32_t) * 641 * (size_t)(ahead ? k.capC : k.capP)))) return rc;
    if ((rc = c2_dalloc(h, s.statout, sizeof(double) * 10 * (size_t)k.capP))) return rc;
    if ((rc = c2_dalloc(h, s.mpairs, sizeof(MapPair) * (size_t)k.capM))) return rc;
    if ((rc = c2_dalloc(h, s.morder, sizeof(int32_t) * (size_t)k.capM))) return rc;
    if ((rc = c2_dalloc(h, s.sitems, sizeof(int4) * (size_t)k.capItems))) return rc;
    if ((rc = c2_dalloc(h, s.apro, sizeof(int32_t) * (size_t)k.capM))) return rc;
    if ((rc = c2_dalloc(h, s.acand, sizeof(int32_t) * (size_t)k.capM))) return rc;
    if ((rc = c2_dalloc(h, s.mstat, sizeof(float) * 4 * (size_t)k.capM))) return rc;
    if ((rc = c2_dalloc(h, s.malntmp, (size_t)k.capM * 3 * k.stride))) return rc;
    if (h->opt.mapaln) {
        if ((rc = c2_dalloc(h, s.mF, sizeof(float) * (size_t)k.capCells))) return rc;
        if ((rc = c2_dalloc(h, s.mB, sizeof(float) * (size_t)k.capCells))) return rc;
        if ((rc = c2_dalloc(h, s.mSd, sizeof(float) * (size_t)k.capCells))) return rc;
        if ((rc = c2_dalloc(h, s.mbt, (size_t)k.capCells))) return rc;
        if ((rc = c2_dalloc(h, s.mhand, sizeof(float) * (size_t)k.capHand))) return rc;
    }
    return CBP_OK;
}

DlBufs dl_bufs(DlQuery& q, DlSlot& s, const DlCaps& k)
{
    DlBufs b{};
    b.cnt = q.cnt.as<int32_t>();
    b.tilecnt = s.tilecnt.as<int32_t>(); b.tileoff = s.tileoff.as<int32_t>();
    b.cand = s.cand.as<int32_t>(); b.pe = s.pe.as<int32_t>(); b.off = s.off.as<int64_t>();
    b.capC = k.capC; b.capW = k.capW;
    b.pass_pro = s.pass_pro.as<int32_t>(); b.passfin = s.passfin.as<int32_t>(); b.statin = s.statin.as<StatIn>();
    b.aln = q.aln.as<float>(); b.epa = q.epa.as<float>(); b.order = q.order.as<int32_t>();
    b.capP = k.capP;
    b.sel = q.sel.as<int32_t>(); b.mpairs = s.mpairs.as<MapPair>(); b.morder = s.morder.as<int32_t>();
    b.sitems = s.sitems.as<int4>(); b.capM = k.capM; b.capItems = k.capItems;
    b.capCells = k.capCells; b.capHand = k.capHand;
    b.apro = s.apro.as<int32_t>(); b.acand = s.acand.as<int32_t>();
    b.res = q.res.as<int32_t>(); b.p1save = q.p1save.as<float>();
    return b;
}

/* per-query constants of the chain (host side, set before any launch) */
struct DlPlan {
    DlCaps k;
    std::vector<StatParams> sp;
    std::vector<int> grid;          /* backtrack-pass queue grid per query */
    std::vector<size_t> lds;
    bool split = false;
};

/* the chain of query i on its slot stream: every stage from the candidate
 * list to the records and the counters' copy to the host (capturable: no
 * host decision, no memset, no allocation) */
int dl_chain(cbp_handle* h, int i, const DlPlan& P)
{
    QWork& w = *h->work[i];
    DlQuery& q = *h->dlq[i];
    DlSlot& sl = h->dls[w.slot];
    const DlCaps& k = P.k;
    hipStream_t st = w.st;
    const DevSet& d = h->dchunk;
    DlBufs b = dl_bufs(q, sl, k);
    const QueryPrep& hp = w.dq->hp;
    DCHK(c2_launch_zero_u32((uint32_t*)q.cnt.p, DL_NCNT, st));
    DCHK(c2_dl_candidates(w.dk, d.nprof, w.scorethld, b, d.len.as<int32_t>(), hp.nstripes, q.dargs.as<P1Args>(),
                          q.tiles.as<P1Tile>(), 0, q.ntiles.as<int32_t>(), st));
    /* the backtrack pass: the query's candidates in pairs from its queue */
    P1QueueArgs qa;
    qa.args = q.dargs.as<P1Args>();
    qa.tiles = q.tiles.as<P1Tile>();
    qa.ntiles = 1;
    qa.ntiles_dev = q.ntiles.as<int32_t>();
    qa.head = (unsigned*)(q.cnt.as<int32_t>() + DL_PAD_);
    qa.botw = hp.nstripes > 1 ? sl.botw.as<float>() : nullptr;
    qa.bot_ld = h->chunk_maxlen;
    ScoreCtx sc;
    c2_fill_scorectx(h, w.dq, &sc);
    /* a query on its own: the candidates' histograms (they need only the
     * pairs) run on a second stream beside the backtrack pass and finalize,
     * and the statistics read the passed pairs' rows through passfin */
    const bool ahead = P.split;
    if (ahead) {
        DCHK(hipEventRecord(h->gfork, st));
        DCHK(hipStreamWaitEvent(h->xstream[0], h->gfork, 0));
        DCHK(c2_launch_hist(sc, sl.cand.as<int32_t>(), k.capC, h->opt.xuninf, sl.counts.as<int32_t>(), h->xstream[0],
                            q.cnt.as<int32_t>() + DL_NC));
        DCHK(hipEventRecord(h->gjoin[0], h->xstream[0]));
    }
    DCHK(c2_launch_phase1q(qa, hp.R, 2, true, P.grid[i], P.lds[i], st));
    DCHK(c2_launch_finalize(sc, sl.cand.as<int32_t>(), k.capC, w.dk, sl.btck.as<uint32_t>(), sl.off.as<int64_t>(),
                            sl.fin.as<FinRec>(), st, q.cnt.as<int32_t>() + DL_NC));
    DCHK(c2_dl_pass(sl.fin.as<FinRec>(), sl.cand.as<int32_t>(), d.len.as<int32_t>(), d.eno.as<float>(),
                    w.scorethld, b, st));
    if (ahead) {
        DCHK(hipStreamWaitEvent(st, h->gjoin[0], 0));
    } else {
        DCHK(c2_launch_hist(sc, sl.pass_pro.as<int32_t>(), k.capP, h->opt.xuninf, sl.counts.as<int32_t>(), st,
                            q.cnt.as<int32_t>() + DL_NPASS));
    }
    DCHK(c2_launch_stats(P.sp[i], h->dnnw.as<float>(), sl.counts.as<int32_t>(),
                         ahead ? sl.passfin.as<int32_t>() : nullptr, sl.statin.as<StatIn>(),
                         k.capP, sl.statout.as<double>(), st, q.cnt.as<int32_t>() + DL_NPASS));
    int mR = 1, mns = 1, mnw = 1;
    c2_map_geometry(hp.Lq, &mR, &mns, &mnw);
    /* the shorter pairs of a split call: fewer, busier waves per pair */
    int mR2 = mR, mns2 = mns, mnw2 = mnw;
    if (P.split && h->map_rmax2 > 0) c2_map_geometry_short(hp.Lq, h->map_rmax2, &mR2, &mns2, &mnw2);
    DlSelArgs sa{};
    sa.len = d.len.as<int32_t>();
    sa.logevthld = w.logevthld;
    sa.cap = h->opt.max_map_pairs;
    sa.mapaln = h->opt.mapaln;
    sa.mR = mR; sa.mns = mns;
    sa.mR2 = mR2; sa.mns2 = mns2;
    sa.modinuse = (h->opt.ssswgt > 0.0f || h->opt.cvswgt > 0.0f) ? 1 : 0;
    sa.split = P.split ? 1 : 0;
    DCHK(c2_dl_order_select(sl.statout.as<double>(), b, sa, st));
    /* strings are NUL-padded to the stride: the whole buffer is output */
    const int64_t alnwords = ((int64_t)k.capM * 3 * k.stride + 3) / 4;
    DCHK(c2_launch_zero_u32(q.malns.as<uint32_t>(), alnwords, st));
    if (h->opt.mapaln) {
        MapArgs ma;
        ma.pairs = sl.mpairs.as<MapPair>(); ma.trans = d.trans.as<float>();
        ma.qtr = w.dq->qtr.as<float>(); ma.qtrb = w.dq->qtrb.as<float>();
        ma.mleft = w.dq->mleft.as<float>(); ma.mdiag = w.dq->mdiag.as<float>();
        ma.F = sl.mF.as<float>(); ma.B = sl.mB.as<float>(); ma.Sd = sl.mSd.as<float>(); ma.hand = sl.mhand.as<float>();
        ma.btck = sl.mbt.as<char>(); ma.stat = sl.mstat.as<float>();
        ma.alns = q.malns.as<char>(); ma.alntmp = sl.malntmp.as<char>(); ma.alnstride = k.stride;
        ma.res = q.res.as<int32_t>(); ma.minpp = h->opt.minpp;
        ma.order = sl.morder.as<int32_t>(); ma.R = mR; ma.ns = mns; ma.nw = mnw;
        ma.sitems = sl.sitems.as<int4>(); ma.nsitems = (int32_t)std::min<int64_t>(k.capItems, 1 << 30);
        ma.dcnt = q.cnt.as<int32_t>();
        ma.pipe = h->map_pipe ? 1 : 0;
        if (P.split) {
            /* the long-pair group on a stream of its own, each group's
             * backward pass beside its forward pass */
            hipStream_t sA = h->xstream[0];
            DCHK(hipEventRecord(h->gfork, st));
            DCHK(hipStreamWaitEvent(sA, h->gfork, 0));
            MapArgs mA = ma, mB = ma;
            mA.group = 0; mB.group = 1;
            mB.R = mR2; mB.ns = mns2; mB.nw = mnw2;
            DCHK(c2_launch_map(sc, mA, k.capM, sA, h->xstream[1], h->mfork[0], h->mjoin[0]));
            DCHK(c2_launch_map(sc, mB, k.capM, st, h->xstream[2], h->mfork[1], h->mjoin[1]));
            DCHK(hipEventRecord(h->gjoin[0], sA));
            DCHK(hipStreamWaitEvent(st, h->gjoin[0], 0));
        } else {
            ma.group = -1;
            DCHK(c2_launch_map(sc, ma, k.capM, st, nullptr, nullptr, nullptr));
        }
    } else {
        DCHK(c2_dl_aln_lists(b, st));
        AlnArgs a;
        a.pro = sl.apro.as<int32_t>(); a.cand = sl.acand.as<int32_t>(); a.keys = w.dk;
        a.btck = sl.btck.as<uint32_t>(); a.btck_off = sl.off.as<int64_t>();
        a.alns = q.malns.as<char>(); a.alntmp = sl.malntmp.as<char>(); a.alnstride = k.stride;
        a.res = q.res.as<int32_t>(); a.dcnt = q.cnt.as<int32_t>();
        DCHK(c2_launch_zero_u32(q.res.as<uint32_t>(), (int64_t)8 * k.capM, st));
        DCHK(c2_launch_aln_final(sc, a, k.capM, st));
    }
    DCHK(c2_dl_results(b, st));
    DCHK(hipMemcpyAsync(q.hcnt, q.cnt.p, sizeof(int32_t) * DL_NCNT, hipMemcpyDeviceToHost, st));
    return CBP_OK;
}

/* phase 1 of the batch and every query's chain: the handle's stream forks
 * to the slot streams after phase 1 and joins them back (capturable) */
int dl_sequence(cbp_handle* h, int nq, int nslots, bool queue, bool timed, const DlPlan& P)
{
    int rc;
    if (queue && (rc = c2_queue_scan_launch(h, nq, timed, false))) return rc;
    DCHK(hipEventRecord(h->gfork, h->stream));
    for (int s = 1; s < nslots; s++) DCHK(hipStreamWaitEvent(c2_slot_stream(h, s), h->gfork, 0));
    for (int i = 0; i < nq; i++) {
        /* without the work queue every query's phase 1 runs on its own slot
         * stream ahead of its chain */
        if (!queue && (rc = c2_phase1_launch(h, *h->work[i], timed, false))) return rc;
        if ((rc = dl_chain(h, i, P))) return rc;
    }
    for (int s = 1; s < nslots; s++) {
        DCHK(hipEventRecord(h->gjoin[s - 1], c2_slot_stream(h, s)));
        DCHK(hipStreamWaitEvent(h->stream, h->gjoin[s - 1], 0));
    }
    return CBP_OK;
}

int dl_build_graph(cbp_handle* h, int nq, int nslots, bool queue, const DlPlan& P, hipGraphExec_t* out)
{
    DCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    const int rc = dl_sequence(h, nq, nslots, queue, false, P);
    hipGraph_t g = nullptr;
    const hipError_t ee = hipStreamEndCapture(h->stream, &g);
    if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
    if (ee != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return c2_fail(h, CBP_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(ee));
    }
    const hipError_t e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) { *out = nullptr; return c2_fail(h, CBP_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(e)); }
    return CBP_OK;
}

}  // namespace

int c2_run_queries_dev(cbp_handle* h, const int32_t* qidx, int32_t nq, const float* scorethld,
                       const float* logevthld, cbp_results* outs)
{
    (void)hipGetLastError();   /* an earlier call's unreported status is not this call's */
    if (!h->have_opt || !h->have_tables || !h->have_queries || !h->have_chunk)
        return c2_fail(h, CBP_ERR_STATE, "options, tables, queries and chunk must be uploaded first");
    if (!h->dnnw.p) return c2_fail(h, CBP_ERR_STATE, "SS18 parameters not uploaded");
    (void)hipSetDevice(h->device);
    for (int i = 0; i < nq; i++)
        if (qidx[i] < 0 || qidx[i] >= h->queries.v.nprof) return c2_fail(h, CBP_ERR_ARG, "query index out of range");
    for (int s = 0; s < C2_NSTREAMS - 1; s++)
        if (!h->xstream[s]) DCHK(hipStreamCreate(&h->xstream[s]));
    int rc;
    if ((rc = c2_ensure_fork_events(h))) return rc;
    for (int g = 0; g < 2; g++)
        if (!h->mfork[g]) {
            DCHK(hipEventCreateWithFlags(&h->mfork[g], hipEventDisableTiming));
            DCHK(hipEventCreateWithFlags(&h->mjoin[g], hipEventDisableTiming));
        }
    while ((int)h->work.size() < nq) h->work.emplace_back(new QWork());
    while ((int)h->dlq.size() < nq) h->dlq.emplace_back(new DlQuery());
    const int nslots = nq < C2_NSTREAMS ? nq : C2_NSTREAMS;
    h->qtick++;
    std::vector<int> lq(nq), nstr(nq);
    for (int i = 0; i < nq; i++) {
        QWork& w = *h->work[i];
        w.qidx = qidx[i];
        w.scorethld = scorethld[i];
        w.logevthld = logevthld[i];
        w.slot = i % C2_NSTREAMS;
        w.st = c2_slot_stream(h, w.slot);
        w.dq = c2_get_query(h, qidx[i], &rc);
        if (!w.dq) return rc;
        lq[i] = w.dq->hp.Lq;
        nstr[i] = w.dq->hp.nstripes;
    }
    DCHK(hipStreamSynchronize(h->stream));
    const bool queue = h->use_queue && nq >= 2;
    /* phase-1 buffers; a grown batch buffer retires the graphs */
    if (queue) {
        const void* old[7] = {h->qkeys.p, h->qwts.p, h->hqkeys, h->qscan.dargs.p, h->qscan.dtiles.p,
                              h->qscan.dhead.p, h->qscan.dbot.p};
        if ((rc = c2_queue_scan_prepare(h, nq))) return rc;
        const void* now[7] = {h->qkeys.p, h->qwts.p, h->hqkeys, h->qscan.dargs.p, h->qscan.dtiles.p,
                              h->qscan.dhead.p, h->qscan.dbot.p};
        if (std::memcmp(old, now, sizeof old) != 0) c2_drop_graphs(h);
    } else {
        for (int i = 0; i < nq; i++) if ((rc = c2_phase1_prepare(h, *h->work[i]))) return rc;
    }
    /* capacities, buffers and the per-query constants */
    DlPlan P;
    P.k = dl_caps(h, lq, nstr);
    P.split = (nq == 1);
    P.sp.resize(nq); P.grid.resize(nq); P.lds.resize(nq);
    const cbp_profiles& ch = h->chunk.v;
    const uint64_t dblen = h->dblen ? h->dblen : (uint64_t)ch.npos;
    const uint64_t nseqs = h->nseqs ? h->nseqs : (uint64_t)ch.nprof;
    int64_t botw = 0;
    std::vector<void*> bufs_before;
    for (int i = 0; i < nq; i++) {
        QWork& w = *h->work[i];
        const QueryPrep& hp = w.dq->hp;
        const size_t lds = c2_phase1_lds_bytes(hp.R, 2, hp.nstripes);
        if (lds > 80 * 1024) return c2_fail(h, CBP_ERR_STATE, "backtrack pass exceeds the LDS");
        P.lds[i] = lds;
        P.grid[i] = c2_phase1q_grid(hp.R, 2, true, lds);
        if (P.grid[i] < 1) return c2_fail(h, CBP_ERR_HIP, "backtrack queue occupancy");
        if (hp.nstripes > 1) botw = std::max<int64_t>(botw, (int64_t)P.grid[i] * 2 * 5 * h->chunk_maxlen);
        uint64_t delta = 0;
        w.sspace = c2_length_adjustment(h->opt, (uint64_t)hp.Lq, dblen, nseqs, &delta);
        StatParams& sp = P.sp[i];
        sp.Lq = hp.Lq;
        sp.qeno = h->queries.v.eno[w.qidx];
        sp.searchspace = (float)w.sspace;
        sp.ssemodel = h->opt.ssemodel;
        sp.ung_lambda = h->opt.ung_lambda; sp.ung_K = h->opt.ung_K;
        sp.gap_lambda = h->opt.gap_lambda; sp.gap_K = h->opt.gap_K;
    }
    /* a buffer that grows moves: graphs that hold the old addresses retire */
    auto bufs = [&]() {
        std::vector<DBuf*> v;
        for (int i = 0; i < nq; i++) {
            DlQuery& q = *h->dlq[i];
            for (DBuf* b : {&q.cnt, &q.aln, &q.epa, &q.order, &q.sel, &q.res, &q.p1save, &q.malns, &q.dargs, &q.tiles, &q.ntiles})
                v.push_back(b);
        }
        for (int s = 0; s < nslots; s++) {
            DlSlot& sl = h->dls[s];
            for (DBuf* b : {&sl.tilecnt, &sl.tileoff, &sl.cand, &sl.pe, &sl.off, &sl.btck, &sl.botw, &sl.fin,
                            &sl.pass_pro, &sl.passfin, &sl.statin, &sl.counts, &sl.statout, &sl.mpairs, &sl.morder,
                            &sl.sitems, &sl.apro, &sl.acand, &sl.mF, &sl.mB, &sl.mSd, &sl.mbt, &sl.mhand,
                            &sl.mstat, &sl.malntmp})
                v.push_back(b);
        }
        return v;
    };
    auto addrs = [&]() {
        std::vector<void*> v;
        for (DBuf* b : bufs()) v.push_back(b->p);
        return v;
    };
    const std::vector<void*> a0 = addrs();
    rc = CBP_OK;
    for (int i = 0; i < nq && !rc; i++) rc = dl_alloc_query(h, *h->dlq[i], P.k);
    for (int s = 0; s < nslots && !rc; s++) rc = dl_alloc_slot(h, h->dls[s], P.k, botw, P.split);
    if (rc == CBP_ERR_MEM) {
        /* the lists' fixed pools do not fit the memory budget: give them back
         * and run the host-decided sequence, which sizes its buffers to the
         * call */
        c2_drop_graphs(h);
        for (DBuf* b : bufs()) b->release();
        h->last_dl_fallbacks = nq;
        return c2_run_queries(h, qidx, nq, scorethld, logevthld, outs);
    }
    if (rc) return rc;
    if (addrs() != a0) c2_drop_graphs(h);
    /* the backtrack pass's arguments (the device writes their segment count) */
    for (int i = 0; i < nq; i++) {
        QWork& w = *h->work[i];
        DlSlot& sl = h->dls[w.slot];
        P1Args a;
        if ((rc = c2_fill_p1args(h, w.dq, &a))) return rc;
        a.keys = w.dk; a.wts = w.dw;
        a.seg_pb = sl.cand.as<int32_t>(); a.seg_pe = sl.pe.as<int32_t>(); a.nseg = 0;
        a.btck = sl.btck.as<uint32_t>(); a.btck_off = sl.off.as<int64_t>();
        if ((rc = c2_dupload(h, h->dlq[i]->dargs, &a, 1, h->stream))) return rc;
    }
    DCHK(hipStreamSynchronize(h->stream));
    /* the whole sequence: eager, or replayed from a graph of the batch */
    bool graphed = false;
    /* one-query calls run eagerly: their graph (non-queue phase 1 and the
     * MAP groups on three extra streams) is switched off */
    if (h->use_graph && nq >= 2) {
        const int32_t glo = (int32_t)(h->chunk_gen & 0x7fffffff), ghi = (int32_t)(h->chunk_gen >> 31);
        if (!h->graphs.empty()) {
            const std::vector<int32_t>& k0 = h->graphs.begin()->first;
            if (k0[0] != glo || k0[1] != ghi) c2_drop_graphs(h);
        }
        std::vector<int32_t> key = {glo, ghi, 2 + (queue ? 1 : 0), nq};
        key.insert(key.end(), qidx, qidx + nq);
        if (!h->graphs.count(key) && h->graphs.size() >= 8) c2_drop_graphs(h);
        cbp_handle::GraphEnt& ge = h->graphs[key];
        if (!ge.exec && ge.seen >= 1 && (rc = dl_build_graph(h, nq, nslots, queue, P, &ge.exec))) return rc;
        if (ge.exec) {
            if (!h->gt0) { DCHK(hipEventCreate(&h->gt0)); DCHK(hipEventCreate(&h->gt1)); }
            DCHK(hipEventRecord(h->gt0, h->stream));
            DCHK(hipGraphLaunch(ge.exec, h->stream));
            DCHK(hipEventRecord(h->gt1, h->stream));
            graphed = true;
        } else {
            ge.seen++;
        }
    }
    if (!graphed && (rc = dl_sequence(h, nq, nslots, queue, true, P))) return rc;
    DCHK(hipStreamSynchronize(h->stream));
    /* counters: a list beyond its capacity sends the call to the host-decided
     * sequence */
    for (int i = 0; i < nq; i++)
        if (h->dlq[i]->hcnt[DL_OVF] != 0) {
            h->last_dl_fallbacks = nq;
            return c2_run_queries(h, qidx, nq, scorethld, logevthld, outs);
        }
    h->last_dl_fallbacks = 0;
    /* results: exactly the lists' sizes */
    h->dlres.resize(nq);
    for (int i = 0; i < nq; i++) {
        QWork& w = *h->work[i];
        DlQuery& q = *h->dlq[i];
        DlResult& r = h->dlres[i];
        const int32_t* c = q.hcnt;
        const int npass = c[DL_NPASS], nsel = c[DL_NSEL];
        int64_t nposits;
        std::memcpy(&nposits, c + DL_NPOSITS, sizeof nposits);
        r.aln.resize((size_t)npass * CBP_NFIELDS); r.epa.resize(npass); r.order.resize(npass);
        r.map.resize(nsel); r.p1save.resize((size_t)3 * nsel);
        if (npass) {
            DCHK(hipMemcpyAsync(r.aln.data(), q.aln.p, sizeof(float) * r.aln.size(), hipMemcpyDeviceToHost, w.st));
            DCHK(hipMemcpyAsync(r.epa.data(), q.epa.p, sizeof(float) * npass, hipMemcpyDeviceToHost, w.st));
            DCHK(hipMemcpyAsync(r.order.data(), q.order.p, sizeof(int32_t) * npass, hipMemcpyDeviceToHost, w.st));
        }
        const int64_t dstride = P.k.stride;
        w.alnstride = nsel ? (int64_t)w.dq->hp.Lq + c[DL_MAXLD] + 1 : 0;
        if (nsel) {
            DCHK(hipMemcpyAsync(r.map.data(), q.sel.p, sizeof(int32_t) * nsel, hipMemcpyDeviceToHost, w.st));
            DCHK(hipMemcpyAsync(r.p1save.data(), q.p1save.p, sizeof(float) * 3 * nsel, hipMemcpyDeviceToHost, w.st));
            const size_t nstr = (size_t)nsel * 3 * dstride;
            if (w.strs_cap < nstr) {
                if (w.strs) { (void)hipHostFree(w.strs); w.strs = nullptr; w.strs_cap = 0; }
                DCHK(hipHostMalloc((void**)&w.strs, nstr, hipHostMallocDefault));
                w.strs_cap = nstr;
            }
            DCHK(hipMemcpyAsync(w.strs, q.malns.p, nstr, hipMemcpyDeviceToHost, w.st));
        }
        w.nposits = nposits;
        w.maxlen = c[DL_MAXLEN];
    }
    for (int s = 0; s < nslots; s++) DCHK(hipStreamSynchronize(c2_slot_stream(h, s)));
    for (int i = 0; i < nq; i++) {
        QWork& w = *h->work[i];
        DlResult& r = h->dlres[i];
        const int npass = (int)r.order.size(), nsel = (int)r.map.size();
        /* strings to the stride of the drawn pairs (rows move down only) */
        const int64_t dstride = P.k.stride;
        for (int64_t row = 0; row < (int64_t)nsel * 3 && w.alnstride < dstride; row++)
            std::memmove(w.strs + row * w.alnstride, w.strs + row * dstride, (size_t)w.alnstride);
        w.p1save = r.p1save;
        cbp_results* out = &outs[i];
        std::memset(out, 0, sizeof *out);
        out->npassed = npass;
        out->nposits = w.nposits;
        out->maxlen = w.maxlen;
        out->nmap = nsel;
        out->alndata = npass ? r.aln.data() : nullptr;
        out->order = npass ? r.order.data() : nullptr;
        out->alns = nsel ? w.strs : nullptr;
        out->alnstride = w.alnstride;
        out->map_rows = nsel ? r.map.data() : nullptr;
        out->epa = npass ? r.epa.data() : nullptr;
        out->scorethld = w.scorethld;
        out->logevthld = w.logevthld;
        out->search_space = (double)w.sspace;
    }
    /* device time of phase 1 (the whole graph when replayed) */
    double ms = 0.0;
    if (graphed || queue) {
        float t = 0.0f;
        if (graphed) DCHK(hipEventElapsedTime(&t, h->gt0, h->gt1));
        else DCHK(hipEventElapsedTime(&t, h->ev0, h->ev1));
        ms = t;
    } else {
        for (int i = 0; i < nq; i++) {
            float t = 0.0f;
            DCHK(hipEventElapsedTime(&t, h->work[i]->e0, h->work[i]->e1));
            ms += t;
        }
    }
    h->last_ms = ms;
    h->last_nlaunch = (graphed || queue) ? 1 : nq;
    return CBP_OK;
}
