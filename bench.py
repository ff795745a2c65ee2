#!/usr/bin/env python3
"""Search throughput of the COMER2 libmycu pipeline on one MI355X.

Workload (plain method at scale): one synthetic query profile of length 300
searched against 200,000 synthetic DB profiles (lengths log-normal, median
200, sigma 0.668, clipped to 30-3000, about 50M positions, sorted longest
first), with homologous segments of the query planted into 4,000 DB profiles
so that the statistics and MAP stages have real hits; MAP realignment is
capped to the best 2,000 pairs by (log E, profile index).  One step is one
full pipeline pass for the query: fused score + phase-1 Viterbi over every DB
cell, backtrack recomputation + finalize for the candidates, FP64 statistics
and E-values, MAP forward/backward/DP/backtrace for the selected pairs, and
the results back on the host.  The DB and the query preparation are resident
in HBM before the timed region; every step ends with the results copied to
the host, so the step boundary is a full device synchronisation.

value = search GCUPS = steps * Lq * DB positions / timed seconds.

The roofline object describes the dominant kernel (fused score + Viterbi):
algorithmic FP32 operations per cell (87, counted from the reference's
recurrences) times the cells of one launch, over the launch's average
duration measured with HIP events on the library's stream; peak is MI355X's
published FP32 rate.  traffic is read from the rocprofv3 counter summary
committed under profiles/ when present.

The CPU baseline runs the oracle's whole-pipeline restatement on a strided
sample of the same DB (OpenMP over profiles) for a bounded time.

The batch object reports 64 queries (lengths 50-1000, homologs planted)
against the same DB: one by one (plain), one batched call with one launch
per query, one batched call whose phase 1 and backtrack pass run from the
device work queue (persistent workgroups, tiles binned by rows per lane),
and the same with phase 1 replayed from a HIP graph; hit lists compared bit
for bit.

The stream object reports the same searches with the DB in host memory,
uploaded in chunks of about 12.6 M positions (4 chunks) through pinned
staging buffers while the previous chunk is searched (cbp_search_db), the
upload included in the time: the one-query search and the 64-query batch,
hit lists compared with the resident-DB runs.

The softdp object reports the differentiable soft DP: forward-only and
forward+adjoint GCUPS on 4,096 pairs (lengths 50-600) and on 256 pairs
(lengths 2,000-5,000, sqrt(L) checkpointing vs all cells kept).

  python bench.py [--gpus 1] [--steps 10] [--warmup 2] [--cpu-seconds 15]
  python bench.py --copy-calib BYTES     one plain device copy (counter scale)
"""
import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FLOP_PER_CELL = 87          # 41 score + 10 SS term + 36 Viterbi (add/max/compare)
PEAK_FP32_TFLOPS = 157.3    # MI355X published FP32 peak
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "phase1_search_c1_counters.md")

DB_SEED, QUERY_SEED, PLANT_SEED = 2, 1002, 3002
N_PROFILES, MEDIAN, SIGMA, MINLEN, MAXLEN = 200_000, 200, 0.668, 30, 3000
QUERY_LEN, N_PLANT, PLANT_KMIN, PLANT_KMAX = 300, 4000, 3.0, 60.0
MAP_CAP = 2000
STREAM_CHUNK_POS = 12_600_000     # about 4 chunks of the C1 DB


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def copy_calib(nbytes):
    import torch
    x = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calib_bytes": nbytes}))


def read_traffic():
    """HBM bytes per launch of the phase-1 kernel from the counter summary."""
    try:
        with open(TRAFFIC_FILE) as f:
            txt = f.read()
    except OSError:
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", txt)
    return float(m.group(1)) if m else None


def _cpu_sample(q, d, orc, pros, thld, logev, sspace, seconds, nthreads):
    """the oracle's whole-pipeline restatement over the listed DB profiles
    (OpenMP over profiles) until done or out of time"""
    cells = mcells = npass = nmap = done = 0
    t0 = time.perf_counter()
    while done < pros.size and time.perf_counter() - t0 < seconds:
        batch = pros[done:done + 1024]
        r = orc.search(q.view, 0, d.view, batch, thld, logev, float(sspace), nthreads)
        cells += r.cells; mcells += r.map_cells; npass += r.npassed; nmap += r.nmap
        done += batch.size
    dt = time.perf_counter() - t0
    return dict(gcups=cells / dt / 1e9, done=done, cells=cells, passed=npass, realigned=nmap, seconds=dt)


def cpu_baseline(q, d, orc, thld, logev, sspace, seconds, nthreads):
    """C1 on the host: the oracle's FP32 restatement of the whole pipeline on
    a strided sample of the DB, on nthreads and on one thread; and C0 (one
    query of 150 positions vs 2,000 profiles, seed 1, MAP of the best 100)
    whole, on nthreads and on one thread"""
    import numpy as np
    from comer2_amd._ffi import SynthDB
    order = np.arange(d.nprof, dtype=np.int32)
    stride = 7
    sample = np.concatenate([order[k::stride] for k in range(stride)])
    allc = _cpu_sample(q, d, orc, sample, thld, logev, sspace, seconds, nthreads)
    one = _cpu_sample(q, d, orc, sample, thld, logev, sspace, max(2.0, seconds / 4), 1)
    # C0: seed 1, query 150, 2,000 profiles (median 180, clip 30-800), MAP top 100
    q0 = SynthDB(1001, 1, 150, 0.0, 150, 150)
    d0 = SynthDB(1, 2000, 180, SIGMA, 30, 800)
    d0.plant(q0, 3001, 40, PLANT_KMIN, PLANT_KMAX)
    t0, l0 = orc.score_threshold(150, d0.npos)
    ss0, _ = orc.length_adjustment(150, d0.npos, d0.nprof)
    every = np.arange(d0.nprof, dtype=np.int32)
    c0n = _cpu_sample(q0, d0, orc, every, t0, l0, ss0, 60.0, nthreads)
    c01 = _cpu_sample(q0, d0, orc, every, t0, l0, ss0, 60.0, 1)
    return dict(value=allc["gcups"], unit="GCUPS", cores=nthreads, kind="port",
                sample="C1: %d of %d DB profiles (stride-%d interleave), %d cells, %d passed, %d realigned, "
                       "%.1f s; oracle FP32 restatement of the whole pipeline, OpenMP over profiles"
                       % (allc["done"], d.nprof, stride, allc["cells"], allc["passed"], allc["realigned"],
                          allc["seconds"]),
                nproc=os.cpu_count(), affinity=len(os.sched_getaffinity(0)),
                omp_num_threads=os.environ.get("OMP_NUM_THREADS"),
                one_core={"value": one["gcups"], "sample": "C1: %d profiles, %d cells, %.1f s"
                          % (one["done"], one["cells"], one["seconds"])},
                c0={"profiles": d0.nprof, "positions": int(d0.npos), "query_len": 150, "seed": 1,
                    "planted": 40, "passed": c0n["passed"], "realigned": c0n["realigned"],
                    "gcups": c0n["gcups"], "seconds": round(c0n["seconds"], 3),
                    "one_core_gcups": c01["gcups"], "one_core_seconds": round(c01["seconds"], 3)})


def batch_bench(d, args):
    """64 queries with lengths uniform in 50-1000 against the same DB: one by
    one (plain per-query pipeline) vs one batched call; hit lists compared."""
    import numpy as np
    from comer2_amd._ffi import SynthDB
    from comer2_amd.host import CuBatchProcessing
    nq = 64
    q = SynthDB(3, nq, 0, 0, 50, 1000, sort_desc=False)
    cbp = CuBatchProcessing(0, 0, 0, options=dict(max_map_pairs=100))
    cbp.TransferQueries(q)
    cbp.TransferCPMData(d)
    cbp.SetDbDetails(d.npos, d.nprof)
    qs = list(range(nq))
    lens = q.lengths()
    cells = int(lens.sum()) * d.npos
    # launch geometry from the autotuner, timed on the median-length query
    mid = int(np.argsort(lens, kind="stable")[nq // 2])
    t = time.perf_counter()
    seg_cols, nw, rmax = cbp.autotune(mid)
    ttune = time.perf_counter() - t
    log("bench: autotuned on query %d (L=%d): %d columns per segment, %d waves per workgroup, "
        "at most %d rows per lane (%.1f s)" % (mid, int(lens[mid]), seg_cols, nw, rmax, ttune))
    # warm-up of every leg's state before its timing: the one-by-one leg
    # prepares all 64 queries and grows the one-query buffers here
    for i in qs:
        cbp.ProcessScoreMatrix(i)
    t = time.perf_counter()
    plain = [cbp.ProcessScoreMatrix(i) for i in qs]
    tp = time.perf_counter() - t
    # one call, one launch per query on the query streams
    cbp.set_queue(False)
    cbp.ProcessBatch(qs)
    t = time.perf_counter()
    launched = cbp.ProcessBatch(qs)
    tl = time.perf_counter() - t
    # one call, phase 1 and the backtrack pass from the device work queue
    cbp.set_queue(True)
    cbp.ProcessBatch(qs)
    t = time.perf_counter()
    batched = cbp.ProcessBatch(qs)
    tb = time.perf_counter() - t
    qms, _ = cbp.last_kernel_ms()
    # the work-queue phase 1 replayed from a HIP graph (captured on the second call)
    cbp.set_graph(True)
    cbp.ProcessBatch(qs)
    cbp.ProcessBatch(qs)
    t = time.perf_counter()
    graphed = cbp.ProcessBatch(qs)
    tg = time.perf_counter() - t
    cbp.set_graph(False)

    def same_hits(xs, ys):
        return all(a["npassed"] == b["npassed"] and a["nmap"] == b["nmap"]
                   and np.array_equal(a["alndata"].view(np.uint32), b["alndata"].view(np.uint32))
                   and np.array_equal(a["order"], b["order"]) and np.array_equal(a["aln_raw"], b["aln_raw"])
                   for a, b in zip(xs, ys))
    same = same_hits(plain, launched) and same_hits(plain, batched) and same_hits(plain, graphed)
    out = {"queries": nq, "query_lengths": [50, 1000], "seed": 3, "cells": cells,
           "plain_s": round(tp, 3), "launched_s": round(tl, 3), "batched_s": round(tb, 3), "graph_s": round(tg, 3),
           "plain_gcups": round(cells / tp / 1e9, 3), "launched_gcups": round(cells / tl / 1e9, 3),
           "batched_gcups": round(cells / tb / 1e9, 3), "graph_gcups": round(cells / tg / 1e9, 3),
           "queue_phase1_ms": round(qms, 3), "queue_phase1_gcups": round(cells / (qms * 1e-3) / 1e9, 3),
           "ratio": round(tp / tb, 4), "ratio_graph": round(tp / tg, 4), "identical_hits": bool(same),
           "passed": int(sum(r["npassed"] for r in batched)), "realigned": int(sum(r["nmap"] for r in batched)),
           "map_cap_per_query": 100,
           "tuning": {"query": mid, "query_len": int(lens[mid]), "seg_cols": seg_cols,
                      "waves_per_workgroup": nw, "max_rows_per_lane": rmax}}
    log("bench: 64 queries plain %.2f s, per-query launches %.2f s, work queue %.2f s (phase 1 %.1f ms), "
        "queue + graph %.2f s, identical %s" % (tp, tl, tb, qms, tg, same))
    cbp.close()
    return out


def stream_bench(d, q, res, nq_batch=64):
    """The whole DB from host memory, uploaded in chunks while the previous
    chunk is searched (cbp_search_db): the one-query search (C1) and the
    64-query batch (C2), each timed with its upload, against the
    resident-DB results (hit lists compared)."""
    import numpy as np
    from comer2_amd._ffi import SynthDB
    from comer2_amd.host import CuBatchProcessing
    out = {}
    cbp = CuBatchProcessing(0, 0, 0, options=dict(max_map_pairs=MAP_CAP))
    cbp.TransferQueries(q)
    chunk = STREAM_CHUNK_POS
    cbp.SearchDB(d, [0], chunk_pos=chunk)            # warm-up: buffers, query preparation
    t = time.perf_counter()
    r = cbp.SearchDB(d, [0], chunk_pos=chunk)[0]
    ts = time.perf_counter() - t
    nch = cbp.last_search_chunks()
    same = (r["npassed"] == res["npassed"] and np.array_equal(r["alndata"].view(np.uint32), res["alndata"].view(np.uint32))
            and np.array_equal(r["order"], res["order"]) and np.array_equal(r["aln_raw"], res["aln_raw"]))
    cells = QUERY_LEN * d.npos
    out["one_query"] = {"chunk_positions": chunk, "chunks": nch, "seconds": round(ts, 4),
                        "gcups_with_upload": round(cells / ts / 1e9, 3), "identical_hits": bool(same)}
    log("bench: streamed one-query search: %d chunks, %.3f s with the upload (%.1f GCUPS), identical %s"
        % (nch, ts, cells / ts / 1e9, same))
    cbp.close()
    qb = SynthDB(3, nq_batch, 0, 0, 50, 1000, sort_desc=False)
    cbp = CuBatchProcessing(0, 0, 0, options=dict(max_map_pairs=100))
    cbp.TransferQueries(qb)
    qs = list(range(nq_batch))
    cbp.SearchDB(d, qs, chunk_pos=chunk)
    t = time.perf_counter()
    rs = cbp.SearchDB(d, qs, chunk_pos=chunk)
    tb = time.perf_counter() - t
    nchb = cbp.last_search_chunks()
    cellsb = int(qb.lengths().sum()) * d.npos
    cbp.TransferCPMData(d)
    cbp.SetDbDetails(d.npos, d.nprof)
    rr = cbp.ProcessBatch(qs)
    sameb = all(a["npassed"] == b["npassed"] and np.array_equal(a["alndata"].view(np.uint32), b["alndata"].view(np.uint32))
                and np.array_equal(a["order"], b["order"]) and np.array_equal(a["aln_raw"], b["aln_raw"])
                for a, b in zip(rs, rr))
    out["batch64"] = {"chunk_positions": chunk, "chunks": nchb,
                      "seconds": round(tb, 4), "gcups_with_upload": round(cellsb / tb / 1e9, 3),
                      "identical_hits": bool(sameb)}
    log("bench: streamed 64-query search: %.3f s with the upload (%.1f GCUPS), identical %s"
        % (tb, cellsb / tb / 1e9, sameb))
    cbp.close()
    return out


def softdp_bench(reps):
    """Soft DP (log-sum-exp forward and hand-written adjoint) at the two
    differentiation workloads: 4,096 pairs with lengths uniform in 50-600
    (all forward cells kept) and 256 pairs with lengths uniform in 2,000-5,000
    (sqrt(L) checkpointing, and all cells kept for comparison).  Scores are
    the pairs' profile-profile score matrices; theta = 0."""
    import torch
    from comer2_amd._ffi import SynthDB
    from comer2_amd.host import CuBatchProcessing
    from comer2_amd.softdp import SoftDP
    out = {}
    for name, seed, n, lo, hi, modes in (("pairs_50_600", 4, 4096, 50, 600, (0,)),
                                         ("pairs_2000_5000", 5, 256, 2000, 5000, (-1, 0))):
        q = SynthDB(seed, n, 0, 0, lo, hi, sort_desc=False)
        d = SynthDB(seed + 1000, n, 0, 0, lo, hi, sort_desc=False)
        cbp = CuBatchProcessing(0)
        cbp.TransferQueries(q)
        cbp.TransferCPMData(d)
        pairs = [(k, k) for k in range(n)]
        rec = {"pairs": n, "lengths": [lo, hi], "seeds": [seed, seed + 1000]}
        for ck in modes:
            torch.cuda.reset_peak_memory_stats()
            sdp = SoftDP(cbp, pairs, ckpt=ck)
            S = sdp.scores()
            th = torch.zeros(5, dtype=torch.float32, device="cuda")
            logZ = sdp.forward(S, th)
            dS, dth, _ = sdp.adjoint(S, th, logZ)
            torch.cuda.synchronize()
            t = time.perf_counter()
            for _ in range(reps):
                logZ = sdp.forward(S, th)
            torch.cuda.synchronize()
            tf = (time.perf_counter() - t) / reps
            t = time.perf_counter()
            for _ in range(reps):
                logZ = sdp.forward(S, th)
                dS, dth, _ = sdp.adjoint(S, th, logZ)
            torch.cuda.synchronize()
            tb = (time.perf_counter() - t) / reps
            key = "all_cells" if ck == 0 else "checkpointed"
            rec[key] = {"cells": sdp.total, "ckpt_steps": "auto" if ck < 0 else 0,
                        "fwd_ms": round(tf * 1e3, 3), "fwd_adj_ms": round(tb * 1e3, 3),
                        "fwd_gcups": round(sdp.total / tf / 1e9, 3),
                        "fwd_adj_gcups": round(sdp.total / tb / 1e9, 3),
                        "ratio": round(tf / tb, 4),
                        "workspace_bytes": int(sdp.ws_bytes),
                        "peak_torch_bytes": int(torch.cuda.max_memory_allocated())}
            log("bench: soft DP %s %s: fwd %.2f ms, fwd+adj %.2f ms" % (name, key, tf * 1e3, tb * 1e3))
            del sdp, S, dS, dth, logZ
            torch.cuda.empty_cache()
        if "checkpointed" in rec and "all_cells" in rec:
            rec["recompute_overhead"] = round(rec["checkpointed"]["fwd_adj_ms"] / rec["all_cells"]["fwd_adj_ms"] - 1.0, 4)
        out[name] = rec
        cbp.close()
    return out


def phase1_variants(cbp, q, d):
    """phase-1 launches of the C1 query (and of a 256-row query: one stripe
    at 4 rows per lane, where neither sweep spills) with the 20-term score
    sums on the vector ALUs and on the matrix cores; keys must be identical"""
    import numpy as np
    from comer2_amd._ffi import SynthDB
    out = {"metric": "phase-1 sweep time, vector vs matrix-core score sums", "unit": "ms", "legs": []}
    for qs in (q, SynthDB(QUERY_SEED + 1, 1, 256, 0.0, 256, 256)):
        cbp.TransferQueries(qs)
        L = int(qs.lengths()[0])
        keys = {}
        for on in (False, True):
            cbp.set_mfma(on)
            keys[on] = cbp.phase1_keys(0)
            ms = []
            for _ in range(3):
                cbp.phase1_keys(0)
                m, n = cbp.last_kernel_ms()
                ms.append(m / max(n, 1))
            leg = dict(query_len=L, matrix_cores=on, kernel_ms=round(min(ms), 3),
                       gcups=round(L * d.npos / (min(ms) * 1e-3) / 1e9, 2))
            out["legs"].append(leg)
            log("bench: phase-1 L=%d %s: %.2f ms" % (L, "matrix cores" if on else "vector ALUs", min(ms)))
        out["legs"][-1]["keys_identical"] = bool(np.array_equal(keys[False], keys[True]))
    cbp.set_mfma(False)
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--nprof", type=int, default=N_PROFILES)
    ap.add_argument("--copy-calib", type=int, default=0)
    ap.add_argument("--no-softdp", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--softdp-reps", type=int, default=3)
    ap.add_argument("--no-stream", action="store_true")
    ap.add_argument("--map-rows", type=int, default=-1,
                    help="rows per lane of the shorter pairs' MAP group (0-5; default: the library's)")
    ap.add_argument("--map-pipe", type=int, default=-1,
                    help="MAP passes streaming pairs through one band per workgroup (1) or one workgroup "
                         "per pair (0); default: the library's")
    ap.add_argument("--phase1-variants", action="store_true",
                    help="only time the phase-1 sweep of the C1 query with its score sums on the vector "
                         "ALUs and on the matrix cores (and a 256-row query, one stripe at 4 rows per lane), "
                         "keys compared; one JSON line")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; the search shards over DB profiles "
              "across GPUs by design but no multi-GPU path is built (--gpus %d refused)" % args.gpus)
        sys.exit(2)
    if args.copy_calib:
        copy_calib(args.copy_calib)
        return

    from comer2_amd._ffi import SynthDB
    from comer2_amd.host import CuBatchProcessing

    t = time.perf_counter()
    d = SynthDB(DB_SEED, args.nprof, MEDIAN, SIGMA, MINLEN, MAXLEN)
    q = SynthDB(QUERY_SEED, 1, QUERY_LEN, 0.0, QUERY_LEN, QUERY_LEN)
    d.plant(q, PLANT_SEED, N_PLANT if args.nprof == N_PROFILES else max(1, args.nprof // 50),
            PLANT_KMIN, PLANT_KMAX)
    # homologs of the 64 batch queries are planted whether or not the batch
    # leg runs, so the one-query workload is the same DB either way
    qb = SynthDB(3, 64, 0, 0, 50, 1000, sort_desc=False)
    d.plant(qb, PLANT_SEED + 1, 20 * 64 if args.nprof == N_PROFILES else max(64, args.nprof // 100),
            PLANT_KMIN, PLANT_KMAX)
    log("bench: generated %d profiles, %d positions in %.1f s" % (d.nprof, d.npos, time.perf_counter() - t))

    cbp = CuBatchProcessing(0, 0, 0, options=dict(max_map_pairs=MAP_CAP))
    if args.map_rows >= 0:
        cbp.set_map_rows(args.map_rows)
    if args.map_pipe >= 0:
        cbp.set_map_pipe(args.map_pipe != 0)
    t = time.perf_counter()
    cbp.TransferQueries(q)
    cbp.TransferCPMData(d)
    cbp.SetDbDetails(d.npos, d.nprof)
    thld, logev = cbp.score_threshold(QUERY_LEN, d.npos)
    log("bench: uploaded in %.1f s; score threshold %.3f, log E threshold %.3f"
        % (time.perf_counter() - t, thld, logev))

    if args.phase1_variants:
        phase1_variants(cbp, q, d)
        return

    for i in range(args.warmup):
        res = cbp.ProcessScoreMatrix(0, scorethld=thld, logevthld=logev)
        log("bench: warmup %d done (%d passed, %d realigned)" % (i, res["npassed"], res["nmap"]))

    kms = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        res = cbp.ProcessScoreMatrix(0, scorethld=thld, logevthld=logev)
        ms, nl = cbp.last_kernel_ms()
        kms.append(ms / max(nl, 1))
    elapsed = time.perf_counter() - t0
    log("bench: %d steps in %.3f s" % (args.steps, elapsed))

    cells = QUERY_LEN * d.npos
    gcups = args.steps * cells / elapsed / 1e9
    kavg = sum(kms) / len(kms)
    achieved = cells * FLOP_PER_CELL / (kavg * 1e-3) / 1e12
    alnlen = res["alnlen"]
    dlen = d.lengths()
    map_cells = int(sum(QUERY_LEN * int(dlen[res["pronr"][r]]) for r in res["map_rows"]))

    out = {
        "metric": "search GCUPS at 1 GPU (plain per-query pipeline: score, Viterbi, statistics, MAP)",
        "value": round(gcups, 3),
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generator; homologs planted)",
        "config": {
            "workload": "1 query (L=300) vs %d synthetic DB profiles (log-normal median %d, "
                        "sigma %.3f, clip %d-%d), full pipeline, MAP capped at %d pairs"
                        % (d.nprof, MEDIAN, SIGMA, MINLEN, MAXLEN, MAP_CAP),
            "db_profiles": d.nprof, "db_positions": int(d.npos), "query_len": QUERY_LEN,
            "seeds": {"db": DB_SEED, "query": QUERY_SEED, "plant": PLANT_SEED},
            "planted": N_PLANT, "score_threshold": round(thld, 4), "log_e_threshold": round(logev, 4),
            "passed": int(res["npassed"]), "realigned": int(res["nmap"]),
            "aligned": int((alnlen[res["map_rows"]] > 0).sum()) if res["nmap"] else 0,
            "map_cells": map_cells, "parallelism": "single GPU",
        },
        "roofline": {
            "kernel": "k_phase1 (fused score + Viterbi)",
            "bound": "valu",
            "achieved": round(achieved, 3),
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
            "traffic": read_traffic(),
            "flop_per_cell": FLOP_PER_CELL,
            "cells_per_launch": cells,
            "avg_launch_ms": round(kavg, 4),
        },
        "cpu_baseline": None,
        "softdp": None,
        "batch": None,
        "stream": None,
    }
    if not args.no_batch:
        out["batch"] = batch_bench(d, args)
    if not args.no_stream:
        out["stream"] = stream_bench(d, q, res)
    if not args.no_softdp:
        out["softdp"] = softdp_bench(args.softdp_reps)
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        from oracle_ffi import Oracle
        from comer2_amd import tables
        orc = Oracle(tables.load_ssss(), tables.load_nn_weights(),
                     dict(tables.default_options(), max_map_pairs=MAP_CAP))
        # every core this process may use, or the share OMP_NUM_THREADS grants
        nthreads = max(1, int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or len(os.sched_getaffinity(0)))
        out["cpu_baseline"] = cpu_baseline(q, d, orc, thld, logev, res["search_space"],
                                           args.cpu_seconds, nthreads)
        cb = out["cpu_baseline"]
        log("bench: cpu baseline C1 %.4f GCUPS on %d threads, %.4f on one; C0 %.4f / %.4f GCUPS (%d passed)"
            % (cb["value"], nthreads, cb["one_core"]["value"], cb["c0"]["gcups"], cb["c0"]["one_core_gcups"],
               cb["c0"]["passed"]))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
