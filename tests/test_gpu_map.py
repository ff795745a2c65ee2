"""MAP realignment on the GPU vs the oracle's direct MAP restatement.

For every realigned pair the alignment length, identities, positives, gaps,
beginning/end coordinates and the three alignment strings must match the FP32
oracle exactly, fed with the same log pair E-value the GPU statistics produced
(the dynamic score parameters depend on it).  The oracle's direct MAP is
itself pinned cell-by-cell against a literal emulation of the reference's
oblique-block schedule.
"""
import numpy as np
import pytest
from comer2_amd._ffi import SynthDB
from gpu_common import oracle, processing

pytestmark = pytest.mark.gpu


def _strings(aln):
    return tuple(bytes(aln[j]).split(b"\0")[0].decode() for j in range(3))


def _planted_sets(seed, qlen=None):
    if qlen is None:
        q = SynthDB(seed, 2, 150, 0.668, 60, 400)
    else:
        q = SynthDB(seed, 1, 0, 0, qlen[0], qlen[1])
    d = SynthDB(seed + 1, 90, 160, 0.668, 40, 600)
    d.plant(q, seed + 2, 14, 3.0, 60.0)
    return q, d


# query lengths 60-400: one band of 1-7 stripes of 64 rows; 560-640: stripes
# of 128 rows; 2600-2700: 320-row stripes in two bands (HBM band hand-off)
@pytest.mark.parametrize("seed,cap,qlen", [(41, 0, None), (51, 5, None), (61, 6, (560, 640)),
                                           (71, 4, (2600, 2700))])
def test_map_matches_oracle(seed, cap, qlen):
    q, d = _planted_sets(seed, qlen)
    orc = oracle()
    cbp = processing(dict(max_map_pairs=cap))
    cbp.TransferQueries(q)
    cbp.TransferCPMData(d)
    cbp.SetDbDetails(d.npos, d.nprof)
    logev = float(np.log(1e-3))
    total = aligned = 0
    for qp in range(q.nprof):
        print(f"map parity: seed={seed} cap={cap} query={qp} len={q.lengths()[qp]}", flush=True)
        res = cbp.ProcessScoreMatrix(qp, scorethld=4.0, logevthld=logev)
        order = res["order"]
        sel = [int(i) for i in order if res["logE"][i] <= logev]
        if cap > 0:
            sel = sel[:cap]
        assert list(res["map_rows"]) == sel
        assert res["nmap"] == len(sel)
        for i in range(res["npassed"]):
            if i not in sel:
                assert res["alnlen"][i] == 0
        for k, row in enumerate(sel):
            dp = int(res["pronr"][row])
            S, M = orc.score_f32(q.view, qp, d.view, dp)
            r, _, _, _, aln = orc.map(q.view, qp, d.view, dp, S, M, np.float32(res["epa"][row]))
            ctx = (seed, qp, dp)
            assert res["alnlen"][row] == r.alnlen, ctx
            assert res["idnts"][row] == r.idts, ctx
            assert res["psts"][row] == r.psts, ctx
            assert res["ngaps"][row] == r.gaps, ctx
            assert res["beg"][row] == ((r.begx << 16) | r.begy), ctx
            assert res["end"][row] == ((r.x << 16) | r.y), ctx
            assert res["alns"][k] == _strings(aln), ctx
            aligned += r.alnlen > 0
        total += len(sel)
    assert total > 0 and aligned > 0


def test_map_edge_lengths():
    """MAP realignment of every pair of edge-length query and profile lengths
    (block and stripe edges, the corner-mask edge), down to 3 query positions
    and 6 profile positions."""
    EQ = [3, 31, 32, 33, 63, 64, 65, 97, 129]
    ED = [6, 7, 31, 32, 33, 60, 61, 64, 65, 100, 177]
    q = SynthDB.from_lengths(121, EQ)
    d = SynthDB.from_lengths(122, ED)
    orc = oracle()
    cbp = processing(dict(max_map_pairs=0))
    cbp.TransferQueries(q)
    cbp.TransferCPMData(d)
    cbp.SetDbDetails(d.npos, d.nprof)
    n = 0
    for qp in range(q.nprof):
        print(f"map edge: Lq={EQ[qp]}", flush=True)
        res = cbp.ProcessScoreMatrix(qp, scorethld=-1.0, logevthld=50.0)
        sel = [int(i) for i in res["order"] if res["logE"][i] <= 50.0]
        assert list(res["map_rows"]) == sel
        for k, row in enumerate(sel):
            dp = int(res["pronr"][row])
            S, M = orc.score_f32(q.view, qp, d.view, dp)
            r, _, _, _, aln = orc.map(q.view, qp, d.view, dp, S, M, np.float32(res["epa"][row]))
            ctx = (EQ[qp], ED[dp])
            assert res["alnlen"][row] == r.alnlen, ctx
            assert res["beg"][row] == ((r.begx << 16) | r.begy), ctx
            assert res["end"][row] == ((r.x << 16) | r.y), ctx
            assert res["alns"][k] == _strings(aln), ctx
            n += 1
    assert n > 0


def test_map_score_chunk_edges():
    """The dynamic-score pass cuts a profile into 256-column items: profiles
    at and around one and two chunk widths give the oracle's alignments."""
    EQ = [40, 100]
    ED = [255, 256, 257, 511, 512, 513, 700]
    q = SynthDB.from_lengths(131, EQ)
    d = SynthDB.from_lengths(132, ED)
    orc = oracle()
    cbp = processing(dict(max_map_pairs=0))
    cbp.TransferQueries(q)
    cbp.TransferCPMData(d)
    cbp.SetDbDetails(d.npos, d.nprof)
    n = 0
    seen = set()
    for qp in range(q.nprof):
        print(f"map chunk edges: Lq={EQ[qp]}", flush=True)
        res = cbp.ProcessScoreMatrix(qp, scorethld=-1.0, logevthld=50.0)
        sel = [int(i) for i in res["order"] if res["logE"][i] <= 50.0]
        assert list(res["map_rows"]) == sel
        for k, row in enumerate(sel):
            dp = int(res["pronr"][row])
            S, M = orc.score_f32(q.view, qp, d.view, dp)
            r, _, _, _, aln = orc.map(q.view, qp, d.view, dp, S, M, np.float32(res["epa"][row]))
            ctx = (EQ[qp], ED[dp])
            assert res["alnlen"][row] == r.alnlen, ctx
            assert res["beg"][row] == ((r.begx << 16) | r.begy), ctx
            assert res["end"][row] == ((r.x << 16) | r.y), ctx
            assert res["alns"][k] == _strings(aln), ctx
            n += 1
            seen.add(ED[dp])
    assert n > 0 and seen == set(ED)


def _same(a, b):
    return (a["npassed"] == b["npassed"] and a["nmap"] == b["nmap"]
            and np.array_equal(a["alndata"].view(np.uint32), b["alndata"].view(np.uint32))
            and np.array_equal(a["order"], b["order"]) and np.array_equal(a["aln_raw"], b["aln_raw"]))


@pytest.mark.parametrize("qlen", [(60, 120), (250, 330), (560, 640), (1000, 1100)])
def test_map_shorter_group_geometry(qlen):
    """One-query calls realign the pairs no longer than half the longest as
    their own group; whatever rows per lane that group takes (1-5, or the
    long pairs' geometry), every record, order and string is the same."""
    q = SynthDB(81, 1, 0, 0, qlen[0], qlen[1])
    d = SynthDB(82, 120, 160, 0.668, 40, 1500)
    d.plant(q, 83, 30, 3.0, 60.0)
    cbp = processing()
    cbp.TransferQueries(q)
    cbp.TransferCPMData(d)
    cbp.SetDbDetails(d.npos, d.nprof)
    out = {}
    for rmax in (0, 1, 2, 3, 4, 5):
        cbp.set_map_rows(rmax)
        print(f"map shorter group: Lq={int(q.lengths()[0])} rmax={rmax}", flush=True)
        out[rmax] = cbp.ProcessScoreMatrix(0, scorethld=6.0, logevthld=np.log(1e-2))
    assert out[0]["nmap"] >= 3
    for rmax in (1, 2, 3, 4, 5):
        assert _same(out[0], out[rmax]), rmax


@pytest.mark.parametrize("qlens", [[70, 300], [319, 512, 900]])
def test_map_streaming_band_equals_per_pair_bands(qlens):
    """The forward and backward passes with each workgroup streaming its
    pairs through one band give the same records, order and strings as one
    workgroup per pair, for one-query calls (both groups) and batches."""
    q = SynthDB.from_lengths(91, qlens)
    d = SynthDB(92, 400, 160, 0.668, 30, 1500)
    d.plant(q, 93, 60, 3.0, 60.0)
    out = {}
    for pipe in (False, True):
        cbp = processing()
        cbp.set_map_pipe(pipe)
        cbp.TransferQueries(q)
        cbp.TransferCPMData(d)
        cbp.SetDbDetails(d.npos, d.nprof)
        print(f"map streaming band: pipe={pipe} lengths={qlens}", flush=True)
        one = [cbp.ProcessScoreMatrix(i, scorethld=6.0, logevthld=np.log(1e-2)) for i in range(q.nprof)]
        batch = cbp.ProcessBatch(list(range(q.nprof)))
        out[pipe] = (one, batch)
        cbp.close()
    assert sum(r["nmap"] for r in out[False][0]) >= 10
    for a, b in zip(out[False][0] + out[False][1], out[True][0] + out[True][1]):
        assert _same(a, b)
