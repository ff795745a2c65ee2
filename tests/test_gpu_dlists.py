"""Stage lists built on the device (cbp_set_device_lists, the default).

With device lists the candidate, pass and drawing lists, the (log E,
profile) order and the MAP plan are ordered compactions and sorts on the
device, so no host decision separates the stages and a HIP graph replays
the whole sequence.  Results must equal the host-built lists' bit for bit:
one-query calls (the long-pair MAP group on its own stream), batches from
the work queue, MAPALN=0, graph replays of the whole sequence, and calls
whose lists exceed small capacities (they run again with host lists).
"""
import numpy as np
import pytest
from comer2_amd._ffi import SynthDB
from gpu_common import processing

pytestmark = pytest.mark.gpu


def _sets():
    q = SynthDB(2502, 5, 300, 0.6, 60, 1000)
    d = SynthDB(2503, 300, 200, 0.668, 30, 3000)
    d.plant(q, 2504, 70, 3.0, 60.0)
    return q, d


def _same(a, b, ctx):
    assert a["npassed"] == b["npassed"] and a["nmap"] == b["nmap"], ctx
    assert a["nposits"] == b["nposits"] and a["maxlen"] == b["maxlen"], ctx
    assert a["search_space"] == b["search_space"], ctx
    assert np.array_equal(a["alndata"].view(np.uint32), b["alndata"].view(np.uint32)), ctx
    assert np.array_equal(a["order"], b["order"]), ctx
    assert np.array_equal(a["epa"].view(np.uint32), b["epa"].view(np.uint32)), ctx
    assert list(a["map_rows"]) == list(b["map_rows"]), ctx
    assert np.array_equal(a["aln_raw"], b["aln_raw"]), ctx


@pytest.mark.parametrize("mapaln,cap", [(1, 8), (1, 0), (0, 6)])
def test_device_lists_equal_host_lists(mapaln, cap):
    q, d = _sets()
    cbp = processing(dict(mapaln=mapaln, max_map_pairs=cap))
    cbp.TransferQueries(q)
    cbp.TransferCPMData(d)
    cbp.SetDbDetails(d.npos, d.nprof)
    qs = list(range(q.nprof))
    cbp.set_device_lists(False)
    host1 = [cbp.ProcessScoreMatrix(i) for i in qs]
    hostb = cbp.ProcessBatch(qs)
    cbp.set_device_lists(True)
    print(f"device lists: mapaln={mapaln} cap={cap}", flush=True)
    drawn = 0
    for i in qs:
        r = cbp.ProcessScoreMatrix(i)
        assert cbp.last_fallbacks() == 0
        _same(r, host1[i], ("one query", mapaln, cap, i))
        drawn += r["nmap"]
    assert drawn >= 10
    devb = cbp.ProcessBatch(qs)
    assert cbp.last_fallbacks() == 0
    for i in qs:
        _same(devb[i], hostb[i], ("batch", mapaln, cap, i))
        _same(devb[i], host1[i], ("batch vs one", mapaln, cap, i))
    # the whole sequence of the batch from a graph: eager on the first call,
    # captured on the second, replayed after; eager work between the replays
    cbp.set_graph(True)
    for rep in range(4):
        g = cbp.ProcessBatch(qs)
        for i in qs:
            _same(g[i], hostb[i], ("graph", rep, i))
        cbp.phase1_keys(0)
    cbp.set_graph(False)


def test_device_lists_fall_back_beyond_capacity():
    q, d = _sets()
    cbp = processing(dict(max_map_pairs=8))
    cbp.TransferQueries(q)
    cbp.TransferCPMData(d)
    cbp.SetDbDetails(d.npos, d.nprof)
    qs = list(range(q.nprof))
    cbp.set_device_lists(False)
    host = cbp.ProcessBatch(qs)
    cbp.set_device_lists(True)
    # candidates, passed pairs, backtrack words and MAP cells each too small
    for caps in ((2, 2, 1 << 27, 1 << 28), (65536, 2, 1 << 27, 1 << 28),
                 (65536, 16384, 4096, 1 << 28), (65536, 16384, 1 << 27, 4096)):
        cbp.set_device_list_caps(*caps)
        print(f"device lists: caps {caps}", flush=True)
        res = cbp.ProcessBatch(qs)
        assert cbp.last_fallbacks() == len(qs), caps
        for i in qs:
            _same(res[i], host[i], ("fallback", caps, i))


def test_device_lists_per_query_launches_fresh_handle():
    """A batch with the work queue off on a fresh handle: every query's
    phase 1 is prepared and launched ahead of its chain (a regression test:
    the per-query maxima of queries after the first were once read before
    any phase 1 had written them)."""
    q, d = _sets()
    cbp = processing(dict(max_map_pairs=8))
    cbp.TransferQueries(q)
    cbp.TransferCPMData(d)
    cbp.SetDbDetails(d.npos, d.nprof)
    qs = list(range(q.nprof))
    cbp.set_queue(False)
    dev = cbp.ProcessBatch(qs)
    cbp.set_device_lists(False)
    host = cbp.ProcessBatch(qs)
    for i in qs:
        _same(dev[i], host[i], ("per-query launches", i))
