"""Whole-DB search in chunks (cbp_search_db) on the GPU.

The DB is cut into chunks of whole profiles, uploaded from pinned staging
buffers by a copy thread while the previous chunk is searched, and each
query's per-chunk hit lists are merged (the reference's chunk loop,
cualn/JobDispatcher.cpp:1103-1134, and AlnWriter::MergeResults,
cualn/AlnWriter.cpp:218-267).  The merged result must equal the search of the
DB as one chunk bit for bit -- records, (log E, profile index) order, drawn
rows, strings -- with the cap on drawn pairs cutting inside some chunks, for
MAPALN 1 and 0; the one-chunk search equals the batch call on the uploaded
chunk and the oracle.
"""
import numpy as np
import pytest
from comer2_amd._ffi import SynthDB
from gpu_common import oracle, processing, search_expect, assert_search_equal

pytestmark = pytest.mark.gpu


def _sets():
    q = SynthDB(2102, 3, 300, 0.5, 100, 700)
    d = SynthDB(2103, 400, 200, 0.668, 30, 3000)
    d.plant(q, 2104, 60, 3.0, 60.0)
    return q, d


def _same(a, b, ctx):
    assert a["npassed"] == b["npassed"], ctx
    assert a["nposits"] == b["nposits"] and a["maxlen"] == b["maxlen"], ctx
    assert a["search_space"] == b["search_space"], ctx
    assert np.array_equal(a["alndata"].view(np.uint32), b["alndata"].view(np.uint32)), ctx
    assert np.array_equal(a["order"], b["order"]), ctx
    assert np.array_equal(a["epa"].view(np.uint32), b["epa"].view(np.uint32)), ctx
    assert list(a["map_rows"]) == list(b["map_rows"]), ctx
    assert a["alns"] == b["alns"], ctx


@pytest.mark.parametrize("mapaln,cap", [(1, 6), (1, 0), (0, 5)])
def test_chunked_search_equals_one_chunk(mapaln, cap):
    q, d = _sets()
    opts = dict(mapaln=mapaln, max_map_pairs=cap)
    cbp = processing(opts)
    cbp.TransferQueries(q)
    qs = [0, 1, 2]
    print(f"stream: mapaln={mapaln} cap={cap} one chunk of {d.npos} positions", flush=True)
    one = cbp.SearchDB(d, qs, chunk_pos=d.npos)
    assert cbp.last_search_chunks() == 1
    chunk = max(3000, d.npos // 4)
    print(f"stream: chunks of at most {chunk} positions", flush=True)
    many = cbp.SearchDB(d, qs, chunk_pos=chunk)
    assert cbp.last_search_chunks() >= 4
    drawn = 0
    for i in qs:
        _same(many[i], one[i], ("chunks vs one", mapaln, cap, i))
        drawn += one[i]["nmap"]
        if cap:
            assert one[i]["nmap"] <= cap
    assert drawn >= 6
    # the one-chunk search is the batch call on the uploaded chunk
    cbp.TransferCPMData(d)
    batch = cbp.ProcessBatch(qs)
    for i in qs:
        _same(one[i], batch[i], ("one vs batch", mapaln, cap, i))
    # and the oracle's, for the first query
    orc = oracle(opts)
    thld, logev = cbp.score_threshold(int(q.lengths()[0]), d.npos)
    exp = search_expect(orc, q, 0, d, thld, logev, d.npos, d.nprof, cap)
    assert_search_equal(many[0], exp, ("stream oracle", mapaln, cap))
