"""Profiles read from COMER text files searched on the GPU.

A query set and a DB (with context vectors, homologs planted) are written in
the text format (tests/profile_text.py), read back by the library
(cbp_proset_read_text) and searched with the CV term on; the hit lists,
records and strings equal the oracle's on the same read profiles.
"""
import numpy as np
import pytest
from comer2_amd._ffi import SynthDB
from comer2_amd.proio import ProfileSet
from gpu_common import oracle, processing, search_expect, assert_search_equal
from profile_text import write_set

pytestmark = pytest.mark.gpu


def test_text_profiles_search_matches_oracle(tmp_path):
    q0 = SynthDB.from_lengths(8001, np.array([120, 260], np.int32)).add_cv(8101)
    d0 = SynthDB(8002, 90, 180, 0.6, 30, 700).add_cv(8102)
    d0.plant(q0, 8003, 24, 3.0, 60.0)
    write_set(tmp_path / "q.pro", q0.view)
    write_set(tmp_path / "db.pro", d0.view)
    q = ProfileSet(tmp_path / "q.pro")
    d = ProfileSet(tmp_path / "db.pro")
    assert q.has_cv and d.has_cv and d.nprof == 90
    opts = dict(cvswgt=0.15, max_map_pairs=10)
    orc = oracle(opts)
    cbp = processing(opts)
    cbp.TransferQueries(q)
    cbp.TransferCPMData(d)
    cbp.SetDbDetails(d.npos, d.nprof)
    thld, logev = 6.0, float(np.log(1e-2))
    drawn = 0
    for qp in range(q.nprof):
        print(f"text profiles: query {qp} ({q.name(qp)})", flush=True)
        res = cbp.ProcessScoreMatrix(qp, thld, logev)
        exp = search_expect(orc, q, qp, d, thld, logev, d.npos, d.nprof, 10)
        assert_search_equal(res, exp, ("text", qp))
        drawn += len(exp["sel"])
        if exp["sel"]:
            _check_output(res, q, qp, d, opts)
    assert drawn >= 5


def _check_output(res, q, qp, d, opts):
    """the plain output of a searched query: one annotation and one record
    per drawn hit, in order, with the drawn strings in their fragments"""
    import ctypes as C
    from comer2_amd import tables
    from comer2_amd.output import format_plain
    o = dict(tables.default_options(), **opts)
    ns = int(d.npos)
    sss = bytes(np.ctypeslib.as_array(C.cast(d.view.sss, C.POINTER(C.c_uint8)), (ns,)))
    q0, ql = int(q.dists()[qp]), int(q.lengths()[qp])
    qsss = bytes(np.ctypeslib.as_array(C.cast(q.view.sss, C.POINTER(C.c_uint8)), (int(q.npos),)))[q0:q0 + ql]
    txt = format_plain(res, o, q.desc(qp), ql, float(q.view.eno[qp]), ["db.pro"], d.npos, d.nprof,
                       [d.desc(i) for i in range(d.nprof)], d.lengths(), [d.view.eno[i] for i in range(d.nprof)],
                       query_sss=qsss, pro_sss=sss, pro_dist=d.dists(), alnwidth=60)
    rows = list(res["map_rows"])
    descs = [d.desc(int(res["pronr"][k])) for k in rows]
    pos = 0
    for k, ds in zip(rows, descs):
        pos = txt.index(">" + ds + "\n", pos)
    assert txt.count("\n>") == len(rows) and txt.count("struct       ") >= 2 * len(rows)
    for j, k in enumerate(rows):
        qs = res["alns"][j][0]
        frag = "".join(l[13:-6].rstrip() for l in txt.split("\n") if l.startswith("Query: "))
        assert qs[:60] in frag
    # the JSON file of the same results parses, and its records carry the
    # drawn strings in the same order
    import json
    from comer2_amd.output import format_json
    js = format_json(res, o, q.desc(qp), ql, float(q.view.eno[qp]), ["db.pro"], d.npos, d.nprof,
                     [d.desc(i) for i in range(d.nprof)], d.lengths(), [d.view.eno[i] for i in range(d.nprof)],
                     query_sss=qsss, pro_sss=sss, pro_dist=d.dists())
    doc = json.loads(js)["comer_search"]
    recs = [h["hit_record"] for h in doc["search_hits"]]
    assert len(recs) == len(rows) == len(doc["search_summary"])
    for j, rec in enumerate(recs):
        a = rec["alignment"]
        assert rec["target_description"] == descs[j]
        assert (a["query_aln"], a["middle"], a["target_aln"]) == tuple(res["alns"][j])
        assert a["query_secstr_available"] == 1 and len(a["query_secstr"]) == len(a["query_aln"])


def test_binary_database_search_equals_text(tmp_path):
    """the same profiles from a binary database (format v2.2) search
    exactly as from the text files, resident and streamed in chunks"""
    from profile_bin import write_bin
    q0 = SynthDB.from_lengths(8201, np.array([150, 330], np.int32)).add_cv(8301)
    d0 = SynthDB(8202, 80, 160, 0.6, 30, 600).add_cv(8302)
    d0.plant(q0, 8203, 20, 3.0, 60.0)
    write_set(tmp_path / "q.pro", q0.view)
    write_set(tmp_path / "db.pro", d0.view)
    q = ProfileSet(tmp_path / "q.pro")
    t = ProfileSet(tmp_path / "db.pro")
    write_bin(tmp_path / "db.bin", t.view, [t.desc(i) for i in range(t.nprof)])
    b = ProfileSet.from_bin(tmp_path / "db.bin")
    opts = dict(cvswgt=0.15, max_map_pairs=10)
    cbp = processing(opts)
    cbp.TransferQueries(q)
    qs = list(range(q.nprof))
    cbp.TransferCPMData(t)
    cbp.SetDbDetails(t.npos, t.nprof)
    rt = cbp.ProcessBatch(qs)
    cbp.TransferCPMData(b)
    rb = cbp.ProcessBatch(qs)
    rs = cbp.SearchDB(b, qs, chunk_pos=b.npos // 2 + 1)
    assert cbp.last_search_chunks() >= 2
    for i in qs:
        for other, ctx in ((rb[i], "binary"), (rs[i], "streamed")):
            a = rt[i]
            assert a["npassed"] == other["npassed"] and list(a["order"]) == list(other["order"]), (ctx, i)
            assert np.array_equal(a["alndata"].view(np.uint32), other["alndata"].view(np.uint32)), (ctx, i)
            assert list(a["alns"]) == list(other["alns"]), (ctx, i)
    assert sum(r["nmap"] for r in rt) >= 3
