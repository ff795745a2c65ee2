"""The JSON output (cbp_format_json) on the CPU.

Known answer: the hand-made result of tests/test_output.py (three passed
pairs, one not drawn) formatted as the reference's JSON file; the expected
text is written here from the format strings of cualn/AlnWriterJSON.cpp and
cupro/CuBatchProcessingFinalizerJSON.cpp: programme, references, query,
database, message, summary entries, hit records (scores, coordinates with
the last one skipping gaps, SS strings, alignment rows, statistics with
"NA" where the reference prints it) and the search statistics.  The text
must also parse as JSON with the same values.  Descriptions
(FormatDescriptionJSON): '"', '\\' and '/' escaped, blanks collapsed, "..."
at the length limit with backslashes just before it turned to '.' too.
Separators: a record followed by ",\\n" unless it is the last of the hits or
of the first NOHITS.
"""
import json
import math
import numpy as np
from comer2_amd import tables
from comer2_amd.output import format_json
from test_output import _result


def _f(x):
    return np.float32(x)


def _expected(opts, sss=None, msg=True):
    E = ['{"comer_search": {\n', '  "program": "comer",\n', '  "version": "2.3",\n',
         '  "references": [\n', '    "Ref one.",\n', '    "Ref two words."\n', '  ],\n',
         '  "query": {\n', '    "length": 12,\n', '    "name": "",\n', '    "description": "query one"\n', '  },\n',
         '  "database": {\n', '    "names": [\n', '      "db1"\n', '    ],\n',
         '    "number_of_profiles": 10,\n', '    "number_of_positions": 1000\n', '  },\n']
    if msg:
        E.append('  "message": "Profiles found below the e-value threshold:",\n')
    E.append('  "search_summary": [\n')
    hits = [("prof A desc", 33.3, -10.0), ("prof C", 20.0, -3.0)]
    for i, (d, sc, le) in enumerate(hits):
        E.append('    {"summary_entry": {\n      "description": "%s",\n      "score": %.0f,\n'
                 '      "evalue": %.2g\n    }}' % (d, _f(sc), math.exp(_f(le))))
        E.append(",\n" if i + 1 < len(hits) else "\n")
    E.append('  ],\n  "search_hits": [\n')
    sspace = _f(12000.0)
    recs = [("prof A desc", 20, 2.0, 33.3, -10.0, 3, 5, 1, 8, (2, 8, 3, 9), "AC-DEFGH", "ACKD-FGH", "A+ D  GH"),
            ("prof C", 30, 7.25, 20.0, -3.0, 0, 0, 0, 5, (1, 5, 1, 5), "MNPQR", "MNPQR", "     ")]
    for i, (d, tl, te, sc, le, idt, pst, gp, n, co, qa, ta, mid) in enumerate(recs):
        ev = math.exp(_f(le))
        bits = _f((_f(math.log(sspace)) - _f(le)) / _f(math.log(2)))
        pv = ev if ev < 0.01 else 1.0 - math.exp(-ev)
        E.append('    {"hit_record": {\n      "target_description": "%s",\n' % d)
        E.append('      "query_length": 12,\n      "query_eno": 3.5,\n      "target_length": %d,\n'
                 '      "target_eno": %.1f,\n      "alignment": {\n' % (tl, te))
        E.append('        "score": %.2f,\n        "bit_score": %.1f,\n        "evalue": %.2g,\n'
                 '        "pvalue": %.2g,\n' % (_f(sc), bits, ev, pv))
        E.append('        "n_identities": %d,\n        "n_positives": %d,\n        "n_gaps": %d,\n'
                 '        "aln_length": %d,\n' % (idt, pst, gp, n))
        E.append('        "query_from": %d,\n        "query_to": %d,\n        "target_from": %d,\n'
                 '        "target_to": %d,\n' % co)
        av = 1 if sss else 0
        qs, ts = sss[i] if sss else ("", "")
        E.append('        "query_secstr_available": %d,\n        "target_secstr_available": %d,\n' % (av, av))
        E.append('        "query_secstr": "%s",\n        "target_secstr": "%s",\n' % (qs, ts))
        E.append('        "query_aln": "%s",\n        "target_aln": "%s",\n        "middle": "%s",\n' % (qa, ta, mid))
        E.append('        "computed_K_ungapped": "0.1000",\n        "computed_lambda_ungapped": "0.3000",\n'
                 '        "estimated_K_gapped": "0.0500",\n        "estimated_lambda_gapped": "0.2500",\n'
                 '        "entropy": 0.4000,\n        "expected_score": -0.5000,\n        "min_score": -5,\n'
                 '        "max_score": 10\n')
        E.append('      }\n    }}')
        E.append(",\n" if i + 1 < len(recs) else "\n")
    E.append('  ],\n')
    E.append('  "search_statistics": {\n    "reference_K_ungapped": %.4f,\n    "reference_lambda_ungapped": %.4f,\n'
             '    "reference_K_gapped": %.4f,\n    "reference_lambda_gapped": %.4f,\n'
             % (_f(opts["ung_K"]), _f(opts["ung_lambda"]), _f(opts["gap_K"]), _f(opts["gap_lambda"])))
    E.append('    "query_length": 12,\n    "database_size": 1000,\n    "effective_query_length": 10,\n'
             '    "effective_database_size": 980,\n    "effective_search_space": 12000\n  }\n}}\n')
    return "".join(E)


def _fmt(opts, res=None, **kw):
    base = dict(query_desc="query one", query_len=12, query_eno=3.5, db_names=["db1"], db_size=1000,
                db_entries=10, pro_desc=["prof A desc", "prof B", "prof C"], pro_len=[20, 15, 30],
                pro_eno=[2.0, 4.5, 7.25], deltalen=2, progname="comer", progversion="2.3",
                references=["Ref one.", "Ref two words."])
    base.update(kw)
    return format_json(res or _result(), opts, **base)


def _diff(got, want):
    return "\n".join("%d: %r | %r" % (i, a, b)
                     for i, (a, b) in enumerate(zip(got.split("\n"), want.split("\n"))) if a != b)


def test_json_output_known_answer():
    opts = dict(tables.default_options(), ssswgt=0.0)
    got = _fmt(opts)
    want = _expected(opts)
    assert got == want, _diff(got, want)
    doc = json.loads(got)["comer_search"]
    assert doc["references"] == ["Ref one.", "Ref two words."]
    assert doc["query"] == {"length": 12, "name": "", "description": "query one"}
    assert [h["summary_entry"]["description"] for h in doc["search_summary"]] == ["prof A desc", "prof C"]
    a = doc["search_hits"][0]["hit_record"]["alignment"]
    assert (a["query_from"], a["query_to"], a["target_from"], a["target_to"]) == (2, 8, 3, 9)
    assert a["query_aln"] == "AC-DEFGH" and a["middle"] == "A+ D  GH"
    assert doc["search_statistics"]["effective_search_space"] == 12000


def test_json_secstr_strings():
    # the strings are drawn under SSSWGT > 0 when both carry SS states
    opts = dict(tables.default_options(), ssswgt=0.12)
    qsss = b"  eehhee  hh"
    psss = (b" e" * 10) + (b"h" * 15) + (b"eh " * 10)
    got = _fmt(opts, query_sss=qsss, pro_sss=psss, pro_dist=[0, 20, 35])
    q1 = "".join(" " if c == "-" else chr(qsss[p]) for c, p in zip("AC-DEFGH", [1, 2, 0, 3, 4, 5, 6, 7]))
    t1 = "".join(" " if c == "-" else chr(psss[p]) for c, p in zip("ACKD-FGH", [2, 3, 4, 5, 0, 6, 7, 8]))
    q2 = "".join(chr(qsss[p]) for p in range(5))
    t2 = "".join(chr(psss[35 + p]) for p in range(5))
    want = _expected(opts, [(q1, t1), (q2, t2)])
    assert got == want, _diff(got, want)
    # without SS on one side the flags are 0 and the strings empty
    assert _fmt(opts, query_sss=qsss) == _expected(opts)


def test_json_descriptions_escape_and_truncate():
    opts = dict(tables.default_options(), ssswgt=0.0)
    got = _fmt(opts, pro_desc=['  a "quoted"\tpath/to\\x  end', "prof B", "y" * 80])
    assert '"description": "a \\"quoted\\" path\\/to\\\\x end"' in got
    assert '"target_description": "a \\"quoted\\" path\\/to\\\\x end"' in got
    # summary descriptions are cut at 75: 72 characters and "..."
    assert '"description": "' + "y" * 72 + '..."' in got
    # a cut through an escape (columns 74 and 75): it falls under the dots
    d = "z" * 73 + "/abc"
    got = _fmt(opts, pro_desc=[d, "prof B", "prof C"])
    assert '"description": "' + "z" * 72 + '...",' in got
    doc = json.loads(_fmt(opts, pro_desc=[d, "prof B", 'q"\\/']))["comer_search"]
    assert doc["search_summary"][1]["summary_entry"]["description"] == 'q"\\/'


def test_json_cut_through_an_escape_exactly():
    # 73 characters then '/': its escape takes columns 74 and 75, the loop
    # stops there; the last three become '.' and the one before them (a
    # plain 'z') stays
    opts = dict(tables.default_options(), ssswgt=0.0)
    got = _fmt(opts, pro_desc=["z" * 73 + "/b", "prof B", "prof C"])
    assert '"description": "' + "z" * 72 + '...",' in got
    # a run of backslashes before the cut turns to '.' as well
    got = _fmt(opts, pro_desc=["z" * 70 + "\\\\" + "abc", "prof B", "prof C"])
    # 70 z, two escaped backslashes (columns 71-74), 'a' (75): cut; the
    # last three and the two backslashes before them become '.'
    assert '"description": "' + "z" * 70 + '.....",' in got


def test_json_no_hits_and_separators():
    opts = dict(tables.default_options(), ssswgt=0.0)
    res = _result()
    res["logevthld"] = -20.0
    got = _fmt(opts, res=res)
    assert '  "message": "No profiles found below an e-value threshold of %g",\n' % math.exp(np.float32(-20.0)) in got
    assert '  "search_summary": [\n  ],\n  "search_hits": [\n  ],\n' in got
    json.loads(got)
    # NOHITS bounds the entries and the records' separators: one entry, two
    # records, the second following "\n" after the first (the reference's
    # writer checks NOHITS for both lists)
    got = _fmt(opts, nohits=1)
    doc_text = got
    assert doc_text.count('"summary_entry"') == 1 and doc_text.count('"hit_record"') == 2
    assert '"max_score": 10\n      }\n    }}\n    {"hit_record"' in doc_text
    # NOALNS < NOHITS: the last record printed keeps its comma
    got = _fmt(opts, noalns=1)
    assert got.count('"hit_record"') == 1 and '    }},\n  ],\n' in got


def test_json_database_names_limit():
    opts = dict(tables.default_options(), ssswgt=0.0)
    long = "d" * 120
    got = _fmt(opts, db_names=[long, "b/c"])
    # a name is cut at its own length + 1: an escape in it pushes it there,
    # and its end becomes "..." (b, \\, /, c: 4 columns)
    assert '      "' + "d" * 97 + '...",\n      "b..."\n' in got
    # 1 KB of scratch: 280 for the section, 15 + name per name, then "..."
    many = ["n%02d" % i for i in range(60)]
    got = _fmt(opts, db_names=many)
    # 1024 - 280 = 744; each 3-character name takes 18: 41 names fit (738)
    assert '      "n40",\n      "..."\n    ],\n' in got and '"n41"' not in got
    # the message needs 360 more: gone once the names took more than 384
    assert '"message"' not in got


def test_json_without_references():
    # no reference list: one empty entry, and nothing is read past the list
    opts = dict(tables.default_options(), ssswgt=0.0)
    got = _fmt(opts, references=None)
    want = _expected(opts).replace('    "Ref one.",\n    "Ref two words."\n', '    ""\n')
    assert got == want, _diff(got, want)
    assert json.loads(got)["comer_search"]["references"] == [""]
