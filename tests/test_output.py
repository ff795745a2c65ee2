"""The plain-text output (cbp_format_plain) on the CPU.

Known answers: a hand-made result of three passed pairs (one not drawn,
so not a hit) formatted with the fragments 5 columns wide and the
statistics footer on; the expected text is written here line by line from
the reference's format strings (cualn/AlnWriterPlain.cpp and
cupro/CuBatchProcessingFinalizerPlain.cpp): the programme header, the
query, the database lines, the hit-list heading padded to 75 columns, one
annotation per hit, each record's description, scores, fragments with
running coordinates (a fragment's last coordinate skips gaps), footer, and
the summary with the effective lengths.  Description wrapping
(FormatDescription): a word that would reach the width starts a line
indented by 13; an annotation longer than 75 ends in "..."; the
"struct" lines carry the SS states along the alignment with ' ' at gaps.
"""
import math
import numpy as np
import pytest
from comer2_amd import _ffi, tables
from comer2_amd.output import format_plain, search_space


def _rec(score, orgpro, psts, beg_xy, logE, alnlen, idts, gaps, K=0.1, lmbd=0.3, Ke=0.05, le=0.25,
         E=-0.5, H=0.4, smin=-5.0, smax=10.0):
    r = np.zeros(_ffi.NFIELDS, np.float32)
    u = r.view(np.uint32)
    r[_ffi.F_SCORE] = score
    u[_ffi.F_ORGPRONO] = orgpro
    r[_ffi.F_PSTVS] = psts
    u[_ffi.F_BEGCOORDS] = (beg_xy[0] << 16) | beg_xy[1]
    r[_ffi.F_EVALUE] = logE
    u[_ffi.F_EPA] = alnlen
    r[_ffi.F_IDNTS] = idts
    r[_ffi.F_NGAPS] = gaps
    r[_ffi.F_K], r[_ffi.F_LMBD], r[_ffi.F_KEST], r[_ffi.F_LMBDEST] = K, lmbd, Ke, le
    r[_ffi.F_E], r[_ffi.F_H], r[_ffi.F_MIN], r[_ffi.F_MAX] = E, H, smin, smax
    return r


def _raw(rows, stride=16):
    a = np.zeros((len(rows), 3, stride), np.uint8)
    for j, trip in enumerate(rows):
        for k, s in enumerate(trip):
            a[j, k, :len(s)] = np.frombuffer(s.encode(), np.uint8)
    return a


def _result():
    recs = [_rec(33.3, 0, 5, (2, 1), -10.0, 8, 3, 1),
            _rec(40.0, 1, 9, (0, 0), -12.0, 0, 0, 0),
            _rec(20.0, 2, 0, (0, 0), -3.0, 5, 0, 0)]
    return dict(npassed=3, nmap=2, alndata=np.stack(recs), order=np.array([1, 0, 2], np.int32),
                map_rows=np.array([0, 2], np.int32),
                aln_raw=_raw([("AC-DEFGH", "A+ D  GH", "ACKD-FGH"), ("MNPQR", "     ", "MNPQR")]),
                logevthld=0.0, search_space=12000.0)


def _expected(opts, sss_rows=None):
    E = []
    E.append("comer 2.3\n\n")
    E.append("Ref one.\nRef two words.\n\n")
    E.append("\n Query (12 positions):\nquery one")
    E.append("\n\n Database(s):\ndb1\n")
    E.append(" " * 13 + "10 profiles\n")
    E.append(" " * 13 + "1000 total positions\n\n\n")
    E.append(" Profiles found below the e-value threshold:".ljust(75) + " %7s %7s\n\n" % ("Score", "E-value"))
    for desc, score, logE in (("prof A desc", 33.3, -10.0), ("prof C", 20.0, -3.0)):
        E.append(desc.ljust(75) + " %7.0f %7.1g\n" % (np.float32(score), math.exp(np.float32(logE))))
    E.append("\n\n\n")
    sspace = np.float32(12000.0)

    def head(desc, slen, seno, score, logE):
        ev = math.exp(np.float32(logE))
        bits = np.float32((np.float32(math.log(sspace)) - np.float32(logE)) / np.float32(math.log(2)))
        pv = ev if ev < 0.01 else 1.0 - math.exp(-ev)
        return (">%s\n" % desc + "  Length/ENO: Query = 12/3.5, Sbjct = %d/%.1f\n\n" % (slen, seno)
                + " Score = %.2f (%.1f bits),  Expect = %.2g, P-value = %.2g\n" % (np.float32(score), bits, ev, pv))

    def footer():
        return ("%25s  %-6s   %-6s\n" % (" ", "K", "Lambda")
                + "%-25s  %6s   %6s\n" % ("Computed  ungapped,", "0.1000", "0.3000")
                + "%-25s  %6s   %6s\n" % ("Estimated gapped,", "0.0500", "0.2500")
                + "Entropy, 0.4000; Expected, -0.5000; Min/Max, -5/10\n\n")

    s1 = sss_rows[0] if sss_rows else None
    E.append(head("prof A desc", 20, 2.0, 33.3, -10.0))
    E.append(" Identities = 3/8 (37%), Positives = 5/8 (62%), Gaps = 1/8 (12%)\n\n")
    if s1: E.append("struct       " + s1[0][0:5] + "\n")
    E.append("Query:     2 AC-DE 5    \n" + " " * 13 + "A+ D \n" + "Sbjct:     3 ACKD- 6    \n")
    if s1: E.append("struct       " + s1[1][0:5] + "\n")
    E.append("\n")
    if s1: E.append("struct       " + s1[0][5:8] + "\n")
    E.append("Query:     6 FGH 8    \n" + " " * 13 + " GH\n" + "Sbjct:     7 FGH 9    \n")
    if s1: E.append("struct       " + s1[1][5:8] + "\n")
    E.append("\n")
    E.append(footer() + "\n\n")
    s2 = sss_rows[1] if sss_rows else None
    E.append(head("prof C", 30, 7.25, 20.0, -3.0) + "\n\n")
    if s2: E.append("struct       " + s2[0] + "\n")
    E.append("Query:     1 MNPQR 5    \n" + " " * 13 + "     \n" + "Sbjct:     1 MNPQR 5    \n")
    if s2: E.append("struct       " + s2[1] + "\n")
    E.append("\n")
    E.append(footer() + "\n\n")
    E.append("%-20s  %-6s   %-6s\n" % ("Reference values of", "K", "Lambda"))
    E.append("%-20s  %6.4f   %6.4f\n" % ("Ungapped", np.float32(opts["ung_K"]), np.float32(opts["ung_lambda"])))
    E.append("%-20s  %6.4f   %6.4f\n\n" % ("Gapped", np.float32(opts["gap_K"]), np.float32(opts["gap_lambda"])))
    E.append("Length of query, 12\nLength of database, 1000\nEffective length of query, 10\n"
             "Effective length of database, 980\nEffective search space, 12000\n\n")
    return "".join(E)


def _fmt(opts, **kw):
    base = dict(query_desc="query one", query_len=12, query_eno=3.5, db_names=["db1"], db_size=1000,
                db_entries=10, pro_desc=["prof A desc", "prof B", "prof C"], pro_len=[20, 15, 30],
                pro_eno=[2.0, 4.5, 7.25], deltalen=2, progname="comer", progversion="2.3",
                references=["Ref one.", "Ref two words."], alnwidth=5)
    base.update(kw)
    return format_plain(_result(), opts, **base)


def test_plain_output_known_answer():
    opts = dict(tables.default_options(), ssswgt=0.0)
    got = _fmt(opts)
    want = _expected(opts)
    assert got == want, "\n".join(
        "%d: %r | %r" % (i, a, b) for i, (a, b) in enumerate(zip(got.split("\n"), want.split("\n"))) if a != b)


def test_plain_output_struct_lines():
    opts = dict(tables.default_options(), ssswgt=0.12)
    # query positions 0..11, DB profiles at 0, 20, 35 (lengths 20, 15, 30)
    qsss = b"  eehhee  hh"
    psss = (b" e" * 10) + (b"h" * 15) + (b"eh " * 10)
    got = _fmt(opts, query_sss=qsss, pro_sss=psss, pro_dist=[0, 20, 35])
    # pair 1: query from position 1 ("AC-DEFGH": gap at column 2), subject
    # from position 2 of profile 0 ("ACKD-FGH": gap at column 4)
    q1 = "".join(" " if c == "-" else chr(qsss[p]) for c, p in zip("AC-DEFGH", [1, 2, 0, 3, 4, 5, 6, 7]))
    t1 = "".join(" " if c == "-" else chr(psss[0 + p]) for c, p in zip("ACKD-FGH", [2, 3, 4, 5, 0, 6, 7, 8]))
    q2 = "".join(chr(qsss[p]) for p in range(5))
    t2 = "".join(chr(psss[35 + p]) for p in range(5))
    want = _expected(opts, [(q1, t1), (q2, t2)])
    assert got == want, "\n".join(
        "%d: %r | %r" % (i, a, b) for i, (a, b) in enumerate(zip(got.split("\n"), want.split("\n"))) if a != b)


def test_descriptions_wrap_and_truncate():
    opts = dict(tables.default_options(), ssswgt=0.0)
    words = "alpha beta gamma delta epsilon zeta eta theta"
    got = _fmt(opts, pro_desc=[words, "prof B", "x" * 100], dscwidth=30)
    # the record: '>' then words; a word that would reach column 30 opens a
    # new line indented by 13
    assert ">alpha beta gamma delta\n             epsilon zeta eta\n             theta\n" in got
    # the annotations never wrap (width = length = 75); a longer one ends in "..."
    assert words.ljust(75) + " " in got
    assert "x" * 72 + "... " in got


def test_no_hits_line_and_limits():
    opts = dict(tables.default_options(), ssswgt=0.0)
    res = _result()
    res["logevthld"] = -20.0
    got = format_plain(res, opts, "q", 12, 3.5, ["db1"], 1000, 10, ["a", "b", "c"], [20, 15, 30],
                       [2.0, 4.5, 7.25], deltalen=2)
    assert " No profiles found below an e-value threshold of %g.\n\n\n" % math.exp(-20.0) in got
    assert ">" not in got
    got = _fmt(opts, nohits=1, noalns=0)
    assert "prof A desc".ljust(75) in got and "prof C".ljust(75) not in got and ">" not in got


def test_search_space_matches_the_oracle():
    from oracle_ffi import lib as olib
    import ctypes as C
    opts = tables.default_options()
    L = olib()
    L.or_length_adjustment.restype = C.c_uint64
    for q, db, n in ((300, 50_037_387, 200_000), (50, 1_000_000, 4_000), (1000, 20_000, 100)):
        ss, dl = search_space(opts, q, db, n)
        o = _ffi.options_struct(opts)
        d = C.c_uint64()
        want = L.or_length_adjustment(C.byref(o), q, db, n, C.byref(d))
        assert ss == want and dl == d.value, (q, db, n)
