"""The COMER text profile reader (cbp_proset_read_text) on the CPU.

* Known answer: a hand-written file of two profiles (the first with HDP
  posteriors, an HDP context line, context vectors and three-probability SS
  lines; the second with none of these and the one-probability SS form and
  the short statistics line) read into fields whose values follow from the
  format's rules (cupro/IOProfileModel.cu:46-1021): value = integer / SCALE;
  transitions as logs, IM II DM DD times 0.4, zero as -32768, ID and DI
  dropped; a profile without context vectors takes the uninformative vector
  (0.17302947, prior 0, norm 0); SS states stored as ' ', 'e', 'h'.
* Round trip: synthetic profiles (with context vectors) written in the
  condensed layout (tests/profile_text.py) read back as their quantized
  values.
* Malformed files are refused with the file and line in the message.
"""
import ctypes as C
import numpy as np
import pytest
from comer2_amd._ffi import SynthDB
from comer2_amd.proio import ProfileSet, ProfileReadError
from profile_text import write_set

_libm = C.CDLL("libm.so.6")
_libm.logf.argtypes = [C.c_float]
_libm.logf.restype = C.c_float


def logf(x):
    return np.float32(_libm.logf(float(np.float32(x))))


def q(i, scale):
    return np.float32(i) / np.float32(scale)


ALPHA = "ARNDCQEGHILKMFPSTWYV"
TINY = """# a comment line
COMER profile v2.0
DESC: tiny test profile
FILE: tiny.pro
CMD: makepro -i tiny.afa -o tiny.pro
LEN: 2
NOS: 12
EFFNOS: 3.5
SCALE: 1000
 A R N D C Q E G H I L K M F P S T W Y V
NULL""" + " 50" * 20 + """
POST""" + " 0" * 20 + """
 900 50 50 600 400 0 1000 0 0
 1000 0 0
1 A 1000""" + " 0" * 19 + """
 800 100 100 500 500 0 250 0 750
 1000 0 0
 50 2
 700 200
 3 5
 CT:2 100 500 400 1 2
 CV: -12345 4500 19""" + "".join(" %d" % v for v in range(-900, 1000, 100)) + """
 SS:E 100 800 100
2 W""" + " 0" * 17 + " 500 500 0" + """
 1000 0 0 0 0 0 0 0 0
 1000 0 0
 0 0
 CT:0 0
 CV: -8000 2000 19""" + " 50" * 19 + """
 SS:h 0 0 1000

K, Lambda:
Computed  ungapped, 0.1000 0.3000
Reference ungapped, 0.1000 0.3000
Entropy, 0.5000; Expected, -1.0000
*
COMER profile v2.0
DESC: second
FILE: second.pro
CMD:
LEN: 1
NOS: 1
EFFNOS: 1.0
SCALE: 1000
 A R N D C Q E G H I L K M F P S T W Y V
NULL""" + " 50" * 20 + """
POST""" + " 0" * 20 + """
 1000 0 0 0 0 0 0 0 0
 1000 0 0
1 C 0 0 0 0 1000""" + " 0" * 15 + """
 1000 0 0 0 0 0 0 0 0
 1000 0 0
 0 0
 SS:C 900
Expected score per pos. is non-negative, 0.1000!
*
"""


def _arr(ptr, n, dt=np.float32):
    return np.ctypeslib.as_array(ptr, (n,)).astype(dt)


def test_known_answer_file(tmp_path):
    path = tmp_path / "tiny.pro"
    path.write_text(TINY)
    S = ProfileSet(path)
    v = S.view
    assert S.nprof == 2 and S.npos == 3
    assert list(S.lengths()) == [2, 1] and list(S.dists()) == [0, 2]
    assert S.desc(0) == "tiny test profile" and S.name(0) == "tiny.pro"
    assert S.desc(1) == "second" and S.name(1) == "second.pro"
    assert list(_arr(v.eno, 2)) == [np.float32(3.5), np.float32(1.0)]
    assert np.all(_arr(v.bkg, 40) == q(50, 1000))
    tr = _arr(v.trans, 7 * 5).reshape(7, 5)      # [7][npos + nprof]
    Z = np.float32(-32768.0)
    f = np.float32(0.4)
    want = {
        0: [logf(q(900, 1000)), logf(q(50, 1000)), logf(q(50, 1000)), logf(q(600, 1000)) * f,
            logf(q(400, 1000)) * f, logf(1.0) * f, Z],
        1: [logf(q(800, 1000)), logf(q(100, 1000)), logf(q(100, 1000)), logf(q(500, 1000)) * f,
            logf(q(500, 1000)) * f, logf(q(250, 1000)) * f, logf(q(750, 1000)) * f],
        2: [np.float32(0.0), Z, Z, Z, Z, Z, Z],
        3: [np.float32(0.0), Z, Z, Z, Z, Z, Z],     # second profile: beginning row
        4: [np.float32(0.0), Z, Z, Z, Z, Z, Z],
    }
    for row, w in want.items():
        assert [np.float32(x) for x in tr[:, row]] == [np.float32(x) for x in w], row
    aa = bytes(np.ctypeslib.as_array(C.cast(v.aa, C.POINTER(C.c_uint8)), (3,)))
    assert aa == b"AWC"
    trg = _arr(v.trgfrq, 20 * 3).reshape(20, 3)
    assert trg[0, 0] == 1.0 and trg[:, 0].sum() == 1.0
    assert trg[ALPHA.index("W"), 1] == q(500, 1000) and trg[ALPHA.index("Y"), 1] == q(500, 1000)
    assert trg[ALPHA.index("C"), 2] == 1.0
    cv = _arr(v.cv, 19 * 3).reshape(19, 3)
    assert [cv[i, 0] for i in range(19)] == [q(x, 1000) for x in range(-900, 1000, 100)]
    assert np.all(cv[:, 1] == q(50, 1000))
    assert np.all(cv[:, 2] == np.float32(0.17302947))
    assert list(_arr(v.cvprior, 3)) == [q(-12345, 1000), q(-8000, 1000), 0.0]
    assert list(_arr(v.cvnorm2, 3)) == [q(4500, 1000), q(2000, 1000), 0.0]
    ss = _arr(v.ssprb, 9).reshape(3, 3)
    assert list(ss[:, 0]) == [q(100, 1000), q(800, 1000), q(100, 1000)]
    assert list(ss[:, 1]) == [0.0, 0.0, 1.0]
    assert list(ss[:, 2]) == [q(900, 1000), 0.0, 0.0]
    sss = bytes(np.ctypeslib.as_array(C.cast(v.sss, C.POINTER(C.c_uint8)), (3,)))
    assert sss == b"eh "


def test_round_trip_of_synthetic_profiles(tmp_path):
    src = SynthDB(5101, 6, 60, 0.5, 20, 200).add_cv(5102)
    path = tmp_path / "syn.pro"
    scale = 10000
    write_set(path, src.view, scale=scale)
    S = ProfileSet(path)
    a, b = src.view, S.view
    assert S.nprof == src.nprof and S.npos == src.npos
    assert list(S.lengths()) == list(src.lengths())
    ns, npr = int(a.npos), int(a.nprof)
    qz = lambda x: np.rint(scale * np.asarray(x, np.float64)).astype(np.int64)
    div = lambda i: np.float32(i) / np.float32(scale)
    for name, n in (("bkg", 20 * npr), ("trgfrq", 20 * ns), ("cvprior", ns), ("cvnorm2", ns), ("cv", 19 * ns)):
        x = _arr(getattr(a, name), n)
        assert np.array_equal(_arr(getattr(b, name), n), div(qz(x))), name
    assert np.array_equal(_arr(b.eno, npr), np.array(["%.1f" % e for e in _arr(a.eno, npr)], np.float32))
    ssa = _arr(a.ssprb, 3 * ns)
    assert np.array_equal(_arr(b.ssprb, 3 * ns), div(np.floor(scale * ssa.astype(np.float64)).astype(np.int64)))
    aa1 = np.ctypeslib.as_array(C.cast(a.aa, C.POINTER(C.c_uint8)), (ns,))
    aa2 = np.ctypeslib.as_array(C.cast(b.aa, C.POINTER(C.c_uint8)), (ns,))
    assert np.array_equal(aa1, aa2)
    sss = bytes(np.ctypeslib.as_array(C.cast(b.sss, C.POINTER(C.c_uint8)), (ns,)))
    assert sss == bytes(b" eh"[int(np.argmax(ssa.reshape(3, ns)[:, i]))] for i in range(ns))
    # transitions: the written probabilities' logs (IM..DD scaled by 0.4)
    ta = _arr(a.trans, 7 * (ns + npr)).reshape(7, ns + npr)
    tb = _arr(b.trans, 7 * (ns + npr)).reshape(7, ns + npr)
    for col in range(0, ns + npr, 7):
        for k in range(7):
            lp = ta[k, col] / 0.4 if k >= 3 else ta[k, col]
            iv = 0 if ta[k, col] <= -32768 else int(np.rint(scale * np.exp(np.float64(lp))))
            want = np.float32(-32768.0) if iv == 0 else (logf(div(iv)) * np.float32(0.4) if k >= 3 else logf(div(iv)))
            assert tb[k, col] == want, (k, col)


@pytest.mark.parametrize("edit,msg", [
    (lambda t: t.replace("COMER profile v2.0", "COMER profile v1.9", 1), "version"),
    (lambda t: t.replace("\n2 W", "\n3 W", 1), "numbering"),
    (lambda t: t.replace("\n*\nCOMER", "\nCOMER", 1), "ending"),
    (lambda t: t.replace(" CV: -8000 2000 19", " XV: -8000 2000 19", 1), "context vector"),
    (lambda t: t.replace("NULL 50", "NULL 90", 1), "background"),
])
def test_malformed_files_are_refused(tmp_path, edit, msg):
    path = tmp_path / "bad.pro"
    path.write_text(edit(TINY))
    with pytest.raises(ProfileReadError) as e:
        ProfileSet(path)
    assert "bad.pro:" in str(e.value) and msg in str(e.value), str(e.value)
