"""The binary profile database reader (cbp_proset_read_bin, format v2.2).

* Known answer: a database of two profiles (lengths 2 and 1) packed here
  byte by byte after the writer's layout (Db::MakeBin_v2_2,
  CalculateAndWriteAddressTable_v2_2, BinaryWriteProfile_v2_2): every field
  holds distinct values that must land in the profile set's arrays, the
  distances follow from the lengths, the descriptions from their end
  addresses.
* Round trip: profiles read from a text file, written as a binary database
  (tests/profile_bin.py) and read back, equal the text-read set bit for bit.
* Damaged files (magic, table, truncation, lengths) are refused.
"""
import ctypes as C
import struct
import numpy as np
import pytest
from comer2_amd._ffi import SynthDB
from comer2_amd.proio import ProfileSet, ProfileReadError
from profile_bin import MAGIC, write_bin
from profile_text import write_set


def _tiny_bytes():
    np_, ns = 2, 3
    f = lambda xs: struct.pack("<%df" % len(xs), *xs)
    secs = [struct.pack("<2i", 2, 1), f([3.5, 1.25])]
    secs += [f([0.01 * (a + 1), 0.02 * (a + 1)]) for a in range(20)]                 # bkg[a][p]
    secs += [f([-1.0 * k - 0.1 * c for c in range(ns + np_)]) for k in range(7)]    # trans[k][col]
    secs += [f([0.001 * a + 0.1 * m for m in range(ns)]) for a in range(20)]          # trg[a][pos]
    secs += [f([0.5 * i - m for m in range(ns)]) for i in range(19)]                  # cv[i][pos]
    secs += [f([-12.5, -8.25, -3.0]), f([4.5, 2.0, 0.0])]                             # prior, norm2
    secs += [f([0.1, 0.2, 0.3]), f([0.8, 0.3, 0.3]), f([0.1, 0.5, 0.4])]              # ss[j][pos]
    secs += [f([0.0] * ns), struct.pack("<3i", -1, -1, -1), b"AWC", b" eh"]
    head = MAGIC + struct.pack("<QQ", np_, ns)
    addr = len(head) + 79 * 8
    table = []
    for s in secs:
        table.append(addr)
        addr += len(s)
    table += [addr, addr + 8 * np_]
    descs = [b"first profile\0", b"second one\0"]
    ends = [addr + 8 * np_ + len(descs[0]), addr + 8 * np_ + len(descs[0]) + len(descs[1])]
    return head + struct.pack("<79Q", *table) + b"".join(secs) + struct.pack("<2Q", *ends) + b"".join(descs)


def _a(ptr, n, dt=np.float32):
    return np.ctypeslib.as_array(ptr, (n,)).astype(dt)


def test_known_answer_database(tmp_path):
    p = tmp_path / "tiny.bin"
    p.write_bytes(_tiny_bytes())
    S = ProfileSet.from_bin(p)
    v = S.view
    assert S.nprof == 2 and S.npos == 3 and S.has_cv
    assert list(S.lengths()) == [2, 1] and list(S.dists()) == [0, 2]
    assert list(_a(v.eno, 2)) == [3.5, 1.25]
    bkg = _a(v.bkg, 40).reshape(20, 2)
    assert np.array_equal(bkg, np.float32([[0.01 * (a + 1), 0.02 * (a + 1)] for a in range(20)]))
    tr = _a(v.trans, 35).reshape(7, 5)
    assert np.array_equal(tr, np.float32([[-1.0 * k - 0.1 * c for c in range(5)] for k in range(7)]))
    trg = _a(v.trgfrq, 60).reshape(20, 3)
    assert np.array_equal(trg, np.float32([[0.001 * a + 0.1 * m for m in range(3)] for a in range(20)]))
    cv = _a(v.cv, 57).reshape(19, 3)
    assert np.array_equal(cv, np.float32([[0.5 * i - m for m in range(3)] for i in range(19)]))
    assert list(_a(v.cvprior, 3)) == [-12.5, -8.25, -3.0] and list(_a(v.cvnorm2, 3)) == [4.5, 2.0, 0.0]
    assert np.array_equal(_a(v.ssprb, 9).reshape(3, 3), np.float32([[0.1, 0.2, 0.3], [0.8, 0.3, 0.3], [0.1, 0.5, 0.4]]))
    assert bytes(np.ctypeslib.as_array(C.cast(v.aa, C.POINTER(C.c_uint8)), (3,))) == b"AWC"
    assert bytes(np.ctypeslib.as_array(C.cast(v.sss, C.POINTER(C.c_uint8)), (3,))) == b" eh"
    assert S.desc(0) == "first profile" and S.desc(1) == "second one"


def test_round_trip_text_to_binary(tmp_path):
    src = SynthDB(6101, 9, 80, 0.5, 20, 300).add_cv(6102)
    write_set(tmp_path / "db.pro", src.view)
    T = ProfileSet(tmp_path / "db.pro")
    write_bin(tmp_path / "db.bin", T.view, [T.desc(i) for i in range(T.nprof)])
    B = ProfileSet.from_bin(tmp_path / "db.bin")
    a, b = T.view, B.view
    ns, npr = int(a.npos), int(a.nprof)
    assert B.nprof == npr and B.npos == ns
    for name, n in (("bkg", 20 * npr), ("eno", npr), ("trgfrq", 20 * ns), ("ssprb", 3 * ns),
                    ("trans", 7 * (ns + npr)), ("cv", 19 * ns), ("cvprior", ns), ("cvnorm2", ns)):
        assert np.array_equal(_a(getattr(a, name), n).view(np.uint32), _a(getattr(b, name), n).view(np.uint32)), name
    for name in ("aa", "sss"):
        x = np.ctypeslib.as_array(C.cast(getattr(a, name), C.POINTER(C.c_uint8)), (ns,))
        y = np.ctypeslib.as_array(C.cast(getattr(b, name), C.POINTER(C.c_uint8)), (ns,))
        assert np.array_equal(x, y), name
    assert [B.desc(i) for i in range(npr)] == [T.desc(i) for i in range(npr)]
    assert list(B.dists()) == list(T.dists())


@pytest.mark.parametrize("damage", ["magic", "table", "truncated", "lengths"])
def test_damaged_databases_are_refused(tmp_path, damage):
    b = bytearray(_tiny_bytes())
    if damage == "magic":
        b[28:31] = b"2.1"
    elif damage == "table":
        off = len(MAGIC) + 16 + 8 * 5
        b[off:off + 8] = struct.pack("<Q", struct.unpack("<Q", bytes(b[off:off + 8]))[0] + 4)
    elif damage == "truncated":
        b = b[:len(b) - 40]
    else:
        t0 = struct.unpack("<Q", bytes(b[len(MAGIC) + 16:len(MAGIC) + 24]))[0]
        b[t0:t0 + 4] = struct.pack("<i", 5)
    p = tmp_path / "bad.bin"
    p.write_bytes(bytes(b))
    with pytest.raises(ProfileReadError) as e:
        ProfileSet.from_bin(p)
    assert "bad.bin" in str(e.value)
