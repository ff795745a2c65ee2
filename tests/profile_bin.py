"""Test helper: profile sets written as COMER binary databases of format
v2.2, the layout of Db::MakeBin_v2_2 (libpro/srcpro/Db.cpp:333-383),
CalculateAndWriteAddressTable_v2_2 and BinaryWriteProfile_v2_2
(libpro/srcpro/PMProfileModel.cpp:1468-1960): the magic and version, the
number of profiles and the database size (8 bytes each), a table of 79
section addresses (8 bytes each), then the sections in the table's order,
each over all profiles; HDP probabilities 0 and indices -1 as placeholders.
"""
import ctypes as C
import struct
import numpy as np

MAGIC = b"COMER profile bin database v2.2"


def _arr(ptr, n, dt):
    return np.ctypeslib.as_array(ptr, (n,)).astype(dt, copy=True)


def write_bin(path, view, descs):
    np_, ns = int(view.nprof), int(view.npos)
    cv = _arr(view.cv, 19 * ns, np.float32) if view.cv else np.full(19 * ns, 0.17302947, np.float32)
    cvp = _arr(view.cvprior, ns, np.float32) if view.cv else np.zeros(ns, np.float32)
    cvn = _arr(view.cvnorm2, ns, np.float32) if view.cv else np.zeros(ns, np.float32)
    aa = bytes(np.ctypeslib.as_array(C.cast(view.aa, C.POINTER(C.c_uint8)), (ns,)))
    sss = bytes(np.ctypeslib.as_array(C.cast(view.sss, C.POINTER(C.c_uint8)), (ns,)))
    secs = [_arr(view.len, np_, np.int32).tobytes(), _arr(view.eno, np_, np.float32).tobytes()]
    bkg = _arr(view.bkg, 20 * np_, np.float32).reshape(20, np_)
    secs += [bkg[a].tobytes() for a in range(20)]
    tr = _arr(view.trans, 7 * (ns + np_), np.float32).reshape(7, ns + np_)
    secs += [tr[k].tobytes() for k in range(7)]
    trg = _arr(view.trgfrq, 20 * ns, np.float32).reshape(20, ns)
    secs += [trg[a].tobytes() for a in range(20)]
    cv = cv.reshape(19, ns)
    secs += [cv[i].tobytes() for i in range(19)]
    secs += [cvp.tobytes(), cvn.tobytes()]
    ss = _arr(view.ssprb, 3 * ns, np.float32).reshape(3, ns)
    secs += [ss[j].tobytes() for j in range(3)]
    secs += [np.zeros(ns, np.float32).tobytes(), np.full(ns, -1, np.int32).tobytes(), aa, sss]
    head = MAGIC + struct.pack("<QQ", np_, ns)
    table_at = len(head)
    addr = table_at + 79 * 8
    table = []
    for s in secs:
        table.append(addr)
        addr += len(s)
    table.append(addr)                     # end addresses of the descriptions
    desc_at = addr + 8 * np_
    table.append(desc_at)                  # the descriptions
    ends, text, a = [], b"", desc_at
    for d in descs:
        b = d.encode() + b"\0"
        a += len(b)
        ends.append(a)
        text += b
    with open(path, "wb") as fh:
        fh.write(head)
        fh.write(struct.pack("<79Q", *table))
        for s in secs:
            fh.write(s)
        fh.write(struct.pack("<%dQ" % np_, *ends))
        fh.write(text)
