"""ctypes view of the C ABI declared in include/comer2_hip.h."""
import ctypes as C
import os
import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_HIP = os.path.join(_HERE, "libcomer2_hip.so")
LIB_SYNTH = os.path.join(_HERE, "libc2synth.so")


def load_hip():
    """libcomer2_hip.so, opened after torch when torch is importable.
    PyTorch-ROCm ships its own HIP runtime under the soname of
    /opt/rocm's (libamdhip64.so.7): with torch loaded first the library binds
    to that runtime, so the process has one HIP runtime.  Opened before
    torch, it binds to /opt/rocm's and a later torch brings a second one, and
    then calls of either see spurious errors (torch finds no GPU, kernel
    launches report an unknown error).  Every loader of the library goes
    through here."""
    if not os.path.exists(LIB_HIP):
        raise OSError("libcomer2_hip.so is not built (run comer2_amd.build)")
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    return C.CDLL(LIB_HIP)


class Profiles(C.Structure):
    _fields_ = [("nprof", C.c_int32), ("npos", C.c_int64),
                ("bkg", C.POINTER(C.c_float)), ("eno", C.POINTER(C.c_float)),
                ("len", C.POINTER(C.c_int32)), ("dist", C.POINTER(C.c_uint32)),
                ("trgfrq", C.POINTER(C.c_float)), ("ssprb", C.POINTER(C.c_float)),
                ("aa", C.POINTER(C.c_char)), ("sss", C.POINTER(C.c_char)),
                ("trans", C.POINTER(C.c_float)), ("cv", C.POINTER(C.c_float)),
                ("cvprior", C.POINTER(C.c_float)), ("cvnorm2", C.POINTER(C.c_float))]


class Options(C.Structure):
    _fields_ = [("ssswgt", C.c_float), ("mapaln", C.c_int32), ("minpp", C.c_float),
                ("evalue", C.c_float), ("ssemodel", C.c_int32), ("xuninf", C.c_int32),
                ("ung_lambda", C.c_float), ("ung_K", C.c_float), ("ung_H", C.c_float),
                ("gap_lambda", C.c_float), ("gap_K", C.c_float), ("gap_alpha", C.c_float),
                ("gap_beta", C.c_float), ("max_map_pairs", C.c_int32), ("cvswgt", C.c_float)]


class Results(C.Structure):
    _fields_ = [("npassed", C.c_int32), ("nposits", C.c_int64), ("maxlen", C.c_int32),
                ("nmap", C.c_int32), ("alndata", C.POINTER(C.c_float)),
                ("order", C.POINTER(C.c_int32)), ("alns", C.POINTER(C.c_char)),
                ("alnstride", C.c_int64), ("map_rows", C.POINTER(C.c_int32)),
                ("epa", C.POINTER(C.c_float)), ("scorethld", C.c_float), ("logevthld", C.c_float),
                ("search_space", C.c_double)]


class SdpPair(C.Structure):
    _fields_ = [("qidx", C.c_int32), ("pronr", C.c_int32), ("soff", C.c_int64)]


NFIELDS = 18
(F_SCORE, F_ORGPRONO, F_PRONEWDST, F_PSTVS, F_BEGCOORDS, F_ENDCOORDS,
 F_LMBDEST, F_KEST, F_EVALUE, F_EPA, F_MIN, F_MAX, F_E, F_H, F_LMBD, F_K,
 F_IDNTS, F_NGAPS) = range(18)


def options_struct(d):
    o = Options()
    for k, _ in Options._fields_:
        setattr(o, k, d.get(k, 0.0) if k == "cvswgt" else d[k])
    return o


def load_synth():
    lib = C.CDLL(LIB_SYNTH)
    lib.c2s_generate.restype = C.c_void_p
    lib.c2s_generate.argtypes = [C.c_uint64, C.c_int32, C.c_double, C.c_double,
                                 C.c_int32, C.c_int32, C.c_int32, C.c_double]
    lib.c2s_view.restype = C.POINTER(Profiles)
    lib.c2s_view.argtypes = [C.c_void_p]
    lib.c2s_generate_lengths.restype = C.c_void_p
    lib.c2s_generate_lengths.argtypes = [C.c_uint64, C.c_int32, C.c_void_p, C.c_double]
    lib.c2s_orgidx.restype = C.POINTER(C.c_int32)
    lib.c2s_orgidx.argtypes = [C.c_void_p]
    lib.c2s_free.argtypes = [C.c_void_p]
    lib.c2s_subset.restype = C.c_void_p
    lib.c2s_subset.argtypes = [C.c_void_p, C.c_int32, C.c_int32]
    lib.c2s_plant.restype = C.c_int32
    lib.c2s_plant.argtypes = [C.c_void_p, C.POINTER(Profiles), C.c_uint64, C.c_int32,
                              C.c_double, C.c_double, C.c_void_p, C.c_void_p]
    lib.c2s_add_cv.restype = None
    lib.c2s_add_cv.argtypes = [C.c_void_p, C.c_uint64]
    lib.c2s_plant_cv.restype = C.c_int32
    lib.c2s_plant_cv.argtypes = [C.c_void_p, C.POINTER(Profiles), C.c_uint64, C.c_int32]
    return lib


class SynthDB:
    """Owned synthetic profile set (see c2_synth.cpp for the distributions)."""
    _lib = None

    def __init__(self, seed, nprof, median, sigma=0.668, minlen=30, maxlen=800,
                 sort_desc=True, conc=20.0, _handle=None):
        if SynthDB._lib is None:
            SynthDB._lib = load_synth()
        self.h = _handle if _handle is not None else SynthDB._lib.c2s_generate(
            seed, nprof, median, sigma, minlen, maxlen, 1 if sort_desc else 0, conc)
        self.view = SynthDB._lib.c2s_view(self.h).contents

    def plant(self, src, seed, nplant, kmin=4.0, kmax=60.0):
        """Plant homologous segments of the profiles of `src` (see c2s_plant);
        with context vectors on both sets, the planted windows take the
        source's vectors (c2s_plant_cv)."""
        ps = np.zeros(nplant, np.int32)
        pp = np.zeros(nplant, np.int32)
        n = SynthDB._lib.c2s_plant(self.h, C.byref(src.view), seed, nplant, kmin, kmax,
                                   ps.ctypes.data_as(C.c_void_p), pp.ctypes.data_as(C.c_void_p))
        if self.has_cv and src.has_cv:
            SynthDB._lib.c2s_plant_cv(self.h, C.byref(src.view), seed, nplant)
        return ps[:n], pp[:n]

    def add_cv(self, seed):
        """Context vectors for every position (see c2s_add_cv)."""
        SynthDB._lib.c2s_add_cv(self.h, seed)
        self.view = SynthDB._lib.c2s_view(self.h).contents
        return self

    @property
    def has_cv(self):
        return bool(self.view.cv)

    @staticmethod
    def from_lengths(seed, lengths, conc=20.0):
        """Profiles of exactly these lengths, in this order."""
        if SynthDB._lib is None:
            SynthDB._lib = load_synth()
        ln = np.ascontiguousarray(lengths, np.int32)
        return SynthDB(0, 0, 0, _handle=SynthDB._lib.c2s_generate_lengths(
            seed, ln.size, ln.ctypes.data_as(C.c_void_p), conc))

    def subset(self, p0, n):
        return SynthDB(0, 0, 0, _handle=SynthDB._lib.c2s_subset(self.h, p0, n))

    @property
    def nprof(self):
        return self.view.nprof

    @property
    def npos(self):
        return self.view.npos

    def lengths(self):
        return np.ctypeslib.as_array(self.view.len, (self.nprof,)).copy()

    def dists(self):
        return np.ctypeslib.as_array(self.view.dist, (self.nprof,)).copy()

    def __del__(self):
        # at interpreter exit the class attribute may already be torn down
        lib = getattr(type(self), "_lib", None) if type(self) is not None else None
        if getattr(self, "h", None) and lib is not None:
            lib.c2s_free(self.h)
            self.h = None
