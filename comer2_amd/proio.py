"""COMER profiles read by the library: text files (cbp_proset_read_text)
and binary databases (cbp_proset_read_bin).

``ProfileSet(path, ...)`` reads one or more files of profiles in the
"COMER profile v2.0" text format (makepro's condensed form), and
``ProfileSet.from_bin(path)`` a binary database of format v2.2 (makedb's),
into a profile set the library owns; ``view`` is the ``cbp_profiles`` for
``CuBatchProcessing.TransferQueries`` / ``TransferCPMData`` / ``SearchDB``,
and ``desc(i)`` / ``name(i)`` the description and file name of profile i
(cupro/IOProfileModel.cu:46-1021 is the reader this follows).
"""
import ctypes as C
import numpy as np
from ._ffi import LIB_HIP, Profiles, load_hip  # noqa: F401

_lib = None


def _L():
    global _lib
    if _lib is None:
        L = load_hip()
        V = C.c_void_p
        L.cbp_proset_create.argtypes = [C.POINTER(V)]
        L.cbp_proset_read_text.argtypes = [V, C.c_char_p]
        L.cbp_proset_read_bin.argtypes = [V, C.c_char_p]
        L.cbp_proset_profiles.argtypes = [V]
        L.cbp_proset_profiles.restype = C.POINTER(Profiles)
        L.cbp_proset_desc.argtypes = [V, C.c_int32]
        L.cbp_proset_desc.restype = C.c_char_p
        L.cbp_proset_name.argtypes = [V, C.c_int32]
        L.cbp_proset_name.restype = C.c_char_p
        L.cbp_proset_error.argtypes = [V]
        L.cbp_proset_error.restype = C.c_char_p
        L.cbp_proset_free.argtypes = [V]
        _lib = L
    return _lib


class ProfileReadError(ValueError):
    pass


class ProfileSet:
    def __init__(self, *paths):
        self.L = _L()
        self.h = C.c_void_p()
        if self.L.cbp_proset_create(C.byref(self.h)) != 0:
            raise MemoryError("cbp_proset_create")
        for p in paths:
            self.read(p)

    def read(self, path):
        """append the profiles of one more text file"""
        if self.L.cbp_proset_read_text(self.h, str(path).encode()) != 0:
            raise ProfileReadError(self.L.cbp_proset_error(self.h).decode())
        return self

    @staticmethod
    def from_bin(path):
        """the profiles of a binary database (format v2.2)"""
        s = ProfileSet()
        if s.L.cbp_proset_read_bin(s.h, str(path).encode()) != 0:
            raise ProfileReadError(s.L.cbp_proset_error(s.h).decode())
        return s

    @property
    def view(self):
        return self.L.cbp_proset_profiles(self.h).contents

    @property
    def nprof(self):
        return self.view.nprof

    @property
    def npos(self):
        return self.view.npos

    @property
    def has_cv(self):
        return bool(self.view.cv)

    def lengths(self):
        return np.ctypeslib.as_array(self.view.len, (self.nprof,)).copy()

    def dists(self):
        return np.ctypeslib.as_array(self.view.dist, (self.nprof,)).copy()

    def desc(self, i):
        return self.L.cbp_proset_desc(self.h, i).decode()

    def name(self, i):
        return self.L.cbp_proset_name(self.h, i).decode()

    def close(self):
        if getattr(self, "h", None):
            self.L.cbp_proset_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
