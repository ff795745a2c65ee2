"""Data tables and default options of the search path.

- ``load_ssss``: the secondary-structure score table (the reference's
  ``src/env/ssss`` format: a header, ``Cardinality = 30`` and a lower triangle
  of 465 scores), expanded to a symmetric 30x30 array; indexing follows
  ``SerializedScoresSM::GetScoreP1E1`` (cupro/SerializedScoresSM.cuh:103-120).
- ``load_nn_weights``: the 93 SS18 model parameters.
- ``load_cvs2s``: the context-vector score map (``src/env/cvs2s``), read and
  smoothed by the library (``cbp_cvs2s_read``).
- ``default_options``: SSSWGT, MAPALN, MINPP, EVAL, SSEMODEL, X_UNINF, CVSWGT
  from src/env/options.txt and the ungapped/gapped statistical parameters from
  src/env/comer.par.  CVSWGT defaults to 0 (profiles without context
  vectors); SCOADJ is not an option here: the HDP score tables are absent.
"""
import ctypes as C
import os
import numpy as np

_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")


def load_ssss(path=None):
    path = path or os.path.join(_DATA, "ssss")
    vals, card = [], None
    with open(path) as fh:
        for line in fh:
            s = line.strip()
            if not s or s.startswith("#") or s.startswith("ssss") or s.startswith("0000"):
                continue
            if s.startswith("Cardinality"):
                card = int(s.split("=")[1])
                continue
            vals.extend(float(t) for t in s.split())
    if card is None or len(vals) != card * (card + 1) // 2:
        raise ValueError("malformed SS score table %s" % path)
    tab = np.zeros((card, card), dtype=np.float32)
    k = 0
    for r in range(card):
        for c in range(r + 1):
            tab[r, c] = tab[c, r] = vals[k]
            k += 1
    return tab


def load_nn_weights(path=None):
    path = path or os.path.join(_DATA, "sse_nn_weights.txt")
    vals = []
    with open(path) as fh:
        for line in fh:
            s = line.split("#", 1)[0].strip()
            if s:
                vals.extend(float(t) for t in s.split())
    if len(vals) != 93:
        raise ValueError("expected 93 SS18 parameters, got %d" % len(vals))
    return np.asarray(vals, dtype=np.float32)


def default_options():
    """options.txt defaults, except CVSWGT (0.15 there): 0 here, the setting
    for profiles without context vectors"""
    return dict(ssswgt=0.12, mapaln=1, minpp=0.28, evalue=10.0, ssemodel=2, xuninf=1,
                ung_lambda=1.116276, ung_K=0.115812, ung_H=0.218508,
                gap_lambda=0.561132, gap_K=0.001488, gap_alpha=0.5, gap_beta=-1.0,
                max_map_pairs=0, cvswgt=0.0)


def load_cvs2s(path=None):
    """The context-vector map (src/env/cvs2s) as the library reads it
    (cbp_cvs2s_read): (keys, smoothed scores, shift).  Host only."""
    path = path or os.path.join(_DATA, "cvs2s")
    from ._ffi import load_hip
    L = load_hip()
    L.cbp_cvs2s_read.argtypes = [C.c_char_p, C.c_void_p, C.c_void_p, C.c_int32,
                                 C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
    card, shift = C.c_int32(), C.c_int32()
    if L.cbp_cvs2s_read(path.encode(), None, None, 0, C.byref(card), C.byref(shift)) != 0:
        raise ValueError("cannot read the context-vector map " + path)
    keys = np.zeros(card.value, np.float32)
    vals = np.zeros(card.value, np.float32)
    if L.cbp_cvs2s_read(path.encode(), keys.ctypes.data_as(C.c_void_p), vals.ctypes.data_as(C.c_void_p),
                        card.value, C.byref(card), C.byref(shift)) != 0:
        raise ValueError("cannot read the context-vector map " + path)
    return keys, vals, shift.value
