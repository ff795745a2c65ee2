"""Shared helpers for GPU parity tests."""
import numpy as np
import pytest
from comer2_amd import tables
from comer2_amd._ffi import SynthDB
from oracle_ffi import Oracle


def oracle(opts=None):
    return Oracle(tables.load_ssss(), tables.load_nn_weights(), dict(tables.default_options(), **(opts or {})))


def expected_keys(orc, q, qp, d):
    from comer2_amd.host import encode_key
    keys = np.zeros(d.nprof, np.uint64)
    for dp in range(d.nprof):
        S, _ = orc.score_f32(q.view, qp, d.view, dp)
        m, _, _ = orc.viterbi_f32(q.view, qp, d.view, dp, S)
        keys[dp] = encode_key(m.maxsc, m.x, m.y)
    return keys


def processing(opts=None):
    from comer2_amd.host import CuBatchProcessing
    return CuBatchProcessing(0, 0, 0, options=opts)
