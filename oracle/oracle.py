"""ctypes front end of the CPU oracle (oracle/sir_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the sagesim_amd package.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build():
    subprocess.check_call(["make", "-s", "-B", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        u32, u64, i64, i32, dbl = (ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int64,
                                   ctypes.c_int32, ctypes.c_double)
        P = ctypes.c_void_p
        L.or_philox2x32_ref.argtypes = [u32, u32, u32]
        L.or_philox2x32_ref.restype = u32
        L.or_rand_uniform_philox.argtypes = [i64, i64, i64, i64]
        L.or_rand_uniform_philox.restype = dbl
        L.or_philox4x32_10.argtypes = [P, P, P]
        L.or_philox4x32_10.restype = None
        L.or_sir_word.argtypes = [u64, i64, i64]
        L.or_sir_word.restype = u32
        L.or_infection_probs.argtypes = [dbl, i64, P]
        L.or_infection_probs.restype = None
        L.or_sir_run.argtypes = [P, i64, i64, P, P, P, P, P, i64, i64, P, P, P, i64, P, ctypes.c_int]
        L.or_sir_run.restype = ctypes.c_int
        L.or_sir_run_fast.argtypes = L.or_sir_run.argtypes
        L.or_sir_run_fast.restype = ctypes.c_int
        L.or_ref_sir_run.argtypes = [i64, P, P, P, P, i32, i64, i64, dbl, dbl]
        L.or_ref_sir_run.restype = None
        _lib = L
    return _lib


class OrModel(ctypes.Structure):
    _fields_ = [("n_states", ctypes.c_int32), ("s_code", ctypes.c_int32),
                ("inf_code", ctypes.c_int32), ("s_target", ctypes.c_int32),
                ("beta", ctypes.c_double), ("trans_p", ctypes.c_double * 4),
                ("trans_next", ctypes.c_int32 * 4), ("seed", ctypes.c_uint64)]


def sir_model(beta, gamma, seed):
    """SIR {S=0, I=1, R=2}."""
    m = OrModel()
    m.n_states, m.s_code, m.inf_code, m.s_target = 3, 0, 1, 1
    m.beta = beta
    m.trans_p[:] = [0.0, gamma, 0.0, 0.0]
    m.trans_next[:] = [0, 2, 2, 3]
    m.seed = seed
    return m


def seir_model(beta, sigma, gamma, seed):
    """SEIR {S=0, E=1, I=2, R=3}."""
    m = OrModel()
    m.n_states, m.s_code, m.inf_code, m.s_target = 4, 0, 2, 1
    m.beta = beta
    m.trans_p[:] = [0.0, sigma, gamma, 0.0]
    m.trans_next[:] = [0, 2, 3, 3]
    m.seed = seed
    return m


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def sir_run(model, row_ptr, col_idx, rng_id, state, ticks, tick0=0, n_local=None,
            history=None, ghost_ids=None, halo_ids=None, halo_vals=None,
            trace=False, nthreads=0, fast=False):
    """Run the oracle; returns (final_state, trace or None).  `state` is not modified.
    fast=True runs or_sir_run_fast, the tuned restatement with the same results."""
    row_ptr = np.ascontiguousarray(row_ptr, dtype=np.int64)
    col_idx = np.ascontiguousarray(col_idx, dtype=np.int32)
    rng_id = np.ascontiguousarray(rng_id, dtype=np.int64)
    st = np.array(state, dtype=np.int8, copy=True)
    n_total = st.shape[0]
    if n_local is None:
        n_local = n_total
    n_halo = 0
    if halo_ids is not None:
        halo_ids = np.ascontiguousarray(halo_ids, dtype=np.int64)
        halo_vals = np.ascontiguousarray(halo_vals, dtype=np.int8)
        ghost_ids = np.ascontiguousarray(ghost_ids, dtype=np.int64)
        n_halo = halo_ids.shape[0]
        assert halo_vals.shape == (ticks, n_halo)
    tr = np.empty((ticks, n_total), dtype=np.int8) if trace else None
    if history is not None:
        assert history.dtype == np.int32 and history.flags.c_contiguous
    fn = lib().or_sir_run_fast if fast else lib().or_sir_run
    rc = fn(ctypes.byref(model), n_local, n_total, _p(row_ptr), _p(col_idx),
            _p(rng_id), _p(st), _p(history), tick0, ticks,
            _p(ghost_ids), _p(halo_ids), _p(halo_vals), n_halo, _p(tr), nthreads)
    if rc != 0:
        raise ValueError("oracle: halo id names no ghost")
    return st, tr


def ref_sir_run(row_ptr, col_idx, logical_id_f32, state_f32, seed, ticks, p_inf, p_rec, tick0=0):
    row_ptr = np.ascontiguousarray(row_ptr, dtype=np.int64)
    col_idx = np.ascontiguousarray(col_idx, dtype=np.int32)
    lid = np.ascontiguousarray(logical_id_f32, dtype=np.float32)
    st = np.array(state_f32, dtype=np.float32, copy=True)
    lib().or_ref_sir_run(st.shape[0], _p(row_ptr), _p(col_idx), _p(lid), _p(st),
                         int(seed), tick0, ticks, p_inf, p_rec)
    return st


def philox4x32_10(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    lib().or_philox4x32_10(_p(c), _p(k), _p(o))
    return o


def infection_probs(beta, kmax):
    out = np.zeros(kmax + 1, dtype=np.float64)
    lib().or_infection_probs(beta, kmax, _p(out))
    return out
