"""ctypes bindings of libsagestep (include/sagestep.h) and libsagegraph.

The product path loads the in-tree libraries built by sagesim_amd/build.py and
fails loudly when they are missing: there is no Python or CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
STEP_LIB_PATH = os.path.join(HERE, "libsagestep.so")
GRAPH_LIB_PATH = os.path.join(HERE, "libsagegraph.so")

SS_OK = 0
SS_ERR = {-1: "SS_ERR_INVALID", -2: "SS_ERR_HIP", -3: "SS_ERR_OOM", -4: "SS_ERR_RANGE",
          -5: "SS_ERR_STATE", -6: "SS_ERR_RTC", -7: "SS_ERR_FULL"}
HISTORY_WIDTH = 32

_c_i32, _c_i64, _c_u64, _c_dbl = ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64, ctypes.c_double
_P = ctypes.c_void_p


class SSModel(ctypes.Structure):
    _fields_ = [("n_states", _c_i32), ("s_code", _c_i32), ("inf_code", _c_i32),
                ("s_target", _c_i32), ("beta", _c_dbl), ("trans_p", _c_dbl * 4),
                ("trans_next", _c_i32 * 4)]


class SSReplicaSpec(ctypes.Structure):
    _fields_ = [("n_local", _c_i64), ("n_ghost", _c_i64), ("nnz", _c_i64),
                ("row_ptr", _P), ("col_idx", _P), ("rng_key_id", _P), ("state", _P),
                ("history", _P), ("model", SSModel), ("seed", _c_u64), ("tick0", _c_i64),
                ("tick_limit", _c_i64)]


class NativeError(RuntimeError):
    """Error reported across the C ABI (code + message)."""

    def __init__(self, code, msg):
        super().__init__(f"{SS_ERR.get(code, code)}: {msg}")
        self.code = code


class LibraryMissing(RuntimeError):
    pass


_step = None
_graph = None


def _proto(lib, name, res, *args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = list(args)


def step_lib():
    """libsagestep.so; raises LibraryMissing when it has not been built."""
    global _step
    if _step is not None:
        return _step
    if not os.path.exists(STEP_LIB_PATH):
        raise LibraryMissing(f"{STEP_LIB_PATH} is not built; run `python -m sagesim_amd.build`")
    L = ctypes.CDLL(STEP_LIB_PATH)
    I, PP = ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)
    _proto(L, "ss_create", I, I, PP)
    _proto(L, "ss_destroy", I, _P)
    _proto(L, "ss_last_error", ctypes.c_char_p, _P)
    _proto(L, "ss_sync", I, _P)
    _proto(L, "ss_device_info", I, _P, _P, _P, _P)
    _proto(L, "ss_replica_create", I, _P, ctypes.POINTER(SSReplicaSpec), _P)
    _proto(L, "ss_replica_destroy", I, _P, _c_i32)
    _proto(L, "ss_replica_footprint", _c_i64, _c_i64, _c_i64, _c_i64, _c_i32, _c_i64)
    _proto(L, "ss_replica_reset", I, _P, _c_i32)
    _proto(L, "ss_replica_tick", I, _P, _c_i32, _P)
    _proto(L, "ss_replica_set_tick", I, _P, _c_i32, _c_i64, _c_i64)
    _proto(L, "ss_replica_bytes", I, _P, _c_i32, _P)
    _proto(L, "ss_set_model", I, _P, _c_i32, ctypes.POINTER(SSModel))
    _proto(L, "ss_set_seed", I, _P, _c_i32, _c_u64)
    _proto(L, "ss_state_download", I, _P, _c_i32, _P, _c_i64)
    _proto(L, "ss_state_upload", I, _P, _c_i32, _P, _c_i64)
    _proto(L, "ss_history_download", I, _P, _c_i32, _P, _c_i64)
    _proto(L, "ss_replica_save", I, _P, _c_i32, _P, _c_i64, _P)
    _proto(L, "ss_replica_restore", I, _P, _c_i32, _P, _c_i64, _P)
    _proto(L, "ss_simulate", I, _P, _c_i64)
    _proto(L, "ss_set_tiling", I, _P, _c_i32, _c_i32, _c_i32, _c_i32, _c_i32)
    _proto(L, "ss_launch_shape", I, _P, _P, _P)
    _proto(L, "ss_map_build", I, _P, _c_i32, _P, _P, _c_i64, _c_dbl)
    _proto(L, "ss_map_insert", I, _P, _c_i32, _P, _P, _c_i64)
    _proto(L, "ss_map_remove", I, _P, _c_i32, _P, _c_i64, _P)
    _proto(L, "ss_map_lookup", I, _P, _c_i32, _P, _P, _c_i64, _P)
    _proto(L, "ss_map_resize", I, _P, _c_i32, _c_i64)
    _proto(L, "ss_map_info", I, _P, _c_i32, _P, _P)
    _proto(L, "ss_map_download", I, _P, _c_i32, _P, _P, _c_i64)
    _proto(L, "ss_halo_bind", I, _P, _c_i32, _P, _c_i64, _P)
    _proto(L, "ss_ghost_discover", I, _P, _P, _c_i64, _P, _c_i64, _c_i32, _c_i32, _P, _P, _P, _c_i64, _P)
    _proto(L, "ss_halo_unpack", I, _P, _c_i32, _P, _c_i64)
    _proto(L, "ss_halo_unpack_ids", I, _P, _c_i32, _P, _P, _c_i64, _P)
    _proto(L, "ss_halo_pack", I, _P, _c_i32, _P, _c_i64, _P)
    _proto(L, "ss_halo_pack_noalloc", I, _P, _c_i32, _P, _c_i64, _P)
    _proto(L, "ss_simulate_halo", I, _P, _c_i32, _c_i64, _P)
    _proto(L, "ss_halo_unpack_2bit", I, _P, _c_i32, _P, _c_i64)
    _proto(L, "ss_halo_pack_2bit", I, _P, _c_i32, _P, _c_i64, _P)
    _proto(L, "ss_simulate_halo_2bit", I, _P, _c_i32, _c_i64, _P)
    _proto(L, "ss_host_alloc", I, _P, _c_i64, _P)
    _proto(L, "ss_host_free", I, _P, _P)
    _proto(L, "ss_set_counting", I, _P, _c_i32)
    _proto(L, "ss_read_counters", I, _P, _P)
    _proto(L, "ss_census", I, _P, _c_i32, _P)
    _proto(L, "ss_copy_probe", I, _P, _c_i64, _c_i32, ctypes.POINTER(ctypes.c_float))
    _proto(L, "ss_gather_probe", I, _P, _c_i64, _c_i64, _c_i32, ctypes.POINTER(ctypes.c_float))
    _proto(L, "ss_set_push", I, _P, ctypes.c_double)
    _proto(L, "ss_replica_push_ok", I, _P, _c_i32, _P)
    _proto(L, "ss_prog_create", I, _P, _c_i64, PP)
    _proto(L, "ss_prog_destroy", I, _P)
    _proto(L, "ss_prog_last_error", ctypes.c_char_p, _P)
    _proto(L, "ss_prog_set_csr", I, _P, _P, _P, _c_i64)
    _proto(L, "ss_prog_set_ids", I, _P, _P, _P, _P, _P)
    _proto(L, "ss_prog_add_prop", I, _P, _c_i32, _c_i64, _c_i32, _P, _P)
    _proto(L, "ss_prog_add_array", I, _P, _c_i32, _c_i64, _c_i64, _c_i32, _P, _P)
    _proto(L, "ss_prog_set_extra", I, _P, _c_i32, _c_i64, _P)
    _proto(L, "ss_prog_get_extra", I, _P, _c_i32, _c_i64, _P)
    _proto(L, "ss_prog_cache_hit", I, _P, _P)
    _proto(L, "ss_prog_set_order", I, _P, _P, _c_i64)
    _proto(L, "ss_prog_save", I, _P, _c_i64, _P, _c_i64, _P)
    _proto(L, "ss_prog_restore", I, _P, _P, _c_i64, _P)
    _proto(L, "ss_prog_prop_upload", I, _P, _c_i32, _P)
    _proto(L, "ss_prog_prop_download", I, _P, _c_i32, _P)
    _proto(L, "ss_prog_set_globals", I, _P, _P, _c_i64)
    _proto(L, "ss_prog_set_seed", I, _P, _c_i64)
    _proto(L, "ss_prog_compile", I, _P, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, _c_i64)
    _proto(L, "ss_prog_set_schedule", I, _P, _c_i32, _P, _P, _P)
    _proto(L, "ss_prog_simulate", I, _P, _c_i64, _c_i64)
    _proto(L, "ss_timer_start", I, _P)
    _proto(L, "ss_timer_stop", I, _P, ctypes.POINTER(ctypes.c_float))
    _step = L
    return L


def graph_lib():
    global _graph
    if _graph is not None:
        return _graph
    if not os.path.exists(GRAPH_LIB_PATH):
        raise LibraryMissing(f"{GRAPH_LIB_PATH} is not built; run `python -m sagesim_amd.build graph`")
    L = ctypes.CDLL(GRAPH_LIB_PATH)
    G = _P
    _proto(L, "sg_last_error", ctypes.c_char_p)
    for name in ("sg_n", "sg_n_cols", "sg_nnz", "sg_n_extra"):
        _proto(L, name, _c_i64, G)
    _proto(L, "sg_row_ptr", _P, G)
    _proto(L, "sg_col_idx", _P, G)
    _proto(L, "sg_extra", _P, G)
    _proto(L, "sg_free", None, G)
    _proto(L, "sg_from_edges", G, _c_i64, _P, _P, _c_i64, ctypes.c_int)
    _proto(L, "sg_er", G, _c_i64, _c_dbl, _c_u64)
    _proto(L, "sg_ba", G, _c_i64, _c_i32, _c_u64)
    _proto(L, "sg_ws", G, _c_i64, _c_i32, _c_dbl, _c_u64)
    _proto(L, "sg_sbm_part", G, _c_i32, _c_i64, _c_dbl, _c_dbl, _c_i32, _c_i32, _c_u64)
    _proto(L, "sg_degree_sort_csr", ctypes.c_int, _c_i64, _P, _P, _P, _P, _P, _P)
    _proto(L, "sg_scramble_ids", None, _P, _P, _c_i64, _c_u64)
    _proto(L, "sg_set_num_threads", None, ctypes.c_int)
    _graph = L
    return L


def ptr(a):
    """Data pointer of a C-contiguous numpy array (or None)."""
    if a is None:
        return None
    assert a.flags.c_contiguous, "array must be C-contiguous"
    return a.ctypes.data_as(ctypes.c_void_p)
