"""Seeded graph generators and CSR helpers (libsagegraph).

The reference builds CSR from ragged Python neighbour containers
(internal_utils.py:111-186) fed one `connect_agents` call at a time
(space.py:107-125).  These helpers produce the same CSR shape (offsets, values)
directly: int64 row_ptr, int32 col_idx, rows sorted ascending, duplicate edges
and self loops removed.
"""
from dataclasses import dataclass, field
import ctypes

import numpy as np

from ._lib import graph_lib, ptr


@dataclass
class CSRGraph:
    row_ptr: np.ndarray            # int64 [n + 1]
    col_idx: np.ndarray            # int32 [nnz]
    n_cols: int = -1               # column id space (> n when ghost columns exist)
    extra: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))

    def __post_init__(self):
        if self.n_cols < 0:
            self.n_cols = self.n

    @property
    def n(self):
        return self.row_ptr.shape[0] - 1

    @property
    def nnz(self):
        return int(self.col_idx.shape[0])

    def degrees(self):
        return np.diff(self.row_ptr)


def _take(h):
    L = graph_lib()
    if not h:
        raise ValueError(L.sg_last_error().decode())
    try:
        n, nnz, ne = L.sg_n(h), L.sg_nnz(h), L.sg_n_extra(h)
        rp = np.ctypeslib.as_array(ctypes.cast(L.sg_row_ptr(h), ctypes.POINTER(ctypes.c_int64)), (n + 1,)).copy()
        if nnz:
            ci = np.ctypeslib.as_array(ctypes.cast(L.sg_col_idx(h), ctypes.POINTER(ctypes.c_int32)), (nnz,)).copy()
        else:
            ci = np.zeros(0, np.int32)
        ex = (np.ctypeslib.as_array(ctypes.cast(L.sg_extra(h), ctypes.POINTER(ctypes.c_int64)), (ne,)).copy()
              if ne else np.zeros(0, np.int64))
        return CSRGraph(rp, ci, int(L.sg_n_cols(h)), ex)
    finally:
        L.sg_free(h)


def erdos_renyi(n, p, seed):
    """G(n, p), undirected."""
    return _take(graph_lib().sg_er(int(n), float(p), int(seed)))


def barabasi_albert(n, m, seed):
    """Preferential attachment from a star on m + 1 nodes; m distinct targets per new node."""
    return _take(graph_lib().sg_ba(int(n), int(m), int(seed)))


def watts_strogatz(n, k, p, seed):
    """Ring lattice of even degree k, each lattice edge rewired with probability p."""
    return _take(graph_lib().sg_ws(int(n), int(k), float(p), int(seed)))


def sbm_part(n_blocks, block_size, deg_in, deg_out, n_parts, part, seed):
    """One part of a stochastic block model: local rows, ghost columns after them;
    `extra` holds the global index of every ghost column."""
    return _take(graph_lib().sg_sbm_part(int(n_blocks), int(block_size), float(deg_in), float(deg_out),
                                         int(n_parts), int(part), int(seed)))


def from_edges(n, u, v, directed=False):
    u = np.ascontiguousarray(u, dtype=np.int64)
    v = np.ascontiguousarray(v, dtype=np.int64)
    assert u.shape == v.shape
    return _take(graph_lib().sg_from_edges(int(n), ptr(u), ptr(v), u.shape[0], 1 if directed else 0))


def degree_sort(g):
    """Relabel rows by descending degree (ties by ascending index), keeping every
    row's order.  Returns (sorted graph, perm, inv) with perm[new] = old and
    inv[old] = new.  Ghost columns (>= n) keep their values."""
    rp = np.ascontiguousarray(g.row_ptr, np.int64)
    ci = np.ascontiguousarray(g.col_idx, np.int32)
    perm = np.empty(g.n, np.int32)
    inv = np.empty(g.n, np.int32)
    new_rp = np.empty(g.n + 1, np.int64)
    new_ci = np.empty(max(ci.shape[0], 1), np.int32)
    rc = graph_lib().sg_degree_sort_csr(g.n, ptr(rp), ptr(ci), ptr(new_rp), ptr(new_ci), ptr(perm), ptr(inv))
    if rc != 0:
        raise ValueError(graph_lib().sg_last_error().decode())
    return CSRGraph(new_rp, new_ci[:ci.shape[0]], g.n_cols, g.extra), perm, inv


def set_threads(n):
    """OpenMP threads used by this (calling) thread's generator calls."""
    graph_lib().sg_set_num_threads(int(n))


def scramble_ids(ids, seed):
    """Bijective map of [0, 2^62) used to give agents random-looking int64 ids."""
    ids = np.ascontiguousarray(ids, np.int64)
    out = np.empty_like(ids)
    graph_lib().sg_scramble_ids(ptr(ids), ptr(out), ids.shape[0], int(seed))
    return out
