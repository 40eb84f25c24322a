"""Ghost topology of one partition, on the GPU.

`discover_ghost_topology` keeps the signature and result of the reference's
function (gpu_kernels.py:35-88): the sorted unique ids of the neighbours that
another rank owns.  `ghost_maps` adds the receive side of
build_communication_maps (gpu_kernels.py:521-850): the ids needed from each
owner rank, in id order, and their ghost slots (ghosts follow the local slots
in ascending id order, model.py:1069-1073).  The host only flattens the
inputs; marking, de-duplication, sorting and grouping by rank run in
libsagestep (ss_ghost_discover).
"""
from dataclasses import dataclass, field
from typing import Dict

import numpy as np

from .engine import Engine


def _flat_ids(all_neighbors):
    """Neighbour ids as int64, padding dropped (NaN and negative entries,
    gpu_kernels.py:50-65)."""
    if isinstance(all_neighbors, np.ndarray) and all_neighbors.ndim == 1 and all_neighbors.dtype.kind in "iu":
        flat = all_neighbors.astype(np.int64, copy=False)
        return flat[flat >= 0]
    if len(all_neighbors) == 0:
        return np.zeros(0, np.int64)
    if isinstance(all_neighbors[0], np.ndarray):
        flat = np.concatenate([np.asarray(r, np.float64).ravel() for r in all_neighbors])
    else:
        flat = np.array([x for row in all_neighbors for x in row], dtype=np.float64)
    if flat.size == 0:
        return np.zeros(0, np.int64)
    return flat[~np.isnan(flat) & (flat >= 0)].astype(np.int64)


def _rank_table(agent2rank):
    """Dense owner table (gpu_kernels.py:67-76): an array as given, or a dict
    spread over [0, max id]."""
    if isinstance(agent2rank, np.ndarray):
        return np.ascontiguousarray(agent2rank, np.int32)
    if not agent2rank:
        return np.zeros(0, np.int32)
    table = np.full(max(agent2rank.keys()) + 1, -1, np.int32)
    keys = np.fromiter(agent2rank.keys(), np.int64, len(agent2rank))
    table[keys] = np.fromiter(agent2rank.values(), np.int32, len(agent2rank))
    return table


@dataclass
class GhostMaps:
    ghost_ids: np.ndarray                                     # ascending; ghost g sits in slot n_local + g
    recv_ids: Dict[int, np.ndarray] = field(default_factory=dict)     # rank -> ids needed from it, ascending
    recv_slots: Dict[int, np.ndarray] = field(default_factory=dict)   # rank -> their ghost slots


def ghost_maps(engine: Engine, all_neighbors, agent2rank, my_rank: int, n_local: int = 0) -> GhostMaps:
    owner = _rank_table(agent2rank)
    nbr = _flat_ids(all_neighbors)
    n_ranks = max(int(owner.max()) + 1 if owner.size else 1, int(my_rank) + 1)
    ids, offsets, index = engine.ghost_discover(nbr, owner, my_rank, n_ranks)
    out = GhostMaps(ids)
    for r in range(n_ranks):
        pos = index[offsets[r]:offsets[r + 1]]
        if pos.size:
            out.recv_ids[r] = ids[pos]
            out.recv_slots[r] = (pos + int(n_local)).astype(np.int32)
    return out


def discover_ghost_topology(all_neighbors, agent2rank, my_rank, engine: Engine = None):
    """As gpu_kernels.py:35-88: the sorted unique ids, among the local agents'
    neighbours, of agents another rank owns (int64)."""
    own = engine is None
    eng = Engine(0) if own else engine
    try:
        return ghost_maps(eng, all_neighbors, agent2rank, my_rank).ghost_ids
    finally:
        if own:
            eng.close()
