"""Ghost discovery and communication maps on the GPU (ss_ghost_discover)
against a numpy brute force of the reference's discover_ghost_topology
(gpu_kernels.py:35-88) and of the per-rank grouping of
build_communication_maps (gpu_kernels.py:521-850), then the send side through
the id -> slot map, and the SBM partition generator's own ghost list."""
import numpy as np
import pytest

from sagesim_amd import Engine, graphs
from sagesim_amd.engine import sir_model
from sagesim_amd.ghosts import discover_ghost_topology, ghost_maps

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(scope="module")
def eng():
    e = Engine(0)
    yield e
    e.close()


def _brute(flat, owner, my_rank):
    """numpy restatement: valid entries, dense rank lookup, other-rank owners, unique."""
    flat = np.asarray(flat, np.float64)
    ids = flat[~np.isnan(flat) & (flat >= 0)].astype(np.int64)
    rk = np.full(ids.shape[0], -1, np.int64)
    ok = ids < owner.shape[0]
    rk[ok] = owner[ids[ok]]
    cross = (rk != my_rank) & (rk >= 0)
    g = np.unique(ids[cross])
    per = {int(r): np.unique(ids[cross & (rk == r)]) for r in np.unique(rk[cross])}
    return g, per


def test_discover_matches_brute_force(eng):
    say("ghost discovery vs numpy")
    rng = np.random.default_rng(12)
    n_ids, n_ranks, my = 300_000, 8, 3
    owner = rng.integers(0, n_ranks, size=n_ids).astype(np.int32)
    owner[rng.random(n_ids) < 0.02] = -1                      # agents with no owner
    local = np.nonzero(owner == my)[0]
    rows = []
    for a in local[:20000]:
        k = int(rng.integers(0, 12))
        nb = rng.integers(0, n_ids + 500, size=k).astype(np.float64)    # some ids beyond the rank table
        if rng.random() < 0.1:
            nb = np.concatenate([nb, [np.nan, -1.0]])          # padding
        rows.append(nb)
    maps = ghost_maps(eng, rows, owner, my, n_local=local.shape[0])
    want, per = _brute(np.concatenate(rows), owner, my)
    assert np.array_equal(maps.ghost_ids, want)
    assert sorted(maps.recv_ids) == sorted(per)
    for r, ids in per.items():
        assert np.array_equal(maps.recv_ids[r], ids)
        assert np.array_equal(maps.ghost_ids[maps.recv_slots[r] - local.shape[0]], ids)
    # the reference's input forms: lists of ids and a dict owner map
    d = {int(i): int(owner[i]) for i in range(0, n_ids, 7) if owner[i] >= 0}
    small = [list(map(int, r[~np.isnan(r)][:5])) for r in rows[:3000]]
    got = discover_ghost_topology(small, d, my, engine=eng)
    dense = np.full(max(d) + 1, -1, np.int32)
    dense[list(d)] = list(d.values())
    want2, _ = _brute(np.array([x for r in small for x in r], np.float64), dense, my)
    assert np.array_equal(got, want2)
    assert discover_ghost_topology([], d, my, engine=eng).shape == (0,)


def test_discover_offsets_with_empty_and_trailing_ranks(eng):
    """recv_offsets straight from ss_ghost_discover: ranks with no ghosts in
    the middle and at the end, and no ghosts at all."""
    owner = np.array([0, 0, 2, 2, 2, 5, 1, 1], np.int32)
    ids, off, idx = eng.ghost_discover(np.array([7, 2, 2, 4, 0, 3, 6], np.int64), owner, 1, 8)
    assert ids.tolist() == [0, 2, 3, 4]              # 6 and 7 are rank 1's own
    assert off.tolist() == [0, 1, 1, 4, 4, 4, 4, 4, 4] and idx.tolist() == [0, 1, 2, 3]
    ids, off, idx = eng.ghost_discover(np.array([7, 2, 5, 0, 3, 6], np.int64), owner, 1, 8)
    assert ids.tolist() == [0, 2, 3, 5]
    assert off.tolist() == [0, 1, 1, 3, 3, 3, 4, 4, 4]
    assert idx.tolist() == [0, 1, 2, 3]
    ids, off, idx = eng.ghost_discover(np.array([6, 7, 6], np.int64), owner, 1, 8)
    assert ids.shape == (0,) and off.tolist() == [0] * 9 and idx.shape == (0,)
    ids, off, idx = eng.ghost_discover(np.zeros(0, np.int64), owner, 1, 3)
    assert ids.shape == (0,) and off.tolist() == [0] * 4


def test_discover_rejects_ranks_beyond_count(eng):
    from sagesim_amd._lib import NativeError
    owner = np.array([0, 1, 9], np.int32)
    with pytest.raises(NativeError):
        eng.ghost_discover(np.array([0, 1, 2], np.int64), owner, 0, 4)


def test_sbm_partition_ghosts_and_send_lists(eng):
    """The SBM part generator's ghost list (remote ids, ascending) is what
    discovery finds from the part's rows in global ids; the send side maps
    requested local ids to slots through the GPU map over all agents."""
    say("SBM part ghosts")
    nb, bs, parts, part = 16, 4000, 4, 1
    g = graphs.sbm_part(nb, bs, 12.0, 2.0, parts, part, 5)
    n_local = g.n
    lo = part * n_local
    col = g.col_idx.astype(np.int64)
    gid = np.where(col < n_local, lo + col, g.extra[np.clip(col - n_local, 0, None)])
    owner = ((np.arange(nb * bs) // bs) // (nb // parts)).astype(np.int32)
    maps = ghost_maps(eng, gid, owner, part, n_local=n_local)
    assert np.array_equal(maps.ghost_ids, g.extra)
    for r, ids in maps.recv_ids.items():
        assert r != part and (owner[ids] == r).all() and (np.diff(ids) > 0).all()
    assert sum(v.shape[0] for v in maps.recv_ids.values()) == g.extra.shape[0]
    # send side: the map over local and ghost agents (model.py:1112-1115)
    rep = eng.add_replica(g.row_ptr, g.col_idx, np.arange(n_local), np.zeros(g.n_cols, np.int8),
                          sir_model(0.1, 0.1), 1, n_ghost=g.n_cols - n_local)
    all_ids = np.concatenate([np.arange(lo, lo + n_local, dtype=np.int64), maps.ghost_ids])
    eng.map_build(rep, all_ids, np.arange(g.n_cols, dtype=np.int32), 0.5)
    req = np.random.default_rng(1).choice(np.arange(lo, lo + n_local), 500, replace=False)
    slots, miss = eng.map_lookup(rep, req)
    assert miss == 0 and np.array_equal(slots, req - lo)
    gs, miss = eng.map_lookup(rep, maps.recv_ids[min(maps.recv_ids)])
    assert miss == 0 and np.array_equal(gs, maps.recv_slots[min(maps.recv_ids)])
    eng.remove_replica(rep)
