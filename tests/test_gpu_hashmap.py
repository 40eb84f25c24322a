"""The device id -> slot map against the restatement of the reference's
GPUHashMap (oracle/hashmap.py, gpu_kernels.py:91-260): after every build,
insert, remove and resize the device's key and value arrays equal the
reference mirror's, slot for slot, and so do sizes and lookups."""
import numpy as np
import pytest

from oracle.hashmap import RefHashMap, home_slot
from sagesim_amd import Engine, graphs
from sagesim_amd._lib import NativeError
from sagesim_amd.engine import sir_model

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(scope="module")
def eng():
    e = Engine(0)
    yield e
    e.close()


def _replica(eng, n_ghost):
    g = graphs.from_edges(10, [], [])
    return eng.add_replica(g.row_ptr, g.col_idx, np.arange(10), np.zeros(n_ghost + 10, np.int8),
                           sir_model(0.1, 0.1), 1, n_ghost=n_ghost)


def _same(eng, rep, ref, where):
    keys, vals = eng.map_arrays(rep)
    cap, size = eng.map_info(rep)
    assert cap == ref.capacity, where
    assert size == ref.size, where
    bad = np.nonzero((keys != ref.keys) | (vals != ref.values))[0]
    assert bad.size == 0, f"{where}: {bad.size} slots differ, first {bad[:5].tolist()}"


def test_map_operations_match_reference_mirror(eng):
    rng = np.random.default_rng(11)
    n_slots = 40000
    rep = _replica(eng, n_slots)
    say("hash map build with repeated ids")
    ids = rng.integers(-(1 << 62), 1 << 62, size=20000, dtype=np.int64)
    ids = ids[(ids != -1) & (ids != -2)]
    dup = rng.choice(ids, 1500)
    all_ids = np.concatenate([ids, dup])
    rng.shuffle(all_ids)
    slots = rng.integers(0, n_slots, size=all_ids.shape[0]).astype(np.int32)
    eng.map_build(rep, all_ids, slots, 0.5)
    ref = RefHashMap(max(64, int(np.ceil(all_ids.shape[0] / 0.5))))
    ref.build_from_arrays(all_ids, slots)
    _same(eng, rep, ref, "build")
    probe = np.concatenate([all_ids[:3000], rng.integers(0, 1 << 62, size=500, dtype=np.int64)])
    got, miss = eng.map_lookup(rep, probe)
    want = np.array([ref.lookup_cpu(x) for x in probe], np.int32)
    assert np.array_equal(got, want) and miss == int((want < 0).sum())

    say("hash map inserts: updates to higher slots, new ids, a repeat within the batch")
    upd = all_ids[:2000]
    new = rng.integers(0, 1 << 62, size=1000, dtype=np.int64)
    batch = np.concatenate([upd, new, upd[:10]])
    bvals = np.concatenate([np.full(2000, n_slots - 1), rng.integers(0, n_slots, 1000), np.arange(10)]).astype(np.int32)
    eng.map_insert(rep, batch, bvals)
    for a, v in zip(batch, bvals):
        ref.insert(a, v)
    _same(eng, rep, ref, "insert")

    say("hash map removes: present, absent, repeated in the batch")
    rm = np.concatenate([all_ids[2000:4000], rng.integers(0, 1 << 62, size=100, dtype=np.int64), all_ids[2000:2010],
                         dup[:20]])
    n_rm = eng.map_remove(rep, rm)
    want_rm = sum(ref.remove(a) for a in rm)
    assert n_rm == want_rm
    _same(eng, rep, ref, "remove")

    say("hash map: insert stops at a tombstone ahead of a live key; remove after update")
    cap = ref.capacity
    free = [s for s in range(cap) if ref.keys[s] == -1 and ref.keys[(s + 1) % cap] == -1]
    h = free[0]
    pair, x = [], 1 << 40
    while len(pair) < 2:
        if home_slot(x, cap) == h:
            pair.append(x)
        x += 1
    a, b = pair
    for ops in ([("ins", [a, b], [7, 8])], [("rm", [a])], [("ins", [b], [9])], [("rm", [b])]):
        for op, ks, *vs in ops:
            ks = np.array(ks, np.int64)
            if op == "ins":
                vv = np.array(vs[0], np.int32)
                eng.map_insert(rep, ks, vv)
                for k, v in zip(ks, vv):
                    ref.insert(k, v)
            else:
                eng.map_remove(rep, ks)
                for k in ks:
                    ref.remove(k)
        _same(eng, rep, ref, f"tombstone case {ops}")
    got, _ = eng.map_lookup(rep, np.array([a, b], np.int64))
    assert got.tolist() == [-1, 8] == [ref.lookup_cpu(a), ref.lookup_cpu(b)]

    say("hash map: five update cycles of 4000 live ids stay within the capacity")
    live = np.array([k for k in ref.keys if k >= 0 or k < -2][:4000], np.int64)
    for cyc in range(5):
        vv = np.full(live.shape[0], 100 + cyc, np.int32)
        eng.map_insert(rep, live, vv)
        for k, v in zip(live, vv):
            ref.insert(k, v)
    _same(eng, rep, ref, "update cycles")

    say("hash map resize")
    new_cap = 2 * ref.capacity + 1
    eng.map_resize(rep, new_cap)
    ref.resize(new_cap)
    _same(eng, rep, ref, "resize")
    got, _ = eng.map_lookup(rep, probe)
    assert np.array_equal(got, np.array([ref.lookup_cpu(x) for x in probe], np.int32))
    eng.remove_replica(rep)


def test_full_map_reports_no_slot(eng):
    say("hash map: a full table refuses the next new id")
    rep = _replica(eng, 100)
    eng.map_build(rep, np.zeros(0, np.int64), np.zeros(0, np.int32), 0.5)
    assert eng.map_info(rep) == (64, 0)
    ids = np.arange(1000, 1064, dtype=np.int64)
    eng.map_insert(rep, ids, np.arange(64, dtype=np.int32))
    assert eng.map_info(rep) == (64, 64)
    eng.map_insert(rep, ids[:5], np.arange(5, dtype=np.int32))       # updates still find their keys
    with pytest.raises(NativeError) as ex:
        eng.map_insert(rep, np.array([5], np.int64), np.array([0], np.int32))
    assert "no free slot" in str(ex.value)
    eng.map_resize(rep, 128)
    eng.map_insert(rep, np.array([5], np.int64), np.array([0], np.int32))
    got, miss = eng.map_lookup(rep, np.concatenate([ids, [5]]).astype(np.int64))
    assert miss == 0 and got[-1] == 0 and got[:5].tolist() == list(range(5))
    eng.remove_replica(rep)
