"""Pins oracle/hashmap.py (the restatement of the reference's GPUHashMap,
gpu_kernels.py:91-260) against exact arithmetic, hand-worked layouts and a
dict model, without reference code.  CPU only."""
import random

import numpy as np
import pytest

from oracle.hashmap import DELETED, EMPTY, RefHashMap, home_slot


def test_home_slot_is_floor_mod_of_wrapped_product():
    # the reference computes np.intp(aid) * np.intp(2654435761) % capacity:
    # a product wrapped to int64, then Python's floor-mod (sign of the divisor)
    cases = [0, 1, 7, 2 ** 31, 2 ** 40 + 3, 2 ** 62, 2 ** 63 - 1, -5, -(2 ** 63)]
    for aid in cases:
        for cap in (64, 97, 1000, 2 ** 20 + 7):
            prod = aid * 2654435761
            wrapped = (prod + 2 ** 63) % 2 ** 64 - 2 ** 63
            want = wrapped - cap * (wrapped // cap)
            assert home_slot(aid, cap) == want
            assert 0 <= want < cap
    # a product that wraps to a negative int64 lands by floor-mod, not by C's %
    aid = 5_000_000_000                # product in [2^63, 2^64)
    wrapped = (aid * 2654435761 + 2 ** 63) % 2 ** 64 - 2 ** 63
    assert wrapped < 0 and home_slot(aid, 97) == wrapped % 97


def _ids_with_home(cap, slot, count, start=0):
    out, a = [], start
    while len(out) < count:
        if home_slot(a, cap) == slot:
            out.append(a)
        a += 1
    return out


def test_build_rounds_differ_from_one_by_one_inserts():
    """a, b share home h and c has home h+1.  Round 1 places a at h and c at
    h+1 while b moves on; round 2 finds h+1 taken; round 3 places b at h+2.
    Inserting a, b, c one by one gives a, b, c at h, h+1, h+2."""
    cap, h = 64, 10
    a, b = _ids_with_home(cap, h, 2)
    (c,) = _ids_with_home(cap, h + 1, 1)
    m = RefHashMap(cap)
    m.build_from_arrays([a, b, c], [100, 101, 102])
    assert m.keys[h:h + 3].tolist() == [a, c, b]
    assert m.values[h:h + 3].tolist() == [100, 102, 101]
    assert m.size == 3
    n = RefHashMap(cap)
    for aid, v in ((a, 100), (b, 101), (c, 102)):
        n.insert(aid, v)
    assert n.keys[h:h + 3].tolist() == [a, b, c]
    for mm in (m, n):
        assert [mm.lookup_cpu(x) for x in (a, b, c)] == [100, 101, 102]


def test_build_repeated_ids_keep_every_occurrence_first_found():
    cap, h = 64, 63                     # the probe wraps from the last slot to 0
    a, = _ids_with_home(cap, h, 1)
    m = RefHashMap(cap)
    m.build_from_arrays([a, a, a], [5, 6, 7])
    assert m.keys[63] == a and m.keys[0] == a and m.keys[1] == a
    assert m.values[[63, 0, 1]].tolist() == [5, 6, 7]
    assert m.size == 3 and m.lookup_cpu(a) == 5


def test_insert_overwrites_and_stops_at_a_tombstone_ahead_of_a_live_key():
    cap, h = 64, 20
    a, b = _ids_with_home(cap, h, 2)
    m = RefHashMap(cap)
    m.build_from_arrays([a, b], [1, 2])          # a at h, b at h+1
    m.insert(b, 3)                               # overwrite in place
    assert m.keys[h + 1] == b and m.values[h + 1] == 3 and m.size == 2
    assert m.remove(a)                           # tombstone at h
    assert m.keys[h] == DELETED and m.values[h] == -1 and m.size == 1
    m.insert(b, 4)                               # stops at the tombstone: a second copy of b
    assert m.keys[h] == b and m.values[h] == 4 and m.keys[h + 1] == b and m.size == 2
    assert m.lookup_cpu(b) == 4
    assert m.remove(b)                           # removes the copy found first
    assert m.lookup_cpu(b) == 3 and m.size == 1
    assert not m.remove(a)
    assert m.lookup_cpu(a) == -1


def test_lookups_follow_a_dict_on_sequences_a_dict_models():
    rng = random.Random(7)
    cap = 512
    ids = rng.sample(range(-10 ** 12, 10 ** 12), 300)
    m = RefHashMap(cap)
    m.build_from_arrays(ids[:100], list(range(100)))
    d = {a: i for i, a in enumerate(ids[:100])}
    pool = ids[100:]
    for step in range(400):
        op = rng.random()
        if op < 0.4 and pool:                    # a new id
            a = pool.pop()
            m.insert(a, step)
            d[a] = step
        elif op < 0.7 and d:                     # remove
            a = rng.choice(sorted(d))
            assert m.remove(a)
            del d[a]
        else:                                    # a miss
            assert m.lookup_cpu(rng.randrange(10 ** 13, 10 ** 14)) == -1
        assert m.size == len(d)
    for a in ids:
        assert m.lookup_cpu(a) == d.get(a, -1)
    # updates of live keys with no tombstones in the table
    m2 = RefHashMap(cap)
    m2.build_from_arrays(ids[:200], list(range(200)))
    d2 = {a: i for i, a in enumerate(ids[:200])}
    for _ in range(300):
        a = rng.choice(ids[:200])
        v = rng.randrange(10 ** 6)
        m2.insert(a, v)
        d2[a] = v
    assert m2.size == 200
    assert all(m2.lookup_cpu(a) == v for a, v in d2.items())


def test_resize_keeps_lookups_and_drops_tombstones():
    rng = random.Random(3)
    ids = rng.sample(range(10 ** 9), 120)
    m = RefHashMap(256)
    m.build_from_arrays(ids, list(range(120)))
    for a in ids[:30]:
        m.remove(a)
    before = {a: m.lookup_cpu(a) for a in ids}
    assert not m.needs_resize()
    m.resize(128)
    assert m.capacity == 128 and m.size == 90 and m.needs_resize()    # 90 > int(128 * 0.7) = 89
    assert not (m.keys == DELETED).any()
    assert {a: m.lookup_cpu(a) for a in ids} == before
    # re-insertion runs in the old slot order: two ids with one home in both
    # capacities come out in their old relative order
    pair, x = [], 0
    while len(pair) < 2:
        if home_slot(x, 64) == 5 and home_slot(x, 128) == 69:
            pair.append(x)
        x += 1
    a, b = pair
    m3 = RefHashMap(64)
    m3.build_from_arrays([b, a], [0, 1])
    assert m3.keys[5:7].tolist() == [b, a]
    m3.resize(128)
    assert m3.keys[69:71].tolist() == [b, a] and m3.values[69:71].tolist() == [0, 1]


@pytest.mark.parametrize("cap", [64, 100])
def test_build_matches_one_by_one_when_homes_are_distinct(cap):
    ids = []
    seen = set()
    a = 0
    while len(ids) < cap // 3:
        h = home_slot(a, cap)
        if h not in seen and (h + 1) % cap not in seen and (h - 1) % cap not in seen:
            seen.add(h)
            ids.append(a)
        a += 1
    m = RefHashMap(cap)
    m.build_from_arrays(ids, range(len(ids)))
    n = RefHashMap(cap)
    for i, x in enumerate(ids):
        n.insert(x, i)
    assert np.array_equal(m.keys, n.keys) and np.array_equal(m.values, n.values)
