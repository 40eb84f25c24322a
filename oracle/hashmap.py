"""Restatement of the reference's GPUHashMap CPU mirror (gpu_kernels.py:91-260).

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the sagesim_amd
package.  Plain Python and numpy, written for being obviously right:

  keys int64 (-1 empty, -2 deleted), values int32 (-1 where no entry);
  home slot = floor-mod of the int64-wrapped product id * 2654435761 by the
  capacity (gpu_kernels.py:112-114, :146), linear probing.

  build_from_arrays  gpu_kernels.py:122-169: rounds; in each round every
                     remaining entry looks at its probe slot, the remaining
                     entry of lowest input index among those at an empty slot
                     takes it, every other remaining entry moves one slot on;
                     size = number of entries (repeated ids included).
  insert             gpu_kernels.py:171-191: probe to the first slot that is
                     empty, deleted or holds the id; size grows when that slot
                     did not hold the id; key and value written.
  remove             gpu_kernels.py:193-206: probe until an empty slot; the
                     first slot holding the id becomes deleted (value -1).
  lookup_cpu         gpu_kernels.py:208-215.
  needs_resize       gpu_kernels.py:217-219: size > int(capacity * 0.7).
  resize             gpu_kernels.py:221-247: live entries re-inserted in old
                     slot order, each at the first empty or deleted slot.

Pinning: tests/test_oracle_hashmap.py checks the home slot against exact
big-integer arithmetic, hand-worked layouts of tiny tables (where the build's
rounds and one-by-one inserts place the same ids differently, and where an
insert stops at a tombstone ahead of a live key) and a dict model of the
lookups on operation sequences for which a dict is exact.
"""
import numpy as np

EMPTY = -1
DELETED = -2
_PRIME = 2654435761


def home_slot(aid, capacity):
    """Floor-mod of the int64-wrapped product (Python integers, no numpy)."""
    p = (int(aid) * _PRIME) & 0xFFFFFFFFFFFFFFFF
    if p >= 1 << 63:
        p -= 1 << 64
    return p % int(capacity)


class RefHashMap:
    def __init__(self, capacity):
        self.capacity = int(capacity)
        self.keys = np.full(self.capacity, EMPTY, np.int64)
        self.values = np.full(self.capacity, -1, np.int32)
        self.size = 0

    def build_from_arrays(self, agent_ids, buffer_indices):
        ids = [int(x) for x in agent_ids]
        vals = [int(x) for x in buffer_indices]
        self.keys[:] = EMPTY
        self.values[:] = -1
        self.size = len(ids)
        probe = {i: home_slot(a, self.capacity) for i, a in enumerate(ids)}   # remaining entries
        while probe:
            winner = {}
            for i in sorted(probe):
                s = probe[i]
                if self.keys[s] == EMPTY and s not in winner:
                    winner[s] = i
            for s, i in winner.items():
                self.keys[s] = ids[i]
                self.values[s] = vals[i]
                del probe[i]
            for i in probe:
                probe[i] = (probe[i] + 1) % self.capacity

    def insert(self, agent_id, buffer_index):
        aid = int(agent_id)
        s = home_slot(aid, self.capacity)
        while self.keys[s] not in (EMPTY, DELETED) and self.keys[s] != aid:
            s = (s + 1) % self.capacity
        if self.keys[s] != aid:
            self.size += 1
        self.keys[s] = aid
        self.values[s] = int(buffer_index)

    def remove(self, agent_id):
        aid = int(agent_id)
        s = home_slot(aid, self.capacity)
        while self.keys[s] != EMPTY:
            if self.keys[s] == aid:
                self.keys[s] = DELETED
                self.values[s] = -1
                self.size -= 1
                return True
            s = (s + 1) % self.capacity
        return False

    def lookup_cpu(self, agent_id):
        aid = int(agent_id)
        s = home_slot(aid, self.capacity)
        while self.keys[s] != EMPTY:
            if self.keys[s] == aid:
                return int(self.values[s])
            s = (s + 1) % self.capacity
        return -1

    def needs_resize(self):
        return self.size > int(self.capacity * 0.7)

    def resize(self, new_capacity):
        old = [(int(k), int(v)) for k, v in zip(self.keys, self.values) if k not in (EMPTY, DELETED)]
        self.capacity = int(new_capacity)
        self.keys = np.full(self.capacity, EMPTY, np.int64)
        self.values = np.full(self.capacity, -1, np.int32)
        self.size = 0
        for k, v in old:
            s = home_slot(k, self.capacity)
            while self.keys[s] not in (EMPTY, DELETED):
                s = (s + 1) % self.capacity
            self.keys[s] = k
            self.values[s] = v
            self.size += 1
