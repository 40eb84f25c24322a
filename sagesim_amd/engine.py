"""Python face of libsagestep: a context of device-resident replicas.

`Engine` is a thin owner of one `ss_ctx`.  `Replica` adds the degree-sorted
relabelling: the engine stores rows in descending-degree order (so one
wavefront takes each high-degree agent and lanes take the rest), while the
caller keeps its own agent order; the RNG stays keyed on the caller's ids, so
results do not depend on the relabelling.
"""
import ctypes
import weakref
from dataclasses import dataclass

import numpy as np

from ._lib import (HISTORY_WIDTH, NativeError, SSModel, SSReplicaSpec, ptr, step_lib)
from .graphs import CSRGraph, degree_sort

S, I, R = 0, 1, 2


def sir_model(beta, gamma):
    """SIR {S=0, I=1, R=2}: S -> I with prob 1-(1-beta)^k, I -> R with prob gamma."""
    m = SSModel()
    m.n_states, m.s_code, m.inf_code, m.s_target = 3, 0, 1, 1
    m.beta = float(beta)
    m.trans_p[:] = [0.0, float(gamma), 0.0, 0.0]
    m.trans_next[:] = [0, 2, 2, 2]
    return m


def seir_model(beta, sigma, gamma):
    """SEIR {S=0, E=1, I=2, R=3}: S -> E (1-(1-beta)^k, k = #I), E -> I (sigma), I -> R (gamma)."""
    m = SSModel()
    m.n_states, m.s_code, m.inf_code, m.s_target = 4, 0, 2, 1
    m.beta = float(beta)
    m.trans_p[:] = [0.0, float(sigma), float(gamma), 0.0]
    m.trans_next[:] = [0, 2, 3, 3]
    return m


class Engine:
    """One libsagestep context on one GPU."""

    def __init__(self, device=0):
        self._L = step_lib()
        h = ctypes.c_void_p()
        rc = self._L.ss_create(int(device), ctypes.byref(h))
        if rc != 0:
            raise NativeError(rc, self._L.ss_last_error(None).decode())
        self._h = h
        self.device = device
        self._pinned = {}                  # address -> byte count of live page-locked buffers
        self._halo_n = {}                  # replica -> records of its bound halo

    # -- plumbing --
    def _ck(self, rc):
        if rc != 0:
            raise NativeError(rc, self._L.ss_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            for addr in list(self._pinned):
                self._L.ss_host_free(self._h, ctypes.c_void_p(addr))
            self._pinned.clear()
            self._L.ss_destroy(self._h)
            self._h = None

    def pinned_empty(self, shape, dtype=np.int8):
        """A numpy array in page-locked host memory (ss_host_alloc): halo values
        passed to simulate_halo from such an array are copied with no staging.
        The memory is released with the array, or when the engine closes."""
        dt = np.dtype(dtype)
        nbytes = int(np.prod(shape, dtype=np.int64)) * dt.itemsize
        p = ctypes.c_void_p()
        self._ck(self._L.ss_host_alloc(self._h, nbytes, ctypes.byref(p)))
        addr = p.value
        self._pinned[addr] = nbytes
        buf = (ctypes.c_char * max(nbytes, 1)).from_address(addr)
        arr = np.frombuffer(buf, dtype=dt, count=nbytes // dt.itemsize).reshape(shape)

        def release(engine=weakref.ref(self), addr=addr):
            e = engine()
            if e is not None and e._h and e._pinned.pop(addr, None) is not None:
                e._L.ss_host_free(e._h, ctypes.c_void_p(addr))
        weakref.finalize(arr, release)
        return arr

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def info(self):
        cus, mem, used = ctypes.c_int32(), ctypes.c_int64(), ctypes.c_int64()
        self._ck(self._L.ss_device_info(self._h, ctypes.byref(cus), ctypes.byref(mem), ctypes.byref(used)))
        return {"num_cus": cus.value, "total_mem": mem.value, "ctx_bytes": used.value}

    # -- replicas --
    def add_replica(self, row_ptr, col_idx, rng_key_id, state, model, seed, n_ghost=0,
                    history=None, tick0=0, tick_limit=0):
        rp = np.ascontiguousarray(row_ptr, np.int64)
        ci = np.ascontiguousarray(col_idx, np.int32)
        ids = np.ascontiguousarray(rng_key_id, np.int64)
        st = np.ascontiguousarray(state, np.int8)
        n_local = rp.shape[0] - 1
        if ids.shape[0] != n_local or st.shape[0] != n_local + n_ghost:
            raise ValueError("rng_key_id must have n_local entries and state n_local + n_ghost")
        hist = None
        if history is not None:
            hist = np.ascontiguousarray(history, np.int32)
            if hist.shape != (n_local, HISTORY_WIDTH):
                raise ValueError(f"history must be ({n_local}, {HISTORY_WIDTH}) int32")
        sp = SSReplicaSpec()
        sp.n_local, sp.n_ghost, sp.nnz = n_local, int(n_ghost), ci.shape[0]
        sp.row_ptr, sp.col_idx, sp.rng_key_id, sp.state = ptr(rp), ptr(ci), ptr(ids), ptr(st)
        sp.history = ptr(hist) if hist is not None else None
        sp.model = model
        sp.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
        sp.tick0 = int(tick0)
        sp.tick_limit = int(tick_limit)
        rep = ctypes.c_int32()
        self._ck(self._L.ss_replica_create(self._h, ctypes.byref(sp), ctypes.byref(rep)))
        return rep.value

    def remove_replica(self, rep):
        self._ck(self._L.ss_replica_destroy(self._h, rep))
        self._halo_n.pop(rep, None)

    def reset(self, rep):
        self._ck(self._L.ss_replica_reset(self._h, rep))

    def tick(self, rep):
        t = ctypes.c_int64()
        self._ck(self._L.ss_replica_tick(self._h, rep, ctypes.byref(t)))
        return t.value

    def set_tick(self, rep, tick, tick_limit=0):
        self._ck(self._L.ss_replica_set_tick(self._h, rep, int(tick), int(tick_limit)))

    def replica_bytes(self, rep):
        b = ctypes.c_int64()
        self._ck(self._L.ss_replica_bytes(self._h, rep, ctypes.byref(b)))
        return b.value

    def set_model(self, rep, model):
        self._ck(self._L.ss_set_model(self._h, rep, ctypes.byref(model)))

    def set_seed(self, rep, seed):
        self._ck(self._L.ss_set_seed(self._h, rep, int(seed) & 0xFFFFFFFFFFFFFFFF))

    def state(self, rep, n_total):
        out = np.empty(n_total, np.int8)
        self._ck(self._L.ss_state_download(self._h, rep, ptr(out), n_total))
        return out

    def set_state(self, rep, st):
        st = np.ascontiguousarray(st, np.int8)
        self._ck(self._L.ss_state_upload(self._h, rep, ptr(st), st.shape[0]))

    def history(self, rep, n_local):
        out = np.empty((n_local, HISTORY_WIDTH), np.int32)
        self._ck(self._L.ss_history_download(self._h, rep, ptr(out), n_local))
        return out

    # -- checkpoint --
    def replica_save(self, rep):
        """The replica's state as a uint8 blob (ss_replica_save)."""
        size = ctypes.c_int64()
        self._ck(self._L.ss_replica_save(self._h, rep, None, 0, ctypes.byref(size)))
        blob = np.zeros(size.value, np.uint8)
        self._ck(self._L.ss_replica_save(self._h, rep, ptr(blob), blob.shape[0], ctypes.byref(size)))
        return blob

    def replica_restore(self, rep, blob):
        """Restore a blob of replica_save into a replica of the same network;
        returns the restored tick."""
        blob = np.ascontiguousarray(blob, np.uint8)
        t = ctypes.c_int64()
        self._ck(self._L.ss_replica_restore(self._h, rep, ptr(blob), blob.shape[0], ctypes.byref(t)))
        return t.value

    # -- run --
    def set_tiling(self, heavy_degree=256, agents_per_lane=16, grid_blocks=0, use_graph=True, hub_degree=1024):
        """Launch shape: rows of degree >= hub_degree get a block, >= heavy_degree a
        wavefront, the rest agents_per_lane agents per lane (hub_degree is raised
        to heavy_degree when below it)."""
        hub_degree = max(int(hub_degree), int(heavy_degree))
        self._ck(self._L.ss_set_tiling(self._h, hub_degree, int(heavy_degree), int(agents_per_lane),
                                       int(grid_blocks), 1 if use_graph else 0))

    def set_push(self, ratio=0.25):
        """Push-direction rule (default 0.25): a tick of a symmetric, ghost-free
        replica writes the next tick's infectious-neighbour counts along the
        rows of its infectious agents when sum(degree of infectious) <= ratio *
        #susceptible; the next tick then decides from the counts and skips idle
        tiles.  0 disables it.  Results are the same either way."""
        self._ck(self._L.ss_set_push(self._h, float(ratio)))

    def push_ok(self, rep):
        out = ctypes.c_int32()
        self._ck(self._L.ss_replica_push_ok(self._h, rep, ctypes.byref(out)))
        return bool(out.value)

    def launch_shape(self):
        nt, g = ctypes.c_int32(), ctypes.c_int32()
        self._ck(self._L.ss_launch_shape(self._h, ctypes.byref(nt), ctypes.byref(g)))
        return nt.value, g.value

    def simulate(self, ticks):
        self._ck(self._L.ss_simulate(self._h, int(ticks)))

    def sync(self):
        self._ck(self._L.ss_sync(self._h))

    # -- measurement --
    def set_counting(self, on):
        self._ck(self._L.ss_set_counting(self._h, 1 if on else 0))

    def read_counters(self):
        """(susceptible agents visited, draws, scans started, CSR entries read,
        counts pushed, replica-ticks in the push direction, agents of the tiles
        that ran); clears them."""
        out = np.zeros(7, np.int64)
        self._ck(self._L.ss_read_counters(self._h, ptr(out)))
        return out

    def census(self, rep):
        """(#S, #I or E, #R or I, #state 3, sum of degrees of susceptible agents)."""
        out = np.zeros(5, np.int64)
        self._ck(self._L.ss_census(self._h, rep, ptr(out)))
        return out

    def copy_probe(self, nbytes, repeats=5):
        """Average ms of a plain device copy of nbytes (the known-size traffic reference)."""
        ms = ctypes.c_float()
        self._ck(self._L.ss_copy_probe(self._h, int(nbytes), int(repeats), ctypes.byref(ms)))
        return ms.value

    def gather_probe(self, span_bytes, n_gathers, repeats=3):
        """Average ms of n_gathers scattered 4-byte reads over a buffer of span_bytes."""
        ms = ctypes.c_float()
        self._ck(self._L.ss_gather_probe(self._h, int(span_bytes), int(n_gathers), int(repeats), ctypes.byref(ms)))
        return ms.value

    def timer_start(self):
        self._ck(self._L.ss_timer_start(self._h))

    def timer_stop(self):
        """Milliseconds between timer_start and now on the context's stream (waits)."""
        ms = ctypes.c_float()
        self._ck(self._L.ss_timer_stop(self._h, ctypes.byref(ms)))
        return ms.value

    # -- ghosts --
    def map_build(self, rep, ids, slots, load_factor=0.5):
        ids = np.ascontiguousarray(ids, np.int64)
        slots = np.ascontiguousarray(slots, np.int32)
        self._ck(self._L.ss_map_build(self._h, rep, ptr(ids), ptr(slots), ids.shape[0], float(load_factor)))

    def map_insert(self, rep, ids, slots):
        ids = np.ascontiguousarray(ids, np.int64)
        slots = np.ascontiguousarray(slots, np.int32)
        self._ck(self._L.ss_map_insert(self._h, rep, ptr(ids), ptr(slots), ids.shape[0]))

    def map_remove(self, rep, ids):
        ids = np.ascontiguousarray(ids, np.int64)
        n = ctypes.c_int64()
        self._ck(self._L.ss_map_remove(self._h, rep, ptr(ids), ids.shape[0], ctypes.byref(n)))
        return n.value

    def map_lookup(self, rep, ids):
        ids = np.ascontiguousarray(ids, np.int64)
        out = np.empty(ids.shape[0], np.int32)
        miss = ctypes.c_int64()
        self._ck(self._L.ss_map_lookup(self._h, rep, ptr(ids), ptr(out), ids.shape[0], ctypes.byref(miss)))
        return out, miss.value

    def map_resize(self, rep, new_capacity):
        self._ck(self._L.ss_map_resize(self._h, rep, int(new_capacity)))

    def map_info(self, rep):
        """(capacity, size) of the replica's id -> slot map."""
        cap, size = ctypes.c_int64(), ctypes.c_int64()
        self._ck(self._L.ss_map_info(self._h, rep, ctypes.byref(cap), ctypes.byref(size)))
        return cap.value, size.value

    def map_arrays(self, rep):
        """Copies of the map's key and value arrays (-1 empty, -2 deleted)."""
        cap, _ = self.map_info(rep)
        keys = np.empty(cap, np.int64)
        vals = np.empty(cap, np.int32)
        self._ck(self._L.ss_map_download(self._h, rep, ptr(keys), ptr(vals), cap))
        return keys, vals

    def ghost_discover(self, nbr_ids, owner, my_rank, n_ranks):
        """Ghost ids (ascending) and, per owner rank r, the positions of r's
        ghosts in them: ghost_ids[index[offsets[r]:offsets[r + 1]]] are the ids
        this rank needs from r (ss_ghost_discover)."""
        nbr = np.ascontiguousarray(nbr_ids, np.int64)
        own = np.ascontiguousarray(owner, np.int32)
        n = ctypes.c_int64()
        self._ck(self._L.ss_ghost_discover(self._h, ptr(nbr), nbr.shape[0], ptr(own), own.shape[0], int(my_rank),
                                           int(n_ranks), None, None, None, 0, ctypes.byref(n)))
        ids = np.empty(n.value, np.int64)
        offsets = np.empty(int(n_ranks) + 1, np.int32)
        index = np.empty(n.value, np.int32)
        self._ck(self._L.ss_ghost_discover(self._h, ptr(nbr), nbr.shape[0], ptr(own), own.shape[0], int(my_rank),
                                           int(n_ranks), ptr(ids), ptr(offsets), ptr(index), ids.shape[0],
                                           ctypes.byref(n)))
        return ids, offsets, index

    def halo_bind(self, rep, ids):
        ids = np.ascontiguousarray(ids, np.int64)
        miss = ctypes.c_int64()
        self._ck(self._L.ss_halo_bind(self._h, rep, ptr(ids), ids.shape[0], ctypes.byref(miss)))
        self._halo_n[rep] = ids.shape[0]
        return miss.value

    def _halo_ticks(self, rep, ticks, buf, per_tick):
        n = self._halo_n.get(rep)
        if n is None:
            raise ValueError(f"replica {rep} has no bound halo")
        need = int(ticks) * per_tick(n)
        if buf.size < need:
            raise ValueError(f"{ticks} ticks of a {n}-record halo need {need} entries, got {buf.size}")

    def halo_unpack(self, rep, vals):
        vals = np.ascontiguousarray(vals, np.int8)
        self._ck(self._L.ss_halo_unpack(self._h, rep, ptr(vals), vals.shape[0]))

    def halo_unpack_ids(self, rep, ids, vals):
        ids = np.ascontiguousarray(ids, np.int64)
        vals = np.ascontiguousarray(vals, np.int8)
        miss = ctypes.c_int64()
        self._ck(self._L.ss_halo_unpack_ids(self._h, rep, ptr(ids), ptr(vals), ids.shape[0], ctypes.byref(miss)))
        return miss.value

    def halo_pack(self, rep, slots):
        slots = np.ascontiguousarray(slots, np.int32)
        out = np.empty(slots.shape[0], np.int8)
        self._ck(self._L.ss_halo_pack(self._h, rep, ptr(slots), slots.shape[0], ptr(out)))
        return out

    def simulate_halo(self, rep, ticks, vals):
        vals = np.ascontiguousarray(vals, np.int8)
        self._halo_ticks(rep, ticks, vals, lambda n: n)
        self._ck(self._L.ss_simulate_halo(self._h, rep, int(ticks), ptr(vals)))

    # 2-bit halo: state codes packed four records a byte (pack_2bit below)
    def halo_unpack_2bit(self, rep, packed, n):
        packed = np.ascontiguousarray(packed, np.uint8)
        if packed.size < (n + 3) // 4:
            raise ValueError(f"{packed.size} bytes hold fewer than {n} records")
        self._ck(self._L.ss_halo_unpack_2bit(self._h, rep, ptr(packed), int(n)))

    def halo_pack_2bit(self, rep, slots):
        slots = np.ascontiguousarray(slots, np.int32)
        out = np.empty((slots.shape[0] + 3) // 4, np.uint8)
        self._ck(self._L.ss_halo_pack_2bit(self._h, rep, ptr(slots), slots.shape[0], ptr(out)))
        return out

    def simulate_halo_2bit(self, rep, ticks, packed):
        """simulate_halo with every tick's halo as 2-bit codes: packed[j] holds
        tick j's records, (n_bound + 3) // 4 bytes (pack_2bit)."""
        packed = np.ascontiguousarray(packed, np.uint8)
        self._halo_ticks(rep, ticks, packed, lambda n: (n + 3) // 4)
        self._ck(self._L.ss_simulate_halo_2bit(self._h, rep, int(ticks), ptr(packed)))


@dataclass
class ReplicaData:
    """Host arrays of a replica in storage order (degree-sorted), ready to upload.
    Building one touches no GPU state, so builder threads can prepare them."""
    graph: CSRGraph
    perm: np.ndarray
    inv: np.ndarray
    state: np.ndarray
    rng_key_id: np.ndarray
    history: np.ndarray = None

    @property
    def n(self):
        return self.graph.n

    @property
    def n_total(self):
        return self.graph.n_cols


def _map_hist(h, table, n):
    h = h.copy()
    local = (h >= 0) & (h < n)
    h[local] = table[h[local]]
    return np.ascontiguousarray(h, np.int32)


def prepare_replica(graph: CSRGraph, state, rng_key_id=None, history=None, sort=True):
    """Degree-sort a graph and carry states, RNG keys and history along."""
    n = graph.n
    state = np.asarray(state, np.int8)
    if state.shape[0] != graph.n_cols:
        raise ValueError("state must cover local and ghost slots")
    if rng_key_id is None:
        rng_key_id = np.arange(n, dtype=np.int64)
    rng_key_id = np.asarray(rng_key_id, np.int64)
    if sort:
        gs, perm, inv = degree_sort(graph)
    else:
        gs, perm, inv = graph, np.arange(n, dtype=np.int32), np.arange(n, dtype=np.int32)
    st = np.concatenate([state[:n][perm], state[n:]])
    hist = None
    if history is not None:
        hist = _map_hist(np.asarray(history, np.int32)[perm], inv, n)
    return ReplicaData(gs, perm, inv, st, np.ascontiguousarray(rng_key_id[perm]), hist)


class Replica:
    """A graph + compartment model placed on an Engine in degree-sorted order.

    States and history are exchanged in the caller's agent order.  Slots at
    and beyond graph.n are ghosts (graph.n_cols - graph.n of them).
    """

    def __init__(self, engine, graph: CSRGraph = None, state=None, model=None, seed=0, rng_key_id=None,
                 history=None, tick0=0, tick_limit=0, sort=True, data: ReplicaData = None):
        if data is None:
            data = prepare_replica(graph, state, rng_key_id, history, sort)
        self.engine = engine
        self.graph, self.perm, self.inv = data.graph, data.perm, data.inv
        self.n = data.n
        self.n_total = data.n_total
        self.n_ghost = self.n_total - self.n
        self.rep = engine.add_replica(data.graph.row_ptr, data.graph.col_idx, data.rng_key_id, data.state, model,
                                      seed, n_ghost=self.n_ghost, history=data.history, tick0=tick0,
                                      tick_limit=tick_limit)

    def states(self):
        st = self.engine.state(self.rep, self.n_total)
        out = np.empty_like(st)
        out[self.perm] = st[:self.n]
        out[self.n:] = st[self.n:]
        return out

    def set_states(self, state):
        state = np.asarray(state, np.int8)
        self.engine.set_state(self.rep, np.concatenate([state[:self.n][self.perm], state[self.n:]]))

    def save(self):
        """Checkpoint blob of the replica (Engine.replica_save)."""
        return self.engine.replica_save(self.rep)

    def restore(self, blob):
        """Restore a checkpoint of a replica of this graph; returns its tick."""
        return self.engine.replica_restore(self.rep, blob)

    def history(self):
        h = self.engine.history(self.rep, self.n)
        out = np.empty_like(h)
        out[self.perm] = h
        return _map_hist(out, self.perm, self.n)

    def tick(self):
        return self.engine.tick(self.rep)

    def reset(self):
        self.engine.reset(self.rep)

    def remove(self):
        self.engine.remove_replica(self.rep)


def pack_2bit(vals):
    """State codes (< 4) to the 2-bit halo format along the last axis: record h
    in bits 2 (h % 4) and 2 (h % 4) + 1 of byte h // 4."""
    v = np.asarray(vals)
    if v.size and (v.min() < 0 or v.max() > 3):
        raise ValueError("2-bit halo codes must lie in [0, 4)")
    n = v.shape[-1]
    pad = (-n) % 4
    u = v.astype(np.uint8)
    if pad:
        u = np.concatenate([u, np.zeros(v.shape[:-1] + (pad,), np.uint8)], axis=-1)
    u = u.reshape(v.shape[:-1] + ((n + pad) // 4, 4))
    return (u[..., 0] | (u[..., 1] << 2) | (u[..., 2] << 4) | (u[..., 3] << 6)).astype(np.uint8)


def unpack_2bit(packed, n):
    """Inverse of pack_2bit: the first n records as int8."""
    p = np.asarray(packed, np.uint8)
    u = np.stack([(p >> s) & 3 for s in (0, 2, 4, 6)], axis=-1).reshape(p.shape[:-1] + (4 * p.shape[-1],))
    return u[..., :n].astype(np.int8)

