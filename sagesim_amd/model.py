"""Model: the reference's user-facing API on the MI355X step path.

Keeps the names, arguments and meaning of sagesim/model.py's public methods
(register_breed, create_agent_of_breed, get/set_agent_property_value,
register_global_property, set_seed, set_agent_logical_id, setup, simulate,
reset).  `setup` (model.py:682-886) sorts agents by breed (agent.py:438-473),
builds the CSR from the space, turns every property into a device tensor
(float32 unless declared otherwise, as internal_utils.py:189-230), writes one
wrapper kernel per priority group around the breeds' HIP step functions and
compiles it with hipRTC.  `simulate` (model.py:961-998) runs the ticks on the
device: each priority group is one launch, then double-buffered properties are
written back.  When the only step function is a built-in compartment step
(sagesim_amd.steps), setup hands the model to the hand-written fused kernel
instead; results are the same either way.
"""
from collections import OrderedDict
import ctypes
from copy import copy
import json
from math import nan
import os
import re

import numpy as np

from . import graphs
from ._lib import NativeError, ptr, step_lib
from .breed import Breed
from .engine import Engine, Replica, seir_model, sir_model

_DTYPES = {
    "float32": ("float", np.float32), "float64": ("double", np.float64), "int64": ("long long", np.int64),
    "int32": ("int", np.int32), "int16": ("short", np.int16), "int8": ("signed char", np.int8),
    "uint8": ("unsigned char", np.uint8),
}


def build_csr_from_ragged(ragged_list):
    """Ragged neighbour lists -> (offsets, values), internal_utils.py:111-164.
    Every row is written in its own iteration order (a set's, a list's)."""
    lengths = [len(r) if isinstance(r, (list, tuple, set, np.ndarray)) else 0 for r in ragged_list]
    offsets = np.zeros(len(lengths) + 1, np.int64)
    np.cumsum(lengths, out=offsets[1:])
    values = np.empty(int(offsets[-1]), np.int32)
    pos = 0
    for r in ragged_list:
        if isinstance(r, set):
            r = list(r)
        n = len(r) if isinstance(r, (list, tuple, np.ndarray)) else 0
        if n:
            values[pos:pos + n] = np.asarray(r, np.int64)
            pos += n
    return offsets, values


def build_csr_values_only(ragged_list, offsets):
    """Values of a CSR whose offsets are known (internal_utils.py:167-186)."""
    return build_csr_from_ragged(ragged_list)[1][: int(offsets[-1])]



# Rows of the hipRTC path that start first (longest first): degree at least
# _LONG_ROW_FACTOR times the priority group's mean and at least _LONG_ROW_MIN.
_LONG_ROW_FACTOR = 8.0
_LONG_ROW_MIN = 64.0

class _Program:
    """Owner of one ss_prog (general step functions)."""

    def __init__(self, engine, n):
        self._L = step_lib()
        self.engine = engine
        h = ctypes.c_void_p()
        rc = self._L.ss_prog_create(engine._h, int(n), ctypes.byref(h))
        if rc:
            raise NativeError(rc, self._L.ss_prog_last_error(None).decode())
        self._h = h

    def ck(self, rc):
        if rc:
            raise NativeError(rc, self._L.ss_prog_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            self._L.ss_prog_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_IDENT = re.compile(r"[A-Za-z_]\w*")


def _split_decl(decl):
    """'int *spikes' -> ('int *', 'spikes'); a bare name is a float buffer."""
    decl = decl.strip()
    if _IDENT.fullmatch(decl):
        return "float *", decl
    m = re.fullmatch(r"(.*?)([A-Za-z_]\w*)\s*", decl)
    if not m or not m.group(1).strip():
        raise ValueError(f"extra kernel parameter {decl!r} is not 'name' or 'type name'")
    return m.group(1).strip(), m.group(2)


class Model:
    def __init__(self, space, threads_per_block: int = 32, step_function_file_path: str = None,
                 verbose_timing: bool = False, agent_slack_factor: float = 1.5, csr_slack_factor: float = 2.0,
                 min_capacity: int = 64, device: int = 0, use_fused: bool = True) -> None:
        # threads_per_block and the slack factors are accepted for signature
        # compatibility (model.py:357-372); launch shapes and capacities are
        # chosen by the library.
        self._threads_per_block = threads_per_block
        self._step_function_file_path = step_function_file_path
        self._verbose_timing = verbose_timing
        self._space = space
        space._agent_factory = self
        self._device = device
        self._use_fused = use_fused
        self._breeds = []
        self._prop_names = ["breed", "locations"]
        self._prop_defaults = {"breed": 0, "locations": None}
        self._prop_dtype = {}
        self._prop_maxdims = {}
        # neighbour visibility by property (agent.py:41-48, :295-309): decides
        # which properties a multi-rank exchange sends; one rank sends nothing
        self._prop_visible = OrderedDict([("breed", True), ("locations", False)])
        self._agent_values = {}          # property -> list of per-agent values (creation order)
        self._agent_breed = []
        self._global_tensors = []
        self._global_names = {}
        self._seed = int(np.random.randint(0, 2 ** 31))
        self._logical_id_map = {}
        self._breed_local_arrays = []    # {name, breed, shape_per_agent, neighbor_visible} (model.py:620-645)
        self.tick = 0
        self._is_setup = False
        self._mode = None
        self._engine = None
        self._prog = None
        self._replica = None
        self._dirty_props = set()
        self._globals_dirty = False
        self._cache = {}
        self._tick_timings = []
        self._last_simulate_ticks = 0
        self._kernel_extra_arrays = []

    # ---- registration ----
    @property
    def verbose_timing(self) -> bool:
        return self._verbose_timing

    @verbose_timing.setter
    def verbose_timing(self, value: bool) -> None:
        self._verbose_timing = value

    def get_space(self):
        return self._space

    def register_breed(self, breed: Breed) -> None:
        if self._agent_breed:
            raise Exception("All breeds must be registered before agents are created!")
        breed._breedidx = len(self._breeds)
        self._breeds.append(breed)
        for name, default in breed.properties.items():
            visible = breed.prop2neighbor_visible.get(name, True)
            if name not in self._prop_defaults:
                self._prop_names.append(name)
                self._agent_values[name] = []
                self._prop_dtype[name] = breed._prop2dtype.get(name, "float32")
                self._prop_maxdims[name] = breed._prop2maxdims.get(name)
                self._prop_visible[name] = visible
            else:
                # visible when any breed marks it visible (agent.py:299-303)
                self._prop_visible[name] = self._prop_visible.get(name, False) or visible
            self._prop_defaults[name] = default

    def set_property_neighbor_visible(self, property_name: str, visible: bool) -> None:
        """As model.py:398-409: before setup only, for a known property."""
        if self._is_setup:
            raise RuntimeError("Cannot change neighbor_visible after setup()")
        if property_name not in self._prop_visible:
            raise ValueError(f"Unknown property: {property_name}")
        self._prop_visible[property_name] = bool(visible)

    def register_breed_local_array(self, name: str, breed: Breed, shape_per_agent: tuple,
                                   neighbor_visible: bool = False) -> int:
        """As model.py:620-645: a float32 array with one row per agent of
        `breed`, each row of shape `shape_per_agent` ((a,) or (a, b)), zero at
        setup.  Step functions take it, after the properties, as the pair
        `array, idx_map` and reach an agent's row as
        `array[idx_map[agent_index]]`.  It is double-buffered like a property
        unless a step function lists its name in `no_double_buffer`."""
        if self._is_setup:
            raise RuntimeError("register breed-local arrays before setup()")
        if breed not in self._breeds:
            raise ValueError(f"breed {breed.name} is not registered")
        if not _IDENT.fullmatch(name) or any(b["name"] == name for b in self._breed_local_arrays):
            raise ValueError(f"breed-local array name {name!r} is not a new identifier")
        shape = tuple(int(x) for x in shape_per_agent)
        if len(shape) not in (1, 2) or min(shape) < 1:
            raise ValueError("shape_per_agent must be (a,) or (a, b) with positive sizes")
        idx = len(self._breed_local_arrays)
        self._breed_local_arrays.append({"name": name, "breed": breed, "shape_per_agent": shape,
                                         "neighbor_visible": bool(neighbor_visible)})
        return idx

    def get_breed_local_array(self, name: str) -> np.ndarray:
        """As model.py:647-650: the whole array (rows of the breed's agents in
        buffer order) as float32 numpy."""
        if not self._is_setup or self._mode != "rtc":
            raise RuntimeError("breed-local arrays live on the device after setup()")
        b = next(i for i, x in enumerate(self._breed_local_arrays) if x["name"] == name)
        aid, _, rows = self._bla_ids[b]
        out = np.empty((rows,) + self._breed_local_arrays[b]["shape_per_agent"], np.float32)
        self._prog.ck(self._prog._L.ss_prog_prop_download(self._prog._h, aid, ptr(out)))
        return out

    def create_agent_of_breed(self, breed: Breed, add_to_space=True, rank: int = None, **kwargs) -> int:
        if breed not in self._breeds:
            raise ValueError(f"Fatal: unregistered breed {breed.name}")
        for k in kwargs:
            if k not in self._agent_values:
                raise ValueError(f"{k} not a property of any breed")
        agent_id = len(self._agent_breed)
        self._agent_breed.append(breed._breedidx)
        for name in self._prop_names[2:]:
            self._agent_values[name].append(kwargs.get(name, copy(self._prop_defaults[name])))
        if add_to_space:
            self._space.add_agent(agent_id)
        return agent_id

    @property
    def num_agents(self):
        return len(self._agent_breed)

    def register_global_property(self, property_name: str, value) -> int:
        idx = len(self._global_tensors)
        self._global_tensors.append(np.atleast_1d(np.asarray(value, dtype=np.float64)).copy())
        self._global_names[property_name] = idx
        self._globals_dirty = True
        return idx

    def set_global_property_value(self, property_name: str, value) -> None:
        idx = self._global_names[property_name]
        self._global_tensors[idx][:] = np.atleast_1d(np.asarray(value, dtype=np.float64))
        self._globals_dirty = True

    def get_global_property_value(self, property_name: str):
        arr = self._global_tensors[self._global_names[property_name]]
        return float(arr.flat[0]) if arr.size == 1 else arr.copy()

    @property
    def globals(self):
        return {n: self._global_tensors[i] for n, i in self._global_names.items()}

    def set_seed(self, seed: int) -> None:
        self._seed = int(seed)
        if self._is_setup:
            self._push_seed()

    def set_agent_logical_id(self, agent_id: int, logical_id: int) -> None:
        self._logical_id_map[int(agent_id)] = int(logical_id)

    # ---- hooks for subclasses (model.py:892-921) ----
    def _get_extra_kernel_config(self) -> dict:
        """Override to extend the wrapper kernels.  Keys:
        'extra_kernel_params': declarations of buffers the kernels see, each
            'name' (a float buffer) or 'type name' (e.g. 'int *spikes'); the
            arrays _prepare_kernel_extras returns fill them, in this order;
        'post_breed_step_code': [(code_lines, once_per_breed[, only_priority])]
            HIP lines run right after a breed's step function (its own-slot
            writes already stored), with `_real_idx`, `thread_local_tick`,
            `breed_id`, `agent_index`, the globals, the extra buffers, `a<k>` /
            `write_a<k>` (property k's read and write buffers, k the property's
            index, breed = 0 and locations = 1) and the breed-local arrays by
            name in scope."""
        return {}

    def _prepare_kernel_extras(self, num_local_agents, sync_ticks) -> tuple:
        """Override to provide the extra kernel buffers of a simulate() call
        (numpy arrays, one per 'extra_kernel_params' entry).  After the run
        their device contents are copied back into these arrays."""
        return ()

    def _process_kernel_extras(self) -> None:
        """Override to read the extra buffers after a simulate() call."""
        pass

    def _has_kernel_extras(self):
        return type(self)._get_extra_kernel_config is not Model._get_extra_kernel_config

    # ---- property access ----
    def _buffer_index(self, agent_id):
        return int(self._inv[agent_id])

    def get_agent_property_value(self, id: int, property_name: str):
        if property_name == "locations":
            return list(self._space.get_location(id))
        if property_name == "breed":
            return self._agent_breed[id]
        if self._is_setup:
            arr = self._device_property(property_name)
            v = arr[self._buffer_index(id)]
            return v.tolist()
        return self._agent_values[property_name][id]

    def set_agent_property_value(self, id: int, property_name: str, value) -> None:
        if property_name not in self._agent_values:
            raise ValueError(f"{property_name} not a property of any breed")
        if self._is_setup:
            arr = self._device_property(property_name).copy()
            arr[self._buffer_index(id)] = value
            self._host_props[property_name] = arr
            self._dirty_props.add(property_name)
            self._cache[property_name] = arr
        else:
            self._agent_values[property_name][id] = value

    def _device_property(self, name):
        if name in self._cache:
            return self._cache[name]
        if self._mode == "fused" and name == self._fused.state:
            st = self._replica.states()
            arr = st.astype(_DTYPES[self._prop_dtype[name]][1])
        elif self._mode == "fused":
            arr = self._host_props[name]
        else:
            j = self._prop_index[name]
            arr = np.empty_like(self._host_props[name])
            self._prog.ck(self._prog._L.ss_prog_prop_download(self._prog._h, j, ptr(arr)))
        self._cache[name] = arr
        return arr

    def _breed_named(self, breed_name):
        for b in self._breeds:
            if b.name == breed_name:
                return b
        raise KeyError(breed_name)

    def get_breed_data(self, breed_name, property_name, local=False):
        """As model.py:464-488 on one rank: the property's rows of the breed's
        agents, in buffer order (agents sorted by breed)."""
        if not self._is_setup:
            raise RuntimeError("call setup() first")
        start, count = self._breed_ranges.get(self._breed_named(breed_name)._breedidx, (0, 0))
        if property_name not in self._agent_values:
            raise KeyError(property_name)
        if count == 0:
            return np.empty((0,), np.float32)
        return np.array(self._device_property(property_name)[start:start + count])

    def get_breed_agent_ids(self, breed_name, local=False):
        """As model.py:490-515 on one rank: the breed's agent ids in buffer
        order, float32 as the reference stores them (model.py:1085-1087)."""
        if not self._is_setup:
            raise RuntimeError("call setup() first")
        start, count = self._breed_ranges.get(self._breed_named(breed_name)._breedidx, (0, 0))
        if count == 0:
            return np.empty((0,), np.float32)
        return np.asarray(self._order[start:start + count]).astype(np.float32)

    # ---- setup ----
    def _priority_groups(self):
        prios = sorted({p for b in self._breeds for p in b.step_funcs})
        groups = []
        for p in prios:
            groups.append({b._breedidx: b.step_funcs[p][0] for b in self._breeds if p in b.step_funcs})
        return prios, groups

    def _property_tensor(self, name, order):
        dtype = _DTYPES[self._prop_dtype[name]][1]
        vals = self._agent_values[name]
        is_row = any(isinstance(v, (list, tuple, np.ndarray)) for v in vals) or isinstance(
            self._prop_defaults[name], (list, tuple, np.ndarray))
        if not is_row:
            return np.asarray([vals[a] for a in order], dtype=dtype), 1
        md = self._prop_maxdims.get(name)
        width = max([len(v) for v in vals if isinstance(v, (list, tuple, np.ndarray))] + [1])
        if md:
            width = max(width, int(np.prod(md)))
        pad = np.nan if np.issubdtype(dtype, np.floating) else 0
        out = np.full((len(order), width), pad, dtype=dtype)
        for b, a in enumerate(order):
            v = vals[a]
            if isinstance(v, (list, tuple, np.ndarray)):
                v = np.asarray(v).ravel()
                out[b, :len(v)] = v
            elif v is not None:
                out[b, 0] = v
        return out, width

    def _csr(self, order, inv):
        """CSR over buffer indices.  Each agent's row lists its neighbours in
        the order its container iterates them (a set's order, or insertion
        order with NetworkSpace(ordered=True); model.py:1143-1148,
        internal_utils.py:111-164).  Edges added with connect_agents_bulk
        follow, in the order given, each undirected edge (u, v) adding v to
        u's row and then u to v's row as connect_agents would; a neighbour
        already in the row is not added again."""
        rows = self._space.neighbor_rows()
        n = len(order)
        src, dst = [], []
        for b, a in enumerate(order):
            nb = rows[a] if a < len(rows) else ()
            if nb:
                src.append(np.full(len(nb), b, np.int64))
                dst.append(inv[np.fromiter(nb, np.int64, len(nb))])
        if not self._space.has_bulk_edges():
            counts = np.zeros(n, np.int64)
            for b, a in enumerate(order):
                nb = rows[a] if a < len(rows) else ()
                counts[b] = len(nb)
            rp = np.zeros(n + 1, np.int64)
            np.cumsum(counts, out=rp[1:])
            ci = np.concatenate(dst).astype(np.int32) if dst else np.zeros(0, np.int32)
            return rp, ci
        for u, v, d in zip(self._space._bulk_u, self._space._bulk_v, self._space._bulk_directed):
            u, v = inv[u], inv[v]
            if d:
                src.append(u)
                dst.append(v)
            else:
                src.append(np.stack([u, v], axis=1).ravel())
                dst.append(np.stack([v, u], axis=1).ravel())
        s = np.concatenate(src) if src else np.zeros(0, np.int64)
        t = np.concatenate(dst) if dst else np.zeros(0, np.int64)
        _, first = np.unique(s * max(n, 1) + t, return_index=True)   # first occurrence of every (row, neighbour)
        keep = np.sort(first)
        keep = keep[np.argsort(s[keep], kind="stable")]
        rp = np.zeros(n + 1, np.int64)
        np.cumsum(np.bincount(s[keep], minlength=n), out=rp[1:])
        return rp, t[keep].astype(np.int32)

    def _fused_spec(self, groups):
        if not self._use_fused or len(groups) != 1 or len(groups[0]) != 1:
            return None
        if self._breed_local_arrays or self._has_kernel_extras():
            return None
        (bidx, fn), = groups[0].items()
        spec = getattr(fn, "fused", None)
        if spec is None or any(b != bidx for b in self._agent_breed):
            return None
        names = [spec.state, spec.beta, spec.gamma] + ([spec.sigma] if spec.kind == "seir" else [])
        if spec.state not in self._agent_values or any(g not in self._global_names for g in names[1:]):
            return None
        return spec

    def _host_layout(self):
        """The host half of setup: breed order (agent.py:438-473), property
        tensors in that order, CSR over buffer indices, logical ids and the
        priority groups.  Touches no device."""
        n = self.num_agents
        order = np.argsort(np.asarray(self._agent_breed, np.int64), kind="stable")
        inv = np.empty(n, np.int64)
        inv[order] = np.arange(n)
        self._order, self._inv = order, inv
        bs = np.asarray(self._agent_breed, np.int64)[order]
        self._breed_ranges = {}
        for b in self._breeds:
            idx = np.nonzero(bs == b._breedidx)[0]
            if idx.size:
                self._breed_ranges[b._breedidx] = (int(idx[0]), int(idx.size))
        self._host_props, self._prop_width = {}, {}
        for name in self._prop_names[2:]:
            self._host_props[name], self._prop_width[name] = self._property_tensor(name, order)
        logical = np.array([self._logical_id_map.get(int(a), int(a)) for a in order], np.int64)
        rp, ci = self._csr(order, inv)
        self._prios, groups = self._priority_groups()
        return order, rp, ci, logical, groups

    def setup(self, use_gpu: bool = True, skip_priority_barriers=False) -> None:
        """As model.py:682-886.  `use_gpu` is accepted and ignored (the reference
        stores it and never reads it, model.py:696).  Every priority group is its
        own launch, so groups are always separated by a grid-wide barrier;
        without barriers (`skip_priority_barriers`) the reference leaves the
        order of groups within a tick unspecified, and this is one of its
        allowed outcomes."""
        self._skip_priority_barriers = skip_priority_barriers
        self.tick = 0
        order, rp, ci, logical, groups = self._host_layout()
        if self._engine is None:
            self._engine = Engine(self._device)
        spec = self._fused_spec(groups)
        if spec is not None:
            self._setup_fused(spec, rp, ci, logical)
        else:
            self._setup_rtc(groups, rp, ci, order, logical)
        self._is_setup = True
        self._cache = {}

    def _compartment(self):
        spec = self._fused
        g = self.get_global_property_value
        if spec.kind == "seir":
            return seir_model(g(spec.beta), g(spec.sigma), g(spec.gamma))
        return sir_model(g(spec.beta), g(spec.gamma))

    def _setup_fused(self, spec, rp, ci, logical):
        self._mode = "fused"
        self._fused = spec
        st = self._host_props[spec.state]
        if st.ndim != 1 or not np.all(np.mod(st, 1) == 0):
            raise ValueError(f"the fused compartment step needs integer scalar states in '{spec.state}'")
        graph = graphs.CSRGraph(np.asarray(rp, np.int64), np.asarray(ci, np.int32))
        if self._replica is not None:
            self._replica.remove()
        self._replica = Replica(self._engine, graph, st.astype(np.int8), self._compartment(),
                                self._seed & 0xFFFFFFFF, rng_key_id=logical)
        self._globals_dirty = False

    def _ctype(self, name):
        return _DTYPES[self._prop_dtype[name]][0]

    def _no_double_buffer(self):
        return set().union(*(b.no_double_buffer_props for b in self._breeds)) if self._breeds else set()

    def generate_source(self, groups=None):
        """HIP source of the step kernels: the breeds' step functions plus one
        `ss_prio_<p>` wrapper per priority group that calls each breed's step
        function with the reference's argument order (model.py:2302-2337):
        tick, agent_index, globals, agent_ids, breeds, locations, properties
        in registration order, then `array, idx_map` per breed-local array."""
        if groups is None:
            _, groups = self._priority_groups()
        lines = []
        # property arrays below 2 GiB: Prop indexes them with 32-bit byte offsets
        props = getattr(self, "_host_props", None) or {}
        if props and all(np.asarray(a).nbytes < (1 << 31) for a in props.values()):
            lines.append("#define SS_RT_OFF32 1")
        lines += ['#include "sagesim_rt.h"', ""]
        seen = set()           # one copy of each source text (several functions may share a file)
        for grp in groups:
            for fn in grp.values():
                if fn.source not in seen:
                    lines.append(fn.source)
                    seen.add(fn.source)
        props = self._prop_names[2:]
        nodb = self._no_double_buffer()
        gdecl, gargs, off = [], [], 0
        for gi, arr in enumerate(self._global_tensors):
            if arr.size == 1:                                  # scalars are passed by value (model.py:2229-2234)
                gdecl.append(f"    const double G{gi} = A->globals[{off}];")
            else:
                gdecl.append(f"    const GlobalArray G{gi}{{A->globals + {off}, {arr.size}}};")
            gargs.append(f"G{gi}")
            off += arr.size
        pdecl, pargs, flush = [], [], []
        for j, name in enumerate(props):
            ct = self._ctype(name)
            r, w = f"(const {ct} *)A->prop_r[{j}]", f"({ct} *)A->prop_w[{j}]"
            if self._prop_width.get(name, 1) == 1:
                if name in nodb:
                    pdecl.append(f"    const Prop<{ct}> P{j}{{{r}, {w}, -1, nullptr}};")
                else:
                    pdecl.append(f"    OwnSlot<{ct}> O{j}{{}};")
                    pdecl.append(f"    const Prop<{ct}> P{j}{{{r}, {w}, i, &O{j}}};")
                    flush.append(f"ss_flush(P{j}, O{j});")
            else:
                pdecl.append(f"    const PropRows<{ct}> P{j}{{{r}, {w}, {self._prop_width[name]}}};")
            pdecl.append(f"    const {ct} *const a{j + 2} = {r}; {ct} *const write_a{j + 2} = {w};")
            pdecl.append(f"    (void)a{j + 2}; (void)write_a{j + 2};")
            pargs.append(f"P{j}")
        bdecl, bargs = [], []
        for k, bla in enumerate(self._breed_local_arrays):
            aid, xid, rows = self._bla_ids[k] if hasattr(self, "_bla_ids") else (0, 0, 1)
            shape = bla["shape_per_agent"]
            r, w = f"(const float *)A->prop_r[{aid}]", f"(float *)A->prop_w[{aid}]"
            if len(shape) == 1:
                bdecl.append(f"    const BlaRows1<float> B{k}{{{r}, {w}, {rows}, {shape[0]}}};")
            else:
                bdecl.append(f"    const BlaRows2<float> B{k}{{{r}, {w}, {rows}, {shape[0]}, {shape[1]}}};")
            bdecl.append(f"    const IdxMap B{k}_idx{{(const int *)A->prop_r[{xid}]}};")
            bdecl.append(f"    const auto &{bla['name']} = B{k}; const auto &{bla['name']}_idx = B{k}_idx;")
            bdecl.append(f"    (void){bla['name']}; (void){bla['name']}_idx;")
            bargs += [f"B{k}", f"B{k}_idx"]
        cfg = self._get_extra_kernel_config() or {}
        xdecl = []
        for k, decl in enumerate(cfg.get("extra_kernel_params", [])):
            ty, nm = _split_decl(decl)
            xdecl.append(f"    {ty} const {nm} = ({ty})A->extra[{k}]; (void){nm};")
        post = cfg.get("post_breed_step_code", [])
        seen_breeds = set()
        names = []
        for p, grp in enumerate(groups):
            kname = f"ss_prio_{p}"
            names.append(kname)
            lines.append(f'extern "C" __global__ void {kname}(const SSProgArgs *__restrict__ A, long long tick, '
                         f"long long start, long long count, const int *__restrict__ order) {{")
            lines.append("    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;")
            lines.append("    if (k >= count) return;")
            lines.append("    const long long i = order ? (long long)order[k] : start + k;")
            lines.append("    const int breed = (int)A->breeds[i];")
            lines.append("    const AgentIndex agent_index{i, A->logical_ids[i], A->seed, A->logical_ids64[i]};")
            lines.append("    const long long _real_idx = i, thread_local_tick = tick; const int breed_id = breed;")
            lines.append("    (void)_real_idx; (void)thread_local_tick; (void)breed_id;")
            lines.append("    const FloatArray agent_ids{A->agent_ids}, breeds{A->breeds};")
            lines.append("    const Locations locations{A->row_ptr, A->col_idx};")
            lines += gdecl + pdecl + bdecl + xdecl
            for b, fn in grp.items():
                args = ", ".join(["tick", "agent_index"] + gargs + ["agent_ids", "breeds", "locations"] + pargs
                                 + bargs)
                body = [f"{fn.name}({args});"] + flush
                for entry in post:
                    code_lines, once = entry[0], entry[1]
                    only = entry[2] if len(entry) > 2 else None
                    if only is not None and p != only:
                        continue
                    if once and b in seen_breeds:
                        continue
                    body += list(code_lines)
                    if once:
                        seen_breeds.add(b)
                lines.append(f"    if (breed == {b}) {{")
                lines += [f"        {x}" for x in body]
                lines.append("        return;")
                lines.append("    }")
            lines.append("}")
            lines.append("")
        return "\n".join(lines), names

    def _setup_rtc(self, groups, rp, ci, order, logical):
        self._mode = "rtc"
        n = len(order)
        if self._prog is not None:
            self._prog.close()
        prog = _Program(self._engine, n)
        L, h = prog._L, prog._h
        prog.ck(L.ss_prog_set_csr(h, ptr(np.ascontiguousarray(rp, np.int64)),
                                  ptr(np.ascontiguousarray(ci, np.int32)), int(len(ci))))
        aids = np.ascontiguousarray(order.astype(np.float32))
        lids = np.ascontiguousarray(logical.astype(np.float32))
        breeds = np.ascontiguousarray(np.asarray(self._agent_breed, np.float32)[order])
        prog.ck(L.ss_prog_set_ids(h, ptr(aids), ptr(lids), ptr(np.ascontiguousarray(logical)), ptr(breeds)))
        nodb = self._no_double_buffer()
        self._prop_index = {}
        for name in self._prop_names[2:]:
            arr = np.ascontiguousarray(self._host_props[name])
            pid = ctypes.c_int32()
            prog.ck(L.ss_prog_add_prop(h, arr.dtype.itemsize, self._prop_width[name], 0 if name in nodb else 1,
                                       ptr(arr), ctypes.byref(pid)))
            self._prop_index[name] = pid.value
        # breed-local arrays: zeroed rows of the breed's agents, and an index
        # map buffer index -> row, -1 for other agents (model.py:1215-1270)
        self._bla_ids = []
        for bla in self._breed_local_arrays:
            start, count = self._breed_ranges.get(bla["breed"]._breedidx, (0, 0))
            rows = max(count, 1)
            aid, xid = ctypes.c_int32(), ctypes.c_int32()
            prog.ck(L.ss_prog_add_array(h, 4, rows, int(np.prod(bla["shape_per_agent"])),
                                        0 if bla["name"] in nodb else 1, None, ctypes.byref(aid)))
            idx_map = np.full(max(n, 1), -1, np.int32)
            idx_map[start:start + count] = np.arange(count, dtype=np.int32)
            prog.ck(L.ss_prog_add_array(h, 4, max(n, 1), 1, 0, ptr(idx_map), ctypes.byref(xid)))
            self._bla_ids.append((aid.value, xid.value, rows))
        self._prog = prog
        self._push_globals()
        src, names = self.generate_source(groups)
        if self._step_function_file_path:
            with open(os.path.abspath(self._step_function_file_path), "w") as f:
                f.write(src)
        log = ctypes.create_string_buffer(1 << 16)
        rc = L.ss_prog_compile(h, src.encode(), "\n".join(names).encode(), log, len(log))
        if rc:
            raise NativeError(rc, L.ss_prog_last_error(h).decode())
        self._push_seed()
        # schedule: priority group p covers the breed-sorted range of its breeds
        # (model.py:1051-1062); within it the long rows (see _LONG_ROW_FACTOR)
        # start first, longest first, and the rest follow in buffer order, so
        # that most threads read their agent's fields from consecutive slots
        deg = np.diff(np.asarray(rp, np.int64))
        kern, start, count, segs = [], [], [], []
        for p, grp in enumerate(groups):
            lo, hi = n, 0
            for b in grp:
                if b in self._breed_ranges:
                    s0, c0 = self._breed_ranges[b]
                    lo, hi = min(lo, s0), max(hi, s0 + c0)
            if hi <= lo:
                lo = hi = 0
            kern.append(p)
            start.append(lo)
            count.append(hi - lo)
            d = deg[lo:hi]
            cut = max(_LONG_ROW_MIN, _LONG_ROW_FACTOR * float(d.mean())) if d.size else 0.0
            long_rows = np.nonzero(d >= cut)[0]
            long_rows = long_rows[np.argsort(-d[long_rows], kind="stable")]
            rest = np.nonzero(d < cut)[0]
            segs.append(lo + np.concatenate([long_rows, rest]))
        k = np.asarray(kern, np.int32)
        s = np.asarray(start, np.int64)
        c = np.asarray(count, np.int64)
        prog.ck(L.ss_prog_set_schedule(h, len(k), ptr(k), ptr(s), ptr(c)))
        ordr = np.ascontiguousarray(np.concatenate(segs) if segs else np.zeros(0), np.int32)
        prog.ck(L.ss_prog_set_order(h, ptr(ordr), ordr.shape[0]))

    def _push_globals(self):
        if self._mode == "fused":
            self._engine.set_model(self._replica.rep, self._compartment())
        elif self._prog is not None:
            g = (np.concatenate([a.ravel() for a in self._global_tensors]) if self._global_tensors
                 else np.zeros(0))
            g = np.ascontiguousarray(g, np.float64)
            self._prog.ck(self._prog._L.ss_prog_set_globals(self._prog._h, ptr(g), g.shape[0]))
        self._globals_dirty = False

    def _push_seed(self):
        if self._mode == "fused":
            self._engine.set_seed(self._replica.rep, self._seed & 0xFFFFFFFF)
        elif self._prog is not None:
            self._prog.ck(self._prog._L.ss_prog_set_seed(self._prog._h, self._seed))

    # ---- run ----
    def _upload_kernel_extras(self, ticks):
        cfg = self._get_extra_kernel_config() or {}
        params = cfg.get("extra_kernel_params", [])
        got = tuple(self._prepare_kernel_extras(self.num_agents, int(ticks)) or ())
        if len(got) != len(params):
            raise ValueError(f"_prepare_kernel_extras returned {len(got)} buffers for {len(params)} parameters")
        self._kernel_extra_arrays = []
        for k, a in enumerate(got):
            arr = a if isinstance(a, np.ndarray) and a.flags.c_contiguous else np.ascontiguousarray(a)
            self._prog.ck(self._prog._L.ss_prog_set_extra(self._prog._h, k, arr.nbytes, ptr(arr)))
            self._kernel_extra_arrays.append((a, arr))

    def _download_kernel_extras(self):
        for k, (orig, arr) in enumerate(self._kernel_extra_arrays):
            if arr.nbytes:
                self._prog.ck(self._prog._L.ss_prog_get_extra(self._prog._h, k, arr.nbytes, ptr(arr)))
            if isinstance(orig, np.ndarray) and orig is not arr and orig.flags.writeable:
                orig[...] = arr.reshape(orig.shape)

    def simulate(self, ticks: int, sync_workers_every_n_ticks: int = 1) -> None:
        """As model.py:961-998 on one rank: all ticks run on the device; with
        verbose_timing, _tick_timings gets the call's device time."""
        if not self._is_setup:
            raise RuntimeError("call setup() before simulate()")
        self._last_simulate_ticks = int(ticks)
        self._tick_timings = []
        self._flush_dirty()
        self._cache = {}
        extras = self._mode == "rtc" and self._has_kernel_extras()
        if extras:
            self._upload_kernel_extras(ticks)
        if self._verbose_timing:
            self._engine.timer_start()
        if self._mode == "fused":
            self._engine.simulate(int(ticks))
            self._engine.sync()
        else:
            self._prog.ck(self._prog._L.ss_prog_simulate(self._prog._h, self.tick, int(ticks)))
        if self._verbose_timing:
            self._tick_timings.append({"ticks": int(ticks), "gpu_kernel_ms": self._engine.timer_stop(),
                                       "mode": self._mode})
        if extras:
            self._download_kernel_extras()
            self._process_kernel_extras()
        self.tick += int(ticks)

    def reset(self) -> None:
        """Bring device properties back to the host lists and free the device
        state (model.py:923-958)."""
        if not self._is_setup:
            return
        for name in self._prop_names[2:]:
            arr = self._device_property(name)
            for b, a in enumerate(self._order):
                v = arr[b]
                self._agent_values[name][a] = v.tolist()
        if self._replica is not None:
            self._replica.remove()
            self._replica = None
        if self._prog is not None:
            self._prog.close()
            self._prog = None
        self._is_setup = False
        self._cache = {}
        self.tick = 0

    # ---- checkpoint (replaces the pickle of model.py:1289-1310) ----
    _CKPT_VERSION = 1

    def _fingerprint(self):
        return {"version": self._CKPT_VERSION, "mode": self._mode, "num_agents": self.num_agents,
                "properties": self._prop_names[2:], "breeds": [b.name for b in self._breeds],
                "breed_local_arrays": [b["name"] for b in self._breed_local_arrays]}

    def save(self, app=None, fpath: str = None) -> None:
        """Checkpoint the model's device state to `fpath` (an .npz written
        with numpy, no pickle): every property, breed-local array, the globals,
        the seed and the tick (model.py:1289-1300 pickles the whole model
        instead).  `app` is the model to save (the reference's argument);
        None or omitted means this model."""
        if isinstance(app, (str, os.PathLike)) and fpath is None:
            app, fpath = None, app
        app = self if app is None else app
        if not app._is_setup:
            raise RuntimeError("save() checkpoints device state: call setup() first")
        app._flush_dirty()
        if app._mode == "rtc":
            size = ctypes.c_int64()
            app._prog.ck(app._prog._L.ss_prog_save(app._prog._h, app.tick, None, 0, ctypes.byref(size)))
            blob = np.zeros(size.value, np.uint8)
            app._prog.ck(app._prog._L.ss_prog_save(app._prog._h, app.tick, ptr(blob), blob.shape[0],
                                                   ctypes.byref(size)))
        else:
            blob = app._engine.replica_save(app._replica.rep)
        g = (np.concatenate([a.ravel() for a in app._global_tensors]) if app._global_tensors else np.zeros(0))
        with open(fpath, "wb") as f:
            np.savez(f, state=blob, globals=np.asarray(g, np.float64),
                     meta=np.frombuffer(json.dumps({**app._fingerprint(), "tick": app.tick,
                                                    "seed": app._seed}).encode(), np.uint8))

    def load(self, fpath: str) -> "Model":
        """Restore a checkpoint written by save() into this model, which must
        be built and set up the same way (same agents, properties, network and
        step functions); returns the model.  The reference unpickles a whole
        model instead (model.py:1302-1310); here the file holds state, not code."""
        if not self._is_setup:
            raise RuntimeError("load() restores device state: call setup() first")
        with np.load(fpath, allow_pickle=False) as z:
            meta = json.loads(bytes(z["meta"]).decode())
            blob = np.ascontiguousarray(z["state"])
            gl = np.asarray(z["globals"], np.float64)
        want = self._fingerprint()
        for key in want:
            if meta.get(key) != want[key]:
                raise ValueError(f"checkpoint {key} {meta.get(key)!r} does not match the model's {want[key]!r}")
        off = 0
        for arr in self._global_tensors:
            arr[:] = gl[off:off + arr.size].reshape(arr.shape)
            off += arr.size
        self._seed = int(meta["seed"])
        self._dirty_props.clear()
        if self._mode == "rtc":
            t = ctypes.c_int64()
            self._prog.ck(self._prog._L.ss_prog_restore(self._prog._h, ptr(blob), blob.shape[0], ctypes.byref(t)))
            self.tick = t.value
        else:
            self.tick = self._engine.replica_restore(self._replica.rep, blob)
        self._globals_dirty = False
        self._cache = {}
        return self

    def _flush_dirty(self):
        if self._globals_dirty:
            self._push_globals()
        for name in list(self._dirty_props):
            arr = np.ascontiguousarray(self._host_props[name])
            if self._mode == "fused":
                if name == self._fused.state:
                    self._replica.set_states(arr.astype(np.int8))
            else:
                self._prog.ck(self._prog._L.ss_prog_prop_upload(self._prog._h, self._prop_index[name], ptr(arr)))
        self._dirty_props.clear()

    @property
    def mode(self):
        """'fused' (built-in kernel) or 'rtc' (compiled step functions) after setup."""
        return self._mode
