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
import ctypes
from copy import copy
from math import nan
import os

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
    Lists keep their order; sets are written in ascending order."""
    lengths = [len(r) if isinstance(r, (list, tuple, set, np.ndarray)) else 0 for r in ragged_list]
    offsets = np.zeros(len(lengths) + 1, np.int64)
    np.cumsum(lengths, out=offsets[1:])
    values = np.empty(int(offsets[-1]), np.int32)
    pos = 0
    for r in ragged_list:
        if isinstance(r, set):
            r = sorted(r)
        n = len(r) if isinstance(r, (list, tuple, np.ndarray)) else 0
        if n:
            values[pos:pos + n] = np.asarray(r, np.int64)
            pos += n
    return offsets, values


def build_csr_values_only(ragged_list, offsets):
    """Values of a CSR whose offsets are known (internal_utils.py:167-186)."""
    return build_csr_from_ragged(ragged_list)[1][: int(offsets[-1])]


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
        self._agent_values = {}          # property -> list of per-agent values (creation order)
        self._agent_breed = []
        self._global_tensors = []
        self._global_names = {}
        self._seed = int(np.random.randint(0, 2 ** 31))
        self._logical_id_map = {}
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
            if name not in self._prop_defaults:
                self._prop_names.append(name)
                self._agent_values[name] = []
                self._prop_dtype[name] = breed._prop2dtype.get(name, "float32")
                self._prop_maxdims[name] = breed._prop2maxdims.get(name)
            self._prop_defaults[name] = default

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

    # ---- property access ----
    def _buffer_index(self, agent_id):
        return int(self._inv[agent_id])

    def get_agent_property_value(self, id: int, property_name: str):
        if property_name == "locations":
            loc = self._space.get_location(id)
            return list(loc) if not isinstance(loc, set) else sorted(loc)
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
        rows = self._space.neighbor_rows()
        n = len(order)
        if self._space.has_bulk_edges():
            us, vs = [], []
            for a in range(n):
                nb = rows[a] if a < len(rows) else ()
                if nb:
                    us.append(np.full(len(nb), a, np.int64))
                    vs.append(np.fromiter(nb, np.int64, len(nb)))
            for u, v, d in zip(self._space._bulk_u, self._space._bulk_v, self._space._bulk_directed):
                us.append(u)
                vs.append(v)
                if not d:
                    us.append(v)
                    vs.append(u)
            u = np.concatenate(us) if us else np.zeros(0, np.int64)
            v = np.concatenate(vs) if vs else np.zeros(0, np.int64)
            g = graphs.from_edges(n, inv[u], inv[v], directed=True)
            return g.row_ptr, g.col_idx
        ragged = []
        for a in order:
            nb = rows[a] if a < len(rows) else ()
            if isinstance(nb, set):
                nb = sorted(nb)
            ragged.append([int(inv[x]) for x in nb])
        return build_csr_from_ragged(ragged)

    def _fused_spec(self, groups):
        if not self._use_fused or len(groups) != 1 or len(groups[0]) != 1:
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

    def generate_source(self, groups=None):
        """HIP source of the step kernels: the breeds' step functions plus one
        `ss_prio_<p>` wrapper per priority group that calls each breed's step
        function with the reference's argument order."""
        if groups is None:
            _, groups = self._priority_groups()
        lines = ['#include "sagesim_rt.h"', ""]
        seen = set()           # one copy of each source text (several functions may share a file)
        for grp in groups:
            for fn in grp.values():
                if fn.source not in seen:
                    lines.append(fn.source)
                    seen.add(fn.source)
        props = self._prop_names[2:]
        gdecl, gargs, off = [], [], 0
        for gi, arr in enumerate(self._global_tensors):
            if arr.size == 1:                                  # scalars are passed by value (model.py:2229-2234)
                gdecl.append(f"    const double G{gi} = A->globals[{off}];")
            else:
                gdecl.append(f"    const GlobalArray G{gi}{{A->globals + {off}, {arr.size}}};")
            gargs.append(f"G{gi}")
            off += arr.size
        pdecl, pargs = [], []
        for j, name in enumerate(props):
            ct = self._ctype(name)
            if self._prop_width[name] == 1:
                pdecl.append(f"    const Prop<{ct}> P{j}{{(const {ct} *)A->prop_r[{j}], ({ct} *)A->prop_w[{j}]}};")
            else:
                pdecl.append(f"    const PropRows<{ct}> P{j}{{(const {ct} *)A->prop_r[{j}], ({ct} *)A->prop_w[{j}], "
                             f"{self._prop_width[name]}}};")
            pargs.append(f"P{j}")
        names = []
        for p, grp in enumerate(groups):
            kname = f"ss_prio_{p}"
            names.append(kname)
            lines.append(f'extern "C" __global__ void {kname}(const SSProgArgs *__restrict__ A, long long tick, '
                         f"long long start, long long count) {{")
            lines.append("    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;")
            lines.append("    if (k >= count) return;")
            lines.append("    const long long i = start + k;")
            lines.append("    const int breed = (int)A->breeds[i];")
            lines.append("    const AgentIndex agent_index{i, A->logical_ids[i], A->seed, A->logical_ids64[i]};")
            lines.append("    const FloatArray agent_ids{A->agent_ids}, breeds{A->breeds};")
            lines.append("    const Locations locations{A->row_ptr, A->col_idx};")
            lines += gdecl + pdecl
            for b, fn in grp.items():
                args = ", ".join(["tick", "agent_index"] + gargs + ["agent_ids", "breeds", "locations"] + pargs)
                lines.append(f"    if (breed == {b}) {{ {fn.name}({args}); return; }}")
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
        nodb = set().union(*(b.no_double_buffer_props for b in self._breeds)) if self._breeds else set()
        self._prop_index = {}
        for name in self._prop_names[2:]:
            arr = np.ascontiguousarray(self._host_props[name])
            pid = ctypes.c_int32()
            prog.ck(L.ss_prog_add_prop(h, arr.dtype.itemsize, self._prop_width[name], 0 if name in nodb else 1,
                                       ptr(arr), ctypes.byref(pid)))
            self._prop_index[name] = pid.value
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
        bs = np.asarray(self._agent_breed, np.int64)[order]
        kern, start, count = [], [], []
        for p, grp in enumerate(groups):
            lo, hi = n, 0
            for b in grp:
                idx = np.nonzero(bs == b)[0]
                if idx.size:
                    lo, hi = min(lo, int(idx[0])), max(hi, int(idx[-1]) + 1)
            kern.append(p)
            start.append(lo if hi > lo else 0)
            count.append(hi - lo if hi > lo else 0)
        k = np.asarray(kern, np.int32)
        s = np.asarray(start, np.int64)
        c = np.asarray(count, np.int64)
        prog.ck(L.ss_prog_set_schedule(h, len(k), ptr(k), ptr(s), ptr(c)))

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
    def simulate(self, ticks: int, sync_workers_every_n_ticks: int = 1) -> None:
        """As model.py:961-998 on one rank: all ticks run on the device."""
        if not self._is_setup:
            raise RuntimeError("call setup() before simulate()")
        if self._globals_dirty:
            self._push_globals()
        for name in self._dirty_props:
            arr = np.ascontiguousarray(self._host_props[name])
            if self._mode == "fused":
                if name == self._fused.state:
                    self._replica.set_states(arr.astype(np.int8))
            else:
                self._prog.ck(self._prog._L.ss_prog_prop_upload(self._prog._h, self._prop_index[name], ptr(arr)))
        self._dirty_props.clear()
        self._cache = {}
        if self._mode == "fused":
            self._engine.simulate(int(ticks))
            self._engine.sync()
        else:
            self._prog.ck(self._prog._L.ss_prog_simulate(self._prog._h, self.tick, int(ticks)))
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

    @property
    def mode(self):
        """'fused' (built-in kernel) or 'rtc' (compiled step functions) after setup."""
        return self._mode
