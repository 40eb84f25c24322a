"""Breed-local arrays (model.py:620-651, step arguments :1994-2007,
write-back :2391-2406) and the kernel-extras hooks (model.py:892-921,
post-step code :2340-2353) on the hipRTC path, against brute-force
restatements written here and the oracle."""
import numpy as np
import pytest

from oracle import oracle as O
from sagesim_amd import Breed, Model, NetworkSpace, StepFunction, graphs
from sagesim_amd.steps import reference_sir_step

pytestmark = pytest.mark.gpu

f32 = np.float32


def say(*a):
    print("[case]", *a, flush=True)


_BLA_SRC = r"""
__device__ void cell_step(auto tick, auto agent_index, auto agent_ids, auto breeds, auto locations, auto level,
                          auto memory, auto memory_idx, auto grid, auto grid_idx, auto scratchpad,
                          auto scratchpad_idx) {
    const int row = memory_idx[agent_index];
    auto mem = memory[row];
    for (int j = len(mem) - 1; j > 0; --j) mem[j] = mem[j - 1];
    mem[0] = level[agent_index];
    memory[row][1] += 0.25f;
    auto sp = scratchpad[scratchpad_idx[agent_index]];
    sp[0] = sp[0] + 1.0f;
    sp[1] = sp[0] * 2.0f;
    level[agent_index] = level[agent_index] + mem[3];
}

__device__ void bus_step(auto tick, auto agent_index, auto agent_ids, auto breeds, auto locations, auto level,
                         auto memory, auto memory_idx, auto grid, auto grid_idx, auto scratchpad,
                         auto scratchpad_idx) {
    const int row = grid_idx[agent_index];
    for (int a = 0; a < len(grid[row]); ++a)
        for (int b = 0; b < len(grid[row][a]); ++b)
            grid[row][a][b] = grid[row][a][b] + (float)(a * 3 + b) * 0.5f + (float)tick;
    auto nb = locations[agent_index];
    float s = 0.0f;
    for (int j = 0; j < len(nb); ++j) {
        const int r = memory_idx[nb[j]];
        if (r >= 0) s += memory[r][0];
    }
    level[agent_index] = s;
}
"""


def test_breed_local_arrays_against_brute_force():
    say("breed-local arrays: (4,), (2, 3), a no_double_buffer (3,)")
    n, ticks = 300, 5
    rng = np.random.default_rng(8)
    cell, bus = Breed("Cell"), Breed("Bus")
    for b in (cell, bus):
        b.register_property("level", 0.0)
    cell.register_step_func(StepFunction("cell_step", _BLA_SRC), priority=0, no_double_buffer=["scratchpad"])
    bus.register_step_func(StepFunction("bus_step", _BLA_SRC), priority=1)
    m = Model(NetworkSpace(ordered=True))
    m.register_breed(cell)
    m.register_breed(bus)
    is_cell = rng.random(n) < 0.7
    level0 = (rng.random(n) * 4).astype(f32)
    for i in range(n):
        m.create_agent_of_breed(cell if is_cell[i] else bus, level=float(level0[i]))
    assert m.register_breed_local_array("memory", cell, (4,)) == 0
    assert m.register_breed_local_array("grid", bus, (2, 3)) == 1
    m.register_breed_local_array("scratchpad", cell, (3,))
    with pytest.raises(ValueError):
        m.register_breed_local_array("memory", cell, (2,))
    for _ in range(1500):
        a, c = rng.integers(0, n, 2)
        m.get_space().connect_agents(int(a), int(c))
    rows = [list(m.get_space().get_location(i)) for i in range(n)]
    cells, buses = np.nonzero(is_cell)[0], np.nonzero(~is_cell)[0]
    crow = {int(a): k for k, a in enumerate(cells)}          # row of an agent: rank within its breed
    brow = {int(a): k for k, a in enumerate(buses)}
    mem = np.zeros((len(cells), 4), f32)
    grid = np.zeros((len(buses), 2, 3), f32)
    sp = np.zeros((len(cells), 3), f32)
    lev = level0.copy()
    for t in range(ticks):
        W_mem, W_grid, W_lev = mem.copy(), grid.copy(), lev.copy()
        for a in cells:                                   # priority 0
            r = crow[int(a)]
            for j in range(3, 0, -1):
                W_mem[r, j] = mem[r, j - 1]
            W_mem[r, 0] = lev[a]
            W_mem[r, 1] = f32(mem[r, 1] + f32(0.25))
            sp[r, 0] = f32(sp[r, 0] + f32(1.0))           # in place: the next line sees it
            sp[r, 1] = f32(sp[r, 0] * f32(2.0))
            W_lev[a] = f32(lev[a] + mem[r, 3])
        for a in buses:                                   # priority 1 reads the tick-start buffers
            r = brow[int(a)]
            for x in range(2):
                for y in range(3):
                    W_grid[r, x, y] = f32(f32(grid[r, x, y] + f32(f32(x * 3 + y) * f32(0.5))) + f32(t))
            s = f32(0.0)
            for j in rows[a]:
                if int(j) in crow:
                    s = f32(s + mem[crow[int(j)], 0])
            W_lev[a] = s
        mem, grid, lev = W_mem, W_grid, W_lev
    m.setup()
    assert m.mode == "rtc"
    m.simulate(ticks)
    assert np.array_equal(m.get_breed_local_array("memory"), mem)
    assert np.array_equal(m.get_breed_local_array("grid"), grid)
    assert np.array_equal(m.get_breed_local_array("scratchpad"), sp)
    got = np.asarray(m._device_property("level"), f32)[m._inv]
    assert np.array_equal(got, lev)
    assert m.get_breed_local_array("grid").shape == (len(buses), 2, 3)
    assert (mem != 0).any() and (grid != 0).any()


class _SpikeModel(Model):
    """The reference SIR step plus kernel extras: an int buffer counting, per
    agent, the ticks it ends infectious (post-step code after every step),
    and a float buffer counting the steps each breed ran (once per breed)."""

    def __init__(self):
        super().__init__(NetworkSpace())
        self.processed = 0

    def _get_extra_kernel_config(self):
        return {"extra_kernel_params": ["int *spikes", "runs"],
                "post_breed_step_code": [
                    (["if ((int)write_a2[_real_idx] == 2) spikes[_real_idx] += 1;"], False),
                    (["atomicAdd(&runs[breed_id], 1.0f);"], True),
                    (["atomicAdd(&runs[3], 1.0f);"], False, 5),           # no priority index 5: never emitted
                ]}

    def _prepare_kernel_extras(self, num_local_agents, sync_ticks):
        self.spikes = np.zeros(num_local_agents, np.int32)
        self.runs = np.zeros(4, np.float32)
        self.sync_ticks = sync_ticks
        return (self.spikes, self.runs)

    def _process_kernel_extras(self):
        self.processed += 1


def test_kernel_extras_post_step_code():
    say("kernel extras")
    g = graphs.erdos_renyi(2000, 0.004, 5)
    b = Breed("SIR")
    b.register_property("state", 1)
    b.register_step_func(reference_sir_step(), priority=0)
    m = _SpikeModel()
    m.register_breed(b)
    m.register_global_property("p_infection", 0.3)
    m.register_global_property("p_recovery", 0.1)
    m.set_seed(91)
    init = np.where(np.random.default_rng(3).random(g.n) < 0.02, 2, 1)
    for i in range(g.n):
        m.create_agent_of_breed(b, state=int(init[i]))
    for i in range(g.n):
        for j in g.col_idx[g.row_ptr[i]:g.row_ptr[i + 1]]:
            m.get_space().connect_agents(i, int(j), directed=True)
    m.setup()
    assert m.mode == "rtc"
    src, _ = m.generate_source()
    assert "runs[3]" not in src and src.count("atomicAdd(&runs[breed_id]") == 1
    ticks = 12
    m.simulate(ticks)
    assert m.processed == 1 and m.sync_ticks == ticks
    st = init.astype(f32)
    spikes = np.zeros(g.n, np.int32)
    for t in range(ticks):
        st = O.ref_sir_run(g.row_ptr, g.col_idx, np.arange(g.n, dtype=f32), st, 91, 1, 0.3, 0.1, tick0=t)
        spikes += (st == 2)
    assert np.array_equal(m.spikes[m._inv], spikes)
    assert m.runs.tolist() == [float(ticks * g.n), 0.0, 0.0, 0.0]
    got = np.array([m.get_agent_property_value(i, "state") for i in range(g.n)], f32)
    assert np.array_equal(got, st)
