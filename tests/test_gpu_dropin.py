"""Drop-in surface of the hipRTC step-function path, each case against a
brute-force restatement of the reference's tick semantics written here
(model.py:2302-2408: priorities in ascending order, every step reading the
tick-start read buffers and writing the write buffers, write-back after the
tick): two breeds created interleaved with steps at three priorities
(agent.py:438-473, model.py:2312-2337), the reference's SIR example with a
ragged NaN-padded 2-D property and a history row (examples/sir/run.py:28-99,
rand_uniform_philox in place of random()), the CSR helpers of utils.py:34-57,
logical ids above 2^24 (model.py:1085-1096), and the code-object cache."""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import philox_py as P
from sagesim_amd import Breed, Model, NetworkSpace, StepFunction, graphs
from sagesim_amd.steps import reference_sir_step

pytestmark = pytest.mark.gpu

f32 = np.float32


def say(*a):
    print("[case]", *a, flush=True)


def _u(seed, tick, agent_id, salt=1):
    """rand_uniform_philox keyed on the float32 logical id (model.py:1934-1951)."""
    return P.rand_uniform_philox(seed, tick, int(f32(agent_id)), salt)


# ---------------- two breeds, three priorities ----------------

_TWO_BREEDS = r"""
__device__ void walker_move(auto tick, auto agent_index, auto p, auto agent_ids, auto breeds, auto locations,
                            auto state, auto energy, auto hits) {
    auto nb = locations[agent_index];
    int k = 0;
    for (int j = 0; j < len(nb); ++j)
        if ((int)state[nb[j]] == 1) ++k;
    const double u = rand_uniform_philox(tick, agent_index, 1);
    if ((int)state[agent_index] == 0 && k > 0 && u < p) state[agent_index] = 1;
    energy[agent_index] += 0.5f * (float)k;
}

__device__ void walker_decay(auto tick, auto agent_index, auto p, auto agent_ids, auto breeds, auto locations,
                             auto state, auto energy, auto hits) {
    energy[agent_index] = energy[agent_index] * 0.9f;
}

__device__ void sensor_step(auto tick, auto agent_index, auto p, auto agent_ids, auto breeds, auto locations,
                            auto state, auto energy, auto hits) {
    auto nb = locations[agent_index];
    int k = 0;
    for (int j = 0; j < len(nb); ++j)
        if ((int)state[nb[j]] == 1) ++k;
    hits[agent_index] += k;
    if (k > 1) state[agent_index] = 2;
}
"""


def _two_breed_model(n, seed, p):
    walker, sensor = Breed("Walker"), Breed("Sensor")
    walker.register_property("state", 0)
    walker.register_property("energy", 1.0)
    sensor.register_property("state", 0)
    sensor.register_property("hits", 0)
    walker.register_step_func(StepFunction("walker_move", _TWO_BREEDS), priority=0)
    walker.register_step_func(StepFunction("walker_decay", _TWO_BREEDS), priority=2)
    sensor.register_step_func(StepFunction("sensor_step", _TWO_BREEDS), priority=1)
    m = Model(NetworkSpace(ordered=True))
    m.register_breed(walker)
    m.register_breed(sensor)
    m.register_global_property("p", p)
    m.set_seed(seed)
    rng = np.random.default_rng(seed)
    kinds = rng.random(n) < 0.6
    state0 = (rng.random(n) < 0.1).astype(int)
    for i in range(n):                       # walkers and sensors interleaved in creation order
        if kinds[i]:
            m.create_agent_of_breed(walker, state=int(state0[i]), energy=float(f32(rng.random() * 3)))
        else:
            m.create_agent_of_breed(sensor, state=int(state0[i]), hits=int(rng.integers(0, 5)))
    g = graphs.erdos_renyi(n, 6.0 / n, seed + 1)
    for i in range(n):
        for j in g.col_idx[g.row_ptr[i]:g.row_ptr[i + 1]]:
            m.get_space().connect_agents(i, int(j), directed=True)
    return m, walker, sensor, kinds


def _two_breed_brute(m, kinds, ticks, p, seed):
    n = len(kinds)
    rows = [list(m.get_space().get_location(i)) for i in range(n)]
    st = np.array([m._agent_values["state"][i] for i in range(n)], f32)
    en = np.array([m._agent_values["energy"][i] for i in range(n)], f32)
    hi = np.array([m._agent_values["hits"][i] for i in range(n)], f32)
    out = []
    for t in range(ticks):
        W = [st.copy(), en.copy(), hi.copy()]
        for i in np.nonzero(kinds)[0]:                  # priority 0: walkers
            k = sum(1 for j in rows[i] if int(st[j]) == 1)
            if int(st[i]) == 0 and k > 0 and _u(seed, t, i) < p:
                W[0][i] = 1
            W[1][i] = f32(en[i] + f32(0.5) * f32(k))
        for i in np.nonzero(~kinds)[0]:                 # priority 1: sensors
            k = sum(1 for j in rows[i] if int(st[j]) == 1)
            W[2][i] = f32(hi[i] + f32(k))
            if k > 1:
                W[0][i] = 2
        for i in np.nonzero(kinds)[0]:                  # priority 2: walkers again, reading the tick start
            W[1][i] = f32(en[i] * f32(0.9))
        st, en, hi = W
        out.append((st.copy(), en.copy(), hi.copy()))
    return out


def test_two_breeds_interleaved_three_priorities():
    say("two breeds, three priorities")
    n, seed, p, ticks = 600, 31, 0.4, 6
    m, walker, sensor, kinds = _two_breed_model(n, seed, p)
    want = _two_breed_brute(m, kinds, ticks, p, seed)
    m.setup()
    assert m.mode == "rtc"
    # buffer order: stable sort by breed (agent.py:438-473)
    assert m._order.tolist() == np.argsort(~kinds, kind="stable").tolist()
    for t in range(ticks):
        m.simulate(1)
        for name, w in zip(("state", "energy", "hits"), want[t]):
            got = np.asarray(m._device_property(name), f32)[m._inv]
            bad = np.nonzero(got != w)[0]
            assert bad.size == 0, (t, name, bad[:5].tolist(), got[bad[:5]].tolist(), w[bad[:5]].tolist())
    assert (want[-1][0] == 2).any() and (want[-1][0] == 1).any()
    # readback by breed (model.py:464-515): buffer order within the breed
    wid = np.nonzero(kinds)[0]
    assert m.get_breed_agent_ids("Walker").tolist() == wid.astype(f32).tolist()
    assert np.array_equal(m.get_breed_data("Walker", "energy"), want[-1][1][wid])
    sid = np.nonzero(~kinds)[0]
    assert np.array_equal(m.get_breed_data("Sensor", "hits"), want[-1][2][sid])
    assert m.get_breed_agent_ids("Sensor").dtype == f32


# ---------------- examples/sir/run.py:28-99 with a ragged 2-D property ----------------

_SIR_EXAMPLE = r"""
__device__ void step_func(auto tick, auto agent_index, auto p_infection, auto p_recovery, auto agent_ids,
                          auto breeds, auto locations, auto state_tensor, auto preventative_measures_tensor,
                          auto state_history_buffer) {
    auto neighbor_indices = locations[agent_index];
    const double rand = rand_uniform_philox(tick, agent_index, 1);
    const float agent_state = state_tensor[agent_index];
    auto agent_preventative_measures = preventative_measures_tensor[agent_index];
    if (agent_state == 2 && rand < p_recovery) {
        state_tensor[agent_index] = 3;
    } else if (agent_state == 1) {
        int i = 0;
        while (i < len(neighbor_indices) && neighbor_indices[i] != -1) {
            const int neighbor_index = neighbor_indices[i];
            const float neighbor_state = state_tensor[neighbor_index];
            auto neighbor_preventative_measures = preventative_measures_tensor[neighbor_index];
            double abs_safety_of_interaction = 0.0;
            for (int n = 0; n < len(agent_preventative_measures); ++n)
                for (int m = 0; m < len(neighbor_preventative_measures); ++m)
                    abs_safety_of_interaction += agent_preventative_measures[n] * neighbor_preventative_measures[m];
            const double normalized_safety_of_interaction =
                abs_safety_of_interaction / (double)(len(agent_preventative_measures) * len(agent_preventative_measures));
            if (neighbor_state == 2 && rand < p_infection * (1 - normalized_safety_of_interaction))
                state_tensor[agent_index] = 2;
            i += 1;
        }
    }
    const long long buffer_idx = tick % len(state_history_buffer[agent_index]);
    state_history_buffer[agent_index][buffer_idx] = state_tensor[agent_index];
}
"""


def test_sir_example_ragged_2d_property_and_history_row():
    say("examples/sir step: NaN-padded 2-D property, history row")
    n, ticks, seed, p_inf, p_rec, width = 500, 8, 77, 0.9, 0.2, 6
    rng = np.random.default_rng(seed)
    b = Breed("SIR")
    b.register_property("state", 1)
    b.register_property("preventative_measures", [0.0] * width)
    b.register_property("state_history_buffer", [])
    b.register_step_func(StepFunction("step_func", _SIR_EXAMPLE), priority=0)
    m = Model(NetworkSpace())
    m.register_breed(b)
    m.register_global_property("p_infection", p_inf)
    m.register_global_property("p_recovery", p_rec)
    m.set_seed(seed)
    state0 = np.where(rng.random(n) < 0.08, 2, 1)
    pm = []
    for i in range(n):
        k = width if rng.random() < 0.9 else int(rng.integers(2, width))    # some rows shorter: NaN padded
        pm.append([float(f32(x)) for x in rng.random(k) * 0.6])
        m.create_agent_of_breed(b, state=int(state0[i]), preventative_measures=pm[i],
                                state_history_buffer=[int(state0[i])] * ticks)
    g = graphs.watts_strogatz(n, 6, 0.2, seed)
    for i in range(n):
        for j in g.col_idx[g.row_ptr[i]:g.row_ptr[i + 1]]:
            m.get_space().connect_agents(i, int(j), directed=True)
    rows = [list(m.get_space().get_location(i)) for i in range(n)]
    PM = np.full((n, width), np.nan, f32)
    for i in range(n):
        PM[i, :len(pm[i])] = pm[i]
    st = state0.astype(f32)
    hist = np.repeat(state0.astype(f32)[:, None], ticks, axis=1)
    for t in range(ticks):
        W = st.copy()
        for i in range(n):
            u = _u(seed, t, i)
            if st[i] == 2 and u < p_rec:
                W[i] = 3
            elif st[i] == 1:
                for j in rows[i]:
                    s = 0.0
                    for a in range(width):
                        for c in range(width):
                            s += float(f32(PM[i, a] * PM[j, c]))
                    norm = s / float(width * width)
                    if st[j] == 2 and u < p_inf * (1 - norm):
                        W[i] = 2
            hist[i, t % ticks] = st[i]
        st = W
    m.setup()
    m.simulate(ticks)
    got = np.asarray(m._device_property("state"), f32)[m._inv]
    assert np.array_equal(got, st), int((got != st).sum())
    got_h = np.asarray(m._device_property("state_history_buffer"), f32)[m._inv]
    assert got_h.shape == (n, ticks) and np.array_equal(got_h, hist)
    got_pm = np.asarray(m._device_property("preventative_measures"), f32)[m._inv]
    assert np.array_equal(np.isnan(got_pm), np.isnan(PM)) and np.array_equal(np.nan_to_num(got_pm), np.nan_to_num(PM))
    assert (st == 2).sum() + (st == 3).sum() > (state0 == 2).sum()


# ---------------- CSR helpers (utils.py:34-57) ----------------

_HELPERS = r"""
__device__ void nb_step(auto tick, auto agent_index, auto agent_ids, auto breeds, auto locations, auto value,
                        auto weight) {
    double acc = 0.0;
    const long long n = get_num_neighbors(agent_index, neighbor_offsets);
    for (long long j = 0; j < n; ++j)
        acc += (double)(j + 1) * (double)weight[get_neighbor(agent_index, j, neighbor_offsets, neighbor_values)];
    double acc2 = 0.0;
    const int n2 = get_num_neighbors(agent_index, locations);
    for (int j = 0; j < n2; ++j) acc2 += (double)weight[get_neighbor(agent_index, j, locations)];
    value[agent_index] = (float)(acc + 0.001 * acc2 + (double)tick + (double)n * 1000.0);
}
"""


def test_neighbor_helpers():
    say("get_num_neighbors / get_neighbor")
    n = 400
    rng = np.random.default_rng(4)
    b = Breed("Node")
    b.register_property("value", 0.0)
    b.register_property("weight", 0.0)
    b.register_step_func(StepFunction("nb_step", _HELPERS), priority=0)
    m = Model(NetworkSpace(ordered=True))
    m.register_breed(b)
    w = rng.random(n).astype(f32)
    for i in range(n):
        m.create_agent_of_breed(b, weight=float(w[i]))
    for _ in range(2000):
        a, c = rng.integers(0, n, 2)
        m.get_space().connect_agents(int(a), int(c), directed=bool(rng.random() < 0.3))
    m.setup()
    m.simulate(3)
    got = np.asarray(m._device_property("value"), f32)[m._inv]
    want = np.empty(n, f32)
    for i in range(n):
        row = list(m.get_space().get_location(i))
        acc = 0.0
        for j, x in enumerate(row):
            acc += float(j + 1) * float(w[x])
        acc2 = 0.0
        for x in row:
            acc2 += float(w[x])
        want[i] = f32(acc + 0.001 * acc2 + 2.0 + len(row) * 1000.0)
    assert np.array_equal(got, want)


# ---------------- logical ids above 2^24 ----------------

def test_reference_step_with_large_logical_ids():
    """set_agent_logical_id with ids above 2^24: draws keyed on the float32
    id (several agents share one), equal to the oracle's restatement."""
    say("logical ids above 2^24")
    g = graphs.erdos_renyi(1500, 0.006, 9)
    b = Breed("SIR")
    b.register_property("state", 1)
    b.register_step_func(reference_sir_step(), priority=0)
    m = Model(NetworkSpace())
    m.register_breed(b)
    m.register_global_property("p_infection", 0.35)
    m.register_global_property("p_recovery", 0.2)
    m.set_seed(555)
    rng = np.random.default_rng(2)
    init = np.where(rng.random(g.n) < 0.03, 2, 1)
    lids = (1 << 24) + rng.permutation(4 * g.n)[:g.n]
    for i in range(g.n):
        m.create_agent_of_breed(b, state=int(init[i]))
        m.set_agent_logical_id(i, int(lids[i]))
    for i in range(g.n):
        for j in g.col_idx[g.row_ptr[i]:g.row_ptr[i + 1]]:
            m.get_space().connect_agents(i, int(j), directed=True)
    m.setup()
    m.simulate(20)
    want = O.ref_sir_run(g.row_ptr, g.col_idx, lids.astype(f32), init.astype(f32), 555, 20, 0.35, 0.2)
    got = np.array([m.get_agent_property_value(i, "state") for i in range(g.n)], f32)
    assert np.array_equal(got, want), int((got != want).sum())
    assert len(set(lids.astype(f32).tolist())) < g.n        # float32 ids collide, as in the reference


# ---------------- code-object cache ----------------

def test_compile_cache_hit_and_skip_priority_barriers():
    say("hipRTC cache")
    import ctypes

    def build():
        b = Breed("C")
        b.register_property("counter", 0)
        b.register_step_func(StepFunction.from_source(
            "__device__ void inc_cache_probe(auto tick, auto agent_index, auto agent_ids, auto breeds, "
            "auto locations, auto counter) { counter[agent_index] += 2; }"), priority=0)
        m = Model(NetworkSpace())
        m.register_breed(b)
        for _ in range(10):
            m.create_agent_of_breed(b)
        m.setup(skip_priority_barriers=True)
        return m
    m1 = build()
    m2 = build()
    hit = ctypes.c_int32()
    m2._prog.ck(m2._prog._L.ss_prog_cache_hit(m2._prog._h, ctypes.byref(hit)))
    assert hit.value == 1
    m2.simulate(4)
    assert [m2.get_agent_property_value(i, "counter") for i in range(10)] == [8] * 10
    m1.simulate(1)
    assert m1.get_agent_property_value(3, "counter") == 2
