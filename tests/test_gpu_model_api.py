"""The reference's Model/Breed/NetworkSpace tests run through hipRTC-compiled
HIP step functions (and the fused kernel for the built-in compartment step):
tests/test_simulation.py:95-159 and tests/test_double_buffer.py:247-312,
plus oracle parity of the reference SIR step and fused-vs-compiled equality."""
import random as stdlib_random

import numpy as np
import pytest

from oracle import oracle as O
from sagesim_amd import Breed, Model, NetworkSpace, StepFunction, graphs
from sagesim_amd.steps import reference_sir_step, seir_step, sir_step

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


# ---------------- tests/test_simulation.py ----------------

class _SIRBreed(Breed):
    def __init__(self):
        super().__init__("SIR")
        self.register_property("state", 1)
        self.register_step_func(reference_sir_step(), priority=0)


class _SIRModel(Model):
    def __init__(self, p_infection=1.0, p_recovery=1.0):
        super().__init__(NetworkSpace())
        self._breed = _SIRBreed()
        self.register_breed(self._breed)
        self.register_global_property("p_infection", p_infection)
        self.register_global_property("p_recovery", p_recovery)

    def create_agent(self, state=1):
        return self.create_agent_of_breed(self._breed, state=state)

    def connect_agents(self, a0, a1):
        self.get_space().connect_agents(a0, a1)


def _chain_model(n, p_inf=1.0, p_rec=1.0, seed=42):
    m = _SIRModel(p_inf, p_rec)
    m.set_seed(seed)
    for _ in range(n):
        m.create_agent(state=1)
    m.set_agent_property_value(0, "state", 2)
    for i in range(n - 1):
        m.connect_agents(i, i + 1)
    return m


def _states(m, n):
    return [int(m.get_agent_property_value(i, "state")) for i in range(n)]


def test_deterministic_sir_chain():
    say("chain wave")
    m = _chain_model(10)
    m.setup(use_gpu=True)
    assert m.mode == "rtc"
    m.simulate(5, sync_workers_every_n_ticks=1)
    assert _states(m, 10) == [3] * 5 + [2] + [1] * 4


def test_seed_reproducibility_and_difference():
    say("seeds")
    a = _chain_model(20, 0.5, 0.3, 999)
    a.setup()
    a.simulate(10)
    b = _chain_model(20, 0.5, 0.3, 999)
    b.setup()
    b.simulate(10)
    assert _states(a, 20) == _states(b, 20)
    c = _chain_model(50, 0.3, 0.2, 111)
    c.setup()
    c.simulate(20)
    d = _chain_model(50, 0.3, 0.2, 222)
    d.setup()
    d.simulate(20)
    assert _states(c, 50) != _states(d, 50)


def test_reference_sir_step_matches_oracle():
    """The compiled reference step equals the oracle's restatement on a random
    graph, every agent, after 15 and 30 ticks (float32 states, Philox-2x32)."""
    say("reference step vs oracle")
    g = graphs.erdos_renyi(3000, 0.004, 12)
    m = _SIRModel(0.3, 0.15)
    m.set_seed(2024)
    rng = np.random.default_rng(1)
    init = np.where(rng.random(g.n) < 0.02, 2, 1)
    for i in range(g.n):
        m.create_agent(state=int(init[i]))
    for i in range(g.n):
        for j in g.col_idx[g.row_ptr[i]:g.row_ptr[i + 1]]:
            m.get_space().connect_agents(i, int(j), directed=True)
    m.setup()
    st = init.astype(np.float32)
    for t0 in (0, 15):
        m.simulate(15)
        st = O.ref_sir_run(g.row_ptr, g.col_idx, np.arange(g.n, dtype=np.float32), st, 2024, 15, 0.3, 0.15,
                           tick0=t0)
        got = np.array([m.get_agent_property_value(i, "state") for i in range(g.n)], np.float32)
        assert np.array_equal(got, st), int((got != st).sum())
    assert (st == 3).sum() > 10


# ---------------- tests/test_double_buffer.py ----------------

_INF_SRC = """
__device__ void infection_step_func(auto tick, auto agent_index, auto p_infection, auto agent_ids, auto breeds,
                                    auto locations, auto state_tensor) {
    auto nb = locations[agent_index];
    const int s = (int)get_this_agent_data_from_tensor(agent_index, state_tensor);
    if (s == 1) {
        bool infected = false;
        for (int i = 0; i < len(nb) && !infected; ++i) {
            if ((int)state_tensor[nb[i]] == 2 && rand_uniform_philox(tick, agent_index, 1) < p_infection) {
                set_this_agent_data_from_tensor(agent_index, state_tensor, 2);
                infected = true;
            }
        }
    }
}
"""
_INF_DUMMY_SRC = """
__device__ void infection_step_func_with_dummy(auto tick, auto agent_index, auto p_infection, auto p_recovery,
                                               auto agent_ids, auto breeds, auto locations, auto state_tensor,
                                               auto dummy_tensor) {
    auto nb = locations[agent_index];
    const int s = (int)get_this_agent_data_from_tensor(agent_index, state_tensor);
    if (s == 1) {
        bool infected = false;
        for (int i = 0; i < len(nb) && !infected; ++i) {
            if ((int)state_tensor[nb[i]] == 2 && rand_uniform_philox(tick, agent_index, 1) < p_infection) {
                set_this_agent_data_from_tensor(agent_index, state_tensor, 2);
                infected = true;
            }
        }
    }
}
__device__ void recovery_step_func(auto tick, auto agent_index, auto p_infection, auto p_recovery, auto agent_ids,
                                   auto breeds, auto locations, auto state_tensor, auto dummy_tensor) {
    dummy_tensor[agent_index] = dummy_tensor[agent_index] + 1;
    const int s = (int)get_this_agent_data_from_tensor(agent_index, state_tensor);
    if (s == 2 && rand_uniform_philox(tick, agent_index, 2) < p_recovery) state_tensor[agent_index] = 3;
}
"""
_DB_SRC = """
__device__ void write_counter_step_func(auto tick, auto agent_index, auto agent_ids, auto breeds, auto locations,
                                        auto counter, auto result) {
    counter[agent_index] = 10;
}
__device__ void read_and_multiply_step_func(auto tick, auto agent_index, auto agent_ids, auto breeds,
                                            auto locations, auto counter, auto result) {
    result[agent_index] = counter[agent_index] * 2;
}
"""


def _hierarchy():
    """tests/test_double_buffer.py:206-221, drawn with the same stdlib random calls."""
    edges = [(0, mid) for mid in range(1, 11)]
    stdlib_random.seed(46)
    end_agents = list(range(11, 111))
    for idx, mid in enumerate(range(1, 11)):
        assigned = end_agents[idx * 10: idx * 10 + 10]
        edges += [(mid, ea) for ea in assigned]
        remaining = [a for a in end_agents if a not in assigned]
        edges += [(mid, ea) for ea in stdlib_random.sample(remaining, min(stdlib_random.randint(1, 10), len(remaining)))]
    return edges


def _hier_model(breed, globals_):
    m = Model(NetworkSpace())
    m.register_breed(breed)
    for k, v in globals_:
        m.register_global_property(k, v)
    for _ in range(111):
        m.create_agent_of_breed(breed, state=1)
    m.set_agent_property_value(0, "state", 2)
    for a, b in _hierarchy():
        m.get_space().connect_agents(a, b)
    return m


def test_si_one_tick_spread():
    say("hierarchy SI")
    b = Breed("Infection")
    b.register_property("state", 1)
    b.register_step_func(StepFunction("infection_step_func", _INF_SRC), priority=0)
    m = _hier_model(b, [("p_infection", 1.0)])
    m.setup()
    m.simulate(1)
    assert m.get_agent_property_value(0, "state") == 2
    assert all(m.get_agent_property_value(a, "state") == 2 for a in range(1, 11))
    assert all(m.get_agent_property_value(a, "state") == 1 for a in range(11, 111))


def test_sir_two_tick_spread_two_priorities():
    say("hierarchy SIR, two priorities")
    b = Breed("Infection")
    b.register_property("state", 1)
    b.register_property("dummy", 0)
    b.register_step_func(StepFunction("infection_step_func_with_dummy", _INF_DUMMY_SRC), priority=0)
    b.register_step_func(StepFunction("recovery_step_func", _INF_DUMMY_SRC), priority=1)
    m = _hier_model(b, [("p_infection", 1.0), ("p_recovery", 1.0)])
    m.setup()
    m.simulate(2)
    assert m.get_agent_property_value(0, "state") == 3
    assert all(m.get_agent_property_value(a, "state") == 3 for a in range(1, 11))
    assert all(m.get_agent_property_value(a, "state") == 2 for a in range(11, 111))
    assert all(m.get_agent_property_value(a, "dummy") == 2 for a in range(111))


@pytest.mark.parametrize("order,nodb,want", [("wr", None, 0), ("wr", ["counter"], 20), ("rw", None, 0),
                                             ("rw", ["counter"], 0)])
def test_double_buffer_semantics(order, nodb, want):
    """tests/test_double_buffer.py:282-312: counter/result = 10/0, 10/20, 10/0, 10/0."""
    say("double buffer", order, nodb)
    b = Breed("DB")
    b.register_property("counter", 0)
    b.register_property("result", 0)
    w = StepFunction("write_counter_step_func", _DB_SRC)
    r = StepFunction("read_and_multiply_step_func", _DB_SRC)
    first, second = (w, r) if order == "wr" else (r, w)
    b.register_step_func(first, priority=0, no_double_buffer=nodb if order == "wr" else None)
    b.register_step_func(second, priority=1, no_double_buffer=nodb if order == "rw" else None)
    m = Model(NetworkSpace())
    m.register_breed(b)
    m.create_agent_of_breed(b, counter=0, result=0)
    m.setup()
    m.simulate(1)
    assert m.get_agent_property_value(0, "counter") == 10
    assert m.get_agent_property_value(0, "result") == want


# ---------------- built-in compartment steps: fused kernel == compiled source ----------------

@pytest.mark.parametrize("kind", ["sir", "seir"])
def test_fused_equals_compiled_and_oracle(kind):
    say("fused vs compiled", kind)
    g = graphs.barabasi_albert(20_000, 4, 77)
    rng = np.random.default_rng(3)
    inf_code = 1 if kind == "sir" else 2
    init = np.where(rng.random(g.n) < 0.01, inf_code, 0).astype(np.int64)

    def build(use_fused):
        b = Breed("Person")
        b.register_property("state", 0, dtype="int8")
        b.register_step_func(sir_step() if kind == "sir" else seir_step(), priority=0)
        m = Model(NetworkSpace(), use_fused=use_fused)
        m.register_breed(b)
        m.register_global_property("beta", 0.06)
        m.register_global_property("gamma", 0.12)
        if kind == "seir":
            m.register_global_property("sigma", 0.25)
        m.set_seed(4321)
        for i in range(g.n):
            m.create_agent_of_breed(b, state=int(init[i]))
        u = np.repeat(np.arange(g.n), np.diff(g.row_ptr))
        m.get_space().connect_agents_bulk(u, g.col_idx, directed=True)
        m.setup()
        return m

    mf, mc = build(True), build(False)
    assert mf.mode == "fused" and mc.mode == "rtc"
    mf.simulate(40)
    mc.simulate(40)
    sf = np.array([mf.get_agent_property_value(i, "state") for i in range(g.n)])
    sc = np.array([mc.get_agent_property_value(i, "state") for i in range(g.n)])
    assert np.array_equal(sf, sc), int((sf != sc).sum())
    om = O.sir_model(0.06, 0.12, 4321) if kind == "sir" else O.seir_model(0.06, 0.25, 0.12, 4321)
    want, _ = O.sir_run(om, g.row_ptr, g.col_idx, np.arange(g.n), init.astype(np.int8), 40)
    assert np.array_equal(sf, want)
    assert (want == (2 if kind == "sir" else 3)).sum() > 0


def test_compile_error_is_reported():
    say("compile error")
    b = Breed("Bad")
    b.register_property("x", 0)
    b.register_step_func(StepFunction.from_source(
        "__device__ void bad(auto tick, auto agent_index, auto agent_ids, auto breeds, auto locations, auto x) "
        "{ x[agent_index] = undefined_symbol; }"), priority=0)
    m = Model(NetworkSpace())
    m.register_breed(b)
    m.create_agent_of_breed(b)
    from sagesim_amd._lib import NativeError
    with pytest.raises(NativeError) as ex:
        m.setup()
    assert "undefined_symbol" in str(ex.value)
