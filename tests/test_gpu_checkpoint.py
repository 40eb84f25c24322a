"""Checkpoint and restore (replacing the pickle of Model.save / Model.load,
model.py:1289-1310): a run saved at tick 50 and restored into a fresh
context finishes at tick 100 bit-identical to the uninterrupted run and to
the oracle, for the fused compartment kernel (engine and Model) and for
hipRTC-compiled step functions."""
import numpy as np
import pytest

from oracle import oracle as O
from sagesim_amd import Breed, Engine, Model, NetworkSpace, Replica, graphs, workloads
from sagesim_amd._lib import NativeError
from sagesim_amd.engine import sir_model
from sagesim_amd.steps import reference_sir_step, sir_step

pytestmark = pytest.mark.gpu

f32 = np.float32


def say(*a):
    print("[case]", *a, flush=True)


def test_replica_checkpoint_fresh_context():
    say("replica checkpoint at tick 50 of BA(200k, 5)")
    g = graphs.barabasi_albert(200_000, 5, 17)
    st0 = workloads.seed_infected(g.n, 200, 3)
    model = sir_model(0.05, 0.1)
    with Engine(0) as e1:
        full = Replica(e1, g, st0, model, 42)
        e1.simulate(100)
        want_full = full.states()
        part = Replica(e1, g, st0, model, 42)
        full.remove()
        e1.simulate(50)
        blob = part.save()
        assert part.tick() == 50
    with Engine(0) as e2:
        rep = Replica(e2, g, st0, model, 42)
        assert rep.restore(blob) == 50 and rep.tick() == 50
        e2.simulate(50)
        got = rep.states()
        # another network is refused
        other = Replica(e2, graphs.barabasi_albert(200_000, 5, 18), st0, model, 42)
        with pytest.raises(NativeError):
            other.restore(blob)
    want, _ = O.sir_run(O.sir_model(0.05, 0.1, 42), g.row_ptr, g.col_idx, np.arange(g.n), st0, 100)
    assert np.array_equal(want_full, want)
    assert np.array_equal(got, want), int((got != want).sum())


def _ref_model(g, init, seed, fused):
    b = Breed("SIR")
    b.register_property("state", 1 if not fused else 0, dtype="float32" if not fused else "int8")
    b.register_step_func(reference_sir_step() if not fused else sir_step(), priority=0)
    m = Model(NetworkSpace())
    m.register_breed(b)
    if fused:
        m.register_global_property("beta", 0.05)
        m.register_global_property("gamma", 0.1)
    else:
        m.register_global_property("p_infection", 0.3)
        m.register_global_property("p_recovery", 0.15)
    m.set_seed(seed)
    for i in range(g.n):
        m.create_agent_of_breed(b, state=int(init[i]))
    u = np.repeat(np.arange(g.n), np.diff(g.row_ptr))
    m.get_space().connect_agents_bulk(u, g.col_idx, directed=True)
    return m


@pytest.mark.parametrize("fused", [False, True])
def test_model_save_load(tmp_path, fused):
    say("Model.save / Model.load", "fused" if fused else "hipRTC")
    g = graphs.erdos_renyi(3000, 0.004, 12)
    rng = np.random.default_rng(1)
    if fused:
        init = workloads.seed_infected(g.n, 30, 4)
    else:
        init = np.where(rng.random(g.n) < 0.02, 2, 1)
    a = _ref_model(g, init, 2024, fused)
    a.setup()
    assert a.mode == ("fused" if fused else "rtc")
    a.simulate(50)
    path = tmp_path / "ckpt.npz"
    a.save(a, str(path))
    a.simulate(50)
    full = np.array([a.get_agent_property_value(i, "state") for i in range(g.n)], f32)
    b = _ref_model(g, init, 7, fused)            # another seed: the checkpoint brings the saved one
    b.setup()
    assert b.load(str(path)) is b and b.tick == 50
    b.simulate(50)
    got = np.array([b.get_agent_property_value(i, "state") for i in range(g.n)], f32)
    assert np.array_equal(got, full)
    if fused:
        want, _ = O.sir_run(O.sir_model(0.05, 0.1, 2024), g.row_ptr, g.col_idx, np.arange(g.n), init, 100)
    else:
        want = O.ref_sir_run(g.row_ptr, g.col_idx, np.arange(g.n, dtype=f32), init.astype(f32), 2024, 100, 0.3, 0.15)
    assert np.array_equal(got, want.astype(f32))
    # a model of another shape is refused
    c = _ref_model(graphs.erdos_renyi(2000, 0.004, 12), init[:2000], 1, fused)
    c.setup()
    with pytest.raises(Exception):
        c.load(str(path))
