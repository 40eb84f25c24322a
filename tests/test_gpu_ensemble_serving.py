"""Batched ensembles and the serving loop give every scenario the states of its
standalone run (and of the CPU oracle), whatever shares the launch."""
import numpy as np
import pytest

from oracle import oracle as O
from sagesim_amd import Engine, Replica
from sagesim_amd.ensemble import Ensemble, Scenario, prepare, run_batched, run_standalone
from sagesim_amd.serving import Server, footprint

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def _mixed_scenarios(count=14, seed=21):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        seir = i % 3 == 1
        out.append(Scenario(n=int(np.exp(rng.uniform(np.log(3000), np.log(60000)))),
                            graph_kind="ba" if i % 2 == 0 else "ws", graph_param=int(rng.integers(2, 9)) if i % 2 == 0 else 10,
                            beta=float(rng.uniform(0.02, 0.1)), gamma=float(rng.uniform(0.05, 0.2)),
                            sigma=float(rng.uniform(0.1, 0.3)) if seir else 0.0,
                            seed=int(rng.integers(1, 2**62)), ticks=int(rng.integers(30, 120)), frac_infected=0.005))
    return out


def _oracle(p):
    sc = p.scenario
    om = (O.seir_model(sc.beta, sc.sigma, sc.gamma, sc.seed) if sc.sigma > 0
          else O.sir_model(sc.beta, sc.gamma, sc.seed))
    out, _ = O.sir_run(om, p.graph.row_ptr, p.graph.col_idx, np.arange(p.graph.n), p.state, sc.ticks)
    return out


@pytest.fixture(scope="module")
def eng():
    e = Engine(0)
    yield e
    e.close()


def test_ensemble_matches_standalone_and_oracle(eng):
    prepared = prepare(_mixed_scenarios(), threads=4)
    say("ensemble", [p.scenario.n for p in prepared])
    ens = Ensemble.build(eng, prepared)
    t = run_batched(eng, ens)
    batched = [ens.states(i) for i in range(len(prepared))]
    ticks = [r.tick() for r in ens.replicas]
    ens.remove()
    assert ticks == [p.scenario.ticks for p in prepared]     # each stopped at its own tick count
    alone, t_alone = run_standalone(eng, prepared)
    say("batched s", round(t, 4), "standalone s", round(t_alone, 4))
    for i, p in enumerate(prepared):
        assert np.array_equal(batched[i], alone[i]), i
    for i in (0, 1, 4):
        assert np.array_equal(batched[i], _oracle(prepared[i])), i


@pytest.mark.parametrize("shape", [(1024, 64, 8), (16, 8, 2), (4096, 128, 16), (1 << 30, 1 << 30, 16)])
def test_ensemble_launch_shapes(eng, shape):
    hub, heavy, apl = shape
    eng.set_tiling(heavy, apl, 0, True, hub_degree=hub)
    prepared = prepare(_mixed_scenarios(6, seed=5), threads=4)
    ens = Ensemble.build(eng, prepared)
    ens.run(ens.ticks_needed)
    for i, p in enumerate(prepared):
        assert np.array_equal(ens.states(i), _oracle(p)), i
    ens.remove()
    eng.set_tiling()


def test_serving_alone_equals_in_mix(eng):
    """Continuous batching: requests arrive while others run; every result is
    the scenario's standalone result."""
    scs = _mixed_scenarios(16, seed=9)
    arrivals = [(0.004 * i, sc) for i, sc in enumerate(scs)]
    say("serving", len(arrivals))

    def make(req):
        sc = req.scenario
        return Replica(eng, model=sc.model(), seed=sc.seed, tick_limit=sc.ticks, data=req.data)

    prepared = prepare([sc for _, sc in arrivals], threads=4)
    budget = 3 * max(footprint(p.graph) for p in prepared)   # several in flight, never all
    srv = Server(eng, make, budget_bytes=budget, chunk_ticks=8, builders=4)
    stats, done = srv.run(arrivals)
    srv.close()
    s = stats.summary()
    say("serving summary", s)
    assert stats.completed == len(arrivals) and stats.max_inflight > 1
    assert stats.max_bytes <= budget
    alone, _ = run_standalone(eng, prepared)
    by_id = {r.rid: r for r in done}
    for i in range(len(arrivals)):
        assert np.array_equal(by_id[i].result, alone[i]), i
