"""The batched ensemble, the SBM partition and the serving loop at the sizes
their configurations name (256 scenarios / 97M agents; 4.0M local agents with
6.28M ghosts; 50k-5M-agent requests), checked the way the smaller tests check
them: the oracle's states (and history) and each scenario run alone."""
import os

import numpy as np
import pytest

from oracle import oracle as O
from sagesim_amd import Engine, Replica, pack_2bit, sir_model, workloads

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


def say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(scope="module")
def eng():
    e = Engine(0)
    yield e
    e.close()


def test_ensemble_256_scenarios(eng):
    """Config D: 256 BA scenarios of 10k-2M agents, 300 ticks, one launch per
    tick; eight replicas (the smallest, the largest and six spread over the
    list) equal their standalone runs, and three equal the oracle."""
    from sagesim_amd.ensemble import Ensemble, ensemble_scenarios, prepare, run_standalone
    scs = ensemble_scenarios(256, ticks=300)
    prepared = prepare(scs, threads=THREADS)
    say("ensemble", len(prepared), sum(p.graph.n for p in prepared), "agents")
    eng.set_tiling()
    ens = Ensemble.build(eng, prepared)
    ens.run(ens.ticks_needed)
    eng.sync()
    sizes = np.array([p.graph.n for p in prepared])
    pick = sorted({int(sizes.argmin()), int(sizes.argmax()), 17, 63, 101, 150, 199, 240})
    batched = {i: ens.states(i) for i in pick}
    ens.remove()
    alone, _ = run_standalone(eng, [prepared[i] for i in pick])
    for i, a in zip(pick, alone):
        assert np.array_equal(a, batched[i]), i
    for i in (int(sizes.argmin()), 17, int(sizes.argmax())):
        p = prepared[i]
        sc = p.scenario
        want, _ = O.sir_run(O.sir_model(sc.beta, sc.gamma, sc.seed), p.graph.row_ptr, p.graph.col_idx,
                            np.arange(p.graph.n), p.state, sc.ticks, nthreads=THREADS)
        say("oracle", i, p.graph.n)
        assert np.array_equal(batched[i], want), i


def test_sbm_partition_stated_size(eng):
    """Config C as stated: part 0 of a 32M-agent SBM (4.0M local agents,
    6.28M ghosts), 2-bit halo every tick, history on: 20 ticks equal the
    oracle's states and history."""
    pw = workloads.sbm_partition(ticks=20)
    g = pw.graph
    say("partition", g.n, g.n_cols - g.n, g.nnz)
    eng.set_tiling()
    rep = Replica(eng, g, pw.state, sir_model(pw.beta, pw.gamma), pw.seed, rng_key_id=pw.local_ids,
                  history=pw.history)
    eng.map_build(rep.rep, pw.ghost_ids, np.arange(g.n, g.n_cols, dtype=np.int32), 0.5)
    assert eng.halo_bind(rep.rep, pw.halo_ids) == 0
    vals = pw.halo_vals(0, pw.ticks)
    hist = pw.history.copy()
    want, _ = O.sir_run(O.sir_model(pw.beta, pw.gamma, pw.seed), g.row_ptr, g.col_idx, pw.local_ids, pw.state,
                        pw.ticks, n_local=g.n, history=hist, ghost_ids=pw.ghost_ids, halo_ids=pw.halo_ids,
                        halo_vals=vals, nthreads=THREADS)
    for t in range(0, pw.ticks, 10):
        eng.simulate_halo_2bit(rep.rep, 10, pack_2bit(vals[t:t + 10]))
    got = rep.states()
    assert np.array_equal(got[:g.n], want[:g.n])
    assert np.array_equal(rep.history(), hist)
    rep.remove()


def test_serving_full_size_requests(eng):
    """Config E's request distribution (SIR/SEIR on BA/WS, 50k-5M agents,
    100-500 ticks) over a 4-second window at 50/s: every request finishes and
    four of them, re-run alone, give the same states."""
    from sagesim_amd.ensemble import Prepared, run_standalone
    from sagesim_amd.serving import Server, serving_requests
    eng.set_tiling()
    arrivals = serving_requests(rate=50.0, duration=4.0, seed=999)
    keep = {0, len(arrivals) // 3, 2 * len(arrivals) // 3, len(arrivals) - 1}
    say("serving", len(arrivals), "requests")

    def make(req):
        sc = req.scenario
        return Replica(eng, model=sc.model(), seed=sc.seed, tick_limit=sc.ticks, data=req.data)

    srv = Server(eng, make, budget_bytes=200 << 30, chunk_ticks=10, builders=max(1, THREADS - 1),
                 keep_results=keep)
    stats, done = srv.run(arrivals)
    srv.close()
    assert stats.completed == len(arrivals) and stats.rejected == 0
    sample = [r for r in done if r.rid in keep]
    assert len(sample) == len(keep)
    prepared = [Prepared(r.scenario, r.scenario.build_graph(), r.scenario.initial_state()) for r in sample]
    alone, _ = run_standalone(eng, prepared)
    for r, a in zip(sample, alone):
        assert np.array_equal(a, r.result), r.rid
