"""The push direction (candidate tags from the infectious agents' rows) gives
the same states as the pull direction and as the oracle, tick by tick, for
every direction rule: never push, the default rule, and push whenever the
replica allows it."""
import numpy as np
import pytest

from oracle import oracle as O
from sagesim_amd import Engine, Replica, graphs, seir_model, sir_model, workloads

pytestmark = pytest.mark.gpu

RULES = (0.0, 0.25, 1.0e6)         # pull only, default, push whenever allowed


def say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(scope="module", params=[4, 16], ids=["apl4", "apl16"])
def engines(request):
    es = [Engine(0) for _ in RULES]
    for e, r in zip(es, RULES):
        e.set_push(r)
        e.set_tiling(64, request.param, 0, True)
    say("agents per lane", request.param)
    yield es
    for e in es:
        e.close()


def first_diff(a, b):
    d = np.nonzero(a != b)[0]
    return None if d.size == 0 else (int(d[0]), int(a[d[0]]), int(b[d[0]]), int(d.size))


def test_ba_push_pull_every_tick(engines):
    """BA(200k, 5) SIR, 150 ticks: the three rules agree after every tick and
    with the oracle's trace; the default rule mixes both directions."""
    w = workloads.ba_sir(n=200_000, m=5, ticks=150, seed=21)
    say("ba push/pull", w.graph.n, w.graph.nnz)
    reps = [Replica(e, w.graph, w.state, w.model, w.seed) for e in engines]
    assert all(e.push_ok(r.rep) for e, r in zip(engines, reps))
    _, trace = O.sir_run(O.sir_model(w.beta, w.gamma, w.seed), w.graph.row_ptr, w.graph.col_idx,
                         np.arange(w.graph.n), w.state, w.ticks, trace=True)
    for e in engines:
        e.set_counting(True)
        e.read_counters()
    for t in range(w.ticks):
        for e in engines:
            e.simulate(1)
        if t % 5 == 4 or t < 12:
            got = [r.states() for r in reps]
            for g in got:
                assert first_diff(g, trace[t]) is None, (t, first_diff(g, trace[t]))
    push_ticks = [int(e.read_counters()[5]) for e in engines]
    for e in engines:
        e.set_counting(False)
    say("push ticks per rule", push_ticks)
    assert push_ticks[0] == 0
    assert 0 < push_ticks[1] < w.ticks
    assert push_ticks[2] == w.ticks
    assert (trace[-1] == 2).sum() > w.graph.n // 10
    for r in reps:
        r.remove()


def test_push_with_history(engines):
    """Last-infector history (first infectious neighbour in CSR order) is the
    oracle's when candidates come from the push direction."""
    g = graphs.erdos_renyi(20_000, 0.0006, 5)
    rng = np.random.default_rng(2)
    st = np.where(rng.random(g.n) < 0.01, 1, 0).astype(np.int8)
    hist0 = np.full((g.n, 32), -1, np.int32)
    ids = np.arange(g.n)
    want_h = hist0.copy()
    want, _ = O.sir_run(O.sir_model(0.08, 0.1, 9), g.row_ptr, g.col_idx, ids, st, 60, history=want_h)
    say("history under push", g.n, g.nnz)
    for e in engines:
        rep = Replica(e, g, st, sir_model(0.08, 0.1), 9, history=hist0)
        e.simulate(60)
        assert first_diff(rep.states(), want) is None
        assert np.array_equal(rep.history(), want_h), np.argwhere(rep.history() != want_h)[:5]
        rep.remove()
    assert (want_h[:, 0] >= 0).sum() > 100


def test_asymmetric_graph_stays_pull(engines):
    """A directed CSR has no mirror entries: the replica never pushes and still
    matches the oracle."""
    rng = np.random.default_rng(4)
    n = 30_000
    u = rng.integers(0, n, 200_000)
    v = rng.integers(0, n, 200_000)
    g = graphs.from_edges(n, u, v, directed=True)
    st = np.where(rng.random(n) < 0.01, 1, 0).astype(np.int8)
    want, _ = O.sir_run(O.sir_model(0.1, 0.1, 3), g.row_ptr, g.col_idx, np.arange(n), st, 50)
    e = engines[2]
    rep = Replica(e, g, st, sir_model(0.1, 0.1), 3)
    assert not e.push_ok(rep.rep)
    e.set_counting(True)
    e.read_counters()
    e.simulate(50)
    assert int(e.read_counters()[5]) == 0
    e.set_counting(False)
    assert first_diff(rep.states(), want) is None
    rep.remove()


def test_seir_ba_bench_tiling_every_rule():
    """SEIR on BA(150k, 4) with the tiling the 10M-agent bench picks (hub
    2048, heavy 512, 16 agents per lane: hub, heavy and long-row light tiles
    of 64-192 agents) under the three direction rules: the states after every
    5th tick equal the oracle's trace."""
    g = graphs.barabasi_albert(150_000, 4, 33)
    st = workloads.seed_infected(g.n, 300, 34, code=2)
    beta, sigma, gamma, seed = 0.06, 0.25, 0.12, 35
    ticks = 90
    _, trace = O.sir_run(O.seir_model(beta, sigma, gamma, seed), g.row_ptr, g.col_idx, np.arange(g.n), st, ticks,
                         trace=True)
    es = [Engine(0) for _ in RULES]
    try:
        reps = []
        for e, r in zip(es, RULES):
            e.set_push(r)
            e.set_tiling(512, 16, 0, True, hub_degree=2048)
            reps.append(Replica(e, g, st, seir_model(beta, sigma, gamma), seed))
        say("seir ba bench tiling", g.n, g.nnz, [e.launch_shape() for e in es])
        for t in range(ticks):
            for e in es:
                e.simulate(1)
            if t % 5 == 4:
                for r in reps:
                    assert first_diff(r.states(), trace[t]) is None, (t, first_diff(r.states(), trace[t]))
        assert (trace[-1] == 3).sum() > g.n // 20
    finally:
        for e in es:
            e.close()


def test_ensemble_mixed_directions(engines):
    """Replicas of different size, model, tick offset and symmetry in one
    launch, always-push rule: each equals its own oracle run."""
    e = engines[2]
    cases = []
    for k, (n, m) in enumerate([(5_000, 2), (40_000, 6), (120_000, 3)]):
        w = workloads.ba_sir(n=n, m=m, beta=0.02 + 0.01 * k, gamma=0.1 + 0.03 * k, ticks=80, seed=30 + k)
        cases.append((w.graph, w.state, w.model, O.sir_model(w.beta, w.gamma, w.seed), w.seed))
    ws = workloads.ws_seir(n=50_000, ticks=80, seed=12)
    cases.append((ws.graph, ws.state, ws.model, O.seir_model(ws.beta, ws.sigma, ws.gamma, ws.seed), ws.seed))
    rng = np.random.default_rng(8)
    n = 20_000
    gd = graphs.from_edges(n, rng.integers(0, n, 120_000), rng.integers(0, n, 120_000), directed=True)
    sd = np.where(rng.random(n) < 0.02, 1, 0).astype(np.int8)
    cases.append((gd, sd, sir_model(0.1, 0.1), O.sir_model(0.1, 0.1, 5), 5))
    reps = [Replica(e, g, st, mdl, seed) for g, st, mdl, _, seed in cases]
    say("ensemble mixed", [g.n for g, *_ in cases])
    e.simulate(80)
    for (g, st, _, om, _), r in zip(cases, reps):
        want, _ = O.sir_run(om, g.row_ptr, g.col_idx, np.arange(g.n), st, 80)
        assert first_diff(r.states(), want) is None, (g.n, first_diff(r.states(), want))
    assert [e.push_ok(r.rep) for r in reps] == [True] * 4 + [False]
    for r in reps:
        r.remove()


def test_tile_rebuild_while_pushing(engines):
    """Adding and removing a replica rebuilds the tile list between ticks of a
    replica that is pushing with tile tags; every state stays the oracle's."""
    e = engines[2]
    wa = workloads.ba_sir(n=100_000, m=4, ticks=160, seed=41)
    wb = workloads.ba_sir(n=30_000, m=3, beta=0.05, ticks=60, seed=42)
    oa = O.sir_model(wa.beta, wa.gamma, wa.seed)
    ob = O.sir_model(wb.beta, wb.gamma, wb.seed)
    ra = Replica(e, wa.graph, wa.state, wa.model, wa.seed)
    e.simulate(60)
    rb = Replica(e, wb.graph, wb.state, wb.model, wb.seed)
    say("tile rebuild", wa.graph.n, wb.graph.n)
    e.simulate(60)
    want_b, _ = O.sir_run(ob, wb.graph.row_ptr, wb.graph.col_idx, np.arange(wb.graph.n), wb.state, 60)
    assert first_diff(rb.states(), want_b) is None
    rb.remove()
    e.simulate(40)
    want_a, _ = O.sir_run(oa, wa.graph.row_ptr, wa.graph.col_idx, np.arange(wa.graph.n), wa.state, 160)
    assert first_diff(ra.states(), want_a) is None
    ra.remove()


def test_si_model_keeps_pushing(engines):
    """SI (gamma = 0): infectious agents never draw, yet their rows are pushed
    every tick; states equal the oracle's under every rule."""
    w = workloads.ba_sir(n=60_000, m=3, beta=0.02, gamma=0.0, ticks=90, seed=51)
    om = O.sir_model(w.beta, 0.0, w.seed)
    _, trace = O.sir_run(om, w.graph.row_ptr, w.graph.col_idx, np.arange(w.graph.n), w.state, w.ticks, trace=True)
    say("si model", w.graph.n)
    for e in engines:
        rep = Replica(e, w.graph, w.state, sir_model(w.beta, 0.0), w.seed)
        for t0 in range(0, w.ticks, 30):
            e.simulate(30)
            assert first_diff(rep.states(), trace[t0 + 29]) is None, (t0, first_diff(rep.states(), trace[t0 + 29]))
        rep.remove()
    assert 0 < (trace[-1] == 1).sum() < w.graph.n


@pytest.mark.parametrize("grid", [7, 16])
def test_wide_tiles_shrunk_chunks(grid):
    """16 agents per lane on a small grid: the light agents take several rounds
    of the grid, so the chunk per wavefront shrinks below 1024 agents (tiles
    of a size that is not a power of two); every rule still gives the
    oracle's states."""
    w = workloads.ba_sir(n=120_000, m=4, beta=0.04, ticks=90, seed=61)
    _, trace = O.sir_run(O.sir_model(w.beta, w.gamma, w.seed), w.graph.row_ptr, w.graph.col_idx,
                         np.arange(w.graph.n), w.state, w.ticks, trace=True)
    say("shrunk chunks", grid)
    for rule in RULES:
        with Engine(0) as e:
            e.set_push(rule)
            e.set_tiling(64, 16, grid, True)
            rep = Replica(e, w.graph, w.state, w.model, w.seed)
            for t0 in range(0, w.ticks, 15):
                e.simulate(15)
                assert first_diff(rep.states(), trace[t0 + 14]) is None, (rule, t0, first_diff(rep.states(), trace[t0 + 14]))
            rep.remove()


@pytest.mark.parametrize("apl", [4, 16])
def test_unsorted_rows(apl):
    """Rows in random order (no degree sort): long rows sit among short ones in
    light tiles and the tile sizes follow them; every rule gives the oracle's
    states."""
    g0 = graphs.barabasi_albert(60_000, 4, 71)
    rng = np.random.default_rng(72)
    p = rng.permutation(g0.n)
    src = np.repeat(np.arange(g0.n), np.diff(g0.row_ptr))
    g = graphs.from_edges(g0.n, p[src], p[g0.col_idx])      # every edge twice: mirrors stay
    st = np.where(rng.random(g.n) < 0.002, 1, 0).astype(np.int8)
    om = O.sir_model(0.05, 0.1, 73)
    _, trace = O.sir_run(om, g.row_ptr, g.col_idx, np.arange(g.n), st, 60, trace=True)
    say("unsorted rows", apl, int(np.diff(g.row_ptr).max()))
    for rule in RULES:
        with Engine(0) as e:
            e.set_push(rule)
            e.set_tiling(64, apl, 0, True)
            rep = Replica(e, g, st, sir_model(0.05, 0.1), 73, sort=False)
            for t0 in range(0, 60, 12):
                e.simulate(12)
                assert first_diff(rep.states(), trace[t0 + 11]) is None, (rule, t0, first_diff(rep.states(), trace[t0 + 11]))
            rep.remove()
    assert (trace[-1] == 2).sum() > g.n // 20


def test_row_of_65535_or_more_stays_pull(engines):
    """Counts are 16-bit: a replica with a row of 65535 entries or more never
    pushes, and its states still equal the oracle's."""
    n = 70_001
    g = graphs.from_edges(n, np.zeros(n - 1, np.int64), np.arange(1, n, dtype=np.int64))
    st = np.zeros(n, np.int8)
    st[1:40] = 1
    want, _ = O.sir_run(O.sir_model(0.02, 0.2, 81), g.row_ptr, g.col_idx, np.arange(n), st, 20)
    e = engines[2]
    rep = Replica(e, g, st, sir_model(0.02, 0.2), 81)
    assert not e.push_ok(rep.rep)
    e.simulate(20)
    assert first_diff(rep.states(), want) is None
    rep.remove()


def test_sirs_counts_reach_every_neighbour(engines):
    """SIRS (R -> S with probability 0.05): a recovered agent becomes
    susceptible again, so pushed counts must reach neighbours that are not
    susceptible when they are pushed.  The three rules agree with the oracle
    every few ticks over a run long enough for agents to return to S."""
    from sagesim_amd._lib import SSModel
    g = graphs.degree_sort(graphs.barabasi_albert(60_000, 4, 5))[0]
    st = workloads.seed_infected(g.n, 300, 9)
    beta, gamma, xi, seed = 0.06, 0.15, 0.05, 77
    m = SSModel()
    m.n_states, m.s_code, m.inf_code, m.s_target = 3, 0, 1, 1
    m.beta = beta
    m.trans_p[:] = [0.0, gamma, xi, 0.0]
    m.trans_next[:] = [0, 2, 0, 0]
    om = O.sir_model(beta, gamma, seed)
    om.trans_p[:] = [0.0, gamma, xi, 0.0]
    om.trans_next[:] = [0, 2, 0, 0]
    ticks = 120
    say("SIRS push", g.n, g.nnz)
    _, trace = O.sir_run(om, g.row_ptr, g.col_idx, np.arange(g.n), st, ticks, trace=True)
    reps = [Replica(e, g, st, m, seed) for e in engines]
    assert all(e.push_ok(r.rep) for e, r in zip(engines, reps))
    for e in engines:
        e.set_counting(True)
        e.read_counters()
    for t in range(ticks):
        for e in engines:
            e.simulate(1)
        if t % 6 == 5:
            for r in reps:
                got = r.states()
                assert first_diff(got, trace[t]) is None, (t, first_diff(got, trace[t]))
    push_ticks = [int(e.read_counters()[5]) for e in engines]
    for e in engines:
        e.set_counting(False)
    say("push ticks per rule", push_ticks)
    assert push_ticks[2] == ticks
    # agents did come back: some susceptible agents had been recovered earlier
    assert (trace[-1] == 0).sum() > 0 and (trace[ticks // 2] == 2).sum() > 0
    for r in reps:
        r.remove()
