"""Bit-exact parity of the fused HIP tick kernels against the CPU oracle.

Every test compares per-agent int8 states (and, where present, the
last-infector history) produced through the C ABI with oracle/sir_oracle.c on
the same seeded inputs.
"""
import numpy as np
import pytest

from oracle import oracle as O
from sagesim_amd import Engine, Replica, graphs, sir_model, seir_model, workloads

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(scope="module")
def eng():
    e = Engine(0)
    yield e
    e.close()


def oracle_model(w):
    if w.sigma > 0:
        return O.seir_model(w.beta, w.sigma, w.gamma, w.seed)
    return O.sir_model(w.beta, w.gamma, w.seed)


def first_diff(a, b):
    d = np.nonzero(a != b)[0]
    return None if d.size == 0 else (int(d[0]), int(a[d[0]]), int(b[d[0]]), int(d.size))


def test_ba_10m_whole_run_matches_oracle(eng):
    """The headline workload at its full size: plain SIR on BA(10M, 5) for 200
    ticks with the tiling the bench's autotuner picks (hub 2048, heavy 512, 16
    agents per lane), final states equal to the oracle's whole run."""
    import os
    w = workloads.ba_sir(n=10_000_000)
    say("ba 10M whole run", w.graph.n, w.graph.nnz)
    eng.set_tiling(heavy_degree=512, agents_per_lane=16, hub_degree=2048)
    try:
        rep = Replica(eng, w.graph, w.state, w.model, w.seed)
        eng.simulate(w.ticks)
        got = rep.states()
        rep.remove()
    finally:
        eng.set_tiling()
    want, _ = O.sir_run(oracle_model(w), w.graph.row_ptr, w.graph.col_idx, np.arange(w.graph.n), w.state, w.ticks,
                        nthreads=min(16, os.cpu_count() or 1))
    assert first_diff(got, want) is None
    assert np.bincount(want, minlength=3)[2] > w.graph.n // 2


def test_er_sir_every_tick(eng):
    """10k-agent G(n, 0.001) SIR, 100 ticks: identical states after every tick."""
    w = workloads.er_sir()
    say("er_sir every tick", w.graph.n, w.graph.nnz)
    rep = Replica(eng, w.graph, w.state, w.model, w.seed)
    _, trace = O.sir_run(oracle_model(w), w.graph.row_ptr, w.graph.col_idx, np.arange(w.graph.n),
                         w.state, w.ticks, trace=True)
    for t in range(w.ticks):
        eng.simulate(1)
        got = rep.states()
        assert first_diff(got, trace[t]) is None, (t, first_diff(got, trace[t]))
    # the epidemic did something
    assert (trace[-1] == 2).sum() > 100
    rep.remove()


@pytest.mark.parametrize("hub,heavy,apl,graph_loop", [(1024, 64, 1, True), (16, 8, 2, True),
                                                      (1 << 30, 1 << 30, 4, False), (64, 16, 1, False),
                                                      (32, 32, 4, True), (8, 8, 1, True), (1024, 64, 8, True),
                                                      (1 << 30, 1 << 30, 8, True), (1024, 64, 16, True),
                                                      (1 << 30, 1 << 30, 16, True), (8, 8, 16, False)])
def test_ba_launch_shapes(eng, hub, heavy, apl, graph_loop):
    """BA(50k, 5), 60 ticks, in one captured loop: every launch shape gives the oracle's states.
    With hub and heavy at 2^30 every row is light, so rows of degree 255 and more (whose
    one-byte degree is the "read the row bounds" mark) go through the light paths."""
    w = workloads.ba_sir(n=50_000, m=5, beta=0.06, ticks=60, seed=3)
    assert int(w.graph.degrees().max()) > 255
    say("ba launch shape", hub, heavy, apl, graph_loop)
    eng.set_tiling(heavy, apl, 0, graph_loop, hub_degree=hub)
    rep = Replica(eng, w.graph, w.state, w.model, w.seed)
    want, _ = O.sir_run(oracle_model(w), w.graph.row_ptr, w.graph.col_idx, np.arange(w.graph.n),
                        w.state, w.ticks)
    eng.simulate(w.ticks)
    got = rep.states()
    assert first_diff(got, want) is None, first_diff(got, want)
    assert rep.tick() == w.ticks
    rep.remove()
    eng.set_tiling()


def test_ba_1m_subset_200_ticks(eng):
    """BA(1M, 5), beta 0.03, gamma 0.1, 0.1% infected, 200 ticks: states compared every 25 ticks."""
    w = workloads.ba_sir(n=1_000_000, m=5, ticks=200)
    say("ba 1M subset", w.graph.nnz)
    rep = Replica(eng, w.graph, w.state, w.model, w.seed)
    st = w.state
    om = oracle_model(w)
    for t0 in range(0, 200, 25):
        st, _ = O.sir_run(om, w.graph.row_ptr, w.graph.col_idx, np.arange(w.graph.n), st, 25, tick0=t0)
        eng.simulate(25)
        got = rep.states()
        assert first_diff(got, st) is None, (t0 + 25, first_diff(got, st))
    counts = np.bincount(st, minlength=3)
    say("final S/I/R", counts.tolist())
    assert counts[2] > 1000
    rep.remove()


def test_seir_ws(eng):
    w = workloads.ws_seir(n=30_000, ticks=80)
    say("seir ws", w.graph.n)
    rep = Replica(eng, w.graph, w.state, w.model, w.seed)
    want, _ = O.sir_run(oracle_model(w), w.graph.row_ptr, w.graph.col_idx, np.arange(w.graph.n),
                        w.state, w.ticks)
    eng.simulate(w.ticks)
    got = rep.states()
    assert first_diff(got, want) is None, first_diff(got, want)
    assert (want == 3).sum() > 0
    rep.remove()


def test_edge_cases(eng):
    """Isolated agents, a single agent, n not a multiple of 64, a hub of degree > 64*3,
    beta = 1 (BFS front), gamma = 1, everything infected."""
    say("edge cases")
    rng = np.random.default_rng(5)
    cases = []
    # star hub with 300 leaves + 37 isolated agents
    n = 338
    u = np.zeros(300, np.int64)
    v = np.arange(1, 301, dtype=np.int64)
    cases.append(("hub", graphs.from_edges(n, u, v), 0.3, 0.2))
    cases.append(("single", graphs.from_edges(1, [], []), 0.5, 0.5))
    cases.append(("chain_beta1", graphs.from_edges(101, np.arange(100), np.arange(1, 101)), 1.0, 1.0))
    g = graphs.erdos_renyi(1000, 0.01, 9)
    cases.append(("er_gamma1", g, 0.2, 1.0))
    shapes = [(1024, 64, 1)] * len(cases) + [(8, 8, 2)] * len(cases) + [(1 << 30, 1 << 30, 16)] * len(cases)
    for (name, g, beta, gamma), (hub, heavy, apl) in zip(cases * 3, shapes):
        eng.set_tiling(heavy, apl, 0, True, hub_degree=hub)
        for init in ("one", "all"):
            st = np.zeros(g.n, np.int8)
            if init == "one":
                st[0] = 1
            else:
                st[:] = 1
            say(name, init, hub, heavy, apl)
            seed = int(rng.integers(1 << 62))
            rep = Replica(eng, g, st, sir_model(beta, gamma), seed)
            want, trace = O.sir_run(O.sir_model(beta, gamma, seed), g.row_ptr, g.col_idx, np.arange(g.n), st, 30,
                                    trace=True)
            for t in range(30):
                eng.simulate(1)
                got = rep.states()
                assert first_diff(got, trace[t]) is None, (name, init, t, first_diff(got, trace[t]))
            rep.remove()
    eng.set_tiling()


@pytest.mark.parametrize("hub,heavy", [(1024, 64), (16, 8)])
def test_history_and_ghosts(eng, hub, heavy):
    """One SBM partition with ghosts refreshed from a halo each tick and a
    last-infector history: states and history identical to the oracle."""
    eng.set_tiling(heavy, 1, 0, True, hub_degree=hub)
    pw = workloads.sbm_partition(n_blocks=16, block_size=4000, deg_in=12, deg_out=2, n_parts=4, part=1,
                                 ticks=40, seed=17)
    g = pw.graph
    say("sbm partition", g.n, g.n_cols, g.nnz)
    rep = Replica(eng, g, pw.state, sir_model(pw.beta, pw.gamma), pw.seed, rng_key_id=pw.local_ids,
                  history=pw.history)
    n_ghost = g.n_cols - g.n
    eng.map_build(rep.rep, pw.ghost_ids, np.arange(g.n, g.n_cols, dtype=np.int32), 0.5)
    assert eng.halo_bind(rep.rep, pw.halo_ids) == 0
    vals = pw.halo_vals(0, pw.ticks)
    hist = pw.history.copy()
    want, _ = O.sir_run(O.sir_model(pw.beta, pw.gamma, pw.seed), g.row_ptr, g.col_idx, pw.local_ids,
                        pw.state, pw.ticks, n_local=g.n, history=hist, ghost_ids=pw.ghost_ids,
                        halo_ids=pw.halo_ids, halo_vals=vals)
    eng.simulate_halo(rep.rep, pw.ticks, vals)
    got = rep.states()
    assert first_diff(got[:g.n], want[:g.n]) is None, first_diff(got[:g.n], want[:g.n])
    gh = rep.history()
    assert np.array_equal(gh, hist), np.argwhere(gh != hist)[:5]
    assert (hist != pw.history).any()
    assert n_ghost > 0
    rep.remove()
    eng.set_tiling()


def test_hash_map_against_dict(eng):
    """The GPU id -> slot map answers like a dict, including misses, removal and reinsertion."""
    say("hash map")
    g = graphs.from_edges(10, [], [])
    rep = eng.add_replica(g.row_ptr, g.col_idx, np.arange(10), np.zeros(5000 + 10, np.int8), sir_model(0.1, 0.1), 1,
                          n_ghost=5000)
    rng = np.random.default_rng(3)
    ids = rng.integers(0, 1 << 62, size=5000, dtype=np.int64)
    ids = np.unique(ids)
    slots = rng.integers(0, 5010, size=ids.shape[0]).astype(np.int32)
    eng.map_build(rep, ids, slots, 0.5)
    probe = np.concatenate([ids, rng.integers(0, 1 << 62, size=1000, dtype=np.int64)])
    got, miss = eng.map_lookup(rep, probe)
    d = dict(zip(ids.tolist(), slots.tolist()))
    want = np.array([d.get(int(x), -1) for x in probe], np.int32)
    assert np.array_equal(got, want)
    assert miss == int((want < 0).sum())
    rm = ids[::3]
    assert eng.map_remove(rep, rm) == rm.shape[0]
    got, _ = eng.map_lookup(rep, ids)
    want = np.where(np.isin(ids, rm), -1, slots)
    assert np.array_equal(got, want)
    eng.map_insert(rep, rm, slots[::3])
    got, _ = eng.map_lookup(rep, ids)
    assert np.array_equal(got, slots)
    eng.remove_replica(rep)


def test_bad_csr_rejected(eng):
    """Out-of-range neighbour slots are refused by the device-side range check,
    with the offending entry named, and the context stays usable."""
    from sagesim_amd._lib import NativeError
    say("bad csr")
    rp = np.array([0, 2, 3], np.int64)
    ci = np.array([1, 7, 0], np.int32)          # 7 is outside the 2 slots
    with pytest.raises(NativeError) as ex:
        eng.add_replica(rp, ci, np.arange(2), np.zeros(2, np.int8), sir_model(0.1, 0.1), 1)
    assert "col_idx[1] = 7" in str(ex.value)
    with pytest.raises(NativeError):
        eng.add_replica(np.array([0, 2, 1], np.int64), np.array([1, 0], np.int32), np.arange(2),
                        np.zeros(2, np.int8), sir_model(0.1, 0.1), 1)
    g = graphs.from_edges(3, [0, 1], [1, 2])
    rep = Replica(eng, g, np.array([1, 0, 0], np.int8), sir_model(1.0, 0.0), 5)
    eng.simulate(2)
    assert rep.states().tolist() == [1, 1, 1]
    rep.remove()


def test_halo_from_page_locked_buffer(eng):
    """Halo values from a buffer of ss_host_alloc (copied with no staging) give
    the same states and history as from an ordinary host array."""
    pw = workloads.sbm_partition(n_blocks=8, block_size=3000, deg_in=10, deg_out=2, n_parts=2, part=0,
                                 ticks=24, seed=19)
    g = pw.graph
    say("halo page-locked", g.n, g.n_cols)
    out = []
    for use_pinned in (False, True):
        rep = Replica(eng, g, pw.state, sir_model(pw.beta, pw.gamma), pw.seed, rng_key_id=pw.local_ids,
                      history=pw.history)
        eng.map_build(rep.rep, pw.ghost_ids, np.arange(g.n, g.n_cols, dtype=np.int32), 0.5)
        assert eng.halo_bind(rep.rep, pw.halo_ids) == 0
        for t in range(0, pw.ticks, 8):
            if use_pinned:
                vals = eng.pinned_empty((8, pw.halo_ids.shape[0]), np.int8)
                pw.halo_vals(t, t + 8, out=vals)
            else:
                vals = pw.halo_vals(t, t + 8)
            eng.simulate_halo(rep.rep, 8, vals)
        out.append((rep.states(), rep.history()))
        rep.remove()
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
    hist = pw.history.copy()
    want, _ = O.sir_run(O.sir_model(pw.beta, pw.gamma, pw.seed), g.row_ptr, g.col_idx, pw.local_ids, pw.state,
                        pw.ticks, n_local=g.n, history=hist, ghost_ids=pw.ghost_ids, halo_ids=pw.halo_ids,
                        halo_vals=pw.halo_vals(0, pw.ticks))
    assert first_diff(out[1][0][:g.n], want[:g.n]) is None
    assert np.array_equal(out[1][1], hist)


def test_halo_2bit_matches_bytes(eng):
    """The 2-bit halo gives the states and history of the byte halo, from an
    ordinary array (staged) and from page-locked memory, with a halo whose
    record count is not a multiple of 4 and one ghost left unbound; and the
    full binding matches the oracle."""
    from sagesim_amd import pack_2bit
    pw = workloads.sbm_partition(n_blocks=8, block_size=3000, deg_in=10, deg_out=2, n_parts=2, part=0,
                                 ticks=24, seed=23)
    g = pw.graph
    for ids in (pw.halo_ids, pw.halo_ids[:-1] if pw.halo_ids.shape[0] % 4 != 1 else pw.halo_ids[:-2]):
        n = ids.shape[0]
        say("halo 2-bit", g.n, g.n_cols, "records", n)
        vals = pw.halo_vals(0, pw.ticks)
        if n < vals.shape[1]:
            # records of the subset halo, in its order (vals columns follow pw.halo_ids)
            vals = vals[:, :n]
        out = []
        for mode in ("bytes", "2bit", "2bit-pinned"):
            rep = Replica(eng, g, pw.state, sir_model(pw.beta, pw.gamma), pw.seed, rng_key_id=pw.local_ids,
                          history=pw.history)
            eng.map_build(rep.rep, pw.ghost_ids, np.arange(g.n, g.n_cols, dtype=np.int32), 0.5)
            assert eng.halo_bind(rep.rep, ids) == 0
            for t in range(0, pw.ticks, 8):
                v = vals[t:t + 8]
                if mode == "bytes":
                    eng.simulate_halo(rep.rep, 8, v)
                elif mode == "2bit":
                    eng.simulate_halo_2bit(rep.rep, 8, pack_2bit(v))
                else:
                    buf = eng.pinned_empty((8, (n + 3) // 4), np.uint8)
                    buf[:] = pack_2bit(v)
                    eng.simulate_halo_2bit(rep.rep, 8, buf)
            out.append((rep.states(), rep.history()))
            rep.remove()
        for k in (1, 2):
            assert first_diff(out[k][0], out[0][0]) is None
            assert np.array_equal(out[k][1], out[0][1])
        if n == pw.halo_ids.shape[0]:
            hist = pw.history.copy()
            want, _ = O.sir_run(O.sir_model(pw.beta, pw.gamma, pw.seed), g.row_ptr, g.col_idx, pw.local_ids,
                                pw.state, pw.ticks, n_local=g.n, history=hist, ghost_ids=pw.ghost_ids,
                                halo_ids=pw.halo_ids, halo_vals=pw.halo_vals(0, pw.ticks))
            assert first_diff(out[1][0][:g.n], want[:g.n]) is None
            assert np.array_equal(out[1][1], hist)


@pytest.mark.parametrize("apl", [16, 4])
def test_halo_page_locked_groups(eng, apl):
    """Page-locked halo input runs in groups of at most 32 ticks, each copied
    ahead in spans of 1, 1, 2, 4, ... ticks: one 75-tick call (groups of 32,
    32 and 11 ticks) as 2-bit codes and 40 + 35 ticks as bytes give the
    oracle's states and history, as does the staged path.  With 16 agents a
    lane the ticks inside a group carry the next tick's refresh; with 4 the
    refresh kernel runs before every tick."""
    pw = workloads.sbm_partition(n_blocks=8, block_size=3000, deg_in=10, deg_out=2, n_parts=2, part=0,
                                 ticks=75, seed=29)
    g = pw.graph
    n = pw.halo_ids.shape[0]
    vals = pw.halo_vals(0, pw.ticks)
    hist = pw.history.copy()
    want, _ = O.sir_run(O.sir_model(pw.beta, pw.gamma, pw.seed), g.row_ptr, g.col_idx, pw.local_ids, pw.state,
                        pw.ticks, n_local=g.n, history=hist, ghost_ids=pw.ghost_ids, halo_ids=pw.halo_ids,
                        halo_vals=vals)
    eng.set_tiling(agents_per_lane=apl)
    try:
        _halo_groups_modes(eng, pw, g, n, vals, want, hist, apl)
    finally:
        eng.set_tiling()


def _halo_groups_modes(eng, pw, g, n, vals, want, hist, apl):
    from sagesim_amd import pack_2bit
    for mode in ("2bit-pinned", "bytes-pinned", "bytes-staged"):
        say("halo groups", mode, "agents per lane", apl, g.n, g.n_cols)
        rep = Replica(eng, g, pw.state, sir_model(pw.beta, pw.gamma), pw.seed, rng_key_id=pw.local_ids,
                      history=pw.history)
        eng.map_build(rep.rep, pw.ghost_ids, np.arange(g.n, g.n_cols, dtype=np.int32), 0.5)
        assert eng.halo_bind(rep.rep, pw.halo_ids) == 0
        if mode == "2bit-pinned":
            buf = eng.pinned_empty((pw.ticks, (n + 3) // 4), np.uint8)
            buf[:] = pack_2bit(vals)
            eng.simulate_halo_2bit(rep.rep, pw.ticks, buf)
        else:
            for t0, t1 in ((0, 40), (40, 75)):
                if mode == "bytes-pinned":
                    buf = eng.pinned_empty((t1 - t0, n), np.int8)
                    buf[:] = vals[t0:t1]
                else:
                    buf = vals[t0:t1]
                eng.simulate_halo(rep.rep, t1 - t0, buf)
        assert first_diff(rep.states()[:g.n], want[:g.n]) is None, mode
        assert np.array_equal(rep.history(), hist), mode
        rep.remove()


def test_halo_refresh_large_with_repeats(eng):
    """A halo of 98k records in a random record order, then a binding with
    three ghosts left out and five records repeated at the end with other
    values (the later record of a slot wins): int8 and 2-bit halos against the
    oracle's states and history, ghosts included."""
    from sagesim_amd import pack_2bit
    pw = workloads.sbm_partition(n_blocks=8, block_size=30000, deg_in=8, deg_out=3, n_parts=2, part=0,
                                 ticks=12, seed=37)
    g = pw.graph
    base_vals = pw.halo_vals(0, pw.ticks)
    cases = [(pw.halo_ids, base_vals)]
    ids2 = np.concatenate([pw.halo_ids[:-3], pw.halo_ids[:5]])
    vals2 = np.concatenate([base_vals[:, :-3], (base_vals[:, :5] + 1) % 3], axis=1).astype(np.int8)
    cases.append((ids2, vals2))
    for ids, vals in cases:
        hist = pw.history.copy()
        want, _ = O.sir_run(O.sir_model(pw.beta, pw.gamma, pw.seed), g.row_ptr, g.col_idx, pw.local_ids, pw.state,
                            pw.ticks, n_local=g.n, history=hist, ghost_ids=pw.ghost_ids, halo_ids=ids, halo_vals=vals)
        for packed in (False, True):
            say("halo large", g.n, g.n_cols, "records", ids.shape[0], "packed", packed)
            rep = Replica(eng, g, pw.state, sir_model(pw.beta, pw.gamma), pw.seed, rng_key_id=pw.local_ids,
                          history=pw.history)
            eng.map_build(rep.rep, pw.ghost_ids, np.arange(g.n, g.n_cols, dtype=np.int32), 0.5)
            assert eng.halo_bind(rep.rep, ids) == 0
            for t in range(0, pw.ticks, 4):
                v = vals[t:t + 4]
                if packed:
                    eng.simulate_halo_2bit(rep.rep, 4, pack_2bit(v))
                else:
                    eng.simulate_halo(rep.rep, 4, v)
            st, hi = rep.states(), rep.history()
            rep.remove()
            assert first_diff(st[:g.n], want[:g.n]) is None
            assert np.array_equal(hi, hist)
            # ghosts: the last tick's halo (later records win), the three unbound ones as in the oracle
            assert first_diff(st[g.n:], want[g.n:]) is None


def test_halo_buffers_too_short_are_refused(eng):
    """simulate_halo and simulate_halo_2bit check the buffer against ticks x
    the bound record count (int8 or packed) before the library reads it, and a
    replica with no bound halo is refused."""
    from sagesim_amd import pack_2bit
    pw = workloads.sbm_partition(n_blocks=8, block_size=1000, deg_in=8, deg_out=2, n_parts=2, part=0,
                                 ticks=4, seed=41)
    g = pw.graph
    rep = Replica(eng, g, pw.state, sir_model(pw.beta, pw.gamma), pw.seed, rng_key_id=pw.local_ids)
    eng.map_build(rep.rep, pw.ghost_ids, np.arange(g.n, g.n_cols, dtype=np.int32), 0.5)
    vals = pw.halo_vals(0, 4)
    with pytest.raises(ValueError):
        eng.simulate_halo(rep.rep, 2, vals[:2])        # no halo bound yet
    assert eng.halo_bind(rep.rep, pw.halo_ids) == 0
    with pytest.raises(ValueError):
        eng.simulate_halo(rep.rep, 4, vals[:3])
    with pytest.raises(ValueError):
        eng.simulate_halo_2bit(rep.rep, 4, pack_2bit(vals[:3]))
    with pytest.raises(ValueError):
        eng.halo_unpack_2bit(rep.rep, pack_2bit(vals[0])[:-1], pw.halo_ids.shape[0])
    eng.simulate_halo_2bit(rep.rep, 4, pack_2bit(vals))   # the full buffer runs
    rep.remove()


def test_halo_2bit_pack_unpack_all_codes(eng):
    """One 2-bit refresh of codes 0..3 (an SEIR replica) sets the ghosts as
    the byte refresh does, and the 2-bit pack of slots equals the packed
    byte pack."""
    from sagesim_amd import pack_2bit
    pw = workloads.sbm_partition(n_blocks=8, block_size=2000, deg_in=8, deg_out=2, n_parts=2, part=0,
                                 ticks=4, seed=29)
    g = pw.graph
    n = pw.halo_ids.shape[0]
    codes = np.random.default_rng(5).integers(0, 4, size=n).astype(np.int8)
    got = []
    for packed in (False, True):
        rep = Replica(eng, g, pw.state, seir_model(0.05, 0.2, 0.1), pw.seed, rng_key_id=pw.local_ids)
        eng.map_build(rep.rep, pw.ghost_ids, np.arange(g.n, g.n_cols, dtype=np.int32), 0.5)
        assert eng.halo_bind(rep.rep, pw.halo_ids) == 0
        if packed:
            eng.halo_unpack_2bit(rep.rep, pack_2bit(codes), n)
        else:
            eng.halo_unpack(rep.rep, codes)
        eng.simulate(2)
        slots = np.random.default_rng(6).integers(0, g.n_cols, size=1001).astype(np.int32)
        got.append((rep.states(), eng.halo_pack(rep.rep, slots), eng.halo_pack_2bit(rep.rep, slots)))
        rep.remove()
    assert first_diff(got[1][0], got[0][0]) is None
    assert np.array_equal(got[1][2], pack_2bit(got[1][1]))
    assert np.array_equal(got[0][2], got[1][2])


def test_plain_ticks_after_halo_ticks(eng):
    """Ticks fed by a halo (whose ghost-only tiles the tick skips), then plain
    ticks on the same replica (every tile runs, ghosts keep the last halo's
    states): states and history equal the oracle's two-phase run."""
    pw = workloads.sbm_partition(n_blocks=8, block_size=20000, deg_in=10, deg_out=3, n_parts=2, part=1,
                                 ticks=24, seed=41)
    g = pw.graph
    t1 = 14
    say("halo then plain", g.n, g.n_cols, g.nnz)
    vals = pw.halo_vals(0, t1)
    hist = pw.history.copy()
    om = O.sir_model(pw.beta, pw.gamma, pw.seed)
    mid, _ = O.sir_run(om, g.row_ptr, g.col_idx, pw.local_ids, pw.state, t1, n_local=g.n, history=hist,
                       ghost_ids=pw.ghost_ids, halo_ids=pw.halo_ids, halo_vals=vals)
    want, _ = O.sir_run(om, g.row_ptr, g.col_idx, pw.local_ids, mid, pw.ticks - t1, tick0=t1, n_local=g.n,
                        history=hist)
    rep = Replica(eng, g, pw.state, sir_model(pw.beta, pw.gamma), pw.seed, rng_key_id=pw.local_ids,
                  history=pw.history)
    eng.map_build(rep.rep, pw.ghost_ids, np.arange(g.n, g.n_cols, dtype=np.int32), 0.5)
    assert eng.halo_bind(rep.rep, pw.halo_ids) == 0
    eng.simulate_halo(rep.rep, t1, vals)
    assert first_diff(rep.states(), mid) is None
    eng.simulate(pw.ticks - t1)
    got = rep.states()
    assert first_diff(got, want) is None, first_diff(got, want)
    assert np.array_equal(rep.history(), hist)
    rep.remove()
