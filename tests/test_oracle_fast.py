"""The tuned CPU restatement (or_sir_run_fast) gives or_sir_run's results bit for bit.

or_sir_run_fast is the CPU baseline's fast leg: it draws first, compares
integer thresholds ceil(p * 2^23), skips rows that cannot reach a threshold and
stops counting at the first one reached.  Every case here runs both functions
on the same inputs and compares whole traces, histories and final states; the
boundary case puts u exactly on p, where a threshold off by one would differ.
"""
import numpy as np
import pytest

from oracle import oracle as O
from sagesim_amd import graphs


def _both(model, g, ids, st, ticks, **kw):
    out = []
    for fast in (False, True):
        hist = np.full((len(g.row_ptr) - 1 if kw.get("n_local") is None else kw["n_local"], 32), -1, np.int32)
        fin, tr = O.sir_run(model, g.row_ptr, g.col_idx, ids, st, ticks, trace=True, history=hist,
                            fast=fast, **kw)
        out.append((fin, tr, hist))
    return out


def _seed_states(n, n_states, frac, rng):
    st = np.zeros(n, np.int8)
    st[rng.choice(n, max(1, int(frac * n)), replace=False)] = 1 if n_states == 3 else 2
    return st


@pytest.mark.parametrize("kind,threads", [("sir", 1), ("sir", 4), ("seir", 4), ("sirs", 4), ("beta1", 2)])
def test_fast_matches_plain_on_ba(kind, threads):
    rng = np.random.default_rng(5)
    g = graphs.barabasi_albert(30000, 4, 17)
    ids = graphs.scramble_ids(np.arange(g.n, dtype=np.int64), 3)
    if kind == "seir":
        om = O.seir_model(0.08, 0.3, 0.1, 1234)
        st = _seed_states(g.n, 4, 0.002, rng)
    else:
        om = O.sir_model(1.0 if kind == "beta1" else 0.05, 0.1, 1234)
        if kind == "sirs":
            om.trans_p[2] = 0.05
            om.trans_next[2] = 0
        st = _seed_states(g.n, 3, 0.002, rng)
    (a, ta, ha), (b, tb, hb) = _both(om, g, ids, st, 60, nthreads=threads)
    assert np.array_equal(ta, tb)
    assert np.array_equal(a, b)
    assert np.array_equal(ha, hb)
    # the epidemic moved, so the comparison covered infections
    assert (ta[-1] != st).sum() > 1000


def test_fast_matches_plain_with_ghosts_and_halo():
    rng = np.random.default_rng(9)
    n_local, n_ghost = 4000, 600
    u = rng.integers(0, n_local, 24000)
    v = rng.integers(0, n_local + n_ghost, 24000)
    full = graphs.from_edges(n_local + n_ghost, u, v)
    rp = full.row_ptr[:n_local + 1].copy()
    ci = full.col_idx[:rp[-1]].copy()
    g = type("G", (), {"row_ptr": rp, "col_idx": ci})
    ghost_ids = rng.permutation(10**6)[:n_ghost] + 10**7
    order = rng.permutation(n_ghost)
    ticks = 30
    halo_vals = (rng.random((ticks, n_ghost)) < 0.2).astype(np.int8)
    ids = np.arange(n_local, dtype=np.int64) * 7 + 1
    st = np.zeros(n_local + n_ghost, np.int8)
    st[rng.choice(n_local, 10, replace=False)] = 1
    res = _both(O.sir_model(0.2, 0.2, 4321), g, ids, st, ticks, n_local=n_local, ghost_ids=ghost_ids,
                halo_ids=ghost_ids[order], halo_vals=halo_vals[:, order], nthreads=3)
    (a, ta, ha), (b, tb, hb) = res
    assert np.array_equal(ta, tb) and np.array_equal(a, b) and np.array_equal(ha, hb)


def test_fast_threshold_at_exact_boundary():
    """u == p does not infect (u < p is strict); above it does, also for p between two u values
    (a threshold rounded down instead of up would fail the m + 0.5 case)."""
    seed, tick, agent_id = 99, 3, 12345
    w = int(O.lib().or_sir_word(seed, tick, agent_id))
    m = w & 0x7FFFFF
    assert 0 < m < 0x7FFFFF
    g = graphs.from_edges(2, np.array([0]), np.array([1]))
    ids = np.array([agent_id, 7], np.int64)
    st = np.array([0, 1], np.int8)
    cases = ((m / 2**23, False), ((m + 1) / 2**23, True), ((m - 1) / 2**23, False),
             ((m + 0.5) / 2**23, True), ((m - 0.5) / 2**23, False))
    for beta, infected in cases:
        om = O.sir_model(beta, 0.0, seed)
        for fast in (False, True):
            fin, _ = O.sir_run(om, g.row_ptr, g.col_idx, ids, st, 1, tick0=tick, fast=fast)
            assert (fin[0] == 1) == infected, (beta, fast)


def test_fast_matches_plain_outside_unit_beta():
    """beta outside [0, 1] (p_k not monotone) takes the plain restatement's path."""
    rng = np.random.default_rng(2)
    g = graphs.erdos_renyi(3000, 0.003, 8)
    ids = np.arange(g.n, dtype=np.int64)
    st = _seed_states(g.n, 3, 0.05, rng)
    (a, ta, _), (b, tb, _) = _both(O.sir_model(1.5, 0.3, 5), g, ids, st, 10, nthreads=2)
    assert np.array_equal(ta, tb) and np.array_equal(a, b)
