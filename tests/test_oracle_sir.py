"""Pins the oracle's tick semantics (oracle/sir_oracle.c).

* The reference's own known answers, run through the oracle's restatement of
  the reference SIR test step (float32 {1,2,3} states, Philox-2x32 draws):
  tests/test_simulation.py:95-115 (chain wave) and
  tests/test_double_buffer.py:247-271 (hierarchical spread).
* Closed forms of the int8 compartment model: beta = 0, gamma = 1,
  beta = 1 BFS front, monotone S -> I -> R.
* Brute force on tiny graphs with exact rational arithmetic (fractions) and
  the big-int Philox transcription, sharing no code with the C oracle.
"""
import random as stdlib_random
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle as O
from oracle import philox_py as P
from sagesim_amd import graphs


# ---------------- the reference's known answers ----------------

def _chain(n):
    return graphs.from_edges(n, np.arange(n - 1), np.arange(1, n))


def test_reference_chain_wave():
    """tests/test_simulation.py:95-115: 10-chain, agent 0 infected, p = 1, 5 ticks
    -> agents 0-4 recovered (3), agent 5 infected (2), 6-9 susceptible (1)."""
    g = _chain(10)
    st = np.ones(10, np.float32)
    st[0] = 2
    out = O.ref_sir_run(g.row_ptr, g.col_idx, np.arange(10, dtype=np.float32), st, 42, 5, 1.0, 1.0)
    assert out.tolist() == [3] * 5 + [2] + [1] * 4


def _hierarchical_edges(total_agents=111):
    """The directed 1 -> 10 -> 100 network of tests/test_double_buffer.py:206-221,
    drawn with the same stdlib random calls (seed 46)."""
    edges = [(0, mid) for mid in range(1, 11)]
    stdlib_random.seed(46)
    end_agents = list(range(11, 111))
    for idx, mid in enumerate(range(1, 11)):
        assigned = end_agents[idx * 10: idx * 10 + 10]
        edges += [(mid, ea) for ea in assigned]
        remaining = [a for a in end_agents if a not in assigned]
        k = min(stdlib_random.randint(1, 10), len(remaining))
        edges += [(mid, ea) for ea in stdlib_random.sample(remaining, k)]
    return total_agents, edges


def _hier_graph():
    n, edges = _hierarchical_edges()
    u, v = zip(*edges)
    # Model.connect_agents -> NetworkSpace.connect_agents(directed=False) (space.py:107-125)
    return graphs.from_edges(n, u, v, directed=False)


def test_reference_hierarchy_si_one_tick():
    """tests/test_double_buffer.py:247-259: after 1 SI tick, root and middle infected, leaves susceptible."""
    g = _hier_graph()
    st = np.ones(g.n, np.float32)
    st[0] = 2
    out = O.ref_sir_run(g.row_ptr, g.col_idx, np.arange(g.n, dtype=np.float32), st, 1234, 1, 1.0, 0.0)
    assert out[0] == 2 and (out[1:11] == 2).all() and (out[11:] == 1).all()


def test_reference_hierarchy_sir_two_ticks():
    """tests/test_double_buffer.py:261-271: after 2 SIR ticks (p = 1), root and middle
    recovered, leaves infected."""
    g = _hier_graph()
    st = np.ones(g.n, np.float32)
    st[0] = 2
    out = O.ref_sir_run(g.row_ptr, g.col_idx, np.arange(g.n, dtype=np.float32), st, 99, 2, 1.0, 1.0)
    assert out[0] == 3 and (out[1:11] == 3).all() and (out[11:] == 2).all()


def _f32_id(x):
    """The reference stores logical ids as float32 (model.py:1085-1096) and the
    RNG keys on int() of that value (model.py:1934-1951, math_utils.py:109)."""
    return int(np.float32(float(x)))


def _brute_ref_sir(rows, logical_ids, state, seed, tick0, ticks, p_inf, p_rec, le=False):
    """tests/test_simulation.py:24-43 with every tick's transitions enumerated
    from the big-int Philox words and compared as exact rationals; returns
    the state after every tick.  le=True compares with <= (a planted mistake)."""
    st = [int(x) for x in state]
    out = []
    below = (lambda a, b: a <= b) if le else (lambda a, b: a < b)
    for t in range(tick0, tick0 + ticks):
        nxt = list(st)
        for r, s in enumerate(st):
            word = P.rand_uniform_philox_words(seed, t, _f32_id(logical_ids[r]), 1)
            u = Fraction(word % 2 ** 23, 2 ** 23)
            if s == 2 and below(u, Fraction(p_rec)):
                nxt[r] = 3
            elif s == 1 and below(u, Fraction(p_inf)) and any(st[j] == 2 for j in rows[r]):
                nxt[r] = 2
        st = nxt
        out.append(list(st))
    return out


def _word_u(seed, t, lid):
    return (P.rand_uniform_philox_words(seed, t, _f32_id(lid), 1) % 2 ** 23) / 2 ** 23


def _fractional_case(case):
    """A graph of at most 8 agents, permuted logical ids (some above 2^24),
    and p_infection / p_recovery set exactly to uniforms that a susceptible
    and an infectious agent draw at the first tick."""
    rng = stdlib_random.Random(100 + case)
    n = rng.randint(4, 8)
    edges = {(i, j) for i in range(n) for j in range(n) if i != j and rng.random() < 0.45}
    rows = [[j for (i, j) in sorted(edges) if i == r] for r in range(n)]
    base = rng.choice([0, 2 ** 24 - 3, 2 ** 24 + 1, 2 ** 31 + 7, 2 ** 33])
    lids = [base + k for k in rng.sample(range(3 * n), n)]
    seed = rng.randrange(2 ** 31)
    state = [rng.choice([1, 1, 2, 3]) for _ in range(n)]
    if 2 not in state:
        state[0] = 2
    tick0 = rng.randrange(0, 5000)
    s_agents = [r for r in range(n) if state[r] == 1 and any(state[j] == 2 for j in rows[r])]
    s_agents = s_agents or [r for r in range(n) if state[r] == 1] or [0]
    i_agent = [r for r in range(n) if state[r] == 2][0]
    p_inf = _word_u(seed, tick0, lids[s_agents[0]]) or 0.5
    p_rec = _word_u(seed, tick0, lids[i_agent]) or 0.5
    return rows, lids, state, seed, tick0, p_inf, p_rec, i_agent


N_FRACTIONAL = 8


@pytest.mark.parametrize("case", range(N_FRACTIONAL))
def test_reference_step_brute_force_at_fractional_p(case):
    """or_ref_sir_run at 0 < p < 1 against the brute force, every tick."""
    rows, lids, state, seed, tick0, p_inf, p_rec, i_agent = _fractional_case(case)
    n, ticks = len(rows), 6
    rp = np.zeros(n + 1, np.int64)
    rp[1:] = np.cumsum([len(r) for r in rows])
    ci = np.array([j for r in rows for j in r], np.int32)
    want = _brute_ref_sir(rows, lids, state, seed, tick0, ticks, p_inf, p_rec)
    lid32 = np.array([np.float32(float(x)) for x in lids], np.float32)
    st = np.array(state, np.float32)
    for k in range(ticks):
        st = O.ref_sir_run(rp, ci, lid32, st, seed, 1, p_inf, p_rec, tick0=tick0 + k)
        assert st.astype(int).tolist() == want[k], (case, k)
    # u == p_recovery at the first tick: the agent stays infectious
    assert want[0][i_agent] == 2


def test_reference_step_brute_force_catches_planted_mistakes():
    """The cases above tell the right rule from `<=` and from draws keyed on
    the slot instead of the float32 logical id: each mistake, planted in the
    brute force, changes at least one case's trajectory."""
    differs_le = differs_slot = 0
    for case in range(N_FRACTIONAL):
        rows, lids, state, seed, tick0, p_inf, p_rec, _ = _fractional_case(case)
        want = _brute_ref_sir(rows, lids, state, seed, tick0, 6, p_inf, p_rec)
        differs_le += _brute_ref_sir(rows, lids, state, seed, tick0, 6, p_inf, p_rec, le=True) != want
        differs_slot += _brute_ref_sir(rows, list(range(len(rows))), state, seed, tick0, 6, p_inf, p_rec) != want
    assert differs_le >= 1 and differs_slot >= 1, (differs_le, differs_slot)


def test_reference_step_float32_ids_share_draws():
    """Ids 2^24 and 2^24 + 1 are one float32 value: both agents draw the same
    uniform (model.py:1085-1096 stores logical ids as float32)."""
    assert _f32_id(2 ** 24 + 1) == 2 ** 24
    lids = np.array([np.float32(2 ** 24), np.float32(2 ** 24 + 1)], np.float32)
    assert lids[0] == lids[1]
    rp = np.array([0, 1, 2], np.int64)
    ci = np.array([1, 0], np.int32)
    for seed in range(40):
        st = O.ref_sir_run(rp, ci, lids, np.array([2, 2], np.float32), seed, 1, 0.5, 0.5)
        assert st[0] == st[1]


def test_reference_seed_reproducibility():
    """tests/test_simulation.py:117-159: same seed -> same states; different seeds differ."""
    g = _chain(50)
    st = np.ones(50, np.float32)
    st[0] = 2
    ids = np.arange(50, dtype=np.float32)
    a = O.ref_sir_run(g.row_ptr, g.col_idx, ids, st, 999, 10, 0.5, 0.3)
    b = O.ref_sir_run(g.row_ptr, g.col_idx, ids, st, 999, 10, 0.5, 0.3)
    assert a.tolist() == b.tolist()
    c = O.ref_sir_run(g.row_ptr, g.col_idx, ids, st, 111, 20, 0.3, 0.2)
    d = O.ref_sir_run(g.row_ptr, g.col_idx, ids, st, 222, 20, 0.3, 0.2)
    assert c.tolist() != d.tolist()


# ---------------- closed forms of the int8 compartment model ----------------

def _run(model, g, st, ticks, **kw):
    return O.sir_run(model, g.row_ptr, g.col_idx, np.arange(g.n), st, ticks, trace=True, **kw)


def test_beta_zero_no_new_infections():
    g = graphs.erdos_renyi(2000, 0.005, 3)
    st = np.zeros(g.n, np.int8)
    st[:50] = 1
    _, tr = _run(O.sir_model(0.0, 0.2, 5), g, st, 30)
    assert ((tr == 1) | (tr == 2)).sum(axis=1).max() == 50
    assert (tr[:, 50:] == 0).all()


def test_gamma_one_recovers_next_tick():
    g = graphs.erdos_renyi(2000, 0.005, 4)
    st = np.zeros(g.n, np.int8)
    st[:50] = 1
    _, tr = _run(O.sir_model(0.3, 1.0, 6), g, st, 10)
    prev = st
    for t in range(10):
        assert (tr[t][prev == 1] == 2).all()          # every I recovers
        new_i = (tr[t] == 1) & (prev == 0)
        assert (tr[t] == 1).sum() == new_i.sum()       # the only I are fresh infections
        prev = tr[t]


def test_beta_one_bfs_front():
    """beta = 1, gamma = 0 on a chain: agent i becomes I exactly at tick i."""
    g = _chain(40)
    st = np.zeros(40, np.int8)
    st[0] = 1
    _, tr = _run(O.sir_model(1.0, 0.0, 7), g, st, 39)
    for t in range(39):
        assert (tr[t][:t + 2] == 1).all() and (tr[t][t + 2:] == 0).all()


def test_monotone_and_conserved():
    g = graphs.barabasi_albert(5000, 3, 8)
    st = np.zeros(g.n, np.int8)
    st[:20] = 1
    _, tr = _run(O.seir_model(0.08, 0.3, 0.15, 9), g, st, 60)
    full = np.vstack([st[None, :], tr])
    assert (np.diff(full.astype(np.int16), axis=0) >= 0).all()    # S->E->I->R only moves up
    assert (np.diff(full.astype(np.int16), axis=0) <= 1).all()    # one compartment per tick


# ---------------- brute force with exact arithmetic ----------------

def _brute(row_ptr, col_idx, ids, st, ticks, beta, trans, nxt, s_code, inf_code, s_target, seed,
           n_local=None, halo=None, history=None):
    """Straight loops, exact p_k = 1 - (1-beta)^k in rationals, draws from the
    big-int Philox transcription."""
    st = [int(x) for x in st]
    n_total = len(st)
    n_local = n_total if n_local is None else n_local
    b = Fraction(1) - Fraction(beta)
    out = []
    for t in range(ticks):
        if halo is not None:
            for slot, val in halo[t].items():
                st[slot] = val
        new = list(st)
        for i in range(n_local):
            m = P.sir_word(seed, t, int(ids[i])) & 0x7FFFFF
            u = Fraction(m, 1 << 23)
            s = st[i]
            if s == s_code:
                nb = [int(c) for c in col_idx[row_ptr[i]:row_ptr[i + 1]]]
                inf = [c for c in nb if st[c] == inf_code]
                k = len(inf)
                if k and u < 1 - b ** k:
                    new[i] = s_target
                    if history is not None:
                        history[i].append(inf[0])
            elif trans[s] > 0 and u < Fraction(trans[s]):
                new[i] = nxt[s]
        st = new
        out.append(list(st))
    return out


@pytest.mark.parametrize("seed", [1, 2, 3, 4, 5, 6])
def test_brute_force_tiny_graphs(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 9))
    ne = int(rng.integers(0, n * 3))
    u = rng.integers(0, n, ne)
    v = rng.integers(0, n, ne)
    g = graphs.from_edges(n, u, v)
    ids = rng.integers(0, 2**40, n)
    beta = float(rng.uniform(0.05, 0.9))
    gamma = float(rng.uniform(0.05, 0.9))
    sigma = float(rng.uniform(0.05, 0.9))
    rseed = int(rng.integers(0, 2**63))
    seir = seed % 2 == 0
    st = rng.integers(0, 4 if seir else 3, n).astype(np.int8)
    if seir:
        om = O.seir_model(beta, sigma, gamma, rseed)
        trans, nxt, inf_code = [0.0, sigma, gamma, 0.0], [0, 2, 3, 3], 2
    else:
        om = O.sir_model(beta, gamma, rseed)
        trans, nxt, inf_code = [0.0, gamma, 0.0, 0.0], [0, 2, 2, 3], 1
    hist = np.full((n, 32), -1, np.int32)
    _, tr = O.sir_run(om, g.row_ptr, g.col_idx, ids, st, 25, trace=True, history=hist)
    bh = [[] for _ in range(n)]
    want = _brute(g.row_ptr, g.col_idx, ids, st, 25, beta, trans, nxt, 0, inf_code, 1, rseed, history=bh)
    assert tr.tolist() == want
    for i in range(n):
        assert hist[i][:len(bh[i])].tolist() == bh[i] and (hist[i][len(bh[i]):] == -1).all()


def test_brute_force_ghosts_and_halo():
    """Ghost slots take the halo's states before every tick, whatever the record order."""
    rng = np.random.default_rng(11)
    n_local, n_ghost = 6, 4
    u = rng.integers(0, n_local, 14)
    v = rng.integers(0, n_local + n_ghost, 14)
    g = graphs.from_edges(n_local + n_ghost, u, v)
    # keep rows of local agents only
    rp = g.row_ptr[:n_local + 1].copy()
    ci = g.col_idx[:rp[-1]].copy()
    ghost_ids = rng.integers(0, 2**50, n_ghost)
    order = rng.permutation(n_ghost)
    halo_ids = ghost_ids[order]
    ticks = 12
    halo_vals = rng.integers(0, 3, (ticks, n_ghost)).astype(np.int8)
    ids = np.arange(n_local) + 1000
    st = np.zeros(n_local + n_ghost, np.int8)
    st[0] = 1
    seed = 77
    _, tr = O.sir_run(O.sir_model(0.5, 0.3, seed), rp, ci, ids, st, ticks, n_local=n_local, trace=True,
                      ghost_ids=ghost_ids, halo_ids=halo_ids, halo_vals=halo_vals)
    halo = [{n_local + int(order[h]): int(halo_vals[t, h]) for h in range(n_ghost)} for t in range(ticks)]
    want = _brute(rp, ci, ids, st, ticks, 0.5, [0.0, 0.3, 0.0, 0.0], [0, 2, 2, 3], 0, 1, 1, seed,
                  n_local=n_local, halo=halo)
    assert tr.tolist() == want
    # an unknown halo id is an error, not a silent skip
    with pytest.raises(ValueError):
        O.sir_run(O.sir_model(0.5, 0.3, seed), rp, ci, ids, st, 1, n_local=n_local, ghost_ids=ghost_ids,
                  halo_ids=np.array([123], np.int64), halo_vals=np.zeros((1, 1), np.int8))
