"""Seeded synthetic SIR/SEIR workloads.

Each function returns plain arrays (graph, initial states, model parameters,
seed) so the same input can be fed to the HIP path and to any checker.
"""
from dataclasses import dataclass

import numpy as np

from . import graphs
from .engine import sir_model, seir_model


@dataclass
class Workload:
    name: str
    graph: graphs.CSRGraph
    state: np.ndarray          # int8 initial states (local + ghost slots)
    model: object              # SSModel
    beta: float
    gamma: float
    sigma: float
    seed: int
    ticks: int
    rng_key_id: np.ndarray = None


def seed_infected(n, count, seed, code=1):
    """int8 states: `count` agents chosen without replacement by seed are `code`, the rest 0."""
    st = np.zeros(n, np.int8)
    if count > 0:
        idx = np.random.Generator(np.random.PCG64(seed)).choice(n, size=count, replace=False)
        st[idx] = code
    return st


def er_sir(n=10_000, p=0.001, beta=0.05, gamma=0.1, frac_infected=0.01, ticks=100, seed=2024):
    """Plain SIR on G(n, p): the reference-CPU-path workload (10k agents, 100 ticks)."""
    g = graphs.erdos_renyi(n, p, seed)
    st = seed_infected(n, int(round(frac_infected * n)), seed + 1)
    return Workload(f"er_sir_n{n}", g, st, sir_model(beta, gamma), beta, gamma, 0.0, seed, ticks)


def ba_sir(n=10_000_000, m=5, beta=0.03, gamma=0.1, frac_infected=0.001, ticks=200, seed=7):
    """Plain SIR on Barabasi-Albert(n, m): the single-GPU workload (10M agents, 200 ticks)."""
    g = graphs.barabasi_albert(n, m, seed)
    st = seed_infected(n, max(1, int(round(frac_infected * n))), seed + 1)
    return Workload(f"ba_sir_n{n}_m{m}", g, st, sir_model(beta, gamma), beta, gamma, 0.0, seed, ticks)


def ws_seir(n=100_000, k=10, p=0.1, beta=0.05, sigma=0.2, gamma=0.1, frac_infected=0.01, ticks=100, seed=11):
    g = graphs.watts_strogatz(n, k, p, seed)
    st = seed_infected(n, max(1, int(round(frac_infected * n))), seed + 1, code=2)
    return Workload(f"ws_seir_n{n}", g, st, seir_model(beta, sigma, gamma), beta, gamma, sigma, seed, ticks)


@dataclass
class PartitionWorkload:
    """One partition's share of an SBM run: local rows, ghost slots, halo feed."""
    graph: graphs.CSRGraph
    state: np.ndarray            # local + ghost initial states
    ghost_ids: np.ndarray        # int64 id of ghost slot n_local + j
    local_ids: np.ndarray        # int64 global id of every local agent (RNG key)
    halo_ids: np.ndarray         # record order of the halo buffer
    history: np.ndarray          # [n_local, 32] int32, -1 padded
    beta: float
    gamma: float
    seed: int
    ticks: int

    def halo_vals(self, t0, t1, out=None):
        """States carried by the halo for ticks [t0, t1): each ghost turns I at a
        seeded tick and R a seeded number of ticks later.  Written into `out`
        (shape (t1 - t0, n_halo), int8) when given."""
        h = self.halo_ids.shape[0]
        rng = np.random.Generator(np.random.PCG64(self.seed + 99))
        t_inf = rng.integers(0, 2 * self.ticks, size=h)
        dur = rng.integers(1, 15, size=h)
        t_rec = t_inf + dur
        v = np.empty((max(0, t1 - t0), h), np.int8) if out is None else out
        assert v.shape == (max(0, t1 - t0), h) and v.dtype == np.int8
        for k, t in enumerate(range(t0, t1)):
            v[k] = t >= t_inf
            v[k] += t >= t_rec
        return v


def sbm_partition(n_blocks=64, block_size=500_000, deg_in=12.0, deg_out=2.0, n_parts=8, part=0,
                  beta=0.05, gamma=0.1, frac_infected=0.001, ticks=150, seed=31):
    g = graphs.sbm_part(n_blocks, block_size, deg_in, deg_out, n_parts, part, seed)
    n_local = g.n
    n_ghost = g.n_cols - g.n
    local_ids = np.arange(part * n_local, (part + 1) * n_local, dtype=np.int64)
    ghost_ids = graphs.scramble_ids(g.extra, seed + 5)
    st = np.zeros(g.n_cols, np.int8)
    st[:n_local] = seed_infected(n_local, max(1, int(round(frac_infected * n_local))), seed + 1)
    halo_ids = ghost_ids[np.random.Generator(np.random.PCG64(seed + 2)).permutation(n_ghost)]
    rng = np.random.Generator(np.random.PCG64(seed + 3))
    lens = rng.integers(0, 4, size=n_local)
    hist = np.full((n_local, 32), -1, np.int32)
    for L in range(1, 4):
        rows = np.nonzero(lens >= L)[0]
        hist[rows, L - 1] = rng.integers(0, g.n_cols, size=rows.shape[0], dtype=np.int32)
    return PartitionWorkload(g, st, ghost_ids, local_ids, halo_ids, hist, beta, gamma, seed, ticks)
