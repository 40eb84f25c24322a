"""Batched ensembles: many independent SIR/SEIR replicas advanced together.

All replicas of an ensemble live in one libsagestep context.  Their tiles form
one list (a block-diagonal CSR in effect: each replica keeps its own row_ptr,
col_idx, states and model parameters, and tiles never cross replicas), so one
persistent launch per tick advances every replica, and a whole tick loop is one
hipGraph replay.  Each replica draws from its own key (seed, local agent id,
tick), so its states are bit-identical to a standalone run of the same
scenario, whatever else shares the launch.
"""
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field
import math
import time

import numpy as np

from . import graphs
from .engine import Engine, Replica, ReplicaData, prepare_replica, seir_model, sir_model
from .workloads import seed_infected


@dataclass
class Scenario:
    """One SIR (sigma == 0) or SEIR scenario on a seeded BA or WS graph."""
    n: int
    graph_kind: str            # "ba" or "ws"
    graph_param: int           # m for BA, k for WS
    beta: float
    gamma: float
    sigma: float
    seed: int
    ticks: int
    frac_infected: float = 0.001
    ws_rewire: float = 0.1

    def model(self):
        return seir_model(self.beta, self.sigma, self.gamma) if self.sigma > 0 else sir_model(self.beta, self.gamma)

    def build_graph(self):
        if self.graph_kind == "ba":
            return graphs.barabasi_albert(self.n, self.graph_param, self.seed)
        if self.graph_kind == "ws":
            return graphs.watts_strogatz(self.n, self.graph_param, self.ws_rewire, self.seed)
        raise ValueError(f"unknown graph kind {self.graph_kind}")

    def initial_state(self):
        code = 2 if self.sigma > 0 else 1
        return seed_infected(self.n, max(1, int(round(self.frac_infected * self.n))), self.seed + 1, code)


def ensemble_scenarios(n_rep=256, lo=10_000, hi=2_000_000, ticks=300, seed=4242):
    """Replicas of log-uniform size in [lo, hi], BA with m ~ U{2..8},
    beta ~ U(0.01, 0.08), gamma ~ U(0.05, 0.2), one seed each."""
    rng = np.random.Generator(np.random.PCG64(seed))
    sizes = np.exp(rng.uniform(math.log(lo), math.log(hi), n_rep)).astype(np.int64)
    ms = rng.integers(2, 9, n_rep)
    betas = rng.uniform(0.01, 0.08, n_rep)
    gammas = rng.uniform(0.05, 0.2, n_rep)
    seeds = rng.integers(1, 2**62, n_rep)
    return [Scenario(int(sizes[i]), "ba", int(ms[i]), float(betas[i]), float(gammas[i]), 0.0,
                     int(seeds[i]), ticks) for i in range(n_rep)]


@dataclass
class Prepared:
    scenario: Scenario
    graph: graphs.CSRGraph
    state: np.ndarray
    gen_s: float = 0.0
    data: ReplicaData = None

    def replica_data(self):
        if self.data is None:
            self.data = prepare_replica(self.graph, self.state)
        return self.data


def prepare(scenarios, threads=8):
    """Build graphs, initial states and their degree-sorted storage order on a
    host thread pool (the generators release the GIL)."""
    def one(sc):
        graphs.set_threads(1)
        t0 = time.perf_counter()
        g = sc.build_graph()
        st = sc.initial_state()
        return Prepared(sc, g, st, time.perf_counter() - t0, prepare_replica(g, st))
    with ThreadPoolExecutor(max_workers=max(1, threads)) as ex:
        return list(ex.map(one, scenarios))


@dataclass
class Ensemble:
    """Replicas of prepared scenarios on one engine; `run` advances all of
    them until each reaches its own tick count."""
    engine: Engine
    replicas: list = field(default_factory=list)
    ticks_needed: int = 0

    @classmethod
    def build(cls, engine, prepared):
        ens = cls(engine)
        for p in prepared:
            sc = p.scenario
            ens.replicas.append(Replica(engine, model=sc.model(), seed=sc.seed, tick_limit=sc.ticks,
                                        data=p.replica_data()))
            ens.ticks_needed = max(ens.ticks_needed, sc.ticks)
        return ens

    def run(self, ticks):
        self.engine.simulate(ticks)

    def reset(self):
        for r in self.replicas:
            r.reset()

    def states(self, i):
        return self.replicas[i].states()

    def remove(self):
        for r in self.replicas:
            r.remove()
        self.replicas = []


def run_standalone(engine, prepared):
    """The plain per-scenario launch loop: each scenario alone on the engine
    (its own tile list and captured tick loop), run to completion.  Only the
    tick loops are timed, after one untimed run that captures the graph, so
    the figure compares launch strategies on resident inputs.
    Returns (final states, seconds)."""
    out = []
    total = 0.0
    for p in prepared:
        sc = p.scenario
        r = Replica(engine, model=sc.model(), seed=sc.seed, data=p.replica_data())
        engine.simulate(sc.ticks)
        r.reset()
        engine.sync()
        t0 = time.perf_counter()
        engine.simulate(sc.ticks)
        engine.sync()
        total += time.perf_counter() - t0
        out.append(r.states())
        r.remove()
    return out, total


def run_batched(engine, ens):
    """All replicas of `ens` in one launch per tick until the longest finishes.
    Timed after one untimed run that captures the graph.  Returns seconds."""
    ticks = ens.ticks_needed
    ens.run(ticks)
    ens.reset()
    engine.sync()
    t0 = time.perf_counter()
    ens.run(ticks)
    engine.sync()
    return time.perf_counter() - t0
