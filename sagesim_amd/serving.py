"""Serving loop: continuous batching of SIR/SEIR scenarios on one GPU.

Requests arrive (Poisson), a host thread pool builds their graphs, and the
loop admits built requests onto the device while their HBM footprint fits the
budget.  Every admitted request becomes a replica of the one engine, so all
in-flight scenarios advance together in the same persistent tick launch; the
loop runs short chunks of ticks, retires replicas that reached their own tick
count (download, free), admits newcomers, and repeats.  A scenario's result
does not depend on what it shared the launch with: each replica draws from
its own (seed, agent, tick) key and its tiles never touch other replicas.
"""
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field
import math
import threading
import time

import numpy as np

from . import graphs
from .engine import prepare_replica
from .ensemble import Scenario


def serving_requests(rate=50.0, duration=600.0, lo=50_000, hi=5_000_000, seed=777):
    """Poisson arrivals at `rate` per second for `duration` seconds.  Each request:
    SIR or SEIR (E -> I with sigma ~ U(0.1, 0.3)), size log-uniform in [lo, hi],
    BA (m in {3..6}) or WS (k = 10, rewire 0.1), 100-500 ticks, its own seed;
    beta ~ U(0.01, 0.08), gamma ~ U(0.05, 0.2)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    t = 0.0
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            break
        seir = bool(rng.integers(0, 2))
        ba = bool(rng.integers(0, 2))
        n = int(math.exp(rng.uniform(math.log(lo), math.log(hi))))
        sc = Scenario(n=n, graph_kind="ba" if ba else "ws", graph_param=int(rng.integers(3, 7)) if ba else 10,
                      beta=float(rng.uniform(0.01, 0.08)), gamma=float(rng.uniform(0.05, 0.2)),
                      sigma=float(rng.uniform(0.1, 0.3)) if seir else 0.0,
                      seed=int(rng.integers(1, 2**62)), ticks=int(rng.integers(100, 501)))
        out.append((t, sc))
    return out


def footprint(graph, with_history=False):
    """Device bytes of a replica, as the library accounts them (ss_replica_footprint)."""
    from ._lib import step_lib
    return int(step_lib().ss_replica_footprint(graph.n, graph.n_cols - graph.n, graph.nnz, 1 if with_history else 0,
                                               int(graph.degrees().max(initial=0))))


@dataclass
class Request:
    rid: int
    arrival: float                 # seconds after start
    scenario: Scenario
    graph: object = None
    state: object = None
    nbytes: int = 0
    admitted: float = None
    done: float = None
    result: np.ndarray = None
    replica: object = None
    data: object = None                # ReplicaData prepared by a builder thread
    deadline: float = None             # seconds after arrival; None: no deadline
    outcome: str = None                # "done", "rejected", "cancelled" or "expired"


@dataclass
class ServeStats:
    completed: int = 0
    submitted: int = 0
    rejected: int = 0
    cancelled: int = 0
    expired: int = 0
    chunks: int = 0
    span_s: float = 0.0
    latencies: list = field(default_factory=list)
    max_inflight: int = 0
    max_bytes: int = 0

    def summary(self):
        lat = np.array(self.latencies) if self.latencies else np.zeros(1)
        return {"completed": self.completed, "submitted": self.submitted, "rejected": self.rejected,
                "cancelled": self.cancelled, "expired": self.expired,
                "scenarios_per_s": self.completed / self.span_s if self.span_s > 0 else 0.0,
                "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
                "max_latency_s": float(lat.max()), "chunks": self.chunks, "max_inflight": self.max_inflight,
                "max_device_bytes": self.max_bytes, "span_s": self.span_s}


class Server:
    """Continuous-batching loop over one engine.

    make_replica(req) -> replica handle with .states(), .tick(), .remove();
    the engine needs .simulate(ticks) and .sync().  Both are injectable so the
    admission and retirement logic runs without a GPU.

    A request leaves early when cancel(rid) is called (from any thread) or when
    its deadline (seconds after arrival: the arrival tuple's third element, or
    `deadline_s` for all) passes before it finishes: wherever it is (not yet
    arrived, being built, waiting for memory, or on the device, where its
    replica is removed and its bytes freed at the next chunk boundary), it
    ends with outcome "cancelled" or "expired" and no result.  The others'
    results do not change: replicas never touch each other's state."""

    def __init__(self, engine, make_replica, budget_bytes=200 << 30, chunk_ticks=10, builders=8,
                 keep_results=True, clock=time.perf_counter, builder_threads=1, deadline_s=None,
                 progress=None, progress_every_s=30.0):
        # keep_results: True (every finished request keeps its states), False
        # (none), or a collection of request ids whose states are kept (long
        # windows: tens of thousands of results would fill host memory)
        self.builder_threads = builder_threads
        self.engine = engine
        self.make_replica = make_replica
        self.budget = int(budget_bytes)
        self.chunk = int(chunk_ticks)
        self.pool = ThreadPoolExecutor(max_workers=max(1, builders))
        self.keep = keep_results
        self.clock = clock
        self.deadline_s = deadline_s
        self.progress = progress
        self.progress_every = progress_every_s
        self._lock = threading.Lock()
        self._cancel = set()

    def cancel(self, rid):
        """Ask the loop to drop request `rid` (thread-safe; a finished request is unaffected)."""
        with self._lock:
            self._cancel.add(int(rid))

    def _build(self, req):
        graphs.set_threads(self.builder_threads)
        sc = req.scenario
        req.graph = sc.build_graph()
        req.state = sc.initial_state()
        req.nbytes = footprint(req.graph)
        req.data = prepare_replica(req.graph, req.state)      # degree sort off the serving thread
        return req

    def _leaving(self, r, now):
        """'cancelled' or 'expired' when request r is to leave now, else None."""
        if r.rid in self._cancel_seen:
            return "cancelled"
        if r.deadline is not None and now > r.arrival + r.deadline:
            return "expired"
        return None

    def run(self, arrivals, realtime=True):
        """Serve (arrival_time, Scenario[, deadline_s]) tuples; returns (stats, finished requests)."""
        stats = ServeStats(submitted=len(arrivals))
        reqs = []
        for i, a in enumerate(arrivals):
            dl = a[2] if len(a) > 2 else self.deadline_s
            reqs.append(Request(i, a[0], a[1], deadline=dl))
        pending = list(reqs)                 # not yet arrived, by arrival time
        building = []                        # (future, request), in arrival order
        ready = []                           # built, waiting for memory (FIFO)
        inflight = []
        finished = []
        used = 0
        t0 = self.clock()
        last_report = 0.0
        pi = 0
        self._cancel_seen = set()

        def leave(r, why, now):
            r.outcome = why
            r.done = now
            r.graph = r.state = r.data = None
            if why == "cancelled":
                stats.cancelled += 1
            else:
                stats.expired += 1
            finished.append(r)

        while pi < len(pending) or building or ready or inflight:
            now = self.clock() - t0
            with self._lock:
                self._cancel_seen |= self._cancel
            # arrivals: hand to the builders (not realtime: everything has arrived)
            while pi < len(pending) and (not realtime or pending[pi].arrival <= now):
                r = pending[pi]
                pi += 1
                why = self._leaving(r, now)
                if why:
                    leave(r, why, now)
                else:
                    building.append((self.pool.submit(self._build, r), r))
            # requests that have not arrived yet may be cancelled too
            if self._cancel_seen and pi < len(pending):
                keep = pending[:pi]
                for r in pending[pi:]:
                    if r.rid in self._cancel_seen:
                        leave(r, "cancelled", now)
                    else:
                        keep.append(r)
                pending = keep
            # collect built requests in arrival order
            while building and building[0][0].done():
                r = building.pop(0)[0].result()
                why = self._leaving(r, now)
                if why:
                    leave(r, why, now)
                elif r.nbytes > self.budget:
                    stats.rejected += 1          # can never fit: refuse instead of blocking the queue
                    r.outcome = "rejected"
                    r.done = self.clock() - t0
                    finished.append(r)
                else:
                    ready.append(r)
            # queued requests that are to leave
            if ready:
                keep = []
                for r in ready:
                    why = self._leaving(r, now)
                    if why:
                        leave(r, why, now)
                    else:
                        keep.append(r)
                ready = keep
            # admission: FIFO while the footprint fits
            while ready and used + ready[0].nbytes <= self.budget:
                r = ready.pop(0)
                r.replica = self.make_replica(r)
                r.admitted = self.clock() - t0
                used += r.nbytes
                inflight.append(r)
            stats.max_inflight = max(stats.max_inflight, len(inflight))
            stats.max_bytes = max(stats.max_bytes, used)
            if inflight:
                self.engine.simulate(self.chunk)
                stats.chunks += 1
                now = self.clock() - t0
                still = []
                for r in inflight:
                    why = self._leaving(r, now)
                    if why is None and r.replica.tick() >= r.scenario.ticks:
                        keep = self.keep if isinstance(self.keep, bool) else r.rid in self.keep
                        r.result = r.replica.states() if keep else None
                        r.replica.remove()
                        r.replica = None
                        r.graph = r.state = r.data = None
                        used -= r.nbytes
                        r.done = self.clock() - t0
                        r.outcome = "done"
                        stats.completed += 1
                        stats.latencies.append(r.done - r.arrival)
                        finished.append(r)
                    elif why:
                        r.replica.remove()
                        r.replica = None
                        used -= r.nbytes
                        leave(r, why, now)
                    else:
                        still.append(r)
                inflight = still
            else:
                # nothing on the device: wait for the next arrival or build
                if building:
                    building[0][0].result()
                elif pi < len(pending) and realtime:
                    time.sleep(max(0.0, min(0.01, pending[pi].arrival - (self.clock() - t0))))
            if self.progress and now - last_report >= self.progress_every:
                last_report = now
                self.progress(now, stats, len(inflight), len(ready), len(building), pi)
        stats.span_s = self.clock() - t0
        return stats, finished

    def close(self):
        self.pool.shutdown(wait=True)
