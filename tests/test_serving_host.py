"""Host-side logic of the serving loop (admission against the budget,
rejection, retirement, latency bookkeeping) with an engine stand-in, and the
request generator: no GPU needed."""
import numpy as np

from sagesim_amd.ensemble import Scenario, ensemble_scenarios
from sagesim_amd.serving import Server, footprint, serving_requests


class FakeReplica:
    def __init__(self, eng, req):
        self.eng, self.req, self.t = eng, req, 0
        eng.live.append(self)

    def tick(self):
        return self.t

    def states(self):
        return np.full(self.req.scenario.n, self.req.rid % 3, np.int8)

    def remove(self):
        self.eng.live.remove(self)


class FakeEngine:
    def __init__(self):
        self.live = []
        self.calls = 0

    def simulate(self, k):
        self.calls += 1
        for r in self.live:
            r.t = min(r.t + k, r.req.scenario.ticks)

    def sync(self):
        pass


def _small_requests(n=25, seed=3):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        sc = Scenario(n=int(rng.integers(500, 3000)), graph_kind="ba" if i % 2 else "ws",
                      graph_param=3 if i % 2 else 10, beta=0.05, gamma=0.1, sigma=0.2 if i % 3 == 0 else 0.0,
                      seed=int(rng.integers(1, 2**40)), ticks=int(rng.integers(20, 60)))
        out.append((0.0005 * i, sc))
    return out


def test_admission_respects_budget_and_everything_completes():
    arrivals = _small_requests()
    fps = [footprint(sc.build_graph()) for _, sc in arrivals]
    budget = int(max(fps) * 2.5)          # at most a few in flight at once
    eng = FakeEngine()
    srv = Server(eng, lambda req: FakeReplica(eng, req), budget_bytes=budget, chunk_ticks=7, builders=2)
    stats, done = srv.run(arrivals)
    srv.close()
    assert stats.completed == len(arrivals) and stats.rejected == 0
    assert stats.max_bytes <= budget
    assert 1 < stats.max_inflight < len(arrivals)
    for r in done:
        assert r.done >= r.admitted >= r.arrival - 1e-9
        assert (r.result == r.rid % 3).all()
    s = stats.summary()
    assert s["p99_latency_s"] >= s["p50_latency_s"] >= 0.0


def test_oversized_request_is_rejected_not_blocking():
    arrivals = _small_requests(6)
    big = Scenario(n=200_000, graph_kind="ba", graph_param=5, beta=0.05, gamma=0.1, sigma=0.0, seed=9, ticks=30)
    arrivals.insert(2, (0.0011, big))
    budget = footprint(arrivals[0][1].build_graph()) * 4
    eng = FakeEngine()
    srv = Server(eng, lambda req: FakeReplica(eng, req), budget_bytes=budget, chunk_ticks=10, builders=2)
    stats, done = srv.run(arrivals)
    srv.close()
    assert stats.rejected == 1 and stats.completed == len(arrivals) - 1
    assert [r.rid for r in done if r.result is None] == [2]


def test_request_generator_is_seeded_and_in_range():
    a = serving_requests(rate=50, duration=4.0, seed=5)
    b = serving_requests(rate=50, duration=4.0, seed=5)
    assert [(t, sc.n, sc.seed) for t, sc in a] == [(t, sc.n, sc.seed) for t, sc in b]
    assert 120 < len(a) < 300
    for t, sc in a:
        assert 0 <= t < 4.0 and 50_000 <= sc.n <= 5_000_000 and 100 <= sc.ticks <= 500
        assert (sc.graph_kind == "ba" and 3 <= sc.graph_param <= 6) or (sc.graph_kind == "ws" and sc.graph_param == 10)
        assert sc.sigma == 0.0 or 0.1 <= sc.sigma <= 0.3


def test_ensemble_scenarios_distribution():
    sc = ensemble_scenarios(256)
    sizes = np.array([s.n for s in sc])
    assert sizes.min() >= 10_000 and sizes.max() <= 2_000_000
    assert {s.graph_param for s in sc} <= set(range(2, 9))
    assert all(0.01 <= s.beta <= 0.08 and 0.05 <= s.gamma <= 0.2 for s in sc)
    assert len({s.seed for s in sc}) == 256


class SimClock:
    """A simulated clock: every reading moves it on by 1 ms (so an idle loop
    reaches the next arrival), and the fake engine adds its run time."""

    def __init__(self):
        self.t = 0.0

    def __call__(self):
        self.t += 1e-3
        return self.t


class TimedEngine(FakeEngine):
    def __init__(self, clock, step):
        super().__init__()
        self.clock, self.step = clock, step

    def simulate(self, k):
        super().simulate(k)
        self.clock.t += self.step


def test_cancel_in_every_stage_and_others_unchanged():
    """Requests cancelled before arrival, while queued and while on the device
    leave with no result, free their bytes, and every other request finishes
    with the result it gets alone."""
    arrivals = _small_requests(12)
    arrivals = [(0.0, sc) for _, sc in arrivals[:4]] + [(5.0 + 0.001 * i, sc) for i, (_, sc) in enumerate(arrivals[4:])]
    clock = SimClock()
    eng = TimedEngine(clock, 0.05)
    fps = [footprint(sc.build_graph()) for _, sc in arrivals]
    budget = int(max(fps) * 2.2)                     # about two in flight: others wait in the queue
    srv = Server(eng, lambda req: FakeReplica(eng, req), budget_bytes=budget, chunk_ticks=5, builders=2,
                 clock=clock)
    srv.cancel(9)                                    # before it arrives
    orig = eng.simulate

    def simulate_and_cancel(k):
        orig(k)
        if eng.calls == 2:
            for r in eng.live[:1]:                   # the first one on the device
                srv.cancel(r.req.rid)
            srv.cancel(3)                            # waiting for memory or in flight
    eng.simulate = simulate_and_cancel
    stats, done = srv.run(arrivals)
    srv.close()
    by = {r.rid: r for r in done}
    assert len(by) == len(arrivals)
    cancelled = {rid for rid, r in by.items() if r.outcome == "cancelled"}
    assert 9 in cancelled and 3 in cancelled and len(cancelled) == stats.cancelled >= 3
    assert stats.completed == len(arrivals) - stats.cancelled and stats.expired == 0
    for rid, r in by.items():
        if r.outcome == "cancelled":
            assert r.result is None and r.replica is None
        else:
            assert r.outcome == "done" and (r.result == rid % 3).all()
    assert eng.live == []                            # every replica removed
    assert stats.max_bytes <= budget


def test_deadline_expires_slow_requests():
    """With a deadline shorter than a long request's run, that request expires
    on the device; short requests still finish, and a per-request deadline in
    the arrival tuple overrides the server's."""
    base = _small_requests(6)
    long_sc = Scenario(n=1000, graph_kind="ws", graph_param=10, beta=0.05, gamma=0.1, sigma=0.0, seed=4, ticks=400)
    arrivals = [(0.0, sc) for _, sc in base] + [(0.0, long_sc), (0.0, long_sc, 1e9)]
    clock = SimClock()
    eng = TimedEngine(clock, 0.05)                   # 5 ticks per 0.05 s: 400 ticks take 4 s
    srv = Server(eng, lambda req: FakeReplica(eng, req), budget_bytes=1 << 40, chunk_ticks=5, builders=2,
                 clock=clock, deadline_s=2.5)
    stats, done = srv.run(arrivals)
    srv.close()
    by = {r.rid: r for r in done}
    assert by[6].outcome == "expired" and by[6].result is None
    assert by[7].outcome == "done"                   # its own deadline is far off
    assert all(by[i].outcome == "done" for i in range(6))
    assert stats.expired == 1 and stats.completed == 7 and eng.live == []


def test_results_kept_only_for_chosen_requests():
    arrivals = _small_requests(8)
    eng = FakeEngine()
    srv = Server(eng, lambda req: FakeReplica(eng, req), budget_bytes=1 << 40, chunk_ticks=10, builders=2,
                 keep_results={1, 5})
    stats, done = srv.run(arrivals)
    srv.close()
    assert stats.completed == len(arrivals)
    assert sorted(r.rid for r in done if r.result is not None) == [1, 5]
    assert all((r.result == r.rid % 3).all() for r in done if r.result is not None)
    assert all(r.outcome == "done" for r in done)
