"""Launch-shape autotuner for the tick kernel.

The tick kernel's cost depends on the degree distribution and on how the
epidemic evolves, so the launch shape (which rows get a block, which a
wavefront, how many agents each lane carries) is picked by timing candidate
shapes on the workload itself: every candidate runs the same tick loop from
the same reset state, timed with HIP events on the engine's stream.  All
shapes produce identical states (the kernels differ only in how work is
mapped), so tuning never changes results.
"""
import itertools

DEFAULT_CANDIDATES = [
    (hub, heavy, apl)
    for hub, heavy, apl in itertools.product((512, 1024, 2048, 4096, 1 << 30), (32, 64, 128, 256, 512), (2, 4, 8, 16))
    if hub >= heavy
]


def time_shape(engine, reset, ticks, shape, repeats=1):
    hub, heavy, apl = shape
    engine.set_tiling(heavy, apl, 0, True, hub_degree=hub)
    reset()
    engine.simulate(ticks)                 # capture + warm
    best = float("inf")
    for _ in range(repeats):
        reset()
        engine.timer_start()
        engine.simulate(ticks)
        best = min(best, engine.timer_stop())
    return best


def autotune(engine, reset, ticks, candidates=None, repeats=1, log=None):
    """Time every candidate (hub_degree, heavy_degree, agents_per_lane) and leave
    the fastest one set on `engine`.  `reset()` must put every replica of the
    engine back to its starting state.  Returns (best_shape, {shape: ms})."""
    results = {}
    for shape in candidates or DEFAULT_CANDIDATES:
        results[shape] = time_shape(engine, reset, ticks, shape, repeats)
        if log:
            log(f"autotune {shape}: {results[shape]:.3f} ms")
    best = min(results, key=results.get)
    hub, heavy, apl = best
    engine.set_tiling(heavy, apl, 0, True, hub_degree=hub)
    reset()
    return best, results
