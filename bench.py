#!/usr/bin/env python3
"""Throughput of the per-tick agent step on one MI355X.

Workload: plain SIR on Barabasi-Albert(10M, m=5) (power-law degrees, mean 10),
int64 row_ptr / int32 col_idx, degree-sorted rows, int8 states, beta = 0.03,
gamma = 0.1, 0.1% initially infected, 200 ticks.  One step = one 200-tick
simulation from the seeded initial state (state reset + hipGraph-captured
tick loop).  Inputs are resident in HBM before the timed region.

Prints one JSON line: agent-ticks/s (whole job), the roofline of the tick
kernel (state-aware algorithmic bytes of each tick over HIP-event launch
time, beside the bytes the kernel's own counters say it touched) and the CPU
oracle timed on this host's cores on windows of the same run (each window
also checked bit-exact against the GPU).  Further legs, reported beside it:
the reference's SIR step function compiled by hipRTC, one SBM partition with
ghosts and a host halo feed (hash-map ids, history), a batched ensemble of
256 scenarios, and a serving loop with continuous batching.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def host_threads():
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit():
        n = min(n, int(env))
    return max(1, n)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def kernel_source_sha():
    """sha256 of the tick kernel's source: a traffic record is only used for the kernel it measured."""
    import hashlib
    with open(os.path.join(ROOT, "sagesim_amd", "csrc", "tick_kernels.hip"), "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()


def read_traffic(path, shape):
    """HBM bytes per tick-kernel launch measured with rocprofv3 --pmc (FETCH_SIZE, WRITE_SIZE),
    from a committed profile summary holding one record per measured launch shape: a line
    'shape: hub=.. heavy=.. apl=.. grid=..' followed by 'kernel_source_sha256: <hex>',
    'traffic_bytes_per_launch: <float>' and 'traffic_bytes_range: <lo> <hi>'.  Returns
    (bytes, (lo, hi), source) for the record of this run's shape and kernel source, or
    (None, None, reason)."""
    want = {k: str(v) for k, v in shape.items()}
    recs, cur = [], None
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("shape:"):
                    cur = {"shape": dict(kv.partition("=")[::2] for kv in line.split(":", 1)[1].split())}
                    recs.append(cur)
                elif cur is not None and line.startswith("kernel_source_sha256:"):
                    cur["sha"] = line.split(":", 1)[1].strip()
                elif cur is not None and line.startswith("traffic_bytes_per_launch:"):
                    cur["bytes"] = float(line.split(":", 1)[1])
                elif cur is not None and line.startswith("traffic_bytes_range:"):
                    cur["range"] = tuple(float(x) for x in line.split(":", 1)[1].split())
    except OSError:
        return None, None, f"no traffic profile at {os.path.relpath(path, ROOT)}"
    sha = kernel_source_sha()
    stale = False
    for r in recs:
        if {k: r["shape"].get(k) for k in want} == want and "bytes" in r:
            if r.get("sha") == sha:
                return r["bytes"], r.get("range"), "rocprofv3 FETCH_SIZE + WRITE_SIZE of this launch shape, calibrated"
            stale = True
    if stale:
        return None, None, (f"the traffic record for launch shape {want} in {os.path.relpath(path, ROOT)} "
                            "measured another version of tick_kernels.hip")
    return None, None, f"no traffic record for launch shape {want} in {os.path.relpath(path, ROOT)}"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=10_000_000)
    ap.add_argument("--m", type=int, default=5)
    ap.add_argument("--ticks", type=int, default=200)
    ap.add_argument("--hub", type=int, default=1024, help="degree from which a row gets a whole block")
    ap.add_argument("--heavy", type=int, default=256, help="degree from which a row gets a whole wavefront")
    ap.add_argument("--apl", type=int, default=16, help="agents per lane of the other rows")
    ap.add_argument("--no-autotune", action="store_true", help="use --hub/--heavy/--apl as given")
    ap.add_argument("--gather-mib", type=int, default=2048,
                    help="span of the random-gather probe (0 skips it): the known-count gather for the counters")
    ap.add_argument("--copy-mib", type=int, default=1024,
                    help="size of the plain-copy reference run (achievable HBM bandwidth; 0 skips it)")
    ap.add_argument("--grid", type=int, default=0, help="blocks of the persistent grid (0: from occupancy)")
    ap.add_argument("--push", type=float, default=0.25,
                    help="push-direction rule ratio (0: always read every neighbour)")
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="budget of the CPU-oracle sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-account", action="store_true", help="skip the work-counter pass")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "tick_kernel_ba10m_traffic.txt"))
    ap.add_argument("--ensemble-replicas", type=int, default=256, help="batched-ensemble leg (0 skips it)")
    ap.add_argument("--ensemble-ticks", type=int, default=300)
    ap.add_argument("--serve-seconds", type=float, default=20.0, help="serving leg: seconds of arrivals (0 skips)")
    ap.add_argument("--serve-rate", type=float, default=50.0, help="serving leg: arrivals per second")
    ap.add_argument("--serve-capacity", type=int, default=200,
                    help="serving leg: requests released at once to measure capacity (0 skips)")
    ap.add_argument("--partition-chunk", type=int, default=10,
                    help="partition leg: ticks per ss_simulate_halo_2bit call (1: a call per tick, as a live "
                         "exchange with peers would make)")
    ap.add_argument("--partition-ticks", type=int, default=150,
                    help="partition leg (one SBM part with ghosts and a halo feed): ticks (0 skips)")
    ap.add_argument("--partition-deg-out", default="2.0,0.17",
                    help="inter-block mean degrees of the partition leg: the stated SBM and the variant "
                         "calibrated to ~0.6M ghosts")
    ap.add_argument("--stepfunc-n", type=int, default=1_000_000,
                    help="step-function leg (the reference's SIR step compiled by hipRTC): agents (0 skips)")
    ap.add_argument("--stepfunc-ticks", type=int, default=100)
    ap.add_argument("--plain", action="store_true", help="only the plain 10M-agent leg")
    ap.add_argument("--legs", default="partition,stepfunc,ensemble,serving",
                    help="legs to run after the plain 10M-agent leg (comma-separated)")
    ap.add_argument("--skip-plain", action="store_true",
                    help="run only the --legs (a development run: the line then has no headline value)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: there is no multi-GPU path (a partition's share of a "
              "sharded run is the partition leg: ghosts through the hash map, a host halo every tick).",
              file=sys.stderr)
        sys.exit(2)

    from sagesim_amd import Engine, Replica, workloads
    legs = {x.strip() for x in args.legs.split(",") if x.strip()}
    if args.skip_plain:
        eng = Engine(0)
        out = {"development_run": "legs only, no headline value", "legs": sorted(legs)}
        _run_legs(eng, args, legs, out)
        print(json.dumps(out), flush=True)
        eng.close()
        return

    t0 = time.perf_counter()
    w = workloads.ba_sir(n=args.n, m=args.m, ticks=args.ticks)
    t_gen = time.perf_counter() - t0
    g = w.graph
    log(f"graph BA({g.n}, {args.m}): nnz {g.nnz}, max degree {int(g.degrees().max())}, {t_gen:.1f}s")

    eng = Engine(0)
    copy_gbs = None
    if args.copy_mib > 0:
        nb = args.copy_mib << 20
        ms = eng.copy_probe(nb, 5)
        copy_gbs = 2 * nb / (ms * 1e-3) / 1e9
        log(f"plain copy of {args.copy_mib} MiB: {ms:.3f} ms, {copy_gbs:.0f} GB/s read+write")
    gather = None
    if args.gather_mib > 0:
        n_g = 1 << 26
        ms = eng.gather_probe(args.gather_mib << 20, n_g, 3)
        gather = {"span_MiB": args.gather_mib, "gathers": n_g, "ms": round(ms, 4),
                  "gathers_per_s": n_g / (ms * 1e-3)}
        log(f"random gather probe: {n_g} 4-byte reads over {args.gather_mib} MiB in {ms:.3f} ms "
            f"({n_g / (ms * 1e-3) / 1e9:.1f} G gathers/s)")
    eng.set_tiling(args.heavy, args.apl, args.grid, True, hub_degree=args.hub)
    eng.set_push(args.push)
    t0 = time.perf_counter()
    rep = Replica(eng, g, w.state, w.model, w.seed)
    eng.sync()
    t_upload = time.perf_counter() - t0
    tuned = None
    if not args.no_autotune:
        from sagesim_amd.autotune import autotune
        t0 = time.perf_counter()
        best, res = autotune(eng, rep.reset, args.ticks)
        args.hub, args.heavy, args.apl = best
        tuned = {"best": list(best), "ms": res[best], "candidates": len(res),
                 "seconds": round(time.perf_counter() - t0, 2)}
        log(f"autotune: {best} {res[best]:.3f} ms per {args.ticks}-tick loop ({len(res)} shapes)")
    n_tiles, grid = eng.launch_shape()
    log(f"upload {t_upload:.2f}s, tiles {n_tiles}, grid {grid}")

    def step():
        rep.reset()
        eng.simulate(args.ticks)

    for _ in range(args.warmup):
        step()
    eng.sync()

    # timed region: K steps; HIP events on the engine's stream around each tick loop
    kern_ms = 0.0
    eng.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        rep.reset()
        eng.timer_start()
        eng.simulate(args.ticks)
        kern_ms += eng.timer_stop()
    eng.sync()
    wall = time.perf_counter() - t0
    agent_ticks = g.n * args.ticks * args.steps
    value = agent_ticks / wall
    launch_us = kern_ms * 1e3 / (args.steps * args.ticks)
    final = rep.states()
    counts = np.bincount(final, minlength=3)[:3].tolist()
    log(f"{args.steps} steps in {wall:.3f}s: {value:.4g} agent-ticks/s, "
        f"tick launch {launch_us:.1f} us, final S/I/R {counts}")

    # Roofline of the tick kernel.  Algorithmic bytes per tick follow the
    # state-aware model B_S(t) = 2N + sum over susceptible agents of (8 + 5 deg)
    # (state read + write, a row_ptr pair and 4 B index + 1 B neighbour state
    # per entry of every susceptible row), taken from the census of each tick
    # of the same run; the kernel's own counters give the bytes it touched.
    roofline = None
    if not args.no_account:
        eng.set_counting(True)
        rep.reset()
        eng.read_counters()
        b_state = touched = 0.0
        push_ticks = 0
        for _ in range(args.ticks):
            cen = eng.census(rep.rep)
            b_state += 2 * g.n + 8 * int(cen[0]) + 5 * int(cen[4])
            eng.simulate(1)
            s_agents, draws, scans, edges, pushed, pt, swept = (int(x) for x in eng.read_counters())
            push_ticks += pt
            # per agent of a tile that ran: state read and write, bitmap bit,
            # and the pushed count on a push tick; per draw: degree (1 B) and key;
            # per scan: row_ptr; per scanned entry: index + bitmap; per push:
            # index + atomic read-modify-write
            touched += (2.125 + (4 if pt else 0)) * swept + 5 * draws + 8 * scans + 5 * edges + 12 * pushed
        eng.set_counting(False)
        per_launch = b_state / args.ticks
        achieved = per_launch / (launch_us * 1e-6) / 1e9
        full_scan = 2 * g.n + 8 * (g.n + 1) + 5 * g.nnz
        shape = {"hub": args.hub, "heavy": args.heavy, "apl": args.apl, "grid": grid}
        traffic, traffic_rng, traffic_src = read_traffic(args.traffic, shape)
        roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": None if traffic is None else round(traffic, 1),
                    "traffic_source": traffic_src,
                    "traffic_range": None if traffic_rng is None else [round(x, 1) for x in traffic_rng],
                    "achieved_counter_GBs": None if traffic is None else
                    round(traffic / (launch_us * 1e-6) / 1e9, 2),
                    "frac_counter": None if traffic is None else
                    round(traffic / (launch_us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4),
                    "alg_bytes_per_launch": round(per_launch, 1),
                    "byte_model": "state-aware: 2N + sum over susceptible agents of (8 + 5 degree)",
                    "touched_bytes_per_launch": round(touched / args.ticks, 1),
                    "achieved_touched_GBs": round(touched / args.ticks / (launch_us * 1e-6) / 1e9, 2),
                    "full_scan_bytes_per_launch": full_scan,
                    "full_scan_equivalent_GBs": round(full_scan / (launch_us * 1e-6) / 1e9, 2),
                    "launch_us": round(launch_us, 2), "push_ticks": push_ticks,
                    "plain_copy_GBs": None if copy_gbs is None else round(copy_gbs, 1),
                    "gather_probe": gather,
                    "kernel": f"tick_kernel<{args.apl}, false, false>"}
        log(f"state-aware {per_launch / 1e6:.2f} MB/launch -> {achieved:.1f} GB/s; touched "
            f"{touched / args.ticks / 1e6:.2f} MB/launch; push ticks {push_ticks}")

    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        threads = host_threads()
        om = O.sir_model(w.beta, w.gamma, w.seed)
        ids = np.arange(g.n, dtype=np.int64)

        def cpu_run(st0, ticks, tick0, nthreads, fast):
            c0 = time.perf_counter()
            got, _ = O.sir_run(om, g.row_ptr, g.col_idx, ids, st0, ticks, tick0=tick0, nthreads=nthreads,
                               fast=fast)
            return got, time.perf_counter() - c0

        def gpu_states(s0, ticks):
            rep.reset()
            eng.simulate(s0)
            st0 = rep.states()
            eng.simulate(ticks)
            return st0, rep.states()

        # Two CPU codes, each checked against the GPU: or_sir_run_fast (draws first, integer
        # thresholds, skips rows that cannot be infected, stops counting early: the tuned
        # baseline) and or_sir_run (every neighbour of every S agent: the plain restatement).
        kinds = (("fast", True), ("plain", False))
        acc = {k: {"time": 0.0, "ticks": 0, "bitexact": True} for k, _ in kinds}
        # windows spread over the run, each started from the GPU's bit-exact state
        starts = [0, args.ticks // 4, args.ticks // 2, 3 * args.ticks // 4]
        win = max(1, min(10, args.ticks // 8))
        budget = args.cpu_seconds
        n_win = 0
        for s0 in starts:
            st0, want = gpu_states(s0, win)
            for k, fast in kinds:
                got, dt = cpu_run(st0, win, s0, threads, fast)
                a = acc[k]
                a["time"] += dt
                a["ticks"] += win
                a["bitexact"] &= bool(np.array_equal(got, want))
                log(f"cpu {k} window ticks [{s0}, {s0 + win}): {dt:.2f}s, match {np.array_equal(got, want)}")
            n_win += 1
            if sum(a["time"] for a in acc.values()) > budget:
                break
        # then, budget permitting, the whole run from the initial state against the GPU's final states
        full = {}
        if sum(a["time"] for a in acc.values()) < budget:
            st0, want = gpu_states(0, args.ticks)
            for k, fast in kinds:
                got, dt = cpu_run(st0, args.ticks, 0, threads, fast)
                acc[k]["time"] += dt
                acc[k]["ticks"] += args.ticks
                full[k] = bool(np.array_equal(got, want))
                log(f"cpu {k} whole run, {args.ticks} ticks: {dt:.2f}s, final states match {full[k]}")
        # one thread on one window of the epidemic's busiest stretch, checked against the GPU too
        s1 = min(args.ticks // 20, max(0, args.ticks - win))
        st0, want = gpu_states(s1, win)
        one = {}
        for k, fast in kinds:
            got, dt1 = cpu_run(st0, win, s1, 1, fast)
            one[k] = {"value": g.n * win / dt1, "cores": 1, "seconds": round(dt1, 2),
                      "sample": f"ticks [{s1}, {s1 + win}) from the GPU state",
                      "bitexact": bool(np.array_equal(got, want))}
            log(f"cpu {k} 1 thread, ticks [{s1}, {s1 + win}): {dt1:.2f}s ({g.n * win / dt1:.3g} agent-ticks/s), "
                f"match {one[k]['bitexact']}")
        sample = (f"C/OpenMP on {threads} threads ({cpu_model()}), {acc['fast']['ticks']} ticks of the same "
                  f"BA({g.n}, {args.m}) run: {win}-tick windows starting at ticks {starts[:n_win]} from the GPU state")
        if full:
            sample += f", then the whole {args.ticks}-tick run from the initial state"
        f, p = acc["fast"], acc["plain"]
        cpu = {"value": g.n * f["ticks"] / f["time"], "unit": "agent-ticks/s", "cores": threads, "kind": "port",
               "code": "oracle/sir_oracle.c or_sir_run_fast (draws first, integer thresholds, no scan of rows "
                       "that cannot be infected, counting stops at the first threshold reached)",
               "one_thread": one["fast"], "sample": sample, "seconds": round(f["time"], 2),
               "windows_bitexact": f["bitexact"], "whole_run_bitexact": full.get("fast"),
               "plain_oracle": {"value": g.n * p["ticks"] / p["time"], "cores": threads,
                                "code": "oracle/sir_oracle.c or_sir_run (every neighbour of every S agent)",
                                "one_thread": one["plain"], "seconds": round(p["time"], 2),
                                "windows_bitexact": p["bitexact"], "whole_run_bitexact": full.get("plain")}}

    out = {
        "metric": "agent-ticks/s (plain SIR, BA 10M agents, 200 ticks)",
        "value": value,
        "unit": "agent-ticks/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8",
        "data": "synthetic (seeded Barabasi-Albert graph and initial infections)",
        "config": {"workload": f"plain SIR, BA(n={g.n}, m={args.m}), beta={w.beta}, gamma={w.gamma}, "
                               f"{args.ticks} ticks, degree-sorted CSR (int64 row_ptr, int32 col_idx)",
                   "agents": g.n, "ticks": args.ticks, "nnz": g.nnz,
                   "tiling": {"hub_degree": args.hub, "heavy_degree": args.heavy, "agents_per_lane": args.apl,
                              "grid": grid, "tiles": n_tiles, "autotune": tuned},
                   "push_ratio": args.push,
                   "graph_gen_s": round(t_gen, 2), "upload_s": round(t_upload, 2),
                   "final_SIR": counts},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if cpu:
        out["speedup_vs_cpu"] = round(value / cpu["value"], 1)
    rep.remove()
    if not args.plain:
        _run_legs(eng, args, legs, out)
    print(json.dumps(out), flush=True)
    eng.close()


def _run_legs(eng, args, legs, out):
    if "partition" in legs and args.partition_ticks > 0:
        for k, d in enumerate(float(x) for x in args.partition_deg_out.split(",") if x.strip()):
            out["partition" if k == 0 else f"partition_deg_out_{d:g}"] = partition_leg(eng, args, d)
    if "stepfunc" in legs and args.stepfunc_n > 0:
        out["stepfunc"] = stepfunc_leg(args)
    if "ensemble" in legs and args.ensemble_replicas > 0:
        out["ensemble"] = ensemble_leg(eng, args)
    if "serving" in legs and args.serve_seconds > 0:
        out["serving"] = serving_leg(eng, args)


def ghost_discovery(eng, g, local_ids, n_ranks=8):
    """The partition's ghost topology found again from its global neighbour
    ids (ss_ghost_discover: ghost ids and per-rank receive lists, host arrays
    in, size query and fill both timed), checked against the ghost ids the
    generator wrote and the owner of every id, and timed beside numpy's
    unique over the same ids (the shape of the reference's own host step)."""
    col = g.col_idx
    gid = np.where(col < g.n, local_ids[np.minimum(col, g.n - 1)], g.extra[np.maximum(col - g.n, 0)])
    owner = (np.arange(g.n * n_ranks, dtype=np.int64) // g.n).astype(np.int32)
    my_rank = int(local_ids[0] // g.n)
    eng.ghost_discover(gid[:1024], owner, my_rank, n_ranks)      # first launches outside the timing
    t0 = time.perf_counter()
    ids, offsets, index = eng.ghost_discover(gid, owner, my_rank, n_ranks)
    t_gpu = time.perf_counter() - t0
    t0 = time.perf_counter()
    want = np.unique(gid[owner[gid] != my_rank])
    t_np = time.perf_counter() - t0
    ok = bool(np.array_equal(ids, want) and np.array_equal(ids, np.sort(g.extra)))
    for r in range(n_ranks):
        pos = index[offsets[r]:offsets[r + 1]]
        ok = ok and bool(np.all(owner[ids[pos]] == r) and np.all(np.diff(pos) > 0))
    ok = ok and int(offsets[-1]) == ids.shape[0]
    log(f"partition: ghost discovery over {gid.shape[0]} entries: {ids.shape[0]} ghosts in {t_gpu * 1e3:.1f} ms "
        f"(numpy unique {t_np * 1e3:.0f} ms), exact {ok}")
    return {"entries": int(gid.shape[0]), "ghosts": int(ids.shape[0]), "ms": round(t_gpu * 1e3, 2),
            "numpy_unique_ms": round(t_np * 1e3, 1), "exact": ok,
            "note": "ss_ghost_discover from host arrays (int64 ids, int32 owner table of "
                    f"{owner.shape[0]} agents), size query and fill calls, PCIe included"}


def partition_leg(eng, args, deg_out=2.0):
    """One partition's share of a 32M-agent stochastic block model (64 blocks,
    mean degree 12 inside a block and 2 across, 8 parts, part 0): local rows,
    ghost slots whose scrambled int64 ids are resolved through the GPU hash map
    (load factor 0.5), every ghost's state refreshed from a seeded host halo
    buffer before every tick, last-infector history on.  The halo crosses the
    host link as 2-bit state codes (ss_simulate_halo_2bit, packed on the host
    before the timed region).  One untimed 10-tick call warms the launch
    path, then the replica is reset; the run is timed in 10-tick calls, halo
    upload included; its first 10 ticks, and then the whole run, are checked
    against the oracle (states and history)."""
    from sagesim_amd import Replica, pack_2bit, sir_model, workloads
    t0 = time.perf_counter()
    pw = workloads.sbm_partition(ticks=args.partition_ticks, deg_out=deg_out)
    g = pw.graph
    n_ghost = g.n_cols - g.n
    t_gen = time.perf_counter() - t0
    eng.set_tiling()
    t0 = time.perf_counter()
    rep = Replica(eng, g, pw.state, sir_model(pw.beta, pw.gamma), pw.seed, rng_key_id=pw.local_ids,
                  history=pw.history)
    eng.map_build(rep.rep, pw.ghost_ids, np.arange(g.n, g.n_cols, dtype=np.int32), 0.5)
    miss = eng.halo_bind(rep.rep, pw.halo_ids)
    eng.sync()
    t_up = time.perf_counter() - t0
    log(f"partition: {g.n} local agents, {n_ghost} ghosts, {g.nnz} entries, built in {t_gen:.1f}s, "
        f"uploaded in {t_up:.1f}s")
    disc = ghost_discovery(eng, g, pw.local_ids)
    chunk = max(1, args.partition_chunk)
    t_run = 0.0
    exact = None
    n_halo = pw.halo_ids.shape[0]
    pinned = eng.pinned_empty((chunk, (n_halo + 3) // 4), np.uint8)   # halo codes need no staging
    pinned[:] = pack_2bit(pw.halo_vals(0, chunk))
    eng.simulate_halo_2bit(rep.rep, chunk, pinned)
    eng.sync()
    rep.reset()
    for t in range(0, pw.ticks, chunk):
        k = min(chunk, pw.ticks - t)
        vals = pw.halo_vals(t, t + k)
        pinned[:k] = pack_2bit(vals)
        eng.sync()
        c0 = time.perf_counter()
        eng.simulate_halo_2bit(rep.rep, k, pinned[:k])
        eng.sync()
        t_run += time.perf_counter() - c0
        if t == 0 and not args.no_cpu:
            from oracle import oracle as O
            hist = pw.history.copy()
            want, _ = O.sir_run(O.sir_model(pw.beta, pw.gamma, pw.seed), g.row_ptr, g.col_idx, pw.local_ids,
                                pw.state, k, n_local=g.n, history=hist, ghost_ids=pw.ghost_ids,
                                halo_ids=pw.halo_ids, halo_vals=vals, nthreads=host_threads())
            exact = bool(np.array_equal(rep.states()[:g.n], want[:g.n]) and np.array_equal(rep.history(), hist))
    final = np.bincount(rep.states()[:g.n], minlength=3).tolist()
    whole = None
    if not args.no_cpu:
        # the whole run on the oracle, fed the same halo (as int8 codes): final states and history
        from oracle import oracle as O
        c0 = time.perf_counter()
        hist = pw.history.copy()
        want, _ = O.sir_run(O.sir_model(pw.beta, pw.gamma, pw.seed), g.row_ptr, g.col_idx, pw.local_ids, pw.state,
                            pw.ticks, n_local=g.n, history=hist, ghost_ids=pw.ghost_ids, halo_ids=pw.halo_ids,
                            halo_vals=pw.halo_vals(0, pw.ticks), nthreads=host_threads())
        whole = bool(np.array_equal(rep.states()[:g.n], want[:g.n]) and np.array_equal(rep.history(), hist))
        log(f"partition: whole {pw.ticks}-tick run on the oracle in {time.perf_counter() - c0:.1f}s, exact {whole}")
    rep.remove()
    rate = g.n * pw.ticks / t_run
    log(f"partition: {pw.ticks} ticks in {t_run * 1e3:.1f} ms ({rate:.3g} local agent-ticks/s), "
        f"halo {(n_halo + 3) // 4 / 1e6:.2f} MB per tick (2-bit), oracle window exact {exact}")
    return {"workload": f"SIR on SBM part 0 of 8 (64 blocks x 500k agents, mean degree 12 in, {deg_out:g} out), "
                        "ghost ids through the hash map, host halo every tick as 2-bit codes, history 32",
            "local_agents": int(g.n), "ghosts": int(n_ghost), "nnz": int(g.nnz), "ticks": pw.ticks,
            "halo_ids_missing": int(miss), "ticks_per_call": chunk, "seconds": round(t_run, 4),
            "ms_per_tick": round(t_run * 1e3 / pw.ticks, 4),
            "local_agent_ticks_per_s": rate, "halo_records": int(n_halo),
            "halo_bytes_per_tick": int((n_halo + 3) // 4), "halo_format": "2-bit state codes, four records a byte",
            "oracle_window_ticks": chunk, "oracle_window_exact": exact, "oracle_whole_run_exact": whole,
            "final_SIR": final, "ghost_discovery": disc,
            "graph_build_s": round(t_gen, 2), "upload_s": round(t_up, 2)}


def stepfunc_leg(args):
    """The general path: the reference's SIR step function (HIP source with
    the reference's argument order, float32 states 1/2/3, the reference's
    Philox-2x32 uniforms) registered through Breed / Model / NetworkSpace,
    compiled by hipRTC, on BA(n, 5).  The first 10 ticks are checked against
    the oracle's restatement of the reference step."""
    from sagesim_amd import Breed, Model, NetworkSpace, graphs
    from sagesim_amd.steps import reference_sir_step
    n, ticks, seed = args.stepfunc_n, args.stepfunc_ticks, 2024
    t0 = time.perf_counter()
    g = graphs.barabasi_albert(n, 5, 17)

    class SIRBreed(Breed):
        def __init__(self):
            super().__init__("SIR")
            self.register_property("state", 1)
            self.register_step_func(reference_sir_step(), priority=0)

    m = Model(NetworkSpace())
    breed = SIRBreed()
    m.register_breed(breed)
    m.register_global_property("p_infection", 0.03)
    m.register_global_property("p_recovery", 0.1)
    m.set_seed(seed)
    rng = np.random.default_rng(5)
    init = np.where(rng.random(n) < 0.001, 2, 1)
    for i in range(n):
        m.create_agent_of_breed(breed, state=int(init[i]))
    src = np.repeat(np.arange(n, dtype=np.int64), np.diff(g.row_ptr))
    m.get_space().connect_agents_bulk(src, g.col_idx.astype(np.int64), directed=True)
    t_build = time.perf_counter() - t0
    t0 = time.perf_counter()
    m.setup()
    t_setup = time.perf_counter() - t0
    exact = None
    if not args.no_cpu:
        from oracle import oracle as O
        m.simulate(10)
        want = O.ref_sir_run(g.row_ptr, g.col_idx, np.arange(n, dtype=np.float32), init.astype(np.float32), seed,
                             10, 0.03, 0.1)
        got = np.asarray(m._device_property("state"), np.float32).reshape(-1)[m._inv]
        exact = bool(np.array_equal(got, want))
    c0 = time.perf_counter()
    m.simulate(ticks)
    dt = time.perf_counter() - c0
    rate = n * ticks / dt
    whole = None
    if not args.no_cpu:
        # all 10 + ticks ticks on the oracle from the initial states
        want = O.ref_sir_run(g.row_ptr, g.col_idx, np.arange(n, dtype=np.float32), init.astype(np.float32), seed,
                             10 + ticks, 0.03, 0.1)
        got = np.asarray(m._device_property("state"), np.float32).reshape(-1)[m._inv]
        whole = bool(np.array_equal(got, want))
    log(f"stepfunc: BA({n}, 5), {ticks} ticks of the hipRTC-compiled reference step in {dt * 1e3:.1f} ms "
        f"({rate:.3g} agent-ticks/s), mode {m.mode}, oracle window exact {exact}, whole run exact {whole}")
    return {"workload": f"reference SIR step function (hipRTC), BA({n}, 5), float32 states, Philox-2x32",
            "agents": n, "nnz": int(g.nnz), "ticks": ticks, "seconds": round(dt, 4), "agent_ticks_per_s": rate,
            "mode": m.mode, "oracle_window_ticks": 10, "oracle_window_exact": exact,
            "oracle_whole_run_ticks": 10 + ticks, "oracle_whole_run_exact": whole,
            "build_s": round(t_build, 2), "setup_s": round(t_setup, 2)}


def ensemble_leg(eng, args):
    """Batched ensemble: replicas of log-uniform size packed into one launch per
    tick and one graph replay, against the per-scenario launch loop."""
    from sagesim_amd.autotune import autotune
    from sagesim_amd.ensemble import Ensemble, ensemble_scenarios, prepare, run_batched, run_standalone
    scs = ensemble_scenarios(args.ensemble_replicas, ticks=args.ensemble_ticks)
    t0 = time.perf_counter()
    prepared = prepare(scs, threads=host_threads())
    t_gen = time.perf_counter() - t0
    total = sum(p.graph.n for p in prepared)
    t0 = time.perf_counter()
    ens = Ensemble.build(eng, prepared)
    eng.sync()
    t_up = time.perf_counter() - t0
    log(f"ensemble: {len(prepared)} replicas, {total} agents, built in {t_gen:.1f}s, uploaded in {t_up:.1f}s")
    cands = [(h, hv, a) for h in (1024, 2048, 4096) for hv in (64, 128, 256) for a in (4, 8, 16)]
    best, res = autotune(eng, ens.reset, args.ensemble_ticks, candidates=cands)
    t_b = run_batched(eng, ens)
    batched = [ens.states(i) for i in range(len(prepared))]
    ens.remove()
    alone, t_a = run_standalone(eng, prepared)
    exact = all(np.array_equal(a, b) for a, b in zip(alone, batched))
    n = len(prepared)
    log(f"ensemble: batched {t_b * 1e3:.1f} ms, per-scenario loop {t_a * 1e3:.1f} ms, "
        f"speedup {t_a / t_b:.2f}, bit-exact {exact}")
    return {"replicas": n, "agents_total": int(total), "ticks": args.ensemble_ticks,
            "scenarios_per_s": n / t_b, "agent_ticks_per_s": total * args.ensemble_ticks / t_b,
            "batched_s": t_b, "per_scenario_loop_s": t_a, "per_scenario_scenarios_per_s": n / t_a,
            "speedup_vs_per_scenario_loop": t_a / t_b, "bitexact_vs_standalone": exact,
            "tiling": list(best), "graph_build_s": round(t_gen, 2), "upload_s": round(t_up, 2)}


def serving_leg(eng, args):
    """Poisson arrivals, host graph builders, admission against a 200 GB HBM
    budget, continuous batching into the shared tick launch; a sample of the
    results is re-run alone and compared."""
    from sagesim_amd import Replica
    from sagesim_amd.ensemble import Prepared, run_standalone
    from sagesim_amd.serving import Server, serving_requests
    eng.set_tiling()
    arrivals = serving_requests(rate=args.serve_rate, duration=args.serve_seconds)

    def make(req):
        sc = req.scenario
        return Replica(eng, model=sc.model(), seed=sc.seed, tick_limit=sc.ticks, data=req.data)

    def progress(now, st, inflight, ready, building, arrived):
        log(f"serving: t={now:.0f}s arrived {arrived}, done {st.completed}, in flight {inflight}, "
            f"queued {ready}, building {building}")

    # the requests re-run alone afterwards are chosen up front; only their
    # states are kept (a 10-minute window finishes ~30k requests)
    rng = np.random.default_rng(1)
    keep = set(rng.choice(len(arrivals), size=min(4, len(arrivals)), replace=False).tolist()) if arrivals else set()
    srv = Server(eng, make, budget_bytes=200 << 30, chunk_ticks=10, builders=max(1, host_threads() - 1),
                 progress=progress, progress_every_s=30.0, keep_results=keep)
    stats, done = srv.run(arrivals)
    srv.close()
    summ = stats.summary()
    sample = [r for r in done if r.rid in keep and r.outcome == "done"]
    prepared = []
    for r in sample:
        sc = r.scenario
        g = sc.build_graph()
        prepared.append(Prepared(sc, g, sc.initial_state()))
    alone, _ = run_standalone(eng, prepared)
    summ["sample_bitexact_alone_vs_mix"] = all(np.array_equal(a, r.result) for a, r in zip(alone, sample))
    summ["sample_size"] = len(sample)
    summ["agent_ticks"] = int(sum(r.scenario.n * r.scenario.ticks for r in done if r.outcome == "done"))
    summ["rate_per_s"] = args.serve_rate
    summ["arrival_window_s"] = args.serve_seconds
    if args.serve_capacity > 0:
        # capacity: every request of a sample there at once (saturating arrivals)
        burst = [(0.0, sc) for _, sc in serving_requests(rate=50.0, duration=args.serve_capacity / 50.0, seed=4242)]
        srv = Server(eng, make, budget_bytes=200 << 30, chunk_ticks=10, builders=max(1, host_threads() - 1),
                     keep_results=False)
        cst, cdone = srv.run(burst, realtime=False)
        srv.close()
        cs = cst.summary()
        summ["capacity"] = {"requests": len(burst), "completed": cs["completed"],
                            "scenarios_per_s": round(cs["scenarios_per_s"], 3), "span_s": round(cs["span_s"], 3),
                            "agent_ticks": int(sum(r.scenario.n * r.scenario.ticks for r in cdone
                                                   if r.outcome == "done"))}
        log(f"serving capacity: {summ['capacity']}")
    log(f"serving: {summ}")
    for k in ("p50_latency_s", "p99_latency_s", "max_latency_s", "scenarios_per_s", "span_s"):
        summ[k] = round(summ[k], 4)
    return summ


if __name__ == "__main__":
    main()
