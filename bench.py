#!/usr/bin/env python3
"""Trajectory-steps/s of the backward Parker modulation workload on one MI355X.

Workload: 1e7 trajectories (64 log-spaced energies 0.1-10 GeV x 156250 each) of backward
Parker transport in the Parker-spiral solar wind with rigidity power-law diffusion, started at
1 au and absorbed at 90 au, counter-based (Philox) Wiener increments.  All 1e7 trajectories are
queued on the device by `begin`; the persistent engine keeps every lane busy, refilling finished
slots from the device work queue.  A bench "step" is one engine launch: the service kernel
(refill, start sampling, event binning) followed by the step kernel (steps_per_launch SDE steps
on every resident lane).  W launches are warm-up, K launches are timed between device
synchronisations; `value` is trajectory-steps executed in the timed launches per second.

Both kernels run: the plain FP64 kernel (the drop-in) and the FP32-step / FP64-accumulate
kernel; `value` is the mixed kernel, the FP64 figures sit beside it.  The roofline entry is the
step kernel of the reported precision: algorithmic flops per trajectory step (counted
mechanically, profiles/flops_per_step.json) x steps in a launch / step-kernel HIP-event time.
The CPU baseline is the FP64 restatement of the reference algorithm on the host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

WORKLOAD = "parker_modulation_64x156250"
N_ENERGIES, PER_ENERGY = 64, 156250
PEAK_TFLOPS = {"fp64": 78.6, "mixed": 157.3}      # MI355X vector FP64 / FP32 peaks
PEAK_HBM_GBPS = 8000.0                              # MI355X HBM3E
PREC_ID = {"fp64": 0, "mixed": 1}


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def load_json(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def gpu_leg(E, cf, prec, args, setup=None, n_traj=N_ENERGIES * PER_ENERGY, autotune=False, lanes=0, steps_per_launch=0):
    s = setup if setup is not None else cf.parker_modulation(n_energies=N_ENERGIES)
    spl = steps_per_launch or args.steps_per_launch
    opts = dict(precision=PREC_ID[prec], steps_per_launch=spl)
    if lanes or args.lanes:
        opts["lanes"] = lanes or args.lanes
    sim = E.Simulation.from_setup(s, **opts)
    tuned = None
    if autotune:
        table = sim.autotune(n_probe=n_traj, seed=2)
        plain = [e for e in table if e[3] == 0]
        best = max(plain, key=lambda e: e[2])
        graph = next((e for e in table if e[3] > 0), None)
        use_graph = bool(graph and graph[2] > 1.02 * best[2])      # the autotuner's own rule
        tuned = {"lanes": best[0], "steps_per_launch": best[1], "candidates": len(plain),
                 "graph_vs_plain": round(graph[2] / best[2], 4) if graph else None, "graph": use_graph}
        log(f"{s.name} {prec}: autotuned lanes {best[0]} steps/launch {best[1]} ({best[2]:.3e} steps/s in the probe); "
            f"as a HIP graph {tuned['graph_vs_plain']}x -> graph {'on' if use_graph else 'off'}")
        spl = best[1]
    # a shorter launch gets proportionally more timed launches, so every leg times about as many steps per slot
    launches = max(args.steps, (args.steps * args.steps_per_launch) // spl)
    sim.begin(n_traj, seed=1)
    w = sim.advance(args.warmup)
    t0 = time.perf_counter()
    st = sim.advance(launches)
    wall = time.perf_counter() - t0
    steps = st.steps_executed - w.steps_executed
    log(f"{s.name} {prec}: lanes {st.lanes} steps/launch {spl} launches {launches} trajectory-steps {steps} wall {wall:.3f}s "
        f"step-kernel {st.step_kernel_ms:.1f} ms service {st.service_kernel_ms:.1f} ms -> {steps / wall:.4e} steps/s")
    # under graph replay the kernels are not timed one by one: the launch time (service + step) stands in
    step_ms = st.step_kernel_ms if st.step_kernel_ms > 0 else st.kernel_ms
    return dict(steps=steps, wall=wall, lanes=st.lanes, step_ms=step_ms, service_ms=st.service_kernel_ms,
                done=st.done, failed=st.failed, tuned=tuned, launches=launches, steps_per_launch=spl)


def roofline(leg, prec, flops, traffic):
    fps = flops["flops_per_step"]
    launches = leg["launches"]
    per_launch_s = leg["step_ms"] * 1e-3 / launches
    achieved = fps * leg["steps"] / launches / per_launch_s / 1e12
    peak = PEAK_TFLOPS[prec]
    t = None
    if traffic and prec in traffic:
        t = traffic[prec].get("bytes_per_launch")
    return {"bound": "valu", "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s", "frac": round(achieved / peak, 4),
            "traffic": t, "kernel": f"parker_step_kernel ({prec})", "avg_launch_ms": round(per_launch_s * 1e3, 3),
            "flops_per_step": round(fps, 1), "transcendentals_per_step": round(flops["transcendentals_per_step"], 1),
            "flops_basis": "the reference algorithm's flops per step (counted on the oracle), not the kernel's executed "
                           "ones: the FP32 kernel's closed-form stencils execute fewer, so its frac can exceed 1"}


def other_workloads(E, cf, args, flops_all):
    """Steady-state throughput of the other two transports at their full trajectory counts: diffusive
    shock acceleration (1e8 trajectories queued) and the guiding centre (1e7), both precisions."""
    out = {}
    for name, setup, n, key, tune in (("dsa_smooth_shock_1e8", cf.dsa_shock(), 100_000_000, "dsa_smooth_shock", True),
                                      ("gc_diff_scatt_1e7", cf.gc_diff_scatt(), 10_000_000, "gc_diff_scatt", False)):
        res = {}
        for p in ("fp64", "mixed"):
            leg = gpu_leg(E, cf, p, args, setup, n, autotune=tune and not args.no_autotune,
                          lanes=args.dsa_lanes if (name.startswith("dsa") and args.no_autotune) else 0,
                          steps_per_launch=args.dsa_steps_per_launch if (name.startswith("dsa") and args.no_autotune) else 0)
            rf = roofline(leg, p, flops_all[key], None)
            res[p] = {"value": leg["steps"] / leg["wall"], "avg_step_kernel_ms": rf["avg_launch_ms"], "roofline_frac": rf["frac"],
                      "achieved_tflops": rf["achieved"], "lanes": leg["lanes"]}
            if leg["tuned"]:
                res[p]["autotuned"] = leg["tuned"]
        res["speedup_mixed_vs_fp64"] = round(res["mixed"]["value"] / res["fp64"]["value"], 3)
        res["flops_per_step"] = round(flops_all[key]["flops_per_step"], 1)
        out[name] = res
    return out


def grid_lanes(E, args):
    """Resident slots of the gridded legs: --grid-waves waves per SIMD on every CU (0 = engine default)."""
    if not args.grid_waves:
        return 0
    cus = int(E.device_info().split("cu=")[1].split("|")[0])
    return cus * 4 * 64 * args.grid_waves


def grid_workloads(E, cf, args, traffic_all):
    """Parker modulation on a gridded solar wind (BackgroundServerCartesian), the reference's
    Cartesian benchmark: its 25^3-block (100^3-zone, 64 MB of FP64 zone records) grid over +-100 au,
    and a 64^3-block (256^3-zone, 1.07 GB of FP64 zone records, 4.3 GB of FP32 cell records) grid that no
    cache holds.  Each step interpolates the grid
    at the new position (kappa is isotropic here, so the six derivative points of div(kappa) need only
    their positions checked): one gather of 8 zones.  The algorithmic gather is 8 zones x 6 values x
    8 B (FP64) or 4 B (FP32) per interpolation; the HBM bytes per step come from the FETCH_SIZE /
    WRITE_SIZE profile of the same legs when present, and the rate from this run's kernel time."""
    from spectrum_amd import cartesian as G
    out = {}
    for nb in args.grid_blocks:
        grid = G.parker_spiral(nb=nb)
        setup = cf.parker_modulation_cartesian(grid)
        name = f"parker_modulation_cartesian_{nb}^3_blocks"
        res = {"zones": 4 * nb, "device_grid_bytes": {"fp64": 64 * (4 * nb) ** 3, "mixed": 8 * 32 * (4 * nb) ** 3}}
        for p in (("fp64", "mixed") if args.precision == "both" else (args.precision,)):
            leg = gpu_leg(E, cf, p, args, setup, 10_000_000, autotune=False,
                          lanes=grid_lanes(E, args), steps_per_launch=args.grid_steps_per_launch)
            per_launch_s = leg["step_ms"] * 1e-3 / leg["launches"]
            steps_per_launch = leg["steps"] / leg["launches"]
            gather = setup.meta["grid_gathers_per_step"] * 8 * 6 * (8 if p == "fp64" else 4)
            r = {"value": leg["steps"] / leg["wall"], "avg_step_kernel_ms": round(per_launch_s * 1e3, 3), "lanes": leg["lanes"],
                 "steps_per_launch": leg["steps_per_launch"], "timed_launches": leg["launches"], "autotuned": leg["tuned"],
                 "gather_bytes_per_step": gather,
                 "gather_GBps": round(gather * steps_per_launch / per_launch_s / 1e9, 1)}
            t = (traffic_all or {}).get(name, {}).get(p)
            if t:
                r["hbm_bytes_per_step"] = round(t["bytes_per_step"], 1)
                r["hbm_GBps"] = round(t["bytes_per_step"] * steps_per_launch / per_launch_s / 1e9, 1)
                r["hbm_frac"] = round(r["hbm_GBps"] / PEAK_HBM_GBPS, 4)
            res[p] = r
        if "fp64" in res and "mixed" in res:
            res["speedup_mixed_vs_fp64"] = round(res["mixed"]["value"] / res["fp64"]["value"], 3)
        out[name] = res
        del grid, setup
    return out


def host_cpus():
    """CPU share of this process (affinity mask, capped by OMP_NUM_THREADS when the environment sets it), the
    machine's logical CPU count (nproc) and the processor model (lscpu)."""
    share = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        share = min(share, int(omp))
    model = ""
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        model = next((l.split(":", 1)[1].strip() for l in out.splitlines() if l.startswith("Model name")), "")
    except Exception:
        pass
    return share, os.cpu_count() or 0, model


def cpu_leg(cf, seconds, threads):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    s = cf.parker_modulation(n_energies=N_ENERGIES)
    o = oracle.Oracle(s)
    batch = threads * 4
    first, steps, n = 0, 0, 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        k, _ = o.run(batch, seed=1, first=first, threads=threads, records=False)
        steps += k
        n += batch
        first += batch
    wall = time.perf_counter() - t0
    share, nproc, model = host_cpus()
    log(f"cpu: {n} trajectories, {steps} steps in {wall:.2f}s on {threads} threads -> {steps / wall:.4e} steps/s")
    return {"value": steps / wall, "unit": "trajectory-steps/s", "cores": threads, "kind": "port",
            "sample": f"{n} whole trajectories (ids 0..{n - 1} of the same workload, seed 1), FP64 oracle, "
                      f"OpenMP over trajectories, {wall:.1f} s",
            "host": {"cpu_share": share, "nproc": nproc, "model": model}}


def pct(a, q):
    import numpy as np
    return round(float(np.percentile(a, q)), 4) if len(a) else None


def serve_leg(E, cf, args):
    """The served mixed stream: Poisson arrivals (mean rate --serve-rate) of modulation-spectrum queries
    (start radius U[1, 80] au, 8-64 energies, log-uniform 1e3-1e5 trajectories per energy times
    --serve-scale) with 20% diffusive-shock-acceleration queries, submitted at their arrival times to a query
    server running the mixed kernel.  Each query's latency is split into queue wait (submission to its first
    trajectory on the device) and service (first trajectory to the last).  The offered load is the arrival
    rate times the measured work per query over the engine's measured rate: above 1 the queue only grows, and
    latency is set by the query's place in it."""
    import numpy as np
    from spectrum_amd.server import Server
    rng = np.random.default_rng(args.serve_seed)
    plan = []
    t = 0.0
    for i in range(args.serve_queries):
        t += rng.exponential(1.0 / args.serve_rate)
        if rng.random() < 0.2:
            n = max(64, int(10.0 ** rng.uniform(3.0, 5.0) * args.serve_scale * 8))
            setup = cf.dsa_shock(name=f"dsa_{i}")
            kind = "dsa"
        else:
            ne = int(rng.integers(8, 65))
            per = 10.0 ** rng.uniform(3.0, 5.0) * args.serve_scale
            n = max(ne, int(ne * per))
            setup = cf.parker_modulation(n_energies=ne, r_start_au=float(rng.uniform(1.0, 80.0)), name=f"mod_{i}")
            kind = "modulation"
        plan.append((t, E.Simulation.from_setup(setup), n, i + 1, kind))
    with Server(precision=E.PREC_MIXED, steps_per_launch=args.serve_steps_per_launch) as srv:
        t0 = time.perf_counter()
        ids = []
        for ta, proto, n, seed, kind in plan:
            delay = t0 + ta - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            ids.append(srv.submit(proto, n, seed=seed))
        for k, q in enumerate(ids):
            srv.wait(q)
            log(f"serve: query {k + 1}/{len(ids)} done")
        makespan = time.perf_counter() - t0
        info = [srv.info(q) for q in ids]
        st = srv.stats()
    lat = np.array([i["latency_s"] for i in info])
    wait = np.array([i["t_start"] - i["t_submit"] for i in info])
    service = np.array([i["t_done"] - i["t_start"] for i in info])
    steps = np.array([i["steps_executed"] for i in info], dtype=float)
    n_traj = sum(p[2] for p in plan)
    engine_rate = steps.sum() / max(st["busy_s"], 1e-9)
    capacity_qps = engine_rate / steps.mean()
    offered = args.serve_rate / capacity_qps
    log(f"serve: {len(ids)} queries ({sum(p[4] == 'dsa' for p in plan)} DSA), {n_traj} trajectories, {int(steps.sum())} steps in {makespan:.2f}s; "
        f"latency p50 {np.percentile(lat, 50):.3f}s p99 {np.percentile(lat, 99):.3f}s (wait p50 {np.percentile(wait, 50):.3f}s, "
        f"service p50 {np.percentile(service, 50):.3f}s); engine busy {st['busy_s']:.2f}s; capacity {capacity_qps:.3f} queries/s, "
        f"offered load {offered:.1f}x capacity")
    return {"queries": len(ids), "dsa_queries": int(sum(p[4] == "dsa" for p in plan)), "arrival_rate_hz": args.serve_rate,
            "trajectories": int(n_traj), "trajectories_per_energy": [round(1e3 * args.serve_scale, 1), round(1e5 * args.serve_scale, 1)],
            "p50_latency_s": pct(lat, 50), "p99_latency_s": pct(lat, 99), "max_latency_s": round(float(lat.max()), 4),
            "p99_is_the_max": len(lat) < 100, "mean_latency_s": round(float(lat.mean()), 4),
            "queue_wait_s": {"p50": pct(wait, 50), "p99": pct(wait, 99)}, "service_s": {"p50": pct(service, 50), "p99": pct(service, 99)},
            "mean_steps_per_query": float(steps.mean()), "engine_steps_per_s_busy": engine_rate, "capacity_queries_per_s": capacity_qps,
            "offered_load_vs_capacity": round(offered, 2), "makespan_s": round(makespan, 3), "engine_busy_s": round(st["busy_s"], 3),
            "trajectory_steps_per_s": float(steps.sum()) / makespan, "precision": "mixed", "steps_per_launch": args.serve_steps_per_launch}


def drain_leg(E, cf, args):
    """Time to solution: --drain-traj trajectories of the headline workload queued and run to completion with
    the mixed kernel, one engine launch at a time.  The lane occupancy of each launch (steps executed over
    slots x steps per launch) shows the steady state while the queue feeds the slots and the drain once it is
    empty, when the longest trajectories run with most slots idle."""
    import numpy as np
    spl = args.steps_per_launch
    sim = E.Simulation.from_setup(cf.parker_modulation(n_energies=N_ENERGIES), precision=PREC_ID["mixed"], steps_per_launch=spl)
    n = args.drain_traj
    sim.begin(n, seed=3)
    t0 = time.perf_counter()
    trace, last = [], 0
    done = 0
    while True:
        st = sim.advance(1)
        now = time.perf_counter() - t0
        trace.append((now, st.done, st.steps_executed - last, st.lanes))
        last = st.steps_executed
        done = st.done
        if done >= n or now > args.drain_seconds:
            break
    wall = time.perf_counter() - t0
    occ = np.array([k / (l * spl) for _, _, k, l in trace])
    tt = np.array([x[0] for x in trace])
    dn = np.array([x[1] for x in trace])
    queue_empty = next((i for i, x in enumerate(trace) if x[1] >= n - st.lanes), len(trace) - 1)
    t_at = lambda f: float(tt[np.argmax(dn >= f * n)]) if (dn >= f * n).any() else None
    res = {"trajectories": n, "completed": int(done), "wall_s": round(wall, 3), "launches": len(trace),
           "trajectory_steps_per_s": float(last) / wall, "steps_per_launch": spl, "lanes": int(trace[-1][3]),
           "time_to_50pct_s": t_at(0.5), "time_to_99pct_s": t_at(0.99), "time_to_all_s": round(wall, 3) if done >= n else None,
           "occupancy_mean_until_queue_empty": round(float(occ[:queue_empty + 1].mean()), 4),
           "occupancy_mean_drain": round(float(occ[queue_empty + 1:].mean()), 4) if queue_empty + 1 < len(occ) else None,
           "drain_s": round(float(tt[-1] - tt[queue_empty]), 3),
           "occupancy_trace": [[round(float(a), 3), round(float(b), 4)] for a, b in zip(tt[:: max(1, len(tt) // 40)], occ[:: max(1, len(tt) // 40)])]}
    log(f"drain: {done}/{n} trajectories in {wall:.2f}s, {res['trajectory_steps_per_s']:.4e} steps/s overall; occupancy "
        f"{res['occupancy_mean_until_queue_empty']} while the queue feeds the slots, {res['occupancy_mean_drain']} over the last "
        f"{res['drain_s']} s")
    return res


def copy_probe(mib):
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    log(f"copy probe: 3 x {mib} MiB device-to-device copies done")


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed engine launches")
    ap.add_argument("--warmup", type=int, default=3, help="untimed engine launches")
    ap.add_argument("--steps-per-launch", type=int, default=2000)
    ap.add_argument("--lanes", type=int, default=0, help="resident trajectory slots (0 = engine default)")
    ap.add_argument("--precision", choices=["both", "fp64", "mixed"], default="both")
    ap.add_argument("--cpu-seconds", type=float, default=30.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-other-workloads", action="store_true", help="skip the DSA and guiding-centre throughput legs")
    ap.add_argument("--no-grid-workloads", action="store_true", help="skip the gridded-background (Cartesian) legs")
    ap.add_argument("--grid-blocks", type=lambda v: [int(x) for x in v.split(",")], default=[25, 64],
                    help="blocks per axis of the gridded legs (comma list; 25 is the reference's Cartesian benchmark)")
    ap.add_argument("--grid-waves", type=int, default=8,
                    help="waves per SIMD of the gridded legs' slots (the autotuner's choice for them; 0 = engine default)")
    ap.add_argument("--grid-steps-per-launch", type=int, default=250,
                    help="steps per slot per launch of the gridded legs (the autotuner's choice for them)")
    ap.add_argument("--grid-only", action="store_true",
                    help="run only the gridded-background legs and print their results (for profiling their kernels)")
    ap.add_argument("--other-only", action="store_true",
                    help="run only the DSA and guiding-centre legs and print their results (for profiling their kernels)")
    ap.add_argument("--dsa-lanes", type=int, default=524288, help="DSA leg's slots without the autotuner (its usual choice)")
    ap.add_argument("--dsa-steps-per-launch", type=int, default=4000, help="DSA leg's steps per launch without the autotuner")
    ap.add_argument("--no-autotune", action="store_true",
                    help="skip the autotuner on the modulation legs (for profiling: the trace then holds the timed launches only)")
    ap.add_argument("--serve-queries", type=int, default=16, help="served-stream leg: number of queries (0 = skip)")
    ap.add_argument("--serve-rate", type=float, default=200.0, help="mean query arrival rate (1/s)")
    ap.add_argument("--serve-scale", type=float, default=1.0,
                    help="trajectories per energy are log-uniform in [1e3, 1e5] times this factor (1: the workload's sizes)")
    ap.add_argument("--drain-traj", type=int, default=1_000_000, help="time-to-solution leg: trajectories run to completion (0 = skip)")
    ap.add_argument("--drain-seconds", type=float, default=90.0, help="time-to-solution leg: stop after this long")
    ap.add_argument("--serve-steps-per-launch", type=int, default=2000)
    ap.add_argument("--serve-seed", type=int, default=7)
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a plain device-to-device copy of this many MiB (3 times) and exit: the reference "
                         "kernel of known traffic that scales the FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.copy_probe_mib > 0:
        copy_probe(args.copy_probe_mib)
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: trajectories are independent and each GPU would take its own id range "
              "(the multi-GPU merge is a per-distribution sum); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.steps < 1 or args.warmup < 0:
        sys.exit("--steps must be >= 1")

    from spectrum_amd import configs as cf
    from spectrum_amd import engine as E
    log("device:", E.device_info())
    flops_all = load_json(os.path.join(ROOT, "profiles", "flops_per_step.json"))["workloads"]
    flops = flops_all[WORKLOAD]
    traffic = load_json(os.path.join(ROOT, "profiles", "traffic_parker_step_kernel.json"))
    traffic_grid = load_json(os.path.join(ROOT, "profiles", "traffic_grid_step_kernel.json"))
    if args.grid_only:
        print(json.dumps({"grid_workloads": grid_workloads(E, cf, args, traffic_grid)}), flush=True)
        return
    if args.other_only:
        print(json.dumps({"other_workloads": other_workloads(E, cf, args, flops_all)}), flush=True)
        return

    precs = ["fp64", "mixed"] if args.precision == "both" else [args.precision]
    legs = {}
    for p in precs:
        legs[p] = gpu_leg(E, cf, p, args, autotune=not args.no_autotune and not args.lanes)
    main_prec = "mixed" if "mixed" in legs else "fp64"
    m = legs[main_prec]
    value = m["steps"] / m["wall"]
    out = {
        "metric": "trajectory-steps/s FP64 vs mixed at equal 3σ MC error; p99 query latency",
        "value": value,
        "unit": "trajectory-steps/s",
        "n_gpus": 1,
        "steps": m["launches"],
        "warmup": args.warmup,
        "ms_per_step": m["wall"] * 1e3 / m["launches"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 step / f64 accumulate" if main_prec == "mixed" else "f64",
        "data": "synthetic (counter-based Philox increments, seed 1)",
        "config": {"workload": WORKLOAD, "n_traj": N_ENERGIES * PER_ENERGY, "energies_GeV": [0.1, 10.0], "transport": "Parker backward",
                   "background": "SolarWind (Parker spiral)", "diffusion": "RigidityMagneticFieldPowerLaw",
                   "lanes": m["lanes"], "steps_per_launch": m["steps_per_launch"],
                   "autotuned": bool(m["tuned"]), "precision": main_prec},
        "roofline": roofline(m, main_prec, flops, traffic),
    }
    if "fp64" in legs:
        f = legs["fp64"]
        out["fp64"] = {"value": f["steps"] / f["wall"], "ms_per_step": f["wall"] * 1e3 / f["launches"],
                       "roofline": roofline(f, "fp64", flops, traffic)}
        if main_prec == "mixed":
            out["speedup_mixed_vs_fp64"] = round(value / out["fp64"]["value"], 3)
    if not args.no_other_workloads:
        out["other_workloads"] = other_workloads(E, cf, args, flops_all)
    if not args.no_grid_workloads:
        out["grid_workloads"] = grid_workloads(E, cf, args, traffic_grid)
    if args.drain_traj > 0:
        out["time_to_solution"] = drain_leg(E, cf, args)
    if args.serve_queries > 0:
        out["serve"] = serve_leg(E, cf, args)
        out["p99_query_latency_s"] = out["serve"]["p99_latency_s"]
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_leg(cf, args.cpu_seconds, max(1, host_cpus()[0]))
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
