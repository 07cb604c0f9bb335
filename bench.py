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
        best = max(table, key=lambda e: e[2])
        tuned = {"lanes": best[0], "steps_per_launch": best[1], "candidates": len(table)}
        log(f"{s.name} {prec}: autotuned lanes {best[0]} steps/launch {best[1]} ({best[2]:.3e} steps/s in the probe)")
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
    return dict(steps=steps, wall=wall, lanes=st.lanes, step_ms=st.step_kernel_ms, service_ms=st.service_kernel_ms,
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
            "flops_per_step": round(fps, 1), "transcendentals_per_step": round(flops["transcendentals_per_step"], 1)}


def other_workloads(E, cf, args, flops_all):
    """Steady-state throughput of the other two transports at their full trajectory counts: diffusive
    shock acceleration (1e8 trajectories queued) and the guiding centre (1e7), both precisions."""
    out = {}
    for name, setup, n, key, tune in (("dsa_smooth_shock_1e8", cf.dsa_shock(), 100_000_000, "dsa_smooth_shock", True),
                                      ("gc_diff_scatt_1e7", cf.gc_diff_scatt(), 10_000_000, "gc_diff_scatt", False)):
        res = {}
        for p in ("fp64", "mixed"):
            leg = gpu_leg(E, cf, p, args, setup, n, autotune=tune)
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
    and a 64^3-block (256^3-zone, 1.07 GB) grid that no cache holds.  Each step interpolates the grid
    at the new position (kappa is isotropic here, so the six derivative points of div(kappa) need only
    their positions checked): one gather of 8 zones.  The algorithmic gather is 8 zones x 6 values x
    8 B (FP64) or 4 B (FP32) per interpolation; the measured HBM traffic per launch comes from the
    FETCH_SIZE / WRITE_SIZE profile when present."""
    from spectrum_amd import cartesian as G
    out = {}
    for nb in args.grid_blocks:
        grid = G.parker_spiral(nb=nb)
        setup = cf.parker_modulation_cartesian(grid)
        name = f"parker_modulation_cartesian_{nb}^3_blocks"
        res = {"zones": 4 * nb, "zone_records_bytes": {"fp64": 64 * (4 * nb) ** 3, "mixed": 32 * (4 * nb) ** 3}}
        for p in (("fp64", "mixed") if args.precision == "both" else (args.precision,)):
            leg = gpu_leg(E, cf, p, args, setup, 10_000_000, autotune=args.autotune and not args.lanes,
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
                r["hbm_bytes_per_launch"] = t["bytes_per_launch"]
                r["hbm_GBps"] = round(t["bytes_per_launch"] / (t["avg_launch_ms"] * 1e-3) / 1e9, 1)
                r["hbm_frac"] = round(r["hbm_GBps"] / PEAK_HBM_GBPS, 4)
            res[p] = r
        if "fp64" in res and "mixed" in res:
            res["speedup_mixed_vs_fp64"] = round(res["mixed"]["value"] / res["fp64"]["value"], 3)
        out[name] = res
        del grid, setup
    return out


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
    log(f"cpu: {n} trajectories, {steps} steps in {wall:.2f}s on {threads} threads -> {steps / wall:.4e} steps/s")
    return {"value": steps / wall, "unit": "trajectory-steps/s", "cores": threads, "kind": "port",
            "sample": f"{n} whole trajectories (ids 0..{n - 1} of the same workload, seed 1), FP64 oracle, "
                      f"OpenMP over trajectories, {wall:.1f} s"}


def serve_leg(E, cf, args):
    """A served stream: Poisson arrivals of modulation-spectrum queries (radius U[1, 80] au, 8-64
    energies, log-uniform trajectories per energy) with 20% diffusive-shock-acceleration queries,
    submitted at their arrival times to a query server running the mixed kernel."""
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
        for q in ids:
            srv.wait(q)
        makespan = time.perf_counter() - t0
        info = [srv.info(q) for q in ids]
        st = srv.stats()
    lat = np.array([i["latency_s"] for i in info])
    steps = sum(i["steps_executed"] for i in info)
    n_traj = sum(p[2] for p in plan)
    log(f"serve: {len(ids)} queries ({sum(p[4] == 'dsa' for p in plan)} DSA), {n_traj} trajectories, {steps} steps in {makespan:.2f}s; "
        f"latency p50 {np.percentile(lat, 50):.3f}s p99 {np.percentile(lat, 99):.3f}s; engine busy {st['busy_s']:.2f}s")
    return {"queries": len(ids), "dsa_queries": int(sum(p[4] == "dsa" for p in plan)), "arrival_rate_hz": args.serve_rate,
            "trajectories": int(n_traj), "trajectories_per_energy": [round(1e3 * args.serve_scale, 1), round(1e5 * args.serve_scale, 1)],
            "p50_latency_s": round(float(np.percentile(lat, 50)), 4), "p99_latency_s": round(float(np.percentile(lat, 99)), 4),
            "mean_latency_s": round(float(lat.mean()), 4), "makespan_s": round(makespan, 3),
            "trajectory_steps_per_s": steps / makespan, "precision": "mixed", "steps_per_launch": args.serve_steps_per_launch}


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
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
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
    ap.add_argument("--autotune", action="store_true",
                    help="autotune the modulation legs' launch shape first (the kernel trace then also holds the probe launches)")
    ap.add_argument("--serve-queries", type=int, default=40, help="served-stream leg: number of queries (0 = skip)")
    ap.add_argument("--serve-rate", type=float, default=200.0, help="mean query arrival rate (1/s)")
    ap.add_argument("--serve-scale", type=float, default=0.01,
                    help="trajectories per energy are log-uniform in [1e3, 1e5] times this factor")
    ap.add_argument("--serve-steps-per-launch", type=int, default=500)
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

    precs = ["fp64", "mixed"] if args.precision == "both" else [args.precision]
    legs = {}
    for p in precs:
        legs[p] = gpu_leg(E, cf, p, args, autotune=args.autotune and not args.lanes)
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
    if args.serve_queries > 0:
        out["serve"] = serve_leg(E, cf, args)
        out["p99_query_latency_s"] = out["serve"]["p99_latency_s"]
    if not args.no_cpu:
        threads = max(1, min(16, len(os.sched_getaffinity(0))))
        out["cpu_baseline"] = cpu_leg(cf, args.cpu_seconds, threads)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
