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
PREC_ID = {"fp64": 0, "mixed": 1}


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def load_json(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def gpu_leg(E, cf, prec, args):
    s = cf.parker_modulation(n_energies=N_ENERGIES)
    opts = dict(precision=PREC_ID[prec], steps_per_launch=args.steps_per_launch)
    if args.lanes:
        opts["lanes"] = args.lanes
    sim = E.Simulation.from_setup(s, **opts)
    sim.begin(N_ENERGIES * PER_ENERGY, seed=1)
    w = sim.advance(args.warmup)
    t0 = time.perf_counter()
    st = sim.advance(args.steps)
    wall = time.perf_counter() - t0
    steps = st.steps_executed - w.steps_executed
    log(f"{prec}: lanes {st.lanes} launches {args.steps} trajectory-steps {steps} wall {wall:.3f}s "
        f"step-kernel {st.step_kernel_ms:.1f} ms service {st.service_kernel_ms:.1f} ms -> {steps / wall:.4e} steps/s")
    return dict(steps=steps, wall=wall, lanes=st.lanes, step_ms=st.step_kernel_ms, service_ms=st.service_kernel_ms,
                done=st.done, failed=st.failed)


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
    flops = load_json(os.path.join(ROOT, "profiles", "flops_per_step.json"))["workloads"][WORKLOAD]
    traffic = load_json(os.path.join(ROOT, "profiles", "traffic_parker_step_kernel.json"))

    precs = ["fp64", "mixed"] if args.precision == "both" else [args.precision]
    legs = {}
    for p in precs:
        legs[p] = gpu_leg(E, cf, p, args)
        legs[p]["launches"] = args.steps
    main_prec = "mixed" if "mixed" in legs else "fp64"
    m = legs[main_prec]
    value = m["steps"] / m["wall"]
    out = {
        "metric": "trajectory-steps/s FP64 vs mixed at equal 3σ MC error; p99 query latency",
        "value": value,
        "unit": "trajectory-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": m["wall"] * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 step / f64 accumulate" if main_prec == "mixed" else "f64",
        "data": "synthetic (counter-based Philox increments, seed 1)",
        "config": {"workload": WORKLOAD, "n_traj": N_ENERGIES * PER_ENERGY, "energies_GeV": [0.1, 10.0], "transport": "Parker backward",
                   "background": "SolarWind (Parker spiral)", "diffusion": "RigidityMagneticFieldPowerLaw",
                   "lanes": m["lanes"], "steps_per_launch": args.steps_per_launch, "precision": main_prec},
        "roofline": roofline(m, main_prec, flops, traffic),
    }
    if "fp64" in legs:
        f = legs["fp64"]
        out["fp64"] = {"value": f["steps"] / f["wall"], "ms_per_step": f["wall"] * 1e3 / args.steps,
                       "roofline": roofline(f, "fp64", flops, traffic)}
        if main_prec == "mixed":
            out["speedup_mixed_vs_fp64"] = round(value / out["fp64"]["value"], 3)
    if not args.no_cpu:
        threads = max(1, min(16, len(os.sched_getaffinity(0))))
        out["cpu_baseline"] = cpu_leg(cf, args.cpu_seconds, threads)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
