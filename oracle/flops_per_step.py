"""Algorithmic operations per trajectory step, counted mechanically (test infrastructure).

Runs each measured workload through liboracle_flops.so (the oracle compiled with an
operation-counting scalar, oracle/flopcount.cpp), checks that its trajectories are
bit-identical to the plain oracle's, and writes the per-step averages to
profiles/flops_per_step.json.  Floating-point operations are add/sub, mul, div and sqrt
(1 flop each); exp/log/pow/tanh/trig calls are counted separately as transcendentals;
comparisons, negations, fabs/fmin/fmax are not flops.

    python oracle/flops_per_step.py            # all workloads
    python oracle/flops_per_step.py NAME ...   # only these, merged into the existing file
"""
from __future__ import annotations

import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

import oracle as O  # noqa: E402
from spectrum_amd import cartesian as G  # noqa: E402
from spectrum_amd import configs as cf  # noqa: E402

FLOP = ("add", "mul", "div", "sqrt")
TRANS = ("exp", "log", "pow", "tanh", "trig")


def workloads():
    """(name, setup, n trajectories counted, seed); the trajectories are ids 0..n-1 of each workload."""
    return [
        ("parker_modulation_64x156250", cf.parker_modulation(n_energies=64), 64, 1),
        ("dsa_smooth_shock", cf.dsa_shock(), 32, 1),
        ("gc_diff_scatt", cf.gc_diff_scatt(), 16, 1),
        ("parker_modulation_cartesian_25^3_blocks", cf.parker_modulation_cartesian(G.parker_spiral(nb=25)), 256, 1),
    ]


def count(setup, n, seed):
    O.flops_reset()
    of = O.Oracle(setup, O.FLOPS_LIB_PATH)
    t0 = time.time()
    steps, rec = of.run(n, seed=seed, threads=1)
    wall = time.time() - t0
    ops = O.flops_read()
    ref = O.Oracle(setup)
    steps_ref, rec_ref = ref.run(n, seed=seed, threads=1)
    assert steps == steps_ref and rec.tobytes() == rec_ref.tobytes(), "counting build diverged from the oracle"
    per = {k: v / steps for k, v in ops.items()}
    return {
        "trajectories": n, "seed": seed, "steps": steps, "count_wall_s": round(wall, 2),
        "flops_per_step": sum(per[k] for k in FLOP),
        "transcendentals_per_step": sum(per[k] for k in TRANS),
        "per_step": {k: round(v, 4) for k, v in per.items()},
    }


def main():
    out_path = os.path.join(ROOT, "profiles", "flops_per_step.json")
    result = {"definition": "add/sub, mul, div, sqrt = 1 flop each, averaged over every step of the counted "
                            "trajectories (start sampling, boundaries and binning included); transcendental "
                            "calls (exp, log, pow, tanh, trig) are listed separately and not charged as flops",
              "command": "make -C oracle && python oracle/flops_per_step.py",
              "workloads": {}}
    only = sys.argv[1:]
    if only and os.path.exists(out_path):
        with open(out_path) as f:
            result["workloads"] = json.load(f)["workloads"]
    for name, setup, n, seed in workloads():
        if only and name not in only:
            continue
        r = count(setup, n, seed)
        print(name, json.dumps(r), flush=True)
        result["workloads"][name] = r
    os.makedirs(os.path.dirname(out_path), exist_ok=True)
    with open(out_path, "w") as f:
        json.dump(result, f, indent=1)
    print("wrote", out_path)


if __name__ == "__main__":
    main()
