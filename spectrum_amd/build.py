"""Builds the HIP engine library in-tree (gfx950) with hipcc.

The product is ``spectrum_amd/libspectrum_amd.so``: the C ABI of include/spectrum_amd.h
over the HIP kernels of spectrum_amd/csrc.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libspectrum_amd.so")
SOURCES = ["spx_kernels_f64.hip", "spx_kernels_f32.hip", "spx_runtime.cpp", "spx_api.cpp", "spx_static.hip"]
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # FP64 kernels keep the reference's rounding: no contraction into multiply-adds
    "-ffp-contract=off",
    # float kernels use the hardware square root / reciprocal paths
    "-fno-hip-fp32-correctly-rounded-divide-sqrt",
    # no packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
# The FP32-step kernels: multiply-adds, reciprocals and approximate library functions are allowed
# (their error is far below the Monte Carlo error they are compared at); NaN checks stay valid.
F32_FLAGS = ["-ffp-contract=fast", "-freciprocal-math", "-fno-signed-zeros", "-fapprox-func"]


def flags_for(src: str) -> list:
    if src.endswith("_f32.hip"):
        return [f for f in HIP_FLAGS if f != "-ffp-contract=off"] + F32_FLAGS
    return HIP_FLAGS


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    """Compile every source to an object in parallel, then link the shared library."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    procs = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, src + ".o")
        objs.append(obj)
        cmd = [hipcc(), *flags_for(src), "-c", "-I", os.path.join(ROOT, "include"), "-o", obj, os.path.join(CSRC, src)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
    failed = False
    for src, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out)
            failed = True
    if failed:
        raise RuntimeError("hipcc failed")
    tmp = LIB + ".tmp"
    cmd = [hipcc(), *HIP_FLAGS, "-shared", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc link failed")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
