"""The FP64 kernels divide several numerators by one denominator with one correctly rounded reciprocal
y = 1 / b and, per numerator, q = a y corrected twice by the exact residual a - b q (spx::div_rn in
spectrum_amd/csrc/spx_vec.h; Markstein's theorem).  The claim is that this rounds exactly as the division
a / b does, so the kernels keep the reference's rounding; it is checked here on the CPU (the harness
tests/native/div_rn.cpp built with hipcc for the host, no contraction) against the host's correctly rounded
IEEE division: random operands over 600 binary orders of magnitude, quotients that are exact or within an
ulp of a representable value (the rounding decisions), and the operand ranges of the transports' fields.
"""
import ctypes as C
import os
import shutil
import subprocess
import zlib

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = "/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else shutil.which("hipcc")


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    if not HIPCC:
        pytest.skip("hipcc not available")
    out = str(tmp_path_factory.mktemp("divrn") / "libdivrn.so")
    subprocess.run([HIPCC, "-O2", "-std=c++17", "-shared", "-fPIC", "-ffp-contract=off", "-o", out,
                    os.path.join(HERE, "native", "div_rn.cpp")], check=True, capture_output=True)
    L = C.CDLL(out)
    L.div_rn_mismatches.restype = C.c_long
    L.div_rn_mismatches.argtypes = [C.c_void_p, C.c_void_p, C.c_long, C.POINTER(C.c_long)]
    L.v3_div.argtypes = [C.c_void_p, C.c_double, C.c_void_p]
    return L


def mismatches(L, a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    first = C.c_long(-1)
    n = L.div_rn_mismatches(a.ctypes.data, b.ctypes.data, len(a), C.byref(first))
    return n, (None if first.value < 0 else (a[first.value], b[first.value]))


@pytest.mark.parametrize("case", ["wide", "fields", "exact_quotients", "near_exact", "mantissa_extremes"])
def test_shared_reciprocal_division_rounds_like_the_division(lib, case):
    rng = np.random.default_rng(zlib.crc32(case.encode()))
    n = 2_000_000
    if case == "wide":                # both signs, 600 binary orders of magnitude
        a = rng.choice([-1.0, 1.0], n) * np.ldexp(rng.uniform(1, 2, n), rng.integers(-300, 300, n))
        b = rng.choice([-1.0, 1.0], n) * np.ldexp(rng.uniform(1, 2, n), rng.integers(-300, 300, n))
    elif case == "fields":            # code-unit positions, fields and norms of the transports
        a = rng.normal(0, 1, n) * 10.0 ** rng.uniform(-6, 6, n)
        b = 10.0 ** rng.uniform(-4, 6, n)
    elif case == "exact_quotients":   # a = q b exactly (27-bit q and b): the quotient is representable
        q = np.ldexp(rng.integers(2 ** 26, 2 ** 27, n).astype(np.float64), -26)
        bb = np.ldexp(rng.integers(2 ** 25, 2 ** 26, n).astype(np.float64), rng.integers(-60, 60, n))
        a, b = q * bb, bb
    elif case == "near_exact":        # one ulp either side of an exact quotient: the rounding decision
        q = np.ldexp(rng.integers(2 ** 26, 2 ** 27, n).astype(np.float64), -26)
        bb = np.ldexp(rng.integers(2 ** 25, 2 ** 26, n).astype(np.float64), rng.integers(-60, 60, n))
        a = np.nextafter(q * bb, np.where(rng.random(n) < 0.5, np.inf, -np.inf))
        b = bb
    else:                             # mantissas at the ends of the binade
        m = np.where(rng.random(n) < 0.5, 1.0 + rng.integers(0, 2 ** 20, n) * 2.0 ** -52, 2.0 - rng.integers(1, 2 ** 20, n) * 2.0 ** -52)
        a = np.ldexp(m, rng.integers(-40, 40, n))
        m2 = np.where(rng.random(n) < 0.5, 1.0 + rng.integers(0, 2 ** 20, n) * 2.0 ** -52, 2.0 - rng.integers(1, 2 ** 20, n) * 2.0 ** -52)
        b = np.ldexp(m2, rng.integers(-40, 40, n))
    bad, first = mismatches(lib, a, b)
    assert bad == 0, f"{bad} quotients differ from the division, first {first}"


def test_vector_division_and_zero(lib):
    out = (C.c_double * 3)()
    v = (C.c_double * 3)(3.0, -0.0, 1.0e-300)
    lib.v3_div(v, 7.0, out)
    assert out[0] == 3.0 / 7.0 and out[1] == 0.0 and out[2] == 1.0e-300 / 7.0
