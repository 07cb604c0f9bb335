"""Exact histogram sums (spectrum_amd/csrc/spx_acc.h), checked on the CPU against Python's math.fsum,
which returns the correctly rounded sum of a list of doubles (Shewchuk's algorithm).

The engine adds every event weight to its bin as 32-bit digits in 64-bit limbs; the value read back must be
the exact sum rounded once, whatever the order of the additions: the order the device's waves happen to
bin events in changes nothing.  The harness tests/native/acc_sum.cpp is compiled here with g++.
"""
import ctypes as C
import math
import os
import shutil
import subprocess
import zlib

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def acc(tmp_path_factory):
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    out = str(tmp_path_factory.mktemp("acc") / "libacc.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-o", out, os.path.join(HERE, "native", "acc_sum.cpp")], check=True)
    L = C.CDLL(out)
    L.acc_sum.restype = C.c_double
    L.acc_sum.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
    L.acc_round_trip.restype = C.c_double
    L.acc_round_trip.argtypes = [C.c_double]
    return L


def acc_sum(L, v, order=None):
    v = np.ascontiguousarray(v, dtype=np.float64)
    o = None if order is None else np.ascontiguousarray(order, dtype=np.int64)
    return L.acc_sum(v.ctypes.data, None if o is None else o.ctypes.data, len(v))


def same(a, b):
    return (math.isnan(a) and math.isnan(b)) or (a == b and math.copysign(1.0, a) == math.copysign(1.0, b)) or (a == b == 0.0)


@pytest.mark.parametrize("case", ["weights", "signed_wide", "cancellation", "tiny", "huge"])
def test_sum_is_the_correctly_rounded_exact_sum(acc, case):
    rng = np.random.default_rng(zlib.crc32(case.encode()))
    n = 20000
    if case == "weights":          # spectrum weights: positive, a few decades
        v = 10.0 ** rng.uniform(-6, 3, n)
    elif case == "signed_wide":    # <dx>-type weights: both signs, 600 binary orders of magnitude
        v = rng.choice([-1.0, 1.0], n) * np.ldexp(rng.uniform(1, 2, n), rng.integers(-300, 300, n))
    elif case == "cancellation":   # large terms that cancel exactly, leaving small ones
        big = np.ldexp(rng.uniform(1, 2, n // 2), rng.integers(40, 60, n // 2))
        v = np.concatenate([big, -big, rng.uniform(-1e-3, 1e-3, 500)])
    elif case == "tiny":           # subnormals and the smallest normals
        v = np.concatenate([np.ldexp(rng.uniform(0, 1, n), -1060), -np.ldexp(rng.uniform(0, 1, 100), -1022), [5e-324, -5e-324, 5e-324]])
    else:                          # near the top of the range, finite sum
        v = np.concatenate([np.ldexp(rng.uniform(1, 2, 64), 1015), -np.ldexp(rng.uniform(1, 2, 64), 1015)])
    want = math.fsum(v.tolist())
    for seed in range(3):
        order = np.random.default_rng(seed).permutation(len(v))
        got = acc_sum(acc, v, order)
        assert got == want, (case, got, want)


def test_ties_round_to_even(acc):
    one = 1.0
    ulp = 2.0 ** -52
    assert acc_sum(acc, [one, ulp / 2]) == one                     # tie, even stays
    assert acc_sum(acc, [one + ulp, ulp / 2]) == one + 2 * ulp      # tie, odd rounds up
    assert acc_sum(acc, [one, ulp / 2, 2.0 ** -200]) == one + ulp   # sticky bit breaks the tie
    assert acc_sum(acc, [one, -ulp / 4]) == one                     # below 1 the tie 1 - 2^-54 goes to 1.0
    assert acc_sum(acc, [one, -ulp / 4, -2.0 ** -200]) == one - ulp / 2


def test_specials_zero_and_overflow(acc):
    assert acc_sum(acc, []) == 0.0
    assert acc_sum(acc, [3.0, -3.0]) == 0.0
    assert acc_sum(acc, [1.0, math.inf]) == math.inf
    assert acc_sum(acc, [1.0, -math.inf]) == -math.inf
    assert math.isnan(acc_sum(acc, [math.inf, -math.inf]))
    assert math.isnan(acc_sum(acc, [1.0, math.nan]))
    assert acc_sum(acc, [1.7e308, 1.7e308]) == math.inf            # the exact sum overflows: rounds to inf
    assert acc_sum(acc, [1.7e308, 1.7e308, -1.7e308]) == 1.7e308    # but not while it fits


def test_round_trip_of_every_kind_of_double(acc):
    rng = np.random.default_rng(1)
    bits = rng.integers(0, 2**63, 20000, dtype=np.int64).view(np.float64)
    vals = [x for x in bits.tolist() if math.isfinite(x)] + [0.0, 5e-324, -5e-324, 2.2250738585072014e-308, 1.7976931348623157e308,
                                                             -1.7976931348623157e308, 1.0, -1.0]
    for x in vals:
        assert same(acc.acc_round_trip(x), x), x
