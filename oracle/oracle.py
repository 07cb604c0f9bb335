"""ctypes wrapper of the CPU oracle (test infrastructure only).

Only the test suite, the smoke check and the CPU-baseline leg of bench.py use
this module.  It loads ``liboracle.so`` built from ``spx_oracle.c`` in this
directory (``make -C oracle``).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
FLOPS_LIB_PATH = os.path.join(HERE, "liboracle_flops.so")   # same source, operation-counting scalar
LIBM_LIB_PATH = os.path.join(HERE, "liboracle_libm.so")     # same source, pow / cbrt one ulp larger
FMA_LIB_PATH = os.path.join(HERE, "liboracle_fma.so")       # same source, multiply-adds contracted (last-bit perturbation)
COMP_LIB_PATH = os.path.join(HERE, "liboracle_comp.so")     # same source, Kahan-compensated sums of t, x, p in the Parker step

RNG_COUNTER, RNG_REPLAY = 0, 1
EX_NAMES = {0: "none", 1: "field", 2: "step_too_small", 3: "step_nan", 4: "max_steps", 5: "replay_exhausted", 6: "boundary"}


class OrcRng(C.Structure):
    _fields_ = [("mode", C.c_int32), ("seed", C.c_uint64), ("normals", C.POINTER(C.c_double)),
                ("uniforms", C.POINTER(C.c_double)), ("normal_stride", C.c_int64), ("uniform_stride", C.c_int64)]


class OrcTraj(C.Structure):
    _fields_ = [("t", C.c_double), ("pos", C.c_double * 3), ("mom", C.c_double * 3), ("amp", C.c_double), ("wgt", C.c_double),
                ("t0", C.c_double), ("pos0", C.c_double * 3), ("mom0", C.c_double * 3),
                ("steps", C.c_int64), ("status", C.c_int32), ("attempts", C.c_int32), ("n_events", C.c_int32),
                ("ev_bin", C.c_int32 * 4), ("ev_weight", C.c_double * 4)]


TRAJ_DTYPE = np.dtype([("t", "f8"), ("pos", "f8", 3), ("mom", "f8", 3), ("amp", "f8"), ("wgt", "f8"),
                       ("t0", "f8"), ("pos0", "f8", 3), ("mom0", "f8", 3), ("steps", "i8"), ("status", "i4"),
                       ("attempts", "i4"), ("n_events", "i4"), ("ev_bin", "i4", 4), ("ev_weight", "f8", 4)], align=True)
assert TRAJ_DTYPE.itemsize == C.sizeof(OrcTraj)

_libs: dict = {}


def lib(path: str = LIB_PATH):
    if path not in _libs:
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: {path} (run `make -C oracle`)")
        L = C.CDLL(path)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.c_int, C.c_int]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_add.argtypes = [C.c_void_p, C.c_int, C.c_char_p, C.c_char_p, C.c_size_t]
        L.orc_set_option.argtypes = [C.c_void_p, C.c_char_p, C.c_double]
        L.orc_run.restype = C.c_int64
        L.orc_run.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.POINTER(OrcRng), C.c_int, C.c_void_p]
        L.orc_distro_nbins.argtypes = [C.c_void_p, C.c_int]
        L.orc_distro_width.argtypes = [C.c_void_p, C.c_int]
        L.orc_get_distro.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.POINTER(C.c_int64)]
        L.orc_reset_distros.argtypes = [C.c_void_p]
        L.orc_discards.restype = C.c_int64
        L.orc_discards.argtypes = [C.c_void_p, C.c_int]
        L.orc_background_fields.argtypes = [C.c_void_p, C.c_double, C.POINTER(C.c_double), C.c_double,
                                            C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_kappa.restype = C.c_double
        L.orc_set_grid.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int32, C.c_void_p]
        L.orc_grid_fields.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_kappa_at.argtypes = [C.c_void_p, C.c_int, C.c_double, C.POINTER(C.c_double), C.c_double, C.c_double]
        L.orc_kappa_at.restype = C.c_double
        L.orc_hyp2f1_11.argtypes = [C.c_double, C.c_double]
        L.orc_gc_vperp.argtypes = [C.c_void_p, C.c_double, C.POINTER(C.c_double), C.c_double, C.c_double, C.POINTER(C.c_double)]
        L.orc_gc_vperp.restype = C.c_int
        L.orc_hyp2f1_11.restype = C.c_double
        L.orc_kappa.argtypes = [C.c_void_p, C.c_int, C.c_double, C.POINTER(C.c_double), C.c_double, C.c_double, C.c_double]
        L.orc_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.orc_counter_normal.restype = C.c_double
        L.orc_counter_normal.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
        L.orc_counter_uniform.restype = C.c_double
        L.orc_counter_uniform.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
        L.orc_const.restype = C.c_double
        L.orc_const.argtypes = [C.c_char_p]
        if hasattr(L, "orc_flops_read"):
            L.orc_flops_read.argtypes = [C.POINTER(C.c_ulonglong), C.c_int]
            L.orc_flops_reset.argtypes = []
            L.orc_flops_reset.restype = None
        _libs[path] = L
    return _libs[path]


class Oracle:
    """One configured oracle simulation (mirrors a configured SimulationWorker)."""

    def __init__(self, setup, path: str = LIB_PATH):
        self.L = L = lib(path)
        self.setup = setup
        self.h = L.orc_create(setup.traj_type, setup.time_flow)
        for role, cls, blob in setup.plugins:
            rc = L.orc_add(self.h, role, cls.encode(), blob, len(blob))
            if rc != 0:
                raise ValueError(f"oracle rejected {cls} (rc={rc})")
        for k, v in setup.options.items():
            if L.orc_set_option(self.h, k.encode(), float(v)) != 0:
                raise ValueError(f"unknown option {k}")
        grid = getattr(setup, "grid", None)
        if grid is not None:
            self._grid = np.ascontiguousarray(grid.data, dtype=np.float64)
            nb = (C.c_int32 * 3)(*grid.nb)
            lo = (C.c_double * 3)(*grid.dom_min)
            hi = (C.c_double * 3)(*grid.dom_max)
            if L.orc_set_grid(self.h, nb, lo, hi, 6, self._grid.ctypes.data) != 0:
                raise ValueError("oracle rejected the grid")

    def __del__(self):
        try:
            self.L.orc_destroy(self.h)
        except Exception:
            pass

    def set_option(self, name, value):
        if self.L.orc_set_option(self.h, name.encode(), float(value)) != 0:
            raise ValueError(name)

    def run(self, n_traj, seed=1, first=0, threads=None, normals=None, uniforms=None, records=True):
        rng = OrcRng()
        keep = []
        if normals is not None:
            normals = np.ascontiguousarray(normals, dtype=np.float64)
            uniforms = np.ascontiguousarray(uniforms if uniforms is not None else np.zeros((normals.shape[0], 1)), dtype=np.float64)
            keep += [normals, uniforms]
            rng.mode = RNG_REPLAY
            rng.normals = normals.ctypes.data_as(C.POINTER(C.c_double))
            rng.uniforms = uniforms.ctypes.data_as(C.POINTER(C.c_double))
            rng.normal_stride = normals.shape[1]
            rng.uniform_stride = uniforms.shape[1]
        else:
            rng.mode = RNG_COUNTER
        rng.seed = seed
        out = np.zeros(n_traj, dtype=TRAJ_DTYPE) if records else None
        nthr = threads or os.cpu_count() or 1
        steps = self.L.orc_run(self.h, first, n_traj, C.byref(rng), nthr, out.ctypes.data if records else None)
        if steps < 0:
            raise RuntimeError(f"orc_run failed ({steps})")
        return int(steps), out

    def distro(self, d):
        L = self.L
        nb = L.orc_distro_nbins(self.h, d)
        w = L.orc_distro_width(self.h, d)
        counts = np.zeros(nb, dtype=np.int64)
        sums = np.zeros(nb * w, dtype=np.float64)
        nev = C.c_int64()
        L.orc_get_distro(self.h, d, counts.ctypes.data, sums.ctypes.data, C.byref(nev))
        return counts, sums.reshape(nb, w) if w > 1 else sums, nev.value

    def reset(self):
        self.L.orc_reset_distros(self.h)

    def discards(self):
        return {EX_NAMES[k]: self.L.orc_discards(self.h, k) for k in range(1, 7)}

    def background_fields(self, t, pos, p):
        P = (C.c_double * 3)(*pos)
        U = (C.c_double * 3)()
        B = (C.c_double * 3)()
        dmax, divU = C.c_double(), C.c_double()
        self.L.orc_background_fields(self.h, t, P, p, U, B, C.byref(dmax), C.byref(divU))
        return np.array(U), np.array(B), dmax.value, divU.value

    def grid_fields(self, pos):
        """Gridded fields at one point: dict of U, B, Bmag, gradU, gradB, gradBmag, dmax, E; None where the
        reference fails (outside the grid)."""
        out = (C.c_double * 32)()
        rc = self.L.orc_grid_fields(self.h, (C.c_double * 3)(*pos), out)
        if rc < 0:
            raise ValueError("no gridded background")
        if rc:
            return None
        a = np.frombuffer(out, dtype=np.float64).copy()
        return {"U": a[0:3], "B": a[3:6], "Bmag": a[6], "gradU": a[7:16].reshape(3, 3), "gradB": a[16:25].reshape(3, 3),
                "gradBmag": a[25:28], "dmax": a[28], "E": a[29:32]}

    def gc_vperp(self, pos, p, mu, t=0.0):
        """(kappa_perp, V_perp) of the guiding centre at a point."""
        out = (C.c_double * 4)()
        if self.L.orc_gc_vperp(self.h, t, (C.c_double * 3)(*pos), p, mu, out):
            raise RuntimeError("field error")
        return out[0], np.array(out[1:4])

    def kappa_at(self, comp, pos, p, mu=0.0, t=0.0):
        """Diffusion component `comp` with the background's own fields and region indicators at pos."""
        return self.L.orc_kappa_at(self.h, comp, t, (C.c_double * 3)(*pos), p, mu)

    def kappa(self, comp, pos, p, mu=0.0, Bmag=1.0, t=0.0):
        P = (C.c_double * 3)(*pos)
        return self.L.orc_kappa(self.h, comp, t, P, p, mu, Bmag)


def hyp2f1_11(c, z):
    """The oracle's 2F1(1, 1; c; z)."""
    return lib().orc_hyp2f1_11(c, z)


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return list(o)


def const(name):
    return lib().orc_const(name.encode())


FLOP_KINDS = ["add", "mul", "div", "sqrt", "cmp", "neg", "exp", "log", "pow", "tanh", "trig", "absminmax"]


def flops_reset():
    lib(FLOPS_LIB_PATH).orc_flops_reset()


def flops_read() -> dict:
    buf = (C.c_ulonglong * 16)()
    n = lib(FLOPS_LIB_PATH).orc_flops_read(buf, 16)
    return {FLOP_KINDS[k]: int(buf[k]) for k in range(n)}
