"""Python host mirror of SPECTRUM's Simulation interface over the HIP engine.

``Simulation`` keeps the reference's driver vocabulary (src/simulation.hh:37-292):
``add_background / add_diffusion / add_source / add_boundary / add_initial /
add_distribution`` take the reference's class name and a DataContainer, ``run``
corresponds to SetTasks + MainLoop, and ``distro`` returns the counts and sums of
DistributionTemplated (src/distribution_templated.cc).  Everything executes in
``libspectrum_amd.so`` (C ABI: include/spectrum_amd.h); there is no CPU fallback,
and a missing library raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from . import configs as cf
from .data_container import DataContainer

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libspectrum_amd.so")

ERRORS = {0: "ok", -1: "bad argument", -2: "unknown class", -3: "container mismatch", -4: "not ready",
          -5: "device error", -6: "wrong call order", -7: "limit exceeded"}
EX_NAMES = {1: "field", 2: "step_too_small", 3: "step_nan", 4: "max_steps", 5: "replay_exhausted", 6: "boundary"}

PREC_FP64, PREC_MIXED = 0, 1


class SpxRng(C.Structure):
    _fields_ = [("mode", C.c_int32), ("pad_", C.c_int32), ("seed", C.c_uint64), ("normals", C.POINTER(C.c_double)),
                ("uniforms", C.POINTER(C.c_double)), ("normal_stride", C.c_int64), ("uniform_stride", C.c_int64)]


class SpxRunStats(C.Structure):
    _fields_ = [("n_traj", C.c_uint64), ("done", C.c_uint64), ("failed", C.c_uint64), ("steps_done", C.c_uint64),
                ("steps_executed", C.c_uint64), ("discards", C.c_uint64 * 8), ("t_min", C.c_double), ("t_max", C.c_double),
                ("wall_s", C.c_double), ("kernel_ms", C.c_double), ("launches", C.c_int32), ("lanes", C.c_int32),
                ("step_kernel_ms", C.c_double), ("service_kernel_ms", C.c_double)]


class SpxTuneEntry(C.Structure):
    _fields_ = [("lanes", C.c_int64), ("steps_per_launch", C.c_int32), ("graph_launches", C.c_int32), ("steps_per_s", C.c_double)]


RECORD_DTYPE = np.dtype([("t", "f8"), ("pos", "f8", 3), ("mom", "f8", 3), ("amp", "f8"), ("wgt", "f8"),
                         ("t0", "f8"), ("pos0", "f8", 3), ("mom0", "f8", 3), ("steps", "i8"), ("status", "i4"),
                         ("attempts", "i4"), ("n_events", "i4"), ("ev_bin", "i4", 4), ("ev_weight", "f8", 4)], align=True)

_lib = None


def lib() -> C.CDLL:
    """Load the engine library; raises when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP engine library missing: {LIB_PATH}; build it with `python -m spectrum_amd.build`")
        L = C.CDLL(LIB_PATH)
        L.spx_create.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_void_p)]
        L.spx_destroy.argtypes = [C.c_void_p]
        L.spx_destroy.restype = None
        L.spx_add_plugin.argtypes = [C.c_void_p, C.c_int, C.c_char_p, C.c_char_p, C.c_size_t]
        L.spx_set_option.argtypes = [C.c_void_p, C.c_char_p, C.c_double]
        L.spx_run.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.POINTER(SpxRng), C.c_void_p, C.POINTER(SpxRunStats)]
        L.spx_begin.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.POINTER(SpxRng), C.c_int]
        L.spx_run_static.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(SpxRunStats)]
        L.spx_advance.argtypes = [C.c_void_p, C.c_int, C.POINTER(SpxRunStats)]
        L.spx_finish.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(SpxRunStats)]
        L.spx_distro_info.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
        L.spx_get_distro.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.POINTER(C.c_int64)]
        L.spx_reset_distros.argtypes = [C.c_void_p]
        L.spx_last_error.argtypes = [C.c_void_p]
        L.spx_last_error.restype = C.c_char_p
        L.spx_device_info.argtypes = [C.c_char_p, C.c_size_t]
        L.spx_set_background_grid.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int32, C.c_void_p]
        L.spx_add_background_file.argtypes = [C.c_void_p, C.c_char_p, C.c_char_p, C.c_size_t, C.c_char_p]
        L.spx_dump_distro.argtypes = [C.c_void_p, C.c_int, C.c_char_p]
        L.spx_restore_distro.argtypes = [C.c_void_p, C.c_int, C.c_char_p]
        L.spx_print_distro1d.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_char_p, C.c_int]
        L.spx_print_distro2d.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_char_p, C.c_int]
        L.spx_set_tasks.argtypes = [C.c_void_p, C.c_int64, C.c_int64, C.c_int64]
        L.spx_distro_file_name.argtypes = [C.c_void_p, C.c_char_p]
        L.spx_restore_distro_at_start.argtypes = [C.c_void_p, C.c_int]
        L.spx_main_loop.argtypes = [C.c_void_p, C.POINTER(SpxRng), C.POINTER(SpxRunStats)]
        L.spx_autotune.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.POINTER(SpxTuneEntry), C.c_int, C.POINTER(C.c_int)]
        _lib = L
    return _lib


def device_info() -> str:
    buf = C.create_string_buffer(256)
    rc = lib().spx_device_info(buf, 256)
    return buf.value.decode() if rc == 0 else ""


@dataclass
class RunStats:
    n_traj: int
    done: int
    failed: int
    steps_done: int
    steps_executed: int
    discards: dict
    t_min: float
    t_max: float
    wall_s: float
    kernel_ms: float
    launches: int
    lanes: int
    step_kernel_ms: float
    service_kernel_ms: float

    @staticmethod
    def of(s: SpxRunStats) -> "RunStats":
        return RunStats(s.n_traj, s.done, s.failed, s.steps_done, s.steps_executed,
                        {EX_NAMES[k]: int(s.discards[k]) for k in range(1, 7)}, s.t_min, s.t_max, s.wall_s,
                        s.kernel_ms, s.launches, s.lanes, s.step_kernel_ms, s.service_kernel_ms)


class SpxError(RuntimeError):
    pass


class Simulation:
    """A configured trajectory simulation on the GPU (the SimulationWorker API of the reference)."""

    def __init__(self, traj_type: int = cf.TRAJ_PARKER, time_flow: int = cf.BACKWARD):
        self.h = C.c_void_p()
        self._check(lib().spx_create(traj_type, time_flow, C.byref(self.h)), "create")
        self.n_dist = 0

    @classmethod
    def from_setup(cls, setup: cf.Setup, **options) -> "Simulation":
        sim = cls(setup.traj_type, setup.time_flow)
        for role, name, blob in setup.plugins:
            sim.add_plugin(role, name, blob)
        if setup.grid is not None:
            sim.set_background_grid(setup.grid)
        for k, v in {**setup.options, **options}.items():
            sim.set_option(k, v)
        return sim

    def __del__(self):
        try:
            if self.h:
                lib().spx_destroy(self.h)
                self.h = C.c_void_p()
        except Exception:
            pass

    def _check(self, rc: int, what: str) -> None:
        if rc != 0:
            msg = ""
            if self.h:
                msg = lib().spx_last_error(self.h).decode()
            raise SpxError(f"{what}: {ERRORS.get(rc, rc)} {msg}")

    # --- SimulationWorker::Add* ---------------------------------------------------------
    def add_plugin(self, role: int, class_name: str, container) -> None:
        blob = container.bytes() if isinstance(container, DataContainer) else bytes(container)
        self._check(lib().spx_add_plugin(self.h, role, class_name.encode(), blob, len(blob)), class_name)
        if role == cf.DISTRIBUTION:
            self.n_dist += 1

    def add_background(self, class_name, container, fname_pattern: str = None) -> None:
        """SimulationWorker::AddBackground (src/simulation.hh:111); a file pattern loads the grid of a
        BackgroundServerCartesian from ``pattern``.info / ``pattern``.out."""
        if fname_pattern is None:
            self.add_plugin(cf.BACKGROUND, class_name, container)
            return
        blob = container.bytes() if isinstance(container, DataContainer) else bytes(container)
        self._check(lib().spx_add_background_file(self.h, class_name.encode(), blob, len(blob), fname_pattern.encode()), class_name)

    def set_background_grid(self, grid) -> None:
        """The grid of a BackgroundServerCartesian from memory (a cartesian.CartesianGrid)."""
        nb = (C.c_int32 * 3)(*grid.nb)
        lo = (C.c_double * 3)(*grid.dom_min)
        hi = (C.c_double * 3)(*grid.dom_max)
        data = np.ascontiguousarray(grid.data, dtype=np.float64)
        self._check(lib().spx_set_background_grid(self.h, nb, lo, hi, 6, data.ctypes.data), "set_background_grid")

    def add_diffusion(self, class_name, container): self.add_plugin(cf.DIFFUSION, class_name, container)
    def add_source(self, class_name, container): self.add_plugin(cf.SOURCE, class_name, container)
    def add_boundary(self, class_name, container): self.add_plugin(cf.BOUNDARY, class_name, container)
    def add_initial(self, class_name, container): self.add_plugin(cf.INITIAL, class_name, container)
    def add_distribution(self, class_name, container): self.add_plugin(cf.DISTRIBUTION, class_name, container)

    def set_option(self, name: str, value) -> None:
        self._check(lib().spx_set_option(self.h, name.encode(), float(value)), f"option {name}")

    # --- runs -----------------------------------------------------------------------------
    @staticmethod
    def _rng(seed: int, normals, uniforms, keep: list) -> SpxRng:
        r = SpxRng()
        r.seed = seed
        if normals is None:
            r.mode = 0
            return r
        normals = np.ascontiguousarray(normals, dtype=np.float64)
        uniforms = np.ascontiguousarray(uniforms if uniforms is not None else np.full((normals.shape[0], 1), 0.5), dtype=np.float64)
        keep += [normals, uniforms]
        r.mode = 1
        r.normals = normals.ctypes.data_as(C.POINTER(C.c_double))
        r.uniforms = uniforms.ctypes.data_as(C.POINTER(C.c_double))
        r.normal_stride = normals.shape[1]
        r.uniform_stride = uniforms.shape[1]
        return r

    def run(self, n_traj: int, seed: int = 1, first: int = 0, normals=None, uniforms=None, records: bool = True):
        """Integrate trajectories first .. first+n_traj-1; returns (RunStats, records or None)."""
        keep: list = []
        rng = self._rng(seed, normals, uniforms, keep)
        st = SpxRunStats()
        rec = np.zeros(n_traj, dtype=RECORD_DTYPE) if records else None
        self._check(lib().spx_run(self.h, first, n_traj, C.byref(rng), rec.ctypes.data if records else None, C.byref(st)), "run")
        return RunStats.of(st), rec

    def begin(self, n_traj: int, seed: int = 1, first: int = 0, normals=None, uniforms=None, records: bool = False) -> None:
        self._keep = []
        rng = self._rng(seed, normals, uniforms, self._keep)
        self._check(lib().spx_begin(self.h, first, n_traj, C.byref(rng), 1 if records else 0), "begin")
        self._n = n_traj

    def advance(self, n_launches: int) -> RunStats:
        st = SpxRunStats()
        self._check(lib().spx_advance(self.h, n_launches, C.byref(st)), "advance")
        return RunStats.of(st)

    def finish(self, records: bool = False):
        st = SpxRunStats()
        rec = np.zeros(self._n, dtype=RECORD_DTYPE) if records else None
        self._check(lib().spx_finish(self.h, rec.ctypes.data if records else None, C.byref(st)), "finish")
        return RunStats.of(st), rec

    # --- SimulationMaster's task loop (src/simulation.cc:496-519, :818-845) --------------------------------
    def set_tasks(self, n_traj: int, batch_size: int = 1, max_traj_per_worker: int = 0) -> None:
        """SetTasks: the main loop's trajectory count and batch size (with the reference's clamps)."""
        self._check(lib().spx_set_tasks(self.h, n_traj, batch_size, max_traj_per_worker), "set_tasks")

    def distro_file_name(self, base: str) -> None:
        """DistroFileName: distributions are written to <base><d>.out (default "distribution_")."""
        self._check(lib().spx_distro_file_name(self.h, base.encode()), "distro_file_name")

    def restore_distro_at_start(self, d: int) -> None:
        """SimulationMaster::RestoreDistro: the main loop starts from <base><d>.out."""
        self._check(lib().spx_restore_distro_at_start(self.h, d), "restore_distro_at_start")

    def main_loop(self, seed: int = 1, normals=None, uniforms=None) -> RunStats:
        """MainLoop: reset (and restore) the distributions, integrate the tasks with progress dumps and
        reports on stderr, dump the final distributions."""
        keep: list = []
        rng = self._rng(seed, normals, uniforms, keep)
        st = SpxRunStats()
        self._check(lib().spx_main_loop(self.h, C.byref(rng), C.byref(st)), "main_loop")
        return RunStats.of(st)

    def autotune(self, n_probe: int, seed: int = 1):
        """Pick resident slots and steps per launch by timing short probe runs; returns the table."""
        table = (SpxTuneEntry * 64)()
        n = C.c_int()
        self._check(lib().spx_autotune(self.h, n_probe, seed, table, 64, C.byref(n)), "autotune")
        return [(table[i].lanes, table[i].steps_per_launch, table[i].steps_per_s, table[i].graph_launches) for i in range(min(n.value, 64))]

    # --- distributions ----------------------------------------------------------------------
    def distro(self, d: int):
        """(counts[int64], sums[n_bins x width], n_events) of distribution d."""
        nb3 = (C.c_int32 * 3)()
        w = C.c_int32()
        self._check(lib().spx_distro_info(self.h, d, nb3, C.byref(w)), "distro_info")
        nb = nb3[0] * nb3[1] * nb3[2]
        counts = np.zeros(nb, dtype=np.int64)
        sums = np.zeros(nb * w.value, dtype=np.float64)
        nev = C.c_int64()
        self._check(lib().spx_get_distro(self.h, d, counts.ctypes.data, sums.ctypes.data, C.byref(nev)), "get_distro")
        return counts, (sums.reshape(nb, w.value) if w.value > 1 else sums), nev.value

    def reset_distros(self) -> None:
        self._check(lib().spx_reset_distros(self.h), "reset")

    # --- distribution files (DistributionTemplated::Dump / Restore / Print1D / Print2D) --------------
    def dump_distro(self, d: int, file_name: str) -> None:
        self._check(lib().spx_dump_distro(self.h, d, file_name.encode()), "dump_distro")

    def restore_distro(self, d: int, file_name: str) -> None:
        self._check(lib().spx_restore_distro(self.h, d, file_name.encode()), "restore_distro")

    def print_distro_1d(self, d: int, ijk: int, file_name: str, phys_units: bool = False) -> None:
        self._check(lib().spx_print_distro1d(self.h, d, ijk, file_name.encode(), int(phys_units)), "print_distro_1d")

    def print_distro_2d(self, d: int, ijk1: int, ijk2: int, file_name: str, phys_units: bool = False) -> None:
        self._check(lib().spx_print_distro2d(self.h, d, ijk1, ijk2, file_name.encode(), int(phys_units)), "print_distro_2d")
