"""Python mirror of the query server (C ABI: spx_server_* in include/spectrum_amd.h).

A served query is a configured simulation (the reference's plugins and options, as in
``Simulation``) plus a trajectory count and a random-stream key.  Queries run concurrently on the
device with per-query counters and histograms; ``Server.result`` returns the same counts and sums
``Simulation.distro`` would for a run of that query alone.  The server replaces the reference's MPI
master/worker request loop (src/simulation.cc:253-305, 743-835).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import configs as cf
from .engine import Simulation, SpxError, lib, ERRORS

STATES = {0: "queued", 1: "running", 2: "done", 3: "failed"}


class ServerOpts(C.Structure):
    _fields_ = [("precision", C.c_int32), ("steps_per_launch", C.c_int32), ("lanes", C.c_int64), ("hbm_budget_bytes", C.c_double),
                ("max_queries", C.c_int32), ("launches_per_poll", C.c_int32)]


class QueryInfo(C.Structure):
    _fields_ = [("state", C.c_int32), ("pad_", C.c_int32), ("n_traj", C.c_uint64), ("done", C.c_uint64), ("failed", C.c_uint64),
                ("steps_done", C.c_uint64), ("steps_executed", C.c_uint64), ("t_submit", C.c_double), ("t_admit", C.c_double),
                ("t_done", C.c_double), ("t_start", C.c_double), ("latency_s", C.c_double), ("footprint_bytes", C.c_int64)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("launches", C.c_uint64), ("steps_executed", C.c_uint64),
                ("busy_s", C.c_double), ("wall_s", C.c_double), ("pools", C.c_int32), ("running", C.c_int32), ("queued", C.c_int32),
                ("pad_", C.c_int32)]


_bound = False


def _bind():
    global _bound
    if _bound:
        return lib()
    L = lib()
    L.spx_server_create.argtypes = [C.POINTER(ServerOpts), C.POINTER(C.c_void_p)]
    L.spx_server_submit.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64, C.POINTER(C.c_int64)]
    L.spx_server_info.argtypes = [C.c_void_p, C.c_int64, C.POINTER(QueryInfo)]
    L.spx_server_wait.argtypes = [C.c_void_p, C.c_int64, C.c_double]
    L.spx_server_result.argtypes = [C.c_void_p, C.c_int64, C.c_int, C.c_void_p, C.c_void_p, C.POINTER(C.c_int64)]
    L.spx_server_release.argtypes = [C.c_void_p, C.c_int64]
    L.spx_server_stats_get.argtypes = [C.c_void_p, C.POINTER(ServerStats)]
    L.spx_server_last_error.argtypes = [C.c_void_p]
    L.spx_server_last_error.restype = C.c_char_p
    L.spx_server_destroy.argtypes = [C.c_void_p]
    L.spx_server_destroy.restype = None
    _bound = True
    return L


class Server:
    """Persistent engine serving many queries at once on device 0."""

    def __init__(self, precision: int = 0, steps_per_launch: int = 500, lanes: int = 0, hbm_budget_bytes: float = 0.0,
                 max_queries: int = 64, launches_per_poll: int = 1):
        L = _bind()
        o = ServerOpts(precision, steps_per_launch, lanes, hbm_budget_bytes, max_queries, launches_per_poll)
        self.h = C.c_void_p()
        rc = L.spx_server_create(C.byref(o), C.byref(self.h))
        if rc != 0:
            raise SpxError(f"server create: {ERRORS.get(rc, rc)}")
        self._shapes = {}

    def _check(self, rc, what):
        if rc != 0:
            msg = lib().spx_server_last_error(self.h).decode() if self.h else ""
            raise SpxError(f"{what}: {ERRORS.get(rc, rc)} {msg}")

    def submit(self, proto, n_traj: int, seed: int = 1, first: int = 0) -> int:
        """Queue a query; `proto` is a configured Simulation or a configs.Setup."""
        sim = proto if isinstance(proto, Simulation) else Simulation.from_setup(proto)
        qid = C.c_int64()
        rc = lib().spx_server_submit(self.h, sim.h, first, n_traj, seed, C.byref(qid))
        if rc != 0:
            raise SpxError(f"submit: {ERRORS.get(rc, rc)} {lib().spx_last_error(sim.h).decode()}")
        shapes = []
        for d in range(sim.n_dist):
            nb3 = (C.c_int32 * 3)()
            w = C.c_int32()
            lib().spx_distro_info(sim.h, d, nb3, C.byref(w))
            shapes.append((nb3[0] * nb3[1] * nb3[2], w.value))
        self._shapes[qid.value] = shapes
        return qid.value

    def info(self, qid: int) -> dict:
        q = QueryInfo()
        self._check(lib().spx_server_info(self.h, qid, C.byref(q)), "info")
        return {"state": STATES[q.state], "n_traj": q.n_traj, "done": q.done, "failed": q.failed, "steps_done": q.steps_done,
                "steps_executed": q.steps_executed, "t_submit": q.t_submit, "t_admit": q.t_admit, "t_start": q.t_start, "t_done": q.t_done,
                "latency_s": q.latency_s, "footprint_bytes": q.footprint_bytes}

    def wait(self, qid: int, timeout: float = -1.0) -> bool:
        rc = lib().spx_server_wait(self.h, qid, timeout)
        if rc == -7:
            return False
        self._check(rc, "wait")
        return True

    def result(self, qid: int, d: int = 0):
        nb, w = self._shapes[qid][d]
        counts = np.zeros(nb, dtype=np.int64)
        sums = np.zeros(nb * w, dtype=np.float64)
        nev = C.c_int64()
        self._check(lib().spx_server_result(self.h, qid, d, counts.ctypes.data, sums.ctypes.data, C.byref(nev)), "result")
        return counts, (sums.reshape(nb, w) if w > 1 else sums), nev.value

    def release(self, qid: int) -> None:
        self._check(lib().spx_server_release(self.h, qid), "release")
        self._shapes.pop(qid, None)

    def stats(self) -> dict:
        s = ServerStats()
        self._check(lib().spx_server_stats_get(self.h, C.byref(s)), "stats")
        return {k: getattr(s, k) for k, _ in ServerStats._fields_ if k != "pad_"}

    def close(self) -> None:
        if self.h:
            lib().spx_server_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def modulation_query(r_au: float, n_energies: int, name: str = "modulation_query") -> cf.Setup:
    """A served modulation-spectrum query: the reference problem observed at radius r_au."""
    return cf.parker_modulation(n_energies=n_energies, r_start_au=r_au, name=name)
