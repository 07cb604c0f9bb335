"""Uniform Cartesian grid data of BackgroundServerCartesian.

The grid is stored as the reference's reader expects it (src/reader_cartesian.cc:25-140): a text
header ``<pattern>.info`` with the blocks per axis, the domain corners and the variable count, and a
binary ``<pattern>.out`` of doubles, block by block (block index x fastest), each block 4x4x4 zones
(zone index x fastest) of 6 values: flow velocity and magnetic field in fluid units (length 1 au,
velocity 100 km/s, magnetic field 1e7 * sqrt(m_p) G; common/physics.hh:20-56).
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass

import numpy as np

from . import configs as cf

BLOCK = 4
NVAR = 6

# Fluid units (common/physics.hh:20-56) and their ratios to the code units.
FLUID_LENGTH = 1.495978707e13
FLUID_VELOCITY = 1.0e7
FLUID_MAGNETIC = 1.0e7 * math.sqrt(1.67262192596e-24)
LENGTH_RATIO = cf.unit_length / FLUID_LENGTH        # code length -> fluid length
VELOCITY_RATIO = cf.unit_velocity / FLUID_VELOCITY
MAGNETIC_RATIO = cf.unit_magnetic / FLUID_MAGNETIC


@dataclass
class CartesianGrid:
    nb: tuple            # blocks per axis
    dom_min: tuple       # domain corners, fluid units (au)
    dom_max: tuple
    data: np.ndarray     # float64, the .out order: [kb][jb][ib][kz][jz][iz][var]

    @property
    def n_zones(self) -> tuple:
        return tuple(BLOCK * n for n in self.nb)

    def zone_array(self) -> np.ndarray:
        """The same values as one array [z][y][x][var] over the whole grid."""
        nbx, nby, nbz = self.nb
        a = self.data.reshape(nbz, nby, nbx, BLOCK, BLOCK, BLOCK, NVAR)
        return a.transpose(0, 3, 1, 4, 2, 5, 6).reshape(nbz * BLOCK, nby * BLOCK, nbx * BLOCK, NVAR)

    def zone_centres(self):
        """Cell-centre coordinates per axis in fluid units, as the reference generator places them
        (benchmarks/main_generate_cartesian_solarwind_background.cc:120-140)."""
        out = []
        for a in range(3):
            bl = (self.dom_max[a] - self.dom_min[a]) / self.nb[a]
            zl = bl / BLOCK
            b = np.repeat(np.arange(self.nb[a]), BLOCK)
            z = np.tile(np.arange(BLOCK), self.nb[a])
            out.append(self.dom_min[a] + 0.5 * zl + b * bl + z * zl)
        return out


def from_zone_array(nb, dom_min, dom_max, zones: np.ndarray) -> CartesianGrid:
    """Grid from a [z][y][x][var] zone array (inverse of CartesianGrid.zone_array)."""
    nbx, nby, nbz = nb
    a = np.asarray(zones, dtype=np.float64).reshape(nbz, BLOCK, nby, BLOCK, nbx, BLOCK, NVAR)
    data = np.ascontiguousarray(a.transpose(0, 2, 4, 1, 3, 5, 6)).reshape(-1)
    return CartesianGrid(tuple(int(n) for n in nb), tuple(map(float, dom_min)), tuple(map(float, dom_max)), data)


def from_function(nb, dom_min, dom_max, fields) -> CartesianGrid:
    """Sample ``fields(x, y, z) -> (U, B)`` (fluid units, arrays of shape (3, ...)) at every zone centre."""
    g = CartesianGrid(tuple(nb), tuple(dom_min), tuple(dom_max), np.zeros(0))
    cx, cy, cz = g.zone_centres()
    Z, Y, X = np.meshgrid(cz, cy, cx, indexing="ij")
    U, B = fields(X, Y, Z)
    zones = np.stack([U[0], U[1], U[2], B[0], B[1], B[2]], axis=-1)
    return from_zone_array(nb, dom_min, dom_max, zones)


def linear(nb, dom_min, dom_max, AU, bU, AB, bB) -> CartesianGrid:
    """U = AU x + bU, B = AB x + bB (fluid units): trilinear interpolation reproduces it exactly."""
    AU, bU, AB, bB = (np.asarray(v, dtype=np.float64) for v in (AU, bU, AB, bB))

    def f(X, Y, Z):
        P = np.stack([X, Y, Z])
        return np.einsum("ij,j...->i...", AU, P) + bU[:, None, None, None], np.einsum("ij,j...->i...", AB, P) + bB[:, None, None, None]
    return from_function(nb, dom_min, dom_max, f)


def parker_spiral(nb=25, corner_au=100.0, V_kms=400.0, B_nT_1au=5.0, period_days=25.0, r_ref_rsun=3.0) -> CartesianGrid:
    """The Parker-spiral solar wind sampled at zone centres, as
    benchmarks/main_generate_cartesian_solarwind_background.cc does (25^3 blocks over +-100 au,
    V = 400 km/s, |B| = 5 nT at 1 au, 25-day rotation, r_ref = 3 R_sun), with the spiral of
    BackgroundSolarWind (src/background_solarwind.cc:111-255, every switch off)."""
    ur0 = V_kms * 1.0e5 / cf.unit_velocity
    r_ref = r_ref_rsun * 6.957e10 / cf.unit_length
    B0 = B_nT_1au * 1.0e-5 / cf.unit_magnetic * (cf.one_au / r_ref) ** 2
    w0 = 2.0 * math.pi / (period_days * 86400.0) * cf.unit_time

    def f(X, Y, Z):
        x, y, z = (v / LENGTH_RATIO for v in (X, Y, Z))          # code units
        r = np.sqrt(x * x + y * y + z * z)
        s = np.sqrt(x * x + y * y)
        cost, sint = z / r, s / r
        with np.errstate(invalid="ignore", divide="ignore"):
            cosp = np.where(s > 0, x / s, 0.0)
            sinp = np.where(s > 0, y / s, 0.0)
        Br = B0 * (r_ref / r) ** 2
        Bp = -Br * sint * (r - r_ref) * w0 / ur0
        U = np.stack([ur0 * x / r, ur0 * y / r, ur0 * z / r]) * VELOCITY_RATIO
        B = np.stack([Br * sint * cosp - Bp * sinp, Br * sint * sinp + Bp * cosp, Br * cost]) * MAGNETIC_RATIO
        return U, B
    c = corner_au
    return from_function((nb, nb, nb), (-c, -c, -c), (c, c, c), f)


def write(pattern: str, g: CartesianGrid) -> None:
    """``pattern``.info and ``pattern``.out in the reference's format."""
    with open(pattern + ".info", "w") as f:
        f.write("".join(f"{n:20d}" for n in g.nb) + "\n")
        f.write("".join(f"{v:20.12g}" for v in g.dom_min) + "\n")
        f.write("".join(f"{v:20.12g}" for v in g.dom_max) + "\n")
        f.write(f"{NVAR:20d}\n")
    np.asarray(g.data, dtype="<f8").tofile(pattern + ".out")


def read(pattern: str) -> CartesianGrid:
    with open(pattern + ".info") as f:
        tok = f.read().split()
    nb = tuple(int(t) for t in tok[0:3])
    lo = tuple(float(t) for t in tok[3:6])
    hi = tuple(float(t) for t in tok[6:9])
    nvar = int(tok[9])
    if nvar != NVAR:
        raise ValueError(f"{pattern}.info: {nvar} variables per zone, expected {NVAR}")
    data = np.fromfile(pattern + ".out", dtype="<f8")
    if data.size != NVAR * BLOCK ** 3 * nb[0] * nb[1] * nb[2]:
        raise ValueError(f"{pattern}.out: {data.size} values for {nb} blocks")
    return CartesianGrid(nb, lo, hi, data)


def exists(pattern: str) -> bool:
    return os.path.exists(pattern + ".info") and os.path.exists(pattern + ".out")
