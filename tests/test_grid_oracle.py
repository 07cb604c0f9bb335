"""Pins of the oracle's gridded background (BackgroundServerCartesian, first-order interpolation) by
what the mathematics fixes, independently of its block-by-block restatement:

* a linear field is reproduced exactly by trilinear interpolation, values and gradients;
* every value, |B| (interpolated as a magnitude) and every gradient equal a global trilinear
  interpolation of the zone array written here with numpy, including stencils that cross block faces;
* the grid's dmax is min(dmax0, zone length); positions outside the domain or whose stencil leaves the
  grid fail; the unit factors are the fluid/code unit ratios;
* the .info/.out format round-trips, and the host library reads it;
* the reference benchmark's own containers (ints where the class reads doubles) are accepted.
"""
import math
import os

import numpy as np
import pytest

from spectrum_amd import cartesian as G
from spectrum_amd import configs as cf
from spectrum_amd import engine as E


def numpy_trilinear(g: G.CartesianGrid, pf):
    """Global trilinear interpolation of [U, B, |B|] and its gradient at pf (fluid units)."""
    Z = g.zone_array()
    Bm = np.sqrt(Z[..., 3] ** 2 + Z[..., 4] ** 2 + Z[..., 5] ** 2)
    V = np.concatenate([Z, Bm[..., None]], axis=-1)                  # [z][y][x][7]
    nz = np.array(g.n_zones)
    zl = (np.array(g.dom_max) - np.array(g.dom_min)) / nz
    s = (np.asarray(pf) - np.array(g.dom_min)) / zl - 0.5
    i0 = np.floor(s).astype(int)
    t = s - i0
    if (i0 < 0).any() or (i0 + 1 >= nz).any():
        return None
    val = np.zeros(7)
    grad = np.zeros((3, 7))
    for k in (0, 1):
        for j in (0, 1):
            for i in (0, 1):
                c = V[i0[2] + k, i0[1] + j, i0[0] + i]
                fx = t[0] if i else 1 - t[0]
                fy = t[1] if j else 1 - t[1]
                fz = t[2] if k else 1 - t[2]
                val += fx * fy * fz * c
                grad[0] += (1 if i else -1) * fy * fz / zl[0] * c
                grad[1] += (1 if j else -1) * fx * fz / zl[1] * c
                grad[2] += (1 if k else -1) * fx * fy / zl[2] * c
    return val, grad, zl


@pytest.fixture(scope="module")
def rough_grid():
    rng = np.random.default_rng(4)
    nb = (3, 2, 4)
    nz = tuple(4 * n for n in nb)
    zones = rng.normal(size=(nz[2], nz[1], nz[0], 6))
    zones[..., 3:6] += 3.0
    return G.from_zone_array(nb, (-10.0, -8.0, -12.0), (14.0, 8.0, 20.0), zones)


def oracle_for(oracle_mod, grid, dmax0=cf.one_au):
    s = cf.parker_modulation_cartesian(grid)
    s.plugins[0] = (cf.BACKGROUND, "BackgroundServerCartesian",
                    cf.background_base(cf.DataContainer(), 0.0, (0, 0, 0), (0, 0, 0), (0, 0, 0), dmax0).bytes())
    return oracle_mod.Oracle(s)


def test_linear_field_is_reproduced(oracle_mod):
    rng = np.random.default_rng(1)
    AU, bU, AB = rng.normal(size=(3, 3)), rng.normal(size=3), rng.normal(size=(3, 3))
    bB = rng.normal(size=3) + 20.0
    g = G.linear((3, 2, 4), (-10, -8, -12), (14, 8, 20), AU, bU, AB, bB)
    o = oracle_for(oracle_mod, g)
    for _ in range(500):
        pf = rng.uniform((-8.5, -6.5, -10.5), (12.5, 6.5, 18.5))
        f = o.grid_fields(pf / G.LENGTH_RATIO)
        np.testing.assert_allclose(f["U"], (AU @ pf + bU) / G.VELOCITY_RATIO, rtol=1e-13, atol=1e-13 * np.abs(bU).max())
        np.testing.assert_allclose(f["B"], (AB @ pf + bB) / G.MAGNETIC_RATIO, rtol=1e-13)
        # gradU[row = d/dx_row][col = component]
        np.testing.assert_allclose(f["gradU"], AU.T / G.VELOCITY_RATIO * G.LENGTH_RATIO, rtol=1e-12, atol=1e-12 * np.abs(AU).max())
        np.testing.assert_allclose(f["gradB"], AB.T / G.MAGNETIC_RATIO * G.LENGTH_RATIO, rtol=1e-12, atol=1e-12 * np.abs(AB).max())


def test_values_and_gradients_equal_a_global_trilinear_interpolation(oracle_mod, rough_grid):
    g = rough_grid
    o = oracle_for(oracle_mod, g, dmax0=1e9)
    rng = np.random.default_rng(2)
    zl = (np.array(g.dom_max) - np.array(g.dom_min)) / np.array(g.n_zones)
    # random points, and points on both sides of every block face (stencils across blocks)
    pts = list(rng.uniform(np.array(g.dom_min) + zl, np.array(g.dom_max) - zl, size=(400, 3)))
    for a in range(3):
        bl = (g.dom_max[a] - g.dom_min[a]) / g.nb[a]
        for b in range(1, g.nb[a]):
            for d in (-0.3, -1e-9, 1e-9, 0.3):
                p = rng.uniform(np.array(g.dom_min) + zl, np.array(g.dom_max) - zl)
                p[a] = g.dom_min[a] + b * bl + d * zl[a]
                pts.append(p)
    for pf in pts:
        f = o.grid_fields(pf / G.LENGTH_RATIO)
        ref = numpy_trilinear(g, pf)
        assert (f is None) == (ref is None)
        if f is None:
            continue
        val, grad, _ = ref
        np.testing.assert_allclose(f["U"], val[0:3] / G.VELOCITY_RATIO, rtol=1e-12, atol=1e-13)
        np.testing.assert_allclose(f["B"], val[3:6] / G.MAGNETIC_RATIO, rtol=1e-12, atol=1e-13)
        np.testing.assert_allclose(f["Bmag"], val[6] / G.MAGNETIC_RATIO, rtol=1e-12)
        np.testing.assert_allclose(f["gradU"], grad[:, 0:3] / G.VELOCITY_RATIO * G.LENGTH_RATIO, rtol=1e-10, atol=1e-12)
        np.testing.assert_allclose(f["gradB"], grad[:, 3:6] / G.MAGNETIC_RATIO * G.LENGTH_RATIO, rtol=1e-10, atol=1e-12)
        np.testing.assert_allclose(f["gradBmag"], grad[:, 6] / G.MAGNETIC_RATIO * G.LENGTH_RATIO, rtol=1e-10, atol=1e-12)
        # E is formed in fluid units and converted; in exact arithmetic it is -(U x B) / c in code units
        Ec = -np.cross(f["U"], f["B"]) / cf.c_code
        np.testing.assert_allclose(f["E"], Ec, rtol=1e-12, atol=1e-14 * np.abs(Ec).max())
        assert f["dmax"] == pytest.approx(zl.min() / G.LENGTH_RATIO, rel=1e-14)


def test_domain_and_stencil_limits(oracle_mod, rough_grid):
    g = rough_grid
    o = oracle_for(oracle_mod, g)
    zl = (np.array(g.dom_max) - np.array(g.dom_min)) / np.array(g.n_zones)
    centre = 0.5 * (np.array(g.dom_min) + np.array(g.dom_max))
    for a in range(3):
        for side, inside in ((g.dom_min[a], 0.51), (g.dom_max[a], -0.51)):
            p = centre.copy()
            p[a] = side + inside * zl[a]                    # stencil within the grid
            assert o.grid_fields(p / G.LENGTH_RATIO) is not None
            p[a] = side + 0.49 * np.sign(inside) * zl[a]    # within half a zone of the face: stencil leaves the grid
            assert o.grid_fields(p / G.LENGTH_RATIO) is None
            p[a] = side - 2.0 * np.sign(inside) * zl[a]     # outside the domain
            assert o.grid_fields(p / G.LENGTH_RATIO) is None
    f = o.grid_fields(centre / G.LENGTH_RATIO)
    assert f["dmax"] == pytest.approx(min(cf.one_au, zl.min() / G.LENGTH_RATIO), rel=1e-14)


def test_unit_factors():
    mp = 1.67262192596e-24
    assert G.LENGTH_RATIO == pytest.approx(1.0e10 / 1.495978707e13, rel=1e-15)
    assert G.VELOCITY_RATIO == pytest.approx(0.1 * 2.99792458e10 / 1.0e7, rel=1e-15)
    # code magnetic unit m_p (0.1 c)^2 / (e * 1e10 cm) over the fluid unit 1e7 sqrt(m_p)
    assert G.MAGNETIC_RATIO == pytest.approx(mp * (0.1 * 2.99792458e10) ** 2 / (4.8032044e-10 * 1.0e10) / (1.0e7 * math.sqrt(mp)), rel=1e-14)
    # 400 km/s is 4 fluid units; 5 nT at 1 au in the generated spiral
    g = G.parker_spiral(nb=5, corner_au=10.0)
    cx, cy, cz = g.zone_centres()
    Z = g.zone_array()
    i = np.argmin(np.abs(cx - 1.0)); j = np.argmin(np.abs(cy)); k = np.argmin(np.abs(cz))
    r = math.sqrt(cx[i] ** 2 + cy[j] ** 2 + cz[k] ** 2)
    assert np.linalg.norm(Z[k, j, i, 0:3]) == pytest.approx(4.0, rel=1e-12)
    Bnt = np.linalg.norm(Z[k, j, i, 3:6]) * 1.0e7 * math.sqrt(mp) * 1e5
    assert Bnt * r ** 2 < 5.0 * 1.5 and Bnt > 0.5


def test_file_format_round_trip(tmp_path, rough_grid):
    pat = str(tmp_path / "grid_3_2_4")
    G.write(pat, rough_grid)
    back = G.read(pat)
    assert back.nb == rough_grid.nb and back.dom_min == rough_grid.dom_min and back.dom_max == rough_grid.dom_max
    np.testing.assert_array_equal(back.data, rough_grid.data)
    # the host library parses the same files (lowering only; no device needed)
    sim = E.Simulation(cf.TRAJ_PARKER, cf.BACKWARD)
    sim.add_background("BackgroundServerCartesian", cf.background_base(cf.DataContainer(), 0.0, (0, 0, 0), (0, 0, 0), (0, 0, 0), 1.0), pat)
    with pytest.raises(E.SpxError):
        sim.add_background("BackgroundServerCartesian", cf.background_base(cf.DataContainer(), 0.0, (0, 0, 0), (0, 0, 0), (0, 0, 0), 1.0),
                           str(tmp_path / "missing"))
    os.remove(pat + ".out")
    with pytest.raises(FileNotFoundError):
        G.read(pat)


def test_benchmark_containers_with_int_records_are_accepted(oracle_mod):
    s = cf.parker_modulation_cartesian(G.parker_spiral(nb=3, corner_au=6.0))
    E.Simulation.from_setup(s)                 # engine lowering
    o = oracle_mod.Oracle(s)                   # oracle lowering
    _, rec = o.run(16, seed=3, threads=4)
    assert rec.shape == (16,)
    # a record longer than the value (a vector where the class reads a double) is still an error
    sim = E.Simulation(cf.TRAJ_PARKER, cf.BACKWARD)
    with pytest.raises(E.SpxError):
        sim.add_plugin(cf.DIFFUSION, "DiffusionParaConstant", cf.DataContainer().insert_vec((1.0, 2.0, 3.0)))


def test_grid_needs_data_and_parker_transport():
    s = cf.parker_modulation_cartesian(None)
    sim = E.Simulation.from_setup(s)
    with pytest.raises(E.SpxError, match="grid"):
        sim.run(1, seed=1)


def test_perturbed_oracle_build_measures_conditioning(oracle_mod):
    """The contracted build differs from the oracle only by rounding: identical steps and statuses, and
    endpoint differences at rounding level on the well-conditioned isotropic case."""
    import test_gpu_parity as P
    s = cf.parker_modulation_cartesian(G.parker_spiral(nb=25), kap_rat=1.0)
    _, a = oracle_mod.Oracle(s).run(128, seed=5, threads=8)
    _, b = oracle_mod.Oracle(s, oracle_mod.FMA_LIB_PATH).run(128, seed=5, threads=8)
    np.testing.assert_array_equal(a["steps"], b["steps"])
    np.testing.assert_array_equal(a["status"], b["status"])
    r = P.sensitivity(a, b)
    assert r.max() < 1e-11 and r.max() > 0.0
