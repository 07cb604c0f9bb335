"""TrajectoryLorentz in the oracle (src/trajectory_lorentz.cc), checked against what the Euler map fixes.

In a uniform field B along z with E = 0 every step has Omega dt = 2 pi / 100 exactly (PhysicalStep, :51-57),
and the Euler step p' = p + dt q v x B / c multiplies p_perp by sqrt(1 + (2 pi / 100)^2) and turns it by
-atan(2 pi / 100), leaving p_z alone: after n steps |p_perp| = |p_perp,0| (1 + (2 pi/100)^2)^(n/2).  The
isotropic initial direction (mu = 2u - 1, then phi = 2 pi u') is checked with Kolmogorov-Smirnov tests,
and a reflecting wall reverses the momentum (TrajectoryBase::ReverseMomentum).
"""
import numpy as np
import pytest

import feature_setups as F
from spectrum_amd import configs as cf
from spectrum_amd import engine as E
from spectrum_amd.data_container import DataContainer as DC

THREADS = 8
THETA = 2.0 * np.pi / 100.0


def uniform_lorentz(initial, boundaries, flow=cf.FORWARD):
    s = cf.Setup("lorentz_uniform", cf.TRAJ_LORENTZ, flow)
    s.add(cf.BACKGROUND, "BackgroundUniform", cf.background_base(DC(), 0.0, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, 3.0), 1.0e6))
    s.add(cf.INITIAL, "InitialTimeFixed", DC().insert_double(0.0))
    s.add(cf.INITIAL, "InitialSpaceFixed", DC().insert_vec((0.0, 0.0, 0.0)))
    s.add(cf.INITIAL, *initial)
    for b in boundaries:
        s.add(cf.BOUNDARY, *b)
    d = F.dist((4, 0, 0), (0.0, 0.0, 0.0), (1.0e9, 0.0, 0.0), (0, 0, 0), (1, 0, 0))
    d.insert_double(1.0).insert_double(0.0).insert_int(1)
    s.add(cf.DISTRIBUTION, "DistributionTimeUniform", d)
    return s


def test_euler_gyration_closed_form(oracle_mod):
    p = cf.mom_from_ekin(10.0 * cf.MeV)
    m0 = (0.6 * p, 0.0, 0.8 * p)
    s = uniform_lorentz(("InitialMomentumFixed", DC().insert_vec(m0)),
                        [("BoundaryPlanePass", cf.boundary(1, [0], ("vec", (0.0, 0.0, 50.0)), ("vec", (0.0, 0.0, 1.0)))),
                         ("BoundaryTimeExpire", cf.boundary(1, [0], ("double", 1.0e9)))])
    o = oracle_mod.Oracle(s)
    _, rec = o.run(1, normals=np.zeros((1, 1)), uniforms=np.full((1, 1), 0.5), threads=1)
    assert rec["status"][0] == 0
    n = int(rec["steps"][0])
    m = rec["mom"][0]
    assert n > 50
    assert m[2] == m0[2]
    pperp = np.hypot(m[0], m[1])
    np.testing.assert_allclose(pperp, 0.6 * p * (1.0 + THETA ** 2) ** (n / 2.0), rtol=1e-11)
    phase = np.arctan2(m[1], m[0])
    want = -n * np.arctan(THETA)
    assert abs(np.angle(np.exp(1j * (phase - want)))) < 1e-9
    assert rec["pos"][0][2] >= 50.0


def test_isotropic_start_and_reflection(oracle_mod):
    p = cf.mom_from_ekin(10.0 * cf.MeV)
    s = uniform_lorentz(("InitialMomentumShell", DC().insert_double(p)),
                        [("BoundaryTimeExpire", cf.boundary(1, [0], ("double", 1.0e-3)))])
    n = 4000
    _, rec = oracle_mod.Oracle(s).run(n, seed=12, threads=THREADS)
    m = rec["mom0"]
    np.testing.assert_allclose(np.linalg.norm(m, axis=1), p, rtol=1e-14)
    mu = np.sort(m[:, 2] / p)
    phi = np.sort(np.mod(np.arctan2(m[:, 1], m[:, 0]), 2.0 * np.pi))
    ks_mu = np.max(np.abs((mu + 1.0) / 2.0 - np.arange(1, n + 1) / n))
    ks_phi = np.max(np.abs(phi / (2.0 * np.pi) - np.arange(1, n + 1) / n))
    assert ks_mu < 1.63 / np.sqrt(n) and ks_phi < 1.63 / np.sqrt(n)
    # a reflecting wall just above the start: every trajectory that reaches it comes back with -p
    s = uniform_lorentz(("InitialMomentumFixed", DC().insert_vec((0.0, 0.0, p))),
                        [("BoundaryPlaneReflect", cf.boundary(-1, [-1], ("vec", (0.0, 0.0, 5.0)), ("vec", (0.0, 0.0, 1.0)))),
                         ("BoundaryPlanePass", cf.boundary(1, [0], ("vec", (0.0, 0.0, -5.0)), ("vec", (0.0, 0.0, 1.0)))),
                         ("BoundaryTimeExpire", cf.boundary(1, [0], ("double", 1.0e9)))])
    _, rec = oracle_mod.Oracle(s).run(2, seed=1, threads=1)
    assert (rec["status"] == 0).all()
    np.testing.assert_array_equal(rec["mom"], -rec["mom0"])
    assert (rec["pos"][:, 2] <= -5.0).all()


def test_lorentz_host_checks():
    p = cf.mom_from_ekin(10.0 * cf.MeV)
    s = uniform_lorentz(("InitialMomentumShell", DC().insert_double(p)), [("BoundaryTimeExpire", cf.boundary(1, [0], ("double", 1.0)))])
    with pytest.raises(E.SpxError, match="device error"):      # no diffusion model needed: past the checks
        E.Simulation.from_setup(s).run(4)
    s.plugins = [x for x in s.plugins if x[0] != cf.BACKGROUND]
    dc = cf.background_base(DC(), 0.0, (0.0, 0.0, 0.0), (0.01, 0.0, 0.0), (0.0, 3.0, 0.0), 10.0)
    dc.insert_vec((-1.0, 0.0, 0.0)).insert_double(0.0).insert_double(4.0).insert_double(1.0).insert_double(0.1)
    s.add(cf.BACKGROUND, "BackgroundSmoothShock", dc)
    with pytest.raises(E.SpxError, match="SolarWind and Uniform"):
        E.Simulation.from_setup(s).run(4)
