"""The three guiding-centre siblings of GuidingDiffScatt in the oracle (TrajectoryGuiding,
TrajectoryGuidingScatt, TrajectoryGuidingDiff; src/trajectory_guiding*.cc), checked against what the
mathematics fixes in a uniform field along z with no flow (no drifts, E = 0):

* Guiding: the centre moves along B at v_par exactly, the momentum does not change, and the
  transport draws no normal random numbers;
* GuidingScatt: x and y never change; <mu(t)> = mu0 exp(-2 D t) for D_mumu = D (1 - mu^2), and
  <z - z0> = v mu0 (1 - exp(-2 D t)) / (2 D);
* GuidingDiff: mu and |p| never change, z - z0 = v_par t, <(x - x0)^2 + (y - y0)^2> = 4 D_perp t, and
  two normal draws per step.

A plausible slip (a missing pitch-angle drift term -2 D mu, a perpendicular step where there is none,
a draw the transport does not make) fails one of these.
"""
import numpy as np
import pytest

import feature_setups as F
from spectrum_amd import configs as cf
from spectrum_amd.data_container import DataContainer as DC

THREADS = 8
B_Z = 3.0
T_END = 300.0


THETA0 = 0.6                                    # InitialMomentumRing's pitch angle (radians)
MU0 = np.cos(THETA0)


def uniform_setup(traj_type, diffusion=None):
    s = cf.Setup("uniform_gc", traj_type, cf.FORWARD)
    s.add(cf.BACKGROUND, "BackgroundUniform", cf.background_base(DC(), 0.0, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, B_Z), 50.0))
    if diffusion is not None:
        s.add(cf.DIFFUSION, *diffusion)
    s.add(cf.INITIAL, "InitialTimeFixed", DC().insert_double(0.0))
    s.add(cf.INITIAL, "InitialSpaceFixed", DC().insert_vec((1.0, 2.0, 0.0)))
    s.add(cf.INITIAL, "InitialMomentumRing", DC().insert_double(cf.mom_from_ekin(10.0 * cf.MeV)).insert_double(THETA0))
    s.add(cf.BOUNDARY, "BoundaryTimeExpire", cf.boundary(1, [0], ("double", T_END)))
    d = F.dist((16, 16, 0), (-500.0, -1.0, 0.0), (500.0, 1.0, 1.0), (0, 0, 0), (1, 0, 0))
    d.insert_double(1.0).insert_double(0.0).insert_int(1)
    s.add(cf.DISTRIBUTION, "DistributionPositionPitchUniform", d)
    return s


def speed():
    return cf.vel_from_mom(cf.mom_from_ekin(10.0 * cf.MeV))


def test_guiding_moves_along_b_and_draws_nothing(oracle_mod):
    o = oracle_mod.Oracle(uniform_setup(cf.TRAJ_GUIDING))
    # replayed streams with a single normal: a transport that drew normals would exhaust them
    n = 16
    _, rec = o.run(n, normals=np.zeros((n, 1)), uniforms=np.full((n, 2), 0.5), threads=THREADS)
    assert (rec["status"] == 0).all()
    np.testing.assert_array_equal(rec["pos"][:, :2], rec["pos0"][:, :2])
    np.testing.assert_array_equal(rec["mom"], rec["mom0"])
    v_par = speed() * MU0
    np.testing.assert_allclose(rec["pos"][:, 2] - rec["pos0"][:, 2], v_par * (rec["t"] - rec["t0"]), rtol=1e-12)
    np.testing.assert_allclose(rec["t"], T_END, rtol=1e-7)


def test_guiding_scatt_relaxes_mu(oracle_mod):
    D = 2.0e-3
    o = oracle_mod.Oracle(uniform_setup(cf.TRAJ_GUIDING_SCATT, ("DiffusionIsotropicConstant", DC().insert_double(D))))
    n = 4000
    _, rec = o.run(n, seed=9, threads=THREADS)
    assert (rec["status"] == 0).all()
    np.testing.assert_array_equal(rec["pos"][:, :2], rec["pos0"][:, :2])
    p = np.linalg.norm(rec["mom"], axis=1)
    mu = rec["mom"][:, 2] / p
    t = rec["t"] - rec["t0"]
    np.testing.assert_allclose(t, T_END, rtol=1e-7)
    want = MU0 * np.exp(-2.0 * D * T_END)
    sig = mu.std() / np.sqrt(n)
    print(f"[scatt] <mu> {mu.mean():.4f} +- {sig:.4f} vs {want:.4f}")
    assert abs(mu.mean() - want) < 5.0 * sig + 0.01       # Euler steps: O(dt) bias, dt D << 1
    dz = rec["pos"][:, 2] - rec["pos0"][:, 2]
    want_z = speed() * MU0 * (1.0 - np.exp(-2.0 * D * T_END)) / (2.0 * D)
    assert abs(dz.mean() - want_z) < 5.0 * dz.std() / np.sqrt(n) + 0.02 * want_z


def test_guiding_diff_spreads_across_b_only(oracle_mod):
    Dp = 0.5
    o = oracle_mod.Oracle(uniform_setup(cf.TRAJ_GUIDING_DIFF, ("DiffusionPerpConstant", DC().insert_double(Dp))))
    n = 4000
    _, rec = o.run(n, seed=4, threads=THREADS)
    assert (rec["status"] == 0).all()
    np.testing.assert_array_equal(rec["mom"], rec["mom0"])
    t = rec["t"] - rec["t0"]
    np.testing.assert_allclose(rec["pos"][:, 2] - rec["pos0"][:, 2], speed() * MU0 * t, rtol=1e-12)
    r2 = np.sum((rec["pos"][:, :2] - rec["pos0"][:, :2]) ** 2, axis=1)
    want = 4.0 * Dp * T_END
    sig = r2.std() / np.sqrt(n)
    print(f"[diff] <r_perp^2> {r2.mean():.2f} +- {sig:.2f} vs {want:.2f}")
    assert abs(r2.mean() - want) < 5.0 * sig
    # two normals per step: a stream two draws short of the step count runs out
    _, one = o.run(1, seed=4, threads=1)
    steps = int(one["steps"][0])
    _, short = o.run(1, normals=np.zeros((1, 2 * steps - 1)), uniforms=np.full((1, 2), 0.5), threads=1)
    assert short["status"][0] != 0
    _, enough = o.run(1, normals=np.zeros((1, 2 * steps)), uniforms=np.full((1, 2), 0.5), threads=1)
    assert enough["status"][0] == 0


def test_guiding_without_diffusion_model_is_ready(oracle_mod):
    from spectrum_amd import engine as E
    sim = E.Simulation.from_setup(uniform_setup(cf.TRAJ_GUIDING))
    with pytest.raises(E.SpxError, match="device error"):     # past the readiness checks, to the device
        sim.run(4, seed=1)
    sim = E.Simulation.from_setup(uniform_setup(cf.TRAJ_GUIDING_DIFF))
    with pytest.raises(E.SpxError, match="no diffusion model"):
        sim.run(4, seed=1)
