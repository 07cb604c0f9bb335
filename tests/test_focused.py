"""TrajectoryFocused in the oracle (src/trajectory_focused.cc), checked against what the physics fixes:

* a uniform field carried by a uniform flow: no gradients, so p and mu never change and the centre moves
  at U + v mu b: pos - pos0 = (U + v mu b) (t - t0);
* one step in the solar wind and across a moving smooth shock against the focused transport equation in its textbook form (Skilling 1971;
  le Roux & Webb 2009), with the gradients taken here by Cartesian central differences of the field and
  the guiding-centre drift written with the curvature (b . grad) b and b . curl b rather than B . grad B
  and B . curl B:
    dp/dt  = p [ (1 - 3 mu^2)/2 bb:grad U - (1 - mu^2)/2 div U - mu b . (U . grad) U / v ]
    dmu/dt = (1 - mu^2)/2 [ v div b + mu div U - 3 mu bb:grad U - 2 b . (U . grad) U / v ]
    dx/dt  = U + v mu b + r_L v [ (1 - mu^2)/2 (b x grad|B|)/|B| + mu^2 b x (b . grad) b + (1 - mu^2)/2 b (b . curl b) ]
  The oracle differentiates in the field-aligned frame at the Larmor radius, so the two agree to the
  truncation of that stencil, (r_L / r)^2 ~ 1e-6;
* a reflecting wall reverses mu (TrajectoryFocused::ReverseMomentum), and mu stays within [-1, 1].
"""
import numpy as np
import pytest

import feature_setups as F
from spectrum_amd import configs as cf
from spectrum_amd import engine as E
from spectrum_amd.data_container import DataContainer as DC

THREADS = 8


def setup(background, pos0, initial, boundaries, flow=cf.FORWARD):
    s = cf.Setup("focused", cf.TRAJ_FOCUSED, flow)
    s.add(cf.BACKGROUND, *background)
    s.add(cf.INITIAL, "InitialTimeFixed", DC().insert_double(0.0))
    s.add(cf.INITIAL, "InitialSpaceFixed", DC().insert_vec(pos0))
    s.add(cf.INITIAL, *initial)
    for b in boundaries:
        s.add(cf.BOUNDARY, *b)
    d = F.dist((4, 0, 0), (-1.0e9, 0.0, 0.0), (1.0e9, 0.0, 0.0), (0, 0, 0), (1, 0, 0))
    d.insert_double(1.0).insert_double(0.0).insert_int(1)
    s.add(cf.DISTRIBUTION, "DistributionTimeUniform", d)
    return s


@pytest.mark.parametrize("flow", [cf.FORWARD, cf.BACKWARD])
def test_uniform_field_drift(oracle_mod, flow):
    B = np.array([3.0, 4.0, 12.0])
    U = np.array([0.02, 0.01, -0.03])
    bg = ("BackgroundUniform", cf.background_base(DC(), 0.0, (0.0, 0.0, 0.0), tuple(U), tuple(B), 2.0))
    T = 300.0 if flow == cf.FORWARD else -300.0
    s = setup(bg, (1.0, 2.0, 3.0), ("InitialMomentumThickShell", DC().insert_double(cf.mom_from_ekin(cf.MeV)).insert_double(cf.mom_from_ekin(100.0 * cf.MeV)).insert_bool(True)),
              [("BoundaryTimeExpire", cf.boundary(1, [0], ("double", T)))], flow)
    n = 32
    _, rec = oracle_mod.Oracle(s).run(n, seed=4, threads=THREADS)
    assert (rec["status"] == 0).all() and (rec["steps"] > 10).all()
    np.testing.assert_array_equal(rec["mom"], rec["mom0"])
    p, mu = rec["mom0"][:, 0], rec["mom0"][:, 1]
    assert (np.abs(mu) <= 1.0).all() and np.std(mu) > 0.3
    v = np.array([cf.vel_from_mom(x) for x in p])
    want = (rec["t"] - rec["t0"])[:, None] * (U[None, :] + (v * mu)[:, None] * (B / 13.0)[None, :])
    np.testing.assert_allclose(rec["pos"] - rec["pos0"], want, rtol=1e-11, atol=1e-11 * np.abs(want).max())


def fields(o, x, t=0.0):
    U, Bv, _, _ = o.background_fields(t, x, 1.0)
    return np.array(U), np.array(Bv)


def smooth_shock():
    """A smooth shock moving at 0.002 along -x, oblique field (BackgroundSmoothShock)."""
    dc = cf.background_base(DC(), 0.0, (0.0, 0.0, 0.0), (0.013343, 0.0, 0.0), (8.0, 6.0, 0.0), 5.0)
    dc.insert_vec((-1.0, 0.0, 0.0)).insert_double(0.002).insert_double(4.0).insert_double(2.5).insert_double(0.1)
    return ("BackgroundSmoothShock", dc)


def jacobians(o, x, h):
    """J[i][j] = d X_j / d x_i by central differences along the Cartesian axes."""
    JU = np.zeros((3, 3)); Jb = np.zeros((3, 3)); gB = np.zeros(3)
    for i in range(3):
        e = np.zeros(3); e[i] = h
        Uf, Bf = fields(o, x + e)
        Ub, Bb = fields(o, x - e)
        JU[i] = (Uf - Ub) / (2 * h)
        Jb[i] = (Bf / np.linalg.norm(Bf) - Bb / np.linalg.norm(Bb)) / (2 * h)
        gB[i] = (np.linalg.norm(Bf) - np.linalg.norm(Bb)) / (2 * h)
    return JU, Jb, gB


@pytest.mark.parametrize("case", ["solarwind", "smoothshock"])
def test_one_step_against_the_focused_transport_equation(oracle_mod, case):
    """In the solar wind the oracle's gradients are its field-aligned stencil; across the smooth shock they
    are the background's analytic ones, and the moving shock adds dU/dt, here a central difference in t."""
    if case == "solarwind":
        bg, x0, t_end, h = ("BackgroundSolarWind", cf.solarwind()), np.array([0.8, 0.5, 0.2]) * cf.one_au, 2.0, 1e-3
    else:
        bg, x0, t_end, h = smooth_shock(), np.array([-0.2, 0.3, 0.1]), 0.01, 1e-4
    s = setup(bg, tuple(x0),
              ("InitialMomentumThickShell", DC().insert_double(cf.mom_from_ekin(1.0 * cf.MeV)).insert_double(cf.mom_from_ekin(100.0 * cf.MeV)).insert_bool(True)),
              [("BoundaryTimeExpire", cf.boundary(1, [0], ("double", t_end)))])
    o = oracle_mod.Oracle(s)
    n = 24
    _, rec = o.run(n, seed=9, threads=THREADS)
    assert (rec["status"] == 0).all() and (rec["steps"] == 1).all()
    U, Bv = fields(o, x0)
    Bm = np.linalg.norm(Bv)
    b = Bv / Bm
    JU, Jb, gB = jacobians(o, x0, h)
    divU = np.trace(JU)
    bbgU = b @ JU @ b
    dUdt = (fields(o, x0, 1e-4)[0] - fields(o, x0, -1e-4)[0]) / 2e-4
    UgU = dUdt + U @ JU                     # dU/dt + (U . grad) U
    divb = np.trace(Jb)
    kappa = b @ Jb                          # (b . grad) b
    curlb = np.array([Jb[1, 2] - Jb[2, 1], Jb[2, 0] - Jb[0, 2], Jb[0, 1] - Jb[1, 0]])
    for k in range(n):
        p, mu = rec["mom0"][k, :2]
        dt = rec["t"][k] - rec["t0"][k]
        v = cf.vel_from_mom(p)
        rL = p * cf.c_code / (cf.charge * Bm)
        dpdt = p * (0.5 * (1 - 3 * mu ** 2) * bbgU - 0.5 * (1 - mu ** 2) * divU - mu * (b @ UgU) / v)
        dmudt = 0.5 * (1 - mu ** 2) * (v * divb + mu * divU - 3 * mu * bbgU - 2 * (b @ UgU) / v)
        vd = rL * v * (0.5 * (1 - mu ** 2) * np.cross(b, gB) / Bm + mu ** 2 * np.cross(b, kappa) + 0.5 * (1 - mu ** 2) * b * (b @ curlb))
        got_p = (rec["mom"][k, 0] - p) / dt
        got_mu = (rec["mom"][k, 1] - mu) / dt
        got_vd = (rec["pos"][k] - rec["pos0"][k]) / dt - (U + v * mu * b)
        scale_p = p * abs(divU)
        assert abs(got_p - dpdt) < 1e-5 * scale_p, (k, got_p, dpdt)
        scale_mu = v * abs(divb) + abs(divU)
        assert abs(got_mu - dmudt) < 1e-5 * scale_mu, (k, got_mu, dmudt)
        assert np.linalg.norm(got_vd - vd) < 1e-4 * rL * v * np.linalg.norm(np.cross(b, gB)) / Bm, (k, got_vd, vd)


def test_pitch_bounds(oracle_mod):
    _, rec = oracle_mod.Oracle(F.focused_spiral()).run(64, seed=5, threads=THREADS)
    assert (rec["status"] == 0).all()
    assert (np.abs(rec["mom"][:, 1]) <= 1.0).all()


def test_reflection_reverses_mu(oracle_mod):
    """One step across a reflecting plane set 0.005 ahead along b: the endpoint without the wall mirrored in
    the plane (TrajectoryBase::HandleBoundaries), mu reversed and p untouched."""
    x0 = np.array([0.8, 0.5, 0.2]) * cf.one_au
    sw = ("BackgroundSolarWind", cf.solarwind())
    time = ("BoundaryTimeExpire", cf.boundary(1, [0], ("double", 0.01)))
    init = ("InitialMomentumFixed", DC().insert_vec((cf.mom_from_ekin(10.0 * cf.MeV), 0.9, 0.0)))
    o = oracle_mod.Oracle(setup(sw, tuple(x0), init, [time]))
    _, Bv = fields(o, x0)
    b = Bv / np.linalg.norm(Bv)
    wall = ("BoundaryPlaneReflect", cf.boundary(-1, [-1], ("vec", tuple(x0 + 0.005 * b)), ("vec", tuple(b))))
    _, r1 = oracle_mod.Oracle(setup(sw, tuple(x0), init, [wall, time])).run(1, seed=1, threads=1)
    _, r2 = o.run(1, seed=1, threads=1)
    assert r1["steps"][0] == 1 and r2["steps"][0] == 1
    assert (r2["pos"][0] - x0 - 0.005 * b) @ b > 0.0          # the step crossed the wall
    o_w = x0 + 0.005 * b
    np.testing.assert_allclose(r1["pos"][0], r2["pos"][0] - 2.0 * ((r2["pos"][0] - o_w) @ b) * b, rtol=1e-14)
    assert r1["mom"][0, 0] == r2["mom"][0, 0]
    assert r1["mom"][0, 1] == -r2["mom"][0, 1] and abs(r2["mom"][0, 1]) < 1.0


def test_focused_host_checks():
    dc = cf.background_base(DC(), 0.0, (0.0, 0.0, 0.0), (0.01, 0.0, 0.0), (0.0, 3.0, 0.0), 10.0)
    dc.insert_vec((-1.0, 0.0, 0.0)).insert_double(0.0).insert_double(4.0)
    s = setup(("BackgroundShock", dc), (0.0, 0.0, 0.0), ("InitialMomentumShell", DC().insert_double(1.0)),
              [("BoundaryTimeExpire", cf.boundary(1, [0], ("double", 10.0)))])
    with pytest.raises(E.SpxError, match="device error"):      # the discontinuous shock is accepted: past the checks
        E.Simulation.from_setup(s).run(4)
    bg = ("BackgroundUniform", cf.background_base(DC(), 0.0, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 1.0))
    s = setup(bg, (0.0, 0.0, 0.0), ("InitialMomentumShell", DC().insert_double(1.0)), [("BoundaryTimeExpire", cf.boundary(1, [0], ("double", 10.0)))])
    with pytest.raises(E.SpxError, match="device error"):      # no diffusion model needed: past the checks
        E.Simulation.from_setup(s).run(4)
