"""TrajectoryFieldline in the oracle (src/trajectory_fieldline.cc), checked against what the geometry fixes:

* a uniform field: the centre moves along bhat at v_par, pos - pos0 = v_par (t - t0) bhat, the momentum
  (0, 0, p_par) never changes and the tracer draws no random numbers;
* the Parker spiral in the ecliptic: z stays exactly 0 (B_z = 0 there), and the traced line follows
  phi(r) - phi(r0) = -(Omega / V) [(r - r0) - r_ref ln(r / r0)] to the Euler tracer's accuracy;
* a reflecting plane reverses the momentum (TrajectoryBase::ReverseMomentum, src/trajectory_base.cc:139-142).
"""
import numpy as np
import pytest

import feature_setups as F
from spectrum_amd import configs as cf
from spectrum_amd import engine as E
from spectrum_amd.data_container import DataContainer as DC

THREADS = 8


def setup(background, pos0, boundaries, flow=cf.FORWARD, p=None):
    s = cf.Setup("fieldline", cf.TRAJ_FIELDLINE, flow)
    s.add(cf.BACKGROUND, *background)
    s.add(cf.INITIAL, "InitialTimeFixed", DC().insert_double(0.0))
    s.add(cf.INITIAL, "InitialSpaceFixed", DC().insert_vec(pos0))
    s.add(cf.INITIAL, "InitialMomentumShell", DC().insert_double(p if p is not None else cf.mom_from_ekin(10.0 * cf.MeV)))
    for b in boundaries:
        s.add(cf.BOUNDARY, *b)
    d = F.dist((4, 0, 0), (0.0, 0.0, 0.0), (1.0e6, 0.0, 0.0), (0, 0, 0), (1, 0, 0))
    d.insert_double(1.0).insert_double(0.0).insert_int(1)
    s.add(cf.DISTRIBUTION, "DistributionTimeUniform", d)
    return s


@pytest.mark.parametrize("flow", [cf.FORWARD, cf.BACKWARD])
def test_uniform_field_straight_line(oracle_mod, flow):
    B = np.array([3.0, 4.0, 12.0])
    bg = ("BackgroundUniform", cf.background_base(DC(), 0.0, (0.0, 0.0, 0.0), (0.01, 0.0, 0.0), tuple(B), 7.0))
    T = 500.0 if flow == cf.FORWARD else -500.0
    s = setup(bg, (1.0, 2.0, 3.0), [("BoundaryTimeExpire", cf.boundary(1, [0], ("double", T)))], flow)
    o = oracle_mod.Oracle(s)
    n = 8
    _, rec = o.run(n, normals=np.zeros((n, 1)), uniforms=np.full((n, 1), 0.5), threads=THREADS)
    assert (rec["status"] == 0).all()
    p = cf.mom_from_ekin(10.0 * cf.MeV)
    np.testing.assert_array_equal(rec["mom"], np.tile([0.0, 0.0, p], (n, 1)))
    v = cf.vel_from_mom(p)
    want = (v * (rec["t"] - rec["t0"]))[:, None] * (B / 13.0)
    np.testing.assert_allclose(rec["pos"] - rec["pos0"], want, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(rec["t"], T, rtol=1e-7)


def test_ecliptic_spiral(oracle_mod):
    bg = ("BackgroundSolarWind", cf.solarwind(dmax0=0.01 * cf.one_au))      # short Euler steps along the line
    s = setup(bg, (cf.one_au, 0.0, 0.0), [("BoundarySphereAbsorb", cf.boundary(1, [0], ("vec", (0.0, 0.0, 0.0)), ("double", 3.0 * cf.one_au))),
                                         ("BoundaryTimeExpire", cf.boundary(1, [0], ("double", 1.0e9)))])
    o = oracle_mod.Oracle(s)
    _, rec = o.run(1, normals=np.zeros((1, 1)), uniforms=np.full((1, 1), 0.5), threads=1)
    assert rec["status"][0] == 0
    x = rec["pos"][0]
    assert x[2] == 0.0
    r = np.linalg.norm(x)
    assert r == pytest.approx(3.0 * cf.one_au, rel=0.01)
    V = 400.0e5 / cf.unit_velocity
    Om = 2.87e-6 * cf.unit_time
    r_ref = 3.0 * 6.957e10 / cf.unit_length
    r0 = cf.one_au
    want = -(Om / V) * ((r - r0) - r_ref * np.log(r / r0))
    got = np.arctan2(x[1], x[0])
    got -= 2.0 * np.pi * np.round((got - want) / (2.0 * np.pi))           # the line winds more than half a turn
    print(f"[fieldline] phi {got:.4f} vs spiral {want:.4f} after {int(rec['steps'][0])} steps")
    assert got == pytest.approx(want, rel=0.01)


def test_reflection_reverses_the_momentum(oracle_mod):
    _, rec = oracle_mod.Oracle(F.fieldline_spiral()).run(64, seed=3, threads=THREADS)
    assert (rec["status"] == 0).all()
    np.testing.assert_array_equal(rec["mom"], -rec["mom0"])
    assert (np.linalg.norm(rec["pos"], axis=1) < 0.5 * cf.one_au).all()    # ran back inward to the inner sphere


def test_fieldline_host_checks():
    bg = ("BackgroundUniform", cf.background_base(DC(), 0.0, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 1.0))
    s = setup(bg, (0.0, 0.0, 0.0), [("BoundaryTimeExpire", cf.boundary(1, [0], ("double", 10.0))),
                                   ("BoundaryMomentumInject", cf.boundary(1, [0], ("double", 1.0)))])
    with pytest.raises(E.SpxError, match="momentum boundaries"):
        E.Simulation.from_setup(s).run(4)
    s = setup(bg, (0.0, 0.0, 0.0), [("BoundaryTimeExpire", cf.boundary(1, [0], ("double", 10.0)))])
    with pytest.raises(E.SpxError, match="device error"):      # no diffusion model needed: past the checks
        E.Simulation.from_setup(s).run(4)
