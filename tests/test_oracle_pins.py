"""Pins of the CPU oracle against closed forms, known-answer vectors and hand calculations.

The reference ships no golden vectors for this path (its only self-checking test is the
DataContainer round trip, tests/main_test_data_container.cc), so every oracle function is
checked against something the mathematics fixes: Philox known-answer vectors (Salmon et al.,
SC'11 / Random123), the Parker-spiral closed form, the tanh shock profile, kappa formulas,
relativistic identities, and an Euler-Maruyama walk recomputed by hand from the same draws.
"""
import math

import numpy as np
import pytest

from spectrum_amd import configs as cf
from spectrum_amd.data_container import DataContainer


def test_code_units_match_reference_expressions(oracle_mod):
    O = oracle_mod
    # c = 10 code units exactly, mass and charge 1 (common/specie.hh:30-85)
    assert O.const("c_code") == 10.0 == cf.c_code
    assert O.const("mass") == 1.0 and O.const("charge") == 1.0
    assert O.const("one_au") == pytest.approx(1495.978707, rel=0, abs=1e-9)
    assert O.const("one_au") == cf.one_au
    assert O.const("unit_energy") == cf.unit_energy
    # 1 GeV proton: m c^2 = 100 code energy units and 1 GeV = 106.58 (0.1 GeV -> 10.658)
    assert O.const("GeV") == pytest.approx(106.58, rel=1e-4)


def test_kinematics_identities():
    for T in (1e-3, 0.1, 10.0, 1065.8):
        p = cf.mom_from_ekin(T)
        assert cf.ekin_from_mom(p) == pytest.approx(T, rel=1e-12)
        v = cf.vel_from_mom(p)
        gamma = 1.0 + T / 100.0
        assert v == pytest.approx(10.0 * math.sqrt(1.0 - 1.0 / gamma**2), rel=1e-12)


# Random123 known-answer vectors for philox4x32 with 10 rounds: (counter, key) -> output
PHILOX_KAT = [
    ((0x00000000, 0x00000000, 0x00000000, 0x00000000), (0x00000000, 0x00000000), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff, 0xffffffff, 0xffffffff, 0xffffffff), (0xffffffff, 0xffffffff), (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0), (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


@pytest.mark.parametrize("ctr,key,out", PHILOX_KAT)
def test_philox_known_answers(oracle_mod, ctr, key, out):
    assert tuple(oracle_mod.philox(ctr, key)) == out


def test_counter_normals_are_standard_normal(oracle_mod):
    L = oracle_mod.lib()
    z = np.array([L.orc_counter_normal(11, t, k) for t in range(200) for k in range(500)])
    n = z.size
    assert abs(z.mean()) < 4.0 / math.sqrt(n)
    assert abs(z.var() - 1.0) < 4.0 * math.sqrt(2.0 / n)
    # even/odd draws of one pair are uncorrelated
    pairs = z.reshape(-1, 2)
    assert abs(np.corrcoef(pairs[:, 0], pairs[:, 1])[0, 1]) < 4.0 / math.sqrt(pairs.shape[0])
    u = np.array([L.orc_counter_uniform(11, t, k) for t in range(100) for k in range(500)])
    assert u.min() > 0.0 and u.max() <= 1.0
    assert abs(u.mean() - 0.5) < 4.0 * math.sqrt(1.0 / 12.0 / u.size)


def parker_spiral(pos, V, B1au_code, r_ref, omega):
    """Parker spiral in the global frame: B = B_r r^ + B_phi phi^, U = V r^ (rotation axis z)."""
    x, y, z = pos
    r = math.sqrt(x * x + y * y + z * z)
    st = math.sqrt(x * x + y * y) / r
    Bref = B1au_code * (cf.one_au / r_ref) ** 2
    Br = Bref * (r_ref / r) ** 2
    Bp = -Br * st * (r - r_ref) * omega / V
    rhat = np.array(pos) / r
    phihat = np.array([-y, x, 0.0]) / math.hypot(x, y)
    return V * rhat, Br * rhat + Bp * phihat


@pytest.mark.parametrize("pos_au", [(1.0, 0.0, 0.0), (3.0, 4.0, 1.0), (-20.0, 7.0, -30.0), (0.5, -0.2, 0.1)])
def test_solarwind_matches_parker_spiral(oracle_mod, pos_au):
    s = cf.parker_modulation()
    o = oracle_mod.Oracle(s)
    pos = [c * cf.one_au for c in pos_au]
    p = cf.mom_from_ekin(cf.GeV)
    U, B, dmax, divU = o.background_fields(0.0, pos, p)
    V = 400.0e5 / cf.unit_velocity
    Ue, Be = parker_spiral(pos, V, 5.0e-5 / cf.unit_magnetic, 3.0 * 6.957e10 / cf.unit_length, 2.87e-6 * cf.unit_time)
    np.testing.assert_allclose(U, Ue, rtol=1e-13, atol=1e-13 * V)
    np.testing.assert_allclose(B, Be, rtol=1e-12, atol=1e-12 * np.linalg.norm(Be))
    r = np.linalg.norm(pos)
    assert dmax == pytest.approx(min(0.1 * r, cf.one_au), rel=1e-15)
    # div U = 2 V / r for a radial wind; the central differences over a Larmor radius are O((r_g/r)^2)
    assert divU == pytest.approx(2.0 * V / r, rel=2e-4)


@pytest.mark.parametrize("x", [-30.0, -2.5, -0.3, 0.0, 0.7, 4.0, 40.0])
def test_smooth_shock_profile(oracle_mod, x):
    s = cf.dsa_shock()
    o = oracle_mod.Oracle(s)
    u1, w = s.meta["u1"], s.meta["width"]
    U, B, dmax, divU = o.background_fields(0.0, [x, 0.3, -0.2], s.meta["p0"])
    ds = -x / w                                   # n = (-1, 0, 0): upstream is x < 0
    a1 = 0.5 * (1.0 + math.tanh(8.0 * ds))
    Ux = u1 * a1 + (u1 / 4.0) * (1.0 - a1)
    assert U[0] == pytest.approx(Ux, rel=1e-14)
    assert U[1] == 0.0 and U[2] == 0.0
    # dU/dx = (u1 - u2) * d a1/dx ; div U = dUx/dx
    dadx = 0.5 * 8.0 * (1.0 - math.tanh(8.0 * ds) ** 2) * (-1.0 / w)
    assert divU == pytest.approx((u1 - u1 / 4.0) * dadx, rel=1e-12, abs=1e-300)


def test_rigidity_powerlaw_kappa(oracle_mod):
    s = cf.parker_modulation()
    o = oracle_mod.Oracle(s)
    for T in (0.1 * cf.GeV, cf.GeV, 10.0 * cf.GeV):
        p = cf.mom_from_ekin(T)
        beta = cf.vel_from_mom(p) / cf.c_code
        R_GV = p * cf.c_code / cf.GV
        kpar = 1.5e22 / cf.unit_diffusion * beta * R_GV**0.5
        assert o.kappa(1, [cf.one_au, 0, 0], p, Bmag=7.0) == pytest.approx(kpar, rel=1e-12)
        assert o.kappa(0, [cf.one_au, 0, 0], p, Bmag=7.0) == pytest.approx(0.02 * kpar, rel=1e-12)


def uniform_walk_setup(U=(0.01, -0.02, 0.005), B=(3.0, 4.0, 0.0), D=2.0, t_limit=1000.0, dmax=5.0, flow=cf.BACKWARD):
    """Uniform background, constant isotropic diffusion, time limit with a binned time-distribution event."""
    s = cf.Setup("uniform_walk", cf.TRAJ_PARKER, flow)
    s.add(cf.BACKGROUND, "BackgroundUniform", cf.background_base(DataContainer(), 0.0, (0, 0, 0), U, B, dmax))
    s.add(cf.DIFFUSION, "DiffusionFullConstant", DataContainer().insert_double(D).insert_double(D))
    s.add(cf.INITIAL, "InitialTimeFixed", DataContainer().insert_double(0.0))
    s.add(cf.INITIAL, "InitialSpaceFixed", DataContainer().insert_vec((1.0, 2.0, 3.0)))
    s.add(cf.INITIAL, "InitialMomentumShell", DataContainer().insert_double(cf.mom_from_ekin(cf.MeV)))
    tmark = -t_limit if flow == cf.BACKWARD else t_limit
    s.add(cf.BOUNDARY, "BoundaryTimeExpire", cf.boundary(1, [0], ("double", tmark)))
    dst = cf.distribution_base(DataContainer(), (10, 0, 0), (tmark * 1.1 if flow == cf.BACKWARD else 0.0, 0, 0),
                               (0.0 if flow == cf.BACKWARD else tmark * 1.1, 0, 0), (0, 0, 0), (0, 0, 0))
    dst.insert_double(1.0).insert_double(0.0).insert_int(1)
    s.add(cf.DISTRIBUTION, "DistributionTimeUniform", dst)
    return s


@pytest.mark.parametrize("flow", [cf.BACKWARD, cf.FORWARD])
def test_euler_walk_by_hand(oracle_mod, flow):
    """Constant coefficients: x_N - x_0 = -/+ sum dt_k U + sum R sqrt(2 D dt_k) n_k, with dt set by
    PhysicalStep and shortened by TimeBoundaryProximityCheck on the last step."""
    U = np.array([0.01, -0.02, 0.005]); B = np.array([3.0, 4.0, 0.0]); D = 2.0; dmax = 5.0; T = 1000.0
    s = uniform_walk_setup(tuple(U), tuple(B), D, T, dmax, flow)
    rng = np.random.default_rng(5)
    n_traj = 4
    normals = rng.standard_normal((n_traj, 3000))
    o = oracle_mod.Oracle(s)
    _, rec = o.run(n_traj, normals=normals, uniforms=np.full((n_traj, 2), 0.5), threads=1)
    # hand calculation
    dt0 = min(0.5 * dmax / np.linalg.norm(U), 0.5 * dmax * dmax / D)
    bhat = B / np.linalg.norm(B)
    e0 = np.cross(bhat, [0.0, 0.0, 1.0]); e0 /= np.linalg.norm(e0)    # smallest |b| component is z
    e1 = np.cross(bhat, e0)
    sign = -1.0 if flow == cf.BACKWARD else 1.0
    for i in range(n_traj):
        t, x, k = 0.0, np.array([1.0, 2.0, 3.0]), 0
        while True:
            remaining = T - abs(t)
            dt = dt0 if dt0 < remaining else max((1.0 + 1e-5) * remaining, 1e-8 * dmax / cf.c_code)
            dW = math.sqrt(dt) * normals[i, 3 * k:3 * k + 3]
            x = x + sign * dt * U + math.sqrt(2 * D) * (dW[0] * e0 + dW[1] * e1 + dW[2] * bhat)
            t += sign * dt
            k += 1
            if abs(t) >= T:
                break
        assert rec["steps"][i] == k
        assert rec["t"][i] == pytest.approx(t, rel=1e-13)
        np.testing.assert_allclose(rec["pos"][i], x, rtol=1e-11, atol=1e-11)
        # the time event is binned: value = t at expiry on 10 linear bins over [-1.1 T, 0] (or [0, 1.1 T])
        lo = -1.1 * T if flow == cf.BACKWARD else 0.0
        expect_bin = int((t - lo) / (1.1 * T / 10))
        assert rec["ev_bin"][i, 0] == expect_bin


def test_diffusion_mean_square_displacement(oracle_mod):
    """<dx dx^T> = 2 kappa_ij t for constant anisotropic diffusion (benchmarks/benchmark_explained.txt:49-63)."""
    D_perp, D_par, T = 0.5, 3.0, 400.0
    s = cf.Setup("msd", cf.TRAJ_PARKER, cf.FORWARD)
    s.add(cf.BACKGROUND, "BackgroundUniform", cf.background_base(DataContainer(), 0.0, (0, 0, 0), (0, 0, 0), (0, 0, 2.0), 50.0))
    s.add(cf.DIFFUSION, "DiffusionFullConstant", DataContainer().insert_double(D_perp).insert_double(D_par))
    s.add(cf.INITIAL, "InitialTimeFixed", DataContainer().insert_double(0.0))
    s.add(cf.INITIAL, "InitialSpaceFixed", DataContainer().insert_vec((0.0, 0.0, 0.0)))
    s.add(cf.INITIAL, "InitialMomentumShell", DataContainer().insert_double(1.0))
    s.add(cf.BOUNDARY, "BoundaryTimeExpire", cf.boundary(1, [0], ("double", T)))
    dst = cf.distribution_base(DataContainer(), (1, 0, 0), (0.0, 0, 0), (2.0 * T, 0, 0), (0, 0, 0), (1, 0, 0))
    s.add(cf.DISTRIBUTION, "DistributionPositionCumulativeOrder2", dst)
    o = oracle_mod.Oracle(s)
    n = 4000
    o.run(n, seed=3, records=False)
    counts, sums, nev = o.distro(0)
    assert counts[0] == n and nev == n
    m = sums[0].reshape(3, 3) / n
    t_end = T * (1.0 + 1e-5)  # the last step lands just past the time boundary
    for i, Dii in enumerate((D_perp, D_perp, D_par)):
        expect = 2.0 * Dii * T
        # var of dx^2 for a Gaussian is 2 sigma^4 -> standard error sqrt(2/n) sigma^2
        assert abs(m[i, i] - expect) < 4.0 * math.sqrt(2.0 / n) * expect + 2.0 * Dii * (t_end - T)
    assert abs(m[0, 1]) < 4.0 * math.sqrt(1.0 / n) * 2 * D_perp * T


def test_spectrum_binning_exact(oracle_mod):
    """Each grid energy lands in its own log bin; counts per bin are exact integers."""
    s = cf.parker_modulation(n_energies=8, t_limit_years=0.2 / 365.0)
    # make the time expiry a binned event so every trajectory adds one event
    plugins = []
    for role, name, blob in s.plugins:
        if name == "BoundaryTimeExpire":
            blob = cf.boundary(1, [0], ("double", -0.2 / 365.0 * cf.one_year)).bytes()
        plugins.append((role, name, blob))
    s.plugins = plugins
    o = oracle_mod.Oracle(s)
    n = 64
    _, rec = o.run(n, seed=9, threads=4)
    counts, sums, nev = o.distro(0)
    binned = rec["ev_bin"][:, 0] >= 0
    # trajectories that reached the inner sphere (action -1) carry no event
    assert nev == binned.sum() >= n - 4
    np.testing.assert_array_equal(counts, np.bincount(rec["ev_bin"][binned, 0], minlength=8))
    np.testing.assert_array_equal(rec["ev_bin"][binned, 0], (np.arange(n) % 8)[binned])
    np.testing.assert_allclose(sums, np.bincount(rec["ev_bin"][binned, 0], weights=rec["ev_weight"][binned, 0], minlength=8), rtol=1e-14)


def legendre_mean(mu, l):
    P = [np.ones_like(mu), mu, 0.5 * (3 * mu**2 - 1), 0.5 * (5 * mu**3 - 3 * mu)]
    return P[l].mean(), P[l].std() / math.sqrt(mu.size)


def test_gc_pitch_angle_isotropization(oracle_mod):
    """Guiding centre in a uniform field, isotropic D_mumu = D0 (1 - mu^2): the Legendre moments of a ring
    relax as <P_l(mu)>(t) = P_l(mu0) exp(-l (l + 1) D0 t); |p| is conserved exactly; the perpendicular
    spread is <x^2 + y^2> = 4 kappa_perp t; the mean parallel displacement is v mu0 (1 - exp(-2 D0 t)) / (2 D0)."""
    mu0 = 0.8
    s = cf.gc_uniform(theta0=math.acos(mu0), t_limit=500.0)
    o = oracle_mod.Oracle(s)
    n = 20000
    _, rec = o.run(n, seed=3, threads=8)
    assert (rec["status"] == 0).all()
    p = np.hypot(rec["mom"][:, 0], rec["mom"][:, 2])
    np.testing.assert_allclose(p, s.meta["p"], rtol=1e-13)
    mu = rec["mom"][:, 2] / p
    D0, T, v = s.meta["D0"], s.meta["t_limit"], s.meta["v"]
    for l, Pl0 in ((1, mu0), (2, 0.5 * (3 * mu0**2 - 1)), (3, 0.5 * (5 * mu0**3 - 3 * mu0))):
        m, e = legendre_mean(mu, l)
        assert abs(m - Pl0 * math.exp(-l * (l + 1) * D0 * T)) < 3.5 * e, (l, m, e)
    r2 = rec["pos"][:, 0] ** 2 + rec["pos"][:, 1] ** 2
    assert abs(r2.mean() - 4 * s.meta["kperp"] * T) < 3.5 * r2.std() / math.sqrt(n)
    z = rec["pos"][:, 2]
    assert abs(z.mean() - v * mu0 * (1 - math.exp(-2 * D0 * T)) / (2 * D0)) < 3.5 * z.std() / math.sqrt(n)


def test_gc_isotropic_parallel_spread(oracle_mod):
    """Isotropic start: <z^2>(t) = (2 v^2 / 3) [t / (2 D0) - (1 - exp(-2 D0 t)) / (4 D0^2)] (the l = 1 mode
    correlation <mu(s) mu(u)> = exp(-2 D0 |s - u|) / 3); for t >> 1 / D0 this is 2 kappa_par t with
    kappa_par = v^2 / (6 D0), the reference's pitch-angle benchmark (benchmarks/main_test_pa_scatt.cc:252)."""
    s = cf.gc_uniform(theta0=None, t_limit=2000.0)
    o = oracle_mod.Oracle(s)
    n = 8000
    _, rec = o.run(n, seed=4, threads=8)
    D0, T, v = s.meta["D0"], s.meta["t_limit"], s.meta["v"]
    z2 = rec["pos"][:, 2] ** 2
    expect = (2 * v * v / 3) * (T / (2 * D0) - (1 - math.exp(-2 * D0 * T)) / (4 * D0 * D0))
    assert abs(z2.mean() - expect) < 3.5 * z2.std() / math.sqrt(n), (z2.mean(), expect)
    mu = rec["mom"][:, 2] / np.hypot(rec["mom"][:, 0], rec["mom"][:, 2])
    for l in (1, 2):
        m, e = legendre_mean(mu, l)
        assert abs(m) < 3.5 * e
