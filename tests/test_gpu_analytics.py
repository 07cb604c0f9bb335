"""The reference's benchmark analytics, run on the HIP engine in both precisions.

* Modulation (benchmarks/main_test_modulation_cartesian_parker.cc and its post-processor
  benchmarks/main_postprocess_modulation_cartesian_parker.cc:9-31): the backward Parker estimate of the
  spectrum at 1 au on the analytic radial wind converges to the solution of the spherically symmetric
  Parker equation with the same boundaries (tests/analytic_modulation.py, a finite-difference solution
  pinned against the Fisk & Axford closed form).  At the reference's own time step the Euler-Maruyama
  estimate carries a step bias (a few percent at 10 MeV, where the 1% momentum-change limit sets the step);
  the run is repeated with the physical step scaled by 1/4 and 1/16 ("step_scale").  If the bias falls
  as a power dt^a with a >= 1/2 (the weak order of Euler-Maruyama, a = 1, or the order of a missed boundary
  crossing, a = 1/2), the bias left at the finest step is at most the change over the last refinement, so
  |m_16 - u| <= |m_4 - m_16| + 3.5 sigma in every energy bin.  The Fisk & Axford formula the reference
  plots is itself an approximation (no convection term, the local spectral slope frozen, an infinite
  domain); the estimate must be as close to it as the exact solution is, within the same bounds, which
  reproduces the benchmark's "small discrepancy at low energies".
* The gridded form of the modulation benchmark (benchmarks/main_test_modulation_cartesian_parker.cc): the same
  problem with the wind of BackgroundServerCartesian, the spiral sampled at zone centres and interpolated
  trilinearly (cartesian.parker_spiral, the reference's 25^3 blocks of 4^3 zones over +-100 au).  Its 2-au
  zones cannot resolve the radial wind near the Sun (at 1 au the interpolation mixes opposite directions), so
  the estimate differs from the Parker-equation solution most at low energies, the "small discrepancy at
  low energies" of the benchmark; as the zones shrink (25, 50, 75, 100 blocks per axis) the estimate must
  approach the solution of the radial problem.
* Diffusive shock acceleration (benchmarks/main_postprocess_diff_shock_acc.cc:9-34,
  benchmarks/benchmark_explained.txt:65-71): the steady-state spectrum at a planar shock with a free-escape
  plane a distance L downstream.  Upstream f = f0 exp(u1 x / kappa); downstream, with f(L) = 0,
  f = f0 (1 - exp(u2 (x - L) / kappa)) / (1 - eps), eps = exp(-u2 L / kappa).  The jump condition
  (u1 - u2)/3 p df0/dp = -[kappa df/dx] gives f0 ~ p^-q with
      q = 3 r / (r - 1) * (1 + eps / (r (1 - eps))),
  4.0187 for r = 4 and L = 4 kappa / u2 (the infinitely distant escape plane gives 4).  The escape rate at
  the plane then gives counts per log-momentum bin ~ p^(3 - q).
"""
import math

import numpy as np
import pytest

import analytic_modulation as A
from spectrum_amd import cartesian as CG
from spectrum_amd import configs as cf
from spectrum_amd import engine as E

pytestmark = pytest.mark.gpu
PRECS = [E.PREC_FP64, E.PREC_MIXED]
NSIG = 3.5


def run(setup, n, prec, seed=1, records=True):
    sim = E.Simulation.from_setup(setup, precision=prec)
    st, rec = sim.run(n, seed=seed, records=records)
    print(f"[run] {setup.name} prec={prec} n={n} steps={st.steps_done} kernel_ms={st.kernel_ms:.1f} failed={st.failed} "
          f"discards={st.discards}", flush=True)
    assert st.done == n and st.failed == 0
    return sim, st, rec


@pytest.fixture(scope="module")
def radial_solution():
    s = cf.parker_modulation_radial(n_energies=8)
    m = s.meta
    ps = np.array([cf.mom_from_ekin(T) for T in m["energies"]])
    kap = lambda r, p: m["kap0"] * (A.ekin(p) / m["T0"]) * (r / m["r0"]) ** 2
    w = lambda p: A.speed(p) * (A.ekin(p) / m["T0"]) ** m["gamma"] / p ** 2
    u = A.pde_solution(kap, w, m["V"], m["inner"], m["outer"], ps, cf.one_au, n_r=3000, ds=2e-3)
    fa = np.array([w(p) * A.fisk_axford(T, cf.one_au, m["kap0"], m["T0"], cf.one_au, 1.0, 2.0, m["V"], 1.8, 1.5)
                   for T, p in zip(m["energies"], ps)])
    return s, u, fa


def run_scaled(setup, n, prec, scale, seed=1):
    sim = E.Simulation.from_setup(setup, precision=prec, step_scale=scale)
    st, rec = sim.run(n, seed=seed)
    print(f"[run] {setup.name} prec={prec} step_scale={scale} n={n} steps={st.steps_done} kernel_ms={st.kernel_ms:.1f}", flush=True)
    assert st.done == n and st.failed == 0
    c, sums, _ = sim.distro(0)
    assert c.sum() == n                      # every trajectory reached a sphere: none was cut by the time limit
    return c, sums, rec


@pytest.mark.parametrize("prec", PRECS)
def test_modulation_spectrum_converges_to_the_parker_equation(prec, radial_solution):
    s, u, fa = radial_solution
    n_e, per = 8, 100000
    n = n_e * per
    e_of = np.arange(n) % n_e
    est = {}
    for scale in (1.0, 0.25, 0.0625):
        c, sums, rec = run_scaled(s, n, prec, scale)
        w = np.where(rec["n_events"] > 0, rec["ev_weight"][:, 0], np.nan)
        assert not np.isnan(w).any()
        est[scale] = []
        for k in range(n_e):
            wk = w[e_of == k]
            mean, err = sums[k] / c[k], wk.std() / math.sqrt(len(wk))
            assert mean == pytest.approx(wk.mean(), rel=1e-12)
            est[scale].append((mean, err))
    for k in range(n_e):
        (m1, e1), (m4, e4), (m16, e16) = est[1.0][k], est[0.25][k], est[0.0625][k]
        bias_bound = abs(m4 - m16)
        noise = NSIG * math.sqrt(2.0 * e16 ** 2 + e4 ** 2)
        print(f"[check] T = {s.meta['energies'][k] / cf.MeV:8.2f} MeV: Parker equation {u[k]:.6g}; estimate / exact - 1 at step x1 "
              f"{m1 / u[k] - 1:+.5f} ({(m1 - u[k]) / e1:+.1f} sigma), x1/4 {m4 / u[k] - 1:+.5f}, x1/16 {m16 / u[k] - 1:+.5f} "
              f"({(m16 - u[k]) / e16:+.1f} sigma, bias bound {bias_bound / u[k]:.5f}); Fisk-Axford / exact - 1 {fa[k] / u[k] - 1:+.4f}", flush=True)
        assert abs(m16 - u[k]) <= bias_bound + noise
        assert abs(m16 - fa[k]) <= abs(u[k] - fa[k]) + bias_bound + noise


def test_gridded_wind_converges_to_the_radial_solution(radial_solution):
    """Mixed kernel, 8 energies x 50000 trajectories, the physical step x 1/16 (its bias below 0.5%, see
    above).  The root mean square over the energies of estimate / exact - 1 must fall with every refinement
    of the grid (the per-energy Monte Carlo errors are 0.1-0.5%, the differences between grids several
    times that) and reach 1.5% or less in every energy on the finest grid (0.5-au zones)."""
    s0, u, _ = radial_solution
    n_e, per = 8, 50000
    n = n_e * per
    rms, worst = [], []
    for nb in (25, 50, 75, 100):
        s = cf.parker_modulation_radial(n_energies=n_e, grid=CG.parker_spiral(nb=nb))
        c, sums, _ = run_scaled(s, n, E.PREC_MIXED, 0.0625)
        dev = sums / c / u - 1.0
        rms.append(float(np.sqrt(np.mean(dev ** 2))))
        worst.append(float(np.abs(dev).max()))
        print(f"[grid] {nb}^3 blocks ({200.0 / (4 * nb):.2f}-au zones): estimate / exact - 1 = "
              + " ".join(f"{d:+.4f}" for d in dev) + f"; rms {rms[-1]:.4f}", flush=True)
    assert all(a > b for a, b in zip(rms, rms[1:])), rms
    assert worst[-1] <= 0.015
    assert rms[0] >= 5.0 * rms[-1]


def dsa_index(counts, p_edges):
    """Least-squares slope of log(counts per log bin) against log p, weighted by Poisson errors, over the
    bins with at least 400 counts; returns (q, error of q) with counts per log bin ~ p^(3 - q)."""
    pc = np.sqrt(p_edges[:-1] * p_edges[1:])
    m = counts >= 400
    x = np.log(pc[m])
    y = np.log(counts[m])
    wt = counts[m].astype(float)
    Am = np.vstack([x, np.ones_like(x)]).T
    cov = np.linalg.inv(Am.T @ (Am * wt[:, None]))
    coef = cov @ (Am.T @ (wt * y))
    return 3.0 - coef[0], math.sqrt(cov[0, 0]), int(m.sum())


def dsa_expected_index(r, u2, L, kappa):
    eps = math.exp(-u2 * L / kappa)
    return 3.0 * r / (r - 1.0) * (1.0 + eps / (r * (1.0 - eps)))


@pytest.mark.parametrize("prec", PRECS)
def test_dsa_index_with_the_escape_plane(prec):
    s = cf.dsa_shock()
    m = s.meta
    q_exp = dsa_expected_index(m["compression"], m["u1"] / m["compression"], m["L_down"], m["kappa"])
    assert q_exp == pytest.approx(4.0187, abs=1e-4)
    n = 400000
    sim, st, _ = run(s, n, prec, records=False)
    counts, _, _ = sim.distro(0)
    edges = m["p0"] * np.logspace(0, 4, 129)
    q, dq, nb = dsa_index(counts, edges)
    print(f"[check] DSA index q = {q:.4f} +- {dq:.4f} over {nb} bins; expected {q_exp:.4f} ({(q - q_exp) / dq:+.2f} sigma; "
          f"without the escape plane 4.0000: {(q - 4.0) / dq:+.2f} sigma)", flush=True)
    assert nb >= 20
    assert abs(q - q_exp) <= NSIG * dq
