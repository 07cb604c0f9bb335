"""Statistical checks of the HIP engine against closed forms and between its two precisions.

Each check runs the plain FP64 kernel and the FP32-step / FP64-accumulate kernel on the same
counter-based streams and compares Monte Carlo estimates with the analytic value within 3.5
standard errors (the bound is written in each test).  The mixed kernel is accepted only where its
estimates agree with the FP64 kernel's within the same Monte Carlo error, which is the sense in which
the two are compared "at equal 3 sigma error".
"""
import math

import numpy as np
import pytest

from spectrum_amd import configs as cf
from spectrum_amd import engine as E

pytestmark = pytest.mark.gpu
PRECS = [E.PREC_FP64, E.PREC_MIXED]
NSIG = 3.5


def run(setup, n, prec, seed=1, records=True, **opts):
    sim = E.Simulation.from_setup(setup, precision=prec, **opts)
    st, rec = sim.run(n, seed=seed, records=records)
    print(f"[run] {setup.name} prec={prec} n={n} steps={st.steps_done} kernel_ms={st.kernel_ms:.1f} failed={st.failed} "
          f"discards={st.discards}", flush=True)
    assert st.done == n
    return sim, st, rec


def within(est, err, expect, what):
    print(f"[check] {what}: {est:.6g} +- {err:.3g} vs {expect:.6g} ({(est - expect) / err if err else 0:+.2f} sigma)", flush=True)
    assert abs(est - expect) <= NSIG * err, what


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("flow", [cf.FORWARD, cf.BACKWARD])
def test_parker_uniform_walk_msd(prec, flow):
    """Constant isotropic kappa = D in a uniform flow: x(T) - x0 -/+ U T ~ N(0, 2 D T) per axis."""
    U = (0.01, -0.02, 0.005)
    s = cf.uniform_walk(D=2.0, U=U, t_limit=1000.0, flow=flow)
    n = 200000
    _, st, rec = run(s, n, prec)
    assert st.failed == 0
    T = np.abs(rec["t"] - rec["t0"])
    sign = 1.0 if flow == cf.FORWARD else -1.0
    d = rec["pos"] - rec["pos0"] - sign * np.outer(T, U)
    for k in range(3):
        within(d[:, k].mean(), d[:, k].std() / math.sqrt(n), 0.0, f"<dx_{k}>")
        v = d[:, k] ** 2
        within(v.mean(), v.std() / math.sqrt(n), 2.0 * 2.0 * T.mean(), f"<dx_{k}^2>")


def converged(results, expect, what):
    """results: {step scale: (estimate, error)} at scales 1, 1/4, 1/16.  The Euler-Maruyama estimate of a
    pitch-angle process carries a step bias (chiefly where the rejected draws at |mu| = 1 act); if the bias
    falls as dt^a with a >= 1/2, the bias left at the finest step is at most the change over the last
    refinement, so |m_16 - exact| <= |m_4 - m_16| + 3.5 sigma."""
    (m1, e1), (m4, e4), (m16, e16) = results[1.0], results[0.25], results[0.0625]
    bound = abs(m4 - m16)
    noise = NSIG * math.sqrt(2.0 * e16 ** 2 + e4 ** 2)
    print(f"[check] {what}: exact {expect:.6g}; step x1 {m1:.6g} ({(m1 - expect) / e1:+.2f} sigma), x1/4 {m4:.6g} "
          f"({(m4 - expect) / e4:+.2f} sigma), x1/16 {m16:.6g} ({(m16 - expect) / e16:+.2f} sigma); bias bound {bound:.3g}", flush=True)
    assert abs(m16 - expect) <= bound + noise, what


@pytest.mark.parametrize("prec", PRECS)
def test_gc_pitch_angle_isotropization(prec):
    """<P_l(mu)> = P_l(mu0) exp(-l(l+1) D0 t) for D_mumu = D0 (1 - mu^2), 1e6 trajectories per step scale."""
    mu0 = 0.8
    s = cf.gc_uniform(theta0=math.acos(mu0), t_limit=500.0)
    n = 1000000
    D0, T = s.meta["D0"], s.meta["t_limit"]
    P0 = {1: mu0, 2: 0.5 * (3 * mu0**2 - 1), 3: 0.5 * (5 * mu0**3 - 3 * mu0)}
    res = {l: {} for l in (1, 2, 3)}
    for scale in (1.0, 0.25, 0.0625):
        _, st, rec = run(s, n, prec, step_scale=scale)
        assert st.failed == 0
        p = np.hypot(rec["mom"][:, 0], rec["mom"][:, 2])
        np.testing.assert_allclose(p, s.meta["p"], rtol=1e-12 if prec == E.PREC_FP64 else 1e-6)
        mu = rec["mom"][:, 2] / p
        P = {1: mu, 2: 0.5 * (3 * mu**2 - 1), 3: 0.5 * (5 * mu**3 - 3 * mu)}
        for l in (1, 2, 3):
            res[l][scale] = (P[l].mean(), P[l].std() / math.sqrt(n))
        if scale == 1.0:   # the perpendicular spread has no step bias: a constant kappa_perp, Gaussian increments
            r2 = rec["pos"][:, 0] ** 2 + rec["pos"][:, 1] ** 2
            within(r2.mean(), r2.std() / math.sqrt(n), 4 * s.meta["kperp"] * T, "<x^2+y^2>")
    for l in (1, 2, 3):
        converged(res[l], P0[l] * math.exp(-l * (l + 1) * D0 * T), f"<P_{l}(mu)>")


@pytest.mark.parametrize("prec", PRECS)
def test_gc_isotropic_parallel_spread(prec):
    s = cf.gc_uniform(theta0=None, t_limit=2000.0)
    n = 200000
    D0, T, v = s.meta["D0"], s.meta["t_limit"], s.meta["v"]
    expect = (2 * v * v / 3) * (T / (2 * D0) - (1 - math.exp(-2 * D0 * T)) / (4 * D0 * D0))
    res = {}
    for scale in (1.0, 0.25, 0.0625):
        _, st, rec = run(s, n, prec, records=True, step_scale=scale)
        z2 = rec["pos"][:, 2] ** 2
        res[scale] = (z2.mean(), z2.std() / math.sqrt(n))
    converged(res, expect, "<z^2>")


def test_mixed_matches_fp64_modulation_spectrum():
    """The reference problem at the benchmark's per-energy scale (backward modulation to 90 au, 64 energies,
    1e6 trajectories per precision): the intensity estimate of every energy bin from the mixed kernel agrees
    with the FP64 kernel's within the Monte Carlo error of the difference, and the mean step count too."""
    n_e, per = 64, 15625
    s = cf.parker_modulation(n_energies=n_e)
    n = n_e * per
    out = {}
    for prec in PRECS:
        sim, st, rec = run(s, n, prec)
        out[prec] = rec
    f, m = out[E.PREC_FP64], out[E.PREC_MIXED]
    bins = np.arange(n) % n_e
    wf = np.where(f["ev_bin"][:, 0] >= 0, f["ev_weight"][:, 0], 0.0)
    wm = np.where(m["ev_bin"][:, 0] >= 0, m["ev_weight"][:, 0], 0.0)
    worst, worst_b = 0.0, 0
    for b in range(n_e):
        sel = bins == b
        k = sel.sum()
        diff = wm[sel] - wf[sel]
        err = max(np.sqrt(wf[sel].var() / k + wm[sel].var() / k), 1e-300)
        z = abs(diff.mean()) / err
        if z > worst:
            worst, worst_b = z, b
        assert z <= NSIG, (b, diff.mean(), err)
    sel = bins == worst_b
    steps_f, steps_m = f["steps"].mean(), m["steps"].mean()
    print(f"[check] worst bin |mixed - fp64| = {worst:.2f} sigma (bin {worst_b}: fp64 {wf[sel].mean():.6g} +- {wf[sel].std() / math.sqrt(sel.sum()):.3g}, "
          f"mixed {wm[sel].mean():.6g} +- {wm[sel].std() / math.sqrt(sel.sum()):.3g}); mean steps fp64 {steps_f:.0f} mixed {steps_m:.0f}", flush=True)
    assert abs(steps_m - steps_f) <= NSIG * math.sqrt(f["steps"].var() / n + m["steps"].var() / n)


def test_mixed_step_arithmetic_bias_on_shared_paths():
    """What the FP32 step arithmetic changes on one path: both kernels replay the same pre-generated N(0,1)
    increments (the counter streams of the two precisions map Philox blocks to normals differently), and
    over 0.1 day (a few hundred steps) the paths stay close, so the paired difference of the spectrum
    weight J(p) at the time limit isolates the FP32 step error from the Monte Carlo noise.  Its mean (the
    systematic part) must be small against the Monte Carlo error of a 1e7-trajectory spectrum."""
    n_e, per = 64, 100
    s = cf.parker_modulation(n_energies=n_e, t_limit_years=0.1 / 365.0, inner_au=None, time_action=0)
    n = n_e * per
    rng = np.random.default_rng(77)
    normals = rng.standard_normal((n, 12000))         # 3 draws per step; the longest of 6400 took 3534 steps (oracle)
    uniforms = rng.random((n, 8))
    out = {}
    for prec in PRECS:
        sim = E.Simulation.from_setup(s, precision=prec)
        st, rec = sim.run(n, normals=normals, uniforms=uniforms)
        print(f"[run] {s.name} prec={prec} n={n} replayed steps={st.steps_done} discards={st.discards}", flush=True)
        assert st.done == n and st.failed == 0
        out[prec] = rec
    f, m = out[E.PREC_FP64], out[E.PREC_MIXED]
    assert (f["ev_bin"][:, 0] >= 0).all() and (m["ev_bin"][:, 0] >= 0).all()
    wf, wm = f["ev_weight"][:, 0], m["ev_weight"][:, 0]
    rel = (wm - wf) / wf
    dpos = np.linalg.norm(m["pos"] - f["pos"], axis=1) / np.linalg.norm(f["pos"], axis=1)
    same_steps = (m["steps"] == f["steps"]).mean()
    bias, bias_err = rel.mean(), rel.std() / math.sqrt(n)
    # a trajectory whose step count differs took one step-size decision the other way and left the shared
    # path (a fraction of a percent); the per-energy bound is taken over those that stayed on it
    bins = np.arange(n) % n_e
    shared = m["steps"] == f["steps"]
    per_bin = max(abs(rel[(bins == b) & shared].mean()) + NSIG * rel[(bins == b) & shared].std() / math.sqrt(((bins == b) & shared).sum())
                  for b in range(n_e))
    print(f"[check] shared paths over 0.1 day: weight rel. difference median |.| {np.median(np.abs(rel)):.2e}, mean {bias:+.2e} +- {bias_err:.1e} "
          f"(all trajectories); worst energy |mean| + 3.5 sigma {per_bin:.2e} (those with equal step counts); position rel. difference median "
          f"{np.median(dpos):.2e}; same step count {same_steps:.4f}", flush=True)
    # the smallest relative Monte Carlo error of a whole-trajectory bin at 156250 per energy (the least
    # modulated energy, 10 GeV) is ~4e-5
    assert same_steps > 0.99
    assert abs(bias) + NSIG * bias_err < 1e-5
    assert per_bin < 1e-5
