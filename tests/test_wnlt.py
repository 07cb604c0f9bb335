"""DiffusionWNLTConstant / DiffusionWNLTRampVLISM (src/diffusion_other.cc:156-330) in the oracle.

* 2F1(1, 1; c; z), which the reference takes from GSL, against scipy.special.hyp2f1 (an independent
  implementation) over the c that ps_index in (1, 3) gives and z up to 1 - 1e-12;
* with A2T = 0 the model is QLT: D_mumu equals DiffusionQLTConstant's bit for bit and kappa_perp is the
  closed form v sqrt(A2L / 8) v (1 - mu^2) / |Omega|; with A2L = 0,
  kappa_perp = v |mu| / k_min sqrt(A2T (q - 1) / (2 (q + 1)));
* the ramp is the constant model with l_max and the amplitudes rescaled: beyond z_sheath it equals the
  constant model, and at z0 = z_nose it equals the constant model with l_max_HP and amplitudes scaled
  by (l_max_HP / l_max)^(q - 1) (z0 = sqrt(r (r + z) / 2), the Rankine half-body coordinate).
The D_mumu term built on 2F1 is the reference's expression; beyond these reductions its parity rests on
the GPU against the oracle.
"""
import math

import numpy as np
import pytest
from scipy.special import hyp2f1

import feature_setups as F
from spectrum_amd import configs as cf
from spectrum_amd import engine as E
from spectrum_amd.data_container import DataContainer as DC

Q = 5.0 / 3.0
LMAX = 0.01 * cf.one_au


def test_hyp2f1_against_scipy(oracle_mod):
    worst = 0.0
    for q in (1.2, 1.5, 5.0 / 3.0, 2.0, 2.5, 2.9, 3.2):
        c = (5.0 + q) / 4.0
        for z in np.concatenate([np.linspace(0.0, 0.999, 120), 1.0 - np.logspace(-12, -3, 30)]):
            got, want = oracle_mod.hyp2f1_11(c, z), hyp2f1(1.0, 1.0, c, z)
            worst = max(worst, abs(got - want) / abs(want))
    assert worst < 1e-12, worst


def model(cls, A2A, l_max, q, A2T, A2L, ramp=None):
    dc = DC().insert_double(A2A).insert_double(l_max).insert_double(q)
    if A2T is not None:
        dc.insert_double(A2T).insert_double(A2L)
    if ramp:
        for v in ramp:
            dc.insert_double(v)
    return cls, dc


def oracle_with(oracle_mod, diffusion, B=(0.0, 0.0, 15.0)):
    s = cf.Setup("wnlt", cf.TRAJ_GUIDING_SCATT, cf.FORWARD)
    s.add(cf.BACKGROUND, "BackgroundUniform", cf.background_base(DC(), 0.0, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), B, 1.0))
    s.add(cf.DIFFUSION, *diffusion)
    s.add(cf.INITIAL, "InitialTimeFixed", DC().insert_double(0.0))
    s.add(cf.INITIAL, "InitialSpaceFixed", DC().insert_vec((0.0, 0.0, 0.0)))
    s.add(cf.INITIAL, "InitialMomentumShell", DC().insert_double(1.0))
    s.add(cf.BOUNDARY, "BoundaryTimeExpire", cf.boundary(1, [0], ("double", 1.0)))
    d = F.dist((4, 0, 0), (0.0, 0, 0), (1.0, 0, 0))
    d.insert_double(1.0).insert_double(0.0).insert_int(1)
    s.add(cf.DISTRIBUTION, "DistributionTimeUniform", d)
    return oracle_mod.Oracle(s)


def test_limits(oracle_mod):
    p = cf.mom_from_ekin(20.0 * cf.MeV)
    v = cf.vel_from_mom(p)
    Bm = 15.0
    gamma = 1.0 / math.sqrt(1.0 - (v / cf.c_code) ** 2)
    Omega = cf.charge * Bm / (gamma * 1.0 * cf.c_code)        # proton: mass 1 in code units
    kmin = 2.0 * math.pi / LMAX
    qlt = oracle_with(oracle_mod, model("DiffusionQLTConstant", 0.05, LMAX, Q, None, None))
    only_L = oracle_with(oracle_mod, model("DiffusionWNLTConstant", 0.05, LMAX, Q, 0.0, 0.2))
    only_T = oracle_with(oracle_mod, model("DiffusionWNLTConstant", 0.05, LMAX, Q, 0.1, 0.0))
    for mu in (-0.9, -0.3, 0.05, 0.5, 0.99):
        pos = (0.0, 0.0, 0.0)
        assert only_L.kappa_at(2, pos, p, mu) == qlt.kappa_at(2, pos, p, mu)
        st2 = 1.0 - mu * mu
        assert only_L.kappa_at(0, pos, p, mu) == pytest.approx(v * math.sqrt(0.2 / 8.0) * v * st2 / Omega, rel=1e-13)
        assert only_T.kappa_at(0, pos, p, mu) == pytest.approx(v * abs(mu) / kmin * math.sqrt(0.1 * (Q - 1) / (2 * (Q + 1))), rel=1e-13)
        # the nonlinear term adds to QLT's D_mumu (A2T > 0, both DT factors positive)
        assert only_T.kappa_at(2, pos, p, mu) > qlt.kappa_at(2, pos, p, mu)


def test_ramp_reduces_to_the_constant_model(oracle_mod):
    p = cf.mom_from_ekin(20.0 * cf.MeV)
    l_hp, z_nose, z_sheath = 0.002 * cf.one_au, 0.5 * cf.one_au, 1.5 * cf.one_au
    ramp = oracle_with(oracle_mod, model("DiffusionWNLTRampVLISM", 0.05, LMAX, Q, 0.1, 0.2, (l_hp, z_nose, z_sheath)))
    const = oracle_with(oracle_mod, model("DiffusionWNLTConstant", 0.05, LMAX, Q, 0.1, 0.2))
    k = (l_hp / LMAX) ** (Q - 1.0)
    nose = oracle_with(oracle_mod, model("DiffusionWNLTConstant", 0.05 * k, l_hp, Q, 0.1 * k, 0.2 * k))
    beyond = (0.0, 0.0, 2.0 * cf.one_au)                     # z0 = 2 au > z_sheath
    at_nose = (0.0, 0.0, z_nose)                             # on the axis, z0 = z
    for mu in (-0.7, 0.2, 0.8):
        for comp in (0, 2):
            assert ramp.kappa_at(comp, beyond, p, mu) == const.kappa_at(comp, beyond, p, mu)
            assert ramp.kappa_at(comp, at_nose, p, mu) == pytest.approx(nose.kappa_at(comp, at_nose, p, mu), rel=1e-12)
    # continuous across z_sheath
    below, above = (0.0, 0.0, z_sheath * (1 - 1e-9)), (0.0, 0.0, z_sheath * (1 + 1e-9))
    assert ramp.kappa_at(2, below, p, 0.4) == pytest.approx(ramp.kappa_at(2, above, p, 0.4), rel=1e-7)


def test_wnlt_host_checks():
    s = F.gc_wnlt_scatt_spiral()
    s.traj_type = cf.TRAJ_PARKER
    with pytest.raises(E.SpxError, match="guiding-centre"):
        E.Simulation.from_setup(s).run(4)
    s = F.gc_wnlt_ramp_diffscatt()
    s.plugins = [x for x in s.plugins if x[0] != cf.BACKGROUND]
    s.add(cf.BACKGROUND, "BackgroundUniform", cf.background_base(DC(), 0.0, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, 15.0), 1.0))
    with pytest.raises(E.SpxError, match="SolarWind background"):
        E.Simulation.from_setup(s).run(4)
    s = F.gc_wnlt_scatt_spiral()
    s.plugins = [x for x in s.plugins if x[0] != cf.DIFFUSION]
    s.add(cf.DIFFUSION, *model("DiffusionWNLTConstant", 0.05, LMAX, 3.0, 0.1, 0.2))
    with pytest.raises(E.SpxError, match="ps_index near 3"):
        E.Simulation.from_setup(s).run(4)
