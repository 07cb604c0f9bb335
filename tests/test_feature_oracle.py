"""CPU checks of the oracle on the feature setups: each plugin class it lowers must do what the
mathematics of that class fixes, independently of the kernels.

* reflecting plane / sphere: no endpoint on the far side;
* time pass / recurrent / expire: the exact number of events per trajectory;
* parallel-only diffusion: the displacement minus advection is parallel to B;
* perpendicular-only diffusion with B along y and U along x: no displacement along y;
* InitialMomentumThickShell with log spacing: log|p0| uniform between the shell radii;
* momentum-inject boundary: every trajectory ends at or above the injection momentum or at the time limit;
* InitialSpaceSphere / InitialSpaceSphereSector: starts on the sphere, uniform in area (the polar cosine and
  the azimuth uniform, Kolmogorov-Smirnov), inside the sector's angles.
"""
import numpy as np
import pytest

import feature_setups as F
from spectrum_amd import configs as cf
from spectrum_amd import engine as E

THREADS = 8


def run(oracle_mod, mk, n, seed=3):
    s = mk()
    o = oracle_mod.Oracle(s)
    _, rec = o.run(n, seed=seed, threads=THREADS)
    assert (rec["status"] == 0).all()
    return s, o, rec


@pytest.mark.parametrize("k", range(len(F.ALL)), ids=F.NAMES)
def test_setup_lowers_and_bins(oracle_mod, k):
    s = F.ALL[k]()
    sim = E.Simulation.from_setup(s)          # the engine's host lowering accepts every class
    assert sim.n_dist == sum(1 for p in s.plugins if p[0] == cf.DISTRIBUTION)
    o = oracle_mod.Oracle(s)
    _, rec = o.run(64, seed=5, threads=THREADS)
    assert (rec["status"] == 0).all()
    for d in range(sim.n_dist):
        c, _, ne = o.distro(d)
        assert ne > 0 and c.sum() > 0


def test_shock_reflect_time_pass_and_thick_shell(oracle_mod):
    _, _, rec = run(oracle_mod, F.shock_flowmom_reflect, 2048)
    assert rec["pos"][:, 0].min() >= -30.0                # reflecting plane at x = -30
    np.testing.assert_allclose(rec["t"], 400.0, rtol=1e-7)  # time expiry
    assert (rec["n_events"] == 2).all()                    # one pass at t=150 plus the expiry
    lp = np.log(np.linalg.norm(rec["mom0"], axis=1))
    lo, hi = np.log(0.03), np.log(0.3)
    assert lp.min() >= lo and lp.max() <= hi
    u = np.sort((lp - lo) / (hi - lo))                     # Kolmogorov-Smirnov against U(0,1)
    ks = np.max(np.abs(u - (np.arange(1, u.size + 1) / u.size)))
    assert ks < 1.63 / np.sqrt(u.size)                     # 99% level


def test_sphere_reflect_and_recurrent_time(oracle_mod):
    _, _, rec = run(oracle_mod, F.solarwind_kinrad_reflect_recurrent, 512)
    r = np.linalg.norm(rec["pos"], axis=1)
    assert r.min() >= 1.1 * cf.one_au * (1 - 1e-12)
    np.testing.assert_allclose(rec["t"], -2.25 * cf.one_day, rtol=1e-7)
    assert (rec["n_events"] == 10).all()                   # 4 recurrent marks + expiry, two distributions


def test_parallel_only_diffusion_stays_on_the_field_line(oracle_mod):
    _, _, rec = run(oracle_mod, F.uniform_para_plane_pass, 512)
    U = np.array([0.002, 0.001, 0.0])
    b = np.array([3.0, 4.0, 12.0]) / 13.0
    dx = rec["pos"] - rec["t"][:, None] * U
    assert np.abs(np.cross(dx, b)).max() < 1e-9 * np.abs(dx @ b).max()


def test_perpendicular_only_diffusion_and_momentum_boundaries(oracle_mod):
    _, _, rec = run(oracle_mod, F.smoothshock_perp_momentum, 96)
    assert np.abs(rec["pos"][:, 1]).max() == 0.0           # B along y: no motion along y
    p = np.linalg.norm(rec["mom"], axis=1)
    p0 = np.linalg.norm(rec["mom0"], axis=1)
    at_limit = np.isclose(rec["t"], 20000.0, rtol=1e-7)
    assert ((p >= 1.6 * p0 * (1 - 1e-12)) | at_limit).all()
    assert (~at_limit).any()                               # some reach the injection momentum


def test_sphere_and_sector_starts(oracle_mod):
    from scipy import stats
    _, _, rec = run(oracle_mod, F.uniform_sphere_start, 4000)
    d = rec["pos0"] - np.array([1.0, 2.0, 3.0])
    r = np.linalg.norm(d, axis=1)
    np.testing.assert_allclose(r, 5.0, rtol=1e-14)
    ct, phi = d[:, 2] / r, np.arctan2(d[:, 1], d[:, 0])
    assert stats.kstest(ct, stats.uniform(loc=-1.0, scale=2.0).cdf).pvalue > 1e-3
    assert stats.kstest(phi, stats.uniform(loc=-np.pi, scale=2 * np.pi).cdf).pvalue > 1e-3
    # the momentum still takes its own uniforms: log p uniform over the shell
    assert stats.kstest(np.log10(rec["mom0"][:, 0]), stats.uniform(loc=np.log10(0.05), scale=1.0).cdf).pvalue > 1e-3
    assert abs(np.corrcoef(ct, np.log10(rec["mom0"][:, 0]))[0, 1]) < 4.0 / np.sqrt(len(ct))
    s = F.gc_sector_start()
    s.plugins = [x for x in s.plugins if not (x[0] == cf.BOUNDARY)]
    s.add(cf.BOUNDARY, "BoundaryTimeExpire", cf.boundary(1, [0], ("double", 1e-6)))
    _, rec = oracle_mod.Oracle(s).run(3000, seed=5, threads=THREADS)
    d = rec["pos0"]
    r = np.linalg.norm(d, axis=1)
    np.testing.assert_allclose(r, cf.one_au, rtol=1e-14)
    ct, phi = d[:, 2] / r, np.arctan2(d[:, 1], d[:, 0])
    lo, hi = np.cos(1.9), np.cos(1.2)
    assert ct.min() >= lo - 1e-12 and ct.max() <= hi + 1e-12 and phi.min() >= -0.3 - 1e-12 and phi.max() <= 0.4 + 1e-12
    assert stats.kstest(ct, stats.uniform(loc=lo, scale=hi - lo).cdf).pvalue > 1e-3
    assert stats.kstest(phi, stats.uniform(loc=-0.3, scale=0.7).cdf).pvalue > 1e-3
