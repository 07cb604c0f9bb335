"""The gridded background (BackgroundServerCartesian) on the GPU: FP64 parity with the oracle on the
reference's Cartesian modulation benchmark (the 25^3-block Parker spiral), on a small grid that
trajectories leave (field-error discards and retries), through the .info/.out files, through the
query server, and the mixed-precision kernel against FP64 within Monte Carlo error.
"""
import os

import numpy as np
import pytest

from spectrum_amd import cartesian as G
from spectrum_amd import configs as cf
from spectrum_amd import engine as E
from spectrum_amd.server import Server
from test_gpu_parity import check_parity, sensitivity

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)


@pytest.fixture(scope="module")
def spiral():
    return G.parker_spiral()


@pytest.mark.parametrize("kap_rat", [1.0, 0.3])
def test_cartesian_modulation_fp64_parity(oracle_mod, spiral, kap_rat):
    # kap_rat 1 (the benchmark): the derivative points of div(kappa) need only their positions checked;
    # kap_rat 0.3: bhat enters div(kappa), so every derivative point interpolates the grid
    s = cf.parker_modulation_cartesian(spiral, kap_rat=kap_rat)
    n = 512
    print(f"[grid] modulation parity n={n} kap_rat={kap_rat}", flush=True)
    sim = E.Simulation.from_setup(s, precision=E.PREC_FP64, lanes=1024, steps_per_launch=500)
    st, rec = sim.run(n, seed=5)
    o = oracle_mod.Oracle(s)
    _, orec = o.run(n, seed=5, threads=THREADS)
    # Near the inner sphere the derivative step (the Larmor radius) is tiny and bhat swings across the
    # 2-au zones, so div(kappa) of the anisotropic case is a difference quotient that amplifies last-bit
    # differences: a trajectory may exceed 1e-10 there by as much as the oracle differs from itself.
    _, prec = oracle_mod.Oracle(s, oracle_mod.FMA_LIB_PATH).run(n, seed=5, threads=THREADS)
    assert (prec["steps"] == orec["steps"]).all()
    check_parity(st, rec, orec, sim, o, n, sens=sensitivity(orec, prec))


def test_guiding_centre_on_the_grid_fp64_parity(oracle_mod, spiral):
    # the guiding centre takes grad B and grad|B| from the stencil (curl bhat, grad-B drift) and E from
    # the interpolated U and B
    s = cf.gc_cartesian(spiral, t_limit_days=0.5)
    n = 256
    print(f"[grid] guiding centre parity n={n}", flush=True)
    sim = E.Simulation.from_setup(s, precision=E.PREC_FP64, lanes=1024, steps_per_launch=500)
    st, rec = sim.run(n, seed=6)
    o = oracle_mod.Oracle(s)
    _, orec = o.run(n, seed=6, threads=THREADS)
    _, prec = oracle_mod.Oracle(s, oracle_mod.FMA_LIB_PATH).run(n, seed=6, threads=THREADS)
    check_parity(st, rec, orec, sim, o, n, sens=sensitivity(orec, prec))
    # and the mixed kernel finishes every trajectory
    st32, _ = E.Simulation.from_setup(s, precision=E.PREC_MIXED).run(4 * n, seed=6, records=False)
    assert st32.done == 4 * n and st32.failed == 0


def test_small_grid_discards_match(oracle_mod):
    # a +-6 au grid inside an 8 au absorbing sphere: most trajectories leave the grid, each exit is a
    # field error that discards the attempt and retries with fresh draws, up to max_attempts
    g = G.parker_spiral(nb=3, corner_au=6.0)
    s = cf.parker_modulation_cartesian(g, outer_au=8.0, inner_au=0.3)
    s.options["max_attempts"] = 3
    n = 256
    print(f"[grid] small grid n={n}", flush=True)
    sim = E.Simulation.from_setup(s, precision=E.PREC_FP64, lanes=512, steps_per_launch=300)
    st, rec = sim.run(n, seed=9)
    o = oracle_mod.Oracle(s)
    _, orec = o.run(n, seed=9, threads=THREADS)
    print(f"[grid] discards {st.discards} failed {st.failed}", flush=True)
    np.testing.assert_array_equal(rec["status"], orec["status"])
    np.testing.assert_array_equal(rec["attempts"], orec["attempts"])
    np.testing.assert_array_equal(rec["steps"], orec["steps"])
    assert st.discards["field"] == o.discards()["field"] > 0
    ok = rec["status"] == 0
    assert ok.any()
    from test_gpu_parity import endpoint_rel
    assert endpoint_rel(rec[ok], orec[ok]).max() <= 1e-10
    for d in range(sim.n_dist):
        np.testing.assert_array_equal(sim.distro(d)[0], o.distro(d)[0])


def test_files_and_memory_give_the_same_run(tmp_path, spiral):
    pat = str(tmp_path / "parker_25_25_25")
    G.write(pat, spiral)
    s = cf.parker_modulation_cartesian(spiral)
    a = E.Simulation.from_setup(s)
    st_a, rec_a = a.run(300, seed=2)
    b = E.Simulation(s.traj_type, s.time_flow)
    for role, name, blob in s.plugins:
        if role == cf.BACKGROUND:
            b.add_background(name, blob, pat)
        else:
            b.add_plugin(role, name, blob)
    st_b, rec_b = b.run(300, seed=2)
    np.testing.assert_array_equal(rec_a["pos"], rec_b["pos"])
    np.testing.assert_array_equal(a.distro(0)[0], b.distro(0)[0])


def test_served_grid_query_equals_single_run(spiral):
    s = cf.parker_modulation_cartesian(spiral)
    single = E.Simulation.from_setup(s)
    single.run(2000, seed=4, records=False)
    with Server() as srv:
        q = srv.submit(s, 2000, seed=4)
        other = srv.submit(cf.parker_modulation(n_energies=8, t_limit_years=2.0 / 365.0, time_action=0), 2000, seed=5)
        assert srv.wait(q, timeout=120.0) and srv.wait(other, timeout=120.0)
        c, sm, ne = srv.result(q, 0)
    c1, s1, n1 = single.distro(0)
    np.testing.assert_array_equal(c, c1)
    assert ne == n1
    np.testing.assert_array_equal(sm, s1)


def test_mixed_matches_fp64_within_mc_error(spiral):
    # Each trajectory is binned at its observation energy (the same uniform draws in both precisions),
    # weighted by the boundary spectrum at its exit energy; the per-bin mean weight is the modulated
    # spectrum.  The FP32 steps draw different normals, so the two estimates are compared as
    # independent samples: |mean64 - mean32| < 5 sigma in every bin.
    s = cf.parker_modulation_cartesian(spiral, n_bins=20)
    n = 100000
    res = {}
    for prec in (E.PREC_FP64, E.PREC_MIXED):
        sim = E.Simulation.from_setup(s, precision=prec)
        st, rec = sim.run(n, seed=11)
        assert st.done == n and st.failed == 0
        res[prec] = rec
    zmax = 0.0
    for b in range(20):
        w64 = res[E.PREC_FP64]["ev_weight"][res[E.PREC_FP64]["ev_bin"][:, 0] == b, 0]
        w32 = res[E.PREC_MIXED]["ev_weight"][res[E.PREC_MIXED]["ev_bin"][:, 0] == b, 0]
        assert w64.size > 100 and w32.size > 100
        se = np.sqrt(w64.var() / w64.size + w32.var() / w32.size)
        zmax = max(zmax, abs(w64.mean() - w32.mean()) / se)
    print(f"[grid] mixed vs fp64 per-bin mean weight: max z {zmax:.2f}", flush=True)
    assert zmax < 5.0
