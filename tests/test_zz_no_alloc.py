"""spx_run_static, the run that allocates nothing: its device state lives in a static device array (one per
GPU), so it must give exactly what spx_run gives for the same trajectories -- counts, event totals and the
exact bin sums bit for bit -- on consecutive calls that reuse the array with a smaller input, and leave the
simulation's own distributions alone.  Called through the C ABI."""
import ctypes as C

import numpy as np
import pytest

from spectrum_amd import configs as cf
from spectrum_amd import engine as E

pytestmark = pytest.mark.gpu


def static_run(sim, n, seed, first=0):
    L = E.lib()
    shapes = [sim.distro(d) for d in range(sim.n_dist)]
    nb = [c.size for c, _, _ in shapes]
    nw = [s.size for _, s, _ in shapes]
    counts = np.full(sum(nb), -1, dtype=np.int64)
    sums = np.full(sum(nw), np.nan)
    nev = np.full(max(sim.n_dist, 1), -1, dtype=np.int64)
    st = E.SpxRunStats()
    rc = L.spx_run_static(sim.h, first, n, seed, counts.ctypes.data, sums.ctypes.data, nev.ctypes.data, C.byref(st))
    if rc != 0:
        raise E.SpxError(f"spx_run_static: {E.ERRORS.get(rc, rc)} {L.spx_last_error(sim.h).decode()}")
    out, oc, os_ = [], 0, 0
    for d in range(sim.n_dist):
        out.append((counts[oc:oc + nb[d]], sums[os_:os_ + nw[d]], int(nev[d])))
        oc += nb[d]
        os_ += nw[d]
    return E.RunStats.of(st), out


def reference_run(setup, prec, n, seed, first=0):
    sim = E.Simulation.from_setup(setup, precision=prec)
    st, _ = sim.run(n, seed=seed, first=first, records=False)
    return st, [(c, s.reshape(-1), ne) for c, s, ne in (sim.distro(d) for d in range(sim.n_dist))]


def same(a, b):
    for (c1, s1, n1), (c2, s2, n2) in zip(a, b):
        np.testing.assert_array_equal(c1, c2)
        np.testing.assert_array_equal(s1, s2)
        assert n1 == n2


@pytest.mark.parametrize("prec", [E.PREC_FP64, E.PREC_MIXED])
def test_static_run_equals_spx_run_twice(prec):
    s = cf.parker_modulation(n_energies=8, t_limit_years=2.0 / 365.0, time_action=0)
    sim = E.Simulation.from_setup(s, precision=prec)
    sim.run(700, seed=9, records=False)                       # the simulation's own histograms, to be left alone
    own = [sim.distro(d) for d in range(sim.n_dist)]
    for n, seed, first in ((20000, 5, 0), (3000, 6, 100)):   # the second call reuses the array with a smaller input
        print(f"[no_alloc] prec={prec} n={n} seed={seed} first={first}", flush=True)
        st, got = static_run(sim, n, seed, first)
        ref_st, ref = reference_run(s, prec, n, seed, first)
        assert st.done == n and st.failed == 0 and st.steps_done == ref_st.steps_done
        assert st.steps_executed == ref_st.steps_executed
        same(got, ref)
        assert sum(ne for _, _, ne in got) > 0
    for (c1, s1, n1), (c2, s2, n2) in zip(own, [sim.distro(d) for d in range(sim.n_dist)]):
        np.testing.assert_array_equal(c1, c2)
        np.testing.assert_array_equal(s1, s2)
        assert n1 == n2


def test_static_run_guiding_centre_and_vector_distribution():
    import feature_setups as F
    s = F.solarwind_kinrad_reflect_recurrent()           # a time distribution and a 3-component <dx> distribution
    sim = E.Simulation.from_setup(s)
    st, got = static_run(sim, 512, 3)
    _, ref = reference_run(s, E.PREC_FP64, 512, 3)
    assert st.done == 512
    same(got, ref)
    s = cf.gc_diff_scatt(t_limit_days=0.2)
    sim = E.Simulation.from_setup(s)
    st, got = static_run(sim, 1024, 4)
    _, ref = reference_run(s, E.PREC_FP64, 1024, 4)
    assert st.done == 1024
    same(got, ref)


def test_static_run_limits():
    s = cf.gc_diff_scatt(n_r=128, n_mu=64)                 # 8192 bins > SPX_STATIC_BINS
    sim = E.Simulation.from_setup(s)
    with pytest.raises(E.SpxError, match="LIMIT|limit"):
        static_run(sim, 16, 1)
