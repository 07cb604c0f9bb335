"""SimulationMaster's task loop on the device (src/simulation.cc:558-643, :699-835): the progress blocks
and partial dumps of DecrementTrajectoryCount, the final dumps of MasterFinish against a plain run of
the same trajectories, trajectory ids that continue over successive main loops, and RestoreDistro's
resume at the start of a loop.
"""
import re

import numpy as np
import pytest

from spectrum_amd import configs as cf
from spectrum_amd import engine as E
from test_distro_files import parse_dump

pytestmark = pytest.mark.gpu

N, BATCH = 20000, 1000


def _setup():
    return cf.parker_modulation(n_energies=8, t_limit_years=3.0 / 365.0, time_action=0)


def _counts_sums(path, nb):
    f = parse_dump(path)
    rest = f["rest"]
    return f["n_events"], np.frombuffer(rest[:4 * nb], dtype="<i4").astype(np.int64), np.frombuffer(rest[4 * nb:12 * nb], dtype="<f8")


def test_main_loop_progress_and_final_dump(tmp_path, capfd):
    sim = E.Simulation.from_setup(_setup())
    sim.set_tasks(N, BATCH)
    sim.distro_file_name(str(tmp_path / "m_"))
    st = sim.main_loop(seed=11)
    err = capfd.readouterr().err
    print(f"[main_loop] done {st.done} failed {st.failed} launches {st.launches}", flush=True)
    assert st.done == N
    assert err.split("\n")[0] == f"Trajectories left: {N}"
    pct = [int(m) for m in re.findall(r"Trajectories completed: +(\d+) %", err)]
    # whole batches of 1000 of 20000: every reported percentage is a multiple of 5, rising to 100
    assert pct and pct[-1] == 100 and all(b > a for a, b in zip(pct, pct[1:])) and all(p % 5 == 0 for p in pct)
    assert len(re.findall(r"Est\. time to finish: +\d+ seconds", err)) == len(pct)
    assert len(re.findall(r"Time allocated for job: +inf \(No WM\)", err)) == len(pct)
    assert len(re.findall(r"% of time was spent computing trajectories\.", err)) == len(pct)
    short = float(re.search(r"Shortest simulated trajectory time = (\S+) s", err).group(1))
    long_ = float(re.search(r"Longest simulated trajectory time = (\S+) s", err).group(1))
    # backward run: elapsed times are negative, the longest the most negative
    assert long_ <= short < 0
    np.testing.assert_allclose(-long_, st.t_max * cf.unit_time, rtol=1e-5)
    np.testing.assert_allclose(-short, st.t_min * cf.unit_time, rtol=1e-5)
    # the final dump equals a plain run of trajectories 0..N-1 with the same seed
    ref = E.Simulation.from_setup(_setup())
    ref.run(N, seed=11, first=0, records=False)
    c, s, nev = ref.distro(0)
    nev_f, c_f, s_f = _counts_sums(tmp_path / "m_0.out", c.size)
    assert nev_f == nev
    np.testing.assert_array_equal(c_f, c)
    np.testing.assert_array_equal(s_f, s)


def test_main_loops_continue_ids_and_resume(tmp_path, capfd):
    sim = E.Simulation.from_setup(_setup())
    sim.distro_file_name(str(tmp_path / "r_"))
    sim.set_tasks(6000, 500)
    sim.main_loop(seed=5)                       # trajectories 0..5999
    sim.main_loop(seed=5)                       # a fresh distribution of trajectories 6000..11999
    nb = 8
    nev2, c2, s2 = _counts_sums(tmp_path / "r_0.out", nb)
    ref = E.Simulation.from_setup(_setup())
    ref.run(6000, seed=5, first=6000, records=False)
    c, s, nev = ref.distro(0)
    assert nev2 == nev
    np.testing.assert_array_equal(c2, c)
    np.testing.assert_array_equal(s2, s)
    # RestoreDistro: the third loop starts from the second's file and adds trajectories 12000..17999
    sim.restore_distro_at_start(0)
    sim.main_loop(seed=5)
    nev3, c3, s3 = _counts_sums(tmp_path / "r_0.out", nb)
    ref.run(6000, seed=5, first=12000, records=False)
    c, s, nev = ref.distro(0)
    assert nev3 == nev
    np.testing.assert_array_equal(c3, c)
    np.testing.assert_allclose(s3, s, rtol=5e-16, atol=0)   # restored from a file of rounded sums
    capfd.readouterr()
