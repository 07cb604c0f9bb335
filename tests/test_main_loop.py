"""SimulationMaster's task loop on the host (src/simulation.cc:496-519, :699-835), checked without a device:
SetTasks' clamps (a negative count runs nothing), the files MasterFinish writes under DistroFileName's
base name, the messages of MasterStart / MasterFinish with the reference's number formatting, and the
errors for runs that need a device.  The progress dumps of a real run are checked in test_gpu_main_loop.py.
"""
import os

import numpy as np
import pytest

from spectrum_amd import configs as cf
from spectrum_amd import engine as E
from test_distro_files import parse_dump


def _lines(err):
    return [l for l in err.split("\n") if l]


def test_zero_tasks_dump_empty_distributions_and_report(tmp_path, capfd):
    sim = E.Simulation.from_setup(cf.parker_modulation(n_energies=8))
    sim.set_tasks(-5, 3)                        # a negative count becomes 0
    sim.distro_file_name(str(tmp_path / "modul_"))
    st = sim.main_loop(seed=1)
    assert st.done == 0
    err = capfd.readouterr().err
    lines = _lines(err)
    assert lines[0] == "Trajectories left: 0"
    assert lines[1].endswith("info message: Simulation completed")
    # backward run: the extremes keep their initial values -1e300 and 0, times unit_time (3.33564 s)
    assert lines[2] == "Shortest simulated trajectory time = -3.33564e+300 s"
    assert lines[3] == "Longest simulated trajectory time = 0 s"
    assert lines[4].startswith("Time per trajectory integration = ") and lines[4].endswith(" ms")
    assert lines[5].startswith("Total time of this simulation: ") and lines[5].endswith(" seconds.")
    assert len(lines[5]) == len("Total time of this simulation: ") + 20 + len(" seconds.")
    f = parse_dump(tmp_path / "modul_0.out")
    assert f["cls"] == "DistributionSpectrumKineticEnergyPowerLaw" and f["n_events"] == 0
    assert not os.path.exists(tmp_path / "modul_1.out")


def test_forward_extremes_and_default_file_name(tmp_path, capfd, monkeypatch):
    monkeypatch.chdir(tmp_path)
    s = cf.parker_modulation(n_energies=4)
    s.time_flow = cf.FORWARD
    sim = E.Simulation.from_setup(s)
    sim.set_tasks(0, 0)
    sim.main_loop()
    lines = _lines(capfd.readouterr().err)
    assert lines[2] == "Shortest simulated trajectory time = 3.33564e+300 s"
    assert os.path.exists(tmp_path / "distribution_0.out")


def test_runs_that_need_the_device_report_it(tmp_path):
    sim = E.Simulation.from_setup(cf.parker_modulation(n_energies=8))
    sim.set_tasks(100, 10)
    sim.distro_file_name(str(tmp_path / "d_"))
    with pytest.raises(E.SpxError, match="device error"):
        sim.main_loop()
    sim.set_tasks(0, 1)
    sim.restore_distro_at_start(0)
    sim.main_loop()                             # no file yet: the distribution starts empty, as Restore leaves it
    assert os.path.exists(tmp_path / "d_0.out")   # written by that loop's MasterFinish
    with pytest.raises(E.SpxError, match="device error"):
        sim.main_loop()                         # restoring it reads into device histograms
    with pytest.raises(E.SpxError, match="bad argument"):
        sim.restore_distro_at_start(1)          # one distribution only
