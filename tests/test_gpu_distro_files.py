"""Distribution files after GPU runs: Print1D read back the way the reference's post-processors read it
(benchmarks/main_postprocess_modulation_cartesian_parker.cc:45-57: two header lines, then value, mean,
sum and count per row), Print2D's collapse, Dump's counts and sums, and checkpoint/resume: a run dumped,
restored into a fresh simulation and continued gives the histograms of the uninterrupted run.
"""
import struct

import numpy as np
import pytest

import feature_setups as F
from spectrum_amd import configs as cf
from spectrum_amd import engine as E
from test_distro_files import parse_dump

pytestmark = pytest.mark.gpu


def read_print1d(path):
    lines = open(path).read().split("\n")
    n_events = int(lines[0].split(":")[1])
    rows = np.array([[float(x) for x in l.split()] for l in lines[2:] if l])
    return n_events, rows


def test_print1d_and_dump_hold_the_run(tmp_path):
    s = cf.parker_modulation(n_energies=16, t_limit_years=5.0 / 365.0, time_action=0)
    sim = E.Simulation.from_setup(s)
    sim.run(20000, seed=3, records=False)
    counts, sums, nev = sim.distro(0)
    print(f"[files] events {nev} binned {counts.sum()}", flush=True)
    assert nev > 0
    sim.print_distro_1d(0, 0, str(tmp_path / "spec.dat"))
    n, rows = read_print1d(tmp_path / "spec.dat")
    assert n == nev
    np.testing.assert_array_equal(rows[:, 3].astype(np.int64), counts)
    np.testing.assert_allclose(rows[:, 2], sums, rtol=1e-9)
    mean = np.where(counts > 0, sums / np.maximum(counts, 1), 0.0)
    np.testing.assert_allclose(rows[:, 1], mean, rtol=1e-9)
    sim.dump_distro(0, str(tmp_path / "spec.out"))
    f = parse_dump(tmp_path / "spec.out")
    assert f["n_events"] == nev
    nb = counts.size
    np.testing.assert_array_equal(np.frombuffer(f["rest"][:4 * nb], dtype="<i4"), counts)
    np.testing.assert_array_equal(np.frombuffer(f["rest"][4 * nb:12 * nb], dtype="<f8"), sums)


def test_print2d_collapses_the_third_dimension(tmp_path):
    s = F.shock_flowmom_reflect()              # PositionMomentumUniform over (x, |p|): 12 x 1 x 10 bins
    sim = E.Simulation.from_setup(s)
    sim.run(4000, seed=5, records=False)
    counts, sums, nev = sim.distro(0)
    c3 = counts.reshape(12, 1, 10)
    s3 = sums.reshape(12, 1, 10)
    sim.print_distro_2d(0, 0, 2, str(tmp_path / "xp.dat"))
    _, rows = read_print1d(tmp_path / "xp.dat")
    assert rows.shape == (120, 5)
    np.testing.assert_array_equal(rows[:, 4].astype(np.int64), c3.sum(axis=1).reshape(-1))
    np.testing.assert_allclose(rows[:, 3], s3.sum(axis=1).reshape(-1), rtol=1e-9, atol=1e-300)
    sim.print_distro_1d(0, 2, str(tmp_path / "p.dat"))
    _, rows = read_print1d(tmp_path / "p.dat")
    np.testing.assert_array_equal(rows[:, 3].astype(np.int64), c3.sum(axis=(0, 1)))


def test_checkpoint_and_resume(tmp_path):
    s = cf.parker_modulation(n_energies=8, t_limit_years=3.0 / 365.0, time_action=0)
    whole = E.Simulation.from_setup(s)
    whole.run(6000, seed=7, first=0, records=False)
    whole.run(6000, seed=7, first=6000, records=False)
    first = E.Simulation.from_setup(s)
    first.run(6000, seed=7, first=0, records=False)
    first.dump_distro(0, str(tmp_path / "ckpt.out"))
    resumed = E.Simulation.from_setup(s)
    resumed.restore_distro(0, str(tmp_path / "ckpt.out"))
    c0, s0, n0 = first.distro(0)
    c1, s1, n1 = resumed.distro(0)
    np.testing.assert_array_equal(c0, c1)
    np.testing.assert_array_equal(s0, s1)
    assert n0 == n1
    resumed.run(6000, seed=7, first=6000, records=False)
    cw, sw, nw = whole.distro(0)
    cr, sr, nr = resumed.distro(0)
    np.testing.assert_array_equal(cw, cr)
    assert nw == nr
    # the checkpoint file holds each sum rounded to a double (the reference's format): resuming adds that
    # rounding once, so the two differ by at most an ulp or two
    np.testing.assert_allclose(sr, sw, rtol=5e-16, atol=0)
