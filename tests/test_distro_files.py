"""Distribution files in the reference's formats (DistributionTemplated::Dump / Restore / Print1D / Print2D,
src/distribution_templated.cc:394-632), checked on the CPU on freshly configured simulations (empty
histograms): the binary layout field by field, the text layout, bin centres (log-centred for log bins),
and Restore's refusals.  The values after a GPU run are checked in test_gpu_distro_files.py.
"""
import struct

import numpy as np
import pytest

import feature_setups as F
from spectrum_amd import configs as cf
from spectrum_amd import engine as E


def parse_dump(path):
    """Read a Dump file the way DistributionTemplated::Restore does."""
    b = open(path, "rb").read()
    o = 0
    def take(fmt):
        nonlocal o
        v = struct.unpack_from(fmt, b, o)
        o += struct.calcsize(fmt)
        return v
    (n,) = take("<I")
    cls = b[o:o + n].decode(); o += n
    (specie,) = take("<B")
    size, length = take("<II")
    container = b[o:o + length]; o += length
    (n_events,) = take("<i")
    return {"cls": cls, "specie": specie, "size": size, "container": container, "n_events": n_events, "rest": b[o:]}


def records(container):
    n, i = 0, 0
    while i < len(container):
        i += 1 + container[i]
        n += 1
    return n


def test_dump_layout(tmp_path):
    s = cf.parker_modulation(n_energies=8)
    sim = E.Simulation.from_setup(s)
    sim.dump_distro(0, str(tmp_path / "d.out"))
    f = parse_dump(tmp_path / "d.out")
    blob = [p for p in s.plugins if p[0] == cf.DISTRIBUTION][0][2]
    assert f["cls"] == "DistributionSpectrumKineticEnergyPowerLaw" and f["specie"] == 0
    assert f["container"] == blob and f["size"] == records(blob)
    assert f["n_events"] == 0
    rest = f["rest"]
    nb = 8
    counts = np.frombuffer(rest[:4 * nb], dtype="<i4")
    sums = np.frombuffer(rest[4 * nb:12 * nb], dtype="<f8")
    tail = struct.unpack("<II", rest[12 * nb:])
    assert (counts == 0).all() and (sums == 0).all() and tail == (0, 0)


def test_dump_width_of_vector_and_matrix_distributions(tmp_path):
    s = F.smoothshock_perp_momentum()          # CumulativeOrder2 (9 doubles per bin) + MomentumUniform
    sim = E.Simulation.from_setup(s)
    sim.dump_distro(0, str(tmp_path / "m.out"))
    f = parse_dump(tmp_path / "m.out")
    assert f["cls"] == "DistributionPositionCumulativeOrder2"
    assert len(f["rest"]) == 6 * 4 + 6 * 9 * 8 + 8


def test_print1d_layout_and_bin_centres(tmp_path):
    s = cf.parker_modulation(n_energies=8)
    sim = E.Simulation.from_setup(s)
    for phys in (False, True):
        p = tmp_path / f"p{int(phys)}.dat"
        sim.print_distro_1d(0, 0, str(p), phys)
        lines = open(p).read().split("\n")
        assert lines[0] == "# Total number of events: 0" and lines[1] == ""
        rows = [l for l in lines[2:] if l]
        assert len(rows) == 8 and all(len(r) == 4 * 20 for r in rows)
        lo, hi = cf.spectrum_bins(0.1 * cf.GeV, 10.0 * cf.GeV, 8)
        l0, l1 = np.log10(lo), np.log10(hi)
        cent = 10.0 ** (l0 + (np.arange(8) + 0.5) * (l1 - l0) / 8) * (cf.unit_energy if phys else 1.0)
        got = np.array([float(r[:20]) for r in rows])
        np.testing.assert_allclose(got, cent, rtol=1e-9)
        assert all(r[20:].split() == ["0", "0", "0"] for r in rows)


def test_print2d_layout(tmp_path):
    s = F.gc_isotropic_only()                  # PositionPitchUniform, 16 x 16 bins
    sim = E.Simulation.from_setup(s)
    p = tmp_path / "p2.dat"
    sim.print_distro_2d(0, 0, 1, str(p))
    rows = [l for l in open(p).read().split("\n")[2:] if l]
    assert len(rows) == 16 * 16 and all(len(r) == 5 * 20 for r in rows)
    # ijk1 outer, ijk2 inner
    first = np.array([[float(r[:20]), float(r[20:40])] for r in rows[:17]])
    assert (first[:16, 0] == first[0, 0]).all() and first[16, 0] > first[0, 0]
    assert np.all(np.diff(first[:16, 1]) > 0)
    # equal dimensions fall back to Print1D
    sim.print_distro_2d(0, 1, 1, str(p))
    assert len([l for l in open(p).read().split("\n")[2:] if l]) == 16


def test_restore_refuses_other_classes_and_binnings(tmp_path):
    a = E.Simulation.from_setup(cf.parker_modulation(n_energies=8))
    a.dump_distro(0, str(tmp_path / "a.out"))
    other = E.Simulation.from_setup(F.uniform_para_plane_pass())    # DistributionTimeUniform
    with pytest.raises(E.SpxError, match="holds"):
        other.restore_distro(0, str(tmp_path / "a.out"))
    wider = E.Simulation.from_setup(cf.parker_modulation(n_energies=16))
    with pytest.raises(E.SpxError, match="binning"):
        wider.restore_distro(0, str(tmp_path / "a.out"))
    with pytest.raises(E.SpxError):
        wider.restore_distro(0, str(tmp_path / "missing.out"))
