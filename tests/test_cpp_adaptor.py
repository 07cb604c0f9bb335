"""The C++ side of the boundary: a driver compiled with g++ against include/spx_simulation_gpu.hpp (the
SimulationGPU adaptor a reference driver uses in place of SimulationWorker) and linked to the engine
library.  It configures the engine through the adaptor's Add* calls from the same class names and
DataContainer bytes as the Python setups.  On the CPU the configuration and the distributions' shapes
must match the Python binding's and a run must report the missing device; on the GPU a run's
distributions must equal the Python binding's bit for bit."""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

from spectrum_amd import configs as cf
from spectrum_amd import engine as E

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    E.lib()                                               # the library is built
    out = str(tmp_path_factory.mktemp("drv") / "driver")
    libdir = os.path.join(ROOT, "spectrum_amd")
    cmd = ["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "native", "driver.cpp"),
           "-L", libdir, "-lspectrum_amd", f"-Wl,-rpath,{libdir}", "-Wl,--allow-shlib-undefined", "-o", out]
    subprocess.run(cmd, check=True)
    return out


def plugin_file(setup, path, **options):
    opts = {**setup.options, **options}
    with open(path, "wb") as f:
        f.write(struct.pack("<iiii", setup.traj_type, setup.time_flow, len(setup.plugins), len(opts)))
        for role, name, blob in setup.plugins:
            f.write(struct.pack("<iI", role, len(name)) + name.encode() + struct.pack("<I", len(blob)) + blob)
        for k, v in opts.items():
            f.write(struct.pack("<I", len(k)) + k.encode() + struct.pack("<d", float(v)))
    return path


def parse(out):
    lines = out.strip().splitlines()
    dists, i = [], 0
    while i < len(lines):
        if lines[i].startswith("distro"):
            _, d, _, nb, _, w, _, ne = lines[i].split()
            nb, w = int(nb), int(w)
            rows = [l.split() for l in lines[i + 1:i + 1 + nb]]
            counts = np.array([int(r[0]) for r in rows], dtype=np.int64)
            sums = np.array([[float(x) for x in r[1:]] for r in rows]).reshape(-1)
            dists.append((counts, sums, int(ne)))
            i += 1 + nb
        else:
            i += 1
    return dists


SETUPS = [lambda: cf.parker_modulation(n_energies=8, t_limit_years=2.0 / 365.0, time_action=0), lambda: cf.gc_diff_scatt(t_limit_days=0.2)]


@pytest.mark.parametrize("k", range(len(SETUPS)))
def test_adaptor_configures_the_engine(driver, tmp_path, k):
    s = SETUPS[k]()
    res = subprocess.run([driver, plugin_file(s, str(tmp_path / "p.bin")), "info"], capture_output=True, text=True)
    assert res.returncode == 0, res.stdout + res.stderr
    got = parse(res.stdout)
    sim = E.Simulation.from_setup(s)
    assert len(got) == sim.n_dist
    for d, (c, sm, ne) in enumerate(got):
        c2, s2, n2 = sim.distro(d)
        assert c.shape == c2.shape and sm.size == s2.size and ne == n2 == 0


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_adaptor_reports_the_missing_device(driver, tmp_path):
    s = SETUPS[0]()
    res = subprocess.run([driver, plugin_file(s, str(tmp_path / "p.bin")), "run", "100", "1"], capture_output=True, text=True)
    assert res.returncode == 3 and "device" in res.stdout


def test_adaptor_rejects_a_bad_class(driver, tmp_path):
    s = SETUPS[0]()
    s.plugins = [(r, n.replace("SphereAbsorb", "SphereAbsorbX"), b) for r, n, b in s.plugins]
    res = subprocess.run([driver, plugin_file(s, str(tmp_path / "p.bin")), "info"], capture_output=True, text=True)
    assert res.returncode == 3 and "unknown" in res.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("k", range(len(SETUPS)))
def test_adaptor_run_equals_the_python_binding(driver, tmp_path, k):
    s = SETUPS[k]()
    n, seed = 4000, 3
    res = subprocess.run([driver, plugin_file(s, str(tmp_path / "p.bin")), "run", str(n), str(seed)], capture_output=True, text=True,
                         timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    got = parse(res.stdout)
    sim = E.Simulation.from_setup(s)
    sim.run(n, seed=seed, records=False)
    for d, (c, sm, ne) in enumerate(got):
        c2, s2, n2 = sim.distro(d)
        np.testing.assert_array_equal(c, c2)
        np.testing.assert_array_equal(sm, s2.reshape(-1))     # %.17g round-trips a double exactly
        assert ne == n2 and ne > 0
