"""Host-side checks of the C ABI and the DataContainer format; no GPU work is launched.

The DataContainer round trip mirrors the reference's only self-check
(tests/main_test_data_container.cc): insert int, double, MultiIndex, GeoVector,
string, vector<int>, vector<double>; read back in order; compare.
"""
import ctypes as C
import os
import re
import struct

import numpy as np
import pytest

from spectrum_amd import configs as cf
from spectrum_amd import engine
from spectrum_amd.data_container import DataContainer, DataReader

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_data_container_round_trip():
    c = DataContainer()
    c.insert_int(10).insert_double(12.3).insert_multiindex((3, 2, 1)).insert_vec((3.0, 2.0, 1.0))
    c.insert_string("abc").insert_int_vector([10, 20, 30]).insert_double_vector([1.5, 3.5, 5.5])
    assert c.n_records == 7
    r = DataReader(c.bytes())
    assert r.read_int() == 10
    assert r.read_double() == 12.3
    assert r.read_multiindex() == (3, 2, 1)
    assert r.read_vec() == (3.0, 2.0, 1.0)
    assert r.read_string() == "abc"
    assert r.read_int_vector() == [10, 20, 30]
    assert r.read_double_vector() == [1.5, 3.5, 5.5]
    with pytest.raises(ValueError):
        r.read_int()


def test_data_container_wire_bytes():
    """[u8 size][raw object bytes] per record, little endian (common/data_container.cc:133-151)."""
    b = DataContainer().insert_int(10).insert_double(1.0).insert_string("ab").bytes()
    assert b == bytes([4]) + struct.pack("<i", 10) + bytes([8]) + struct.pack("<d", 1.0) + bytes([2]) + b"ab"
    with pytest.raises(ValueError):
        DataContainer().insert_double_vector([0.0] * 32)      # 256 bytes: the u8 size field cannot hold it
    with pytest.raises(ValueError):
        DataContainer.records(bytes([9, 1, 2]))


def header_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"\b(spx_[a-z0-9_]+)\s*\(", txt))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    L = engine.lib()
    names = header_symbols()
    assert len(names) >= 13
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def new_sim(traj=cf.TRAJ_PARKER, flow=cf.BACKWARD):
    L = engine.lib()
    h = C.c_void_p()
    assert L.spx_create(traj, flow, C.byref(h)) == 0
    return L, h


def test_create_rejects_bad_arguments():
    L = engine.lib()
    h = C.c_void_p()
    assert L.spx_create(9, 0, C.byref(h)) == -1
    assert L.spx_create(0, 3, C.byref(h)) == -1
    assert L.spx_create(0, 0, None) == -1


@pytest.mark.parametrize("setup", [cf.parker_modulation(), cf.dsa_shock(), cf.gc_diff_scatt()], ids=lambda s: s.name)
def test_workload_setups_lower_on_host(setup):
    """Every plugin of the measured workloads is accepted from its reference DataContainer bytes."""
    L, h = new_sim(setup.traj_type, setup.time_flow)
    try:
        for role, name, blob in setup.plugins:
            rc = L.spx_add_plugin(h, role, name.encode(), blob, len(blob))
            assert rc == 0, (name, L.spx_last_error(h))
        for k, v in setup.options.items():
            assert L.spx_set_option(h, k.encode(), float(v)) == 0
        nb3 = (C.c_int32 * 3)()
        w = C.c_int32()
        assert L.spx_distro_info(h, 0, nb3, C.byref(w)) == 0
        assert nb3[0] * nb3[1] * nb3[2] > 0 and w.value >= 1
    finally:
        L.spx_destroy(h)


def test_error_codes():
    L, h = new_sim()
    try:
        sw = cf.parker_modulation().plugins[0]
        assert sw[1] == "BackgroundSolarWind"
        # unknown class name (the factory lookup failure of the reference)
        assert L.spx_add_plugin(h, cf.BACKGROUND, b"BackgroundNoSuchThing", sw[2], len(sw[2])) == -2
        assert b"unknown" in L.spx_last_error(h)
        # container too short for the class parameters
        assert L.spx_add_plugin(h, cf.BACKGROUND, b"BackgroundSolarWind", sw[2][:20], 20) == -3
        # unknown role / option / bad option value
        assert L.spx_add_plugin(h, 17, b"BackgroundSolarWind", sw[2], len(sw[2])) == -1
        assert L.spx_set_option(h, b"no_such_option", 1.0) == -1
        assert L.spx_set_option(h, b"block", 100.0) == -1
        # incomplete setup: IsSimmulationReady fails before any device work
        assert L.spx_add_plugin(h, cf.BACKGROUND, b"BackgroundSolarWind", sw[2], len(sw[2])) == 0
        rng = engine.SpxRng()
        assert L.spx_begin(h, 0, 16, C.byref(rng), 0) == -4
        assert b"initial" in L.spx_last_error(h) or b"diffusion" in L.spx_last_error(h)
        # advance / finish without begin
        st = engine.SpxRunStats()
        assert L.spx_advance(h, 1, C.byref(st)) == -6
        assert L.spx_finish(h, None, C.byref(st)) == -6
        # distribution index out of range
        nb3 = (C.c_int32 * 3)()
        w = C.c_int32()
        assert L.spx_distro_info(h, 0, nb3, C.byref(w)) == -1
    finally:
        L.spx_destroy(h)


def test_limits():
    L, h = new_sim()
    try:
        s = cf.parker_modulation()
        bnds = [p for p in s.plugins if p[0] == cf.BOUNDARY]
        n_ok = 0
        for _ in range(3):
            for role, name, blob in bnds:
                rc = L.spx_add_plugin(h, role, name.encode(), blob, len(blob))
                if rc == 0:
                    n_ok += 1
                else:
                    assert rc == -7
        assert n_ok == 4
    finally:
        L.spx_destroy(h)


def test_python_mirror_raises_on_errors():
    sim = engine.Simulation(cf.TRAJ_PARKER, cf.BACKWARD)
    with pytest.raises(engine.SpxError, match="unknown class"):
        sim.add_background("BackgroundNoSuchThing", DataContainer().insert_double(1.0))
    with pytest.raises(engine.SpxError, match="not ready"):
        sim.run(4)


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_begin_without_device_reports_device_error():
    s = cf.parker_modulation(n_energies=2)
    sim = engine.Simulation.from_setup(s)
    with pytest.raises(engine.SpxError, match="device error"):
        sim.run(4)
    assert engine.device_info() == ""


def test_oracle_and_library_read_the_same_bytes(oracle_mod):
    """The oracle consumes the identical DataContainer bytes (configuration parity of the two paths)."""
    s = cf.dsa_shock()
    o = oracle_mod.Oracle(s)
    _, rec = o.run(8, seed=3, threads=1)
    assert (rec["status"] == 0).all()
    p0 = s.meta["p0"]
    np.testing.assert_allclose(rec["mom0"][:, 0], p0, rtol=1e-15)


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_server_without_device_reports_device_error():
    from spectrum_amd.server import Server
    with pytest.raises(engine.SpxError, match="device error"):
        Server()


def test_server_argument_checks():
    from spectrum_amd import server as S
    L = S._bind()
    h = C.c_void_p()
    assert L.spx_server_create(None, None) == -1
    assert L.spx_server_submit(None, None, 0, 1, 1, None) == -1
    assert L.spx_server_wait(None, 1, 0.0) == -1
    assert L.spx_server_result(None, 1, 0, None, None, None) == -1
    L.spx_server_destroy(None)
