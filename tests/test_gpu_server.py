"""The query server: continuous batching of different queries on shared slots, admission by device
memory, and results that are the same alone, in a mix, and through the single-run interface.

Counts, event totals and bin sums must be identical: sums are exact fixed-point accumulations, rounded
once when read, so the order in which the waves of a mix bin their events cannot change them.
"""
import time

import numpy as np
import pytest

from spectrum_amd import configs as cf
from spectrum_amd import engine as E
from spectrum_amd.server import Server

pytestmark = pytest.mark.gpu


def queries():
    return [
        (cf.parker_modulation(n_energies=8, r_start_au=5.0, t_limit_years=3.0 / 365.0, time_action=0, name="mod_r5"), 3000, 11),
        (cf.parker_modulation(n_energies=16, r_start_au=30.0, t_limit_years=3.0 / 365.0, time_action=0, name="mod_r30"), 5000, 12),
        (cf.dsa_shock(t_limit_days=0.5, src_ratio=1.2), 2000, 13),
        (cf.gc_uniform(theta0=0.6, t_limit=200.0), 4000, 14),
    ]


def single_run(setup, n, seed, prec=E.PREC_FP64):
    sim = E.Simulation.from_setup(setup, precision=prec)
    st, _ = sim.run(n, seed=seed, records=False)
    assert st.done == n
    return [sim.distro(d) for d in range(sim.n_dist)]


def same(a, b, what):
    for (c1, s1, n1), (c2, s2, n2) in zip(a, b):
        np.testing.assert_array_equal(c1, c2, err_msg=what)
        assert n1 == n2, what
        np.testing.assert_array_equal(s1, s2, err_msg=what)


def served(srv, setup, qid):
    return [srv.result(qid, d) for d in range(len(srv._shapes[qid]))]


@pytest.mark.parametrize("prec", [E.PREC_FP64, E.PREC_MIXED])
def test_alone_mix_and_single_run_agree(prec):
    qs = queries()
    ref = [single_run(s, n, seed, prec) for s, n, seed in qs]
    # each query alone on a server
    with Server(precision=prec) as srv:
        for (s, n, seed), r in zip(qs, ref):
            qid = srv.submit(s, n, seed=seed)
            assert srv.wait(qid, timeout=120.0)
            same(served(srv, s, qid), r, f"alone {s.name}")
    # all of them at once, twice over with a second set of keys in between
    with Server(precision=prec) as srv:
        ids = []
        for (s, n, seed) in qs:
            ids.append(srv.submit(s, n, seed=seed))
            ids.append(srv.submit(s, n // 2, seed=seed + 1000))
        for qid in ids:
            assert srv.wait(qid, timeout=180.0)
        for k, ((s, n, seed), r) in enumerate(zip(qs, ref)):
            same(served(srv, s, ids[2 * k]), r, f"mix {s.name}")
        st = srv.stats()
        print(f"[server] {st}", flush=True)
        assert st["completed"] == len(ids) and st["pools"] == 3


def test_admission_by_memory_budget():
    s, n, seed = queries()[0]
    with Server(hbm_budget_bytes=1.0) as srv:          # room for one query at a time
        ids = [srv.submit(s, n, seed=seed + i) for i in range(3)]
        for qid in ids:
            assert srv.wait(qid, timeout=120.0)
        info = [srv.info(q) for q in ids]
        print("[server] serial", [(i["t_admit"], i["t_done"]) for i in info], flush=True)
        for a, b in zip(info, info[1:]):
            assert b["t_admit"] >= a["t_done"]
    with Server(hbm_budget_bytes=1e9) as srv:
        ids = [srv.submit(s, n, seed=seed + i) for i in range(3)]
        for qid in ids:
            assert srv.wait(qid, timeout=120.0)
        info = [srv.info(q) for q in ids]
        assert info[2]["t_admit"] < info[0]["t_done"]     # admitted together


def test_continuous_batching_short_query_overtakes_long_one():
    long_q = cf.parker_modulation(n_energies=16, r_start_au=1.0, t_limit_years=40.0 / 365.0, time_action=0, name="long")
    short_q = cf.parker_modulation(n_energies=4, r_start_au=10.0, t_limit_years=0.5 / 365.0, time_action=0, name="short")
    with Server(precision=E.PREC_MIXED) as srv:
        a = srv.submit(long_q, 2000000, seed=1)
        time.sleep(0.3)
        assert srv.info(a)["state"] == "running"      # the short query arrives while the long one runs
        b = srv.submit(short_q, 2000, seed=2)
        assert srv.wait(b, timeout=120.0)
        ib = srv.info(b)
        ia = srv.info(a)
        print(f"[server] long {ia['state']} short latency {ib['latency_s']:.3f}s", flush=True)
        assert ia["state"] in ("running", "done")
        assert srv.wait(a, timeout=300.0)
        ia = srv.info(a)
        assert ib["t_done"] < ia["t_done"] and ib["t_submit"] > ia["t_submit"]
