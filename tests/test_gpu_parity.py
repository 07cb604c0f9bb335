"""Parity of the HIP engine (through the C ABI) with the FP64 CPU oracle.

Criteria (FP64 kernel): identical Wiener increments on both sides (counter-based
Philox streams keyed by trajectory id, or replayed pre-generated N(0,1) arrays),
then
  * per-trajectory step counts and exit status identical,
  * endpoints (t, x, p) within 1e-10 relative,
  * distribution counts identical (bit-exact binning), bin sums within 1e-10 relative.
The residual differences come only from the device and host libm returning
different last bits for acos/pow/exp/tanh; every other operation follows the
reference's expression order with contraction disabled.
"""
import os

import numpy as np
import pytest

from spectrum_amd import configs as cf
from spectrum_amd import engine as E

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)
ENDPOINT_RTOL = 1e-10


def rel_diff(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.maximum(np.abs(b), 1e-300)
    return np.abs(a - b) / scale


def endpoint_rel(rec, orec):
    dp = np.linalg.norm(rec["pos"] - orec["pos"], axis=1) / np.linalg.norm(orec["pos"], axis=1).clip(1e-300)
    dm = np.linalg.norm(rec["mom"] - orec["mom"], axis=1) / np.linalg.norm(orec["mom"], axis=1).clip(1e-300)
    dt = np.abs(rec["t"] - orec["t"]) / np.maximum(np.abs(orec["t"]), np.abs(orec["t"] - orec["t0"])).clip(1e-300)
    return np.maximum(np.maximum(dp, dm), dt)


def sensitivity(orec, prec):
    """Per-trajectory distance between the oracle and its last-bit-perturbed build (multiply-adds
    contracted): how strongly each trajectory amplifies rounding differences."""
    return endpoint_rel(prec, orec)


def check_parity(sim_stats, rec, orec, sim, o, n, ev_weight_atol=0.0, sens=None):
    """`ev_weight_atol` bounds event weights that are ill-conditioned functions of the endpoint
    (a squared position component near zero): an (n, 4) array or a scalar, added to the relative test.
    `sens` (per trajectory, from `sensitivity`) lets a trajectory that amplifies last-bit differences
    exceed the 1e-10 endpoint bound by as much as 100 times the oracle's own disagreement with its
    perturbed build; steps, statuses and bins stay exact."""
    print(f"[parity] n={n} steps={int(rec['steps'].sum())} discards={sim_stats.discards}", flush=True)
    assert sim_stats.done == n and sim_stats.failed == 0
    np.testing.assert_array_equal(rec["status"], orec["status"])
    np.testing.assert_array_equal(rec["steps"], orec["steps"])
    np.testing.assert_array_equal(rec["attempts"], orec["attempts"])
    np.testing.assert_array_equal(rec["n_events"], orec["n_events"])
    np.testing.assert_array_equal(rec["ev_bin"], orec["ev_bin"])
    r = endpoint_rel(rec, orec)
    tol = np.full(n, ENDPOINT_RTOL) if sens is None else np.maximum(ENDPOINT_RTOL, 100.0 * sens)
    print(f"[parity] endpoint rel max {r.max():.3e} median {np.median(r):.3e}; above 1e-10: {int((r > ENDPOINT_RTOL).sum())} "
          f"(their oracle sensitivity {'-' if sens is None else np.array2string(sens[r > ENDPOINT_RTOL], precision=2)})", flush=True)
    assert (r <= tol).all()
    dw = np.abs(rec["ev_weight"] - orec["ev_weight"])
    wtol = tol[:, None] * np.abs(orec["ev_weight"])
    assert (dw <= wtol + ev_weight_atol).all(), dw.max()
    assert sim_stats.steps_done == int(orec["steps"].sum())
    for d in range(sim.n_dist):
        c, s, ne = sim.distro(d)
        oc, osum, one = o.distro(d)
        np.testing.assert_array_equal(c, oc)
        assert ne == one
        atol = 1e-300
        if sens is not None:   # one event per trajectory and distribution: a bin sum inherits its trajectories' bounds
            extra = np.zeros(osum.shape[0])
            hit = orec["ev_bin"][:, d] >= 0
            np.add.at(extra, orec["ev_bin"][hit, d], (tol[hit] - ENDPOINT_RTOL) * np.abs(orec["ev_weight"][hit, d]))
            atol = (extra if osum.ndim == 1 else extra[:, None]) + 1e-300
        assert (np.abs(s - osum) <= ENDPOINT_RTOL * np.abs(osum) + atol).all()


@pytest.fixture(scope="module")
def device():
    info = E.device_info()
    print("[device]", info, flush=True)
    assert info, "no HIP device visible to the engine"
    return info


@pytest.mark.parametrize("flow_days", [2.0, 20.0])
def test_parker_modulation_fp64_counter(device, oracle_mod, flow_days):
    s = cf.parker_modulation(n_energies=16, t_limit_years=flow_days / 365.0, time_action=0)
    n = 512
    sim = E.Simulation.from_setup(s, precision=E.PREC_FP64, lanes=1024, steps_per_launch=300)
    st, rec = sim.run(n, seed=7)
    o = oracle_mod.Oracle(s)
    _, orec = o.run(n, seed=7, threads=THREADS)
    check_parity(st, rec, orec, sim, o, n)


def test_parker_modulation_fp64_full_trajectories(device, oracle_mod):
    """Whole backward trajectories to the 90 au sphere (the reference problem), 48 trajectories of ~1.7e5
    steps.  Over that many steps some trajectories amplify last-bit differences: the oracle's own build
    with contracted multiply-adds moves them by up to 8e-11, so each trajectory is held to
    max(1e-10, 100 x its spread) while steps, statuses and bins stay exact."""
    s = cf.parker_modulation(n_energies=16)
    n = 48
    sim = E.Simulation.from_setup(s, precision=E.PREC_FP64, lanes=256, steps_per_launch=4000)
    st, rec = sim.run(n, seed=1)
    o = oracle_mod.Oracle(s)
    _, orec = o.run(n, seed=1, threads=THREADS)
    _, prec = oracle_mod.Oracle(s, oracle_mod.FMA_LIB_PATH).run(n, seed=1, threads=THREADS)
    check_parity(st, rec, orec, sim, o, n, sens=sensitivity(orec, prec))


def test_parker_modulation_fp64_replay(device, oracle_mod):
    """Identical pre-generated N(0,1) increments and uniforms fed to both paths."""
    s = cf.parker_modulation(n_energies=8, t_limit_years=1.0 / 365.0, time_action=0)
    n = 256
    rng = np.random.default_rng(2024)
    normals = rng.standard_normal((n, 15000))      # 3 draws per step; a 1-day trajectory takes < 4000 steps
    uniforms = rng.random((n, 8))
    sim = E.Simulation.from_setup(s, precision=E.PREC_FP64, lanes=512, steps_per_launch=250)
    st, rec = sim.run(n, normals=normals, uniforms=uniforms)
    o = oracle_mod.Oracle(s)
    _, orec = o.run(n, normals=normals, uniforms=uniforms, threads=THREADS)
    check_parity(st, rec, orec, sim, o, n)


def test_dsa_smooth_shock_fp64_counter(device, oracle_mod):
    """Forward ParkerSource at the smooth shock: injection source term, plane absorber, momentum bins."""
    s = cf.dsa_shock(t_limit_days=1.0, src_ratio=1.2)
    n = 192
    sim = E.Simulation.from_setup(s, precision=E.PREC_FP64, lanes=512, steps_per_launch=2000)
    st, rec = sim.run(n, seed=11)
    o = oracle_mod.Oracle(s)
    _, orec = o.run(n, seed=11, threads=THREADS)
    check_parity(st, rec, orec, sim, o, n)
    assert (rec["amp"] != 0).any()


@pytest.mark.parametrize("prec", [E.PREC_FP64, E.PREC_MIXED])
@pytest.mark.parametrize("case", ["parker", "gc"])
def test_results_do_not_depend_on_launch_shape(device, prec, case):
    """Counter RNG is keyed by trajectory id: lanes, block, steps per launch and graph capture change nothing.
    Launches of 97 and 7 steps stop streams at every offset inside a Philox block (3 normals per step), so a
    resumed stream regenerates the block's tail."""
    if case == "parker":
        s = cf.parker_modulation(n_energies=8, t_limit_years=2.0 / 365.0, time_action=0)
        n = 1000
    else:
        s = cf.gc_diff_scatt(t_limit_days=0.05)
        n = 256
    runs = []
    for opts in (dict(lanes=256, steps_per_launch=97, block=64), dict(lanes=4096, steps_per_launch=1000, block=256),
                 dict(lanes=512, steps_per_launch=200, block=128, graph_launches=4), dict(lanes=256, steps_per_launch=7, block=256)):
        sim = E.Simulation.from_setup(s, precision=prec, **opts)
        st, rec = sim.run(n, seed=5)
        runs.append((st, rec, sim.distro(0)))
    for st, rec, (c, sm, ne) in runs[1:]:
        np.testing.assert_array_equal(rec, runs[0][1])
        np.testing.assert_array_equal(c, runs[0][2][0])
        np.testing.assert_array_equal(sm, runs[0][2][1])   # exact fixed-point sums: no order dependence
        assert st.steps_done == runs[0][0].steps_done


def uniform_no_libm_setup(B=(3.0, 4.0, 12.0)):
    """Parker in a uniform oblique field with a flow, constant kappa_par != kappa_perp, a reflecting sphere,
    an absorbing plane and a time limit, binned by exit time: every operation of its step is +, -, *, /,
    sqrt or a comparison (no pow, exp, log or trigonometric function, with the increments replayed)."""
    s = cf.Setup("uniform_no_libm", cf.TRAJ_PARKER, cf.FORWARD)
    s.add(cf.BACKGROUND, "BackgroundUniform", cf.background_base(cf.DataContainer(), 0.0, (0.0, 0.0, 0.0), (0.002, 0.001, -0.0005),
                                                                 B, 5.0))
    s.add(cf.DIFFUSION, "DiffusionFullConstant", cf.DataContainer().insert_double(0.3).insert_double(2.5))
    s.add(cf.INITIAL, "InitialTimeFixed", cf.DataContainer().insert_double(0.0))
    s.add(cf.INITIAL, "InitialSpaceFixed", cf.DataContainer().insert_vec((0.5, -0.25, 1.0)))
    s.add(cf.INITIAL, "InitialMomentumFixed", cf.DataContainer().insert_vec((cf.mom_from_ekin(cf.MeV), 0.0, 0.0)))
    s.add(cf.BOUNDARY, "BoundarySphereReflect", cf.boundary(-1, [-1], ("vec", (0.0, 0.0, 0.0)), ("double", 30.0)))
    s.add(cf.BOUNDARY, "BoundaryPlaneAbsorb", cf.boundary(1, [0], ("vec", (0.0, 0.0, 20.0)), ("vec", (0.0, 0.6, 0.8))))
    s.add(cf.BOUNDARY, "BoundaryTimeExpire", cf.boundary(1, [0], ("double", 900.0)))
    d = cf.distribution_base(cf.DataContainer(), (12, 0, 0), (0.0, 0.0, 0.0), (960.0, 0.0, 0.0), (0, 0, 0), (0, 0, 0))
    d.insert_double(1.0).insert_double(0.5).insert_int(1)
    s.add(cf.DISTRIBUTION, "DistributionTimeUniform", d)
    return s


def test_fp64_bit_identical_without_transcendentals(device, oracle_mod):
    """Where the step uses only +, -, *, /, sqrt and comparisons (uniform_no_libm_setup, increments replayed), the
    FP64 kernel rounds every operation as the oracle does (the oracle is built without contraction, every
    operation of the kernel's reference order is correctly rounded): endpoints, steps, bins and histogram sums
    are identical bit for bit.  This pins the kernel's divisions through shared reciprocals and its square
    roots to the host's correctly rounded ones over every step of every trajectory."""
    s = uniform_no_libm_setup()
    n = 512
    rng = np.random.default_rng(77)
    normals = rng.standard_normal((n, 6000))
    uniforms = rng.random((n, 8))
    sim = E.Simulation.from_setup(s, precision=E.PREC_FP64, lanes=512, steps_per_launch=300)
    st, rec = sim.run(n, normals=normals, uniforms=uniforms)
    o = oracle_mod.Oracle(s)
    _, orec = o.run(n, normals=normals, uniforms=uniforms, threads=THREADS)
    assert st.done == n and st.failed == 0
    print(f"[bitexact] steps {int(rec['steps'].sum())}, statuses {np.unique(rec['status'])}, events {np.bincount(rec['n_events'])}", flush=True)
    for f in orec.dtype.names:
        np.testing.assert_array_equal(rec[f], orec[f], err_msg=f)
    c, sm, ne = sim.distro(0)
    oc, osm, one = o.distro(0)
    np.testing.assert_array_equal(c, oc)
    np.testing.assert_array_equal(sm, osm)


def test_zero_field_is_a_field_error_in_both_precisions(oracle_mod):
    """|B| below sp_tiny is an invalid field (src/background_base.cc:447): in a uniform zero field every
    start attempt is discarded as a field error and each trajectory is abandoned after max_attempts, in the
    oracle, the FP64 kernel and the mixed one (whose FP32 |B| = b2 rsq(b2) is 0 * inf = NaN there)."""
    s = uniform_no_libm_setup(B=(0.0, 0.0, 0.0))
    s.options["max_attempts"] = 3
    n = 64
    o = oracle_mod.Oracle(s)
    _, orec = o.run(n, seed=3, threads=THREADS)
    assert (orec["status"] != 0).all() and (orec["attempts"] == 3).all()
    for prec in (E.PREC_FP64, E.PREC_MIXED):
        sim = E.Simulation.from_setup(s, precision=prec, lanes=256, steps_per_launch=100)
        st, rec = sim.run(n, seed=3)
        print(f"[zero field] precision {prec}: discards {st.discards} failed {st.failed}", flush=True)
        np.testing.assert_array_equal(rec["status"], orec["status"])
        np.testing.assert_array_equal(rec["attempts"], orec["attempts"])
        assert st.discards["field"] == o.discards()["field"] == 3 * n


def f32_counter_normals(oracle_mod, seed, traj, count):
    """Normal draws 0..count-1 of a trajectory's FP32-step counter stream, restated on the host from the
    stream's definition: draw k is word k % 4 of Philox4x32-10 block k / 4 (key = seed, counter = (k / 4,
    2 << 28, traj)); Box-Muller on the word pairs (w0, w1) and (w2, w3), cosine branch first, with the radius
    uniform (float(w) + 1/2) 2^-32 and the angle uniform ((w >> 8) + 1) 2^-24."""
    out = np.empty(4 * ((count + 3) // 4))
    key = (seed & 0xFFFFFFFF, seed >> 32)
    for b in range(len(out) // 4):
        w = oracle_mod.philox((b & 0xFFFFFFFF, ((b >> 32) & 0x0FFFFFFF) | (2 << 28), traj & 0xFFFFFFFF, traj >> 32), key)
        for h in range(2):
            a, c = int(w[2 * h]), int(w[2 * h + 1])
            u1 = float(np.float32(float(np.float32(a)) * 2.0 ** -32 + 2.0 ** -33))
            u2 = ((c >> 8) + 1) * 2.0 ** -24
            r, ang = np.sqrt(-2.0 * np.log(u1)), 2.0 * np.pi * u2
            out[4 * b + 2 * h], out[4 * b + 2 * h + 1] = r * np.cos(ang), r * np.sin(ang)
    return out[:count]


def counter_uniforms(oracle_mod, seed, traj, count):
    """Uniform draws 0..count-1 of a trajectory's counter stream: the 53-bit uniform of words (w0, w1) of
    Philox block k with the uniform tag (1 << 28) in the counter."""
    key = (seed & 0xFFFFFFFF, seed >> 32)
    out = []
    for k in range(count):
        w = oracle_mod.philox((k, 1 << 28, traj & 0xFFFFFFFF, traj >> 32), key)
        x = (int(w[1]) << 32) | int(w[0])
        out.append(((x >> 11) + 1) * 2.0 ** -53)
    return np.array(out)


def test_mixed_counter_stream_follows_its_definition(device, oracle_mod):
    """The FP32 step's counter normals are the stream restated above, in order: the same run with those
    draws replayed from the host ends at the same points (within the FP32 step's rounding and its fast
    log / sincos), while draws taken from the wrong word, block or branch would move every endpoint by the
    size of its displacement.  Launches of 5 steps resume streams at every offset inside a block."""
    s = cf.parker_modulation(n_energies=4, t_limit_years=0.5 / 365.0, time_action=0)
    n, seed = 64, 9
    sim = E.Simulation.from_setup(s, precision=E.PREC_MIXED, lanes=256, steps_per_launch=5)
    st, rec = sim.run(n, seed=seed)
    assert st.done == n and st.failed == 0
    count = 3 * int(rec["steps"].max()) + 16
    normals = np.stack([f32_counter_normals(oracle_mod, seed, i, count) for i in range(n)])
    uniforms = np.stack([counter_uniforms(oracle_mod, seed, i, 8) for i in range(n)])
    sim2 = E.Simulation.from_setup(s, precision=E.PREC_MIXED, lanes=256, steps_per_launch=1000)
    st2, rec2 = sim2.run(n, normals=normals, uniforms=uniforms)
    assert st2.done == n and st2.failed == 0
    np.testing.assert_array_equal(rec2["status"], rec["status"])
    np.testing.assert_array_equal(rec2["pos0"], rec["pos0"])
    same = rec2["steps"] == rec["steps"]
    disp = np.linalg.norm(rec["pos"] - rec["pos0"], axis=1)
    d = np.linalg.norm(rec2["pos"] - rec["pos"], axis=1) / disp
    print(f"[stream] steps equal {int(same.sum())}/{n}; endpoint difference / displacement: median {np.median(d[same]):.2e} "
          f"max {d[same].max():.2e}; mean displacement {disp.mean() / cf.one_au:.3f} au", flush=True)
    assert same.mean() >= 0.95
    assert np.all(d[same] < 1e-3)


@pytest.mark.parametrize("axis", [(0.0, 0.0, 1.0), (0.0, 0.0, -1.0), (0.0, 0.6, 0.8)])
def test_mixed_guiding_centre_follows_fp64_in_every_basis(device, axis):
    """The guiding centre in the spiral with every Wiener increment zero (replayed) and the pitch cosines
    spread over (-1, 1): streaming, the drifts and the mirror force (which turns the inward-moving
    particles within 0.02 day) are deterministic.  The FP32 step evaluates the spiral in the global
    frame (diagonal bases) or through the rotation (a tilted axis); both must follow the FP64 kernel."""
    s = cf.gc_diff_scatt(t_limit_days=0.02, sw=cf.solarwind(axis=axis))
    n = 64
    normals = np.zeros((n, 3000))
    uniforms = np.tile(np.linspace(0.02, 0.98, n)[:, None], (1, 8))
    recs = {}
    for prec in (E.PREC_FP64, E.PREC_MIXED):
        sim = E.Simulation.from_setup(s, precision=prec, lanes=256, steps_per_launch=500)
        st, rec = sim.run(n, normals=normals, uniforms=uniforms)
        assert st.done == n and st.failed == 0
        recs[prec] = rec
    a, b = recs[E.PREC_FP64], recs[E.PREC_MIXED]
    np.testing.assert_array_equal(b["status"], a["status"])
    same = b["steps"] == a["steps"]
    dx = np.linalg.norm(b["pos"] - a["pos"], axis=1) / cf.one_au
    dm = np.linalg.norm(b["mom"] - a["mom"], axis=1) / np.linalg.norm(a["mom"], axis=1)
    turned = np.sign(a["mom"][:, 2]) != np.sign(a["mom0"][:, 2])
    print(f"[drift-gc] axis {axis}: steps equal {int(same.sum())}/{n}; mirrored {int(turned.sum())}; FP32 - FP64 endpoint max "
          f"{dx[same].max():.2e} au, momentum {dm[same].max():.2e}", flush=True)
    assert same.mean() >= 0.9 and turned.any()
    assert np.all(dx[same] < 1e-4)
    assert np.all(dm[same] < 1e-4)


@pytest.mark.parametrize("axis", [(0.0, 0.0, 1.0), (0.0, 0.0, -1.0), (0.0, 0.6, 0.8)])
def test_mixed_drift_follows_fp64_in_every_basis(device, axis):
    """With every Wiener increment zero (replayed), a backward Parker step moves by dt (div kappa - U) and
    changes p by the adiabatic term: the drift alone, about 1 au in 0.1 day at 10 GeV.  The FP32 step takes
    div U from the radial wind's closed-form stencil (any basis) and div kappa from the spiral's closed-form
    b b stencil when the rotation axis is along +z or -z (the winding sense in the global frame flips with
    the axis), and from the generic stencil of the rotated field otherwise; the FP64 kernel evaluates the
    reference's numerical derivatives of the field itself.  Both must end at the same points to FP32
    accuracy; a wrong winding sense mirrors the azimuthal drift."""
    s = cf.parker_modulation(n_energies=16, t_limit_years=0.1 / 365.0, time_action=0, sw=cf.solarwind(axis=axis))
    n = 64
    normals = np.zeros((n, 3000))
    uniforms = np.full((n, 8), 0.5)
    recs = {}
    for prec in (E.PREC_FP64, E.PREC_MIXED):
        sim = E.Simulation.from_setup(s, precision=prec, lanes=256, steps_per_launch=500)
        st, rec = sim.run(n, normals=normals, uniforms=uniforms)
        assert st.done == n and st.failed == 0
        recs[prec] = rec
    a, b = recs[E.PREC_FP64], recs[E.PREC_MIXED]
    np.testing.assert_array_equal(b["status"], a["status"])
    same = b["steps"] == a["steps"]
    disp = np.linalg.norm(a["pos"] - a["pos0"], axis=1)
    dx = np.linalg.norm(b["pos"] - a["pos"], axis=1) / disp
    dp = np.abs(b["mom"][:, 0] - a["mom"][:, 0]) / np.abs(a["mom"][:, 0] - a["mom0"][:, 0])
    print(f"[drift] axis {axis}: steps equal {int(same.sum())}/{n}; displacement {disp.min() / cf.one_au:.3f}-{disp.max() / cf.one_au:.3f} au; "
          f"FP32 - FP64 endpoint / displacement max {dx[same].max():.2e}; momentum change max {dp[same].max():.2e}", flush=True)
    assert same.mean() >= 0.9
    assert np.all(dx[same] < 1e-3)
    assert np.all(dp[same] < 1e-3)


def test_first_offset_selects_the_same_streams(device):
    """Trajectories first..first+n-1 of a larger run are reproduced by a run started at `first`."""
    s = cf.parker_modulation(n_energies=4, t_limit_years=1.0 / 365.0, time_action=0)
    sim = E.Simulation.from_setup(s, lanes=256, steps_per_launch=400)
    _, rec = sim.run(300, seed=3)
    sim2 = E.Simulation.from_setup(s, lanes=256, steps_per_launch=400)
    _, rec2 = sim2.run(100, seed=3, first=120)
    np.testing.assert_array_equal(rec2, rec[120:220])


def test_full_size_modulation_binning(device):
    """1e7 trajectories (64 energies x 156250) with a one-day limit binned at expiry: each energy bin gets
    exactly its 156250 trajectories, and every trajectory ends at the time boundary."""
    n_e, per = 64, 156250
    s = cf.parker_modulation(n_energies=n_e, t_limit_years=1.0 / 365.0, inner_au=None, time_action=0)
    sim = E.Simulation.from_setup(s, precision=E.PREC_FP64)
    st, _ = sim.run(n_e * per, seed=1, records=False)
    print(f"[full] steps {st.steps_done} kernel_ms {st.kernel_ms:.1f} launches {st.launches}", flush=True)
    assert st.done == n_e * per and st.failed == 0
    c, sm, ne = sim.distro(0)
    assert ne == n_e * per
    np.testing.assert_array_equal(c, np.full(n_e, per))
    assert st.t_min == pytest.approx(st.t_max, rel=1e-6)
    assert np.all(sm > 0)


def test_gc_diff_scatt_fp64_counter(device, oracle_mod):
    """Guiding centre with perpendicular diffusion and pitch-angle scattering in the Parker spiral
    (numerical grad B, curl b, E x B and gradient drifts, rejection-sampled pitch-angle steps)."""
    s = cf.gc_diff_scatt(t_limit_days=0.3)
    n = 256
    sim = E.Simulation.from_setup(s, precision=E.PREC_FP64, lanes=512, steps_per_launch=1500)
    st, rec = sim.run(n, seed=13)
    o = oracle_mod.Oracle(s)
    _, orec = o.run(n, seed=13, threads=THREADS)
    check_parity(st, rec, orec, sim, o, n)


def test_gc_uniform_ring_fp64_counter(device, oracle_mod):
    s = cf.gc_uniform(theta0=0.6435011087932844, t_limit=300.0)
    n = 512
    sim = E.Simulation.from_setup(s, precision=E.PREC_FP64, lanes=512, steps_per_launch=700)
    st, rec = sim.run(n, seed=17)
    o = oracle_mod.Oracle(s)
    _, orec = o.run(n, seed=17, threads=THREADS)
    check_parity(st, rec, orec, sim, o, n)


def test_autotune_picks_a_shape_and_keeps_results(device):
    """The autotuner times 18 launch shapes and the best one as a HIP graph on a probe run, keeps the fastest, and leaves the results
    of later runs unchanged (counter streams are keyed by trajectory id, not by slot)."""
    s = cf.parker_modulation(n_energies=8, t_limit_years=2.0 / 365.0, time_action=0)
    ref = E.Simulation.from_setup(s, precision=E.PREC_MIXED)
    _, rec_ref = ref.run(2000, seed=9)
    sim = E.Simulation.from_setup(s, precision=E.PREC_MIXED)
    table = sim.autotune(n_probe=2_000_000, seed=1)
    print("[autotune]", [(l, k, f"{v:.3e}", g) for l, k, v, g in table], flush=True)
    assert len(table) == 19 and all(v > 0 for _, _, v, _ in table)
    assert [g for *_, g in table] == [0] * 18 + [2]       # the last probe replays the best shape as a graph
    c, sm, ne = sim.distro(0)
    assert ne == 0 and c.sum() == 0            # the probe's events were discarded
    _, rec = sim.run(2000, seed=9)
    np.testing.assert_array_equal(rec, rec_ref)


@pytest.mark.parametrize("case", ["parker", "gc"])
def test_step_scale_fp64_counter(device, oracle_mod, case):
    """The step-scale option (the physical time step times 1/4, for convergence studies) changes the
    kernel and the oracle alike: Parker modulation and the guiding centre with pitch-angle scattering."""
    if case == "parker":
        s = cf.parker_modulation(n_energies=8, t_limit_years=1.0 / 365.0, time_action=0)
        n, lanes, spl = 256, 512, 500
    else:
        s = cf.gc_uniform(theta0=0.6435011087932844, t_limit=100.0)
        n, lanes, spl = 256, 512, 700
    sim = E.Simulation.from_setup(s, precision=E.PREC_FP64, lanes=lanes, steps_per_launch=spl, step_scale=0.25)
    st, rec = sim.run(n, seed=29)
    o = oracle_mod.Oracle(s)
    o.set_option("step_scale", 0.25)
    _, orec = o.run(n, seed=29, threads=THREADS)
    check_parity(st, rec, orec, sim, o, n)
    o1 = oracle_mod.Oracle(s)
    _, orec1 = o1.run(n, seed=29, threads=THREADS)
    assert rec["steps"].sum() > 3 * orec1["steps"].sum()     # the smaller step took effect
