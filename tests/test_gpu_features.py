"""FP64 parity with the oracle over the feature matrix: every background, diffusion, boundary, initial
condition and distribution class the engine lowers, in Parker, ParkerSource and guiding-centre
transport, forward and backward in time.  Same criteria as the modulation parity tests (identical
counter streams; identical step counts, statuses, event bins and histogram counts; endpoints and bin
sums within 1e-10 relative).  The mixed-precision kernel runs the same setups and must finish every
trajectory with its events binned into the same histograms' totals within sampling error.
"""
import os

import numpy as np
import pytest

import feature_setups as F
from spectrum_amd import engine as E
from test_gpu_parity import check_parity, sensitivity

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)
N = 256


@pytest.mark.parametrize("k", range(len(F.ALL)), ids=F.NAMES)
def test_feature_fp64_parity(oracle_mod, k):
    s = F.ALL[k]()
    print(f"[features] {F.NAMES[k]} fp64 n={N}", flush=True)
    sim = E.Simulation.from_setup(s, precision=E.PREC_FP64, lanes=512, steps_per_launch=400)
    st, rec = sim.run(N, seed=21)
    o = oracle_mod.Oracle(s)
    _, orec = o.run(N, seed=21, threads=THREADS)
    sens = None
    if F.NAMES[k] in F.ILL_CONDITIONED:
        _, prec = oracle_mod.Oracle(s, oracle_mod.FMA_LIB_PATH).run(N, seed=21, threads=THREADS)
        _, lrec = oracle_mod.Oracle(s, oracle_mod.LIBM_LIB_PATH).run(N, seed=21, threads=THREADS)
        sens = np.maximum(sensitivity(orec, prec), sensitivity(orec, lrec))
    check_parity(st, rec, orec, sim, o, N, ev_weight_atol=F.weight_atol(F.NAMES[k], orec, 1e-10), sens=sens)


@pytest.mark.parametrize("k", range(len(F.ALL)), ids=F.NAMES)
def test_feature_mixed_runs(k):
    s = F.ALL[k]()
    print(f"[features] {F.NAMES[k]} mixed n={4 * N}", flush=True)
    ref = E.Simulation.from_setup(s, precision=E.PREC_FP64)
    st64, _ = ref.run(4 * N, seed=22, records=False)
    sim = E.Simulation.from_setup(s, precision=E.PREC_MIXED)
    st, _ = sim.run(4 * N, seed=22, records=False)
    assert st.done == 4 * N and st.failed == 0
    for d in range(sim.n_dist):
        c, _, ne = sim.distro(d)
        c64, _, ne64 = ref.distro(d)
        # event totals follow the trajectory count and the boundaries' structure, not the last bits
        assert abs(ne - ne64) <= max(5, 5 * np.sqrt(ne64)), (d, ne, ne64)
        assert abs(int(c.sum()) - int(c64.sum())) <= max(5, 5 * np.sqrt(c64.sum())), (d, c.sum(), c64.sum())
