"""The engine's per-kernel timing (jssp_engine_set_profiling,
jssp_engine_kernel_time, jssp_engine_kernel_launches), which bench.py's
roofline figures come from: a profiled solve gives the same result as an
unprofiled one, every step launch is timed once, the per-launch durations
and algorithmic bytes add up to the totals, and the bytes follow the
per-launch formulas from the solve's own per-level counts."""
import numpy as np
import pytest

import beamsearch_jssp_gpu_amd as bs

pytestmark = pytest.mark.gpu

SCORE, FILTER, RANK = 0, 1, 2


@pytest.mark.parametrize("phases", [1, 2])
def test_profiled_solve_times_every_launch(engine, phases):
    inst = bs.generate_instance(3, 20, 10, 1, 99)
    K, L = 256, 200
    plain = engine.solve(inst, K, mode=bs.MODE_ICH_DET, phases=phases, seed=9)
    engine.set_profiling(True)
    print("profiled solve, phases", phases, flush=True)
    r = engine.solve(inst, K, mode=bs.MODE_ICH_DET, phases=phases, seed=9, trace=True)
    totals = [engine.kernel_time(k) for k in (SCORE, FILTER, RANK)]
    launches = [engine.kernel_launches(k) for k in (SCORE, FILTER, RANK)]
    engine.set_profiling(False)
    assert r.makespan == plain.makespan and (r.sequence == plain.sequence).all()
    # Single-phase levels launch expand, filter and rank; two-phase levels
    # (fused) launch filter and rank, the rank kernel scoring the next level.
    # The last level's filter and rank only record the best child: untimed.
    want = {SCORE: L if phases == 1 else 0, FILTER: L - 1, RANK: L - 1}
    S = 2 * 20 + 10 + 1
    pc = r.pcount.astype(np.float64)
    for k, ((ms_tot, n, b_tot), (ms, b)) in enumerate(zip(totals, launches)):
        assert n == want[k] == len(ms) == len(b), (k, n, len(ms))
        if n == 0:
            continue
        assert (ms > 0).all()
        assert abs(float(ms.astype(np.float64).sum()) - ms_tot) <= 1e-3 * ms_tot + 1e-3
        assert abs(float(b.sum()) - b_tot) <= 1e-6 * b_tot
        if k == RANK:
            # 4S bytes read and written per survivor plus its 8-byte record,
            # and (fused) the next level's 12 bytes per child (key word, record).
            base = 8.0 * S * pc[1:L] + 8.0 * pc[1:L]
            assert (b >= base - 1e-6).all()
            if phases == 1:
                np.testing.assert_allclose(b, base)
        if k == FILTER:
            assert (b >= 8.0 * pc[1:L]).all()
