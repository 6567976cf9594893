"""The no-allocation entry point (jssp_solve_noalloc, C ABI): its results
equal jssp_solve's on the same requests, on consecutive calls whose second
input is smaller than the first (the shared scratch array keeps the first
call's contents), and sizes beyond the scratch array are refused with
JSSP_ERR_LIMIT."""
import numpy as np
import pytest

import beamsearch_jssp_gpu_amd as bs

pytestmark = pytest.mark.gpu


def _same(a, b):
    assert a.makespan == b.makespan and a.found == b.found and a.expanded_states == b.expanded_states
    assert a.levels == b.levels
    if a.found:
        np.testing.assert_array_equal(a.sequence, b.sequence)


def test_noalloc_equals_solve_large_then_small():
    eng = bs.Engine(0)
    cases = [
        (bs.generate_instance(1, 50, 20, 1, 99), 4096, bs.MODE_ICH_DET, 2, 12345),  # the headline solve
        (bs.generate_instance(7, 12, 6, 1, 99), 64, bs.MODE_PMS_GPU, 2, 3),         # smaller, right after
        (bs.generate_instance(8, 20, 10, 1, 99), 256, bs.MODE_ICH_DET, 1, 9),       # single-phase
        (bs.generate_instance(9, 10, 5, 5, 99), 32, bs.MODE_PMS_GPU, 1, 4),
    ]
    for inst, K, mode, phases, seed in cases:
        print("noalloc", inst.nbjob, "x", inst.nbmachine, "K", K, "mode", mode, "phases", phases, flush=True)
        want = eng.solve(inst, K, mode=mode, phases=phases, seed=seed)
        got = eng.solve_noalloc(inst, K, mode=mode, phases=phases, seed=seed)
        _same(got, want)
    eng.close()


def test_noalloc_bound_and_limits():
    eng = bs.Engine(0)
    inst = bs.generate_instance(7, 12, 6, 1, 99)
    free = eng.solve(inst, 64, mode=bs.MODE_PMS_GPU, seed=3)
    tight = eng.solve_noalloc(inst, 64, mode=bs.MODE_PMS_GPU, seed=3, best_known=free.makespan)
    assert not tight.found and tight.makespan == free.makespan
    _same(eng.solve_noalloc(inst, 64, mode=bs.MODE_PMS_GPU, seed=3, best_known=free.makespan + 1),
          eng.solve(inst, 64, mode=bs.MODE_PMS_GPU, seed=3, best_known=free.makespan + 1))
    big = bs.generate_instance(1, 50, 20, 1, 99)
    with pytest.raises(bs.JsspError) as e:  # single-phase 50x20 K=4096 needs ~130 MB
        eng.solve_noalloc(big, 4096, mode=bs.MODE_ICH_DET, phases=1, seed=1)
    assert e.value.code == 2
    with pytest.raises(bs.JsspError) as e:
        eng.solve_noalloc(inst, 64, mode=bs.MODE_ICH_EXACT, seed=1)
    assert e.value.code == 1
    # still usable after the refusals
    _same(eng.solve_noalloc(inst, 64, mode=bs.MODE_ICH_DET, seed=5), eng.solve(inst, 64, mode=bs.MODE_ICH_DET, seed=5))
    eng.close()
