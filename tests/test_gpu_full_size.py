"""Size-independent properties at the benchmark shapes, where the oracle is
too slow to run: 50x20 with K=4096 (about 1.6e8 children per solve), the
16-beam multi-seed search and the 256-instance batch.

* single-phase and two-phase solves of the same request give the same
  beams: same makespan, sequence and expanded count (they share no
  selection or materialisation code path);
* the sequence re-simulates to the makespan and respects the lower bound;
* repeated solves are bit-identical;
* the multi-seed search equals its beams solved one by one;
* the batch equals its instances solved one by one."""
import numpy as np
import pytest

import beamsearch_jssp_gpu_amd as bs

pytestmark = pytest.mark.gpu


def _same(a, b):
    assert a.makespan == b.makespan and a.found == b.found and a.expanded_states == b.expanded_states
    if a.found:
        np.testing.assert_array_equal(a.sequence, b.sequence)


@pytest.mark.parametrize("mode", [bs.MODE_ICH_DET, bs.MODE_PMS_GPU])
def test_50x20_k4096_phases_agree(engine, mode):
    inst = bs.generate_instance(1, 50, 20, 1, 99)
    print("50x20 K=4096 mode", mode, flush=True)
    two = engine.solve(inst, 4096, mode=mode, phases=2, seed=12345)
    one = engine.solve(inst, 4096, mode=mode, phases=1, seed=12345)
    _same(two, one)
    again = engine.solve(inst, 4096, mode=mode, phases=2, seed=12345)
    _same(two, again)
    assert two.found
    assert bs.validate_sequence(inst, two.sequence) == two.makespan
    assert two.makespan >= bs.lower_bound(inst)
    # every level keeps min(children, K) beams: at least one child per level
    assert two.expanded_states >= 1000 and two.levels == 1000


def test_16_beams_equal_single_solves(engine):
    inst = bs.generate_instance(3, 50, 20, 1, 99)
    K, S = 4096, 16
    print("16 beams", flush=True)
    got = engine.solve(inst, K, mode=bs.MODE_PMS_GPU, seed=1, n_seeds=S)
    singles = [engine.solve(inst, K, mode=bs.MODE_PMS_GPU, seed=1 + i) for i in range(S)]
    best = min(range(S), key=lambda i: (singles[i].makespan if singles[i].found else bs.NO_MAKESPAN, i))
    assert got.makespan == singles[best].makespan
    assert got.best_seed_index == best
    assert got.expanded_states == sum(r.expanded_states for r in singles)
    np.testing.assert_array_equal(got.sequence, singles[best].sequence)


def test_256_instance_batch_equals_single_solves(engine):
    rng = np.random.default_rng(2026)
    insts, params = [], []
    for t in range(256):
        n, m = int(rng.choice([10, 15, 20, 30, 50])), int(rng.choice([5, 10, 15, 20]))
        insts.append(bs.generate_instance(2026 * 1000 + t, n, m, 1, 99))
        params.append(dict(beam_width=int(rng.choice([256, 1024])), mode=bs.MODE_PMS_GPU, phases=2,
                           best_known=bs.NO_MAKESPAN, seed=t))
    print("batch of 256", flush=True)
    got = engine.solve_batch(insts, params)
    for i, (inst, p) in enumerate(zip(insts, params)):
        _same(got[i], engine.solve(inst, **p))
