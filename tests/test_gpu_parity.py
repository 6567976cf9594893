"""Device path vs the oracle, bit-exact: makespan, found flag, sequence,
expanded-state count, and the full per-level survivor trail (parent rank and
job of every survivor in rank order) -- i.e. every beam of every level.

All calls go through the C ABI (include/jssp.h) via the package's ctypes layer.
"""
import numpy as np
import pytest

import beamsearch_jssp_gpu_amd as bs
from testlib import (Inst, golden_instances, load_json, lower_bound, orc_bruteforce, orc_gpu, orc_ich,
                     orc_pms, orc_simulate, random_instance, splitmix_instance)

pytestmark = pytest.mark.gpu


def _binst(i: Inst) -> bs.JSSPInstance:
    return bs.JSSPInstance(i.machine, i.duration, i.name)


def _check_det(engine, inst, K, mode, phases, seed, bk=bs.NO_MAKESPAN, flags=0):
    print(f"case {inst.name} n={inst.n} m={inst.m} K={K} mode={mode} phases={phases} seed={seed} bk={bk} "
          f"flags={flags}", flush=True)
    want = orc_gpu(inst, K, mode, bk, seed)
    got = engine.solve(_binst(inst), K, mode=mode, phases=phases, best_known=bk, seed=seed, trace=True,
                       flags=flags)
    assert got.makespan == want["makespan"]
    assert got.found == want["found"]
    assert got.expanded_states == want["expanded"]
    L = inst.n * inst.m
    np.testing.assert_array_equal(got.pcount[:L], want["pcount"][:L])
    for lvl in range(L - 1):
        c = want["pcount"][lvl + 1]
        np.testing.assert_array_equal(got.trail[lvl * K: lvl * K + c], want["trail"][lvl * K: lvl * K + c],
                                      err_msg=f"level {lvl}")
    if want["found"]:
        np.testing.assert_array_equal(got.sequence, want["sequence"])
        assert orc_simulate(inst, got.sequence) == got.makespan or got.makespan == bk
        assert bs.validate_sequence(_binst(inst), got.sequence) == orc_simulate(inst, got.sequence)
    return got


@pytest.mark.parametrize("phases", [1, 2])
@pytest.mark.parametrize("mode", [bs.MODE_ICH_DET, bs.MODE_PMS_GPU])
def test_det_modes_10x5(engine, mode, phases):
    for s in (1, 2, 3):
        inst = splitmix_instance(s, 10, 5, 5, 99)
        _check_det(engine, inst, 64, mode, phases, seed=12345 + s)


@pytest.mark.parametrize("phases", [1, 2])
@pytest.mark.parametrize("mode", [bs.MODE_ICH_DET, bs.MODE_PMS_GPU])
def test_det_modes_20x20_k1024(engine, mode, phases):
    inst = splitmix_instance(1, 20, 20, 1, 99)
    _check_det(engine, inst, 1024, mode, phases, seed=12345)


# 0: exact digit, tile-listed ties, next level scored inside the rank kernel;
# UNFUSED: the same with a separate score kernel; GENERIC_DIGIT: score-bit digit
# and the exact threshold cut; + RADIX_CUT: the radix-select cut.
SELECT_PATHS = [0, bs.FLAG_UNFUSED, bs.FLAG_GENERIC_DIGIT, bs.FLAG_GENERIC_DIGIT | bs.FLAG_RADIX_CUT]


@pytest.mark.parametrize("flags", SELECT_PATHS)
@pytest.mark.parametrize("mode", [bs.MODE_ICH_DET, bs.MODE_PMS_GPU])
def test_det_selection_paths(engine, mode, flags):
    """Every selection / scoring path (exact digit with the fused or the
    separate score kernel; generic digit + exact score histograms; generic
    digit + radix cut) gives the oracle's beams."""
    for s in (5, 6):
        inst = splitmix_instance(s, 30, 10, 1, 99)
        _check_det(engine, inst, 512, mode, 2, seed=s, flags=flags)
        _check_det(engine, inst, 64, mode, 1, seed=s + 100, flags=flags)


@pytest.mark.parametrize("K", [1024, 2048, 4096])
def test_det_histogram_copy_counts(engine, K):
    """Beams of 1024, 2048 and 4096 merge their score histograms into one,
    two and four copies (the copy count follows K); each gives the oracle's
    beams in both modes and both phase forms."""
    inst = splitmix_instance(7, 15, 8, 1, 99)
    for mode in (bs.MODE_ICH_DET, bs.MODE_PMS_GPU):
        _check_det(engine, inst, K, mode, 2, seed=77)
        _check_det(engine, inst, K, mode, 1, seed=78)


def test_det_equal_durations(engine):
    """Unit durations: massive score ties, resolved by child index only."""
    inst = splitmix_instance(3, 12, 6, 1, 1)
    inst.name = "unit12x6"
    for mode in (bs.MODE_ICH_DET, bs.MODE_PMS_GPU):
        for flags in SELECT_PATHS:
            _check_det(engine, inst, 300, mode, 2, seed=1, flags=flags)


def test_det_zero_and_wide_durations(engine):
    """Zero durations (scores are pure noise) and durations up to 10^6 (a
    lower makespan no longer always scores lower, so the generic digit is used)."""
    rng = np.random.default_rng(21)
    z = Inst(np.array([rng.permutation(4) for _ in range(6)], np.int32), np.zeros((6, 4), np.int32), "zero6x4")
    w = random_instance(rng, 10, 6, dmin=1, dmax=1_000_000)
    w.name = "wide10x6"
    for inst in (z, w):
        for mode in (bs.MODE_ICH_DET, bs.MODE_PMS_GPU):
            _check_det(engine, inst, 200, mode, 2, seed=3)


@pytest.mark.parametrize("mode", [bs.MODE_ICH_DET, bs.MODE_PMS_GPU])
def test_det_odd_shapes(engine, mode):
    rng = np.random.default_rng(7)
    shapes = [(1, 1, 4), (1, 6, 3), (6, 1, 5), (2, 2, 1), (3, 4, 2), (65, 3, 8), (100, 2, 16), (128, 2, 4),
              (7, 9, 1000)]
    for (n, m, K) in shapes:
        inst = random_instance(rng, n, m)
        inst.name = f"rand{n}x{m}"
        for phases in (1, 2):
            _check_det(engine, inst, K, mode, phases, seed=int(rng.integers(0, 2**31)))


def test_det_repeated_machines(engine):
    """Jobs that visit a machine more than once (the reference does not forbid it)."""
    rng = np.random.default_rng(11)
    for mode in (bs.MODE_ICH_DET, bs.MODE_PMS_GPU):
        inst = random_instance(rng, 12, 6, perm=False)
        inst.name = "repeat12x6"
        _check_det(engine, inst, 128, mode, 2, seed=5)


def test_det_pruning(engine):
    """best_known prunes children with cmax >= bound (gpu/v1/solver.cu:120)."""
    inst = splitmix_instance(4, 15, 8, 1, 99)
    free = orc_gpu(inst, 128, bs.MODE_PMS_GPU, bs.NO_MAKESPAN, 3, trace=False)["makespan"]
    for bk in (free + 50, free + 1, free, free - 20, lower_bound(inst)):
        for phases in (1, 2):
            g = _check_det(engine, inst, 128, bs.MODE_PMS_GPU, phases, seed=3, bk=bk)
            assert g.makespan <= bk


def test_det_wide_beam_is_exhaustive(engine):
    """With K above every level's child count, operation-level search keeps
    every state and must reach the brute-force optimum."""
    rng = np.random.default_rng(3)
    for t in range(6):
        inst = random_instance(rng, 3, 3)
        opt = orc_bruteforce(inst)
        got = engine.solve(_binst(inst), 4096, mode=bs.MODE_ICH_DET, phases=2, seed=t)
        assert got.makespan == opt
        assert orc_simulate(inst, got.sequence) == opt


def test_ich_exact_matches_reference_goldens(engine):
    g = load_json("golden_ref.json")
    insts = golden_instances()
    n = 0
    for c in g["cases"]:
        if c["solver"] != "ich":
            continue
        inst = insts[c["instance"]]
        print("ich", c["instance"], c["K"], flush=True)
        r = engine.solve(_binst(inst), c["K"], mode=bs.MODE_ICH_EXACT, seed=c["seed"])
        assert r.makespan == c["makespan"], c
        ms, seq, ex = orc_ich(inst, c["K"], c["seed"])
        assert r.expanded_states == ex
        np.testing.assert_array_equal(r.sequence, seq)
        n += 1
    assert n > 50


def test_pms_exact_matches_reference_goldens(engine):
    g = load_json("golden_ref.json")
    insts = golden_instances()
    n = 0
    for c in g["cases"]:
        if c["solver"] != "pms":
            continue
        inst = insts[c["instance"]]
        print("pms", c["instance"], c["K"], c["best_known"], flush=True)
        r = engine.solve(_binst(inst), c["K"], mode=bs.MODE_PMS_EXACT, best_known=c["best_known"], seed=c["seed"])
        assert r.makespan == c["makespan"], c
        ms, ex = orc_pms(inst, c["K"], c["best_known"], c["seed"])
        assert r.expanded_states == ex
        if r.found:
            assert orc_simulate(inst, r.sequence) == r.makespan
        n += 1
    assert n > 50


def test_ich_exact_10x5_k64_config(engine):
    """The 10x5 K=64 operation-level configuration on 32 seeded instances."""
    g = load_json("golden_ref.json")
    insts = golden_instances()
    cases = [c for c in g["cases"] if c["solver"] == "ich" and c["instance"].startswith("r10x5")]
    assert len(cases) == 32
    for c in cases:
        r = engine.solve(_binst(insts[c["instance"]]), 64, mode=bs.MODE_ICH_EXACT, seed=c["seed"])
        assert r.makespan == c["makespan"]


def test_reference_signature_shims(engine):
    inst = splitmix_instance(9, 8, 4, 1, 99)
    b = _binst(inst)
    seq = np.zeros(32, np.int32)
    ms = bs.run_gpu_beam_search_v1(b, 32, bs.NO_MAKESPAN, 77, seq)
    want = orc_gpu(inst, 32, bs.MODE_PMS_GPU, bs.NO_MAKESPAN, 77, trace=False)
    assert ms == want["makespan"]
    np.testing.assert_array_equal(seq, want["sequence"])
    assert bs.run_gpu_beam_search_v2(b, 32, bs.NO_MAKESPAN, 77) == ms
    assert bs.beam_search_ich(b, 16, seed=12345) == orc_ich(inst, 16, 12345)[0]
    assert bs.beam_search_pms(b, 16, bs.NO_MAKESPAN, seed=5) == orc_pms(inst, 16, bs.NO_MAKESPAN, 5)[0]


def test_limits_and_argument_errors(engine):
    """The largest shapes the engine accepts run and re-simulate (255
    machines; 65536-wide beam; 128 jobs); one past any limit is an error code,
    not a crash."""
    rng = np.random.default_rng(11)
    inst = random_instance(rng, 3, 255)                       # 765 levels, 255 machines
    r = engine.solve(_binst(inst), 8, mode=bs.MODE_ICH_DET, seed=1)
    w = orc_gpu(inst, 8, bs.MODE_ICH_DET, bs.NO_MAKESPAN, 1, trace=False)
    assert r.makespan == w["makespan"] and r.expanded_states == w["expanded"]
    inst = random_instance(rng, 12, 3)                        # K = 65536 over 36 levels
    for mode in (bs.MODE_ICH_DET, bs.MODE_PMS_GPU):
        r = engine.solve(_binst(inst), 65536, mode=mode, seed=2)
        w = orc_gpu(inst, 65536, mode, bs.NO_MAKESPAN, 2, trace=False)
        assert r.makespan == w["makespan"] and r.expanded_states == w["expanded"]
        np.testing.assert_array_equal(r.sequence, w["sequence"])
    too_big = [(random_instance(rng, 129, 2), 4), (random_instance(rng, 2, 256), 4),
               (random_instance(rng, 4, 4), 65537)]
    for bad, K in too_big:
        with pytest.raises(bs.JsspError) as e:
            engine.solve(_binst(bad), K, mode=bs.MODE_ICH_DET)
        assert e.value.code == 2  # JSSP_ERR_LIMIT
    with pytest.raises(bs.JsspError):
        engine.solve(_binst(random_instance(rng, 4, 4)), 0, mode=bs.MODE_ICH_DET)
