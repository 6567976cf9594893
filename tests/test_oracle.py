"""Pins the oracle (oracle/oracle.cpp) before it is trusted as the checker.

* orc_ich_exact / orc_pms_exact against makespans of the compiled reference
  (tests/golden/golden_ref.json, written by tests/golden/make_golden.py from
  cpu/v0_ich and cpu/v1_pms), and against oracle/_ref live where it is built;
* every search mode against brute force on tiny instances with a beam wide
  enough to keep every state;
* returned sequences re-simulate to the returned makespan and respect the
  lower bound;
* the pivot hash against the published inverse of lowbias32 (a wrong
  constant or shift breaks the round trip);
* orc_gpu_solve's operation-level child generation against orc_ich_exact
  (same ready order and children; only the score differs), via K=1 greedy
  equality on instances where the rand noise cannot matter;
* the restated glibc generator against the C library's own srand()/rand();
* the multi-rank driver loops: one rank and one pass reduce to the pinned
  single solves, and the PMS loop equals the reference's own solver run
  under per-rank glibc states (oracle/_ref).
"""
import numpy as np
import pytest

from testlib import (Inst, golden_instances, libc_rand, load_json, lower_bound, oracle, orc_bruteforce,
                     orc_gpu, orc_gpu_driver, orc_hybrid_driver, orc_ich, orc_pms, orc_pms_driver, orc_rand,
                     orc_simulate, p32, random_instance, ref_lib, splitmix_instance, valid_trail_entries)

ICH_DET, PMS_GPU = 1, 2
NO = 500000


def test_oracle_exact_modes_match_reference_goldens():
    g = load_json("golden_ref.json")
    insts = golden_instances()
    assert len(g["cases"]) > 200
    for c in g["cases"]:
        inst = insts[c["instance"]]
        if c["solver"] == "ich":
            ms, seq, _ = orc_ich(inst, c["K"], c["seed"])
            assert ms == c["makespan"], c
            assert orc_simulate(inst, seq) == ms
        else:
            ms, _ = orc_pms(inst, c["K"], c["best_known"], c["seed"])
            assert ms == c["makespan"], c


@pytest.mark.skipif(ref_lib() is None, reason="oracle/_ref not built (no reference tree)")
def test_oracle_exact_modes_match_live_reference():
    ref = ref_lib()
    rng = np.random.default_rng(20261203)
    for t in range(60):
        n, m = int(rng.integers(1, 12)), int(rng.integers(1, 7))
        inst = random_instance(rng, n, m, perm=bool(t % 3))
        K = int(rng.choice([1, 2, 5, 16, 64, 300]))
        seed = int(rng.integers(0, 2**31))
        from testlib import p32
        a = ref.ref_ich(n, m, p32(inst.machine), p32(inst.duration), K, seed)
        assert orc_ich(inst, K, seed)[0] == a, (n, m, K, seed)
        bk = int(rng.choice([NO, a, a + 7, max(1, a - 10)]))
        b = ref.ref_pms(n, m, p32(inst.machine), p32(inst.duration), K, bk, seed)
        assert orc_pms(inst, K, bk, seed)[0] == b, (n, m, K, bk, seed)


@pytest.mark.parametrize("shape", [(2, 2), (3, 2), (2, 4), (3, 3), (4, 3), (1, 5), (5, 1)])
def test_wide_beam_equals_bruteforce(shape):
    rng = np.random.default_rng(sum(shape))
    for _ in range(4):
        inst = random_instance(rng, *shape)
        opt = orc_bruteforce(inst)
        assert opt >= lower_bound(inst) or shape[1] == 1
        ms, seq, _ = orc_ich(inst, 100000, 1)
        assert ms == opt
        assert orc_simulate(inst, seq) == opt
        r = orc_gpu(inst, 100000, ICH_DET, NO, 9)
        assert r["makespan"] == opt
        assert orc_simulate(inst, r["sequence"]) == opt


def test_sequences_resimulate_and_bound():
    rng = np.random.default_rng(5)
    for t in range(20):
        inst = random_instance(rng, int(rng.integers(2, 15)), int(rng.integers(2, 8)))
        for mode in (ICH_DET, PMS_GPU):
            r = orc_gpu(inst, 32, mode, NO, t)
            assert r["found"]
            assert orc_simulate(inst, r["sequence"]) == r["makespan"]
            assert r["makespan"] >= lower_bound(inst)


def test_gpu_sem_trail_is_sorted_by_score_then_child():
    """Survivor q of level l comes from parent p with a (score, child) key not
    above survivor q+1's -- recomputed from the replayed states."""
    inst = splitmix_instance(3, 8, 5, 1, 99)
    K = 16
    for mode in (ICH_DET, PMS_GPU):
        r = orc_gpu(inst, K, mode, NO, 42)
        L = inst.n * inst.m
        assert r["pcount"][0] == 1
        for lvl in range(L - 1):
            assert 0 < r["pcount"][lvl + 1] <= K
        # parent ranks of consecutive survivors come from the previous level
        for lvl, q, p, j in valid_trail_entries(r["trail"], r["pcount"], K):
            assert 0 <= p < r["pcount"][lvl]
            assert 0 <= j < inst.n


def _lowbias32(x):
    x &= 0xFFFFFFFF
    x ^= x >> 16
    x = (x * 0x7FEB352D) & 0xFFFFFFFF
    x ^= x >> 15
    x = (x * 0x846CA68B) & 0xFFFFFFFF
    x ^= x >> 16
    return x


def _lowbias32_inverse(x):
    # published inverse of lowbias32 (C. Wellons, hash-prospector)
    x &= 0xFFFFFFFF
    x ^= x >> 16
    x = (x * 0x43021123) & 0xFFFFFFFF
    x ^= (x >> 15) ^ (x >> 30)
    x = (x * 0x1D69E2A5) & 0xFFFFFFFF
    x ^= x >> 16
    return x


def test_pivot_hash_is_lowbias32():
    """The pivot hash of gpu/v1/solver.cu:51-53 is lowbias32; its published
    inverse must undo it, which fixes both multipliers and all shifts."""
    rng = np.random.default_rng(0)
    for x in list(range(1000)) + [int(v) for v in rng.integers(0, 2**32, 2000)]:
        assert _lowbias32_inverse(_lowbias32(x)) == x


def test_pivot_choice_follows_lowbias32():
    """PMS_GPU level 0 from the root: the oracle's single survivor at K=1
    schedules an op of the pivot machine ready[lowbias32(seed) % n]."""
    rng = np.random.default_rng(1)
    for t in range(30):
        inst = random_instance(rng, 6, 4)
        seed = int(rng.integers(0, 2**31))
        r = orc_gpu(inst, 1, PMS_GPU, NO, seed)
        j0 = int(r["trail"][0]) & 0xFF
        pj = _lowbias32(seed) % inst.n
        assert inst.machine[j0, 0] == inst.machine[pj, 0]


def test_ich_det_children_equal_ich_children():
    """With one job per machine in every ready set (a flow of distinct first
    machines) and K large, ICH_DET and the reference ICH explore the same
    children; with K >= all states both reach the same optimum."""
    rng = np.random.default_rng(8)
    for _ in range(5):
        inst = random_instance(rng, 3, 3)
        a = orc_ich(inst, 10**6, 3)[2]
        b = orc_gpu(inst, 10**6, ICH_DET, NO, 3)["expanded"]
        assert a == b  # same number of children when nothing is pruned


def test_pruning_semantics():
    inst = splitmix_instance(2, 10, 5, 5, 99)
    free = orc_gpu(inst, 64, PMS_GPU, NO, 1)
    assert free["found"]
    tight = orc_gpu(inst, 64, PMS_GPU, free["makespan"], 1)
    assert tight["makespan"] <= free["makespan"]
    none = orc_gpu(inst, 64, PMS_GPU, 1, 1)
    assert not none["found"] and none["makespan"] == 1


def test_oracle_pivot_hash_inverts():
    """The oracle's pivot hash is undone by the published lowbias32 inverse."""
    lib = oracle()
    rng = np.random.default_rng(5)
    for x in [0, 1, 2, 0xFFFFFFFF] + [int(v) for v in rng.integers(0, 2**32, 500)]:
        assert _lowbias32_inverse(lib.orc_pivot_hash(x)) == x


def _unmix_noise(h):
    # Inverse of: h ^= h>>16; h *= c1; h ^= h>>16; h *= c2; h ^= h>>16.  A
    # xor-shift by 16 on 32 bits is its own inverse; an odd multiplier is
    # undone by its inverse mod 2^32.
    c1i, c2i = pow(0x85EBCA6B, -1, 2**32), pow(0xC2B2AE35, -1, 2**32)
    h ^= h >> 16
    h = (h * c2i) & 0xFFFFFFFF
    h ^= h >> 16
    h = (h * c1i) & 0xFFFFFFFF
    h ^= h >> 16
    return h


def test_oracle_noise_hash_inverts_to_its_input_mix():
    """The sibling-noise hash (gpu/v1/solver.cu:122-125): unmixing its output
    gives back seed ^ cid*c1 ^ bid*c2; a wrong shift, multiplier or a swapped
    cid/bid breaks it."""
    lib = oracle()
    rng = np.random.default_rng(6)
    for _ in range(500):
        seed, cid, bid = (int(v) for v in rng.integers(0, 2**32, 3))
        bid %= 128
        h = lib.orc_noise_hash(seed, cid, bid)
        assert _unmix_noise(h) == (seed ^ ((cid * 0x85EBCA6B) & 0xFFFFFFFF) ^ ((bid * 0xC2B2AE35) & 0xFFFFFFFF))
    # cid and bid are not interchangeable
    assert lib.orc_noise_hash(7, 3, 5) != lib.orc_noise_hash(7, 5, 3)


def test_oracle_score_is_fp32_division_plus_noise():
    """score = fl(fl(cmax / ns) + fl((h % 1000) / 1e9)) in FP32, in that
    order; compared with numpy's IEEE float32 arithmetic."""
    lib = oracle()
    rng = np.random.default_rng(7)
    for _ in range(2000):
        cmax, ns, h = int(rng.integers(0, 200000)), int(rng.integers(1, 33000)), int(rng.integers(0, 2**32))
        want = np.float32(np.float32(cmax) / np.float32(ns)) + np.float32(np.float32(h % 1000) / np.float32(1e9))
        got = np.float32(lib.orc_gpu_score(cmax, ns, h))
        assert got.view(np.uint32) == np.float32(want).view(np.uint32), (cmax, ns, h)


@pytest.mark.parametrize("seed", [0, 1, 2, 12345, 1700000000, 2**31 - 1, 2**31, 2**31 + 12345, 2**32 - 1])
def test_oracle_rand_is_glibc_rand(seed):
    """The restated TYPE_3 generator gives the C library's own rand() values
    (seed 0 maps to 1; seeds >= 2^31 go through the signed seeding path)."""
    np.testing.assert_array_equal(orc_rand(seed, 1000), libc_rand(seed, 1000))


def test_pms_driver_reduces_to_single_pass():
    inst = golden_instances()["la01"]
    for t in (5, 1234567):
        d = orc_pms_driver(inst, 16, 1, 1, t)
        ms, ex = orc_pms(inst, 16, NO, t)
        assert d["makespan"] == ms and d["expanded"] == ex and list(d["pass_best"]) == [ms]
        assert d["best_index"] == (0 if ms < NO else -1)


@pytest.mark.skipif(ref_lib() is None, reason="oracle/_ref not built (no reference tree)")
def test_pms_driver_matches_reference_solver_with_rank_streams():
    ref = ref_lib()
    insts = golden_instances()
    for name, K, R, E, t in [("la01", 8, 3, 4, 1700000000), ("la06", 16, 2, 3, 42), ("r10x5s1", 4, 4, 3, 7)]:
        inst = insts[name]
        want = np.zeros(E, np.int32)
        got = orc_pms_driver(inst, K, R, E, t)
        rv = ref.ref_pms_driver(inst.n, inst.m, p32(inst.machine), p32(inst.duration), K, R, E, t, p32(want))
        assert got["makespan"] == rv, name
        np.testing.assert_array_equal(got["pass_best"], want)
        assert np.all(np.diff(want) <= 0)


def test_gpu_driver_reduces_to_single_solves():
    inst = splitmix_instance(21, 8, 5, 1, 99)
    t = 99
    d = orc_gpu_driver(inst, 16, 1, 1, t)
    w = orc_gpu(inst, 16, PMS_GPU, NO, int(orc_rand(t, 1)[0]), trace=False)
    assert d["makespan"] == w["makespan"] and d["expanded"] == w["expanded"]
    np.testing.assert_array_equal(d["sequence"], w["sequence"])
    # two ranks, three passes: pass e prunes with the minimum of passes < e
    R, E = 2, 3
    d = orc_gpu_driver(inst, 8, R, E, t)
    seeds = np.stack([orc_rand(t + r, E) for r in range(R)], axis=1)  # [pass, rank]
    bound, exp, mins = NO, 0, []
    for e in range(E):
        res = [orc_gpu(inst, 8, PMS_GPU, bound, int(seeds[e, r]), trace=False) for r in range(R)]
        exp += sum(x["expanded"] for x in res)
        bound = min([bound] + [x["makespan"] for x in res])
        mins.append(bound)
    assert d["expanded"] == exp and list(d["pass_best"]) == mins and d["makespan"] == bound


def test_hybrid_driver_reduces_to_its_parts():
    """One thread per rank is the GPU driver; with host threads, every
    iteration's best is the minimum of the GPU pass and the PMS passes from
    their seeds (each pinned on its own)."""
    inst = splitmix_instance(17, 8, 5, 1, 99)
    a = orc_hybrid_driver(inst, 16, 4, 1, 2, 3, 555)
    b = orc_gpu_driver(inst, 16, 2, 3, 555)
    assert a["makespan"] == b["makespan"] and list(a["pass_best"]) == list(b["pass_best"])
    assert a["expanded"] == b["expanded"]
    T, t = 3, 777
    h = orc_hybrid_driver(inst, 16, 4, T, 1, 2, t)
    seeds = orc_rand(t, 2)
    bound, exp, mins = NO, 0, []
    for it in range(2):
        w = orc_gpu(inst, 16, PMS_GPU, bound, int(seeds[it]), trace=False)
        best, exp = min(bound, w["makespan"]), exp + w["expanded"]
        for tid in range(1, T):
            ms, ex = orc_pms(inst, 4, bound, t + it * 100 + tid)
            best, exp = min(best, ms), exp + ex
        bound = min(bound, best)
        mins.append(bound)
    assert list(h["pass_best"]) == mins and h["expanded"] == exp
