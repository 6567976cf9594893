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
* orc_gpu_solve's survivors, level by level, against an independent Python
  replay of the GPU semantics (pinned hashes, numpy float32 scores, numpy's
  stable sort): they must be exactly the min(V, K) smallest (score bits,
  child index) keys in ascending order; planted convention mistakes (level
  seed, cid/bid, order, tie order) must make the replay disagree; a level-0
  known-answer test fixes the noise hash's call convention;
* orc_gpu_solve's operation-level child count against orc_ich_exact's;
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


def _noise_hash(seed, cid, bid):
    """gpu/v1/solver.cu:122-125 in Python (pinned below by its inverse)."""
    h = (seed ^ ((cid * 0x85EBCA6B) & 0xFFFFFFFF) ^ ((bid * 0xC2B2AE35) & 0xFFFFFFFF)) & 0xFFFFFFFF
    h ^= h >> 16
    h = (h * 0x85EBCA6B) & 0xFFFFFFFF
    h ^= h >> 16
    h = (h * 0xC2B2AE35) & 0xFFFFFFFF
    h ^= h >> 16
    return h


def _score_bits(cmax, ns, h):
    """FP32 fl(fl(cmax / ns) + fl((h % 1000) / 1e9)) as its bit pattern, in
    numpy's IEEE float32 (gpu/v1/solver.cu:121-127)."""
    base = np.float32(np.float32(cmax) / np.float32(ns))
    noise = np.float32(np.float32(h % 1000) / np.float32(1e9))
    return int(np.float32(base + noise).view(np.uint32))


class _Convention:
    """The call convention of the GPU reference's per-level hashes and
    selection.  The defaults are the reference's (gpu/v1/solver.cu:51, 122,
    250, 255-261); the other values are planted mistakes that the replay must
    catch."""

    def __init__(self, level_seed=True, cid_is_parent=True, descending=False, ties_reversed=False):
        self.level_seed, self.cid_is_parent = level_seed, cid_is_parent
        self.descending, self.ties_reversed = descending, ties_reversed


def _replay_check(inst, K, mode, best_known, seed, r, conv=_Convention()):
    """Replays orc_gpu_solve's trail with an independent Python statement of
    the GPU semantics and returns the first disagreement (None if none).

    At every level it rebuilds the parents from the previous level's
    survivors, generates every child (ICH_DET: every ready job in ready order;
    PMS_GPU: the ready jobs on the machine of ready[lowbias32(seed+level+p) %
    r], gpu/v1/solver.cu:48-69), drops those with cmax >= best_known
    (:120), keys the rest (FP32 score bits << 32 | p*n + slot, noise
    hash(seed+level, cid = parent rank p, bid = slot), :121-127) and checks
    that the trail's survivors are exactly the min(V, K) smallest keys in
    ascending order -- the prefix of CUB's stable ascending SortPairs
    (:255-261).  At the last level the returned makespan and sequence must be
    the complete child minimising (cmax, child index)."""
    n, m = inst.n, inst.m
    L, B = n * m, n
    mach, dur = inst.machine, inst.duration
    beam = [dict(mr=[0] * m, jr=[0] * n, nx=[0] * n, ready=list(range(n)), cmax=0)]
    trail, pcount = r["trail"], r["pcount"]
    expanded = 0
    best = None
    for lvl in range(L):
        lseed = (seed + lvl) & 0xFFFFFFFF if conv.level_seed else seed & 0xFFFFFFFF
        kids = []
        for p, st in enumerate(beam):
            ready = st["ready"]
            if not ready:
                continue
            if mode == ICH_DET:
                slots = list(ready)
            else:
                pj = ready[_lowbias32(lseed + p) % len(ready)]
                tmc = mach[pj, st["nx"][pj]]
                slots = [jj for jj in ready if mach[jj, st["nx"][jj]] == tmc]
            for s, j in enumerate(slots):
                k = st["nx"][j]
                mc = int(mach[j, k])
                pred = st["jr"][j] if m > 1 else (st["jr"][j - 1] if j > 0 else 0)
                f = max(st["mr"][mc], pred) + int(dur[j, k])
                cmax = max(st["cmax"], f)
                expanded += 1
                if not cmax < best_known:
                    continue
                x = p * B + s
                if lvl == L - 1:
                    if best is None or (cmax, x) < best[:2]:
                        best = (cmax, x, p, j)
                    continue
                cid, bid = (p, s) if conv.cid_is_parent else (s, p)
                bits = _score_bits(cmax, lvl + 1, _noise_hash(lseed, cid, bid))
                kids.append(((bits << 32) | ((0xFFFFFFFF - x) if conv.ties_reversed else x), x, p, j, f, mc))
        if lvl == L - 1:
            break
        # numpy's stable sort of the 64-bit keys (or the mutant's order)
        keys = np.array([c[0] for c in kids], np.uint64)
        order = np.argsort(-keys.astype(np.float64) if conv.descending else keys, kind="stable")
        keep = min(len(kids), K)
        if pcount[lvl + 1] != keep:
            return f"level {lvl}: {pcount[lvl + 1]} survivors, want {keep}"
        nxt = []
        for q in range(keep):
            _, x, p, j, f, mc = kids[order[q]]
            t = int(trail[lvl * K + q])
            if (t >> 8, t & 0xFF) != (p, j):
                return f"level {lvl} survivor {q}: (parent, job) {(t >> 8, t & 0xFF)}, want {(p, j)}"
            st = beam[p]
            ch = dict(mr=list(st["mr"]), jr=list(st["jr"]), nx=list(st["nx"]), ready=list(st["ready"]),
                      cmax=max(st["cmax"], f))
            ch["mr"][mc] = f
            ch["jr"][j] = f
            ch["nx"][j] += 1
            pos = ch["ready"].index(j)
            if mode == ICH_DET:       # erase, keeping order (cpu/v0_ich/solver.cpp:29)
                del ch["ready"][pos]
            else:                     # swap-remove (gpu/v1/solver.cu:101-106)
                ch["ready"][pos] = ch["ready"][-1]
                ch["ready"].pop()
            if ch["nx"][j] < m:       # append the job's next op (gpu_snapshot.cuh:67-69)
                ch["ready"].append(j)
            nxt.append(ch)
        beam = nxt
    if expanded != r["expanded"]:
        return f"expanded {r['expanded']}, want {expanded}"
    if best is None:
        return None if (not r["found"] and r["makespan"] == best_known) else "found a child where none exists"
    if not r["found"] or r["makespan"] != min(best[0], best_known):
        return f"makespan {r['makespan']}, want {best[0]}"
    # the sequence ends with the best child's job; the rest is its trail walk
    jobs, q = [best[3]], best[2]
    for lvl in range(L - 2, -1, -1):
        t = int(trail[lvl * K + q])
        jobs.append(t & 0xFF)
        q = t >> 8
    jobs.reverse()
    cnt = [0] * n
    want = []
    for j in jobs:
        want.append(j * m + cnt[j] + 1)
        cnt[j] += 1
    if list(r["sequence"]) != want:
        return "sequence differs from the best child's trail walk"
    return None


# (instance, K, mode, best_known, seed): wide and narrow beams, pruning, m = 1
_REPLAY_CASES = [
    (splitmix_instance(3, 8, 5, 1, 99), 16, ICH_DET, NO, 42),
    (splitmix_instance(3, 8, 5, 1, 99), 16, PMS_GPU, NO, 42),
    (splitmix_instance(11, 10, 5, 5, 99), 64, ICH_DET, NO, 12345),
    (splitmix_instance(11, 10, 5, 5, 99), 64, PMS_GPU, NO, 12345),
    (splitmix_instance(12, 6, 6, 1, 9), 500, ICH_DET, NO, 7),          # equal-score runs
    (splitmix_instance(12, 6, 6, 1, 9), 500, PMS_GPU, NO, 2**32 - 3),  # level seed wraps
    (splitmix_instance(13, 12, 4, 1, 99), 24, PMS_GPU, 0, 99),         # bound set below
    (splitmix_instance(14, 7, 1, 1, 50), 8, ICH_DET, NO, 5),            # m = 1 predecessor rule
    (splitmix_instance(15, 5, 5, 3, 3), 40, ICH_DET, NO, 1),            # every duration equal
]


@pytest.mark.parametrize("case", range(len(_REPLAY_CASES)))
def test_gpu_sem_survivors_are_the_k_smallest_keys_in_order(case):
    """orc_gpu_solve's survivors at every level are exactly the min(V, K)
    smallest (score bits, child index) keys, ascending, recomputed from the
    replayed parents with the pinned hashes and numpy float32 scores."""
    inst, K, mode, bk, seed = _REPLAY_CASES[case]
    if bk == 0:  # a bound that prunes: just above the unpruned makespan's half-way cmax
        free = orc_gpu(inst, K, mode, NO, seed)
        bk = free["makespan"] - 1
    r = orc_gpu(inst, K, mode, bk, seed)
    assert _replay_check(inst, K, mode, bk, seed, r) is None


@pytest.mark.parametrize("mutant", [dict(level_seed=False), dict(cid_is_parent=False), dict(descending=True),
                                    dict(ties_reversed=True)])
def test_replay_catches_planted_convention_mistakes(mutant):
    """The replay is sensitive: each planted mistake in the convention (level
    seed without +level, cid and bid swapped, descending order, equal
    scores ordered by descending child index) disagrees with the oracle on some replay case."""
    conv = _Convention(**mutant)
    caught = False
    for inst, K, mode, bk, seed in _REPLAY_CASES:
        if bk == 0:
            continue
        r = orc_gpu(inst, K, mode, bk, seed)
        if _replay_check(inst, K, mode, bk, seed, r, conv) is not None:
            caught = True
            break
    assert caught, mutant


def test_noise_hash_level0_known_answers():
    """Level 0 of ICH_DET from the root: every first op lasts 1, so all n
    children have cmax 1 and score fl(1 + fl((h % 1000) / 1e9)); their order
    is fixed by the noise hash with seed = the solve seed (level 0), cid = 0
    (the root's rank) and bid = the slot (= job).  The expected order is
    computed here from the pinned Python hash and compared with the oracle's
    level-0 trail."""
    n, m = 32, 3
    mach = np.array([[j % m, (j + 1) % m, (j + 2) % m] for j in range(n)], np.int32)
    dur = np.full((n, m), 5, np.int32)
    dur[:, 0] = 1
    from testlib import Inst
    inst = Inst(mach, dur)
    for seed in (0, 1, 12345, 2**31 + 17, 2**32 - 1):
        want = sorted(range(n), key=lambda s: (_score_bits(1, 1, _noise_hash(seed, 0, s)), s))
        # the KAT distinguishes the convention: several distinct score values
        assert len({_score_bits(1, 1, _noise_hash(seed, 0, s)) for s in range(n)}) > 3
        r = orc_gpu(inst, n, ICH_DET, NO, seed)
        got = [int(r["trail"][q]) for q in range(n)]
        assert [t >> 8 for t in got] == [0] * n
        assert [t & 0xFF for t in got] == want, seed


def test_python_noise_hash_matches_oracle_and_inverts():
    lib = oracle()
    rng = np.random.default_rng(9)
    for _ in range(300):
        seed, cid, bid = (int(v) for v in rng.integers(0, 2**32, 3))
        bid %= 128
        h = _noise_hash(seed, cid, bid)
        assert h == lib.orc_noise_hash(seed, cid, bid)
        assert _unmix_noise(h) == (seed ^ ((cid * 0x85EBCA6B) & 0xFFFFFFFF) ^ ((bid * 0xC2B2AE35) & 0xFFFFFFFF))


def test_gpu_sem_trail_parents_in_range():
    """Every trail entry names a parent of the previous level and a job."""
    inst = splitmix_instance(3, 8, 5, 1, 99)
    K = 16
    for mode in (ICH_DET, PMS_GPU):
        r = orc_gpu(inst, K, mode, NO, 42)
        L = inst.n * inst.m
        assert r["pcount"][0] == 1
        for lvl in range(L - 1):
            assert 0 < r["pcount"][lvl + 1] <= K
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
