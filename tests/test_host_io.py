"""Host-side logic of the library (no device): instance files, the resolver,
the generator, the validator and the lower bound."""
import os
import shutil

import numpy as np
import pytest

import beamsearch_jssp_gpu_amd as bs
from testlib import GOLDEN, load_json, lower_bound, orc_simulate, splitmix_instance


def test_parser_matches_reference_parser_on_lawrence():
    law = load_json("lawrence.json")
    inst = bs.parse_instance(os.path.join(GOLDEN, "instances", "la01.txt"))
    assert (inst.nbjob, inst.nbmachine, inst.nboperation) == (10, 5, 52)
    assert inst.machine.ravel().tolist() == law["la01"]["machine"]
    assert inst.duration.ravel().tolist() == law["la01"]["duration"]


def test_parser_taillard_formats():
    a = bs.parse_instance(os.path.join(GOLDEN, "instances", "ta01"))
    assert a.machine.shape == (15, 15)
    b = bs.parse_instance(os.path.join(GOLDEN, "instances", "ta71.txt"))
    assert b.machine.shape == (100, 20)
    for inst in (a, b):
        for row in inst.machine:
            assert sorted(row.tolist()) == list(range(inst.nbmachine))


def test_parser_errors(tmp_path):
    with pytest.raises(bs.JsspError) as e:
        bs.parse_instance(str(tmp_path / "missing.txt"))
    assert e.value.code == 5
    bad = tmp_path / "trunc.txt"
    bad.write_text("3 2\n0 5 1 6\n1 4\n")
    with pytest.raises(bs.JsspError) as e:
        bs.parse_instance(str(bad))
    assert e.value.code == 1
    hdr = tmp_path / "hdr.txt"
    hdr.write_text("x y\n")
    with pytest.raises(bs.JsspError):
        bs.parse_instance(str(hdr))


def test_resolver_search_order(tmp_path, monkeypatch):
    """cpu/common/instance_parser.cpp:8-34: a direct path first, then
    ./laXX.txt, ./taXX.txt, then the lawrence/taillard folders."""
    work = tmp_path / "a" / "b"
    work.mkdir(parents=True)
    (tmp_path / "lawrence_instances").mkdir()
    (tmp_path / "taillard_instances").mkdir()
    src = os.path.join(GOLDEN, "instances", "la01.txt")
    shutil.copy(src, tmp_path / "lawrence_instances" / "la07.txt")
    shutil.copy(os.path.join(GOLDEN, "instances", "ta01"), tmp_path / "taillard_instances" / "ta07")
    shutil.copy(os.path.join(GOLDEN, "instances", "ta01"), tmp_path / "taillard_instances" / "ta09")
    monkeypatch.chdir(work)
    # Lawrence wins over Taillard for the same id (../../lawrence before ../../taillard)
    assert bs.resolve_instance_file("07") == "../../lawrence_instances/la07.txt"
    assert bs.resolve_instance_file("09") == "../../taillard_instances/ta09"
    shutil.copy(src, work / "la07.txt")
    assert bs.resolve_instance_file("07") == "./la07.txt"
    assert bs.resolve_instance_file(src) == src
    assert bs.resolve_instance_file("99") is None


def test_generator_matches_fixture_stream():
    g = load_json("golden_ref.json")["randoms"]
    for name, d in g.items():
        inst = bs.generate_instance(d["seed"], d["n"], d["m"], d["dmin"], d["dmax"])
        assert inst.machine.ravel().tolist() == d["machine"], name
        assert inst.duration.ravel().tolist() == d["duration"], name
    big = bs.generate_instance(1, 50, 20, 1, 99)
    ref = splitmix_instance(1, 50, 20, 1, 99)
    assert np.array_equal(big.machine, ref.machine) and np.array_equal(big.duration, ref.duration)
    assert big.duration.min() >= 1 and big.duration.max() <= 99
    for row in big.machine:
        assert sorted(row.tolist()) == list(range(20))


def test_validator_and_lower_bound():
    inst = bs.generate_instance(3, 6, 4, 1, 20)
    rng = np.random.default_rng(1)
    for _ in range(50):
        # random feasible sequence
        left = [4] * 6
        nxt = [0] * 6
        seq = []
        while len(seq) < 24:
            j = int(rng.choice([k for k in range(6) if left[k]]))
            seq.append(j * 4 + nxt[j] + 1)
            nxt[j] += 1
            left[j] -= 1
        ms = bs.validate_sequence(inst, np.array(seq))
        assert ms == orc_simulate(inst, np.array(seq, np.int32))
        assert ms >= bs.lower_bound(inst)
    bad = np.arange(24, 0, -1, dtype=np.int32)
    assert bs.validate_sequence(inst, bad) == -1
    assert bs.lower_bound(inst) == lower_bound(inst)


@pytest.mark.parametrize("seed", [0, 1, 12345, 1700000003, 2**31 + 99, 2**32 - 1])
def test_rank_rand_streams_match_libc(seed):
    """jssp_rand_stream (the private per-rank streams of jssp_run_passes)
    gives srand(seed); rand()... of the C library and of the restated
    generator, and leaves the process-wide stream where it was."""
    from testlib import libc_rand, orc_rand
    import ctypes
    libc = ctypes.CDLL(None)
    libc.rand.restype = ctypes.c_int
    libc.srand(777)
    got = bs.rand_stream(seed, 500)
    after = libc.rand()
    np.testing.assert_array_equal(got, orc_rand(seed, 500))
    np.testing.assert_array_equal(got, libc_rand(seed, 500))
    assert after == int(libc_rand(777, 1)[0])


def test_lawrence_runner_skips_missing_files(tmp_path, capsys):
    """With no instance files in reach the runner prints the reference's
    header and one SKIP line per id, and touches no device."""
    from beamsearch_jssp_gpu_amd import drivers
    cwd = os.getcwd()
    try:
        assert drivers.main(["bench_gpu_v1", "--dir", str(tmp_path), "--time-seed", "1"]) == 0
    finally:
        os.chdir(cwd)
    out = capsys.readouterr().out.splitlines()
    assert out[1].split() == ["Kbest", "Instance", "Jobs", "Machines", "Ops", "BestSpan", "Time(s)"]
    skips = [l for l in out if l.startswith("SKIP")]
    assert skips[0] == "SKIP la01  (file not found)" and len(skips) == 40
    assert [drivers.lawrence_beam_width(i) for i in (0, 9, 10, 39)] == [1000, 1000, 5000, 5000]


def test_host_pms_matches_reference_goldens():
    """The host BS-PMS (the CPU side of the hybrid drivers) gives the compiled
    reference's makespans (tests/golden/golden_ref.json), the oracle's
    expanded counts, and sequences that re-simulate to the makespan."""
    from testlib import golden_instances, orc_pms
    g = load_json("golden_ref.json")
    insts = golden_instances()
    cases = [c for c in g["cases"] if c["solver"] == "pms"]
    assert len(cases) > 50
    for c in cases:
        inst = insts[c["instance"]]
        b = bs.JSSPInstance(inst.machine, inst.duration)
        r = bs.host_solve_pms(b, c["K"], c["best_known"], c["seed"])
        assert r.makespan == c["makespan"], c
        ms, ex = orc_pms(inst, c["K"], c["best_known"], c["seed"])
        assert r.expanded_states == ex, c
        if r.found:
            assert orc_simulate(inst, r.sequence) == r.makespan, c


def test_host_pms_random_shapes_match_oracle():
    from testlib import orc_pms, random_instance
    rng = np.random.default_rng(99)
    for t in range(40):
        n, m = int(rng.integers(1, 14)), int(rng.integers(1, 8))
        inst = random_instance(rng, n, m, perm=bool(t % 2))
        K = int(rng.choice([1, 3, 16, 100]))
        seed = int(rng.integers(0, 2**32))
        bk = int(rng.choice([500000, 150, 400]))
        r = bs.host_solve_pms(bs.JSSPInstance(inst.machine, inst.duration), K, bk, seed)
        ms, ex = orc_pms(inst, K, bk, seed)
        assert (r.makespan, r.expanded_states) == (ms, ex), (n, m, K, seed, bk)


REF_TREE = "/root/reference"


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF_TREE, "lawrence_instances")),
                    reason="the reference tree (instance files) is not present")
def test_parser_matches_reference_parser_on_every_instance_file():
    """All 120 instance files the reference ships (40 Lawrence, 80 Taillard)
    parse to the same n, m, machines and durations with this parser and with
    the reference's own parse_instance (compiled into oracle/_ref)."""
    import ctypes
    from testlib import ref_lib
    ref = ref_lib()
    if ref is None:
        pytest.skip("oracle/_ref not built")
    ref.ref_parse.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                              ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), ctypes.c_int]
    files = []
    for d in ("lawrence_instances", "taillard_instances"):
        files += sorted(os.path.join(REF_TREE, d, f) for f in os.listdir(os.path.join(REF_TREE, d)))
    assert len(files) == 120
    for path in files:
        inst = bs.parse_instance(path)
        cap = inst.nbjob * inst.nbmachine
        n, m = ctypes.c_int(), ctypes.c_int()
        mach = np.zeros(cap, np.int32)
        dur = np.zeros(cap, np.int32)
        got = ref.ref_parse(path.encode(), ctypes.byref(n), ctypes.byref(m),
                            mach.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                            dur.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), cap)
        assert got == cap and (n.value, m.value) == (inst.nbjob, inst.nbmachine), path
        np.testing.assert_array_equal(inst.machine.ravel(), mach, err_msg=path)
        np.testing.assert_array_equal(inst.duration.ravel(), dur, err_msg=path)
