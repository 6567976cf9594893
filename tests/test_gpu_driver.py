"""The reference's multi-rank pass drivers on one device (jssp_run_passes,
jssp_solve_seeded): each result must equal the oracle's restatement of the
driver loop (gpu/v1/main.cu:55-71, cpu/v1_pms/main.cpp:55-73) run on the
same rank streams."""
import numpy as np
import pytest

import beamsearch_jssp_gpu_amd as bs
from testlib import golden_instances, orc_gpu, orc_gpu_driver, orc_pms_driver, orc_rand, splitmix_instance

pytestmark = pytest.mark.gpu


def _b(i):
    return bs.JSSPInstance(i.machine, i.duration, i.name)


@pytest.mark.parametrize("phases", [1, 2])
def test_gpu_driver_equals_oracle_loop(engine, phases):
    inst = splitmix_instance(31, 15, 10, 1, 99)
    K, R, E, t = 64, 4, 3, 1700000000
    print("gpu driver", phases, flush=True)
    got, pass_best = engine.run_passes(_b(inst), K, R, E, t, mode=bs.MODE_PMS_GPU, phases=phases)
    want = orc_gpu_driver(inst, K, R, E, t)
    assert got.makespan == want["makespan"]
    np.testing.assert_array_equal(pass_best, want["pass_best"])
    assert got.expanded_states == want["expanded"]
    assert got.best_seed_index == max(0, want["best_index"])
    assert got.found == (want["best_index"] >= 0)
    if got.found:
        np.testing.assert_array_equal(got.sequence, want["sequence"])
        assert bs.validate_sequence(_b(inst), got.sequence) == got.makespan


def test_pms_exact_driver_equals_oracle_loop(engine):
    inst = golden_instances()["la01"]
    K, R, E, t = 16, 3, 3, 4242
    print("pms driver", flush=True)
    got, pass_best = engine.run_passes(_b(inst), K, R, E, t, mode=bs.MODE_PMS_EXACT)
    want = orc_pms_driver(inst, K, R, E, t)
    assert got.makespan == want["makespan"]
    np.testing.assert_array_equal(pass_best, want["pass_best"])
    assert got.expanded_states == want["expanded"]
    assert got.best_seed_index == max(0, want["best_index"])
    assert bs.validate_sequence(_b(inst), got.sequence) == got.makespan


def test_seeded_passes_equal_single_solves(engine):
    """jssp_solve_seeded: beam i of pass e uses seeds[e][i] and prunes with
    the best of the passes before e -- the same as single solves."""
    inst = splitmix_instance(12, 12, 8, 1, 99)
    seeds = np.array([[5, 2**31 + 7, 999], [123456789, 0, 2**32 - 1]], dtype=np.int64)
    K = 32
    for mode in (bs.MODE_PMS_GPU, bs.MODE_ICH_DET):
        got = engine.solve_seeded(_b(inst), K, seeds, mode=mode)
        bound, best_idx, best_seq, exp = bs.NO_MAKESPAN, -1, None, 0
        for e in range(seeds.shape[0]):
            pb = bound
            for i in range(seeds.shape[1]):
                r = engine.solve(_b(inst), K, mode=mode, best_known=pb, seed=int(seeds[e, i]))
                exp += r.expanded_states
                if r.found and r.makespan < bound:
                    bound, best_idx, best_seq = r.makespan, e * seeds.shape[1] + i, r.sequence
        assert got.makespan == bound and got.expanded_states == exp
        assert got.best_seed_index == max(0, best_idx)
        np.testing.assert_array_equal(got.sequence, best_seq)
    # and one of them against the oracle directly
    w = orc_gpu(inst, K, bs.MODE_PMS_GPU, bs.NO_MAKESPAN, int(seeds[0, 1]), trace=False)
    r = engine.solve(_b(inst), K, mode=bs.MODE_PMS_GPU, seed=int(seeds[0, 1]))
    assert r.makespan == w["makespan"] and r.expanded_states == w["expanded"]


def test_driver_seeds_are_rank_streams(engine):
    """One rank, one pass: the solve seed is the first rand() after
    srand(time_seed), as gpu/v1/main.cu:60-63."""
    inst = splitmix_instance(3, 10, 6, 1, 99)
    t = 31337
    got, _ = engine.run_passes(_b(inst), 32, 1, 1, t, mode=bs.MODE_PMS_GPU, phases=2)
    want = engine.solve(_b(inst), 32, mode=bs.MODE_PMS_GPU, seed=int(orc_rand(t, 1)[0]))
    assert got.makespan == want.makespan and got.expanded_states == want.expanded_states
    np.testing.assert_array_equal(got.sequence, want.sequence)


def test_driver_argument_errors(engine):
    inst = splitmix_instance(3, 6, 4, 1, 99)
    for kw in (dict(mode=bs.MODE_ICH_DET), dict(mode=bs.MODE_PMS_GPU, ranks=0), dict(mode=bs.MODE_PMS_GPU,
                                                                                       iterations=0)):
        args = dict(ranks=2, iterations=1, mode=bs.MODE_PMS_GPU)
        args.update(kw)
        with pytest.raises(bs.JsspError):
            engine.run_passes(_b(inst), 8, args["ranks"], args["iterations"], 1, mode=args["mode"])


def test_lawrence_runner_rows_equal_oracle(engine, tmp_path, capsys):
    """bench_gpu_v2 / bench_pms on la01-la02 (instance files written from the
    fixture): every BestSpan equals the oracle's driver loop with the
    runner's K rule and seeds."""
    import os
    from beamsearch_jssp_gpu_amd import drivers
    from testlib import load_json
    law = load_json("lawrence.json")
    for name in ("la01", "la02"):
        d = law[name]
        with open(tmp_path / f"{name}.txt", "w") as f:
            f.write(f"{d['n']} {d['m']}\n")
            for j in range(d["n"]):
                row = range(j * d["m"], (j + 1) * d["m"])
                f.write(" ".join(f"{d['machine'][i]} {d['duration'][i]}" for i in row) + "\n")
    insts = golden_instances()
    cwd = os.getcwd()
    for cmd, iters in (("bench_gpu_v2", 3), ("bench_pms", 2)):
        print(cmd, flush=True)
        try:
            assert drivers.main([cmd, "--dir", str(tmp_path), "--only", "01,02", "--iterations", str(iters),
                                 "--time-seed", "20240601", "--ranks", "2"]) == 0
        finally:
            os.chdir(cwd)
        rows = [l.split() for l in capsys.readouterr().out.splitlines() if l.startswith("1000 ")]
        assert [r[1] for r in rows] == ["./la01.txt", "./la02.txt"]
        for r, name in zip(rows, ("la01", "la02")):
            assert r[2:5] == ["10", "5", "52"]
            if cmd == "bench_pms":
                want = orc_pms_driver(insts[name], 1000, 2, iters, 20240601)["makespan"]
            else:
                want = orc_gpu_driver(insts[name], 1000, 2, iters, 20240601)["makespan"]
            assert int(r[5]) == want, (cmd, name)


@pytest.mark.parametrize("phases", [1, 2])
def test_hybrid_driver_equals_oracle_loop(engine, phases):
    """jssp_run_hybrid (hybrid/v1/main.cu loop): device passes batched per
    iteration, host BS-PMS passes on threads; equal to the oracle's loop."""
    from testlib import orc_hybrid_driver
    inst = splitmix_instance(41, 12, 6, 1, 99)
    Kg, accel, T, R, E, t = 64, 8, 3, 2, 3, 1700000123
    print("hybrid driver", phases, flush=True)
    got, pass_best = engine.run_hybrid(_b(inst), Kg, accel, T, R, E, t, phases=phases)
    want = orc_hybrid_driver(inst, Kg, Kg // accel, T, R, E, t)
    assert got.makespan == want["makespan"]
    np.testing.assert_array_equal(pass_best, want["pass_best"])
    assert got.expanded_states == want["expanded"]
    assert got.found
    assert bs.validate_sequence(_b(inst), got.sequence) == got.makespan


def test_hybrid_iteration_is_min_of_its_passes(engine):
    inst = splitmix_instance(43, 10, 5, 1, 99)
    Kg, Kc, T, bk, gseed, cbase = 32, 4, 4, bs.NO_MAKESPAN, 987654, 42000
    got = engine.hybrid_iteration(_b(inst), Kg, Kc, bk, gseed, cbase, T, phases=2)
    g = engine.solve(_b(inst), Kg, mode=bs.MODE_PMS_GPU, phases=2, best_known=bk, seed=gseed)
    hs = [bs.host_solve_pms(_b(inst), Kc, bk, cbase + tid) for tid in range(1, T)]
    parts = [g.makespan] + [h.makespan for h in hs]
    assert got.makespan == min([bk] + parts)
    assert got.expanded_states == g.expanded_states + sum(h.expanded_states for h in hs)
    first = [i for i, v in enumerate(parts) if v == got.makespan][0]
    assert got.best_seed_index == first
    assert bs.validate_sequence(_b(inst), got.sequence) == got.makespan


def test_hybrid_cli_equals_oracle(engine, tmp_path, capsys):
    """python -m beamsearch_jssp_gpu_amd.drivers hybrid_v2 la01 64 8 2
    --threads 3 (hybrid/v2_twophases/main.cu arguments): the printed best
    makespan equals the oracle's hybrid loop."""
    import os
    from beamsearch_jssp_gpu_amd import drivers
    from testlib import load_json, orc_hybrid_driver
    d = load_json("lawrence.json")["la01"]
    with open(tmp_path / "la01.txt", "w") as f:
        f.write(f"{d['n']} {d['m']}\n")
        for j in range(d["n"]):
            f.write(" ".join(f"{d['machine'][i]} {d['duration'][i]}" for i in range(j * d["m"], (j + 1) * d["m"]))
                    + "\n")
    cwd = os.getcwd()
    try:
        assert drivers.main(["hybrid_v2", "01", "64", "8", "2", "--dir", str(tmp_path), "--threads", "3",
                             "--time-seed", "9"]) == 0
    finally:
        os.chdir(cwd)
    out = capsys.readouterr().out
    line = [l for l in out.splitlines() if l.startswith("Best Makespan")][0]
    want = orc_hybrid_driver(golden_instances()["la01"], 64, 8, 3, 1, 2, 9)["makespan"]
    assert int(line.split(":")[1]) == want
    assert "GPU Beam: 64  |  CPU Beam: 8  |  Accel Factor: 8" in out


def test_host_pms_equals_device_pms_exact(engine):
    """The host BS-PMS (hybrid workers) and the device-expanded PMS_EXACT mode
    replay the same reference search: same makespan, expanded count and
    sequence, including edge shapes (one job, one machine) and a tight bound."""
    from testlib import random_instance
    rng = np.random.default_rng(5)
    cases = [(1, 5), (6, 1), (8, 4), (12, 7)]
    for t, (n, m) in enumerate(cases):
        inst = random_instance(rng, n, m)
        b = _b(inst)
        for K, bk, seed in ((1, bs.NO_MAKESPAN, 3), (16, bs.NO_MAKESPAN, 77 + t), (64, 0, 5)):
            h = bs.host_solve_pms(b, K, bk, seed)
            d = engine.solve(b, K, mode=bs.MODE_PMS_EXACT, best_known=bk, seed=seed)
            assert (h.makespan, h.found, h.expanded_states) == (d.makespan, d.found, d.expanded_states), (n, m, K)
            if h.found:
                np.testing.assert_array_equal(h.sequence, d.sequence)


def test_width_study_rows_equal_oracle(engine, tmp_path, capsys):
    """width_study (cpu/profiling/time_width_study.cpp:20-81): instances la10,
    la15, la21, la31, one row per width with the columns Width, Instance,
    Jobs, Machines, Ops, BestSpan, Time(s); width w's ranks seed
    srand(time + rank + w) and run the PMS pass loop from 500000.  BestSpan
    equals the oracle's PMS driver loop with those seeds."""
    import os
    from beamsearch_jssp_gpu_amd import drivers
    from testlib import load_json
    law = load_json("lawrence.json")
    names = ["la10", "la15", "la21", "la31"]
    assert [f"la{i}" for i in drivers.WIDTH_STUDY_IDS] == names
    assert drivers.WIDTH_STUDY_WIDTHS == [100, 200, 400, 800, 1600, 3200, 6400, 12800, 25600]
    for name in names:
        d = law[name]
        with open(tmp_path / f"{name}.txt", "w") as f:
            f.write(f"{d['n']} {d['m']}\n")
            for j in range(d["n"]):
                row = range(j * d["m"], (j + 1) * d["m"])
                f.write(" ".join(f"{d['machine'][i]} {d['duration'][i]}" for i in row) + "\n")
    widths, iters, ranks, t = [100, 200], 2, 2, 1717171717
    cwd = os.getcwd()
    print("width_study", flush=True)
    try:
        assert drivers.main(["width_study", "--dir", str(tmp_path), "--widths", ",".join(map(str, widths)),
                             "--iterations", str(iters), "--time-seed", str(t), "--ranks", str(ranks)]) == 0
    finally:
        os.chdir(cwd)
    out = capsys.readouterr().out.splitlines()
    head = [l.split() for l in out if l.startswith("Width")]
    assert head == [["Width", "Instance", "Jobs", "Machines", "Ops", "BestSpan", "Time(s)"]]
    rows = [l.split() for l in out if l[:1].isdigit()]
    assert len(rows) == len(names) * len(widths)
    insts = golden_instances()
    k = 0
    for name in names:
        for w, K in enumerate(widths):
            r = rows[k]
            k += 1
            d = law[name]
            assert r[0] == str(K) and r[1] == f"./{name}.txt"
            assert r[2:5] == [str(d["n"]), str(d["m"]), str(d["n"] * d["m"] + 2)]
            want = orc_pms_driver(insts[name], K, ranks, iters, t + w)["makespan"]
            assert int(r[5]) == want, (name, K)
            assert float(r[6]) >= 0.0
    assert "Time vs Beam Width study finished." in out[-1]
