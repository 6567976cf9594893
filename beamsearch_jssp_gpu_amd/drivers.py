"""The reference's command-line drivers and benchmark runners on the engine.

    python -m beamsearch_jssp_gpu_amd.drivers ich      [instance] [beam_width]
    python -m beamsearch_jssp_gpu_amd.drivers pms      [instance] [beam_width]            [--ranks R]
    python -m beamsearch_jssp_gpu_amd.drivers gpu_v1   [instance] [beam_width] [iters]    [--ranks R]
    python -m beamsearch_jssp_gpu_amd.drivers gpu_v2   [instance] [beam_width] [iters]    [--ranks R]
    python -m beamsearch_jssp_gpu_amd.drivers hybrid_v1 | hybrid_v2
                                  [instance] [gpu_beam_width] [accel_factor] [iters]      [--ranks R --threads T]
    python -m beamsearch_jssp_gpu_amd.drivers bench_pms | bench_gpu_v1 | bench_gpu_v2     [--ranks R]
    python -m beamsearch_jssp_gpu_amd.drivers bench_hybrid_v1 | bench_hybrid_v2           [--ranks R --threads T]
    python -m beamsearch_jssp_gpu_amd.drivers width_study                                 [--ranks R]

Each command follows one reference program: arguments, defaults, the
multi-rank pass loop and the printed lines and table columns
(cpu/v0_ich/main.cpp, cpu/v1_pms/main.cpp, gpu/v1/main.cu,
gpu/v2_twophases/main.cu, hybrid/v1/main.cu, hybrid/v2_twophases/main.cu,
gpu/v1/benchmark.cu, cpu/v1_pms/benchmark.cpp, hybrid/v1/benchmark.cu,
cpu/profiling/time_width_study.cpp).  `--ranks R` stands for `mpirun -np R`:
the R ranks run on the one device, rank r seeded srand(time_seed + r).
`--time-seed` fixes the value the reference reads from time(NULL), so runs
repeat exactly.  `--dir D` resolves instance names from D (the reference
resolves relative to its working directory).  Times are wall-clock seconds
(the reference prints clock() CPU seconds).
"""
from __future__ import annotations

import argparse
import os
import sys
import time
from typing import Optional

from . import (MODE_PMS_EXACT, MODE_PMS_GPU, JSSPInstance, beam_search_ich, default_engine,
               parse_instance, resolve_instance_file)

# gpu/v1/benchmark.cu:30-36, cpu/v1_pms/benchmark.cpp:25-31
LAWRENCE_IDS = [f"{i:02d}" for i in range(1, 41)]
# cpu/profiling/time_width_study.cpp:20-26
WIDTH_STUDY_IDS = ["10", "15", "21", "31"]
WIDTH_STUDY_WIDTHS = [100, 200, 400, 800, 1600, 3200, 6400, 12800, 25600]


def lawrence_beam_width(idx: int) -> int:
    """The K rule of the Lawrence runners: 1000 for la01-la10, 5000 after
    (gpu/v1/benchmark.cu:52, cpu/v1_pms/benchmark.cpp:48)."""
    return 1000 if idx < 10 else 5000


def run_passes(inst: JSSPInstance, beam_width: int, ranks: int, iterations: int, time_seed: int, mode: int,
               phases: int = 2):
    """The multi-rank pass loop; returns (global best, per-pass best)."""
    res, pass_best = default_engine().run_passes(inst, beam_width, ranks, iterations, time_seed, mode=mode,
                                                 phases=phases)
    return res.makespan, pass_best


def _resolve(arg: str) -> Optional[str]:
    return resolve_instance_file(arg)


def _row(beam_width: int, path: str, inst: JSSPInstance, best: int, secs: float) -> str:
    # "%-8d %-30s %-6d %-9d %-8d %-12d %-10.2f" (gpu/v1/benchmark.cu:73-76)
    return (f"{beam_width:<8d} {path:<30s} {inst.nbjob:<6d} {inst.nbmachine:<9d} {inst.nboperation:<8d} "
            f"{best:<12d} {secs:<10.2f}")


def _header(first: str) -> str:
    return (f"\n{first:<8s} {'Instance':<30s} {'Jobs':<6s} {'Machines':<9s} {'Ops':<8s} {'BestSpan':<12s} "
            f"{'Time(s)':<10s}\n" + "-" * 79)


def cmd_ich(a) -> int:
    """cpu/v0_ich/main.cpp: srand(12345); beam_search_ich(inst, K)."""
    path = "./la33.txt"
    if a.instance is not None:
        path = _resolve(a.instance)
        if path is None:
            print(f"Error: cannot locate instance file for: {a.instance}", file=sys.stderr)
            return 1
    K = a.beam_width if a.beam_width is not None else 5000
    if K < 1:
        print("Error: beam_width must be >= 1", file=sys.stderr)
        return 1
    inst = parse_instance(path)
    print(f"  Instance File : {path}\n  Jobs          : {inst.nbjob}\n  Machines      : {inst.nbmachine}\n"
          f"  Beam Width (K): {K}")
    t0 = time.perf_counter()
    best = beam_search_ich(inst, K, seed=12345)
    secs = time.perf_counter() - t0
    print(f"  ===========================\n  Scheduled Ops    : {inst.nbjob * inst.nbmachine}\n"
          f"  Best Makespan    : {best}\n  Execution Time   : {secs:g} seconds")
    return 0


def cmd_pms(a) -> int:
    """cpu/v1_pms/main.cpp: R ranks x 100 iterations of beam_search_pms."""
    path = "./la33.txt"
    if a.instance is not None:
        path = _resolve(a.instance)
        if path is None:
            print(f"Error: cannot locate instance file for: {a.instance}", file=sys.stderr)
            return 1
    K = a.beam_width if a.beam_width is not None else 5000
    if K < 1:
        print("Error: beam_width must be >= 1", file=sys.stderr)
        return 1
    iters = a.iterations if a.iterations is not None else 100
    inst = parse_instance(path)
    print(f"  Instance : {path}  ({inst.nbjob} jobs x {inst.nbmachine} machines)")
    print(f"  Beam     : {K}  |  Ranks : {a.ranks}  |  Iterations : {iters}")
    t0 = time.perf_counter()
    best, pass_best = run_passes(inst, K, a.ranks, iters, a.time_seed, MODE_PMS_EXACT)
    secs = time.perf_counter() - t0
    for it in range(0, iters, 10):
        print(f"  iter {it:4d} | global best = {int(pass_best[it])}")
    print(f"  ===========================\n  Best Makespan  : {best}\n  Execution Time : {secs:.2f} seconds")
    return 0


def cmd_gpu(a, phases: int) -> int:
    """gpu/v1/main.cu (K default 1000) and gpu/v2_twophases/main.cu (5000)."""
    path = "./la33.txt"
    if a.instance is not None:
        path = _resolve(a.instance)
        if path is None:
            print(f"Error: cannot locate instance file: {a.instance}", file=sys.stderr)
            return 1
    K = a.beam_width if a.beam_width is not None else (1000 if phases == 1 else 5000)
    iters = a.iterations if a.iterations is not None else 100
    inst = parse_instance(path)
    bar = "=" * 28 if phases == 1 else "=" * 39
    print("=== GPU Beam Search (v1) ===" if phases == 1 else "=== GPU Beam Search (v2 Two-Phases) ===")
    print(f"Instance: {path}  |  Jobs: {inst.nbjob}  |  Machines: {inst.nbmachine}")
    print(f"Beam Width (K): {K}  |  MPI Ranks: {a.ranks}  |  GPUs/Node: 1  |  Iterations: {iters}")
    print(bar)
    t0 = time.perf_counter()
    best, _ = run_passes(inst, K, a.ranks, iters, a.time_seed, MODE_PMS_GPU, phases)
    secs = time.perf_counter() - t0
    print(bar)
    print(f"Best Makespan  : {best}\nExecution Time : {secs:.2f} seconds")
    return 0


def cmd_bench(a, solver: str) -> int:
    """The Lawrence tables: gpu/v1/benchmark.cu (50 iterations per
    instance) and cpu/v1_pms/benchmark.cpp (100)."""
    mode, phases = (MODE_PMS_EXACT, 2) if solver == "pms" else (MODE_PMS_GPU, 1 if solver == "gpu_v1" else 2)
    iters = a.iterations if a.iterations is not None else (100 if solver == "pms" else 50)
    ids = a.only.split(",") if a.only else LAWRENCE_IDS
    print(_header("Kbest"))
    for idx, iid in enumerate(LAWRENCE_IDS):
        if iid not in ids:
            continue
        path = _resolve(iid)
        if path is None:
            print(f"SKIP la{iid}  (file not found)")
            continue
        inst = parse_instance(path)
        K = a.beam_width if a.beam_width is not None else lawrence_beam_width(idx)
        t0 = time.perf_counter()
        best, _ = run_passes(inst, K, a.ranks, iters, a.time_seed, mode, phases)
        print(_row(K, path, inst, best, time.perf_counter() - t0), flush=True)
    print("\nAll Lawrence instances finished." if solver != "pms" else "\nAll instances finished.")
    return 0


def hybrid_threads(a) -> int:
    """OpenMP threads per rank: --threads, else OMP_NUM_THREADS, else the
    host's thread count (omp_get_max_threads in hybrid/v1/main.cu:54)."""
    if a.threads:
        return a.threads
    env = os.environ.get("OMP_NUM_THREADS")
    return max(1, int(env)) if env else max(1, os.cpu_count() or 1)


def cmd_hybrid(a, phases: int) -> int:
    """hybrid/v1/main.cu and hybrid/v2_twophases/main.cu: instance,
    gpu_beam_width (5000), accel_factor (34), iterations (100)."""
    path = "./la33.txt"
    if a.instance is not None:
        path = _resolve(a.instance)
        if path is None:
            print(f"Error: cannot locate instance file: {a.instance}", file=sys.stderr)
            return 1
    Kg = max(1, a.beam_width) if a.beam_width is not None else 5000
    accel = max(1, a.iters) if a.iters is not None else 34
    iters = max(1, a.extra) if a.extra is not None else (a.iterations if a.iterations is not None else 100)
    Kc = max(1, Kg // accel)
    T = hybrid_threads(a)
    inst = parse_instance(path)
    bar = "=" * 40 if phases == 1 else "=" * 51
    print("=== Hybrid CPU+GPU Beam Search (v1) ===" if phases == 1 else
          "=== Hybrid CPU+GPU Beam Search (v2 Two-Phases) ===")
    print(f"Instance: {path}  |  Jobs: {inst.nbjob}  |  Machines: {inst.nbmachine}")
    print(f"GPU Beam: {Kg}  |  CPU Beam: {Kc}  |  Accel Factor: {accel}")
    print(f"MPI Ranks: {a.ranks}  |  GPUs/Node: 1  |  OpenMP Threads/Rank: {T}  |  Iterations: {iters}")
    print(bar)
    t0 = time.perf_counter()
    res, pass_best = default_engine().run_hybrid(inst, Kg, accel, T, a.ranks, iters, a.time_seed, phases=phases)
    secs = time.perf_counter() - t0
    for it in range(19, iters, 20):
        print(f"  iter {it + 1:3d} / {iters:3d} | best makespan = {int(pass_best[it])}")
    print(bar)
    print(f"Best Makespan  : {res.makespan}\nExecution Time : {secs:.2f} seconds")
    return 0


def cmd_bench_hybrid(a, phases: int) -> int:
    """hybrid/v1/benchmark.cu, hybrid/v2_twophases/benchmark.cu: the Lawrence
    table, 50 iterations, accel factor 34."""
    iters = a.iterations if a.iterations is not None else 50
    T = hybrid_threads(a)
    ids = a.only.split(",") if a.only else LAWRENCE_IDS
    print(_header("Kbest"))
    for idx, iid in enumerate(LAWRENCE_IDS):
        if iid not in ids:
            continue
        path = _resolve(iid)
        if path is None:
            print(f"SKIP la{iid}  (file not found)")
            continue
        inst = parse_instance(path)
        Kg = lawrence_beam_width(idx)
        t0 = time.perf_counter()
        res, _ = default_engine().run_hybrid(inst, Kg, 34, T, a.ranks, iters, a.time_seed, phases=phases)
        print(_row(Kg, path, inst, res.makespan, time.perf_counter() - t0), flush=True)
    print("\nAll Lawrence instances finished.")
    return 0


def cmd_width_study(a) -> int:
    """cpu/profiling/time_width_study.cpp: PMS, 100 iterations per width,
    rank r of width w seeded srand(time + r + w)."""
    iters = a.iterations if a.iterations is not None else 100
    widths = [int(x) for x in a.widths.split(",")] if a.widths else WIDTH_STUDY_WIDTHS
    print(_header("Width"))
    for iid in WIDTH_STUDY_IDS:
        path = _resolve(iid)
        if path is None:
            print(f"SKIP la{iid} (file not found)")
            continue
        inst = parse_instance(path)
        print(f"\n==== Instance {path} ====")
        for w, K in enumerate(widths):
            t0 = time.perf_counter()
            best, _ = run_passes(inst, K, a.ranks, iters, a.time_seed + w, MODE_PMS_EXACT)
            print(_row(K, path, inst, best, time.perf_counter() - t0), flush=True)
    print("\nTime vs Beam Width study finished.")
    return 0


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m beamsearch_jssp_gpu_amd.drivers", description=__doc__,
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("command", choices=["ich", "pms", "gpu_v1", "gpu_v2", "hybrid_v1", "hybrid_v2", "bench_pms",
                                        "bench_gpu_v1", "bench_gpu_v2", "bench_hybrid_v1", "bench_hybrid_v2",
                                        "width_study"])
    ap.add_argument("instance", nargs="?")
    ap.add_argument("beam_width", nargs="?", type=int)
    ap.add_argument("iters", nargs="?", type=int, help="gpu_*: iterations; hybrid_*: accel factor")
    ap.add_argument("extra", nargs="?", type=int, help="hybrid_*: iterations")
    ap.add_argument("--threads", type=int, default=0, help="hybrid_*: threads per rank (OpenMP threads)")
    ap.add_argument("--ranks", type=int, default=1, help="MPI world size replayed on the device")
    ap.add_argument("--time-seed", type=int, default=None, help="value of time(NULL) (default: now)")
    ap.add_argument("--iterations", type=int, default=None, help="passes (default: the program's own)")
    ap.add_argument("--dir", default=None, help="directory instance names are resolved from")
    ap.add_argument("--only", default=None, help="bench_*: comma-separated Lawrence ids (e.g. 01,02)")
    ap.add_argument("--widths", default=None, help="width_study: comma-separated beam widths")
    a = ap.parse_args(argv)
    if a.time_seed is None:
        a.time_seed = int(time.time())
    if a.iters is not None and a.iterations is None and not a.command.startswith("hybrid"):
        a.iterations = a.iters
    if a.ranks < 1:
        print("Error: --ranks must be >= 1", file=sys.stderr)
        return 1
    if a.dir:
        os.chdir(a.dir)
    if a.command.startswith("bench_") and (a.instance is not None):
        # the Lawrence runners take no positional arguments
        ap.error(f"{a.command} takes no instance argument")
    if a.command == "ich":
        return cmd_ich(a)
    if a.command == "pms":
        return cmd_pms(a)
    if a.command in ("gpu_v1", "gpu_v2"):
        return cmd_gpu(a, 1 if a.command == "gpu_v1" else 2)
    if a.command in ("hybrid_v1", "hybrid_v2"):
        return cmd_hybrid(a, 1 if a.command == "hybrid_v1" else 2)
    if a.command in ("bench_hybrid_v1", "bench_hybrid_v2"):
        return cmd_bench_hybrid(a, 1 if a.command == "bench_hybrid_v1" else 2)
    if a.command.startswith("bench_"):
        return cmd_bench(a, a.command[len("bench_"):])
    return cmd_width_study(a)


if __name__ == "__main__":
    sys.exit(main())
