#!/usr/bin/env python3
"""Benchmark of the beam step on one MI355X.

Headline workload: 50 jobs x 20 machines (seeded, durations uniform [1,99]),
beam width K=4096, operation-level branching (ICH_DET), two-phase path
(score every child, select the K survivors, materialise only those).  A step
is one complete solve (1000 levels, ~10^8 expanded states) with the instance
already uploaded; value = expanded states / second over the timed steps.

Also reported on the same line:
  roofline      dominant kernel of that solve: algorithmic bytes per launch /
                mean launch duration (HIP events on the engine's stream, from a
                profiled solve of the same workload after the timed steps)
  expand_single_phase   the same for the single-phase expand kernel, which
                writes every child snapshot (the HBM-bound kernel)
  cpu_baseline  the reference's own BS-ICH (cpu/v0_ich, compiled into
                oracle/_ref) on the same instance, one thread, K=1024
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0
KIND_NAMES = ["expand_score", "select_filter", "rank_materialize"]
KERNEL_SYMBOLS = ["k_score", "k_filter", "k_rank"]


def roofline(eng, inst, K, mode, phases, seed):
    eng.set_profiling(True)
    eng.solve(inst, K, mode=mode, phases=phases, seed=seed)
    kinds = []
    for k in range(len(KIND_NAMES)):
        ms, n, b = eng.kernel_time(k)
        kinds.append((k, ms, n, b))
    eng.set_profiling(False)
    return kinds


def roofline_obj(kind):
    k, ms, n, b = kind
    avg_s = ms / 1e3 / max(1, n)
    per_launch = b / max(1, n)
    achieved = per_launch / avg_s / 1e9 if avg_s > 0 else 0.0
    return {"bound": "hbm", "kernel": KERNEL_SYMBOLS[k], "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "avg_launch_us": round(avg_s * 1e6, 3), "launches": n,
            "algorithmic_bytes_per_launch": round(per_launch, 1)}


def cpu_baseline(inst, eng, bs, K=1024, seed=12345):
    """The reference BS-ICH on this host, one thread, same instance, K=1024.
    The number of states it expands is the GPU ICH_EXACT solve's count (the
    same trajectory; its makespan is checked equal)."""
    ref_so = os.path.join(REPO, "oracle", "_ref", "libjssp_ref.so")
    n, m = inst.nbjob, inst.nbmachine
    I32P = ctypes.POINTER(ctypes.c_int32)
    mp, dp = inst.machine.ctypes.data_as(I32P), inst.duration.ctypes.data_as(I32P)
    exact = eng.solve(inst, K, mode=bs.MODE_ICH_EXACT, seed=seed)
    if os.path.exists(ref_so):
        lib = ctypes.CDLL(ref_so)
        lib.ref_ich.argtypes = [ctypes.c_int, ctypes.c_int, I32P, I32P, ctypes.c_int, ctypes.c_uint]
        kind, label = "reference", "cpu/v0_ich beam_search_ich compiled from the reference sources (g++ -O3)"
        t0 = time.perf_counter()
        ms = lib.ref_ich(n, m, mp, dp, K, seed)
        dt = time.perf_counter() - t0
    else:
        sys.path.insert(0, os.path.join(REPO, "tests"))
        from testlib import oracle
        lib = oracle()
        ex = ctypes.c_longlong()
        kind, label = "port", "oracle/oracle.cpp orc_ich_exact (restatement of cpu/v0_ich)"
        t0 = time.perf_counter()
        ms = lib.orc_ich_exact(n, m, mp, dp, K, seed, 1, None, ctypes.byref(ex))
        dt = time.perf_counter() - t0
    return {"value": round(exact.expanded_states / dt, 1), "unit": "states/s", "cores": 1, "kind": kind,
            "sample": f"one full solve of the same 50x20 instance at K={K} (srand({seed})), {label}; "
                      f"{exact.expanded_states} states in {dt:.2f} s; makespan {ms} == GPU ICH_EXACT "
                      f"{exact.makespan}: {ms == exact.makespan}",
            "cpu_model": _cpu_model()}


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--jobs", type=int, default=50)
    ap.add_argument("--machines", type=int, default=20)
    ap.add_argument("--beam", type=int, default=4096)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-beam", type=int, default=1024)
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one MI355X: this build runs on a single GPU (independent solves would shard "
              "one per GPU, see DESIGN.md); --gpus must be 1", file=sys.stderr)
        sys.exit(2)

    import beamsearch_jssp_gpu_amd as bs
    n, m, K, seed = args.jobs, args.machines, args.beam, 12345
    inst = bs.generate_instance(1, n, m, 1, 99)
    eng = bs.Engine(0)
    mode, phases = bs.MODE_ICH_DET, 2
    for _ in range(args.warmup):
        eng.solve(inst, K, mode=mode, phases=phases, seed=seed)
    t0 = time.perf_counter()
    res = [eng.solve(inst, K, mode=mode, phases=phases, seed=seed) for _ in range(args.steps)]
    dt = time.perf_counter() - t0
    states = sum(r.expanded_states for r in res)
    assert len({(r.makespan, tuple(r.sequence)) for r in res}) == 1, "non-deterministic result"
    assert bs.validate_sequence(inst, res[0].sequence) == res[0].makespan

    kinds = roofline(eng, inst, K, mode, phases, seed)
    dom = max(kinds, key=lambda k: k[1])
    single = roofline(eng, inst, K, mode, 1, seed)
    line = {
        "metric": "expanded states/s at 50x20 K=4096; batched solves/s + p99 latency vs plain",
        "value": round(states / dt, 1),
        "unit": "states/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": f"two-phase operation-level beam search (ICH_DET), {n}x{m}, K={K}, "
                               f"durations U[1,99], instance seed 1, solve seed {seed}",
                   "jobs": n, "machines": m, "beam_width": K, "levels": n * m,
                   "expanded_states_per_solve": res[0].expanded_states, "makespan": res[0].makespan,
                   "device_ms_per_solve": round(sum(r.device_ms for r in res) / len(res), 3)},
        "roofline": roofline_obj(dom),
        "kernels": {KIND_NAMES[k[0]]: {"total_ms": round(k[1], 3), "launches": k[2],
                                       "avg_us": round(k[1] * 1e3 / max(1, k[2]), 3)} for k in kinds},
        "expand_single_phase": roofline_obj(single[0]),
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(inst, eng, bs, K=args.cpu_beam, seed=seed)
    eng.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
