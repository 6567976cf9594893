#!/usr/bin/env python3
"""Benchmark of the beam step on one MI355X.

Headline workload: 50 jobs x 20 machines (seeded, durations uniform [1,99]),
beam width K=4096, operation-level branching (ICH_DET), two-phase path
(score every child, select the K survivors, materialise only those).  A step
is one complete solve (1000 levels, ~1.65e8 expanded states) with the
instance already uploaded; value = expanded states / second over the timed
steps.

Also reported on the same JSON line:
  roofline         dominant kernel of that solve: algorithmic bytes per launch
                   / mean launch duration (HIP events on the engine's stream,
                   from a profiled solve of the same workload); `traffic` is
                   the HBM bytes per launch measured with rocprofv3 counters
                   (read from profiles/traffic_50x20_k4096.json when present)
  kernels          per-kernel mean launch time and algorithmic GB/s
  expand_single_phase  the single-phase expand kernel (writes every child)
  multiseed        16 machine-focused (PMS_GPU) beams sharing one device bound,
                   as one batch, vs the same 16 solves one after another
  batched          256 uneven instances in one batched solve vs the plain
                   one-instance-at-a-time loop (same results checked)
  small_configs    10x5 K=64 BS-ICH (ICH_EXACT) over 32 instances and 20x20
                   K=1024 single-phase PMS_GPU / ICH_DET over 16 instances
  serving          Poisson arrivals into the continuous-batching server and
                   into the plain one-at-a-time loop, both run in real time
                   on the same request list: solves/s, p50/p99 at 2x the
                   plain loop's capacity and (load_sweep) at 0.5x, 1x and 4x;
                   then every request queued at once (the server's capacity
                   vs the plain loop's, 1 / mean service time)
  instance_seeds_1_8  the headline solve on the instances of generator seeds
                   1..8 (states/s over all eight, makespans)
  cpu_baseline     the reference's own BS-ICH (cpu/v0_ich) compiled into
                   oracle/_ref from the reference sources (or the oracle's
                   restatement when that library is absent), one thread, one
                   solve of the same instance at K=2048 (~15 s), beside the
                   GPU's bit-exact ICH_EXACT solve of it; and (all_cores) the
                   reference's BS-PMS (cpu/v1_pms) on every instance of the
                   batch and serving workloads, one process per core
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
# Two-phase solves of this workload score each level inside the previous
# level's rank kernel (k_rank_fused_b), so the score kind has no launches.
KERNEL_SYMBOLS = {1: ["k_expand_b", "k_filter_b", "k_rank_b"], 2: ["k_score_b", "k_filter_b", "k_rank_fused_b"]}
TRAFFIC_FILE = os.path.join(REPO, "profiles", "traffic_50x20_k4096.json")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def profile_kinds(eng, inst, K, mode, phases, seed):
    eng.set_profiling(True)
    eng.solve(inst, K, mode=mode, phases=phases, seed=seed)
    kinds = []
    for k in range(len(KIND_NAMES)):
        ms, n, b = eng.kernel_time(k)
        kinds.append((k, ms, n, b))
    eng.set_profiling(False)
    return kinds


def full_width(eng, k):
    """The kernel's rate over the full-width levels of the last profiled
    solve (algorithmic bytes at least 95% of the widest level's), beside the
    mean over all its launches: the narrow first and last levels are mostly
    fixed per-launch cost."""
    ms, b = eng.kernel_launches(k)
    ok = ms > 0
    if not ok.any():
        return None
    wide = ok & (b >= 0.95 * b[ok].max())
    t = float(ms[wide].sum()) / 1e3
    gbs = float(b[wide].sum()) / t / 1e9
    return {"levels": int(wide.sum()), "of": int(ok.sum()), "avg_launch_us": round(t / wide.sum() * 1e6, 3),
            "algorithmic_bytes_per_launch": round(float(b[wide].mean()), 1), "achieved": round(gbs, 1),
            "frac": round(gbs / HBM_PEAK_GBS, 4)}


def roofline_obj(kind, phases, traffic=None):
    k, ms, n, b = kind
    avg_s = ms / 1e3 / max(1, n)
    per_launch = b / max(1, n)
    achieved = per_launch / avg_s / 1e9 if avg_s > 0 else 0.0
    sym = KERNEL_SYMBOLS[phases][k]
    tr = None
    if traffic and sym in traffic.get("kernels", {}):
        tr = traffic["kernels"][sym].get("hbm_bytes_per_launch")
    return {"bound": "hbm", "kernel": sym, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": tr,
            "avg_launch_us": round(avg_s * 1e6, 3), "launches": n,
            "algorithmic_bytes_per_launch": round(per_launch, 1)}


def load_traffic():
    try:
        with open(TRAFFIC_FILE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def same(a, b):
    return a.makespan == b.makespan and ((a.sequence is None and b.sequence is None) or
                                         (a.sequence is not None and b.sequence is not None and
                                          (a.sequence == b.sequence).all()))


def multiseed(eng, bs, inst, K, seed, beams):
    """16 PMS_GPU beams with a shared device bound, as one batch, vs the same
    solves one after another with the bound carried the same way (one pass:
    every beam prunes with the initial bound)."""
    eng.solve(inst, K, mode=bs.MODE_PMS_GPU, seed=seed, n_seeds=beams)  # warm-up
    t0 = time.perf_counter()
    r = eng.solve(inst, K, mode=bs.MODE_PMS_GPU, seed=seed, n_seeds=beams)
    t_multi = time.perf_counter() - t0
    t0 = time.perf_counter()
    plain = [eng.solve(inst, K, mode=bs.MODE_PMS_GPU, seed=seed + i) for i in range(beams)]
    t_plain = time.perf_counter() - t0
    best = min(p.makespan for p in plain)
    out = {"workload": f"PMS_GPU 50x20 K={K}, {beams} beams (seeds {seed}..{seed + beams - 1}), one device bound",
           "beams_per_s": round(beams / t_multi, 2), "plain_beams_per_s": round(beams / t_plain, 2),
           "speedup": round(t_plain / t_multi, 2), "makespan": r.makespan, "plain_best_makespan": best,
           "equal": r.makespan == best, "expanded_states_per_s": round(r.expanded_states / t_multi, 1)}
    # Two passes of beams/2: the second pass prunes with the first pass's
    # best (the device bound read at init), against the same loop run one
    # solve at a time with the bound passed as best_known.
    half = beams // 2
    eng.solve(inst, K, mode=bs.MODE_PMS_GPU, seed=seed, n_seeds=half, n_passes=2)  # warm-up
    t0 = time.perf_counter()
    r2 = eng.solve(inst, K, mode=bs.MODE_PMS_GPU, seed=seed, n_seeds=half, n_passes=2)
    t_two = time.perf_counter() - t0
    t0 = time.perf_counter()
    p1 = [eng.solve(inst, K, mode=bs.MODE_PMS_GPU, seed=seed + i) for i in range(half)]
    b1 = min(p.makespan for p in p1)
    p2 = [eng.solve(inst, K, mode=bs.MODE_PMS_GPU, seed=seed + half + i, best_known=b1) for i in range(half)]
    t_two_plain = time.perf_counter() - t0
    b2 = min([b1] + [p.makespan for p in p2 if p.found])
    unbounded = sum(p.expanded_states for p in plain[half:beams])
    out["two_passes"] = {
        "workload": f"{half} beams per pass, 2 passes; pass 2 prunes with pass 1's best ({b1})",
        "beams_per_s": round(beams / t_two, 2), "plain_beams_per_s": round(beams / t_two_plain, 2),
        "speedup": round(t_two_plain / t_two, 2), "makespan": r2.makespan, "plain_best_makespan": b2,
        "equal": r2.makespan == b2,
        "pass2_expanded_with_bound": sum(p.expanded_states for p in p2), "pass2_expanded_without_bound": unbounded}
    return out


def batch_workload(bs, count=256, gen_seed=2026):
    from beamsearch_jssp_gpu_amd import workloads
    specs = workloads.batch_specs(count, gen_seed)
    insts = [bs.generate_instance(s["gen_seed"], s["n"], s["m"], workloads.DMIN, workloads.DMAX) for s in specs]
    params = [dict(beam_width=s["K"], mode=bs.MODE_PMS_GPU, phases=2, best_known=bs.NO_MAKESPAN, seed=s["seed"])
              for s in specs]
    return insts, params


def batched(eng, bs, count, use_graph=True):
    insts, params = batch_workload(bs, count)
    eng.solve_batch(insts, params, use_graph=use_graph)  # warm-up: autotunes the lane shapes, captures the graph
    t0 = time.perf_counter()
    rb = eng.solve_batch(insts, params, use_graph=use_graph)
    t_batch = time.perf_counter() - t0
    log("batched: the same instances one at a time")
    t0 = time.perf_counter()
    rp = [eng.solve(i, p["beam_width"], mode=p["mode"], seed=p["seed"]) for i, p in zip(insts, params)]
    t_plain = time.perf_counter() - t0
    eq = all(same(a, b) for a, b in zip(rb, rp))
    return {"workload": f"{count} instances, n in {{10,15,20,30,50}}, m in {{5,10,15,20}}, K in {{256,1024}}, "
                        "durations U[1,99], generator seed 2026, PMS_GPU two-phase",
            "solves_per_s": round(count / t_batch, 2), "plain_solves_per_s": round(count / t_plain, 2),
            "speedup": round(t_plain / t_batch, 2), "identical_results": eq,
            "expanded_states_per_s": round(sum(r.expanded_states for r in rb) / t_batch, 1)}


def serve_workload(bs, count, seed=7):
    from beamsearch_jssp_gpu_amd import workloads
    specs, gaps = workloads.serve_specs(count, seed)  # gaps: unit-rate Poisson process, scaled below
    reqs = [(bs.generate_instance(s["gen_seed"], s["n"], s["m"], workloads.DMIN, workloads.DMAX), s["K"])
            for s in specs]
    return reqs, gaps


def pct(v, q):
    v = sorted(v)
    return v[min(len(v) - 1, max(0, int(round(q * (len(v) - 1)))))]


def plain_realtime(eng, bs, reqs, arr, plain):
    """The plain loop in real time on the same arrivals: one request at a
    time, each solved as soon as it has arrived and the previous one is done.
    Returns (latencies, span, results equal to the reference solves)."""
    lat, t0 = [], time.perf_counter()
    eq = True
    for i, (inst, K) in enumerate(reqs):
        wait = arr[i] - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        r = eng.solve(inst, K, mode=bs.MODE_PMS_GPU, seed=12345)
        lat.append(time.perf_counter() - t0 - arr[i])
        eq = eq and r.makespan == plain[i].makespan and (r.sequence == plain[i].sequence).all()
    span = max(a + l for a, l in zip(arr, lat)) - arr[0]
    return lat, span, eq


def serve_poisson(eng, bs, reqs, gaps, plain, dur, budget_gib, load):
    """One Poisson run at `load` x the plain loop's capacity (1 / its mean
    service time): the plain loop and the server each run in real time on
    the same arrivals, and their latencies (completion - arrival) are
    measured."""
    count = len(reqs)
    mean_service = sum(dur) / len(dur)
    rate = load / mean_service
    arr, t = [], 0.0
    for g in gaps:
        t += g / rate
        arr.append(t)
    lat_plain, plain_span, eq_plain = plain_realtime(eng, bs, reqs, arr, plain)
    srv = bs.Server(eng, int(budget_gib * 2**30), max_slots=64)
    tickets, done, lat = [None] * count, {}, [0.0] * count
    nxt = 0
    t0 = time.perf_counter()
    while len(done) < count:
        now = time.perf_counter() - t0
        while nxt < count and arr[nxt] <= now:
            tickets[nxt] = srv.submit(reqs[nxt][0], reqs[nxt][1], mode=bs.MODE_PMS_GPU, seed=12345)
            nxt += 1
        active = srv.step(16)
        now = time.perf_counter() - t0
        for i in range(nxt):
            if i not in done:
                r = srv.poll(tickets[i])
                if r is not None:
                    done[i] = r
                    lat[i] = now - arr[i]
        if active == 0 and len(done) == nxt and nxt < count:
            time.sleep(max(0.0, arr[nxt] - (time.perf_counter() - t0)))
    served_span = max(arr[i] + lat[i] for i in range(count)) - arr[0]
    srv.close()
    eq = eq_plain and all(done[i].makespan == plain[i].makespan and (done[i].sequence == plain[i].sequence).all()
                          for i in range(count))
    return {"load": load, "arrival_rate_per_s": round(rate, 1),
            "solves_per_s": round(count / served_span, 2), "plain_solves_per_s": round(count / plain_span, 2),
            "p50_ms": round(pct(lat, 0.5) * 1e3, 2), "p99_ms": round(pct(lat, 0.99) * 1e3, 2),
            "plain_p50_ms": round(pct(lat_plain, 0.5) * 1e3, 2),
            "plain_p99_ms": round(pct(lat_plain, 0.99) * 1e3, 2), "identical_results": eq}


def serving(eng, bs, count, budget_gib, load, sweep=(0.5, 1.0, 4.0)):
    reqs, gaps = serve_workload(bs, count)
    # Plain loop: solve times measured one by one.
    for inst, K in reqs[:2]:
        eng.solve(inst, K, mode=bs.MODE_PMS_GPU, seed=12345)
    plain, dur = [], []
    for inst, K in reqs:
        t0 = time.perf_counter()
        plain.append(eng.solve(inst, K, mode=bs.MODE_PMS_GPU, seed=12345))
        dur.append(time.perf_counter() - t0)
    mean_service = sum(dur) / len(dur)
    log(f"serving: plain service times measured ({mean_service * 1e3:.1f} ms mean); Poisson run at {load}x")
    main = serve_poisson(eng, bs, reqs, gaps, plain, dur, budget_gib, load)
    runs = []
    for x in sweep:
        log(f"serving: Poisson run at {x}x")
        runs.append(serve_poisson(eng, bs, reqs, gaps, plain, dur, budget_gib, x))
    eq = main["identical_results"] and all(r["identical_results"] for r in runs)
    # Saturation: every request queued at t=0, so the server's capacity shows
    # rather than the arrival rate (the plain loop's capacity is 1/mean service).
    log("serving: every request queued at once")
    srv = bs.Server(eng, int(budget_gib * 2**30), max_slots=64)
    t0 = time.perf_counter()
    tk = [srv.submit(inst, K, mode=bs.MODE_PMS_GPU, seed=12345) for inst, K in reqs]
    srv.drain()
    sat_span = time.perf_counter() - t0
    sat = [srv.poll(x) for x in tk]
    srv.close()
    eq = eq and all(sat[i].makespan == plain[i].makespan and (sat[i].sequence == plain[i].sequence).all()
                    for i in range(count))
    return {"workload": f"{count} requests, Poisson arrivals (seed 7) at {load}x the plain loop's capacity "
                        f"({main['arrival_rate_per_s']}/s), n in [10,100], m in [5,20], K in {{256,1024,4096}}, "
                        f"PMS_GPU, HBM budget {budget_gib} GiB, 64 slots",
            "solves_per_s": main["solves_per_s"], "plain_solves_per_s": main["plain_solves_per_s"],
            "p50_ms": main["p50_ms"], "p99_ms": main["p99_ms"],
            "plain_p50_ms": main["plain_p50_ms"], "plain_p99_ms": main["plain_p99_ms"],
            "saturated_solves_per_s": round(count / sat_span, 2),
            "saturated_speedup": round((count / sat_span) / (1.0 / mean_service), 2),
            "load_sweep": runs,
            "identical_results": eq}


def _ref_pms_job(job):
    """Worker of the all-cores baseline (a fresh process: no GPU state): the
    reference's own BS-PMS, srand(seed); beam_search_pms(inst, K, 500000)."""
    ref_so, n, m, mach, dur, K, seed = job
    lib = ctypes.CDLL(ref_so)
    I32P = ctypes.POINTER(ctypes.c_int32)
    lib.ref_pms.argtypes = [ctypes.c_int, ctypes.c_int, I32P, I32P, ctypes.c_int, ctypes.c_int, ctypes.c_uint]
    a = (ctypes.c_int32 * (n * m))(*mach)
    b = (ctypes.c_int32 * (n * m))(*dur)
    t0 = time.perf_counter()
    ms = lib.ref_pms(n, m, a, b, K, 500000, seed)
    return ms, time.perf_counter() - t0


def cpu_all_cores(bs, insts, ks, cores):
    """The reference CPU BS-PMS (cpu/v1_pms, the CPU form of the machine-focused
    search the batch and the server run) on every instance of a workload,
    one process per core, longest first: solves/s over the wall time."""
    import multiprocessing
    from concurrent.futures import ProcessPoolExecutor
    ref_so = os.path.join(REPO, "oracle", "_ref", "libjssp_ref.so")
    if not os.path.exists(ref_so):
        return None
    jobs = [(ref_so, i.nbjob, i.nbmachine, i.machine.ravel().tolist(), i.duration.ravel().tolist(), K, 12345)
            for i, K in zip(insts, ks)]
    order = sorted(range(len(jobs)), key=lambda j: -jobs[j][1] * jobs[j][1] * jobs[j][5])
    # Workers start clean (spawn) and never touch the device; a worker that
    # dies raises here instead of hanging the pool.
    with ProcessPoolExecutor(cores, mp_context=multiprocessing.get_context("spawn")) as pool:
        list(pool.map(_ref_pms_job, [jobs[order[-1]]] * cores))  # start every worker first
        t0 = time.perf_counter()
        res = list(pool.map(_ref_pms_job, [jobs[j] for j in order], chunksize=1))
        wall = time.perf_counter() - t0
    return {"solves_per_s": round(len(jobs) / wall, 3), "wall_s": round(wall, 2),
            "cpu_s": round(sum(r[1] for r in res), 1), "longest_solve_s": round(max(r[1] for r in res), 2),
            "solves": len(jobs)}


def cpu_cores():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(inst, eng, bs, K=2048, seed=12345):
    """The reference BS-ICH on this host, one thread, same instance.  The
    number of states it expands is the GPU ICH_EXACT solve's count (the same
    trajectory: its makespan is checked equal)."""
    ref_so = os.path.join(REPO, "oracle", "_ref", "libjssp_ref.so")
    n, m = inst.nbjob, inst.nbmachine
    I32P = ctypes.POINTER(ctypes.c_int32)
    mp, dp = inst.machine.ctypes.data_as(I32P), inst.duration.ctypes.data_as(I32P)
    t0 = time.perf_counter()
    exact = eng.solve(inst, K, mode=bs.MODE_ICH_EXACT, seed=seed)
    exact_wall = time.perf_counter() - t0
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
                      f"{exact.makespan}: {ms == exact.makespan}. K={K} rather than the headline's 4096 "
                      f"keeps the sample near 15 s (K=4096 takes about twice as long; the rate per state "
                      f"changes little with K)",
            "cpu_model": _cpu_model(),
            "gpu_ich_exact": {"workload": f"the same solve on the GPU (ICH_EXACT: the reference's rand() "
                                          f"stream and heap replayed, bit-exact), K={K}",
                              "device_ms": round(exact.device_ms, 2), "wall_ms": round(exact_wall * 1e3, 2),
                              "states_per_s": round(exact.expanded_states / (exact_wall), 1)}}


def small_configs(eng, bs):
    """The two small configurations on the GPU: BS-ICH on 10x5 (durations
    U[5,99], instance seeds 1..32, K=64; ICH_EXACT, the reference's own
    rand()/heap semantics) and single-phase PMS_GPU / ICH_DET on 20x20
    (U[1,99], seeds 1..16, K=1024).  Wall time per solve includes the upload
    and readback."""
    out = {}
    for name, (n, m, dmin, dmax, seeds, K, mode, phases) in {
            "ich_exact_10x5_k64": (10, 5, 5, 99, 32, 64, bs.MODE_ICH_EXACT, 2),
            "pms_gpu_20x20_k1024_single_phase": (20, 20, 1, 99, 16, 1024, bs.MODE_PMS_GPU, 1),
            "ich_det_20x20_k1024_single_phase": (20, 20, 1, 99, 16, 1024, bs.MODE_ICH_DET, 1)}.items():
        insts = [bs.generate_instance(sd, n, m, dmin, dmax) for sd in range(1, seeds + 1)]
        for inst in insts[:2]:
            eng.solve(inst, K, mode=mode, phases=phases, seed=12345)
        t0 = time.perf_counter()
        res = [eng.solve(inst, K, mode=mode, phases=phases, seed=12345) for inst in insts]
        dt = time.perf_counter() - t0
        states = sum(r.expanded_states for r in res)
        out[name] = {"instances": seeds, "ms_per_solve": round(dt / seeds * 1e3, 3),
                     "device_ms_per_solve": round(sum(r.device_ms for r in res) / seeds, 3),
                     "states_per_s": round(states / dt, 1), "makespans": [r.makespan for r in res]}
    return out


def cpu_small_config(bs):
    """The reference's own BS-ICH (oracle/_ref) on the 10x5 configuration, one
    thread: the CPU side of small_configs' first entry."""
    ref_so = os.path.join(REPO, "oracle", "_ref", "libjssp_ref.so")
    if not os.path.exists(ref_so):
        return None
    lib = ctypes.CDLL(ref_so)
    I32P = ctypes.POINTER(ctypes.c_int32)
    lib.ref_ich.argtypes = [ctypes.c_int, ctypes.c_int, I32P, I32P, ctypes.c_int, ctypes.c_uint]
    insts = [bs.generate_instance(sd, 10, 5, 5, 99) for sd in range(1, 33)]
    t0 = time.perf_counter()
    ms = [lib.ref_ich(10, 5, i.machine.ctypes.data_as(I32P), i.duration.ctypes.data_as(I32P), 64, 12345)
          for i in insts]
    dt = time.perf_counter() - t0
    return {"workload": "BS-ICH 10x5, durations U[5,99], instance seeds 1..32, K=64, srand(12345)",
            "kind": "reference", "cores": 1, "ms_per_solve": round(dt / 32 * 1e3, 3), "makespans": ms}


def _cpu_model():
    """The host CPU as /proc/cpuinfo names it (the first processor): model
    name, family / model / stepping, clock, and the logical CPUs this process
    may run on."""
    fields = {}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if not line.strip():
                    break
                k, _, v = line.partition(":")
                fields[k.strip()] = v.strip()
    except OSError:
        return "unknown"
    name = fields.get("model name", "unknown")
    extra = [f"{k} {fields[k]}" for k in ("cpu family", "model", "stepping") if k in fields]
    if "cpu MHz" in fields:
        extra.append(f"{float(fields['cpu MHz']):.0f} MHz")
    try:
        extra.append(f"{len(os.sched_getaffinity(0))} logical CPUs available")
    except (AttributeError, OSError):
        pass
    return f"{name} ({', '.join(extra)})" if extra else name


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--jobs", type=int, default=50)
    ap.add_argument("--machines", type=int, default=20)
    ap.add_argument("--beam", type=int, default=4096)
    ap.add_argument("--phases", type=int, default=2, choices=[1, 2])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-extras", action="store_true", help="headline and roofline only")
    ap.add_argument("--batch-only", action="store_true", help="the batched measurement only")
    ap.add_argument("--no-graph", action="store_true",
                    help="batches enqueue their step kernels directly instead of replaying captured graphs")
    ap.add_argument("--serving-only", action="store_true",
                    help="only the serving measurement (its JSON object), e.g. under a profiler")
    ap.add_argument("--cpu-beam", type=int, default=2048)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--requests", type=int, default=64)
    ap.add_argument("--serve-load", type=float, default=2.0)
    ap.add_argument("--budget-gib", type=float, default=8.0)
    ap.add_argument("--tune-table", default=None, metavar="PATH",
                    help="batch-shape tuning table to load (and extend) instead of tuning in memory; a "
                         "profile of the batch can then leave the tuning runs out")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run three device-to-device copies of MIB MiB (reference point for the "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        nb = args.copy_calibration * 2**20
        a = torch.ones(nb // 4, dtype=torch.int32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": nb, "copy_bytes_written": nb, "copies": 3}), flush=True)
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one MI355X: the engine drives one GPU per process (independent solves "
              "would be spread one per GPU); --gpus must be 1", file=sys.stderr)
        sys.exit(2)

    import beamsearch_jssp_gpu_amd as bs
    if args.batch_only:
        eng = bs.Engine(0)
        if args.tune_table:
            eng.set_autotune(True, args.tune_table)
        out = batched(eng, bs, args.batch, use_graph=not args.no_graph)
        eng.close()
        print(json.dumps({"batched": out, "use_graph": not args.no_graph}), flush=True)
        return
    if args.serving_only:
        eng = bs.Engine(0)
        out = serving(eng, bs, args.requests, args.budget_gib, args.serve_load)
        eng.close()
        print(json.dumps({"serving": out}), flush=True)
        return
    n, m, K, seed = args.jobs, args.machines, args.beam, 12345
    inst = bs.generate_instance(1, n, m, 1, 99)
    eng = bs.Engine(0)
    mode, phases = bs.MODE_ICH_DET, args.phases
    log(f"headline: {n}x{m} K={K} phases={phases}, {args.warmup} warm-up + {args.steps} timed solves")
    for _ in range(args.warmup):
        eng.solve(inst, K, mode=mode, phases=phases, seed=seed)
    t0 = time.perf_counter()
    res = [eng.solve(inst, K, mode=mode, phases=phases, seed=seed) for _ in range(args.steps)]
    wall = time.perf_counter() - t0
    # The timed region of each solve is its device time (HIP events on the
    # engine's stream, from the root's init to the sequence backtrack): the
    # instance is already in HBM then.  The wall-clock rate, which also
    # includes uploading the op table and descriptor and reading the result
    # back over PCIe, is reported beside it.
    dt = sum(r.device_ms for r in res) / 1e3
    states = sum(r.expanded_states for r in res)
    assert len({(r.makespan, tuple(r.sequence)) for r in res}) == 1, "non-deterministic result"
    assert bs.validate_sequence(inst, res[0].sequence) == res[0].makespan

    log("profiled solve (HIP events per launch)")
    traffic = load_traffic()
    kinds = profile_kinds(eng, inst, K, mode, phases, seed)
    dom = max(kinds, key=lambda k: k[1])
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
        "value_pcie_inclusive": round(states / wall, 1),
        "config": {"workload": f"{'two' if phases == 2 else 'single'}-phase operation-level beam search "
                               f"(ICH_DET), {n}x{m}, K={K}, durations U[1,99], instance seed 1, solve seed {seed}",
                   "jobs": n, "machines": m, "beam_width": K, "levels": n * m,
                   "expanded_states_per_solve": res[0].expanded_states, "makespan": res[0].makespan,
                   "device_ms_per_solve": round(sum(r.device_ms for r in res) / len(res), 3)},
        "roofline": roofline_obj(dom, phases, traffic),
        "kernels": {KIND_NAMES[k[0]]: {"symbol": KERNEL_SYMBOLS[phases][k[0]], "total_ms": round(k[1], 3),
                                       "launches": k[2], "avg_us": round(k[1] * 1e3 / max(1, k[2]), 3),
                                       "algorithmic_bytes_per_launch": round(k[3] / max(1, k[2]), 1),
                                       "algorithmic_GBps": round(k[3] / (k[1] / 1e3) / 1e9, 1) if k[1] else 0}
                    for k in kinds if k[2] > 0},
    }
    line["roofline"]["full_width_levels"] = full_width(eng, dom[0])
    if not args.no_extras:
        log("headline over instance seeds 1..8")
        rs8 = []
        for sd in range(1, 9):
            i8 = bs.generate_instance(sd, n, m, 1, 99)
            r8 = eng.solve(i8, K, mode=mode, phases=phases, seed=seed)
            assert bs.validate_sequence(i8, r8.sequence) == r8.makespan
            rs8.append(r8)
        line["instance_seeds_1_8"] = {
            "workload": f"the headline solve on {n}x{m} instances of generator seeds 1..8",
            "states_per_s": round(sum(r.expanded_states for r in rs8) / (sum(r.device_ms for r in rs8) / 1e3), 1),
            "ms_per_solve": round(sum(r.device_ms for r in rs8) / len(rs8), 3),
            "makespans": [r.makespan for r in rs8]}
        if phases == 2:
            log("single-phase expand kernel")
            line["expand_single_phase"] = roofline_obj(profile_kinds(eng, inst, K, mode, 1, seed)[0], 1, traffic)
            line["expand_single_phase"]["full_width_levels"] = full_width(eng, 0)
        log("multi-seed: 16 PMS beams")
        line["multiseed"] = multiseed(eng, bs, inst, K, 1, 16)
        log(f"batched: {args.batch} uneven instances")
        line["batched"] = batched(eng, bs, args.batch)
        log(f"serving: {args.requests} Poisson requests")
        line["serving"] = serving(eng, bs, args.requests, args.budget_gib, args.serve_load)
        log("small configurations: 10x5 K=64 BS-ICH, 20x20 K=1024 single-phase")
        line["small_configs"] = small_configs(eng, bs)
    if not args.no_cpu:
        log("CPU baseline (reference BS-ICH, one thread)")
        line["cpu_baseline"] = cpu_baseline(inst, eng, bs, K=args.cpu_beam, seed=seed)
        small = cpu_small_config(bs)
        if small is not None:
            if "small_configs" in line:
                gpu_ms = line["small_configs"]["ich_exact_10x5_k64"]["makespans"]
                small["equal_makespans"] = small["makespans"] == gpu_ms
            line["cpu_baseline"]["small_config"] = small
        if not args.no_extras:
            cores = cpu_cores()
            log(f"CPU baseline: reference BS-PMS on {cores} cores, the batch and the serving workloads")
            bi, bp = batch_workload(bs, args.batch)
            allc = {"cores": cores, "kind": "reference",
                    "solver": "cpu/v1_pms beam_search_pms compiled from the reference sources, srand(12345), "
                              "best_known 500000, one process per core",
                    "batch": cpu_all_cores(bs, bi, [p["beam_width"] for p in bp], cores)}
            sreqs, _ = serve_workload(bs, args.requests)
            allc["serving"] = cpu_all_cores(bs, [r[0] for r in sreqs], [r[1] for r in sreqs], cores)
            line["cpu_baseline"]["all_cores"] = allc
    eng.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
