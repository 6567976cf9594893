"""Per-phase time split of one level of the single-solve beam step (filter
kernel and fused rank kernel), from the phase-stamp build of the library.

    python -m beamsearch_jssp_gpu_amd.build --stamps      # builds lib_stamps/ (host, once)
    python -m beamsearch_jssp_gpu_amd.phase_stamps [--mode ich|pms] [--jobs 50 --machines 20 --beam 4096]

The stamp build writes wall-clock (100 MHz) stamps at the phase boundaries
of levels 100..107 of a solve: thread 0 of every filter block (entry, after
the histogram scan, after the threshold, after the tile pass) and lane 0 of
every rank wave (entry, keys staged, survivor ranked, child built, children
of the next level scored, block partials, histogram flushed).  The script
solves the workload three times on the stamp build and prints, per level,
each phase's mean / median / max duration over the waves (rank) or blocks
(filter), the kernels' spans and the gap between them, as a markdown table.
The stamps cost a few instructions per wave; the product library carries
none of them.
"""
from __future__ import annotations

import argparse
import ctypes
import os
import sys

import numpy as np

from . import build as _build

RANK_PHASES = ["keys staged", "survivor ranked", "child built", "next level scored", "block partials",
               "histogram flushed"]
FILTER_PHASES = ["histogram scan", "threshold", "tile pass"]
LEVEL0, LEVELS = 100, 8


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--mode", choices=["ich", "pms"], default="ich")
    ap.add_argument("--jobs", type=int, default=50)
    ap.add_argument("--machines", type=int, default=20)
    ap.add_argument("--beam", type=int, default=4096)
    a = ap.parse_args(argv)
    if not os.path.exists(_build.STAMPS_LIB):
        print(f"{_build.STAMPS_LIB} is missing: run `python -m beamsearch_jssp_gpu_amd.build --stamps`",
              file=sys.stderr)
        return 1
    import beamsearch_jssp_gpu_amd as bs
    bs.library_path = lambda: _build.STAMPS_LIB  # the stamp build, not the product library
    lib = bs.load_library()
    lib.jssp_debug_phase_stamps.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    mode = bs.MODE_ICH_DET if a.mode == "ich" else bs.MODE_PMS_GPU
    inst = bs.generate_instance(1, a.jobs, a.machines, 1, 99)
    eng = bs.Engine(0)
    for _ in range(3):
        r = eng.solve(inst, a.beam, mode=mode, phases=2, seed=12345)
    rank = np.zeros((LEVELS, 256, 16, 8), np.uint64)
    filt = np.zeros((LEVELS, 512, 4), np.uint64)
    rc = lib.jssp_debug_phase_stamps(rank.ctypes.data, filt.ctypes.data)
    eng.close()
    if rc:
        print("reading the stamps failed", file=sys.stderr)
        return 1
    R, F = rank.astype(np.int64), filt.astype(np.int64)
    us = 0.01  # one tick = 10 ns
    print(f"{a.jobs}x{a.machines} K={a.beam} {a.mode.upper()} two-phase, makespan {r.makespan}, "
          f"{r.device_ms:.2f} ms per solve; levels {LEVEL0}..{LEVEL0 + LEVELS - 1}, microseconds\n")
    print("| level | filter span | gap | rank span | " + " | ".join(RANK_PHASES) + " | " +
          " | ".join("filter " + p for p in FILTER_PHASES) + " |")
    print("|" + "---|" * (4 + len(RANK_PHASES) + len(FILTER_PHASES)))
    for L in range(LEVELS):
        f, w = F[L], R[L]
        fb = f[:, 0] > 0
        wv = w[:, :, 0] > 0
        if not fb.any() or not wv.any():
            continue
        f0, f3 = f[fb, 0].min(), f[fb, 3].max()
        r0, r6 = w[:, :, 0][wv].min(), w[:, :, 6][wv].max()
        cells = [f"{(f3 - f0) * us:.2f}", f"{(r0 - f3) * us:.2f}", f"{(r6 - r0) * us:.2f}"]
        for k in range(len(RANK_PHASES)):
            d = (w[:, :, k + 1] - w[:, :, k])[wv] * us
            cells.append(f"{d.mean():.2f} / {np.median(d):.2f} / {d.max():.2f}")
        for k in range(len(FILTER_PHASES)):
            d = (f[fb, k + 1] - f[fb, k]) * us
            cells.append(f"{d.mean():.2f}")
        print(f"| {LEVEL0 + L} | " + " | ".join(cells) + " |")
    print("\nrank phases: mean / median / max over the waves; filter phases: mean over the blocks")
    return 0


if __name__ == "__main__":
    sys.exit(main())
