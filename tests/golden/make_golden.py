"""Writes the golden fixtures of tests/golden/ from the compiled reference.

Run in the build container (needs the reference tree and oracle/_ref):

    make -C oracle && python tests/golden/make_golden.py

Outputs (data only -- inputs and the reference's outputs):
  lawrence.json      the 40 Lawrence instances (lawrence_instances/la01-40.txt)
                     as 0-based machine / duration arrays, parsed by the
                     reference parser (cpu/common/instance_parser.cpp:37-93)
  golden_ref.json    makespans of the reference's own beam_search_ich
                     (cpu/v0_ich/solver.cpp) after srand(seed), and of one
                     beam_search_pms pass (cpu/v1_pms/solver.cpp) after
                     srand(seed), on Lawrence and on seeded random instances
  instances/         three raw files copied byte for byte (la01.txt, ta01,
                     ta71.txt) for the parser tests
"""
from __future__ import annotations

import ctypes
import json
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"
I32P = ctypes.POINTER(ctypes.c_int32)


def p32(a):
    return np.ascontiguousarray(a, dtype=np.int32).ctypes.data_as(I32P)


def splitmix_instance(seed: int, n: int, m: int, dmin: int, dmax: int):
    """Same stream as jssp_generate_instance (beamsearch_jssp_gpu_amd/csrc/host_io.cpp)."""
    mask = (1 << 64) - 1
    s = seed & mask

    def nxt():
        nonlocal s
        s = (s + 0x9E3779B97F4A7C15) & mask
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
        return z ^ (z >> 31)

    mach = np.zeros((n, m), np.int32)
    dur = np.zeros((n, m), np.int32)
    span = dmax - dmin + 1
    for j in range(n):
        perm = list(range(m))
        for k in range(m - 1, 0, -1):
            r = nxt() % (k + 1)
            perm[k], perm[r] = perm[r], perm[k]
        for k in range(m):
            mach[j, k] = perm[k]
            dur[j, k] = dmin + nxt() % span
    return mach, dur


def main():
    ref = ctypes.CDLL(os.path.join(REPO, "oracle", "_ref", "libjssp_ref.so"))
    ref.ref_ich.restype = ctypes.c_int
    ref.ref_pms.restype = ctypes.c_int
    # Lawrence instances through the reference parser.
    law = {}
    for i in range(1, 41):
        path = f"{REF}/lawrence_instances/la{i:02d}.txt"
        n, m = ctypes.c_int(), ctypes.c_int()
        mach = np.zeros(4096, np.int32)
        dur = np.zeros(4096, np.int32)
        tot = ref.ref_parse(path.encode(), ctypes.byref(n), ctypes.byref(m), p32(mach), p32(dur), 4096)
        assert tot == n.value * m.value, path
        law[f"la{i:02d}"] = {"n": n.value, "m": m.value, "machine": mach[:tot].tolist(),
                             "duration": dur[:tot].tolist()}
    with open(os.path.join(HERE, "lawrence.json"), "w") as f:
        json.dump(law, f, separators=(",", ":"))

    os.makedirs(os.path.join(HERE, "instances"), exist_ok=True)
    for src in ("lawrence_instances/la01.txt", "taillard_instances/ta01", "taillard_instances/ta71.txt"):
        shutil.copyfile(os.path.join(REF, src), os.path.join(HERE, "instances", os.path.basename(src)))

    cases = []

    def add_ich(name, n, m, mach, dur, K, seed):
        ms = ref.ref_ich(n, m, p32(mach), p32(dur), K, seed)
        cases.append({"solver": "ich", "instance": name, "K": K, "seed": seed, "makespan": ms})

    def add_pms(name, n, m, mach, dur, K, bk, seed):
        ms = ref.ref_pms(n, m, p32(mach), p32(dur), K, bk, seed)
        cases.append({"solver": "pms", "instance": name, "K": K, "seed": seed, "best_known": bk, "makespan": ms})

    for name, d in law.items():
        n, m = d["n"], d["m"]
        mach, dur = np.array(d["machine"], np.int32), np.array(d["duration"], np.int32)
        for K in (1, 8, 64):
            add_ich(name, n, m, mach, dur, K, 12345)
        if int(name[2:]) <= 20:
            add_ich(name, n, m, mach, dur, 256, 12345)
            add_pms(name, n, m, mach, dur, 64, 500000, 7)
            add_pms(name, n, m, mach, dur, 256, 500000, 11)
    # Seeded random instances (generator stream of jssp_generate_instance).
    randoms = {}
    for s in range(1, 33):
        mach, dur = splitmix_instance(s, 10, 5, 5, 99)
        name = f"r10x5s{s}"
        randoms[name] = {"seed": s, "n": 10, "m": 5, "dmin": 5, "dmax": 99,
                         "machine": mach.ravel().tolist(), "duration": dur.ravel().tolist()}
        add_ich(name, 10, 5, mach, dur, 64, 12345)
        add_pms(name, 10, 5, mach, dur, 64, 500000, 12345)
        # a tight bound: the ICH result, so pruning is exercised
        add_pms(name, 10, 5, mach, dur, 64, cases[-2]["makespan"], 12345)
    for s in range(1, 5):
        mach, dur = splitmix_instance(s, 20, 20, 1, 99)
        name = f"r20x20s{s}"
        randoms[name] = {"seed": s, "n": 20, "m": 20, "dmin": 1, "dmax": 99,
                         "machine": mach.ravel().tolist(), "duration": dur.ravel().tolist()}
        add_ich(name, 20, 20, mach, dur, 64, 12345)
        add_pms(name, 20, 20, mach, dur, 128, 500000, 99)
    with open(os.path.join(HERE, "golden_ref.json"), "w") as f:
        json.dump({"source": "reference cpu/v0_ich + cpu/v1_pms compiled by oracle/Makefile (g++ -O3 -std=c++11)",
                   "randoms": randoms, "cases": cases}, f, separators=(",", ":"))
    print(len(cases), "golden cases")


if __name__ == "__main__":
    sys.exit(main())
