"""Test helpers: the oracle (CPU restatement, oracle/oracle.cpp) through ctypes,
fixture loading and instance builders.  The oracle is only ever the checker."""
from __future__ import annotations

import ctypes
import json
import os
import subprocess

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(REPO, "tests", "golden")
ORACLE_SO = os.path.join(REPO, "oracle", "build", "liboracle.so")
REF_SO = os.path.join(REPO, "oracle", "_ref", "libjssp_ref.so")
I32P = ctypes.POINTER(ctypes.c_int32)


def p32(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(I32P)


_oracle = None


def oracle():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            subprocess.run(["make", "-C", os.path.join(REPO, "oracle"), "oracle"], check=True,
                           stdout=subprocess.DEVNULL)
        lib = ctypes.CDLL(ORACLE_SO)
        ll = ctypes.POINTER(ctypes.c_longlong)
        ip = ctypes.POINTER(ctypes.c_int)
        lib.orc_ich_exact.argtypes = [ctypes.c_int, ctypes.c_int, I32P, I32P, ctypes.c_int, ctypes.c_uint,
                                      ctypes.c_int, I32P, ll]
        lib.orc_pms_exact.argtypes = [ctypes.c_int, ctypes.c_int, I32P, I32P, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_uint, ctypes.c_int, ll]
        lib.orc_gpu_solve.argtypes = [ctypes.c_int, ctypes.c_int, I32P, I32P, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_uint, I32P, ll, ip, I32P, I32P]
        lib.orc_simulate.argtypes = [ctypes.c_int, ctypes.c_int, I32P, I32P, I32P]
        lib.orc_bruteforce.argtypes = [ctypes.c_int, ctypes.c_int, I32P, I32P]
        lib.orc_pivot_hash.argtypes = [ctypes.c_uint32]
        lib.orc_pivot_hash.restype = ctypes.c_uint32
        lib.orc_noise_hash.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]
        lib.orc_noise_hash.restype = ctypes.c_uint32
        lib.orc_gpu_score.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint32]
        lib.orc_gpu_score.restype = ctypes.c_float
        lib.orc_glibc_rand.argtypes = [ctypes.c_uint, ctypes.c_int, I32P]
        lib.orc_glibc_rand.restype = None
        lib.orc_pms_driver.argtypes = [ctypes.c_int, ctypes.c_int, I32P, I32P, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_uint, I32P, ip, ll]
        lib.orc_gpu_driver.argtypes = [ctypes.c_int, ctypes.c_int, I32P, I32P, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_uint, I32P, ip, ll, I32P]
        lib.orc_hybrid_driver.argtypes = [ctypes.c_int, ctypes.c_int, I32P, I32P, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint, I32P, ll]
        _oracle = lib
    return _oracle


def ref_lib():
    """The compiled reference (oracle/_ref), or None where it was not built."""
    if not os.path.exists(REF_SO):
        return None
    lib = ctypes.CDLL(REF_SO)
    lib.ref_ich.argtypes = [ctypes.c_int, ctypes.c_int, I32P, I32P, ctypes.c_int, ctypes.c_uint]
    lib.ref_pms.argtypes = [ctypes.c_int, ctypes.c_int, I32P, I32P, ctypes.c_int, ctypes.c_int, ctypes.c_uint]
    lib.ref_pms_driver.argtypes = [ctypes.c_int, ctypes.c_int, I32P, I32P, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_uint, I32P]
    return lib


class Inst:
    """Minimal instance holder usable without the product library."""

    def __init__(self, machine, duration, name=""):
        self.machine = np.ascontiguousarray(machine, dtype=np.int32)
        self.duration = np.ascontiguousarray(duration, dtype=np.int32)
        if self.machine.ndim == 1:
            raise ValueError("2-D arrays expected")
        self.name = name

    @property
    def n(self):
        return self.machine.shape[0]

    @property
    def m(self):
        return self.machine.shape[1]


def splitmix_instance(seed: int, n: int, m: int, dmin: int, dmax: int) -> Inst:
    """Independent Python statement of the generator stream (splitmix64,
    Fisher-Yates per job, then the job's durations)."""
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
    for j in range(n):
        perm = list(range(m))
        for k in range(m - 1, 0, -1):
            r = nxt() % (k + 1)
            perm[k], perm[r] = perm[r], perm[k]
        for k in range(m):
            mach[j, k] = perm[k]
            dur[j, k] = dmin + nxt() % (dmax - dmin + 1)
    return Inst(mach, dur, f"gen{n}x{m}s{seed}")


def random_instance(rng: np.random.Generator, n: int, m: int, dmin=1, dmax=99, perm=True) -> Inst:
    if perm:
        mach = np.array([rng.permutation(m) for _ in range(n)], dtype=np.int32)
    else:  # machines may repeat inside a job
        mach = rng.integers(0, m, size=(n, m)).astype(np.int32)
    dur = rng.integers(dmin, dmax + 1, size=(n, m)).astype(np.int32)
    return Inst(mach, dur)


def load_json(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def golden_instances():
    out = {}
    for k, d in load_json("lawrence.json").items():
        out[k] = Inst(np.array(d["machine"]).reshape(d["n"], d["m"]), np.array(d["duration"]).reshape(d["n"], d["m"]), k)
    for k, d in load_json("golden_ref.json")["randoms"].items():
        out[k] = Inst(np.array(d["machine"]).reshape(d["n"], d["m"]), np.array(d["duration"]).reshape(d["n"], d["m"]), k)
    return out


# ---- oracle wrappers --------------------------------------------------------
def orc_ich(inst: Inst, K: int, seed: int):
    ex = ctypes.c_longlong()
    seq = np.zeros(inst.n * inst.m, np.int32)
    ms = oracle().orc_ich_exact(inst.n, inst.m, p32(inst.machine), p32(inst.duration), K, seed, 1, p32(seq),
                                ctypes.byref(ex))
    return ms, seq, ex.value


def orc_pms(inst: Inst, K: int, best_known: int, seed: int):
    ex = ctypes.c_longlong()
    ms = oracle().orc_pms_exact(inst.n, inst.m, p32(inst.machine), p32(inst.duration), K, best_known, seed, 1,
                                ctypes.byref(ex))
    return ms, ex.value


def orc_gpu(inst: Inst, K: int, mode: int, best_known: int, seed: int, trace=True):
    L = inst.n * inst.m
    seq = np.zeros(L, np.int32)
    ex = ctypes.c_longlong()
    found = ctypes.c_int()
    trail = np.full(max(0, L - 1) * K, -1, np.int32) if trace else None
    pcount = np.zeros(L + 1, np.int32) if trace else None
    ms = oracle().orc_gpu_solve(inst.n, inst.m, p32(inst.machine), p32(inst.duration), K, mode, best_known,
                                seed & 0xFFFFFFFF, p32(seq), ctypes.byref(ex), ctypes.byref(found),
                                p32(trail) if trace else None, p32(pcount) if trace else None)
    return {"makespan": ms, "found": bool(found.value), "sequence": seq if found.value else None,
            "expanded": ex.value, "trail": trail, "pcount": pcount}


def orc_rand(seed: int, count: int) -> np.ndarray:
    out = np.zeros(count, np.int32)
    oracle().orc_glibc_rand(seed & 0xFFFFFFFF, count, p32(out))
    return out


def libc_rand(seed: int, count: int) -> np.ndarray:
    """srand(seed) then `count` rand() calls of the C library itself."""
    libc = ctypes.CDLL(None)
    libc.srand.argtypes = [ctypes.c_uint]
    libc.rand.restype = ctypes.c_int
    libc.srand(seed & 0xFFFFFFFF)
    return np.array([libc.rand() for _ in range(count)], np.int32)


def orc_pms_driver(inst: Inst, K: int, ranks: int, iterations: int, time_seed: int):
    pb = np.zeros(iterations, np.int32)
    bi, ex = ctypes.c_int(), ctypes.c_longlong()
    ms = oracle().orc_pms_driver(inst.n, inst.m, p32(inst.machine), p32(inst.duration), K, ranks, iterations,
                                 time_seed & 0xFFFFFFFF, p32(pb), ctypes.byref(bi), ctypes.byref(ex))
    return {"makespan": ms, "pass_best": pb, "best_index": bi.value, "expanded": ex.value}


def orc_gpu_driver(inst: Inst, K: int, ranks: int, iterations: int, time_seed: int):
    pb = np.zeros(iterations, np.int32)
    seq = np.zeros(inst.n * inst.m, np.int32)
    bi, ex = ctypes.c_int(), ctypes.c_longlong()
    ms = oracle().orc_gpu_driver(inst.n, inst.m, p32(inst.machine), p32(inst.duration), K, ranks, iterations,
                                 time_seed & 0xFFFFFFFF, p32(pb), ctypes.byref(bi), ctypes.byref(ex), p32(seq))
    return {"makespan": ms, "pass_best": pb, "best_index": bi.value, "expanded": ex.value,
            "sequence": seq if bi.value >= 0 else None}


def orc_hybrid_driver(inst: Inst, gpu_k: int, cpu_k: int, threads: int, ranks: int, iterations: int,
                      time_seed: int):
    pb = np.zeros(iterations, np.int32)
    ex = ctypes.c_longlong()
    ms = oracle().orc_hybrid_driver(inst.n, inst.m, p32(inst.machine), p32(inst.duration), gpu_k, cpu_k, threads,
                                    ranks, iterations, time_seed & 0xFFFFFFFF, p32(pb), ctypes.byref(ex))
    return {"makespan": ms, "pass_best": pb, "expanded": ex.value}


def orc_simulate(inst, seq) -> int:
    s = np.ascontiguousarray(seq, np.int32)
    n, m = inst.machine.shape
    return oracle().orc_simulate(n, m, p32(inst.machine), p32(inst.duration), p32(s))


def orc_bruteforce(inst: Inst) -> int:
    return oracle().orc_bruteforce(inst.n, inst.m, p32(inst.machine), p32(inst.duration))


def lower_bound(inst) -> int:
    jl = inst.duration.sum(axis=1).max()
    ml = max(int(inst.duration[inst.machine == k].sum()) for k in range(inst.machine.shape[1]))
    return int(max(jl, ml))


def valid_trail_entries(trail, pcount, K):
    """Yields (level, q, parent, job) for every written trail entry."""
    L1 = len(pcount) - 2
    for lvl in range(L1):
        for q in range(pcount[lvl + 1]):
            t = int(trail[lvl * K + q])
            yield lvl, q, t >> 8, t & 0xFF
