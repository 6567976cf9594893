"""MI355X beam-search engine for job-shop scheduling (Python host side).

Thin ctypes layer over the C ABI in ``include/jssp.h`` (library
``beamsearch_jssp_gpu_amd/lib/libjssp_hip.so``).  It mirrors the reference's
solver API -- ``beam_search_ich`` (cpu/v0_ich/solver.h:6), ``beam_search_pms``
(cpu/v1_pms/solver.h:6), ``run_gpu_beam_search_v1`` / ``_v2``
(gpu/v1/solver.cuh:6-12, gpu/v2_twophases/solver.cuh:6-12) and
``parse_instance`` / ``resolve_instance_file`` (cpu/common/instance_parser.h)
-- so callers keep the same names, argument meanings and sentinel (500000).

There is no CPU fallback: every solve runs the HIP library, and importing
the solver functions fails loudly when the library has not been built.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

__all__ = [
    "NO_MAKESPAN", "MODE_ICH_EXACT", "MODE_ICH_DET", "MODE_PMS_GPU", "MODE_PMS_EXACT",
    "JSSPInstance", "Result", "Engine", "JsspError", "load_library", "library_path",
    "generate_instance", "parse_instance", "resolve_instance_file", "validate_sequence",
    "lower_bound", "beam_search_ich", "beam_search_pms", "run_gpu_beam_search_v1",
    "run_gpu_beam_search_v2", "default_engine", "Server", "rand_stream", "host_solve_pms",
]

NO_MAKESPAN = 500000
MODE_ICH_EXACT, MODE_ICH_DET, MODE_PMS_GPU, MODE_PMS_EXACT = 0, 1, 2, 3
FLAG_GLOBAL_RAND = 0x1
FLAG_RADIX_CUT = 0x2
FLAG_GENERIC_DIGIT = 0x4
FLAG_UNFUSED = 0x8
LANES_FUSED = 1

_HERE = os.path.dirname(os.path.abspath(__file__))


def library_path() -> str:
    return os.path.join(_HERE, "lib", "libjssp_hip.so")


class JsspError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"jssp error {code}: {msg}")
        self.code = code


_I32P = ctypes.POINTER(ctypes.c_int32)


class _Instance(ctypes.Structure):
    _fields_ = [("n_jobs", ctypes.c_int32), ("n_machines", ctypes.c_int32),
                ("machine", _I32P), ("duration", _I32P)]


class _Params(ctypes.Structure):
    _fields_ = [("beam_width", ctypes.c_int32), ("mode", ctypes.c_int32), ("phases", ctypes.c_int32),
                ("best_known", ctypes.c_int32), ("seed", ctypes.c_uint32), ("n_seeds", ctypes.c_int32),
                ("n_passes", ctypes.c_int32), ("flags", ctypes.c_uint32)]


class _Result(ctypes.Structure):
    _fields_ = [("makespan", ctypes.c_int32), ("found", ctypes.c_int32), ("sequence", _I32P),
                ("expanded_states", ctypes.c_int64), ("levels", ctypes.c_int32),
                ("best_seed_index", ctypes.c_int32), ("device_ms", ctypes.c_double)]


class _DriverParams(ctypes.Structure):
    _fields_ = [("beam_width", ctypes.c_int32), ("mode", ctypes.c_int32), ("phases", ctypes.c_int32),
                ("ranks", ctypes.c_int32), ("iterations", ctypes.c_int32), ("time_seed", ctypes.c_uint32)]


class _HybridParams(ctypes.Structure):
    _fields_ = [("gpu_beam_width", ctypes.c_int32), ("accel_factor", ctypes.c_int32), ("phases", ctypes.c_int32),
                ("threads", ctypes.c_int32), ("ranks", ctypes.c_int32), ("iterations", ctypes.c_int32),
                ("time_seed", ctypes.c_uint32)]


class _ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_int64), ("max_slots", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]


_lib = None

# Every exported symbol of include/jssp.h with its ctypes signature.
_SIGNATURES = {
    "jssp_engine_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    "jssp_engine_destroy": (None, [ctypes.c_void_p]),
    "jssp_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
    "jssp_solve": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(_Instance), ctypes.POINTER(_Params),
                                  ctypes.POINTER(_Result)]),
    "jssp_solve_trace": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(_Instance), ctypes.POINTER(_Params),
                                        ctypes.POINTER(_Result), _I32P, _I32P]),
    "jssp_solve_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(_Instance), ctypes.POINTER(_Params),
                                          ctypes.POINTER(_Result)]),
    "jssp_solve_batch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(_Instance),
                                        ctypes.POINTER(_Params), ctypes.POINTER(_Result)]),
    "jssp_solve_batch_ex": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(_Instance),
                                           ctypes.POINTER(_Params), ctypes.POINTER(_Result), ctypes.c_int32]),
    "jssp_solve_seeded": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(_Instance), ctypes.POINTER(_Params),
                                         ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(_Result)]),
    "jssp_run_passes": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(_Instance), ctypes.POINTER(_DriverParams),
                                       ctypes.POINTER(_Result), _I32P]),
    "jssp_rand_stream": (None, [ctypes.c_uint32, ctypes.c_int32, _I32P]),
    "jssp_host_solve_pms": (ctypes.c_int, [ctypes.POINTER(_Instance), ctypes.c_int32, ctypes.c_int32,
                                           ctypes.c_uint32, ctypes.POINTER(_Result)]),
    "jssp_hybrid_iteration": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(_Instance), ctypes.c_int32,
                                             ctypes.c_int32, ctypes.c_int32, ctypes.c_uint32, ctypes.c_int32,
                                             ctypes.c_int32, ctypes.c_uint32, ctypes.POINTER(_Result)]),
    "jssp_run_hybrid": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(_Instance), ctypes.POINTER(_HybridParams),
                                       ctypes.POINTER(_Result), _I32P]),
    "jssp_engine_set_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_char_p]),
    "jssp_engine_tuned_lanes": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
    "jssp_engine_tuned_shape": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                               _I32P, _I32P, _I32P, _I32P]),
    "jssp_engine_set_profiling": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "jssp_engine_kernel_time": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_double),
                                               ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_double)]),
    "jssp_engine_level_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_uint32),
                                               ctypes.c_int32]),
    "jssp_engine_kernel_launches": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_float),
                                                   ctypes.POINTER(ctypes.c_double), ctypes.c_int32]),
    "jssp_server_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(_ServerConfig),
                                          ctypes.POINTER(ctypes.c_void_p)]),
    "jssp_server_destroy": (None, [ctypes.c_void_p]),
    "jssp_serve_submit": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(_Instance), ctypes.POINTER(_Params),
                                         ctypes.POINTER(ctypes.c_uint64)]),
    "jssp_serve_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]),
    "jssp_serve_poll": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(_Result)]),
    "jssp_serve_cancel": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64]),
    "jssp_serve_drain": (ctypes.c_int, [ctypes.c_void_p]),
    "jssp_serve_times": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(ctypes.c_double),
                                        ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
    "jssp_resolve_instance_file": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int32]),
    "jssp_parse_instance_file": (ctypes.c_int, [ctypes.c_char_p, _I32P, _I32P, _I32P, _I32P, ctypes.c_int32]),
    "jssp_generate_instance": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                              ctypes.c_int32, _I32P, _I32P]),
    "jssp_validate_sequence": (ctypes.c_int, [ctypes.POINTER(_Instance), _I32P, _I32P]),
    "jssp_lower_bound": (ctypes.c_int32, [ctypes.POINTER(_Instance)]),
    "jssp_version": (ctypes.c_char_p, []),
}


def load_library():
    """Loads the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    path = library_path()
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: run `python -m beamsearch_jssp_gpu_amd.build` first "
                           "(there is no CPU fallback)")
    lib = ctypes.CDLL(path)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _p32(a: np.ndarray):
    return a.ctypes.data_as(_I32P)


@dataclass
class JSSPInstance:
    """Job-major instance: machine[j, k] (0-based) and duration[j, k] of op k of job j.

    Equivalent of the reference's JSSPInstance (cpu/common/jssp_types.h:16-23),
    whose op ids are 1-based and job-major: op = j*m + k + 1.
    """
    machine: np.ndarray
    duration: np.ndarray
    name: str = ""

    def __post_init__(self):
        self.machine = np.ascontiguousarray(self.machine, dtype=np.int32)
        self.duration = np.ascontiguousarray(self.duration, dtype=np.int32)
        if self.machine.ndim == 1:
            raise ValueError("machine must be 2-D (n_jobs, n_machines)")
        if self.machine.shape != self.duration.shape:
            raise ValueError("machine and duration shapes differ")

    @property
    def nbjob(self) -> int:
        return int(self.machine.shape[0])

    @property
    def nbmachine(self) -> int:
        return int(self.machine.shape[1])

    @property
    def nboperation(self) -> int:  # cpu/common/instance_parser.cpp:51
        return self.nbjob * self.nbmachine + 2

    def _c(self) -> _Instance:
        return _Instance(self.nbjob, self.nbmachine, _p32(self.machine), _p32(self.duration))


@dataclass
class Result:
    makespan: int
    found: bool
    sequence: Optional[np.ndarray]
    expanded_states: int
    levels: int
    device_ms: float
    best_seed_index: int = 0
    trail: Optional[np.ndarray] = None
    pcount: Optional[np.ndarray] = None


def _params(beam_width, mode, phases, best_known, seed, n_seeds=1, n_passes=1, flags=0) -> _Params:
    return _Params(int(beam_width), int(mode), int(phases), int(best_known), int(seed) & 0xFFFFFFFF,
                   int(n_seeds), int(n_passes), int(flags))


class Engine:
    """One HIP stream + device arena (jssp_engine)."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        h = ctypes.c_void_p()
        rc = self.lib.jssp_engine_create(device, ctypes.byref(h))
        if rc != 0:
            raise JsspError(rc, f"cannot create an engine on device {device}")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.jssp_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int):
        if rc != 0:
            raise JsspError(rc, self.lib.jssp_last_error(self.h).decode())

    def solve(self, inst: JSSPInstance, beam_width: int, mode: int = MODE_PMS_GPU, phases: int = 2,
              best_known: int = NO_MAKESPAN, seed: int = 0, n_seeds: int = 1, n_passes: int = 1,
              flags: int = 0, trace: bool = False) -> Result:
        L = inst.nbjob * inst.nbmachine
        seq = np.zeros(L, dtype=np.int32)
        r = _Result()
        r.sequence = _p32(seq)
        ci = inst._c()
        pr = _params(beam_width, mode, phases, best_known, seed, n_seeds, n_passes, flags)
        trail = pcount = None
        if trace:
            trail = np.full(max(0, L - 1) * beam_width, -1, dtype=np.int32)
            pcount = np.zeros(L + 1, dtype=np.int32)
            self._check(self.lib.jssp_solve_trace(self.h, ctypes.byref(ci), ctypes.byref(pr), ctypes.byref(r),
                                                  _p32(trail), _p32(pcount)))
        else:
            self._check(self.lib.jssp_solve(self.h, ctypes.byref(ci), ctypes.byref(pr), ctypes.byref(r)))
        return Result(int(r.makespan), bool(r.found), seq if r.found else None, int(r.expanded_states),
                      int(r.levels), float(r.device_ms), int(r.best_seed_index), trail, pcount)

    def solve_noalloc(self, inst: JSSPInstance, beam_width: int, mode: int = MODE_PMS_GPU, phases: int = 2,
                      best_known: int = NO_MAKESPAN, seed: int = 0, flags: int = 0) -> Result:
        """jssp_solve_noalloc: the same solve with no allocation (workspace in
        the library's static per-device scratch array)."""
        L = inst.nbjob * inst.nbmachine
        seq = np.zeros(L, dtype=np.int32)
        r = _Result()
        r.sequence = _p32(seq)
        ci = inst._c()
        pr = _params(beam_width, mode, phases, best_known, seed, 1, 1, flags)
        self._check(self.lib.jssp_solve_noalloc(self.h, ctypes.byref(ci), ctypes.byref(pr), ctypes.byref(r)))
        return Result(int(r.makespan), bool(r.found), seq if r.found else None, int(r.expanded_states),
                      int(r.levels), float(r.device_ms), int(r.best_seed_index))

    def solve_batch(self, insts: Sequence[JSSPInstance], params: Sequence[dict],
                    use_graph: bool = True) -> list[Result]:
        """Batched solve; params[i] holds keyword arguments of _params
        (beam_width, mode, phases, best_known, seed)."""
        n = len(insts)
        cis = (_Instance * n)(*[i._c() for i in insts])
        prs = (_Params * n)(*[_params(**p) for p in params])
        rs = (_Result * n)()
        seqs = []
        for i, inst in enumerate(insts):
            s = np.zeros(inst.nbjob * inst.nbmachine, dtype=np.int32)
            seqs.append(s)
            rs[i].sequence = _p32(s)
        self._check(self.lib.jssp_solve_batch_ex(self.h, n, cis, prs, rs, 1 if use_graph else 0))
        return [Result(int(r.makespan), bool(r.found), seqs[i] if r.found else None, int(r.expanded_states),
                       int(r.levels), float(r.device_ms), int(r.best_seed_index)) for i, r in enumerate(rs)]

    def solve_seeded(self, inst: JSSPInstance, beam_width: int, seeds, mode: int = MODE_PMS_GPU,
                     phases: int = 2, best_known: int = NO_MAKESPAN, flags: int = 0) -> Result:
        """Multi-pass solve with explicit seeds: seeds[e][i] is the random_seed
        of beam i in pass e (array of shape (passes, beams)); pass e prunes
        with the best makespan of the passes before it."""
        sd = np.ascontiguousarray(np.asarray(seeds, dtype=np.int64) & 0xFFFFFFFF, dtype=np.uint32)
        if sd.ndim != 2 or sd.size == 0:
            raise ValueError("seeds must be a non-empty 2-D array (passes, beams)")
        E, S = sd.shape
        seq = np.zeros(inst.nbjob * inst.nbmachine, dtype=np.int32)
        r = _Result()
        r.sequence = _p32(seq)
        ci = inst._c()
        pr = _params(beam_width, mode, phases, best_known, 0, S, E, flags)
        self._check(self.lib.jssp_solve_seeded(self.h, ctypes.byref(ci), ctypes.byref(pr),
                                               sd.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), ctypes.byref(r)))
        return Result(int(r.makespan), bool(r.found), seq if r.found else None, int(r.expanded_states),
                      int(r.levels), float(r.device_ms), int(r.best_seed_index))

    def run_passes(self, inst: JSSPInstance, beam_width: int, ranks: int, iterations: int, time_seed: int,
                   mode: int = MODE_PMS_GPU, phases: int = 2):
        """The reference's multi-rank pass loop (gpu/v1/main.cu:55-71,
        cpu/v1_pms/main.cpp:55-73) with `ranks` ranks, rank r seeded
        srand(time_seed + r).  Returns (Result, per-pass global best)."""
        seq = np.zeros(inst.nbjob * inst.nbmachine, dtype=np.int32)
        pass_best = np.zeros(max(1, int(iterations)), dtype=np.int32)
        r = _Result()
        r.sequence = _p32(seq)
        ci = inst._c()
        d = _DriverParams(int(beam_width), int(mode), int(phases), int(ranks), int(iterations),
                          int(time_seed) & 0xFFFFFFFF)
        self._check(self.lib.jssp_run_passes(self.h, ctypes.byref(ci), ctypes.byref(d), ctypes.byref(r),
                                             _p32(pass_best)))
        res = Result(int(r.makespan), bool(r.found), seq if r.found else None, int(r.expanded_states),
                     int(r.levels), float(r.device_ms), int(r.best_seed_index))
        return res, pass_best

    def hybrid_iteration(self, inst: JSSPInstance, gpu_beam_width: int, cpu_beam_width: int, best_known: int,
                         gpu_seed: int, cpu_seed_base: int, threads: int, phases: int = 1) -> Result:
        """One hybrid CPU+GPU iteration (hybrid/v1/solver.cu:11-44): the device
        pass with gpu_seed and threads-1 host BS-PMS passes seeded
        cpu_seed_base + tid."""
        seq = np.zeros(inst.nbjob * inst.nbmachine, dtype=np.int32)
        r = _Result()
        r.sequence = _p32(seq)
        ci = inst._c()
        self._check(self.lib.jssp_hybrid_iteration(self.h, ctypes.byref(ci), int(gpu_beam_width),
                                                   int(cpu_beam_width), int(best_known), int(gpu_seed) & 0xFFFFFFFF,
                                                   int(phases), int(threads), int(cpu_seed_base) & 0xFFFFFFFF,
                                                   ctypes.byref(r)))
        return Result(int(r.makespan), bool(r.found), seq if r.found else None, int(r.expanded_states),
                      int(r.levels), float(r.device_ms), int(r.best_seed_index))

    def run_hybrid(self, inst: JSSPInstance, gpu_beam_width: int, accel_factor: int, threads: int, ranks: int,
                   iterations: int, time_seed: int, phases: int = 1):
        """The hybrid drivers (hybrid/v1/main.cu:45-85); returns (Result,
        per-iteration global best)."""
        seq = np.zeros(inst.nbjob * inst.nbmachine, dtype=np.int32)
        pass_best = np.zeros(max(1, int(iterations)), dtype=np.int32)
        r = _Result()
        r.sequence = _p32(seq)
        ci = inst._c()
        hp = _HybridParams(int(gpu_beam_width), int(accel_factor), int(phases), int(threads), int(ranks),
                           int(iterations), int(time_seed) & 0xFFFFFFFF)
        self._check(self.lib.jssp_run_hybrid(self.h, ctypes.byref(ci), ctypes.byref(hp), ctypes.byref(r),
                                             _p32(pass_best)))
        res = Result(int(r.makespan), bool(r.found), seq if r.found else None, int(r.expanded_states),
                     int(r.levels), float(r.device_ms), int(r.best_seed_index))
        return res, pass_best

    def set_autotune(self, on: bool, table_path: Optional[str] = None):
        self._check(self.lib.jssp_engine_set_autotune(self.h, 1 if on else 0,
                                                      table_path.encode() if table_path else None))

    def tuned_lanes(self, n_jobs: int, n_machines: int, beam_width: int) -> int:
        return int(self.lib.jssp_engine_tuned_lanes(self.h, n_jobs, n_machines, beam_width))

    def tuned_shape(self, n_jobs: int, n_machines: int, beam_width: int) -> Optional[dict]:
        """The class's tuned batch shape (lanes, rank_waves, rank_batch,
        filter_blocks), or None if not tuned yet."""
        v = [ctypes.c_int32() for _ in range(4)]
        if not self.lib.jssp_engine_tuned_shape(self.h, n_jobs, n_machines, beam_width, *[ctypes.byref(x) for x in v]):
            return None
        return dict(zip(("lanes", "rank_waves", "rank_batch", "filter_blocks"), (x.value for x in v)))

    def set_profiling(self, on: bool):
        self._check(self.lib.jssp_engine_set_profiling(self.h, 1 if on else 0))

    def level_stats(self, level: int) -> dict:
        out = (ctypes.c_uint32 * 10)()
        self._check(self.lib.jssp_engine_level_stats(self.h, level, out, 10))
        keys = ["kept", "survivors", "lo", "shift", "T1", "below", "tie", "need", "sure_ctr", "tie_ctr"]
        return dict(zip(keys, list(out)))

    def kernel_time(self, kind: int):
        ms, n, b = ctypes.c_double(), ctypes.c_int64(), ctypes.c_double()
        self._check(self.lib.jssp_engine_kernel_time(self.h, kind, ctypes.byref(ms), ctypes.byref(n),
                                                     ctypes.byref(b)))
        return ms.value, n.value, b.value

    def kernel_launches(self, kind: int):
        """Per launch of `kind` in the last profiled solve, in level order:
        (durations in ms, algorithmic bytes) as numpy arrays."""
        n = self.lib.jssp_engine_kernel_launches(self.h, kind, None, None, 0)
        if n < 0:
            raise ValueError(f"kernel kind {kind} is not 0, 1 or 2")
        ms = np.zeros(n, dtype=np.float32)
        b = np.zeros(n, dtype=np.float64)
        if n:
            self.lib.jssp_engine_kernel_launches(self.h, kind, ms.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                                 b.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n)
        return ms, b


_default_engine: Optional[Engine] = None


def default_engine() -> Engine:
    global _default_engine
    if _default_engine is None:
        _default_engine = Engine(0)
    return _default_engine


# ---- host-side helpers (no device needed) --------------------------------
def generate_instance(seed: int, n_jobs: int, n_machines: int, dmin: int = 1, dmax: int = 99) -> JSSPInstance:
    lib = load_library()
    m = np.zeros((n_jobs, n_machines), dtype=np.int32)
    d = np.zeros((n_jobs, n_machines), dtype=np.int32)
    rc = lib.jssp_generate_instance(int(seed) & 0xFFFFFFFFFFFFFFFF, n_jobs, n_machines, dmin, dmax, _p32(m), _p32(d))
    if rc != 0:
        raise JsspError(rc, "generate_instance: bad arguments")
    return JSSPInstance(m, d, name=f"gen{n_jobs}x{n_machines}s{seed}")


def resolve_instance_file(arg: str) -> Optional[str]:
    lib = load_library()
    buf = ctypes.create_string_buffer(4096)
    return buf.value.decode() if lib.jssp_resolve_instance_file(arg.encode(), buf, 4096) else None


def parse_instance(path: str) -> JSSPInstance:
    lib = load_library()
    n, m = ctypes.c_int32(), ctypes.c_int32()
    rc = lib.jssp_parse_instance_file(path.encode(), ctypes.byref(n), ctypes.byref(m), None, None, 0)
    if rc != 0:
        raise JsspError(rc, f"cannot read {path}")
    mach = np.zeros((n.value, m.value), dtype=np.int32)
    dur = np.zeros((n.value, m.value), dtype=np.int32)
    rc = lib.jssp_parse_instance_file(path.encode(), ctypes.byref(n), ctypes.byref(m), _p32(mach), _p32(dur),
                                      n.value * m.value)
    if rc != 0:
        raise JsspError(rc, f"malformed instance file {path}")
    return JSSPInstance(mach, dur, name=os.path.basename(path))


def rand_stream(seed: int, count: int) -> np.ndarray:
    """The first `count` values of glibc rand() after srand(seed), computed
    without touching the process-wide generator."""
    out = np.zeros(max(0, int(count)), dtype=np.int32)
    if count > 0:
        load_library().jssp_rand_stream(int(seed) & 0xFFFFFFFF, int(count), _p32(out))
    return out


def host_solve_pms(inst: JSSPInstance, beam_width: int, best_known: int = NO_MAKESPAN, seed: int = 0) -> Result:
    """BS-PMS on the host (cpu/v1_pms/solver.cpp:8-70): srand(seed) then one
    pass, from a private rand() stream; the CPU side of the hybrid drivers.
    Needs no device."""
    lib = load_library()
    seq = np.zeros(inst.nbjob * inst.nbmachine, dtype=np.int32)
    r = _Result()
    r.sequence = _p32(seq)
    ci = inst._c()
    rc = lib.jssp_host_solve_pms(ctypes.byref(ci), int(beam_width), int(best_known), int(seed) & 0xFFFFFFFF,
                                 ctypes.byref(r))
    if rc != 0:
        raise JsspError(rc, "host_solve_pms: bad arguments")
    return Result(int(r.makespan), bool(r.found), seq if r.found else None, int(r.expanded_states),
                  int(r.levels), 0.0, 0)


def validate_sequence(inst: JSSPInstance, seq: np.ndarray) -> int:
    """Makespan of a 1-based op sequence, or -1 if infeasible."""
    lib = load_library()
    s = np.ascontiguousarray(seq, dtype=np.int32)
    out = ctypes.c_int32()
    ci = inst._c()
    rc = lib.jssp_validate_sequence(ctypes.byref(ci), _p32(s), ctypes.byref(out))
    if rc != 0:
        raise JsspError(rc, "validate_sequence")
    return int(out.value)


def lower_bound(inst: JSSPInstance) -> int:
    ci = inst._c()
    return int(load_library().jssp_lower_bound(ctypes.byref(ci)))


# ---- the reference's solver entry points ------------------------------------
def beam_search_ich(inst: JSSPInstance, beam_width: int, seed: Optional[int] = None) -> int:
    """cpu/v0_ich/solver.cpp:8-54.  With seed=None the process-global rand()
    stream continues (as the reference); otherwise srand(seed) first."""
    flags = FLAG_GLOBAL_RAND if seed is None else 0
    return default_engine().solve(inst, beam_width, MODE_ICH_EXACT, 2, NO_MAKESPAN, seed or 0,
                                  flags=flags).makespan


def beam_search_pms(inst: JSSPInstance, beam_width: int, best_known: int, seed: Optional[int] = None) -> int:
    """cpu/v1_pms/solver.cpp:8-70 (one pass)."""
    flags = FLAG_GLOBAL_RAND if seed is None else 0
    return default_engine().solve(inst, beam_width, MODE_PMS_EXACT, 2, best_known, seed or 0,
                                  flags=flags).makespan


def _gpu(inst, beam_width, best_known, random_seed, out_best_sequence, phases):
    r = default_engine().solve(inst, beam_width, MODE_PMS_GPU, phases, best_known, random_seed)
    if out_best_sequence is not None and r.found:
        out_best_sequence[: len(r.sequence)] = r.sequence
    return r.makespan


def run_gpu_beam_search_v1(inst: JSSPInstance, beam_width: int, best_known: int, random_seed: int,
                           out_best_sequence: Optional[np.ndarray] = None) -> int:
    """gpu/v1/solver.cu:178-298 (single-phase)."""
    return _gpu(inst, beam_width, best_known, random_seed, out_best_sequence, 1)


def run_gpu_beam_search_v2(inst: JSSPInstance, beam_width: int, best_known: int, random_seed: int,
                           out_best_sequence: Optional[np.ndarray] = None) -> int:
    """gpu/v2_twophases/solver.cu:256-386 (two-phase)."""
    return _gpu(inst, beam_width, best_known, random_seed, out_best_sequence, 2)


class Server:
    """Continuous-batching solve server (jssp_server): requests are admitted
    as slots and HBM budget free up and advance together, one level per step."""

    def __init__(self, engine: Engine, hbm_budget_bytes: int, max_slots: int = 64):
        self.engine = engine
        self.lib = engine.lib
        cfg = _ServerConfig(int(hbm_budget_bytes), int(max_slots), 0)
        h = ctypes.c_void_p()
        engine._check(self.lib.jssp_server_create(engine.h, ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h
        self._keep = {}

    def close(self):
        if getattr(self, "h", None):
            self.lib.jssp_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, inst: JSSPInstance, beam_width: int, mode: int = MODE_PMS_GPU, best_known: int = NO_MAKESPAN,
               seed: int = 0, flags: int = 0) -> int:
        ci = inst._c()
        pr = _params(beam_width, mode, 2, best_known, seed, flags=flags)
        t = ctypes.c_uint64()
        self.engine._check(self.lib.jssp_serve_submit(self.h, ctypes.byref(ci), ctypes.byref(pr), ctypes.byref(t)))
        self._keep[t.value] = inst
        return int(t.value)

    def step(self, levels: int = 16) -> int:
        a = ctypes.c_int32()
        self.engine._check(self.lib.jssp_serve_step(self.h, int(levels), ctypes.byref(a)))
        return int(a.value)

    def poll(self, ticket: int) -> Optional[Result]:
        inst = self._keep[ticket]
        seq = np.zeros(inst.nbjob * inst.nbmachine, dtype=np.int32)
        r = _Result()
        r.sequence = _p32(seq)
        rc = self.lib.jssp_serve_poll(self.h, ctypes.c_uint64(ticket), ctypes.byref(r))
        if rc == 7:
            return None
        self.engine._check(rc)
        return Result(int(r.makespan), bool(r.found), seq if r.found else None, int(r.expanded_states),
                      int(r.levels), float(r.device_ms))

    def cancel(self, ticket: int):
        self.engine._check(self.lib.jssp_serve_cancel(self.h, ctypes.c_uint64(ticket)))

    def drain(self):
        self.engine._check(self.lib.jssp_serve_drain(self.h))

    def times(self, ticket: int):
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self.engine._check(self.lib.jssp_serve_times(self.h, ctypes.c_uint64(ticket), ctypes.byref(a),
                                                     ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value
