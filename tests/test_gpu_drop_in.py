"""The literal drop-in: a driver compiled against include/jssp_compat.hpp only
(tests/drop_in/ref_style_driver.cpp, linked to libjssp_hip.so by the build)
calls the reference's solver functions the way the reference's drivers do
(cpu/v0_ich/main.cpp:39-42, cpu/v1_pms/main.cpp:62-63, gpu/v1/main.cu:62-64,
gpu/v2_twophases/main.cu:62-64).  Its results are checked against

* the compiled reference's golden makespans (tests/golden/golden_ref.json)
  and the oracle for beam_search_ich / beam_search_pms,
* the C library's rand() stream: the shims consume the caller's process-wide
  generator like the reference's solvers (cpu/v0_ich/solver.cpp:35,
  cpu/v1_pms/solver.cpp:32,52), so the next rand() after each call must equal
  the next rand() after the oracle's restatement (which draws from the same
  C library generator) -- and after oracle/_ref's reference solver where that
  is built,
* orc_gpu_solve for run_gpu_beam_search_v1/_v2 (makespan and
  out_best_sequence, which stays untouched when nothing beats best_known)."""
import ctypes
import json
import os
import subprocess

import numpy as np
import pytest

from testlib import (GOLDEN, golden_instances, load_json, orc_gpu, orc_ich, orc_pms, p32, ref_lib)

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(REPO, "tests", "drop_in", "build", "ref_style_driver")
LA01 = os.path.join(GOLDEN, "instances", "la01.txt")
PMS_GPU, NO = 2, 500000


def _libc_next_rand():
    libc = ctypes.CDLL(None)
    libc.rand.restype = ctypes.c_int
    return libc.rand()


def _run(*args):
    out = subprocess.run([DRIVER, *[str(a) for a in args]], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    return json.loads(out.stdout)


def test_driver_links_and_checks_arguments():
    """Built, linked (the library resolves through the relative rpath) and
    prints its usage on bad arguments without touching the device."""
    assert os.path.exists(DRIVER), "built by beamsearch_jssp_gpu_amd.build"
    out = subprocess.run([DRIVER], capture_output=True, text=True, timeout=60)
    assert out.returncode == 2 and "usage" in out.stderr


def test_process_wide_rand_stream_host_check():
    """The engine's process-wide rand() stream draws the caller's next values
    and leaves the generator where those rand() calls would, even when other
    code reseeds or draws from it meanwhile (tests/drop_in/rand_stream_check.cpp)."""
    exe = os.path.join(REPO, "tests", "drop_in", "build", "rand_stream_check")
    out = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0 and out.stdout.strip() == "ok", out.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("K,srand_seed,gpu_seed", [(64, 12345, 777), (8, 12345, 3), (256, 12345, 2**31 - 5)])
def test_drop_in_driver_matches_reference_and_oracle(K, srand_seed, gpu_seed):
    print("ref_style_driver la01 K", K, flush=True)
    d = _run(LA01, K, srand_seed, gpu_seed, NO)
    inst = golden_instances()["la01"]
    assert (d["n"], d["m"], d["nboperation"]) == (10, 5, 52)
    gold = {(c["solver"], c["K"], c["seed"]): c["makespan"] for c in load_json("golden_ref.json")["cases"]
            if c["instance"] == "la01" and c.get("best_known", NO) == NO}
    if ("ich", K, srand_seed) in gold:
        assert d["ich"] == gold[("ich", K, srand_seed)]
    # BS-ICH: makespan and the caller's rand() stream after the call
    ms, _, _ = orc_ich(inst, K, srand_seed)   # srand(seed) + the reference's rand() calls, in this process
    assert d["ich"] == ms
    assert d["ich_next_rand"] == _libc_next_rand()
    # BS-PMS with the same srand() seed
    pms, _ = orc_pms(inst, K, NO, srand_seed)
    assert d["pms"] == pms
    assert d["pms_next_rand"] == _libc_next_rand()
    ref = ref_lib()
    if ref is not None:
        assert ref.ref_ich(inst.n, inst.m, p32(inst.machine), p32(inst.duration), K, srand_seed) == d["ich"]
        assert d["ich_next_rand"] == _libc_next_rand()
        assert ref.ref_pms(inst.n, inst.m, p32(inst.machine), p32(inst.duration), K, NO, srand_seed) == d["pms"]
        assert d["pms_next_rand"] == _libc_next_rand()
    # GPU solvers: gpu/v1 and gpu/v2 semantics
    w = orc_gpu(inst, K, PMS_GPU, NO, gpu_seed, trace=False)
    assert d["v1"] == d["v2"] == d["v1_plain"] == w["makespan"]
    np.testing.assert_array_equal(d["v1_seq"], w["sequence"])
    np.testing.assert_array_equal(d["v2_seq"], w["sequence"])


@pytest.mark.gpu
def test_drop_in_driver_bound_semantics():
    """With best_known at or below every reachable makespan the GPU solvers
    return best_known and leave out_best_sequence as it was; BS-PMS prunes
    every child and finds nothing (500000)."""
    inst = golden_instances()["la01"]
    w = orc_gpu(inst, 32, PMS_GPU, NO, 11, trace=False)
    bound = w["makespan"]
    d = _run(LA01, 32, 5, 11, bound)
    assert d["v1"] == d["v2"] == bound
    assert d["v1_seq"] == [-7] * 50 and d["v2_seq"] == [-7] * 50
    assert d["pms"] == orc_pms(inst, 32, bound, 5)[0]
    d = _run(LA01, 32, 5, 11, bound + 1)
    assert d["v1"] == d["v2"] == bound
    np.testing.assert_array_equal(d["v2_seq"], w["sequence"])
