"""Builds libjssp_hip.so (HIP kernels for gfx950 + host engine + C ABI) in-tree.

    python -m beamsearch_jssp_gpu_amd.build

Every HIP translation unit is compiled with hipcc for gfx950 only, with the
packed-FP32 instruction forms disabled and IEEE-correct FP32 division (the
score must round exactly like the host restatement).
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libjssp_hip.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # The command processor loads the first kernel arguments (the descriptor
    # and block map pointers) into SGPRs before a wave starts, so a kernel's
    # first descriptor load does not wait for a kernarg load: -0.5 ms per
    # 1000-level solve.
    "-mllvm", "-amdgpu-kernarg-preload-count=16",
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", f"-I{ROCM}/include", "-D__HIP_PLATFORM_AMD__"]

HIP_SOURCES = ["level_kernels.hip", "engine.hip", "batch.hip", "serve.hip", "driver.hip", "noalloc.hip"]
CXX_SOURCES = ["host_io.cpp", "compat.cpp", "host_pms.cpp"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout)
        raise RuntimeError(f"build step failed: {os.path.basename(cmd[-1])}")
    # keep real warnings visible, drop hipcc's note about the disabled feature
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if out:
        sys.stderr.write(out + "\n")


# The phase-stamp build (beamsearch_jssp_gpu_amd/phase_stamps.py): the same
# library with the kernels' wall-clock phase stamps compiled in, kept apart.
STAMPS_OBJ = os.path.join(HERE, "_build_stamps")
STAMPS_LIBDIR = os.path.join(HERE, "lib_stamps")
STAMPS_LIB = os.path.join(STAMPS_LIBDIR, "libjssp_hip.so")


def build(verbose: bool = False, stamps: bool = False) -> str:
    """Compiles every source and links the shared library; returns its path.
    stamps=True builds the phase-stamp variant into lib_stamps/ instead."""
    obj_dir, lib_dir, lib = (STAMPS_OBJ, STAMPS_LIBDIR, STAMPS_LIB) if stamps else (OBJ, LIBDIR, LIB)
    extra = ["-DJSSP_PHASE_STAMPS"] if stamps else []
    os.makedirs(obj_dir, exist_ok=True)
    os.makedirs(lib_dir, exist_ok=True)
    jobs = []
    objs = []
    for src in HIP_SOURCES:
        if not os.path.exists(os.path.join(CSRC, src)):
            continue
        o = os.path.join(obj_dir, src + ".o")
        objs.append(o)
        jobs.append([HIPCC, *HIP_FLAGS, *extra, "-x", "hip", "-c", "-o", o, os.path.join(CSRC, src)])
    for src in CXX_SOURCES:
        if not os.path.exists(os.path.join(CSRC, src)):
            continue
        o = os.path.join(obj_dir, src + ".o")
        objs.append(o)
        jobs.append(["g++", *CXX_FLAGS, "-c", "-o", o, os.path.join(CSRC, src)])
    with ThreadPoolExecutor(max_workers=min(4, len(jobs))) as ex:
        list(ex.map(_run, jobs))
    tmp = lib + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs,
          f"-L{ROCM}/lib", "-lamdhip64", "-Wl,-rpath," + f"{ROCM}/lib"])
    os.replace(tmp, lib)
    if not stamps:
        build_drop_in_driver()
    if verbose:
        print(lib)
    return lib


REPO = os.path.dirname(HERE)
DRIVER_SRC = os.path.join(REPO, "tests", "drop_in", "ref_style_driver.cpp")
DRIVER = os.path.join(REPO, "tests", "drop_in", "build", "ref_style_driver")
RAND_CHECK_SRC = os.path.join(REPO, "tests", "drop_in", "rand_stream_check.cpp")
RAND_CHECK = os.path.join(REPO, "tests", "drop_in", "build", "rand_stream_check")


def build_drop_in_driver() -> None:
    """Links the reference-style test driver (tests/drop_in) against the
    library through include/jssp_compat.hpp only, as a reference driver would
    link (the rpath is relative so it runs from any copy of the tree), and
    the host check of the process-wide rand() stream."""
    if not os.path.exists(DRIVER_SRC):
        return
    os.makedirs(os.path.dirname(DRIVER), exist_ok=True)
    _run(["g++", "-O2", "-std=c++11", "-Wall", f"-I{os.path.join(REPO, 'include')}", "-o", DRIVER, DRIVER_SRC,
          f"-L{LIBDIR}", "-ljssp_hip", "-Wl,-rpath,$ORIGIN/../../../beamsearch_jssp_gpu_amd/lib"])
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-o", RAND_CHECK, RAND_CHECK_SRC])


if __name__ == "__main__":
    build(verbose=True, stamps="--stamps" in sys.argv[1:])
