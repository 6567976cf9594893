"""The C ABI library loads and exports every symbol include/jssp.h declares;
the C++ compatibility header (reference signatures) compiles and links
against it; device entry points fail with an error code, not an exit, when
no GPU is present."""
import os
import re
import subprocess

import pytest

import beamsearch_jssp_gpu_amd as bs

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_c_symbols():
    text = open(os.path.join(REPO, "include", "jssp.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(jssp_[a-z_0-9]+)\s*\(", text)))


def test_exports_every_declared_symbol():
    names = _declared_c_symbols()
    assert len(names) >= 20
    out = subprocess.run(["nm", "-D", "--defined-only", bs.library_path()], stdout=subprocess.PIPE, text=True,
                         check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    lib = bs.load_library()
    for n in names:
        assert hasattr(lib, n)
    assert set(names) == set(bs._SIGNATURES), "ctypes signature table out of sync with jssp.h"


def test_compat_symbols_exported():
    out = subprocess.run(["nm", "-DC", "--defined-only", bs.library_path()], stdout=subprocess.PIPE, text=True,
                         check=True).stdout
    for sig in ["beam_search_ich(JSSPInstance const&, int)",
                "beam_search_pms(JSSPInstance const&, int, int)",
                "run_gpu_beam_search_v1(JSSPInstance const&, int, int, int, int*)",
                "run_gpu_beam_search_v2(JSSPInstance const&, int, int, int, int*)",
                "run_hybrid_iteration_v1(JSSPInstance const&, int, int, int, int, int)",
                "run_hybrid_iteration_v2(JSSPInstance const&, int, int, int, int, int)",
                "parse_instance(char const*, JSSPInstance&)",
                "resolve_instance_file(char const*, char*, int)"]:
        assert sig in out, sig


def test_reference_style_driver_compiles_and_parses(tmp_path):
    """A driver written against the reference's API (cpu/v0_ich/main.cpp
    shape) builds with jssp_compat.hpp and runs its host-side part."""
    src = tmp_path / "driver.cpp"
    src.write_text(r'''
#include "jssp_compat.hpp"
#include <cstdio>
int main(int argc, char** argv) {
  char path[512];
  if (!resolve_instance_file(argv[1], path, sizeof(path))) return 3;
  JSSPInstance inst;
  if (!parse_instance(path, inst)) return 4;
  std::printf("%d %d %d %zu %zu %d\n", inst.nbjob, inst.nbmachine, inst.nboperation,
              inst.operations.size(), inst.opr_mach.size(), inst.operations[1].num_mc);
  return 0;
}
''')
    exe = tmp_path / "driver"
    libdir = os.path.dirname(bs.library_path())
    subprocess.run(["g++", "-std=c++11", "-I", os.path.join(REPO, "include"), str(src), "-o", str(exe),
                    "-L", libdir, "-ljssp_hip", f"-Wl,-rpath,{libdir}"], check=True)
    out = subprocess.run([str(exe), os.path.join(REPO, "tests", "golden", "instances", "la01.txt")],
                         stdout=subprocess.PIPE, text=True, check=True).stdout.split()
    # la01: 10 jobs, 5 machines, 52 ops incl. dummies, first op on machine 1 (1-based 2)
    assert out == ["10", "5", "52", "52", "6", "2"]


def test_c_header_compiles_as_c(tmp_path):
    src = tmp_path / "t.c"
    src.write_text('#include "jssp.h"\nint main(void){jssp_params p={0}; (void)p; return (int)jssp_lower_bound(0) + 1;}\n')
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(REPO, "include"), "-c", str(src),
                    "-o", str(tmp_path / "t.o")], check=True)


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="checks the no-device error path")
def test_engine_create_without_device_is_an_error():
    with pytest.raises(bs.JsspError) as e:
        bs.Engine(0)
    assert e.value.code == 3
