"""C-ABI checks that need no GPU: the library loads, exports every entry point
declared in include/smoke.h, the host-only entry points behave, and the
compat header smokeSimulation.cuh declares the reference's nine functions
with their exact signatures (project/smokeSimulation.cuh:4-17)."""
import ctypes
import os
import re
import shutil
import subprocess
import tempfile

import pytest

import smokesim

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "smoke.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(smoke_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for required in ("smoke_create", "smoke_destroy", "smoke_step", "smoke_set_solids",
                     "smoke_add_source", "smoke_add_obstacle", "smoke_move_object",
                     "smoke_read_density", "smoke_legacy_simulate"):
        assert required in names
    assert len(names) >= 40


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(smokesim.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_library_exports_nothing_else():
    nm = shutil.which("nm")
    if nm is None:
        pytest.skip("nm not available")
    out = subprocess.run([nm, "-D", "--defined-only", smokesim.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    assert exported == set(declared_functions())


def test_host_only_entry_points():
    p = smokesim.default_params()
    assert p["iterations"] == 30 and p["projection"] == smokesim.PROJ_REFERENCE
    assert abs(p["gravity"] - (-9.82)) < 1e-6 and p["buoyancy"] == 2.0
    assert p["max_velocity"] == 3.0 and p["sor_omega"] == 1.9
    assert smokesim.footprint_bytes(2, 80, 80) == 0
    big = smokesim.footprint_bytes(256, 256, 256)
    assert big > 11 * 4 * 257 ** 3
    lib = smokesim.load()
    h = ctypes.c_void_p()
    st = lib.smoke_create(2, 5, 5, None, ctypes.byref(h))
    assert st == 1 and not h.value
    assert b"dims" in lib.smoke_last_error()


def test_compat_header_has_the_reference_signatures():
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    src = r"""
#include "smokeSimulation.cuh"
#include <type_traits>
static_assert(std::is_same<decltype(&getGPUProperties), void (*)(void)>::value, "");
static_assert(std::is_same<decltype(&simulate), void (*)(float*, float)>::value, "");
static_assert(std::is_same<decltype(&initializeVolume),
              void (*)(float*, unsigned int, unsigned int, unsigned int)>::value, "");
static_assert(std::is_same<decltype(&deleteVolume), void (*)()>::value, "");
static_assert(std::is_same<decltype(&addObstacle),
              int (*)(float, float, float, float, float, float, float)>::value, "");
static_assert(std::is_same<decltype(&addSmokeSource), int (*)(float, float, float, float)>::value, "");
static_assert(std::is_same<decltype(&updateObjectPos), void (*)(int, float, float, float)>::value, "");
static_assert(std::is_same<decltype(&getBuoyancy), float* (*)()>::value, "");
static_assert(std::is_same<decltype(&getGravity), float* (*)()>::value, "");
int main() { return 0; }
"""
    tmp = tempfile.mkdtemp(prefix="smoke_abi_")
    path = os.path.join(tmp, "compat_sig_check.cpp")
    with open(path, "w") as f:
        f.write(src)
    r = subprocess.run([gxx, "-std=c++11", "-fsyntax-only", "-I", os.path.join(ROOT, "include"), path],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_compiled_caller_links_with_the_host_compiler_alone():
    # The C++ caller of the nine functions (tests/drop_in/app_driver.cpp)
    # builds with g++ and no HIP headers or flags, and resolves every symbol
    # against libsmokesim.so (run on the GPU by test_gpu_drop_in.py).
    gxx = shutil.which("g++")
    lib = os.path.join(ROOT, "smoke-simulation_amd", "smokesim", "libsmokesim.so")
    if gxx is None or not os.path.exists(lib):
        pytest.skip("g++ or the built library not available")
    out = os.path.join(tempfile.mkdtemp(prefix="smoke_dropin_"), "app_driver")
    r = subprocess.run([gxx, "-std=c++11", "-Wall", "-Wextra", "-Werror", "-I",
                        os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "drop_in", "app_driver.cpp"),
                        "-L", os.path.dirname(lib), "-lsmokesim", "-Wl,--no-undefined", "-o", out],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    nm = subprocess.run(["nm", "-u", "-C", out], capture_output=True, text=True).stdout
    assert "hip" not in nm.lower()
    for name in ("smoke_legacy_initialize", "smoke_legacy_simulate", "smoke_legacy_delete",
                 "smoke_legacy_add_source", "smoke_legacy_add_obstacle"):
        assert name in nm
