"""The library's own kernel sources and host code, built for the CPU emulator
under AddressSanitizer (tests/emu), run small single-simulation, Jacobi-plan,
batch (with graph replay), serving-loop and read-back cases and must give the
FP32 oracle's bits.  This checks every kernel index against its allocation on
the CPU and covers the serving loop's logic without a GPU.  The cases that use
several streams run again with stream work deferred until a synchronisation
needs it (newest stream first) and blocks in reverse order: a missing wait
between streams, a host buffer reused before its copy ran, or a result that
depends on block order then fails here."""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EMU = os.path.join(HERE, "emu")
CXX = "/opt/rocm/llvm/bin/clang++"


def _asan_runtime():
    try:
        p = subprocess.run([CXX, "-print-file-name=libclang_rt.asan-x86_64.so"], capture_output=True,
                           text=True, timeout=60)
    except OSError:
        return None
    path = p.stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


@pytest.fixture(scope="module")
def emu_lib(tmp_path_factory):
    asan = _asan_runtime() if os.path.exists(CXX) else None
    if asan is None:
        pytest.skip("host clang++ with an AddressSanitizer runtime is not available")
    out = str(tmp_path_factory.mktemp("emu_build"))
    r = subprocess.run(["make", "-s", "-j8", "-C", EMU, f"OUT={out}"], capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    return os.path.join(out, "libsmokesim_emu.so"), asan


@pytest.mark.parametrize("case", ["single", "norms", "plans", "batch", "server", "readback", "vorticity",
                                  "advection", "refine_groups", "proxy", "step_graph"])
def test_emulated_case_matches_oracle(emu_lib, case):
    lib, asan = emu_lib
    env = dict(os.environ, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0")
    r = subprocess.run([sys.executable, os.path.join(EMU, "emu_cases.py"), lib, case], env=env,
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0 and f"ok {case}" in r.stdout, (r.stdout[-4000:] + r.stderr[-8000:])


@pytest.mark.parametrize("case", ["single", "batch", "server", "readback", "norms"])
def test_emulated_case_with_deferred_streams(emu_lib, case):
    lib, asan = emu_lib
    env = dict(os.environ, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0",
               EMU_DEFERRED="1", EMU_REVERSE_BLOCKS="1")
    r = subprocess.run([sys.executable, os.path.join(EMU, "emu_cases.py"), lib, case], env=env,
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0 and f"ok {case}" in r.stdout, (r.stdout[-4000:] + r.stderr[-8000:])
