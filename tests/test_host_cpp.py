"""Host-side C++ checks, built and run on the CPU plain, under
AddressSanitizer + UndefinedBehaviorSanitizer and under ThreadSanitizer:

- tests/cpp/parallel_check.cpp: the front end's thread control
  (include/ipf/parallel.hpp), the analogue of the reference's
  tests/test_parallel.cpp:23-50 -- coverage for any thread count, nested
  regions serial, first exception rethrown, concurrent callers;
- tests/cpp/arena_check.cpp: the serving loop's placement of request buffers
  in its budget-sized block (csrc/arena.hpp) against a brute-force model;
- tests/cpp/abi_check.cpp: the C ABI's host code (csrc/runtime.hip: argument
  checks in the reference's order, the Netpbm header parser, the tuning
  table, the Gaussian weights) compiled with AddressSanitizer and
  UndefinedBehaviorSanitizer on the host pass, leak detection on.
"""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
PROGRAMS = {
    "parallel": (ROOT / "tests" / "cpp" / "parallel_check.cpp", "PARALLEL OK"),
    "arena": (ROOT / "tests" / "cpp" / "arena_check.cpp", "ARENA OK"),
}
VARIANTS = {
    "plain": ["-O2"],
    "asan_ubsan": ["-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
                   "-fno-omit-frame-pointer"],
    "tsan": ["-O1", "-g", "-fsanitize=thread"],
}


@pytest.mark.parametrize("variant", sorted(VARIANTS))
@pytest.mark.parametrize("program", sorted(PROGRAMS))
def test_host_program(program, variant, tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    src, ok = PROGRAMS[program]
    exe = tmp_path / f"{program}_{variant}"
    cmd = ["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-pthread", *VARIANTS[variant],
           "-I", str(ROOT / "include"), str(src), "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and ok in r.stdout, r.stdout + r.stderr


def test_runtime_host_code_under_asan_ubsan():
    import os

    from imageprocessingflex_amd import build as b
    exe = b.build_abi_check_asan()
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0 and "ABI OK" in r.stdout, r.stdout + r.stderr


def test_cmake_target_links_the_front_end(tmp_path):
    # configure this repository with CMake and build the check programs
    # against ipf::ipf; the ABI check runs on the CPU
    if shutil.which("cmake") is None:
        pytest.skip("cmake not available")
    bdir = tmp_path / "b"
    subprocess.run(["cmake", "-S", str(ROOT), "-B", str(bdir), "-DCMAKE_BUILD_TYPE=Release"], check=True,
                   capture_output=True, text=True, timeout=300)
    r = subprocess.run(["cmake", "--build", str(bdir), "--target", "abi_check", "frontend_check", "-j", "4"],
                       capture_output=True, text=True, timeout=1800)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    run = subprocess.run([str(bdir / "abi_check")], capture_output=True, text=True, timeout=120)
    assert run.returncode == 0 and "ABI OK" in run.stdout, run.stdout + run.stderr
    run = subprocess.run([str(bdir / "frontend_check"), "errors"], capture_output=True, text=True, timeout=120)
    assert run.returncode == 0 and "ERRORS OK" in run.stdout, run.stdout + run.stderr
