"""Host-side C++ checks, built and run on the CPU plain, under
AddressSanitizer + UndefinedBehaviorSanitizer and under ThreadSanitizer:

- tests/cpp/parallel_check.cpp: the front end's thread control
  (include/ipf/parallel.hpp), the analogue of the reference's
  tests/test_parallel.cpp:23-50 -- coverage for any thread count, nested
  regions serial, first exception rethrown, concurrent callers;
- tests/cpp/arena_check.cpp: the serving loop's placement of request buffers
  in its budget-sized block (csrc/arena.hpp) against a brute-force model.
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
