"""Host thread control of the front end (include/ipf/parallel.hpp), the
analogue of the reference's tests/test_parallel.cpp:23-50: coverage for any
thread count, nested regions serial, first exception rethrown, concurrent
callers. Built and run plain, under AddressSanitizer + UndefinedBehaviorSanitizer
and under ThreadSanitizer (CPU only)."""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
SRC = ROOT / "tests" / "cpp" / "parallel_check.cpp"

VARIANTS = {
    "plain": ["-O2"],
    "asan_ubsan": ["-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
                   "-fno-omit-frame-pointer"],
    "tsan": ["-O1", "-g", "-fsanitize=thread"],
}


@pytest.mark.parametrize("variant", sorted(VARIANTS))
def test_parallel_header(variant, tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = tmp_path / f"parallel_{variant}"
    cmd = ["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-pthread", *VARIANTS[variant],
           "-I", str(ROOT / "include"), str(SRC), "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "PARALLEL OK" in r.stdout, r.stdout + r.stderr
