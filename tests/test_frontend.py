"""The header-only C++ front end (include/ipf/*.hpp) used the way the
reference's users call ipf: argument checks on the CPU, results on the GPU
against the oracle."""
from __future__ import annotations

import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
EXE = ROOT / "tests" / "cpp" / "frontend_check"


@pytest.fixture(scope="module")
def exe():
    if not EXE.exists():
        import __graft_entry__ as ge
        ge._build_frontend_check()
    return EXE


def test_frontend_argument_errors(exe):
    r = subprocess.run([str(exe), "errors"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "ERRORS OK" in r.stdout, r.stdout + r.stderr


@pytest.mark.gpu
def test_frontend_results_match_oracle(exe, oracle, tmp_path):
    r = subprocess.run([str(exe), "run", str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "RUN OK" in r.stdout, r.stdout + r.stderr
    u8 = np.fromfile(tmp_path / "in_u8.bin", np.uint8).reshape(61, 97, 3)
    f = np.fromfile(tmp_path / "in_f32.bin", np.float32).reshape(59, 83)

    def out(i, dt, shape):
        return np.fromfile(tmp_path / f"o{i}.bin", dt).reshape(shape)

    U = 2.0 ** -24
    assert np.array_equal(out(1, np.uint8, u8.shape), oracle.gaussian_blur(u8, 2.0))
    want = oracle.gaussian_blur(u8, 1.5, out_dtype=np.float32)
    assert np.all(np.abs(out(2, np.float32, u8.shape) - want.astype(np.float64)) <= 23 * U * 255 * 1.01 + U * want)
    want = oracle.sep_filter(f, [1, 2, 3], [4, 5, 6, 7, 8], None, 0, 0.5, 0.25)
    assert np.all(np.abs(out(3, np.float32, f.shape) - want.astype(np.float64)) <= 13 * U * (6 * 30 + 0.5) * 1.01 + U * np.abs(want))
    assert np.array_equal(out(4, np.uint8, u8.shape), oracle.box_filter(u8, 5, 3, True, 2))
    want = oracle.box_filter(f, 15, 15)
    assert np.all(np.abs(out(5, np.float32, f.shape) - want.astype(np.float64)) <= 1e-12 * 225 + U * want)
    k = np.array([[0, -1, 0], [-1, 5, -1], [0, -1, 0]], np.float32)
    assert np.array_equal(out(6, np.uint8, u8.shape), oracle.filter2d(u8, k, 1, 1.0))
    assert np.array_equal(out(7, np.uint8, u8.shape), oracle.median(u8, 5))
    assert np.array_equal(out(8, np.uint8, u8.shape), oracle.median(u8, 3, 3))
    assert np.array_equal(out(9, np.uint8, u8.shape), oracle.bilateral(u8, 9, 75.0, 75.0))
    assert np.array_equal(out(10, np.float32, f.shape), oracle.gaussian_blur(f, 3.0))
