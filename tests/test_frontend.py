"""The header-only C++ front end (include/ipf/*.hpp) used the way the
reference's users call ipf: argument checks on the CPU, results on the GPU
against the oracle."""
from __future__ import annotations

import subprocess
from pathlib import Path

import numpy as np
import pytest

from tests.fixtures import phase_tolerance

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


def test_frontend_codecs_on_host(exe, tmp_path):
    # encodePNM / encodePFM and .npy files of the C++ front end: bytes equal the
    # restatement of the reference's encoders (pinned to its own output), numpy
    # reads and writes the arrays
    from oracle import pyoracle as po
    np.save(tmp_path / "in.npy", np.arange(24, dtype=np.int16).reshape(2, 3, 4) - 7)
    r = subprocess.run([str(exe), "codec", str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "CODEC OK" in r.stdout, r.stdout + r.stderr
    rgb = (np.arange(13 * 7 * 3) * 37 + 5).astype(np.uint8).reshape(7, 13, 3)
    g16 = (np.arange(36) * 2741 + 11).astype(np.uint16).reshape(4, 9)
    f = (np.arange(90, dtype=np.float32) * np.float32(0.37) - np.float32(3.0)).reshape(5, 6, 3)
    assert (tmp_path / "rgb_bin.ppm").read_bytes() == po.pnm_encode(rgb, True)
    assert (tmp_path / "rgb_txt.ppm").read_bytes() == po.pnm_encode(rgb, False)
    assert (tmp_path / "g16_bin.pgm").read_bytes() == po.pnm_encode(g16, True)
    assert (tmp_path / "g16_txt.pgm").read_bytes() == po.pnm_encode(g16, False)
    assert (tmp_path / "f.pfm").read_bytes() == po.pfm_encode(f)
    assert np.array_equal(np.load(tmp_path / "rgb.npy"), rgb)
    assert np.array_equal(np.load(tmp_path / "g16.npy"), g16)
    assert np.array_equal(np.load(tmp_path / "f.npy"), f)
    got = np.fromfile(tmp_path / "npy_in.bin", np.float64).reshape(2, 3, 4)
    assert np.array_equal(got, np.arange(24).reshape(2, 3, 4) - 7.0)


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
    # derivative and sharpening filters
    assert np.array_equal(out(11, np.float32, u8.shape), oracle.sobel(u8, 1, 0))
    assert np.array_equal(out(12, np.float32, f.shape), oracle.scharr(f, 0, 1, 2.0, 1))
    assert np.array_equal(out(13, np.float32, u8.shape), oracle.laplacian(u8, 3))
    gx, gy = oracle.sobel(f, 1, 0), oracle.sobel(f, 0, 1)
    assert np.array_equal(out(14, np.float32, f.shape), oracle.magnitude(gx, gy))
    want = oracle.phase(gx, gy)
    assert np.all(np.abs(out(15, np.float32, f.shape) - want.astype(np.float64)) <= phase_tolerance(gx, gy, want, True))
    assert np.array_equal(out(16, np.uint8, u8.shape), oracle.unsharp_mask(u8, 1.5, 0.8, 2.0))
    assert np.array_equal(out(17, np.uint8, u8.shape), oracle.sharpen(u8, 0.5))
    assert np.array_equal(out(18, np.uint8, u8.shape), oracle.emboss(u8))
    gray = np.ascontiguousarray(u8[:, :, 1])
    assert np.array_equal(out(19, np.uint8, gray.shape), oracle.adaptive_threshold(gray, 255, 1, 1, 7, 3.0))
    assert np.array_equal(out(20, np.float32, u8.shape), oracle.prewitt(u8, 1, 0, 4))
    assert np.array_equal(out(21, np.float32, u8.shape),
                          oracle.magnitude(oracle.sobel(u8, 1, 0), oracle.sobel(u8, 0, 1)))
    # pyramids: 97 x 61 -> 49 x 31 -> 25 x 16 -> 13 x 8 -> 7 x 4 -> 4 x 2 -> 2 x 1 (stops: height 1)
    assert np.array_equal(out(22, np.uint8, (31, 49, 3)), oracle.pyr_down(u8))
    assert np.array_equal(out(23, np.float32, (118, 166)), oracle.pyr_up(f))
    assert "LEVELS 7" in r.stdout
    # feature callers
    assert np.array_equal(out(24, np.uint8, u8.shape[:2]), oracle.canny(u8, 40.0, 120.0))
    assert np.array_equal(out(25, np.uint8, f.shape), oracle.canny(f, 0.5, 1.5, 5, True))
    assert np.array_equal(out(26, np.uint8, u8.shape[:2]), oracle.canny_auto(u8))
    assert np.array_equal(out(27, np.float32, gray.shape), oracle.corner_harris(gray, 3, 3, 0.04))
    assert np.array_equal(out(28, np.float32, f.shape), oracle.corner_min_eigen(f, 5, 5))
    assert np.array_equal(out(29, np.uint8, u8.shape), oracle.filter2d_fft(u8, k, 1, 1.0))
    k2 = np.outer(oracle.gaussian_kernel_1d(2.0), oracle.gaussian_kernel_1d(2.0)).astype(np.float32)
    assert np.array_equal(out(30, np.float64, f.shape), oracle.filter2d_fft(f, k2, 0, 0.5, 0.25, np.float64))
    assert np.array_equal(out(31, np.float64, (62, 98, 3)), oracle.integral_image(u8))
    disc = np.fromfile(tmp_path / "in_disc.bin", np.uint8).reshape(61, 97)
    want = oracle.hough_circles(disc, 8, 24)
    assert len(want) >= 1 and abs(want[0, 0] - 48) < 1 and abs(want[0, 1] - 30) < 1 and abs(want[0, 2] - 14) < 1
    assert np.array_equal(np.fromfile(tmp_path / "o33.bin", np.float64).reshape(-1, 4), want)
    assert np.array_equal(out(32, np.float64, (60, 84)), oracle.integral_image(f, True))
    level = u8
    for i in range(7):
        assert np.array_equal(np.fromfile(tmp_path / f"p{i}.bin", np.uint8).reshape(level.shape), level), i
        level = oracle.pyr_down(level)


def test_frontend_motion_blur_kernel_known_answers(exe):
    import json
    from tests import fixtures
    kat = json.loads((fixtures.GOLDEN / "kat.json").read_text())["motion"]
    args = [str(x) for n, a, _ in kat for x in (n, repr(a))]
    r = subprocess.run([str(exe), "motion", *args], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.split("\n")
    for (n, a, want), line in zip(kat, lines):
        vals = line.split()
        side = int(vals[0])
        got = np.array([float(v) for v in vals[1:]], np.float32)
        assert side * side == len(want) and np.array_equal(got, np.array(want, np.float32)), (n, a)
