"""ipfg_sep_filter_noalloc (the separable filter that allocates nothing)
against ipfg_sep_filter on the same device inputs, through the C ABI: two
calls in a row per case, the second on a smaller image, bit for bit (both
run the same certified kernels). Also its argument errors."""
from __future__ import annotations

import zlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _both(ipf, img, kx, ky, border, precision=0, out_dtype=None, delta=0.0, bval=0.0, anchor=None):
    d = ipf.DeviceImage.from_array(img)
    want = ipf.sep_filter2d(d, kx, ky, border, delta=delta, border_value=bval, out_dtype=out_dtype,
                            precision=precision, anchor=anchor).download()
    out = ipf.DeviceImage.like(d, out_dtype)
    ipf.sep_filter2d_noalloc(d, out, kx, ky, border, delta=delta, border_value=bval, precision=precision,
                             anchor=anchor)
    ipf.synchronize()
    return out.download(), want


CASES = [
    # (name, big shape, small shape, dtype, sigma or explicit taps, border, out dtype, precision)
    ("gauss2_u8_rgb", (600, 517, 3), (97, 61, 3), np.uint8, 2.0, 3, None, 0),
    ("gauss3_f32", (1000, 900), (64, 48), np.float32, 3.0, 2, None, 0),
    ("gauss8_f32_wrap", (700, 333), (50, 70), np.float32, 8.0, 4, None, 0),
    ("gauss1_u8_4ch_const", (257, 300, 4), (9, 13, 4), np.uint8, 1.0, 0, None, 0),
    ("gauss2_u8_to_f32", (300, 301), (31, 17), np.uint8, 2.0, 1, np.float32, 0),
    ("gauss2_u16", (200, 150, 3), (20, 11, 3), np.uint16, 2.0, 3, None, 0),
    ("gauss2_f32_fp32", (400, 380), (40, 38), np.float32, 2.0, 3, None, 2),
    ("sobel_u8", (300, 200), (30, 20), np.uint8, "sobel", 3, np.float32, 0),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_noalloc_matches_sep_filter(ipf, oracle, case):
    name, big, small, dt, k, border, out_dtype, precision = case
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    if k == "sobel":
        kx, ky = np.array([-1.0, 0.0, 1.0]), np.array([1.0, 2.0, 1.0])
    else:
        kx = ky = oracle.gaussian_kernel_1d(k)
    for shape in (big, small):
        if dt == np.float32:
            img = rng.random(shape).astype(np.float32)
        else:
            img = rng.integers(0, np.iinfo(dt).max + 1, shape).astype(dt)
        got, want = _both(ipf, img, kx, ky, border, precision, out_dtype, bval=7.0)
        assert got.dtype == want.dtype and got.shape == want.shape
        assert np.array_equal(got, want), (name, shape, int((got != want).sum()))
        if precision == 0:  # DEFAULT is the reference's result
            ref = oracle.sep_filter(img, kx, ky, border=border, bval=7.0, out_dtype=out_dtype)
            assert np.array_equal(got, ref), (name, shape)


def test_noalloc_heavy_fixups_and_anchor(ipf, oracle):
    # weights 0.5: every uint8 output of an odd pair sum sits on a rounding tie
    # and is recomputed in the reference's order; an off-centre anchor and delta
    rng = np.random.default_rng(5)
    kx, ky = np.array([0.5, 0.5]), np.array([0.25, 0.5, 0.25])
    for shape in ((333, 257), (40, 21)):
        img = rng.integers(0, 256, shape).astype(np.uint8)
        got, want = _both(ipf, img, kx, ky, 4, delta=0.5, anchor=(0, 2))
        assert np.array_equal(got, want)
        assert np.array_equal(got, oracle.sep_filter(img, kx, ky, anchor=(0, 2), border=4, delta=0.5))


def test_noalloc_rejects_what_it_cannot_do(ipf):
    rng = np.random.default_rng(6)
    img = rng.integers(0, 256, (32, 32), dtype=np.uint8)
    d, o = ipf.DeviceImage.from_array(img), ipf.DeviceImage.from_array(img)
    k = np.ones(3)
    with pytest.raises(ipf.IpfError, match="device-resident"):
        ipf.sep_filter2d_noalloc(ipf._desc(img), o, k, k)
    with pytest.raises(ipf.DeviceError, match="at most 128 taps"):
        ipf.sep_filter2d_noalloc(d, o, np.ones(129) / 129, k)
    with pytest.raises(ipf.DeviceError, match="no EXACT mode"):
        ipf.sep_filter2d_noalloc(d, o, k, k, precision=1)
    with pytest.raises(ipf.IpfError, match="in place"):
        ipf.sep_filter2d_noalloc(d, d, k, k)
    with pytest.raises(ipf.IpfError, match="separable filter: empty kernel"):
        ipf.sep_filter2d_noalloc(d, o, np.zeros(0), k)
    # int32 images need the exact two-pass path
    i32 = ipf.DeviceImage.from_array(img.astype(np.int32))
    with pytest.raises(ipf.DeviceError, match="exact two-pass path"):
        ipf.sep_filter2d_noalloc(i32, ipf.DeviceImage.like(i32), k, k)
