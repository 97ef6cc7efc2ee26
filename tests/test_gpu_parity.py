"""GPU parity: libipfg (gfx950 kernels through the C ABI) against the
reference-generated fixtures and the pinned oracle.

Bar: integer outputs and the EXACT precision mode are bit-identical to the
reference; float outputs of the FP32 path stay within the bound computed by
tests.fixtures.float_tolerance (written next to the formula)."""
from __future__ import annotations

import gc
import hashlib
import json
import threading

import numpy as np
import pytest

from tests import fixtures

pytestmark = pytest.mark.gpu

CASES = fixtures.load_cases()
IDS = [fixtures.case_id(c) for c in CASES]


def _exact_expected(case, precision) -> bool:
    """Where libipfg must reproduce the reference bit for bit.

    DEFAULT accumulates in FP64 and certifies every integer and float32
    output, EXACT replays the reference's operation order; the exceptions are
    box sums of float images in DEFAULT (FP64 window sums instead of the
    reference's left-to-right running sums) and the float bilateral filter,
    whose range weights use the device exp(), and phase, whose angle comes
    from an FP64 atan2 rounded to float in place of the host's atan2f."""
    c, inp, want, _ = case
    if c["op"] == "median":
        return True
    if c["op"] == "bilateral" and inp.dtype.kind == "f":
        return False
    if c["op"] == "phase":
        return False
    if precision == 1:
        return True
    if c["op"] == "box":
        return inp.dtype.kind in "ui"
    return True


@pytest.mark.parametrize("precision", [0, 1], ids=["default", "exact"])
@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_fixture(ipf, case, precision):
    out = fixtures.run_gpu(ipf, case, precision)
    want = case[2]
    assert out.dtype == want.dtype and out.shape == want.shape
    if _exact_expected(case, precision):
        assert np.array_equal(out, want, equal_nan=True), (
            f"{(out != want).sum()} samples differ, max {np.abs(out.astype(np.float64) - want).max()}")
    else:
        tol = fixtures.float_tolerance(case)
        err = np.abs(out.astype(np.float64) - want.astype(np.float64))
        assert np.all(err <= tol), f"max err {err.max()} vs tol {np.max(tol)}"


# ----------------------------------------------------------------- edge shapes
EDGE_SHAPES = [(0, 5, 1), (5, 0, 3), (1, 1, 1), (1, 9, 3), (13, 1, 1), (2, 2, 2), (65, 63, 1), (31, 130, 4),
               (17, 67, 5), (129, 257, 3)]


@pytest.mark.parametrize("shape", EDGE_SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_edge_shapes_vs_oracle(ipf, oracle, shape):
    rng = np.random.default_rng(sum(shape) + 7)
    u8 = rng.integers(0, 256, shape, dtype=np.uint8)
    f = rng.random(shape).astype(np.float32)
    u16 = rng.integers(0, 65536, shape, dtype=np.uint16)
    i32 = rng.integers(-1000, 1000, shape).astype(np.int32)
    f64 = rng.random(shape)
    for b in range(5):
        for a in (u8, u16, i32, f64):
            assert np.array_equal(ipf.gaussian_blur(a, 1.7, border=b), oracle.gaussian_blur(a, 1.7, border=b))
        assert np.array_equal(ipf.gaussian_blur(f, 2.0, border=b), oracle.gaussian_blur(f, 2.0, border=b))
        g = ipf.gaussian_blur(f, 2.0, border=b, precision=2)
        assert np.abs(g - oracle.gaussian_blur(f, 2.0, border=b)).max(initial=0) <= 3e-6
        for a in (u8, u16, i32):
            assert np.array_equal(ipf.box_filter(a, 5, 7, True, b, 2.0), oracle.box_filter(a, 5, 7, True, b, 2.0))
        K = rng.normal(size=(5, 3)).astype(np.float32)
        assert np.array_equal(ipf.filter2d(u8, K, b, 1.0, (-1, -1), 9.0), oracle.filter2d(u8, K, b, 1.0, (-1, -1), 9.0))
        assert np.array_equal(ipf.bilateral_filter(u8, 5, 30.0, 2.0, b), oracle.bilateral(u8, 5, 30.0, 2.0, b))
        if b:
            for k in (3, 5, 7):
                for a in (u8, u16, i32, f, f64):
                    assert np.array_equal(ipf.median_blur(a, k, b), oracle.median(a, k, b))


def test_types_and_outputs_vs_oracle(ipf, oracle):
    rng = np.random.default_rng(5)
    shape = (45, 70, 3)
    imgs = {
        np.uint8: rng.integers(0, 256, shape, dtype=np.uint8),
        np.uint16: rng.integers(0, 65536, shape, dtype=np.uint16),
        np.int16: rng.integers(-32768, 32768, shape).astype(np.int16),
        np.int32: rng.integers(-5000, 5000, shape).astype(np.int32),
        np.float32: rng.random(shape).astype(np.float32),
        np.float64: rng.random(shape),
    }
    kx, ky = rng.normal(size=5), rng.normal(size=9)
    for ti, a in imgs.items():
        for to in imgs:
            want = oracle.sep_filter(a, kx, ky, None, 3, 0.5, 0.0, to)
            # DEFAULT (certified FP64) and EXACT both reproduce the reference
            assert np.array_equal(ipf.sep_filter2d(a, kx, ky, 3, 0.5, 0.0, to), want), (ti, to)
            assert np.array_equal(ipf.sep_filter2d(a, kx, ky, 3, 0.5, 0.0, to, precision=1), want), (ti, to)
            if ti in (np.uint8, np.uint16, np.int16, np.float32) and to == np.float32:
                # FP32: |fp32 - ref| <= (kw + kh + 5) u sum|kx| sum|ky| X + u |ref|
                got = ipf.sep_filter2d(a, kx, ky, 3, 0.5, 0.0, to, precision=2)
                x = np.abs(a.astype(np.float64)).max()
                tol = 19 * 2.0 ** -24 * (np.abs(kx).sum() * np.abs(ky).sum() * x + 0.5) * 1.01 + 2.0 ** -24 * np.abs(want)
                assert np.all(np.abs(got.astype(np.float64) - want) <= tol), (ti, to)
        if ti != np.float64:
            assert np.array_equal(ipf.median_blur(a, 5, 1), oracle.median(a, 5, 1))
        if np.issubdtype(np.dtype(ti), np.integer):
            assert np.array_equal(ipf.bilateral_filter(a, 5, 40.0, 3.0), oracle.bilateral(a, 5, 40.0, 3.0))
        K = rng.normal(size=(3, 4)).astype(np.float32)
        if ti != np.float64:
            assert np.array_equal(ipf.filter2d(a, K, 1, 0.25), oracle.filter2d(a, K, 1, 0.25)), ti


def test_row_stride_views(ipf, oracle):
    rng = np.random.default_rng(11)
    big = rng.integers(0, 256, (50, 90, 3), dtype=np.uint8)
    view = big[3:47, 5:77]  # rows with a stride larger than their width
    want = oracle.gaussian_blur(np.ascontiguousarray(view), 2.0)
    assert np.array_equal(ipf.gaussian_blur(view, 2.0), want)
    assert np.array_equal(ipf.median_blur(view, 5), oracle.median(np.ascontiguousarray(view), 5))
    cv = np.ascontiguousarray(view)
    assert np.array_equal(ipf.sobel(view, 1, 0), oracle.sobel(cv, 1, 0))
    assert np.array_equal(ipf.sharpen(view, 0.5), oracle.sharpen(cv, 0.5))
    assert np.array_equal(ipf.unsharp_mask(view, 1.0, 1.2, 1.0), oracle.unsharp_mask(cv, 1.0, 1.2, 1.0))


def test_in_place_device_calls_of_the_register_kernels(ipf, oracle):
    # src and dst the same device buffer: the call reads a copy of the source
    import ctypes as C
    from imageprocessingflex_amd import _lib
    lib = _lib.load()
    rng = np.random.default_rng(12)
    img = rng.integers(0, 256, (120, 256, 3), dtype=np.uint8)
    calls = ((lambda dsc, h: lib.ipfg_sharpen(dsc, dsc, 1.0, 0, h), oracle.sharpen(img, 1.0)),
             (lambda dsc, h: lib.ipfg_unsharp_mask(dsc, dsc, 1.0, 1.5, 0.0, 0, h), oracle.unsharp_mask(img, 1.0, 1.5, 0.0)))
    for fn, ref in calls:
        s = ipf.Stream()
        d = ipf.DeviceImage.from_array(img, s)
        desc = d.descriptor()
        _lib.check(fn(C.byref(desc), s.handle))
        assert np.array_equal(d.download(s).reshape(img.shape), ref)


def test_device_resident_async_and_in_place(ipf, oracle):
    rng = np.random.default_rng(12)
    img = rng.integers(0, 256, (300, 410, 3), dtype=np.uint8)
    s = ipf.Stream()
    d = ipf.DeviceImage.from_array(img, s)
    outs = [ipf.gaussian_blur(d, sig, stream=s) for sig in (0.7, 2.0, 5.0)]
    med = ipf.median_blur(d, 5, stream=s)
    s.synchronize()
    for sig, o in zip((0.7, 2.0, 5.0), outs):
        assert np.array_equal(o.download(s), oracle.gaussian_blur(img, sig))
    assert np.array_equal(med.download(s), oracle.median(img, 5))
    # dst aliasing src on the device reads from a private copy
    from imageprocessingflex_amd import _lib
    import ctypes as C
    desc = d.descriptor()
    kx = np.array([0.25, 0.5, 0.25])
    _lib.check(_lib.load().ipfg_sep_filter(C.byref(desc), C.byref(desc), kx.ctypes.data_as(C.POINTER(C.c_double)), 3,
                                           kx.ctypes.data_as(C.POINTER(C.c_double)), 3, 1, 1, 3, 0.0, 0.0, 0,
                                           s.handle))
    s.synchronize()
    assert np.array_equal(d.download(s), oracle.sep_filter(img, kx, kx))


def test_cached_events_outlive_destroyed_streams(ipf, oracle):
    # Parameter uploads (pinned ring slots), parameter-cache and scratch
    # events recorded on a stream that is then destroyed are recycled by
    # later calls on other streams: the ring wraps (64 slots) over slots last
    # used by the destroyed stream, and the results stay the reference's.
    rng = np.random.default_rng(14)
    img = rng.integers(0, 256, (48, 72, 3), dtype=np.uint8)
    sigmas = [0.5 + 0.013 * i for i in range(80)]  # distinct weights: one upload each
    s = ipf.Stream()
    d = ipf.DeviceImage.from_array(img, s)
    for sig in sigmas[:70]:
        ipf.gaussian_blur(d, sig, stream=s)
    s.synchronize()
    del d, s
    gc.collect()
    for _ in range(16):  # reuse the destroyed stream's memory
        t = ipf.Stream()
        del t
    gc.collect()
    s2 = ipf.Stream()
    d2 = ipf.DeviceImage.from_array(img, s2)
    outs = [(sig, ipf.gaussian_blur(d2, sig + 3.0, stream=s2)) for sig in sigmas]
    s2.synchronize()
    for sig, o in outs[::9]:
        assert np.array_equal(o.download(s2), oracle.gaussian_blur(img, sig + 3.0)), sig


def test_stream_forget_before_caller_destroys_its_stream(ipf, oracle):
    # a stream the caller created itself (hipStreamCreate from the runtime the
    # library uses) carries calls with fresh parameter uploads;
    # ipfg_stream_forget detaches the library's cached events from it before
    # the caller destroys it, and later calls recycle those ring slots
    import ctypes as C
    from imageprocessingflex_amd import _lib
    lib = _lib.load()
    hip = C.CDLL("libamdhip64.so")
    hip.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    hip.hipStreamDestroy.argtypes = [C.c_void_p]
    rng = np.random.default_rng(15)
    img = rng.integers(0, 256, (40, 56), dtype=np.uint8)
    d = ipf.DeviceImage.from_array(img)
    o = ipf.DeviceImage.like(d)
    h = C.c_void_p()
    assert hip.hipStreamCreate(C.byref(h)) == 0
    sd, od = d.descriptor(), o.descriptor()
    for i in range(70):
        _lib.check(lib.ipfg_gaussian_blur(C.byref(sd), C.byref(od), 0.6 + 0.011 * i, 0.0, 0, 3, 0, h))
    _lib.check(lib.ipfg_stream_forget(h))
    assert np.array_equal(o.download(), oracle.gaussian_blur(img, 0.6 + 0.011 * 69))
    assert hip.hipStreamDestroy(h) == 0
    s2 = ipf.Stream()
    for i in range(70):
        got = ipf.gaussian_blur(d, 5.0 + 0.007 * i, stream=s2)
    s2.synchronize()
    assert np.array_equal(got.download(s2), oracle.gaussian_blur(img, 5.0 + 0.007 * 69))


def test_concurrent_host_threads(ipf, oracle):
    rng = np.random.default_rng(13)
    imgs = [rng.integers(0, 256, (120 + 7 * i, 200 - 5 * i, 1 + i % 4), dtype=np.uint8) for i in range(8)]
    want = [oracle.bilateral(im, 9, 40.0, 4.0) for im in imgs]
    got = [None] * 8
    errs = []

    def work(i):
        try:
            for _ in range(3):
                got[i] = ipf.bilateral_filter(imgs[i], 9, 40.0, 4.0)
        except Exception as e:  # pragma: no cover
            errs.append(e)

    ts = [threading.Thread(target=work, args=(i,)) for i in range(8)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errs
    for g, w in zip(got, want):
        assert np.array_equal(g, w)


@pytest.mark.parametrize("shape", [(1024, 1024), (333, 517, 3), (77, 1100, 4)])
def test_certification_heavy_fixups_sep(ipf, oracle, shape):
    # kernel [0.5, 0.5]: every sample whose 2x2 sum is 2 mod 4 sits exactly on
    # a rounding tie, so about a quarter of the samples are recomputed in the
    # reference's order inside the launch. The fractional Constant value keeps
    # the call off the exact-integer-sum route (which certifies nothing).
    rng = np.random.default_rng(14)
    img = rng.integers(0, 256, shape, dtype=np.uint8)
    k = np.array([0.5, 0.5])
    got = ipf.sep_filter2d(img, k, k, 0, 0.0, 0.5)
    assert np.array_equal(got, oracle.sep_filter(img, k, k, border=0, bval=0.5))
    assert ipf.last_fixups() > img.size // 8


@pytest.mark.parametrize("shape", [(1024, 1024), (61, 203, 3)])
def test_certification_heavy_fixups_filter2d(ipf, oracle, shape):
    rng = np.random.default_rng(16)
    img = rng.integers(0, 256, shape, dtype=np.uint8)
    kern = np.array([[0.5, 0.5], [0.5, 0.5]], dtype=np.float32) * 0.5
    got = ipf.filter2d(img, kern, 0, border_value=0.5)
    assert np.array_equal(got, oracle.filter2d(img, kern, border=0, bval=0.5))
    assert ipf.last_fixups() > img.size // 8


def test_non_finite_and_extreme_float_inputs(ipf, oracle):
    # NaN, infinities, values near FLT_MAX and subnormals: certified launches
    # must flag every sample they reach and reproduce the reference's result
    rng = np.random.default_rng(31)
    img = rng.random((150, 333)).astype(np.float32)
    img[10, 20] = np.nan
    img[70, 300] = np.inf
    img[140, 5] = -np.inf
    img[80:90, 100:110] = 3.0e38
    img[100, 150:160] = 1e-41
    kx, ky = rng.normal(size=7), rng.normal(size=5)
    for b in range(5):
        for s in (1.0, 3.0):
            assert np.array_equal(ipf.gaussian_blur(img, s, border=b), oracle.gaussian_blur(img, s, border=b),
                                  equal_nan=True), (s, b)
        for to in (np.float32, np.uint8, np.int16):
            got = ipf.sep_filter2d(img, kx, ky, b, 0.25, 7.0, to)
            assert np.array_equal(got, oracle.sep_filter(img, kx, ky, None, b, 0.25, 7.0, to), equal_nan=True), (b, to)


def test_certified_calls_reuse_stream_state(ipf, oracle):
    # The per-stream counters are reset by each launch: alternate ops that
    # flag samples with ones that do not, on one stream.
    rng = np.random.default_rng(15)
    img = rng.integers(0, 256, (96, 130, 3), dtype=np.uint8)
    s = ipf.Stream()
    d = ipf.DeviceImage.from_array(img, s)
    k = np.array([0.5, 0.5])
    outs = []
    for i in range(6):
        outs.append(ipf.sep_filter2d(d, k, k, 0, 0.0, 0.5, stream=s) if i % 2 == 0
                    else ipf.gaussian_blur(d, 1.3, stream=s))
    s.synchronize()
    for i, o in enumerate(outs):
        want = oracle.sep_filter(img, k, k, border=0, bval=0.5) if i % 2 == 0 else oracle.gaussian_blur(img, 1.3)
        assert np.array_equal(o.download(s), want), i


# ------------------------------------------------------------- full-size runs
FULL = json.loads((fixtures.GOLDEN / "fullsize.json").read_text())
BNAME = {0: "constant", 1: "replicate", 2: "reflect", 3: "reflect101", 4: "wrap"}


# ------------------------------------------------ exact integer (dyadic) sums
def test_dyadic_kernels_on_integer_images_bit_exact(ipf, oracle):
    # integer images with weights m * 2^-f have exact sums in FP64 (and, when
    # small, in FP32): those calls skip certification and must still match
    rng = np.random.default_rng(53)
    imgs = [rng.integers(0, 256, (203, 311, 3), dtype=np.uint8),
            rng.integers(0, 65536, (170, 250), dtype=np.uint16),
            rng.integers(-32768, 32768, (150, 190, 4)).astype(np.int16)]
    seps = [([0.25, 0.5, 0.25], [0.25, 0.5, 0.25], 0.0), ([-1, 0, 1], [1, 2, 1], 0.5),
            ([1, 4, 6, 4, 1], [0.0625, 0.25, 0.375, 0.25, 0.0625], -3.0), ([3, -10.5, 3], [1, 2, 1], 1.25),
            ([2.0 ** 20, 1, -2.0 ** 20], [1, 2, 1], 0.0)]
    for img in imgs:
        for to in (None, np.float32):
            for kx, ky, delta in seps:
                for b, bval in ((0, 7.0), (0, 2.5), (1, 0.0), (3, 0.0), (4, 0.0)):
                    got = ipf.sep_filter2d(img, kx, ky, b, delta, bval, to)
                    want = oracle.sep_filter(img, kx, ky, None, b, delta, bval, to)
                    assert np.array_equal(got, want), (img.dtype, to, kx, ky, delta, b, bval)
        for k in (np.array([[0, 1, 0], [1, -4, 1], [0, 1, 0]], np.float32),
                  np.array([[0.5, 0, -0.5], [0.25, 0.125, 0.25], [-1, 0, 1], [2, 0, 2]], np.float32)):
            for to in (None, np.float32):
                for b, bval in ((0, 3.0), (3, 0.0), (2, 0.0)):
                    got = ipf.filter2d(img, k, b, 0.75, (-1, -1), bval, to)
                    want = oracle.filter2d(img, k, b, 0.75, (-1, -1), bval, to)
                    assert np.array_equal(got, want), (img.dtype, to, k.shape, b)


@pytest.mark.parametrize("shape", [(150, 256, 3), (97, 64), (70, 128, 4), (5, 48, 3), (33, 16), (1, 32, 4),
                                   (2, 1024, 1), (64, 80, 3)], ids=lambda s: "x".join(map(str, s)))
def test_small_exact_stencils_vs_oracle(ipf, oracle, shape):
    # 3x3 / 5x5 kernels with exact FP32 sums on uint8 rows that are whole
    # 16-byte groups take the register-window stencil: interior groups, image
    # edges and every border mode against the reference order
    rng = np.random.default_rng(sum(shape) * 7)
    img = rng.integers(0, 256, shape, dtype=np.uint8)
    ks = [np.array([[1, 2, 1], [2, 4, 2], [1, 2, 1]], np.float32) / 16,
          np.array([[0, -1, 0], [-1, 5, -1], [0, -1, 0]], np.float32),
          np.arange(25, dtype=np.float32).reshape(5, 5) - 12.0,
          np.ones((5, 5), np.float32) / 32]
    for k in ks:
        for to in (None, np.float32):
            for b, bval in ((0, 9.0), (1, 0.0), (2, 0.0), (3, 0.0), (4, 0.0)):
                got = ipf.filter2d(img, k, b, -2.5, (-1, -1), bval, to)
                assert np.array_equal(got, oracle.filter2d(img, k, b, -2.5, (-1, -1), bval, to)), (k.shape, to, b)
        # an off-centre anchor and a fractional delta at the FP32 limit stay on the other kernels
        got = ipf.filter2d(img, k, 3, 0.0, (0, 1))
        assert np.array_equal(got, oracle.filter2d(img, k, 3, 0.0, (0, 1)))
    for dx, dy, kk in ((1, 0, 3), (0, 1, 5), (2, 0, 1), (1, 1, 3), (0, 2, 5)):
        assert np.array_equal(ipf.sobel(img, dx, dy, kk, 0.5, 4), oracle.sobel(img, dx, dy, kk, 0.5, 4))
    for kk in (3, 5):
        for b in (0, 3):
            assert np.array_equal(ipf.laplacian(img, kk, 1.5, b), oracle.laplacian(img, kk, 1.5, b)), (kk, b)
    d = ipf.DeviceImage.from_array(img)
    out = ipf.sharpen(d, 0.75)
    assert np.array_equal(out.download().reshape(img.shape), oracle.sharpen(img, 0.75))


@pytest.mark.parametrize("shape", [(211, 320, 3), (97, 256), (64, 96, 4), (7, 48, 3), (40, 20, 1), (3, 400, 4)],
                         ids=lambda s: "x".join(map(str, s)))
def test_unsharp_u8_single_launch_vs_oracle(ipf, oracle, shape):
    # uint8 unsharpMask with windows up to 13 taps: FP32 blur and certified
    # combine in one launch (interior, frame rows and columns, recomputed
    # doubtful samples); larger sigmas take the two-pass route
    rng = np.random.default_rng(sum(shape) * 11)
    img = rng.integers(0, 256, shape, dtype=np.uint8)
    if img.ndim == 3 and img.shape[0] > 20:
        img[5:15, 8:30] = 128  # a flat patch: blur == sample up to rounding
    for sigma, amount, th in ((0.5, 1.0, 0.0), (1.0, 1.5, 0.0), (1.5, -0.6, 3.0), (2.0, 2.0, 1.0), (2.6, 0.8, 0.0)):
        got = ipf.unsharp_mask(img, sigma, amount, th)
        assert np.array_equal(got, oracle.unsharp_mask(img, sigma, amount, th)), (sigma, amount, th)


# ------------------------------------------- derivative and sharpening filters
def test_derivative_filters_vs_oracle(ipf, oracle):
    rng = np.random.default_rng(41)
    imgs = [rng.integers(0, 256, (300, 517, 3), dtype=np.uint8),
            (rng.random((257, 411)) * 2 - 1).astype(np.float32),
            rng.integers(-30000, 30000, (190, 263, 2)).astype(np.int16)]
    for img in imgs:
        for b in (0, 1, 3, 4):
            for dx, dy, k in ((1, 0, 3), (0, 1, 5), (1, 1, 3), (2, 0, 1), (0, 2, 7), (1, 2, 31), (3, 0, 9)):
                got, want = ipf.sobel(img, dx, dy, k, 0.75, b), oracle.sobel(img, dx, dy, k, 0.75, b)
                assert np.array_equal(got, want), (img.dtype, b, dx, dy, k)
            for dx, dy in ((1, 0), (0, 1)):
                assert np.array_equal(ipf.scharr(img, dx, dy, 2.0, b), oracle.scharr(img, dx, dy, 2.0, b))
                assert np.array_equal(ipf.prewitt(img, dx, dy, b), oracle.prewitt(img, dx, dy, b))
            for k in (1, 3, 7):
                assert np.array_equal(ipf.laplacian(img, k, 0.5, b), oracle.laplacian(img, k, 0.5, b)), (b, k)


def test_magnitude_bit_exact_and_phase_within_an_ulp(ipf, oracle):
    rng = np.random.default_rng(43)
    n = 1 << 20
    gx = (rng.standard_normal(n) * 10.0 ** rng.integers(-38, 38, n)).astype(np.float32).reshape(512, 2048)
    gy = (rng.standard_normal(n) * 10.0 ** rng.integers(-38, 38, n)).astype(np.float32).reshape(512, 2048)
    gy[0, :1024] = gx[0, :1024] * rng.random(1024).astype(np.float32)  # near-equal magnitudes
    special = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 1e-45, -1e-45, 3.4e38, -3.4e38, 1.0, -1.0], np.float32)
    sx, sy = np.meshgrid(special, special)
    gx[1, :sx.size], gy[1, :sy.size] = sx.ravel(), sy.ravel()
    assert np.array_equal(ipf.magnitude(gx, gy), oracle.magnitude(gx, gy), equal_nan=True)
    for deg in (True, False):
        got, want = ipf.phase(gx, gy, deg), oracle.phase(gx, gy, deg)
        assert np.array_equal(np.isnan(got), np.isnan(want))
        ok = ~np.isnan(want)
        err = np.abs(got[ok].astype(np.float64) - want[ok])
        assert np.all(err <= fixtures.phase_tolerance(gx, gy, want, deg)[ok]), err.max()
        assert np.all((got[ok] >= 0) & (got[ok] <= (360.0 if deg else 2 * np.pi)))
    img = rng.integers(0, 256, (211, 307, 3), dtype=np.uint8)
    assert np.array_equal(ipf.gradient_magnitude(img, 5),
                          oracle.magnitude(oracle.sobel(img, 1, 0, 5), oracle.sobel(img, 0, 1, 5)))


def test_sharpening_and_adaptive_threshold_vs_oracle(ipf, oracle):
    rng = np.random.default_rng(47)
    imgs = [rng.integers(0, 256, (300, 517, 3), dtype=np.uint8),
            rng.random((257, 411)).astype(np.float32) * 255,
            rng.integers(0, 65536, (190, 263), dtype=np.uint16)]
    for img in imgs:
        for s, a, th in ((1.0, 1.0, 0.0), (2.2, 0.7, 3.0), (6.0, 1.5, 0.5)):
            assert np.array_equal(ipf.unsharp_mask(img, s, a, th), oracle.unsharp_mask(img, s, a, th)), (s, a, th)
        for st in (1.0, 0.3, 2.5):
            assert np.array_equal(ipf.sharpen(img, st), oracle.sharpen(img, st)), st
        if img.dtype == np.uint8:
            # flat regions: the blur equals the sample up to FP64 rounding, so
            # the FP32 blur cannot decide the threshold and every such sample
            # is recomputed in the reference's order
            flat = np.full((64, 97, 3), 77, np.uint8)
            flat[20:40, 30:60] = 200
            flat[50:, :10] = img[:14, :10]
            for s, a, th in ((1.0, 1.5, 0.0), (3.0, -0.7, 0.0), (2.0, 1.0, 2.0)):
                assert np.array_equal(ipf.unsharp_mask(flat, s, a, th), oracle.unsharp_mask(flat, s, a, th)), (s, a, th)
        assert np.array_equal(ipf.emboss(img), oracle.emboss(img))
        if img.ndim == 2:
            for m in (0, 1):
                for tt in (0, 1):
                    for bs in (3, 11, 31):
                        got = ipf.adaptive_threshold(img, 177.0, m, tt, bs, -1.5)
                        assert np.array_equal(got, oracle.adaptive_threshold(img, 177.0, m, tt, bs, -1.5)), (m, tt, bs)


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_full_c1_gaussian_bit_exact(ipf):
    from imageprocessingflex_amd import workloads as wl
    out = ipf.gaussian_blur(wl.c1_image(), 2.0)
    assert _sha(out) == FULL["c1_gauss2_reflect101_u8"]
    print("c1 fix-ups", ipf.last_fixups())


def test_full_c3_median_and_bilateral_bit_exact(ipf):
    from imageprocessingflex_amd import workloads as wl
    img = wl.polygon_image()
    d = ipf.DeviceImage.from_array(img)
    for k in (3, 5):
        for b in (1, 2, 3, 4):
            assert _sha(ipf.median_blur(d, k, b).download()) == FULL[f"c3_median{k}_{BNAME[b]}"], (k, b)
    out = ipf.bilateral_filter(d, 9, 75.0, 75.0).download()
    assert _sha(out) == FULL["c3_bilateral9_75_75_reflect101"]


def test_full_c4_filter2d_u8_bit_exact(ipf):
    from imageprocessingflex_amd import workloads as wl
    u8, _ = wl.c4_images()
    k, _ = wl.c4_kernel()
    assert _sha(ipf.filter2d(u8, k)) == FULL["c4_filter2d31_reflect101_u8"]
    print("c4 fix-ups", ipf.last_fixups())


def test_full_size_derivative_workload_vs_oracle(ipf, oracle):
    # the derivative / sharpening calls at 3840 x 2160 (RGB and gray uint8),
    # against the oracle (all host threads)
    from imageprocessingflex_amd import workloads as wl
    rgb = wl.uniform_u8((2160, 3840, 3), seed=7, stream=0)
    gray = wl.uniform_u8((2160, 3840), seed=7, stream=1)
    assert np.array_equal(ipf.sobel(rgb, 1, 0, 3), oracle.sobel(rgb, 1, 0, 3))
    assert np.array_equal(ipf.sobel(rgb, 0, 1, 5), oracle.sobel(rgb, 0, 1, 5))
    assert np.array_equal(ipf.laplacian(rgb, 3), oracle.laplacian(rgb, 3))
    assert np.array_equal(ipf.unsharp_mask(rgb, 1.0, 1.0, 0.0), oracle.unsharp_mask(rgb, 1.0, 1.0, 0.0))
    assert np.array_equal(ipf.sharpen(rgb, 1.0), oracle.sharpen(rgb, 1.0))
    for m in (0, 1):
        assert np.array_equal(ipf.adaptive_threshold(gray, 255.0, m, 0, 11, 2.0),
                              oracle.adaptive_threshold(gray, 255.0, m, 0, 11, 2.0)), m


# -------------------------------------------------------------------- pyramids
PYR_SHAPES = [(1, 1, 1), (1, 2, 1), (2, 1, 3), (3, 5, 3), (8, 13, 3), (33, 27, 1), (47, 65, 4), (130, 257, 2),
              (255, 513, 1)]


def _pyr_images(rng, shape):
    f = (rng.normal(size=shape) * 1e3).astype(np.float32)
    flat = f.reshape(-1)
    if flat.size >= 8:  # specials: NaN, infinities, signed zeros, a denormal
        flat[[0, 3, 5, 6, 7]] = [np.nan, np.inf, -np.inf, -0.0, 1e-42]
    return [rng.integers(0, 256, shape, dtype=np.uint8), rng.integers(0, 65536, shape, dtype=np.uint16),
            rng.integers(-32768, 32768, shape).astype(np.int16),
            rng.integers(-2 ** 31, 2 ** 31, shape, dtype=np.int64).astype(np.int32), f, rng.normal(size=shape) * 1e6]


def _same_bits(got, want):
    """Bit-identical, NaNs compared as NaNs (their payloads are not part of the contract)."""
    assert got.shape == want.shape and got.dtype == want.dtype
    if want.dtype.kind != "f":
        return np.array_equal(got, want)
    nan = np.isnan(want)
    u = np.dtype(f"u{want.dtype.itemsize}")
    return np.array_equal(np.isnan(got), nan) and np.array_equal(np.where(nan, 0, got).view(u),
                                                                 np.where(nan, 0, want).view(u))


@pytest.mark.parametrize("shape", PYR_SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_pyramid_steps_bit_exact_vs_oracle(ipf, oracle, shape):
    rng = np.random.default_rng(sum(shape) + 11)
    for a in _pyr_images(rng, shape):
        assert _same_bits(ipf.pyr_down(a), oracle.pyr_down(a)), (shape, a.dtype, "down")
        assert _same_bits(ipf.pyr_up(a), oracle.pyr_up(a)), (shape, a.dtype, "up")


def test_pyramid_device_resident_row_stride_and_levels(ipf, oracle):
    rng = np.random.default_rng(23)
    wide = rng.integers(0, 256, (61, 140, 3), dtype=np.uint8)
    view = wide[:, 7:104]  # rows 420 bytes apart, 291 bytes used
    want = oracle.pyr_down(np.ascontiguousarray(view))
    assert np.array_equal(ipf.pyr_down(view), want)
    d = ipf.DeviceImage.from_array(np.ascontiguousarray(view))
    assert np.array_equal(ipf.pyr_down(d).download(), want)
    assert np.array_equal(ipf.pyr_up(d).download(), oracle.pyr_up(np.ascontiguousarray(view)))
    levels = ipf.build_pyramid(d, 9)
    assert [(x.width, x.height) for x in levels] == [(97, 61), (49, 31), (25, 16), (13, 8), (7, 4), (4, 2), (2, 1)]
    level = np.ascontiguousarray(view)
    for x in levels:
        got = x.download() if isinstance(x, ipf.DeviceImage) else x
        assert np.array_equal(got, level)
        level = oracle.pyr_down(level)


def test_full_size_pyramids_vs_reference(ipf):
    # 3840 x 2160 RGB uint8 (the medianBlur workload's image): SHA-256 of the
    # reference's own outputs
    from imageprocessingflex_amd import workloads as wl
    img = ipf.DeviceImage.from_array(wl.polygon_image())
    assert _sha(ipf.pyr_down(img).download()) == FULL["c3_pyrdown_u8"]
    assert _sha(ipf.pyr_up(img).download()) == FULL["c3_pyrup_u8"]


def test_full_c2_exact_mode_bit_exact_and_fast_within_bound(ipf):
    from imageprocessingflex_amd import workloads as wl
    img = wl.sinusoid_image()
    d = ipf.DeviceImage.from_array(img)
    U = 2.0 ** -24
    for s in (1.0, 3.0, 8.0):
        ks = 2 * max(1, int(3 * s + 0.5)) + 1
        for b in range(5):
            exact = ipf.gaussian_blur(d, s, border=b, precision=1).download()
            assert _sha(exact) == FULL[f"c2_gauss{s:g}_{BNAME[b]}_f32"], (s, b)
            default = ipf.gaussian_blur(d, s, border=b).download()
            assert _sha(default) == FULL[f"c2_gauss{s:g}_{BNAME[b]}_f32"], (s, b)
            fast = ipf.gaussian_blur(d, s, border=b, precision=2).download()
            # FP32: |fp32 - ref| <= (2 ks + 5) u X + u |ref|, X = 1 (image in [0, 1], kernel sums 1)
            tol = (2 * ks + 5) * U * 1.01 + U * np.abs(exact)
            assert np.all(np.abs(fast.astype(np.float64) - exact) <= tol), (s, b)
    for b in range(5):
        exact = ipf.blur(d, 15, b, precision=1).download()
        assert _sha(exact) == FULL[f"c2_box15_{BNAME[b]}_f32"], b
        fast = ipf.blur(d, 15, b).download()
        assert np.all(np.abs(fast.astype(np.float64) - exact) <= 1e-12 * 225 + U * np.abs(exact)), b


# ------------------------------------------------------------- autotuner
def test_every_strip_height_gives_the_reference_result(ipf, oracle):
    # the autotuner may pick any candidate strip height: each must be exact
    rng = np.random.default_rng(21)
    cases = [(rng.integers(0, 256, (203, 301, 3), dtype=np.uint8), 2.0),
             (rng.random((157, 263)).astype(np.float32), 3.0)]
    try:
        for img, sigma in cases:
            want = oracle.gaussian_blur(img, sigma)
            h, w = img.shape[:2]
            cn = 1 if img.ndim == 2 else img.shape[2]
            dt = 0 if img.dtype == np.uint8 else 4
            k = 2 * max(1, int(3 * sigma + 0.5)) + 1
            wl, hl = int(np.ceil(np.log2(w))), int(np.ceil(np.log2(h)))
            # 16-row steps at any height; for one-channel float32, 8-row steps (1 << 20 + height)
            extra = tuple(1048576 + v for v in (8, 16, 32, 64, 128, 256)) if img.ndim == 2 else ()
            for sh in (16, 32, 48, 64, 96, 128, 192, 256, 384, 512) + extra:
                ipf.tune_clear()
                ipf.tune_import(f"sep {dt} {dt} {cn} {k} {k} 1 {wl} {hl} {sh}\n")
                got = ipf.gaussian_blur(img, sigma)
                assert np.array_equal(got, want), (img.dtype, sh)
    finally:
        ipf.tune_clear()


TILE_SH = 1 << 22  # strip heights TILE_SH + rows select the one-pass tile kernel (uint8 -> uint8)


@pytest.mark.parametrize("cn", [1, 3, 4])
def test_tile_kernel_gives_the_reference_result(ipf, oracle, cn):
    # every kernel size the tile kernel takes, every border, signed weights
    # (outputs saturate at both ends), a delta, a fractional Constant value,
    # and images smaller than one tile, ragged at both edges, or one row tall
    rng = np.random.default_rng(60 + cn)
    shapes = [(1, 1), (1, 40), (5, 3), (37, 70), (67, 131), (150, 203)]
    try:
        for k in range(3, 17, 2):
            kx = rng.normal(0.0, 1.0, k) / k
            ky = rng.normal(0.0, 1.0, k) + 1.0 / k
            for i, (h, w) in enumerate(shapes):
                img = rng.integers(0, 256, (h, w, cn), dtype=np.uint8)
                wl, hl = int(np.ceil(np.log2(w))), int(np.ceil(np.log2(h)))
                for b in range(5):
                    delta = 3.25 if (k + i + b) % 2 else 0.0
                    want = oracle.sep_filter(img, kx, ky, None, b, delta, 17.5)
                    for rows in (16, 32):
                        ipf.tune_clear()
                        ipf.tune_import(f"sep 0 0 {cn} {k} {k} 1 {wl} {hl} {TILE_SH + rows}\n")
                        got = ipf.sep_filter2d(img, kx, ky, b, delta, 17.5)
                        assert np.array_equal(got, want), (k, h, w, b, rows)
    finally:
        ipf.tune_clear()


def test_tile_kernel_gaussian_and_recompute_count(ipf, oracle):
    # gaussian weights on an interior-heavy image, with the recompute count:
    # the float-rounded sums leave a few doubtful samples per million
    rng = np.random.default_rng(70)
    img = rng.integers(0, 256, (333, 517, 3), dtype=np.uint8)
    try:
        for sigma, k in ((0.8, 5), (2.0, 13), (2.4, 15)):
            want = oracle.gaussian_blur(img, sigma)
            for rows in (16, 32):
                ipf.tune_clear()
                ipf.tune_import(f"sep 0 0 3 {k} {k} 1 10 9 {TILE_SH + rows}\n")
                got = ipf.gaussian_blur(img, sigma)
                assert np.array_equal(got, want), (sigma, rows)
                assert ipf.last_fixups() < img.size // 1000, (sigma, rows, ipf.last_fixups())
    finally:
        ipf.tune_clear()


def test_tile_kernel_full_c1_hash(ipf):
    from imageprocessingflex_amd import workloads as wl
    try:
        img = wl.c1_image()
        for rows in (16, 32):
            ipf.tune_clear()
            ipf.tune_import(f"sep 0 0 3 13 13 1 11 11 {TILE_SH + rows}\n")
            assert _sha(ipf.gaussian_blur(img, 2.0)) == FULL["c1_gauss2_reflect101_u8"], rows
    finally:
        ipf.tune_clear()


def test_autotuned_call_matches_untuned(ipf):
    rng = np.random.default_rng(22)
    img = rng.integers(0, 256, (1100, 1000, 1), dtype=np.uint8)
    try:
        ipf.autotune(False)
        ipf.tune_clear()
        plain = ipf.gaussian_blur(img, 2.5)
        ipf.autotune(True)
        tuned = ipf.gaussian_blur(img, 2.5)
        assert "sep 0 0 1 17 17 1" in ipf.tune_export()
        again = ipf.gaussian_blur(img, 2.5)
        assert np.array_equal(tuned, plain) and np.array_equal(again, plain)
    finally:
        ipf.autotune(True)
        ipf.tune_clear()


# ------------------------------------------------------------- size-specialised kernels
@pytest.mark.parametrize("ks", list(range(3, 51, 2)))
def test_size_specialised_kernels_vs_oracle(ipf, oracle, ks):
    # gaussian windows of every specialised size, on uint8 1/3/4-channel and
    # float32 images, each under a different border mode
    rng = np.random.default_rng(100 + ks)
    k = oracle.gaussian_kernel_1d(0.0, ks)
    cases = [rng.integers(0, 256, (83, 150), dtype=np.uint8),
             rng.integers(0, 256, (61, 97, 3), dtype=np.uint8),
             rng.integers(0, 256, (45, 70, 4), dtype=np.uint8),
             rng.random((90, 130)).astype(np.float32)]
    for i, img in enumerate(cases):
        b = (ks // 2 + i) % 5
        got = ipf.sep_filter2d(img, k, k, b)
        want = oracle.sep_filter(img, k, k, border=b)
        assert np.array_equal(got, want), (img.dtype, img.shape, b)


# ------------------------------------------------------------- MFMA filter2D (split precision)
def _mfma_bound(kernel, xmax, precision, float_input):
    # input split error + weight split error + FP32 accumulation over every product
    k = np.asarray(kernel, np.float64)
    nprod = k.size * (3 if precision == 3 else 2) + (k.size if float_input else 0)
    xs = (2.0 ** -16 if precision == 3 else 2.0 ** -22) if float_input else 0.0
    ws = 2.0 ** -24 if precision == 3 else 2.0 ** -22
    return (xs + ws + 2 * nprod * 2.0 ** -24) * np.abs(k).sum() * xmax


@pytest.mark.parametrize("precision", [3, 4], ids=["bf16x3", "fp16x2"])
@pytest.mark.parametrize("shape,ks,anchor", [((97, 131), (31, 31), (-1, -1)), ((64, 64, 3), (7, 5), (1, 3)),
                                             ((70, 45), (3, 3), (-1, -1)), ((33, 200), (9, 17), (8, 0))])
def test_filter2d_mfma_error_within_bound(ipf, oracle, precision, shape, ks, anchor):
    rng = np.random.default_rng(31)
    kernel = rng.standard_normal((ks[1], ks[0]))
    kernel = (kernel / kernel.sum()).astype(np.float32)
    for dt, scale in ((np.uint8, 255), (np.float32, 1.0), (np.float32, 1e-3), (np.float32, 1e3)):
        img = (rng.integers(0, 256, shape).astype(dt) if dt == np.uint8
               else (rng.random(shape) * scale).astype(np.float32))
        for b in (0, 3, 4):
            want = oracle.filter2d(img, kernel, border=b, anchor=anchor, out_dtype=np.float64)
            got = ipf.filter2d(img, kernel, b, anchor=anchor, out_dtype=np.float32, precision=precision)
            err = np.abs(got.astype(np.float64) - want)
            bound = _mfma_bound(kernel, float(np.abs(img).max()), precision, dt != np.uint8) + \
                2.0 ** -24 * np.abs(want).max()
            assert err.max() <= bound, (dt, scale, b, err.max(), bound)


@pytest.mark.parametrize("precision", [3, 4], ids=["bf16x3", "fp16x2"])
def test_filter2d_mfma_u8_output_close_to_reference(ipf, oracle, precision):
    # uint8 -> uint8: results differ from the reference only at rounding ties
    rng = np.random.default_rng(32)
    img = rng.integers(0, 256, (160, 190), dtype=np.uint8)
    kernel = rng.standard_normal((31, 31))
    kernel = (kernel / kernel.sum()).astype(np.float32)
    want = oracle.filter2d(img, kernel, border=3)
    got = ipf.filter2d(img, kernel, 3, precision=precision)
    d = np.abs(got.astype(int) - want.astype(int))
    assert d.max() <= 1 and (d > 0).mean() < 0.01


# ------------------------------------------------------------- batches and serving
def _mixed_jobs(rng, n):
    # up to 600 x 520 x 3: many strips and tiles per job
    shapes = [(17, 23, 1), (64, 64, 3), (130, 257, 4), (300, 200, 1), (5, 700, 3), (96, 33, 4), (1, 1, 1),
              (255, 129, 3), (40, 41, 1), (600, 520, 3), (530, 520, 1)]
    out = []
    for i in range(n):
        h, w, cn = shapes[i % len(shapes)]
        img = rng.integers(0, 256, (h, w, cn) if cn > 1 else (h, w), dtype=np.uint8)
        kind = i % 4
        if kind in (0, 1):
            sigma = [0.5, 2.0, 5.3, 8.0, 1.1][i % 5]
            out.append((img, "gaussian", dict(sigma=sigma, border=(i // 4) % 5)))
        elif kind == 2:
            out.append((img, "median", dict(ksize=5 if i % 8 == 2 else 3, border=1 + (i // 4) % 4)))
        else:
            out.append((img, "bilateral", dict(d=9 if i % 8 == 3 else 5, sigma_color=75.0 if i % 8 == 3 else 20.0,
                                               sigma_space=75.0 if i % 8 == 3 else 10.0, border=(i // 4) % 5)))
    return out


def _oracle_result(oracle, img, op, kw):
    if op == "gaussian":
        return oracle.gaussian_blur(img, kw["sigma"], border=kw["border"])
    if op == "median":
        return oracle.median(img, kw["ksize"], kw["border"])
    return oracle.bilateral(img, kw["d"], kw["sigma_color"], kw["sigma_space"], kw["border"])


@pytest.mark.parametrize("graph", [True, False], ids=["graph", "eager"])
def test_batch_matches_single_image_results(ipf, oracle, graph):
    rng = np.random.default_rng(41)
    items = _mixed_jobs(rng, 36)
    want = [_oracle_result(oracle, img, op, kw) for img, op, kw in items]
    s = ipf.Stream()
    srcs = [ipf.DeviceImage.from_array(img, s) for img, _, _ in items]
    b = ipf.Batch(64, graph=graph)
    for rep in range(3):  # replays of the same captured graph, with other work in between
        dsts = [ipf.DeviceImage.like(d) for d in srcs]
        jobs = [ipf.job(d, o, op, **kw) for d, o, (_, op, kw) in zip(srcs, dsts, items)]
        b.run(jobs, stream=s)
        ipf.gaussian_blur(srcs[1], 1.5, stream=s)
        s.synchronize()
        for i, (o, w_) in enumerate(zip(dsts, want)):
            assert np.array_equal(o.download(s).reshape(w_.shape), w_), (rep, i, items[i][1], items[i][2])
    b.close()


def test_batch_rejects_unsupported_jobs(ipf):
    s = ipf.Stream()
    a = ipf.DeviceImage.from_array(np.zeros((8, 8), np.uint8), s)
    o = ipf.DeviceImage.like(a)
    b = ipf.Batch(4)
    with pytest.raises(ipf.IpfError, match="medianBlur: Constant border is not supported"):
        b.run([ipf.job(a, o, "median", ksize=3, border=0)], stream=s)
    f = ipf.DeviceImage.from_array(np.zeros((8, 8), np.float32), s)
    with pytest.raises(ipf.IpfError, match="uint8"):
        b.run([ipf.job(f, ipf.DeviceImage.like(f), "gaussian", sigma=1.0)], stream=s)
    b.close()


@pytest.mark.parametrize("budget_mb,graph", [(256, True), (3, True), (64, False)], ids=["roomy", "tight", "eager"])
def test_server_results_same_alone_and_in_the_mix(ipf, oracle, budget_mb, graph):
    # every request's result equals the single-image (reference) result, whatever
    # it was batched with and however the device budget throttled admission
    rng = np.random.default_rng(42)
    items = _mixed_jobs(rng, 45)
    srv = ipf.Server(budget_mb << 20, max_batch=16, graph=graph)
    outs = []
    for img, op, kw in items:
        out = np.empty_like(img)
        outs.append(out)
        srv.submit(img, out, op, **kw)
    srv.drain()
    done = srv.poll()
    assert sorted(i for i, _ in done) == list(range(len(items)))
    assert all(lat >= 0 for _, lat in done)
    batches, biggest = srv.stats()
    assert 1 <= batches <= len(items) and 1 <= biggest <= 16
    for i, ((img, op, kw), out) in enumerate(zip(items, outs)):
        assert np.array_equal(out, _oracle_result(oracle, img, op, kw)), (i, op, kw)
    srv.close()


def test_server_rejects_request_above_budget(ipf):
    srv = ipf.Server(1 << 10, max_batch=4)
    img = np.zeros((64, 64), np.uint8)
    with pytest.raises(ipf.IpfError, match="larger than the device budget"):
        srv.submit(img, np.empty_like(img), "gaussian", sigma=1.0)
    srv.close()


# ------------------------------------------------------------------------ ssim
def _ssim_tol(n, smap):
    """Summation-order bound on the mean: the reference's running sum and the device
    tree sum each stay within (n + 64) * 2^-53 * sum|map| of the exact sum."""
    return 2.0 * (n + 64) * 2.0 ** -53 * float(np.mean(np.abs(smap.astype(np.float64)))) + 1e-300


SSIM_SHAPES = [(1, 1, 1), (3, 40, 1), (11, 11, 3), (10, 30, 2), (37, 53, 4), (130, 257, 1), (271, 97, 3)]


@pytest.mark.parametrize("shape", SSIM_SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_ssim_map_bit_exact_and_mean_vs_oracle(ipf, oracle, shape):
    rng = np.random.default_rng(sum(shape) + 5)
    for a in _pyr_images(rng, shape):
        if a.dtype.kind == "f":
            a = np.nan_to_num(a, nan=0.0, posinf=1.0, neginf=-1.0)
            b = (a + rng.normal(size=shape) * 0.1 * (np.abs(a).max() + 1)).astype(a.dtype)
        else:
            info = np.iinfo(a.dtype)
            b = np.clip(a.astype(np.float64) + rng.normal(size=shape) * 50, info.min, info.max).astype(a.dtype)
        want_v, want_m = oracle.ssim(a, b)
        got_v, got_m = ipf.ssim(a, b, return_map=True)
        assert _same_bits(got_m, want_m), (shape, a.dtype)
        assert abs(got_v - want_v) <= _ssim_tol(a.size, want_m), (shape, a.dtype, got_v, want_v)
        assert ipf.ssim(a, b) == got_v


def test_ssim_full_size_device_resident(ipf, oracle):
    # 1920x1080x3 uint8 (the gaussianBlur workload's image) against a noisy copy: map bit-exact, mean within the summation-order bound.
    from imageprocessingflex_amd import workloads as wl
    a = wl.uniform_u8((1080, 1920, 3), seed=1, stream=0)
    b = np.clip(a.astype(np.int16) + (wl.uniform_u8(a.shape, seed=1, stream=1) >> 3) - 16, 0, 255).astype(np.uint8)
    want_v, want_m = oracle.ssim(a, b)
    da, db = ipf.DeviceImage.from_array(a), ipf.DeviceImage.from_array(b)
    got_v, dm = ipf.ssim(da, db, return_map=True)
    assert _same_bits(dm.download(), want_m)
    assert abs(got_v - want_v) <= _ssim_tol(a.size, want_m), (got_v, want_v)
    assert ipf.ssim(da, da) == 1.0


def test_ssim_rejects_mismatched_images(ipf):
    a = np.zeros((12, 12), np.uint8)
    with pytest.raises(ipf.IpfError, match="ssim: images must be non-empty and have the same shape"):
        ipf.ssim(a, np.zeros((12, 13), np.uint8))
    with pytest.raises(ipf.IpfError, match="ssim: images must be non-empty and have the same shape"):
        ipf.ssim(np.zeros((0, 4), np.uint8), np.zeros((0, 4), np.uint8))


# ------------------------------------------------------------- one-channel float32 strip kernel
@pytest.mark.parametrize("ks", [3, 5, 7, 13, 15, 19, 25, 27, 33, 49])
def test_float_gray_strip_kernel_edges_vs_oracle(ipf, oracle, ks):
    # widths on both sides of the 128-column strip, heights below and above the
    # window, every border (Constant with a non-zero value), a delta, box sums
    rng = np.random.default_rng(300 + ks)
    k = oracle.gaussian_kernel_1d(0.0, ks)
    shapes = [(1, 1), (3, 127), (37, 128), (64, 129), (ks + 5, 255), (131, 257), (260, 384)]
    for i, shape in enumerate(shapes):
        img = (rng.random(shape) * 4.0 - 1.0).astype(np.float32)
        b = i % 5
        got = ipf.sep_filter2d(img, k, k, b, delta=0.25, border_value=0.75)
        want = oracle.sep_filter(img, k, k, border=b, delta=0.25, bval=0.75)
        assert np.array_equal(got, want), (shape, b)
        got = ipf.box_filter(img, ks, ks, True, b, 0.75)
        want = oracle.box_filter(img, ks, ks, True, b, 0.75)
        err = np.abs(got.astype(np.float64) - want)
        assert np.all(err <= 1e-12 * ks * ks + 2.0 ** -24 * np.abs(want)), (shape, b, err.max())


def test_float_gray_strip_kernel_ties_and_specials(ipf, oracle):
    # dyadic inputs with weights 1/4, 1/2, 1/4 put many results exactly on float
    # rounding ties (recomputed in the reference's order); inf and NaN inputs
    rng = np.random.default_rng(310)
    img = (rng.integers(0, 1 << 20, (300, 333)) * 2.0 ** -20).astype(np.float32)
    k = np.array([0.25, 0.5, 0.25])
    for b in range(5):
        assert np.array_equal(ipf.sep_filter2d(img, k, k, b), oracle.sep_filter(img, k, k, border=b)), b
    spec = rng.random((140, 150)).astype(np.float32)
    spec[10, 20] = np.inf
    spec[70, 140] = -np.inf
    spec[100, 3] = np.nan
    kg = oracle.gaussian_kernel_1d(1.5)
    got = ipf.sep_filter2d(spec, kg, kg, 4)
    want = oracle.sep_filter(spec, kg, kg, border=4)
    assert np.array_equal(got, want, equal_nan=True)


# ------------------------------------------------------ feature callers
def test_full_size_feature_callers_vs_reference(ipf):
    # canny / cannyAuto on the 3840 x 2160 RGB polygon image and the corner
    # responses on its first channel: SHA-256 of the reference's own outputs
    from imageprocessingflex_amd import workloads as wl
    img = wl.polygon_image()
    d = ipf.DeviceImage.from_array(img)
    assert _sha(ipf.canny(d, 50.0, 150.0).download()) == FULL["c3_canny_50_150_u8"]
    assert _sha(ipf.canny_auto(d).download()) == FULL["c3_canny_auto_u8"]
    g = ipf.DeviceImage.from_array(np.ascontiguousarray(img[..., 0]))
    assert _sha(ipf.corner_harris(g, 3, 3, 0.04).download()) == FULL["c3g_harris_u8"]
    assert _sha(ipf.corner_min_eigen_val(g, 3, 3).download()) == FULL["c3g_min_eigen_u8"]


def _spiral(n, gap=4, strong=True):
    """A square spiral ridge (value 60, one pixel wide, turns `gap` apart) on
    0, optionally with a strong 255 block at its outer end: with the L2
    magnitude hysteresis must follow the ridge's edges all the way in."""
    img = np.zeros((n, n), np.uint8)
    x0, y0, x1, y1 = 2, 2, n - 3, n - 3
    while x0 < x1 and y0 < y1:
        img[y0, x0:x1 + 1] = 60
        img[y0:y1 + 1, x1] = 60
        img[y1, x0:x1 + 1] = 60
        img[y0 + gap:y1 + 1, x0] = 60
        x0, y0, x1, y1 = x0 + gap, y0 + gap, x1 - gap, y1 - gap
        if x0 < x1 and y0 < y1:
            img[y0 - gap:y0 + 1, x0] = 60  # the link into the next turn
    if strong:
        img[0:3, 0:3] = 255
    return img


@pytest.mark.parametrize("n", [61, 257, 1030])
def test_canny_hysteresis_long_paths_vs_oracle(ipf, oracle, n):
    # a path of weak candidates thousands of pixels long, crossing many
    # workgroups, must be kept end to end (one strong blob at its end) or
    # dropped end to end (none)
    for strong in (True, False):
        img = _spiral(n, 4, strong)
        for l2 in (0, 1):
            got = ipf.canny(img, 20.0, 400.0, 3, l2)
            want = oracle.canny(img, 20.0, 400.0, 3, l2)
            assert np.array_equal(got, want), (n, strong, l2, int((got != want).sum()))
            if strong and l2:
                assert (got == 255).sum() > n * n // 3  # the spiral's edges, not just the block's


@pytest.mark.parametrize("shape", [(1, 1), (1, 37), (29, 1), (2, 2), (130, 517, 3), (333, 250, 4), (700, 901)],
                         ids=lambda s: "x".join(map(str, s)))
def test_canny_random_images_vs_oracle(ipf, oracle, shape):
    from imageprocessingflex_amd import workloads as wl
    rng = np.random.default_rng(sum(shape))
    img = wl.uniform_u8(shape, seed=71)
    blobs = rng.integers(0, 4, shape).astype(np.uint8) * 60  # coarse structure plus noise
    img = np.clip(img // 4 + blobs, 0, 255).astype(np.uint8)
    for lo, hi, k, l2 in ((20.0, 60.0, 3, 0), (50.0, 150.0, 3, 1), (300.0, 900.0, 5, 0), (0.0, 0.0, 3, 0),
                          (-5.0, 1e9, 3, 1), (2000.0, 6000.0, 7, 1)):
        got = ipf.canny(img, lo, hi, k, l2)
        want = oracle.canny(img, lo, hi, k, l2)
        assert np.array_equal(got, want), (lo, hi, k, l2, int((got != want).sum()))
    f = (img.astype(np.float32) / 255.0)
    for s in (1.4, 0.8, 2.5):
        assert np.array_equal(ipf.canny_auto(img, s), oracle.canny_auto(img, s)), s
        assert np.array_equal(ipf.canny_auto(f, s), oracle.canny_auto(f, s)), s


def test_canny_repeatable_and_in_place(ipf, oracle):
    # the union-find's visiting order varies from run to run; the edge map
    # must not, and an in-place call on the device reads the original image
    from imageprocessingflex_amd import workloads as wl
    img = wl.polygon_image(540, 960, 1)
    want = oracle.canny(img, 30.0, 90.0)
    d = ipf.DeviceImage.from_array(img)
    for _ in range(5):
        assert np.array_equal(ipf.canny(d, 30.0, 90.0).download(), want)
    from imageprocessingflex_amd import _lib
    import ctypes as C
    desc = d.descriptor()
    _lib.check(_lib.load().ipfg_canny(C.byref(desc), C.byref(desc), 30.0, 90.0, 3, 0, None))
    ipf.synchronize()
    assert np.array_equal(d.download(), want)


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.int16, np.float32])
def test_corner_responses_vs_oracle(ipf, oracle, dtype):
    from imageprocessingflex_amd import workloads as wl
    base = wl.polygon_image(300, 411, 1, seed=9).astype(np.float64)
    img = {np.uint8: base, np.uint16: base * 250, np.int16: base * 120 - 15000,
           np.float32: base / 255.0}[dtype].astype(dtype)
    for blk in (1, 2, 3, 5, 9):
        for k in (1, 3, 5, 7):
            assert np.array_equal(ipf.corner_harris(img, blk, k, 0.04), oracle.corner_harris(img, blk, k, 0.04)), (blk, k)
            assert np.array_equal(ipf.corner_min_eigen_val(img, blk, k), oracle.corner_min_eigen(img, blk, k)), (blk, k)


# --------------------------------------------------------------- filter2DFFT
def test_full_c4_filter2d_fft_vs_reference(ipf):
    # 4096 x 4096 uint8 and float32 with the 31 x 31 kernel: 8192 x 8192-point
    # FP64 transforms; SHA-256 of the reference's own outputs
    from imageprocessingflex_amd import workloads as wl
    u8, f32 = wl.c4_images()
    k, _ = wl.c4_kernel()
    assert _sha(ipf.filter2d_fft(ipf.DeviceImage.from_array(u8), k).download()) == FULL["c4_filter2dfft31_reflect101_u8"]
    assert _sha(ipf.filter2d_fft(ipf.DeviceImage.from_array(f32), k).download()) == FULL["c4_filter2dfft31_reflect101_f32"]


@pytest.mark.parametrize("shape", [(1, 1), (1, 300), (300, 1), (2, 2, 3), (37, 23, 4), (129, 257, 3), (600, 70),
                                   (70, 1100)], ids=lambda s: "x".join(map(str, s)))
def test_filter2d_fft_vs_oracle(ipf, oracle, shape):
    # one- and two-pass axes (up to 2^11 points), rows shorter than a tile,
    # every border, odd and even kernel sides, every output type
    from imageprocessingflex_amd import workloads as wl
    rng = np.random.default_rng(sum(shape))
    u8 = wl.uniform_u8(shape, seed=12)
    f = wl.uniform_f32(shape, seed=13) * 8 - 4
    for ks in ((3, 3), (1, 7), (6, 4), (17, 15)):
        k = rng.standard_normal(ks).astype(np.float32)
        for b in range(5):
            for img, to in ((u8, None), (u8, np.float32), (f, None), (f, np.float64), (u8, np.int16)):
                got = ipf.filter2d_fft(img, k, b, 0.5, 3.0, to)
                want = oracle.filter2d_fft(img, k, b, 0.5, 3.0, to)
                assert np.array_equal(got, want), (ks, b, img.dtype, to)


def test_filter2d_fft_device_strided_and_repeat(ipf, oracle):
    from imageprocessingflex_amd import workloads as wl
    img = wl.uniform_u8((260, 300, 3), seed=14)
    view = img[7:250, 11:290]
    k = np.arange(1, 10, dtype=np.float32).reshape(3, 3) / 45
    want = oracle.filter2d_fft(np.ascontiguousarray(view), k, 4)
    assert np.array_equal(ipf.filter2d_fft(view, k, 4), want)
    d = ipf.DeviceImage.from_array(np.ascontiguousarray(view))
    s = ipf.Stream()
    outs = [ipf.filter2d_fft(d, k, 4, stream=s) for _ in range(3)]
    s.synchronize()
    for o in outs:
        assert np.array_equal(o.download(), want)


# ------------------------------------------------------------ integralImage
def test_full_c2_integral_vs_reference(ipf):
    # 7680 x 4320 float32: both tables (sums and squared sums) bit for bit
    from imageprocessingflex_amd import workloads as wl
    d = ipf.DeviceImage.from_array(wl.sinusoid_image())
    assert _sha(ipf.integral_image(d).download()) == FULL["c2_integral_f32"]
    assert _sha(ipf.integral_image(d, True).download()) == FULL["c2_integral_sq_f32"]


def test_integral_image_types_and_sums(ipf, oracle):
    from imageprocessingflex_amd import workloads as wl
    for shape in ((1, 1), (3, 700, 4), (513, 2), (129, 257, 3)):
        for img in (wl.uniform_u8(shape, seed=21), wl.uniform_f32(shape, seed=22) * 2 - 1,
                    (wl.uniform_f32(shape, seed=23) * 4e9 - 2e9).astype(np.int32)):
            for sq in (False, True):
                got = ipf.integral_image(img, sq)
                assert np.array_equal(got, oracle.integral_image(img, sq)), (shape, img.dtype, sq)
    img = wl.uniform_u8((200, 300, 3), seed=24)
    t = ipf.integral_image(img[5:150, 7:260])  # a strided view
    assert np.array_equal(t, oracle.integral_image(np.ascontiguousarray(img[5:150, 7:260])))
    assert ipf.integral_sum(t, 10, 20, 30, 40, 2) == img[5:150, 7:260][20:60, 10:40, 2].astype(np.int64).sum()


# -------------------------------------------------------------- houghCircles
def test_full_hough_circles_1080p_vs_reference(ipf):
    # 1920 x 1080 disc scene: the reference's circles, bit for bit
    from imageprocessingflex_amd import workloads as wl
    got = ipf.hough_circles(wl.circles_scene(1080, 1920, 24, 7, 3.0), 20, 120, 100.0, 30, 10.0, 0.6)
    assert len(got) >= 10
    assert _sha(got) == FULL["hough_1080p_circles"]


@pytest.mark.parametrize("args", [(8, 60, 100.0, 20, 10.0, 0.6), (15, 40, 80.0, 12, 8.0, 0.5),
                                  (30, 90, 120.0, 40, 20.0, 0.7), (8, 60, 100.0, 20, 1.5, 0.6),
                                  (15, 40, 80.0, 25, 0.0, 0.5)], ids=str)
def test_hough_circles_vs_oracle(ipf, oracle, args):
    from imageprocessingflex_amd import workloads as wl
    for img in (wl.circles_scene(360, 480, 8, 11, 6.0), wl.circles_scene(301, 257, 5, 12, 14.0)):
        got = ipf.hough_circles(img, *args)
        want = oracle.hough_circles(img, *args)
        assert np.array_equal(got, want), (args, img.shape, got[:3], want[:3])
    # device-resident input, float pixels
    f = (wl.circles_scene(200, 240, 4, 13, 5.0).astype(np.float32) / 255).astype(np.float32)
    a = list(args)
    a[2] /= 255.0
    got = ipf.hough_circles(ipf.DeviceImage.from_array(f), *a)
    assert np.array_equal(got, oracle.hough_circles(f, *a))


# ------------------------------------------------------------ Netpbm / PFM ingest
def test_decode_pnm_full_size_vs_restatement(ipf):
    # 3840 x 2160 RGB files in every binary layout, decoded on the device and
    # on the host restatement of the reference's decoder
    from imageprocessingflex_amd import workloads as wl
    from oracle import pyoracle as po
    img = wl.polygon_image()
    g16 = (img[..., 1].astype(np.uint16) * 257 - 3).astype(np.uint16)
    for b, dt in ((po.pnm_encode(img), np.uint8), (po.pnm_encode(img), np.uint16),
                  (po.pnm_encode(g16), np.uint16), (po.pnm_encode(g16), np.uint8),
                  (b"P5 3840 2160 1000\n" + (g16 // 66).astype(">u2").tobytes(), np.uint8)):
        want = po.pnm_decode(b, dt)
        assert np.array_equal(ipf.decode_pnm(b, dt), want)
        d = ipf.decode_pnm(b, dt, device=True)
        assert np.array_equal(d.download(), want)
    f = (img.astype(np.float32) / 255).astype(np.float32)
    for b in (po.pfm_encode(f), b"PF\n3840 2160\n1.0\n" + f[::-1].astype(">f4").tobytes()):
        assert np.array_equal(ipf.decode_pnm(b), po.pfm_decode(b))
    bits = np.packbits(img[..., 0] > 128, axis=1)
    b = b"P4\n3840 2160\n" + bits.tobytes()
    assert np.array_equal(ipf.decode_pnm(b), po.pnm_decode(b))
