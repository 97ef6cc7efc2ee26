"""CPU checks of the C ABI: libipfg.so loads, exports every function that
include/ipfg.h declares, and rejects bad arguments with the reference's
messages before touching a device. No kernel runs here."""
from __future__ import annotations

import ctypes as C
import json
import re
from pathlib import Path

import numpy as np
import pytest

import imageprocessingflex_amd as ipf
from tests import fixtures

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "ipfg.h"


def declared_functions():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(ipfg_\w+)\s*\(", text)))


def test_header_declares_expected_entry_points():
    names = declared_functions()
    for n in ("ipfg_sep_filter", "ipfg_gaussian_blur", "ipfg_box_filter", "ipfg_filter2d", "ipfg_median",
              "ipfg_bilateral", "ipfg_init", "ipfg_last_error"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from imageprocessingflex_amd import _lib
    lib = C.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding covers exactly the header
    assert sorted(_lib.SIGNATURES) == declared_functions()
    _lib.load()


def test_gaussian_kernel_host_matches_oracle(oracle):
    import imageprocessingflex_amd as ipf
    for s, k in ((0.5, 0), (1.0, 0), (2.0, 0), (8.0, 0), (0.0, 7), (1.3, 9)):
        assert np.array_equal(ipf.gaussian_kernel_1d(s, k), oracle.gaussian_kernel_1d(s, k))
    with pytest.raises(ipf.IpfError, match="ipf: gaussianKernel1D: ksize must be odd"):
        ipf.gaussian_kernel_1d(1.0, 4)
    with pytest.raises(ipf.IpfError, match="ipf: gaussianKernel1D: need sigma > 0 or ksize > 0"):
        ipf.gaussian_kernel_1d(0.0, 0)


def test_motion_blur_kernel_known_answers():
    kat = json.loads((fixtures.GOLDEN / "kat.json").read_text())["motion"]
    for n, a, want in kat:
        got = ipf.motion_blur_kernel(n, a)
        assert got.dtype == np.float32 and np.array_equal(got.ravel(), np.array(want, np.float32)), (n, a)
    with pytest.raises(ipf.IpfError, match="length must be >= 1"):
        ipf.motion_blur_kernel(0, 10.0)


def _raw_pyr(name, src, dst):
    """A pyramid step through the C ABI with a caller-chosen destination."""
    import ctypes as C

    import imageprocessingflex_amd as ipf
    from imageprocessingflex_amd import _lib
    s, d = ipf._desc(src), ipf._desc(dst)
    _lib.check(getattr(_lib.load(), name)(C.byref(s), C.byref(d), None))


def test_argument_errors_precede_device_work():
    import imageprocessingflex_amd as ipf
    img = np.zeros((4, 5), np.uint8)
    cases = [
        (lambda: ipf.median_blur(img, 3, ipf.BorderMode.CONSTANT), "ipf: medianBlur: Constant border is not supported"),
        (lambda: ipf.median_blur(img, 4), "ipf: medianBlur: ksize must be odd and positive"),
        (lambda: ipf.median_blur(img, -1), "ipf: medianBlur: ksize must be odd and positive"),
        (lambda: ipf.bilateral_filter(img, 5, 0.0, 3.0), "ipf: bilateralFilter: sigmas must be positive"),
        (lambda: ipf.bilateral_filter(img, 5, 3.0, -1.0), "ipf: bilateralFilter: sigmas must be positive"),
        (lambda: ipf.box_filter(img, 0, 3), "ipf: boxFilter: kernel size must be positive"),
        (lambda: ipf.sep_filter2d(img, [], [1.0]), "ipf: separable filter: empty kernel"),
        (lambda: ipf.filter2d(img, np.ones((3, 3)), anchor=(3, 1)), "ipf: filter2D: anchor outside kernel"),
        (lambda: ipf.filter2d(img, np.ones((0, 3))), "ipf: filter2D: kernel must be a non-empty single-channel image"),
        (lambda: ipf.gaussian_blur(img, 0.0), "ipf: gaussianKernel1D: need sigma > 0 or ksize > 0"),
        (lambda: ipf.gaussian_blur(img, 1.0, ksize=4), "ipf: gaussianKernel1D: ksize must be odd"),
        (lambda: ipf.sobel(img, 0, 0), "ipf: sobel: need dx + dy > 0"),
        (lambda: ipf.sobel(img, 1, 0, 33), "ipf: derivKernel: ksize must be odd in [3, 31]"),
        (lambda: ipf.sobel(img, 0, 3, 1), "ipf: derivKernel: order > 2 requires ksize >= 3"),
        (lambda: ipf.sobel(img, 3, 0, 3), "ipf: derivKernel: invalid derivative order"),
        (lambda: ipf.scharr(img, 1, 1), "ipf: scharr: (dx, dy) must be (1,0) or (0,1)"),
        (lambda: ipf.prewitt(img, 2, 0), "ipf: prewitt: (dx, dy) must be (1,0) or (0,1)"),
        (lambda: ipf.laplacian(img, 2), "ipf: derivKernel: ksize must be odd in [3, 31]"),
        (lambda: ipf.magnitude(np.zeros((2, 3)), np.zeros((3, 2))), "ipf: magnitude: shape mismatch"),
        (lambda: ipf.phase(np.zeros((2, 3)), np.zeros((2, 3, 2))), "ipf: phase: shape mismatch"),
        (lambda: ipf.adaptive_threshold(np.zeros((4, 4, 3), np.uint8), 255, 0, 0, 3, 0.0),
         "ipf: adaptiveThreshold: expected a single-channel image"),
        (lambda: ipf.adaptive_threshold(img, 255, 0, 0, 1, 0.0), "ipf: adaptiveThreshold: blockSize must be odd and >= 3"),
        (lambda: ipf.adaptive_threshold(img, 255, 1, 2, 5, 0.0), "ipf: adaptiveThreshold: type must be Binary or BinaryInv"),
        (lambda: ipf.build_pyramid(img, 0), "ipf: buildPyramid: levels must be >= 1"),
        (lambda: _raw_pyr("ipfg_pyr_down", img, np.zeros((3, 3), np.uint8)),
         "ipfg: pyrDown: dst must be (w + 1) / 2 x (h + 1) / 2"),
        (lambda: _raw_pyr("ipfg_pyr_up", img, np.zeros((8, 10), np.uint16)),
         "ipfg: pyrUp: dst must have the pixel type of src"),
        (lambda: _raw_pyr("ipfg_pyr_up", img, np.zeros((8, 10, 2), np.uint8)), "ipfg: pyrUp: dst must be 2w x 2h"),
        (lambda: ipf.canny(img, 5.0, 3.0), "ipf: canny: lowThreshold must be <= highThreshold"),
        (lambda: ipf.canny(img, 5.0, 3.0, 4), "ipf: canny: lowThreshold must be <= highThreshold"),
        (lambda: ipf.canny(img, 1.0, 3.0, 4), "ipf: derivKernel: ksize must be odd in [3, 31]"),
        (lambda: ipf.canny_auto(img, 0.0), "ipf: gaussianKernel1D: need sigma > 0 or ksize > 0"),
        (lambda: ipf.canny_auto(np.zeros((4, 5, 5), np.uint8), 0.0),
         "ipf: rgbToGray: expected a 3-channel (RGB) or 4-channel (RGBA) image"),
        (lambda: ipf.corner_harris(np.zeros((4, 5, 3), np.uint8), 0, 4),
         "ipf: corner detection: expected a single-channel image"),
        (lambda: ipf.corner_harris(img, 0, 4), "ipf: derivKernel: ksize must be odd in [3, 31]"),
        (lambda: ipf.corner_min_eigen_val(img, 0, 3), "ipf: boxFilter: kernel size must be positive"),
        (lambda: ipf.corner_min_eigen_val(img, -2, 5), "ipf: boxFilter: kernel size must be positive"),
    ]
    for fn, msg in cases:
        with pytest.raises(ipf.IpfError) as e:
            fn()
        assert str(e.value) == msg


def test_reference_check_order():
    # medianBlur checks ksize before the border (filter.hpp:304-305)
    import imageprocessingflex_amd as ipf
    with pytest.raises(ipf.IpfError, match="ksize must be odd"):
        ipf.median_blur(np.zeros((3, 3), np.uint8), 2, ipf.BorderMode.CONSTANT)


def test_tuning_table_round_trip():
    # the table functions are host-only: no device is touched
    import imageprocessingflex_amd as ipf
    ipf.tune_clear()
    assert ipf.tune_export() == ""
    text = "sep 0 0 3 13 13 1 11 11 64\nsep 4 4 1 49 49 1 13 13 128\n"
    assert ipf.tune_import(text) == 2
    got = ipf.tune_export()
    assert sorted(got.splitlines()) == sorted(text.splitlines())
    assert ipf.tune_import("# comment\n\nsep 0 0 3 13 13 1 11 11 96\n") == 1
    assert "sep 0 0 3 13 13 1 11 11 96" in ipf.tune_export()
    for bad in ("sep 0 0 3 13 13 1 11 11\n", "sep 0 0 3 13 13 1 11 11 0\n", "sep a b c\n"):
        with pytest.raises((ipf.IpfError, ipf.DeviceError), match="malformed tuning table"):
            ipf.tune_import(bad)
    ipf.tune_clear()
    assert ipf.tune_export() == ""


def test_pnm_encoders_match_reference_bytes():
    # encodePNM / encodePFM run on the host: the package's bytes equal the
    # reference encoder's for every fixture image
    n = 0
    for case in fixtures.load_cases():
        if case[0]["op"] in ("pnm_enc", "pfm_enc"):
            got = fixtures.run_gpu(ipf, case)
            assert np.array_equal(got, case[2]), fixtures.case_id(case)
            n += 1
    assert n >= 10


def test_pnm_header_errors_precede_device_work():
    cases = [
        (b"Q5 1 1 255\n\x00", "ipf: PNM: not a P1-P6 netpbm file"),
        (b"P5 0 1 255\n", "ipf: PNM: invalid image dimensions"),
        (b"P5 2 x", "ipf: PNM: malformed header or pixel data"),
        (b"P5 99999999999 1 255\n", "ipf: PNM: number too large"),
        (b"P5 2 1 70000\n\x00\x00", "ipf: PNM: invalid maxval"),
        (b"P5 2 1 255\n\x00", "ipf: PNM: unexpected end of data (truncated file?)"),
        (b"P5 2 1 255", "ipf: PNM: malformed header"),
        (b"P1 2 1 02", "ipf: PNM: malformed P1 pixel data"),
        (b"Pf 2 1 abc\n", "ipf: PFM: invalid scale value"),
        (b"Pf 2 1 0\n", "ipf: PFM: invalid scale value"),
        (b"PF 2 1 -1.0\n\x00\x00", "ipf: PFM: unexpected end of data (truncated file?)"),
        (b"PF 20000000 1 -1.0\n", "ipf: PFM: image dimensions too large"),
    ]
    for b, msg in cases:
        with pytest.raises(ipf.IpfError) as e:
            ipf.decode_pnm(b)
        assert str(e.value) == msg, (b, str(e.value))
