"""TEST INFRASTRUCTURE -- NOT PART OF THE PRODUCT.

numpy front end for the CPU oracle (oracle/libipf_oracle.so, the restatement
of ipf's filtering path) and, where it was built, the reference-built checker
(oracle/_ref/libipf_ref.so). Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py use this module.

    o = Oracle("oracle")      # the restatement
    r = Oracle("ref")         # compiled from the reference headers
    out = o.gaussian_blur(img, 2.0)
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ORACLE_SO = HERE / "libipf_oracle.so"
REF_SO = HERE / "_ref" / "libipf_ref.so"

U8, U16, I16, I32, F32, F64 = range(6)
CONSTANT, REPLICATE, REFLECT, REFLECT101, WRAP = range(5)
BORDERS = (CONSTANT, REPLICATE, REFLECT, REFLECT101, WRAP)

NP2DT = {np.dtype(np.uint8): U8, np.dtype(np.uint16): U16, np.dtype(np.int16): I16,
         np.dtype(np.int32): I32, np.dtype(np.float32): F32, np.dtype(np.float64): F64}
DT2NP = {v: k for k, v in NP2DT.items()}


class OracleError(RuntimeError):
    pass


def build(ref: bool = True) -> None:
    """Compiles the oracle (and the reference checker when the reference tree exists)."""
    target = "all" if ref else "oracle"
    subprocess.run(["make", "-s", "-C", str(HERE), target], check=True)


def ref_available() -> bool:
    return REF_SO.exists()


_D, _I, _V = C.c_double, C.c_int, C.c_void_p
_PD = C.POINTER(C.c_double)
_PF = C.POINTER(C.c_float)


class Oracle:
    def __init__(self, which: str = "oracle"):
        if which == "oracle":
            path, pre = ORACLE_SO, "ipfo_"
        elif which == "ref":
            path, pre = REF_SO, "ipfr_"
        else:
            raise ValueError(which)
        if not path.exists():
            raise OracleError(f"{path} not built (make -C oracle)")
        self.which = which
        self.lib = C.CDLL(str(path))
        self.p = pre
        sig = {
            "last_error": (C.c_char_p, []),
            "set_num_threads": (None, [_I]),
            "get_num_threads": (_I, []),
            "border_index": (_I, [_I, _I, _I]),
            "saturate": (_I, [_D, _I, _V]),
            "gaussian_kernel_1d": (_I, [_D, _I, _PD, _I]),
            "sep_filter": (_I, [_V, _I, _I, _I, _I, _V, _I, _PD, _I, _PD, _I, _I, _I, _I, _D, _D]),
            "box_filter": (_I, [_V, _I, _I, _I, _I, _V, _I, _I, _I, _I, _I, _D]),
            "filter2d": (_I, [_V, _I, _I, _I, _I, _V, _I, _PF, _I, _I, _I, _I, _I, _D, _D]),
            "hough_circles": (_I, [_V, _I, _I, _I, _I, _I, _I, _D, _I, _D, _D, _PD, _I, C.POINTER(_I)]),
            "integral_image": (_I, [_V, _I, _I, _I, _I, _V, _I]),
            "filter2d_fft": (_I, [_V, _I, _I, _I, _I, _V, _I, _PF, _I, _I, _I, _D, _D]),
            "median": (_I, [_V, _I, _I, _I, _I, _V, _I, _I]),
            "bilateral": (_I, [_V, _I, _I, _I, _I, _V, _I, _D, _D, _I]),
            "sobel": (_I, [_V, _I, _I, _I, _I, _V, _I, _I, _I, _D, _I]),
            "scharr": (_I, [_V, _I, _I, _I, _I, _V, _I, _I, _D, _I]),
            "prewitt": (_I, [_V, _I, _I, _I, _I, _V, _I, _I, _I]),
            "laplacian": (_I, [_V, _I, _I, _I, _I, _V, _I, _D, _I]),
            "magnitude": (_I, [_V, _V, _I, _I, _I, _V]),
            "phase": (_I, [_V, _V, _I, _I, _I, _V, _I]),
            "unsharp": (_I, [_V, _I, _I, _I, _I, _V, _D, _D, _D]),
            "sharpen": (_I, [_V, _I, _I, _I, _I, _V, _D]),
            "emboss": (_I, [_V, _I, _I, _I, _I, _V]),
            "pyr_down": (_I, [_V, _I, _I, _I, _I, _V]),
            "pyr_up": (_I, [_V, _I, _I, _I, _I, _V]),
            "ssim": (_I, [_V, _V, _I, _I, _I, _I, _V, C.POINTER(_D)]),
            "adaptive_threshold": (_I, [_V, _I, _I, _I, _V, _D, _I, _I, _I, _D]),
            "motion_kernel": (_I, [_I, _D, _PF, _I]),
            "canny": (_I, [_V, _I, _I, _I, _I, _V, _D, _D, _I, _I]),
            "canny_auto": (_I, [_V, _I, _I, _I, _I, _V, _D]),
            "corner_harris": (_I, [_V, _I, _I, _I, _I, _V, _I, _I, _D]),
            "corner_min_eigen": (_I, [_V, _I, _I, _I, _I, _V, _I, _I]),
        }
        for name, (res, args) in sig.items():
            f = getattr(self.lib, pre + name)
            f.restype, f.argtypes = res, args
        if which == "ref":
            self.lib.ipfr_last_seconds.restype = C.c_double
            self.lib.ipfr_last_seconds.argtypes = []

    def _f(self, name):
        return getattr(self.lib, self.p + name)

    def _check(self, rc):
        if rc != 0:
            raise OracleError(self._f("last_error")().decode())

    def last_seconds(self) -> float:
        """Reference checker only: duration of the last ipf call, excluding copies."""
        return self.lib.ipfr_last_seconds()

    def set_num_threads(self, n: int) -> None:
        self._f("set_num_threads")(n)

    def get_num_threads(self) -> int:
        return self._f("get_num_threads")()

    def border_index(self, p, n, mode):
        return self._f("border_index")(p, n, mode)

    def saturate(self, v: float, dtype) -> object:
        dt = NP2DT[np.dtype(dtype)]
        out = np.zeros(1, DT2NP[dt])
        self._check(self._f("saturate")(float(v), dt, out.ctypes.data))
        return out[0]

    def gaussian_kernel_1d(self, sigma, ksize=0):
        n = self._f("gaussian_kernel_1d")(float(sigma), int(ksize), None, 0)
        if n < 0:
            self._check(n)
        out = np.empty(n, np.float64)
        self._f("gaussian_kernel_1d")(float(sigma), int(ksize), out.ctypes.data_as(_PD), n)
        return out

    @staticmethod
    def _shape(a):
        a = np.ascontiguousarray(a)
        if a.ndim == 2:
            return a, a.shape[0], a.shape[1], 1
        return a, a.shape[0], a.shape[1], a.shape[2]

    def sep_filter(self, src, kx, ky, anchor=None, border=REFLECT101, delta=0.0, bval=0.0, out_dtype=None):
        src, h, w, cn = self._shape(src)
        kx = np.ascontiguousarray(kx, np.float64)
        ky = np.ascontiguousarray(ky, np.float64)
        ax, ay = anchor if anchor is not None else (len(kx) // 2, len(ky) // 2)
        out = np.empty(src.shape, out_dtype or src.dtype)
        self._check(self._f("sep_filter")(src.ctypes.data, w, h, cn, NP2DT[src.dtype], out.ctypes.data,
                                          NP2DT[out.dtype], kx.ctypes.data_as(_PD), len(kx),
                                          ky.ctypes.data_as(_PD), len(ky), ax, ay, border, delta, bval))
        return out

    def gaussian_blur(self, src, sigma_x, sigma_y=0.0, ksize=0, border=REFLECT101, out_dtype=None):
        if sigma_y <= 0:
            sigma_y = sigma_x
        kx = self.gaussian_kernel_1d(sigma_x, ksize)
        ky = self.gaussian_kernel_1d(sigma_y, ksize)
        return self.sep_filter(src, kx, ky, None, border, 0.0, 0.0, out_dtype)

    def box_filter(self, src, kw, kh, normalize=True, border=REFLECT101, bval=0.0, out_dtype=None):
        src, h, w, cn = self._shape(src)
        out = np.empty(src.shape, out_dtype or src.dtype)
        self._check(self._f("box_filter")(src.ctypes.data, w, h, cn, NP2DT[src.dtype], out.ctypes.data,
                                          NP2DT[out.dtype], kw, kh, int(bool(normalize)), border, bval))
        return out

    def filter2d(self, src, kernel, border=REFLECT101, delta=0.0, anchor=(-1, -1), bval=0.0, out_dtype=None):
        src, h, w, cn = self._shape(src)
        k = np.ascontiguousarray(kernel, np.float32)
        kh, kw = k.shape
        out = np.empty(src.shape, out_dtype or src.dtype)
        self._check(self._f("filter2d")(src.ctypes.data, w, h, cn, NP2DT[src.dtype], out.ctypes.data,
                                        NP2DT[out.dtype], k.ctypes.data_as(_PF), kw, kh, anchor[0], anchor[1],
                                        border, delta, bval))
        return out

    def hough_circles(self, src, min_r, max_r, canny_high=100.0, center_votes=20, min_dist=10.0, min_coverage=0.6):
        """houghCircles (features.hpp:388-504): (n, 4) float64 rows (cx, cy, radius, votes)."""
        src, h, w, cn = self._shape(src)
        cap = 256
        while True:
            out = np.zeros((cap, 4), np.float64)
            n = C.c_int(0)
            self._check(self._f("hough_circles")(src.ctypes.data, w, h, cn, NP2DT[src.dtype], int(min_r), int(max_r),
                                                 float(canny_high), int(center_votes), float(min_dist),
                                                 float(min_coverage), out.ctypes.data_as(_PD), cap, C.byref(n)))
            if n.value <= cap:
                return out[:n.value].copy()
            cap = n.value

    def integral_image(self, src, squared=False):
        """integralImage (filter.hpp:522-538): float64 (h + 1, w + 1[, cn]) table."""
        a = np.ascontiguousarray(src)
        src, h, w, cn = self._shape(src)
        out = np.empty((h + 1, w + 1) if a.ndim == 2 else (h + 1, w + 1, cn), np.float64)
        self._check(self._f("integral_image")(src.ctypes.data, w, h, cn, NP2DT[src.dtype], out.ctypes.data,
                                              int(bool(squared))))
        return out

    def filter2d_fft(self, src, kernel, border=REFLECT101, delta=0.0, bval=0.0, out_dtype=None):
        """filter2DFFT (fft.hpp:196-216): filter2D through FP64 radix-2 FFTs; anchor at the centre."""
        src, h, w, cn = self._shape(src)
        k = np.ascontiguousarray(kernel, np.float32)
        kh, kw = k.shape
        out = np.empty(src.shape, out_dtype or src.dtype)
        self._check(self._f("filter2d_fft")(src.ctypes.data, w, h, cn, NP2DT[src.dtype], out.ctypes.data,
                                            NP2DT[out.dtype], k.ctypes.data_as(_PF), kw, kh, border, delta, bval))
        return out

    def median(self, src, ksize, border=REPLICATE):
        src, h, w, cn = self._shape(src)
        out = np.empty_like(src)
        self._check(self._f("median")(src.ctypes.data, w, h, cn, NP2DT[src.dtype], out.ctypes.data, ksize, border))
        return out

    # ---- derivative and sharpening filters (filter.hpp:388-567)
    def _float_op(self, name, src, *args):
        src, h, w, cn = self._shape(src)
        out = np.empty(src.shape, np.float32)
        self._check(self._f(name)(src.ctypes.data, w, h, cn, NP2DT[src.dtype], out.ctypes.data, *args))
        return out

    def sobel(self, src, dx, dy, ksize=3, scale=1.0, border=REFLECT101):
        return self._float_op("sobel", src, dx, dy, ksize, float(scale), border)

    def scharr(self, src, dx, dy, scale=1.0, border=REFLECT101):
        return self._float_op("scharr", src, dx, dy, float(scale), border)

    def prewitt(self, src, dx, dy, border=REFLECT101):
        return self._float_op("prewitt", src, dx, dy, border)

    def laplacian(self, src, ksize=1, scale=1.0, border=REFLECT101):
        return self._float_op("laplacian", src, ksize, float(scale), border)

    def magnitude(self, gx, gy):
        gx, h, w, cn = self._shape(np.asarray(gx, np.float32))
        gy = np.ascontiguousarray(gy, np.float32)
        out = np.empty(gx.shape, np.float32)
        self._check(self._f("magnitude")(gx.ctypes.data, gy.ctypes.data, w, h, cn, out.ctypes.data))
        return out

    def phase(self, gx, gy, degrees=True):
        gx, h, w, cn = self._shape(np.asarray(gx, np.float32))
        gy = np.ascontiguousarray(gy, np.float32)
        out = np.empty(gx.shape, np.float32)
        self._check(self._f("phase")(gx.ctypes.data, gy.ctypes.data, w, h, cn, out.ctypes.data, int(bool(degrees))))
        return out

    def _same_op(self, name, src, *args):
        src, h, w, cn = self._shape(src)
        out = np.empty_like(src)
        self._check(self._f(name)(src.ctypes.data, w, h, cn, NP2DT[src.dtype], out.ctypes.data, *args))
        return out

    def unsharp_mask(self, src, sigma=1.0, amount=1.0, threshold=0.0):
        return self._same_op("unsharp", src, float(sigma), float(amount), float(threshold))

    def sharpen(self, src, strength=1.0):
        return self._same_op("sharpen", src, float(strength))

    def emboss(self, src):
        return self._same_op("emboss", src)

    def _pyr(self, name, src, up):
        src = np.ascontiguousarray(src)
        h, w = src.shape[:2]
        cn = src.shape[2] if src.ndim == 3 else 1
        oh, ow = (2 * h, 2 * w) if up else ((h + 1) // 2, (w + 1) // 2)
        out = np.empty((oh, ow) + src.shape[2:], src.dtype)
        self._check(self._f(name)(src.ctypes.data, w, h, cn, NP2DT[src.dtype], out.ctypes.data))
        return out

    def pyr_down(self, src):
        """pyrDown (geometry.hpp:501-512)."""
        return self._pyr("pyr_down", src, False)

    def pyr_up(self, src):
        """pyrUp (geometry.hpp:514-524)."""
        return self._pyr("pyr_up", src, True)

    def ssim(self, a, b):
        """ssim (metrics.hpp:46-78): (mean over the valid region, float32 map)."""
        a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
        if a.shape != b.shape or a.dtype != b.dtype:
            raise OracleError("ssim: images must have the same shape and type")
        h, w = a.shape[:2]
        cn = a.shape[2] if a.ndim == 3 else 1
        m = np.empty(a.shape, np.float32)
        r = C.c_double(0.0)
        self._check(self._f("ssim")(a.ctypes.data, b.ctypes.data, w, h, cn, NP2DT[a.dtype], m.ctypes.data,
                                    C.byref(r)))
        return r.value, m

    def adaptive_threshold(self, src, maxval, method, ttype, block, c):
        src = np.ascontiguousarray(src)
        h, w = src.shape[:2]
        out = np.empty_like(src)
        self._check(self._f("adaptive_threshold")(src.ctypes.data, w, h, NP2DT[src.dtype], out.ctypes.data,
                                                  float(maxval), int(method), int(ttype), int(block), float(c)))
        return out

    def motion_blur_kernel(self, length, angle):
        n = self._f("motion_kernel")(int(length), float(angle), None, 0)
        if n < 0:
            self._check(n)
        out = np.empty((n, n), np.float32)
        self._f("motion_kernel")(int(length), float(angle), out.ctypes.data_as(_PF), n * n)
        return out

    def bilateral(self, src, d, sigma_color, sigma_space, border=REFLECT101):
        src, h, w, cn = self._shape(src)
        out = np.empty_like(src)
        self._check(self._f("bilateral")(src.ctypes.data, w, h, cn, NP2DT[src.dtype], out.ctypes.data, d,
                                         sigma_color, sigma_space, border))
        return out

    # ---- feature callers of the filters (features.hpp:26-160)
    def canny(self, src, low, high, ksize=3, l2=False):
        """canny (features.hpp:26-90): uint8 edge map (255 / 0), one channel."""
        src, h, w, cn = self._shape(src)
        out = np.empty((h, w), np.uint8)
        self._check(self._f("canny")(src.ctypes.data, w, h, cn, NP2DT[src.dtype], out.ctypes.data, float(low),
                                     float(high), int(ksize), int(bool(l2))))
        return out

    def canny_auto(self, src, sigma=1.4):
        """cannyAuto (features.hpp:92-101)."""
        src, h, w, cn = self._shape(src)
        out = np.empty((h, w), np.uint8)
        self._check(self._f("canny_auto")(src.ctypes.data, w, h, cn, NP2DT[src.dtype], out.ctypes.data,
                                          float(sigma)))
        return out

    def corner_harris(self, src, block=3, ksize=3, k=0.04):
        """cornerHarris (features.hpp:128-139): float32 response."""
        src, h, w, cn = self._shape(src)
        out = np.empty((h, w), np.float32)
        self._check(self._f("corner_harris")(src.ctypes.data, w, h, cn, NP2DT[src.dtype], out.ctypes.data,
                                             int(block), int(ksize), float(k)))
        return out

    def corner_min_eigen(self, src, block=3, ksize=3):
        """cornerMinEigenVal (features.hpp:141-153): float32 response."""
        src, h, w, cn = self._shape(src)
        out = np.empty((h, w), np.float32)
        self._check(self._f("corner_min_eigen")(src.ctypes.data, w, h, cn, NP2DT[src.dtype], out.ctypes.data,
                                                int(block), int(ksize)))
        return out


# ------------------------------------------------------------------ Netpbm / PFM codecs
# Test infrastructure: a numpy restatement of the reference's codec/pnm.hpp
# (decodePNM 58-107, encodePNM 109-141, decodePFM 143-176, encodePFM 178-197),
# pinned in tests/test_oracle.py against bytes and images the reference itself
# produced (tests/golden), and the reference-built codecs (RefCodec) that
# wrote them.
class PnmError(ValueError):
    pass


def _pnm_fail(msg):
    raise PnmError("ipf: " + msg)


class _Tok:
    """The Netpbm header tokenizer: '#' comments to the end of the line, any
    whitespace between tokens (pnm.hpp:16-55)."""

    def __init__(self, b: bytes, pos: int):
        self.b, self.pos = b, pos

    def skip(self):
        b = self.b
        while self.pos < len(b):
            if b[self.pos] == 0x23:  # '#'
                while self.pos < len(b) and b[self.pos] not in (0x0A, 0x0D):
                    self.pos += 1
            elif chr(b[self.pos]).isspace():
                self.pos += 1
            else:
                break

    def number(self):
        self.skip()
        b = self.b
        if self.pos >= len(b) or not chr(b[self.pos]).isdigit():
            _pnm_fail("PNM: malformed header or pixel data")
        v = 0
        while self.pos < len(b) and chr(b[self.pos]).isdigit():
            v = v * 10 + b[self.pos] - 48
            self.pos += 1
            if v > (1 << 31):
                _pnm_fail("PNM: number too large")
        return v

    def bit(self):
        self.skip()
        if self.pos >= len(self.b) or self.b[self.pos] not in (0x30, 0x31):
            _pnm_fail("PNM: malformed P1 pixel data")
        self.pos += 1
        return self.b[self.pos - 1] - 48

    def word(self):
        self.skip()
        s = self.pos
        while self.pos < len(self.b) and not chr(self.b[self.pos]).isspace():
            self.pos += 1
        return self.b[s:self.pos].decode("latin-1")


def _check_size(w, h, fmt):
    if w <= 0 or h <= 0:
        _pnm_fail(f"{fmt}: invalid image dimensions")
    if w > (1 << 24) or h > (1 << 24) or w * h > (1 << 30):
        _pnm_fail(f"{fmt}: image dimensions too large")


def pnm_decode(b: bytes, dtype=np.uint8) -> np.ndarray:
    """decodePNM<T>: (h, w) or (h, w, 3), samples rescaled from maxval to T's range."""
    b = bytes(b)
    if len(b) < 3 or b[0] != 0x50 or not (0x31 <= b[1] <= 0x36):
        _pnm_fail("PNM: not a P1-P6 netpbm file")
    kind = b[1] - 48
    t = _Tok(b, 2)
    w, h = t.number(), t.number()
    _check_size(w, h, "PNM")
    bitmap = kind in (1, 4)
    maxval = 1 if bitmap else t.number()
    if maxval < 1 or maxval > 65535:
        _pnm_fail("PNM: invalid maxval")
    cn = 3 if kind in (3, 6) else 1
    n = w * h * cn
    tmax = float(np.iinfo(dtype).max)

    def scale(v):
        v = np.minimum(np.asarray(v, np.int64), maxval).astype(np.float64)
        return ((v * tmax) / float(maxval) + 0.5).astype(dtype)

    if kind == 1:
        out = np.array([0 if t.bit() else int(tmax) for _ in range(n)], dtype)
    elif kind in (2, 3):
        out = scale([t.number() for _ in range(n)])
    else:
        if t.pos >= len(b) or not chr(b[t.pos]).isspace():
            _pnm_fail("PNM: malformed header")
        pos = t.pos + 1

        def take(k):
            if k > len(b) - pos:
                _pnm_fail("PNM: unexpected end of data (truncated file?)")
            return np.frombuffer(b, np.uint8, k, pos)

        if kind == 4:
            rb = (w + 7) // 8
            bits = np.unpackbits(take(rb * h).reshape(h, rb), axis=1)[:, :w]
            out = np.where(bits.astype(bool), 0, int(tmax)).astype(dtype).ravel()
        elif maxval < 256:
            out = scale(take(n))
        else:
            r = take(2 * n).astype(np.int64)
            out = scale((r[0::2] << 8) | r[1::2])
    return out.reshape((h, w, 3) if cn == 3 else (h, w))


def pfm_decode(b: bytes) -> np.ndarray:
    """decodePFM: float32, rows stored bottom-up; byte order from the scale's sign."""
    b = bytes(b)
    if len(b) < 3 or b[0] != 0x50 or b[1] not in (0x46, 0x66):
        _pnm_fail("PFM: not a PFM file")
    cn = 3 if b[1] == 0x46 else 1
    t = _Tok(b, 2)
    w, h = t.number(), t.number()
    _check_size(w, h, "PFM")
    word = t.word()
    try:
        sc = float(word)
    except ValueError:
        _pnm_fail("PFM: invalid scale value")
    if sc == 0:
        _pnm_fail("PFM: invalid scale value")
    pos = t.pos
    if pos >= len(b):
        _pnm_fail("PFM: truncated")
    pos += 1
    n = w * h * cn
    if 4 * n > len(b) - pos:
        _pnm_fail("PFM: unexpected end of data (truncated file?)")
    a = np.frombuffer(b, "<f4" if sc < 0 else ">f4", n, pos).astype(np.float32)
    a = a.reshape(h, w * cn)[::-1]
    return np.ascontiguousarray(a.reshape((h, w, 3) if cn == 3 else (h, w)))


def _to_1_or_3(img):
    if img.ndim == 2:
        return img
    c = img.shape[2]
    if c == 1 or c == 2:
        return img[..., 0]
    if c == 3:
        return img
    return img[..., :3]


def pnm_encode(img: np.ndarray, binary: bool = True) -> bytes:
    """encodePNM: P5/P6 (binary) or P2/P3; uint8 maxval 255, uint16 maxval 65535."""
    if img.size == 0:
        _pnm_fail("encodePNM: empty image")
    a = np.ascontiguousarray(_to_1_or_3(img))
    one = a.ndim == 2
    maxval = 255 if a.dtype == np.uint8 else 65535
    hdr = f"P{(5 if binary else 2) if one else (6 if binary else 3)}\n{a.shape[1]} {a.shape[0]}\n{maxval}\n"
    if binary:
        body = a.astype(">u2").tobytes() if a.dtype == np.uint16 else a.tobytes()
        return hdr.encode() + body
    per = a.shape[1] * (1 if one else 3)
    flat = a.ravel()
    parts = [str(int(v)) + ("\n" if (i + 1) % per == 0 else " ") for i, v in enumerate(flat)]
    return hdr.encode() + "".join(parts).encode()


def pfm_encode(img: np.ndarray) -> bytes:
    """encodePFM: little-endian (scale -1.0), rows bottom-up, samples as float."""
    if img.ndim == 3 and img.shape[2] not in (1, 3) or img.ndim not in (2, 3):
        _pnm_fail("encodePFM: image must have 1 or 3 channels")
    if img.size == 0:
        _pnm_fail("encodePFM: empty image")
    cn = 3 if img.ndim == 3 and img.shape[2] == 3 else 1
    h, w = img.shape[:2]
    hdr = f"{'PF' if cn == 3 else 'Pf'}\n{w} {h}\n-1.0\n"
    return hdr.encode() + np.ascontiguousarray(img.astype("<f4")[::-1]).tobytes()


class RefCodec:
    """The reference's own Netpbm / PFM codecs (oracle/_ref)."""

    def __init__(self):
        if not REF_SO.exists():
            raise OracleError(f"{REF_SO} not built (make -C oracle)")
        self.lib = C.CDLL(str(REF_SO))
        self.lib.ipfr_last_error.restype = C.c_char_p

    def _err(self, rc):
        if rc != 0:
            raise PnmError(self.lib.ipfr_last_error().decode())

    def decode_pnm(self, b: bytes, dtype=np.uint8):
        w, h, cn = C.c_int(), C.c_int(), C.c_int()
        self._err(self.lib.ipfr_decode_pnm(b, C.c_size_t(len(b)), int(dtype == np.uint16), None, C.c_size_t(0),
                                           C.byref(w), C.byref(h), C.byref(cn)))
        out = np.empty((h.value, w.value, cn.value), dtype)
        self._err(self.lib.ipfr_decode_pnm(b, C.c_size_t(len(b)), int(dtype == np.uint16), out.ctypes.data_as(_V),
                                           C.c_size_t(out.size), C.byref(w), C.byref(h), C.byref(cn)))
        return out[..., 0] if cn.value == 1 else out

    def decode_pfm(self, b: bytes):
        w, h, cn = C.c_int(), C.c_int(), C.c_int()
        self._err(self.lib.ipfr_decode_pfm(b, C.c_size_t(len(b)), None, C.c_size_t(0), C.byref(w), C.byref(h),
                                           C.byref(cn)))
        out = np.empty((h.value, w.value, cn.value), np.float32)
        self._err(self.lib.ipfr_decode_pfm(b, C.c_size_t(len(b)), out.ctypes.data_as(_V), C.c_size_t(out.size),
                                           C.byref(w), C.byref(h), C.byref(cn)))
        return out[..., 0] if cn.value == 1 else out

    def _encode(self, fn, img, *args):
        a = np.ascontiguousarray(img)
        h, w = a.shape[:2]
        cn = 1 if a.ndim == 2 else a.shape[2]
        need = C.c_size_t(0)
        self._err(fn(a.ctypes.data_as(_V), w, h, cn, NP2DT[a.dtype], *args, None, C.c_size_t(0), C.byref(need)))
        buf = C.create_string_buffer(need.value)
        self._err(fn(a.ctypes.data_as(_V), w, h, cn, NP2DT[a.dtype], *args, buf, need, C.byref(need)))
        return buf.raw[:need.value]

    def encode_pnm(self, img, binary=True):
        return self._encode(self.lib.ipfr_encode_pnm, img, int(bool(binary)))

    def encode_pfm(self, img):
        return self._encode(self.lib.ipfr_encode_pfm, img)
