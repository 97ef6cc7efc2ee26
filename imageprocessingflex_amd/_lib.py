"""ctypes binding of libipfg.so (include/ipfg.h).

The library is loaded from this package directory; there is no fallback: if
libipfg.so is missing or does not load, every call raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libipfg.so"

# ipfg_dtype
U8, U16, I16, I32, F32, F64 = range(6)
# ipfg_border (ipf::BorderMode order)
CONSTANT, REPLICATE, REFLECT, REFLECT101, WRAP = range(5)
# ipfg_precision
PREC_DEFAULT, PREC_EXACT, PREC_FP32, PREC_BF16X3, PREC_FP16X2 = range(5)
# ipfg_status
OK, ERR_ARG, ERR_DEVICE, ERR_NOMEM, ERR_UNSUPPORTED, ERR_STATE = range(6)
# ipfg_op
OP_GAUSSIAN, OP_MEDIAN, OP_BILATERAL = range(3)


class IpfgImage(C.Structure):
    _fields_ = [
        ("data", C.c_void_p),
        ("width", C.c_int),
        ("height", C.c_int),
        ("channels", C.c_int),
        ("dtype", C.c_int),
        ("row_stride_bytes", C.c_size_t),
        ("on_device", C.c_int),
    ]


class IpfgJob(C.Structure):
    _fields_ = [
        ("src", IpfgImage),
        ("dst", IpfgImage),
        ("op", C.c_int),
        ("border", C.c_int),
        ("sigma", C.c_double),
        ("ksize", C.c_int),
        ("d", C.c_int),
        ("sigma_color", C.c_double),
        ("sigma_space", C.c_double),
    ]


_P = C.POINTER
_IMG = _P(IpfgImage)
_JOB = _P(IpfgJob)
_D = C.c_double
_I = C.c_int
_V = C.c_void_p

# name -> (restype, argtypes); every symbol declared in include/ipfg.h
SIGNATURES = {
    "ipfg_init": (_I, [_I]),
    "ipfg_shutdown": (_I, []),
    "ipfg_last_error": (C.c_char_p, []),
    "ipfg_abi_version": (_I, []),
    "ipfg_last_fixups": (_I, [_P(C.c_longlong)]),
    "ipfg_autotune": (_I, [_I]),
    "ipfg_tune_export": (_I, [C.c_char_p, C.c_size_t, _P(C.c_size_t)]),
    "ipfg_tune_import": (_I, [C.c_char_p, _P(_I)]),
    "ipfg_tune_clear": (_I, []),
    "ipfg_set_num_threads": (None, [_I]),
    "ipfg_get_num_threads": (_I, []),
    "ipfg_malloc": (_I, [_P(_V), C.c_size_t]),
    "ipfg_free": (_I, [_V]),
    "ipfg_memcpy_h2d": (_I, [_V, _V, C.c_size_t, _V]),
    "ipfg_memcpy_d2h": (_I, [_V, _V, C.c_size_t, _V]),
    "ipfg_memset": (_I, [_V, _I, C.c_size_t, _V]),
    "ipfg_stream_create": (_I, [_P(_V)]),
    "ipfg_stream_destroy": (_I, [_V]),
    "ipfg_stream_synchronize": (_I, [_V]),
    "ipfg_stream_forget": (_I, [_V]),
    "ipfg_device_synchronize": (_I, []),
    "ipfg_gaussian_kernel_1d": (_I, [_D, _I, _P(_D), _I]),
    "ipfg_sep_filter": (_I, [_IMG, _IMG, _P(_D), _I, _P(_D), _I, _I, _I, _I, _D, _D, _I, _V]),
    "ipfg_sep_filter_noalloc": (_I, [_IMG, _IMG, _P(_D), _I, _P(_D), _I, _I, _I, _I, _D, _D, _I, _V]),
    "ipfg_gaussian_blur": (_I, [_IMG, _IMG, _D, _D, _I, _I, _I, _V]),
    "ipfg_box_filter": (_I, [_IMG, _IMG, _I, _I, _I, _I, _D, _I, _V]),
    "ipfg_filter2d": (_I, [_IMG, _IMG, _P(C.c_float), _I, _I, _I, _I, _I, _D, _D, _I, _V]),
    "ipfg_median": (_I, [_IMG, _IMG, _I, _I, _V]),
    "ipfg_bilateral": (_I, [_IMG, _IMG, _I, _D, _D, _I, _I, _V]),
    "ipfg_sobel": (_I, [_IMG, _IMG, _I, _I, _I, _D, _I, _I, _V]),
    "ipfg_scharr": (_I, [_IMG, _IMG, _I, _I, _D, _I, _I, _V]),
    "ipfg_prewitt": (_I, [_IMG, _IMG, _I, _I, _I, _I, _V]),
    "ipfg_laplacian": (_I, [_IMG, _IMG, _I, _D, _I, _I, _V]),
    "ipfg_magnitude": (_I, [_IMG, _IMG, _IMG, _V]),
    "ipfg_phase": (_I, [_IMG, _IMG, _IMG, _I, _V]),
    "ipfg_unsharp_mask": (_I, [_IMG, _IMG, _D, _D, _D, _I, _V]),
    "ipfg_sharpen": (_I, [_IMG, _IMG, _D, _I, _V]),
    "ipfg_emboss": (_I, [_IMG, _IMG, _I, _V]),
    "ipfg_adaptive_threshold": (_I, [_IMG, _IMG, _D, _I, _I, _I, _D, _I, _V]),
    "ipfg_pyr_down": (_I, [_IMG, _IMG, _V]),
    "ipfg_pyr_up": (_I, [_IMG, _IMG, _V]),
    "ipfg_ssim": (_I, [_IMG, _IMG, _IMG, _P(_D), _V]),
    "ipfg_hough_circles": (_I, [_IMG, _I, _I, _D, _I, _D, _D, _P(_D), _I, _P(_I), _V]),
    "ipfg_pnm_info": (_I, [C.c_char_p, C.c_size_t, _P(_I), _P(_I), _P(_I), _P(_I)]),
    "ipfg_decode_pnm": (_I, [C.c_char_p, C.c_size_t, _IMG, _V]),
    "ipfg_integral_image": (_I, [_IMG, _IMG, _I, _V]),
    "ipfg_filter2d_fft": (_I, [_IMG, _IMG, _P(C.c_float), _I, _I, _I, _D, _D, _V]),
    "ipfg_canny": (_I, [_IMG, _IMG, _D, _D, _I, _I, _V]),
    "ipfg_canny_auto": (_I, [_IMG, _IMG, _D, _V]),
    "ipfg_corner_harris": (_I, [_IMG, _IMG, _I, _I, _D, _V]),
    "ipfg_corner_min_eigen_val": (_I, [_IMG, _IMG, _I, _I, _V]),
    "ipfg_batch_create": (_I, [_P(_V), _I, _I]),
    "ipfg_batch_run": (_I, [_V, _JOB, _I, _V]),
    "ipfg_batch_destroy": (_I, [_V]),
    "ipfg_server_create": (_I, [_P(_V), C.c_size_t, _I, _I]),
    "ipfg_server_submit": (_I, [_V, _JOB, _P(C.c_longlong)]),
    "ipfg_server_submit_many": (_I, [_V, _JOB, _I, _P(C.c_longlong)]),
    "ipfg_server_poll": (_I, [_V, _P(C.c_longlong), _P(_D), _I, _P(_I)]),
    "ipfg_server_drain": (_I, [_V]),
    "ipfg_server_stats": (_I, [_V, _P(C.c_longlong), _P(_I)]),
    "ipfg_server_memory": (_I, [_V, _P(C.c_size_t), _P(C.c_size_t)]),
    "ipfg_server_destroy": (_I, [_V]),
}

_lib = None


class IpfError(RuntimeError):
    """An argument check of the reference failed (ipf::Error)."""


class DeviceError(RuntimeError):
    """The HIP runtime or the library failed."""


def load(path: os.PathLike | str | None = None) -> C.CDLL:
    """Loads libipfg.so and binds every declared symbol (raises if anything is missing)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"{p} is not built; run imageprocessingflex_amd.build.build() first")
    lib = C.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)  # AttributeError if the symbol is absent
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def check(rc: int) -> None:
    if rc == OK:
        return
    msg = load().ipfg_last_error().decode()
    if rc == ERR_ARG:
        raise IpfError(msg)
    raise DeviceError(f"status {rc}: {msg}")
