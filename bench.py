"""Benchmark of the MI355X filtering path (one GPU).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c2|c1|c3|c4|c5|deriv|pyr|features]
                    [--no-cpu] [--no-extra] [--cpu-budget S]

The default run (c2) prints the c2 headline line with the other
configurations measured in the same run under "configs" (c1, c3, c4: the
same measurement with 5 steps; c5: a 3000-request prefix of the serving
stream), each with its own roofline and CPU baseline.

Workloads (one step = every call listed once; inputs resident in HBM; enough
input/output sets rotate that the working set exceeds the 256 MB Infinity
Cache):
  c2 (default)  7680x4320x1 float32 (16 seeded sinusoids + N(0, 0.05) noise,
                clamped to [0, 1]): gaussianBlur sigma 1, 3, 8 and blur 15x15,
                each under all five border modes (20 calls).
  c1            1920x1080x3 uint8 uniform: gaussianBlur sigma 2, Reflect101.
  c3            3840x2160x3 uint8 (64 Voronoi regions + N(0, 8) + 5% salt and
                pepper): medianBlur 3 and 5 (Replicate), bilateralFilter(9, 75, 75).
  c4            4096x4096x1 uint8 and float32 uniform: filter2D with a dense
                31x31 N(0, 1) kernel normalised to sum 1, Reflect101.
  deriv         3840x2160x3 uint8 uniform: sobel (dx 3x3, dy 5x5), scharr,
                laplacian (ksize 1 and 3) with float32 output, unsharpMask,
                sharpen, emboss; 3840x2160x1 uint8: adaptiveThreshold (mean and
                Gaussian, block 11).
  pyr           3840x2160x3 uint8 polygons + noise (the c3 image): pyrDown and
                pyrUp (exact FP64 in the reference's operation order); pixels
                are counted at the input.
  features      1920x1080x1 uint8 disc scene: canny, cannyAuto, cornerHarris,
                cornerMinEigenVal, integralImage, houghCircles; filter2DFFT
                31x31 on its float32 version. Bytes count the input and the
                output image (houghCircles: the input only).
All calls use the default precision: FP64 accumulation, certified bit-identical
to the reference (median is compare-only).

Prints ONE JSON line: value = Gpix/s over the whole step (the step's calls
back to back, wall clock between two device synchronisations), the roofline
of the dominant kernel (algorithmic bytes per launch / mean launch time from
HIP events around 4 back-to-back launches on the launch stream), per-filter
Gpix/s and roofline fraction, the latency of one call on an idle stream, and
the reference's multithreaded CPU path on this host (one call per filter
group, one warm-up then the median of 5 runs, 3 for filter2D), with the host
CPU model and its physical and logical core counts.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import statistics
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
MFMA_PEAK_TFLOPS = 2500.0  # dense bf16 / fp16
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 vector and FP64 matrix (dense)
BORDERS = ["constant", "replicate", "reflect", "reflect101", "wrap"]
METRIC = "Gpix/s + % HBM roofline per filter; p50/p99 latency, batched vs plain"


class HipEvents:
    """hipEvent timing on an explicit stream (the one the kernels run on)."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so")
        self.hip.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.hip.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.hip.hipEventSynchronize.argtypes = [C.c_void_p]
        self.hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        self.hip.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]

    def make(self, n):
        evs = []
        for _ in range(n):
            e = C.c_void_p()
            assert self.hip.hipEventCreate(C.byref(e)) == 0
            evs.append(e)
        return evs

    def record(self, ev, stream):
        assert self.hip.hipEventRecord(ev, stream) == 0

    def ms(self, a, b):
        self.hip.hipEventSynchronize(b)
        out = C.c_float()
        assert self.hip.hipEventElapsedTime(C.byref(out), a, b) == 0
        return out.value


class Call:
    """One filter call of a step: `launch(lib, src, dst, stream)` enqueues it
    on device descriptors; `cpu(oracle, img)` runs the reference's CPU path."""

    def __init__(self, name, group, image, out_dtype, launch, cpu, kernel, flops=0.0, out_shape=None):
        self.name, self.group, self.image, self.out_dtype = name, group, image, out_dtype
        self.launch, self.cpu, self.kernel, self.flops = launch, cpu, kernel, flops
        # out_shape(h, w, c) -> (h', w') when the output's size differs from the input's
        self.out_shape = out_shape

    def out_key(self):
        return (self.image, np.dtype(self.out_dtype).str, self.name if self.out_shape else "")

    def out_image(self, ipf, src):
        if self.out_shape is None:
            return ipf.DeviceImage.like(src, self.out_dtype)
        h, w = self.out_shape(src.height, src.width, src.channels)
        return ipf.DeviceImage(h, w, src.channels, self.out_dtype)

    def out_size(self, im):
        if self.out_shape is None:
            return im.size
        h, w = self.out_shape(*im.shape[:2], im.shape[2] if im.ndim == 3 else 1)
        return h * w * (im.shape[2] if im.ndim == 3 else 1)


def _c2():
    from imageprocessingflex_amd import workloads as wl
    img = wl.sinusoid_image(4320, 7680, seed=2)
    calls = []
    for name, sigma in (("gauss_s1", 1.0), ("gauss_s3", 3.0), ("gauss_s8", 8.0), ("box15", None)):
        for b in range(5):
            if sigma is None:
                launch = (lambda b: lambda lib, s, d, st: lib.ipfg_box_filter(s, d, 15, 15, 1, b, 0.0, 0, st))(b)
                cpu = (lambda b: lambda o, im: o.box_filter(im, 15, 15, True, b))(b)
                kern = "sep1_kernel<15, CH, true> (one-channel float32 strip kernel, unit weights)"
            else:
                launch = (lambda b, sg: lambda lib, s, d, st: lib.ipfg_gaussian_blur(s, d, sg, 0.0, 0, b, 0, st))(b, sigma)
                cpu = (lambda b, sg: lambda o, im: o.gaussian_blur(im, sg, border=b))(b, sigma)
                ks = 2 * max(1, int(3 * sigma + 0.5)) + 1
                kern = f"sep1_kernel<{ks}, CH, false> (one-channel float32 strip kernel)"
            # one FMA per tap and output sample in each pass (halo rows not counted)
            taps = 30 if sigma is None else 2 * (2 * max(1, int(3 * sigma + 0.5)) + 1)
            calls.append(Call(f"{name}/{BORDERS[b]}", name, "img", np.float32, launch, cpu, kern,
                              flops=2.0 * taps * img.shape[0] * img.shape[1]))
    return {"images": {"img": img}, "calls": calls, "dtype": "f64",
            "workload": "gaussianBlur sigma 1/3/8 + blur 15x15, x 5 border modes, 7680x4320x1 float32"}


def _c1():
    from imageprocessingflex_amd import workloads as wl
    img = wl.c1_image()
    calls = [Call("gauss_s2/reflect101", "gauss_s2", "img", np.uint8,
                  lambda lib, s, d, st: lib.ipfg_gaussian_blur(s, d, 2.0, 0.0, 0, 3, 0, st),
                  lambda o, im: o.gaussian_blur(im, 2.0, border=3),
                  "tile_u8_kernel<3, 13, TY> (one-pass tile kernel; the autotuner picks TY or a strip kernel)")]
    return {"images": {"img": img}, "calls": calls, "dtype": "f64",
            "workload": "gaussianBlur sigma 2 (13 taps), Reflect101, 1920x1080x3 uint8"}


def _c3():
    from imageprocessingflex_amd import workloads as wl
    img = wl.polygon_image()
    calls = [
        Call("median3/replicate", "median3", "img", np.uint8,
             lambda lib, s, d, st: lib.ipfg_median(s, d, 3, 1, st), lambda o, im: o.median(im, 3, 1),
             "median_kernel<unsigned char, 3>"),
        Call("median5/replicate", "median5", "img", np.uint8,
             lambda lib, s, d, st: lib.ipfg_median(s, d, 5, 1, st), lambda o, im: o.median(im, 5, 1),
             "median_kernel<unsigned char, 5>"),
        Call("bilateral9/reflect101", "bilateral9", "img", np.uint8,
             lambda lib, s, d, st: lib.ipfg_bilateral(s, d, 9, 75.0, 75.0, 3, 0, st),
             lambda o, im: o.bilateral(im, 9, 75.0, 75.0, 3),
             "bilateral_u8_kernel<3, false, double, true>"),
    ]
    return {"images": {"img": img}, "calls": calls, "dtype": "f64",
            "workload": "medianBlur 3x3 and 5x5 (Replicate) + bilateralFilter(9, 75, 75) (Reflect101), "
                        "3840x2160x3 uint8 polygons + noise"}


def _c4():
    from imageprocessingflex_amd import workloads as wl
    u8, f32 = wl.c4_images(4096)
    k, ksum = wl.c4_kernel(31)
    kp = np.ascontiguousarray(k, np.float32)
    kptr = kp.ctypes.data_as(C.POINTER(C.c_float))
    flops = 2.0 * 4096 * 4096 * 31 * 31
    names = {0: "", 3: "_bf16x3", 4: "_fp16x2"}
    kern = {0: "f2d_fast_kernel<T, double, false, true>", 3: "f2d_mfma_kernel<T, true, 3, XS>",
            4: "f2d_mfma_kernel<T, false, 2, XS>"}
    calls = []
    for prec in (0, 4, 3):
        for name, key, dt in (("filter2d_u8", "u8", np.uint8), ("filter2d_f32", "f32", np.float32)):
            calls.append(Call(f"{name}{names[prec]}/reflect101", f"{name}{names[prec]}", key, dt,
                              lambda lib, s, d, st, prec=prec: lib.ipfg_filter2d(s, d, kptr, 31, 31, 15, 15, 3, 0.0,
                                                                                  0.0, prec, st),
                              lambda o, im: o.filter2d(im, kp, border=3), kern[prec], flops=flops))

    def accuracy(ipf):
        """max / RMS error of every precision against the FP64 oracle (filter2D<double>)."""
        from oracle.pyoracle import Oracle, ref_available
        o = Oracle("ref") if ref_available() else Oracle("oracle")
        o.set_num_threads(0)
        out = {}
        for key, img in (("u8", u8), ("f32", f32)):
            ref = o.filter2d(img, kp, border=3, out_dtype=np.float64)
            ref_u8 = o.filter2d(img, kp, border=3) if key == "u8" else None
            for prec in (0, 4, 3):
                got = ipf.filter2d(img, kp, 3, out_dtype=np.float32, precision=prec).astype(np.float64)
                err = got - ref
                e = {"max_abs_err": float(np.abs(err).max()), "rms_err": float(np.sqrt(np.mean(err * err))),
                     "vs": "FP64 oracle (filter2D<double>), float32 output"}
                if ref_u8 is not None:
                    g8 = ipf.filter2d(img, kp, 3, precision=prec)
                    e["u8_mismatches"] = int(np.count_nonzero(g8 != ref_u8))
                    e["u8_max_diff"] = int(np.abs(g8.astype(int) - ref_u8.astype(int)).max())
                out[f"filter2d_{key}{names[prec]}"] = e
        return out

    return {"images": {"u8": u8, "f32": f32}, "calls": calls, "dtype": "f64 / bf16x3 / fp16x2", "roof": "mfma",
            "workload": "filter2D dense 31x31 float32 kernel (N(0,1), sum 1), Reflect101, 4096x4096x1 uint8 and "
                        "float32; default (FP64, certified), FP16X2 and BF16X3 MFMA",
            "keep": kp, "post": accuracy, "kernel_raw_sum": ksum}


def _deriv():
    from imageprocessingflex_amd import workloads as wl
    rgb = wl.uniform_u8((2160, 3840, 3), seed=7, stream=0)
    gray = wl.uniform_u8((2160, 3840), seed=7, stream=1)
    f = np.float32
    u8 = np.uint8
    calls = [
        Call("sobel_dx3", "sobel_dx3", "rgb", f, lambda lib, s, d, st: lib.ipfg_sobel(s, d, 1, 0, 3, 1.0, 3, 0, st),
             lambda o, im: o.sobel(im, 1, 0, 3), "stencil_u8_kernel<float, 3, 3, 3, true>"),
        Call("sobel_dy5", "sobel_dy5", "rgb", f, lambda lib, s, d, st: lib.ipfg_sobel(s, d, 0, 1, 5, 1.0, 3, 0, st),
             lambda o, im: o.sobel(im, 0, 1, 5), "stencil_u8_kernel<float, 5, 5, 3, true>"),
        Call("scharr_dx", "scharr_dx", "rgb", f, lambda lib, s, d, st: lib.ipfg_scharr(s, d, 1, 0, 1.0, 3, 0, st),
             lambda o, im: o.scharr(im, 1, 0), "stencil_u8_kernel<float, 3, 3, 3, true>"),
        Call("laplacian1", "laplacian1", "rgb", f, lambda lib, s, d, st: lib.ipfg_laplacian(s, d, 1, 1.0, 3, 0, st),
             lambda o, im: o.laplacian(im, 1), "stencil_u8_kernel<float, 3, 3, 3, false>"),
        Call("laplacian3", "laplacian3", "rgb", f, lambda lib, s, d, st: lib.ipfg_laplacian(s, d, 3, 1.0, 3, 0, st),
             lambda o, im: o.laplacian(im, 3), "stencil_u8_kernel<float, 3, 3, 3, false> (combined kernel)"),
        Call("unsharp_s1", "unsharp_s1", "rgb", u8,
             lambda lib, s, d, st: lib.ipfg_unsharp_mask(s, d, 1.0, 1.0, 0.0, 0, st),
             lambda o, im: o.unsharp_mask(im, 1.0, 1.0, 0.0),
             "unsharp_u8_kernel<7, 3>"),
        Call("sharpen", "sharpen", "rgb", u8, lambda lib, s, d, st: lib.ipfg_sharpen(s, d, 1.0, 0, st),
             lambda o, im: o.sharpen(im, 1.0), "stencil_u8_kernel<unsigned char, 3, 3, 3, false>"),
        Call("emboss", "emboss", "rgb", u8, lambda lib, s, d, st: lib.ipfg_emboss(s, d, 0, st),
             lambda o, im: o.emboss(im), "stencil_u8_kernel<unsigned char, 3, 3, 3, false>"),
        Call("adaptive_mean11", "adaptive_mean11", "gray", u8,
             lambda lib, s, d, st: lib.ipfg_adaptive_threshold(s, d, 255.0, 0, 0, 11, 2.0, 0, st),
             lambda o, im: o.adaptive_threshold(im, 255.0, 0, 0, 11, 2.0),
             "sep_plane_kernel<unsigned char, unsigned char, double, float, 1, 2, true, 11> + adaptive_rows_kernel"),
        Call("adaptive_gauss11", "adaptive_gauss11", "gray", u8,
             lambda lib, s, d, st: lib.ipfg_adaptive_threshold(s, d, 255.0, 1, 0, 11, 2.0, 0, st),
             lambda o, im: o.adaptive_threshold(im, 255.0, 1, 0, 11, 2.0),
             "sep_plane_kernel<unsigned char, unsigned char, double, float, 1, 2, true, 11> + adaptive_rows_kernel"),
    ]
    return {"images": {"rgb": rgb, "gray": gray}, "calls": calls, "dtype": "f64",
            "workload": "sobel/scharr/laplacian (float out), unsharpMask/sharpen/emboss on 3840x2160x3 uint8; "
                        "adaptiveThreshold mean/gaussian 11 on 3840x2160x1 uint8"}


def _pyr():
    from imageprocessingflex_amd import workloads as wl
    img = wl.polygon_image()
    u8 = np.uint8
    calls = [
        Call("pyr_down", "pyr_down", "img", u8, lambda lib, s, d, st: lib.ipfg_pyr_down(s, d, st),
             lambda o, im: o.pyr_down(im), "pyr_down_kernel<unsigned char>",
             out_shape=lambda h, w, c: ((h + 1) // 2, (w + 1) // 2)),
        Call("pyr_up", "pyr_up", "img", u8, lambda lib, s, d, st: lib.ipfg_pyr_up(s, d, st),
             lambda o, im: o.pyr_up(im), "pyr_up_kernel<unsigned char>", out_shape=lambda h, w, c: (2 * h, 2 * w)),
    ]
    return {"images": {"img": img}, "calls": calls, "dtype": "f64",
            "workload": "pyrDown + pyrUp, 3840x2160x3 uint8 polygons + noise (pixels counted at the input)"}


def _features():
    from imageprocessingflex_amd import workloads as wl
    gray = wl.circles_scene(1080, 1920, 24, 7, 3.0)
    grayf = (gray.astype(np.float32) / np.float32(255)).astype(np.float32)
    k, _ = wl.c4_kernel(31)
    kp = np.ascontiguousarray(k, np.float32)
    kptr = kp.ctypes.data_as(C.POINTER(C.c_float))
    circles = np.zeros((256, 4), np.float64)
    count = C.c_int(0)
    u8, f32, f64 = np.uint8, np.float32, np.float64

    def hough(lib, s, d, st):  # returns after the circles are in host memory
        return lib.ipfg_hough_circles(s, 20, 120, 100.0, 30, 10.0, 0.6, circles.ctypes.data_as(C.POINTER(C.c_double)),
                                      256, C.byref(count), st)

    calls = [
        Call("canny", "canny", "gray", u8, lambda lib, s, d, st: lib.ipfg_canny(s, d, 50.0, 150.0, 3, 0, st),
             lambda o, im: o.canny(im, 50.0, 150.0), "canny_nms_kernel + union-find hysteresis"),
        Call("canny_auto", "canny_auto", "gray", u8, lambda lib, s, d, st: lib.ipfg_canny_auto(s, d, 1.4, st),
             lambda o, im: o.canny_auto(im, 1.4), "gaussian + sobel + select_* (70th percentile) + canny"),
        Call("corner_harris", "corner_harris", "gray", f32,
             lambda lib, s, d, st: lib.ipfg_corner_harris(s, d, 3, 3, 0.04, st),
             lambda o, im: o.corner_harris(im, 3, 3, 0.04), "sobel + tensor_rows_kernel + tensor_cols_kernel"),
        Call("corner_min_eigen", "corner_min_eigen", "gray", f32,
             lambda lib, s, d, st: lib.ipfg_corner_min_eigen_val(s, d, 3, 3, st),
             lambda o, im: o.corner_min_eigen(im, 3, 3), "sobel + tensor_rows_kernel + tensor_cols_kernel"),
        Call("integral", "integral", "gray", f64, lambda lib, s, d, st: lib.ipfg_integral_image(s, d, 0, st),
             lambda o, im: o.integral_image(im), "integral_rows_kernel + integral_cols_kernel",
             out_shape=lambda h, w, c: (h + 1, w + 1)),
        Call("filter2d_fft31", "filter2d_fft31", "grayf", f32,
             lambda lib, s, d, st: lib.ipfg_filter2d_fft(s, d, kptr, 31, 31, 3, 0.0, 0.0, st),
             lambda o, im: o.filter2d_fft(im, kp, border=3), "fft_stages_kernel (radix-2, bit-exact)",
             flops=2.0 * 1080 * 1920 * 31 * 31),
        Call("hough_circles", "hough_circles", "gray", u8, hough,
             lambda o, im: o.hough_circles(im, 20, 120, 100.0, 30, 10.0, 0.6),
             "canny + hough votes / peaks / refine (synchronous: circles copied to the host)",
             out_shape=lambda h, w, c: (1, 1)),
    ]
    return {"images": {"gray": gray, "grayf": grayf}, "calls": calls, "dtype": "f64",
            "workload": "canny(50, 150), cannyAuto(1.4), cornerHarris(3, 3, 0.04), cornerMinEigenVal(3, 3), "
                        "integralImage, houghCircles(20..120) on a 1920x1080x1 uint8 disc scene; filter2DFFT "
                        "31x31 on its float32 version",
            "keep": (kp, circles, count)}


WORKLOADS = {"c1": _c1, "c2": _c2, "c3": _c3, "c4": _c4, "c5": None, "deriv": _deriv, "pyr": _pyr,
             "features": _features}

# Device-memory budget of the batched server: fixed, 4 GB of the 288 GB (the
# largest request needs 2 x 67 MB). Measured on the 3000-request prefix:
# 256 MB -> 2.6 x plain, 1 GB -> 3.1 x, 4 GB -> 3.4 x; batches above 256
# requests lost throughput.
C5_BUDGET = 4096 << 20
C5_MAX_BATCH = 256


def _pct(v, q):
    v = sorted(v)
    return v[min(len(v) - 1, int(round(q * (len(v) - 1))))] if v else None


def run_c5(args, ipf):
    """Serving stream: plain per-image launches vs the batched server.

    Requests (project RNG, seed 6): Poisson arrivals, W and H independently
    log-uniform in [64, 4096], channels in {1, 3, 4}, uint8; 50% gaussianBlur
    (sigma ~ U[0.5, 8], Reflect101), 25% medianBlur 5 (Replicate), 25%
    bilateralFilter(9, 75, 75) (Reflect101). Sources are device-resident
    views of seeded uniform banks (4096 x 4096 x cn); outputs are device
    buffers. Throughput is measured closed-loop (all requests available);
    latency open-loop at 0.8x and 1.2x the plain path's capacity.
    """
    from imageprocessingflex_amd import workloads as wl
    n = args.requests
    reqs = wl.serving_stream(n, rate=1.0, seed=6)  # unit-rate arrival times, scaled per run
    t0 = time.time()
    banks = {cn: ipf.DeviceImage.from_array(wl.squeeze(wl.uniform_u8((4096, 4096, cn), seed=6, stream=cn)))
             for cn in (1, 3, 4)}
    outbank = {cn: ipf.DeviceImage.like(banks[cn]) for cn in (1, 3, 4)}
    print(f"# c5 banks ready in {time.time() - t0:.1f} s", file=sys.stderr, flush=True)
    lib = ipf._lib.load()
    views = []
    for r in reqs:
        oy = (r.rid * 7919) % (4096 - r.height + 1)
        ox = (r.rid * 104729) % (4096 - r.width + 1)
        views.append((ipf.device_view(banks[r.channels], oy, ox, r.height, r.width),
                      ipf.device_view(outbank[r.channels], oy, ox, r.height, r.width)))
    px = sum(r.height * r.width for r in reqs)
    stream = ipf.Stream()
    ev = HipEvents()
    ipf.autotune(False)  # plain calls of many shape classes: built-in launch shapes, no first-call measuring

    def plain_call(i):
        r = reqs[i]
        s, d = views[i]
        if r.op == "gaussian":
            rc = lib.ipfg_gaussian_blur(C.byref(s), C.byref(d), r.sigma, 0.0, 0, 3, 0, stream.handle)
        elif r.op == "median5":
            rc = lib.ipfg_median(C.byref(s), C.byref(d), 5, 1, stream.handle)
        else:
            rc = lib.ipfg_bilateral(C.byref(s), C.byref(d), 9, 75.0, 75.0, 3, 0, stream.handle)
        ipf._lib.check(rc)

    # ---- plain path, closed loop: one blocking call per image (the drop-in
    # API's contract), kernel time from events between calls
    for i in range(min(64, n)):  # warm-up: parameter caches
        plain_call(i)
    stream.synchronize()
    evs = ev.make(n + 1)
    ipf.synchronize()
    w0 = time.perf_counter()
    ev.record(evs[0], stream.handle)
    for i in range(n):
        plain_call(i)
        ev.record(evs[i + 1], stream.handle)
        stream.synchronize()
    plain_wall = time.perf_counter() - w0
    dur = [ev.ms(evs[i], evs[i + 1]) for i in range(n)]
    plain_rate = n / plain_wall
    # the same calls enqueued back to back without waiting (one stream)
    ipf.synchronize()
    w0 = time.perf_counter()
    for i in range(n):
        plain_call(i)
    stream.synchronize()
    async_wall = time.perf_counter() - w0
    print(f"# plain: {plain_wall:.2f} s ({px / plain_wall / 1e9:.2f} Gpix/s), back-to-back {async_wall:.2f} s",
          file=sys.stderr, flush=True)

    def plain_latency(rate):
        # a single FIFO queue in front of the measured per-call times (+ the
        # mean host overhead per call)
        over = max(0.0, plain_wall - sum(dur) / 1e3) / n
        free, lat = 0.0, []
        for r, d in zip(reqs, dur):
            a = r.arrival / rate
            start = max(a, free)
            free = start + d / 1e3 + over
            lat.append((free - a) * 1e3)
        return lat

    def server_run(rate, graph=True):
        srv = ipf.Server(args.c5_budget_mb << 20, max_batch=args.c5_max_batch, graph=graph)
        jobs = []
        for r, (s, _) in zip(reqs, views):
            if r.op == "gaussian":
                jobs.append(ipf.job(s, None, "gaussian", sigma=r.sigma, border=3))
            elif r.op == "median5":
                jobs.append(ipf.job(s, None, "median", ksize=5, border=1))
            else:
                jobs.append(ipf.job(s, None, "bilateral", d=9, sigma_color=75.0, sigma_space=75.0, border=3))
        # warm-up batch (graph capture, parameter caches)
        srv.submit_jobs(jobs[:min(64, n)])
        srv.drain()
        srv.poll(1 << 16)
        lat = []
        w0 = time.perf_counter()
        if rate is None:
            srv.submit_jobs(jobs)
        else:
            i = 0
            while i < n:
                now = time.perf_counter() - w0
                j = i
                while j < n and reqs[j].arrival / rate <= now:
                    j += 1
                if j > i:
                    srv.submit_jobs(jobs[i:j])
                    i = j
                else:
                    lat += [l for _, l in srv.poll(1 << 16)]
                    time.sleep(min(0.0005, max(0.0, reqs[i].arrival / rate - now)))
        srv.drain()
        wall = time.perf_counter() - w0
        lat += [l for _, l in srv.poll(1 << 16)]
        batches, biggest = srv.stats()
        srv.close()
        return wall, lat, batches, biggest

    bw, _, bb, bmax = server_run(None)
    print(f"# batched: {bw:.2f} s, {px / bw / 1e9:.2f} Gpix/s, {bb} batches", file=sys.stderr, flush=True)
    loads = {}
    for f in (0.8, 1.2):
        rate = f * plain_rate
        wall, lat, batches, _ = server_run(rate)
        pl = plain_latency(rate)
        loads[f"{f}x_plain_capacity"] = {
            "arrival_rate_req_s": round(rate, 1),
            "batched": {"p50_ms": round(_pct(lat, 0.5), 3), "p99_ms": round(_pct(lat, 0.99), 3),
                        "gpix_s": round(px / wall / 1e9, 2), "batches": batches},
            "plain_fifo_model": {"p50_ms": round(_pct(pl, 0.5), 3), "p99_ms": round(_pct(pl, 0.99), 3)},
        }
    value = px / bw / 1e9
    cpu = None if args.no_cpu else c5_cpu_baseline(reqs, banks, args.cpu_budget)
    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Gpix/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 1,
        "ms_per_step": round(bw * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded project RNG; device-resident source banks)",
        "config": {"workload": f"serving stream, {n} requests: 50% gaussianBlur sigma~U[0.5,8], 25% median 5x5, "
                               "25% bilateral d=9; W,H log-uniform [64,4096]; cn in {1,3,4}; uint8",
                   "device_budget_bytes": args.c5_budget_mb << 20, "max_batch": args.c5_max_batch, "requests": n,
                   "total_pixels": px},
        "batched_vs_plain": {"batched_gpix_s": round(value, 2), "plain_gpix_s": round(px / plain_wall / 1e9, 2),
                             "speedup": round(plain_wall / bw, 2), "batches": bb, "largest_batch": bmax,
                             "plain_kernel_ms_total": round(sum(dur), 2),
                             "plain_back_to_back_gpix_s": round(px / async_wall / 1e9, 2),
                             "speedup_vs_back_to_back": round(async_wall / bw, 2),
                             "plain": "one blocking call per image (ipfg_* on device buffers + stream sync)"},
        "latency": loads,
        "cpu_baseline": cpu,
    }
    return line


def c5_cpu_baseline(reqs, banks, budget_s):
    """The reference's CPU path on the serving stream's requests in arrival
    order (each op with all host threads), on the prefix that fits the budget."""
    try:
        from oracle.pyoracle import Oracle, ref_available
        o = Oracle("ref") if ref_available() else Oracle("oracle")
    except Exception as e:
        return {"value": None, "unit": "Gpix/s", "cores": os.cpu_count(), "kind": None, "sample": f"unavailable: {e}"}
    o.set_num_threads(0)
    host = {cn: b.download() for cn, b in banks.items()}
    t_total, px, done = 0.0, 0, 0
    for r in reqs:
        oy = (r.rid * 7919) % (4096 - r.height + 1)
        ox = (r.rid * 104729) % (4096 - r.width + 1)
        img = np.ascontiguousarray(host[r.channels][oy:oy + r.height, ox:ox + r.width])
        t0 = time.perf_counter()
        if r.op == "gaussian":
            o.gaussian_blur(img, r.sigma, border=3)
        elif r.op == "median5":
            o.median(img, 5, 1)
        else:
            o.bilateral(img, 9, 75.0, 75.0, 3)
        t_total += time.perf_counter() - t0
        px += r.height * r.width
        done += 1
        if t_total > budget_s:
            break
    return {"value": round(px / t_total / 1e9, 5), "unit": "Gpix/s", "cores": o.get_num_threads(),
            "kind": "reference" if o.which == "ref" else "port", "host": host_cpu(),
            "sample": f"first {done} requests of the stream in arrival order, {t_total:.1f} s (prefix "
                      f"bounded by time; the rate extrapolates to the whole stream)"}


def host_cpu():
    """Model name and physical / logical core counts of this host (from /proc/cpuinfo)."""
    model, cores = None, set()
    phys = core = None
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "model name" and model is None:
                model = v
            elif k == "physical id":
                phys = v
            elif k == "core id":
                core = v
            elif not k and phys is not None:
                cores.add((phys, core))
                phys = core = None
        if phys is not None:
            cores.add((phys, core))
    except OSError:
        pass
    return {"model": model, "physical_cores": len(cores) or None, "logical_cpus": os.cpu_count(),
            "affinity_cpus": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None}


def cpu_baseline(wd, runs: int = 5):
    """The reference's CPU path (oracle/_ref when built, else the oracle
    restatement) with all host threads: for one call of each filter group of
    the step (the Reflect101 / first call), one warm-up run then the median of
    `runs` timed runs; value = the step's pixels over the step's time with each
    call costing its group's median."""
    try:
        from oracle.pyoracle import Oracle, ref_available
        o = Oracle("ref") if ref_available() else Oracle("oracle")
    except Exception as e:  # the CPU leg is a reported baseline, not the product
        return {"value": None, "unit": "Gpix/s", "cores": os.cpu_count(), "kind": None, "sample": f"unavailable: {e}"}
    o.set_num_threads(0)
    threads = o.get_num_threads()
    def cpu_group(c):  # the GPU precision variants of a call run the same reference code
        return c.group.replace("_fp16x2", "").replace("_bf16x3", "")

    reps = {}
    for c in wd["calls"]:
        g = cpu_group(c)
        if g not in reps or "reflect101" in c.name:
            reps[g] = c
    med, t_all = {}, 0.0
    for g, c in reps.items():
        im = wd["images"][c.image]
        c.cpu(o, im)  # warm-up
        ts = []
        for _ in range(runs):
            t0 = time.perf_counter()
            c.cpu(o, im)
            ts.append(time.perf_counter() - t0)
        med[g] = statistics.median(ts)
        t_all += sum(ts)
    step_s = sum(med[cpu_group(c)] for c in wd["calls"])
    step_px = sum(wd["images"][c.image].shape[0] * wd["images"][c.image].shape[1] for c in wd["calls"])
    return {"value": round(step_px / step_s / 1e9, 5), "unit": "Gpix/s", "cores": threads,
            "kind": "reference" if o.which == "ref" else "port",
            "host": host_cpu(),
            "sample": f"{len(reps)} calls ({', '.join(c.name for c in reps.values())}), each 1 warm-up + median "
                      f"of {runs} runs, {threads} threads, {t_all:.1f} s timed; the step's other border modes "
                      f"are costed at their filter's median",
            "median_ms": {g: round(v * 1e3, 2) for g, v in med.items()}}


def traffic_from_profiles(group: str):
    p = ROOT / "profiles" / "traffic.json"
    if not p.exists():
        return None
    d = json.loads(p.read_text())
    return d.get(group, {}).get("bytes_per_launch")


def calibrate_counters(nbytes: int) -> None:
    """Kernels of known traffic for the HBM byte counters (rocprofv3 --pmc
    FETCH_SIZE / WRITE_SIZE): torch's fill of an nbytes float tensor writes
    nbytes, its sum reads them once. Run outside every timed region."""
    import torch
    t = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda")
    t.fill_(1.0)
    float(t.sum())
    del t
    torch.cuda.synchronize()


def measure(wd, ipf, steps: int, warmup: int, cpu: bool, cpu_runs: int = 5):
    """One workload: the step's calls back to back (value), per-call launch
    times from HIP events (per filter, roofline of the dominant one), call
    latency on an idle stream, and the CPU baseline. Returns the JSON dict."""
    stream = ipf.Stream()
    # rotating device sets: > 256 MB of inputs + outputs in flight
    set_bytes = sum(im.nbytes * 2 for im in wd["images"].values())
    nsets = max(3, math.ceil(300e6 / set_bytes))
    sets = [{k: ipf.DeviceImage.from_array(im, stream) for k, im in wd["images"].items()} for _ in range(nsets)]
    outs = [{c.out_key(): c.out_image(ipf, s[c.image]) for c in wd["calls"]} for s in sets]
    calls = wd["calls"]
    ev = HipEvents()
    evs = ev.make(2 * len(calls))
    lib = ipf._lib.load()
    k = 0

    def run(c):
        nonlocal k
        s = sets[k % nsets][c.image].descriptor()
        d = outs[k % nsets][c.out_key()].descriptor()
        ipf._lib.check(c.launch(lib, C.byref(s), C.byref(d), stream.handle))
        k += 1

    for _ in range(warmup):
        for c in calls:
            run(c)
    stream.synchronize()
    ipf.synchronize()
    # timed region: the step's calls back to back, nothing recorded between them
    t_start = time.perf_counter()
    for _ in range(steps):
        for c in calls:
            run(c)
    stream.synchronize()
    ipf.synchronize()
    wall = time.perf_counter() - t_start
    ms_per_step = wall * 1e3 / steps
    step_px = sum(wd["images"][c.image].shape[0] * wd["images"][c.image].shape[1] for c in calls)
    value = step_px / (ms_per_step * 1e-3) / 1e9

    # launch durations: REP launches of each call (rotating sets) between one
    # pair of HIP events on the launch stream, repeated `steps` times
    REP = 4
    per_call = {c.name: [] for c in calls}
    for _ in range(steps):
        for i, c in enumerate(calls):
            ev.record(evs[2 * i], stream.handle)
            for _r in range(REP):
                run(c)
            ev.record(evs[2 * i + 1], stream.handle)
        for i, c in enumerate(calls):
            per_call[c.name].append(ev.ms(evs[2 * i], evs[2 * i + 1]) / REP)
    # call latency: one call between a pair of events, the device idle before it
    lat = []
    for _ in range(steps):
        for i, c in enumerate(calls):
            stream.synchronize()
            ev.record(evs[0], stream.handle)
            run(c)
            ev.record(evs[1], stream.handle)
            lat.append(ev.ms(evs[0], evs[1]))
    stream.synchronize()

    per_filter = {}
    for c in calls:
        im = wd["images"][c.image]
        g = per_filter.setdefault(c.group, {"ms": [], "bytes": im.nbytes + c.out_size(im) * np.dtype(c.out_dtype).itemsize,
                                            "px": im.shape[0] * im.shape[1], "flops": c.flops, "kernel": c.kernel})
        g["ms"] += per_call[c.name]
    summary = {}
    for name, g in per_filter.items():
        mean = statistics.mean(g["ms"])
        gbs = g["bytes"] / (mean * 1e-3) / 1e9
        summary[name] = {"mean_ms": round(mean, 4), "gpix_s": round(g["px"] / (mean * 1e-3) / 1e9, 2),
                         "gb_s": round(gbs, 1), "hbm_roofline_frac": round(gbs / HBM_PEAK_GBS, 4)}
        if g["flops"]:
            summary[name]["tflop_s"] = round(g["flops"] / (mean * 1e-3) / 1e12, 2)
    dom = max(summary, key=lambda n: summary[n]["mean_ms"])
    all_ms = sorted(lat)
    p50 = all_ms[len(all_ms) // 2]
    p99 = all_ms[min(len(all_ms) - 1, int(round(0.99 * (len(all_ms) - 1))))]
    achieved = summary[dom]["gb_s"]
    roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_from_profiles(dom),
            "kernel": per_filter[dom]["kernel"], "filter": dom,
            "algorithmic_bytes_per_launch": per_filter[dom]["bytes"], "mean_launch_ms": summary[dom]["mean_ms"]}
    fl = per_filter[dom]["flops"]
    if fl:
        # the FP64 arithmetic of the certified paths against MI355X's FP64
        # peak (vector and matrix peaks coincide): 1 FMA = 2 FLOP
        tf = summary[dom]["tflop_s"]
        peak = MFMA_PEAK_TFLOPS if ("bf16" in dom or "fp16" in dom) else FP64_PEAK_TFLOPS
        compute = {"achieved": tf, "peak": peak, "unit": "TFLOP/s", "frac": round(tf / peak, 4),
                   "algorithmic_flops_per_launch": fl}
        if wd.get("roof") == "mfma":
            roof.update(compute, bound="mfma", hbm_frac=round(achieved / HBM_PEAK_GBS, 4))
        else:
            roof["fp64_compute"] = compute
    # a plain device-to-device copy of the largest input, for context and as
    # the calibration kernel of the HBM byte counters
    big = max(sets[0].values(), key=lambda d: d.nbytes)
    dst_big = ipf.DeviceImage.like(big)
    for _ in range(2):
        assert ev.hip.hipMemcpyAsync(dst_big.ptr, big.ptr, big.nbytes, 3, stream.handle) == 0
    ev.record(evs[0], stream.handle)
    for _ in range(5):
        assert ev.hip.hipMemcpyAsync(dst_big.ptr, big.ptr, big.nbytes, 3, stream.handle) == 0
    ev.record(evs[1], stream.handle)
    copy_ms = ev.ms(evs[0], evs[1]) / 5
    copy = {"bytes_per_launch": 2 * big.nbytes, "mean_ms": round(copy_ms, 4),
            "gb_s": round(2 * big.nbytes / (copy_ms * 1e-3) / 1e9, 1)}
    calibrate_counters(big.nbytes)
    cpu_line = cpu_baseline(wd, cpu_runs) if cpu else None
    extra = {}
    if "post" in wd and cpu:
        extra["accuracy"] = wd["post"](ipf)
    mfma = {n: {"algorithmic_tflop_s": v["tflop_s"], "peak_tflop_s": MFMA_PEAK_TFLOPS,
                "frac": round(v["tflop_s"] / MFMA_PEAK_TFLOPS, 4)}
            for n, v in summary.items() if "tflop_s" in v and ("bf16" in n or "fp16" in n)}
    if mfma:
        extra["mfma_roofline"] = mfma
    return {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Gpix/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": wd["dtype"],
        "data": "synthetic (seeded project RNG, generated in-process)",
        "config": {"workload": wd["workload"], "calls_per_step": len(calls),
                   "precision": "default: FP64 accumulation, certified bit-identical to the reference",
                   "rotating_sets": nsets},
        "roofline": roof,
        "cpu_baseline": cpu_line,
        "per_filter": summary,
        "latency_ms": {"p50": round(p50, 4), "p99": round(p99, 4),
                       "what": "one filter call on an idle stream, HIP events around the call"},
        "launch_timing": f"mean of {REP} back-to-back launches between two HIP events, x {steps}",
        "device_copy": copy,
        **extra,
    }


# The other configurations reported beside the headline (default run): the
# same measurement with fewer steps; c5 on a prefix of the serving stream.
# A c1 step is one ~40 us call: 100 steps keep the final synchronisation and
# the first launch from weighing on its value.
EXTRA_CONFIGS = ("c1", "c3", "c4", "c5")
EXTRA_STEPS = {"c1": 100, "c3": 5, "c4": 5}
EXTRA_C5_REQUESTS = 3000


def _brief(line: dict) -> dict:
    """A config's entry in the headline line: its numbers without the metric header."""
    drop = {"metric", "unit", "n_gpus", "higher_is_better", "scaling", "vs_baseline", "data"}
    return {k: v for k, v in line.items() if k not in drop}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c2", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="c2 only: skip the other configurations")
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="c5: seconds of CPU baseline")
    ap.add_argument("--requests", type=int, default=20000, help="c5: number of requests")
    ap.add_argument("--c5-budget-mb", type=int, default=C5_BUDGET >> 20, help="c5: device-memory budget of the server")
    ap.add_argument("--c5-max-batch", type=int, default=C5_MAX_BATCH, help="c5: requests per batch at most")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build runs on a single MI355X (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)

    import imageprocessingflex_amd as ipf

    ipf.init(0)
    if args.workload == "c5":
        print(json.dumps(run_c5(args, ipf)), flush=True)
        return
    t0 = time.time()
    wd = WORKLOADS[args.workload]()
    print(f"# {args.workload} inputs ready in {time.time() - t0:.1f} s", file=sys.stderr, flush=True)
    cpu_runs = 3 if args.workload == "c4" else 5
    line = measure(wd, ipf, args.steps, args.warmup, not args.no_cpu, cpu_runs)
    if args.workload == "c2" and not args.no_extra:
        # every other configuration in the same run (its own roofline and CPU baseline)
        del wd
        configs = {}
        for name in EXTRA_CONFIGS:
            t0 = time.time()
            if name == "c5":
                sub = argparse.Namespace(**{**vars(args), "requests": EXTRA_C5_REQUESTS, "cpu_budget": 5.0})
                configs[name] = _brief(run_c5(sub, ipf))
            else:
                w = WORKLOADS[name]()
                configs[name] = _brief(measure(w, ipf, EXTRA_STEPS[name], 1, not args.no_cpu, 3 if name == "c4" else 5))
                del w
            print(f"# {name} measured in {time.time() - t0:.1f} s", file=sys.stderr, flush=True)
        line["configs"] = configs
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
