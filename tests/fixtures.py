"""Loads the golden fixtures written by tests/golden/make_goldens.py and runs a
case through any implementation with the ipf-style Python API."""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden"
U = 2.0 ** -24  # FP32 unit roundoff


def load_cases():
    data = np.load(GOLDEN / "fixtures.npz", allow_pickle=False)
    cases = json.loads((GOLDEN / "manifest.json").read_text())
    out = []
    for c in cases:
        i = c["id"]
        extra = {k: data[f"c{i}_{k}"] for k in c["extra"]}
        out.append((c, data[f"c{i}_in"], data[f"c{i}_out"], extra))
    return out


def case_id(case) -> str:
    c, inp, _, _ = case
    p = c["params"]
    tag = "-".join(f"{k}={v}" for k, v in sorted(p.items()))
    return f"{c['id']}:{c['op']}:{inp.dtype}:{'x'.join(map(str, inp.shape))}:{tag}"


CODEC_OPS = ("pnm_dec", "pnm_enc", "pfm_enc")


def _run_codec(case, gpu_ipf=None):
    """Netpbm / PFM cases: the numpy restatement of the reference's codecs, or
    (gpu_ipf) the package (decoding unpacks on the GPU)."""
    from oracle import pyoracle as po
    c, inp, _, _ = case
    p = c["params"]
    if c["op"] == "pnm_dec":
        b = inp.tobytes()
        if gpu_ipf is not None:
            return gpu_ipf.decode_pnm(b, np.dtype(p["dtype"]))
        return po.pfm_decode(b) if p["dtype"] == "float32" else po.pnm_decode(b, np.dtype(p["dtype"]))
    if c["op"] == "pnm_enc":
        b = (gpu_ipf.encode_pnm if gpu_ipf is not None else po.pnm_encode)(inp, bool(p["binary"]))
    else:
        b = (gpu_ipf.encode_pfm if gpu_ipf is not None else po.pfm_encode)(inp)
    return np.frombuffer(b, np.uint8)


def run_oracle(o, case):
    """Runs a fixture case through the oracle (oracle.pyoracle.Oracle)."""
    c, inp, _, ex = case
    p = c["params"]
    if c["op"] in CODEC_OPS:
        return _run_codec(case)
    if c["op"] in ("pyr_down", "pyr_up"):
        return getattr(o, c["op"])(inp)
    if c["op"] == "ssim":
        return o.ssim(inp, ex["b"])[1]
    if c["op"] == "integral":
        return o.integral_image(inp, bool(p["squared"]))
    if c["op"] in FEATURE_OPS:
        return _run_feature(o, case)
    out_dt = np.dtype(p["out"]) if "out" in p else None
    b = p["border"]
    if c["op"] == "gaussian":
        return o.gaussian_blur(inp, p["sigma"], p.get("sigma_y", 0.0), 0, b, out_dt)
    if c["op"] == "sep":
        return o.sep_filter(inp, ex["kx"], ex["ky"], tuple(p["anchor"]) if "anchor" in p else None, b, p["delta"],
                            p["bval"], out_dt)
    if c["op"] == "box":
        return o.box_filter(inp, p["kw"], p["kh"], p["normalize"], b, p["bval"], out_dt)
    if c["op"] == "filter2d":
        return o.filter2d(inp, ex["k"], b, p.get("delta", 0.0), tuple(p.get("anchor", (-1, -1))), p["bval"], out_dt)
    if c["op"] == "filter2d_fft":
        return o.filter2d_fft(inp, ex["k"], b, p["delta"], p["bval"], out_dt)
    if c["op"] == "median":
        return o.median(inp, p["k"], b)
    if c["op"] == "bilateral":
        return o.bilateral(inp, p["d"], p["sc"], p["ss"], b)
    return _run_composite(o, case)


FEATURE_OPS = ("canny", "canny_auto", "harris", "min_eigen", "hough")


def _run_feature(m, case, gpu=False):
    """canny / cannyAuto / cornerHarris / cornerMinEigenVal cases; m is the
    oracle or the package (whose min-eigen function carries the reference's
    name)."""
    c, inp, _, _ = case
    p = c["params"]
    if c["op"] == "canny":
        return m.canny(inp, p["low"], p["high"], p["ksize"], bool(p["l2"]))
    if c["op"] == "canny_auto":
        return m.canny_auto(inp, p["sigma"])
    if c["op"] == "harris":
        return m.corner_harris(inp, p["block"], p["ksize"], p["k"])
    if c["op"] == "hough":
        return m.hough_circles(inp, p["rmin"], p["rmax"], p["high"], p["votes"], p["min_dist"], p["coverage"])
    fn = m.corner_min_eigen_val if gpu else m.corner_min_eigen
    return fn(inp, p["block"], p["ksize"])


def _run_composite(m, case, **kw):
    """The derivative and sharpening cases; m is the oracle or the package,
    which name these functions alike (kw: the package's precision)."""
    c, inp, _, ex = case
    p = c["params"]
    b = p["border"]
    if c["op"] == "sobel":
        return m.sobel(inp, p["dx"], p["dy"], p["ksize"], p["scale"], b, **kw)
    if c["op"] == "scharr":
        return m.scharr(inp, p["dx"], p["dy"], p["scale"], b, **kw)
    if c["op"] == "prewitt":
        return m.prewitt(inp, p["dx"], p["dy"], b, **kw)
    if c["op"] == "laplacian":
        return m.laplacian(inp, p["ksize"], p["scale"], b, **kw)
    if c["op"] == "unsharp":
        return m.unsharp_mask(inp, p["sigma"], p["amount"], p["threshold"], **kw)
    if c["op"] == "sharpen":
        return m.sharpen(inp, p["strength"], **kw)
    if c["op"] == "emboss":
        return m.emboss(inp, **kw)
    if c["op"] == "adaptive":
        return m.adaptive_threshold(inp, p["maxval"], p["method"], p["type"], p["block"], p["c"], **kw)
    if c["op"] == "magnitude":
        return m.magnitude(inp, ex["gy"])
    if c["op"] == "phase":
        return m.phase(inp, ex["gy"], bool(p["degrees"]))
    raise ValueError(c["op"])


def run_gpu(ipf, case, precision=0):
    """Runs a fixture case through libipfg (the imageprocessingflex_amd API)."""
    c, inp, _, ex = case
    p = c["params"]
    if c["op"] in CODEC_OPS:
        return _run_codec(case, ipf)
    if c["op"] in ("pyr_down", "pyr_up"):
        return getattr(ipf, c["op"])(inp)
    if c["op"] == "ssim":
        return ipf.ssim(inp, ex["b"], return_map=True)[1]
    if c["op"] == "integral":
        return ipf.integral_image(inp, bool(p["squared"]))
    if c["op"] in FEATURE_OPS:
        return _run_feature(ipf, case, gpu=True)
    out_dt = np.dtype(p["out"]) if "out" in p else None
    b = p["border"]
    if c["op"] == "gaussian":
        return ipf.gaussian_blur(inp, p["sigma"], p.get("sigma_y", 0.0), 0, b, out_dt, precision)
    if c["op"] == "sep":
        return ipf.sep_filter2d(inp, ex["kx"], ex["ky"], b, p["delta"], p["bval"], out_dt, precision,
                                tuple(p["anchor"]) if "anchor" in p else None)
    if c["op"] == "box":
        return ipf.box_filter(inp, p["kw"], p["kh"], p["normalize"], b, p["bval"], out_dt, precision)
    if c["op"] == "filter2d":
        return ipf.filter2d(inp, ex["k"], b, p.get("delta", 0.0), tuple(p.get("anchor", (-1, -1))), p["bval"],
                            out_dt, precision)
    if c["op"] == "filter2d_fft":
        return ipf.filter2d_fft(inp, ex["k"], b, p["delta"], p["bval"], out_dt)
    if c["op"] == "median":
        return ipf.median_blur(inp, p["k"], b)
    if c["op"] == "bilateral":
        return ipf.bilateral_filter(inp, p["d"], p["sc"], p["ss"], b, precision)
    return _run_composite(ipf, case, **({} if c["op"] in ("magnitude", "phase") else {"precision": precision}))


def float_tolerance(case) -> np.ndarray | float:
    """Worst-case |FP32 result - FP64 reference| for float outputs of the fast path.

    sep/gaussian: (kw + kh + 5) u * (sum|kx| sum|ky| X + |delta|) + u |ref|
    box         : FP64 window sums: 1e-12 * kw * kh * X + u |ref|
    filter2D    : (taps + 5) u * (sum|k| X + |delta|) + u |ref|
    bilateral   : (2 taps + 10) u * X (weights, two positive sums, one division)
    phase       : one float ulp of atan2(gy, gx), times 180/pi in degrees,
                  plus one ulp of the result
    where X bounds |input| (and the Constant border value), u = 2^-24.
    """
    c, inp, ref, ex = case
    p = c["params"]
    X = float(np.abs(inp.astype(np.float64)).max(initial=0.0))
    X = max(X, abs(p.get("bval", 0.0)) if p.get("border") == 0 else 0.0)
    r = np.abs(ref.astype(np.float64))
    if c["op"] in ("gaussian", "sep"):
        if c["op"] == "gaussian":
            from math import exp  # noqa: F401
            ks = 2 * max(1, int(3.0 * p["sigma"] + 0.5)) + 1
            kw = kh = ks
            sx = sy = 1.0
            delta = 0.0
        else:
            kw, kh = len(ex["kx"]), len(ex["ky"])
            sx, sy = np.abs(ex["kx"]).sum(), np.abs(ex["ky"]).sum()
            delta = p["delta"]
        return (kw + kh + 5) * U * (sx * sy * X + abs(delta)) * 1.01 + U * r
    if c["op"] == "box":
        return 1e-12 * p["kw"] * p["kh"] * X + U * r
    if c["op"] == "filter2d":
        k = ex["k"]
        taps = int((k != 0).sum())
        return (taps + 5) * U * (np.abs(k.astype(np.float64)).sum() * X + abs(p.get("delta", 0.0))) * 1.01 + U * r
    if c["op"] == "phase":
        return phase_tolerance(inp, ex["gy"], ref, bool(p["degrees"]))
    if c["op"] == "bilateral":
        r_ = p["d"] // 2 if p["d"] > 0 else int(round(p["ss"] * 1.5))
        taps = sum(1 for j in range(-r_, r_ + 1) for i in range(-r_, r_ + 1) if i * i + j * j <= r_ * r_)
        return (2 * taps + 10) * U * X * 1.01 + U * r
    return 0.0


def phase_tolerance(gx, gy, ref, degrees) -> np.ndarray:
    """One float ulp of the angle atan2(gy, gx) (scaled to degrees) plus one
    ulp of the wrapped result: the bound for an atan2 that differs from the
    host's atan2f by at most an ulp."""
    t = np.abs(np.arctan2(gy.astype(np.float64), gx.astype(np.float64))).astype(np.float32)
    scale = 180.0 / np.pi if degrees else 1.0
    return (np.spacing(t).astype(np.float64) * scale + np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64)) * 1.01
