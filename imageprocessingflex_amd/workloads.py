"""Seeded synthetic inputs for the filtering workloads.

A counter-based SplitMix64 generator gives the same bits on every host and
compiler (the reference's tests use std::mt19937 with std::uniform_*
distributions, whose mapping is toolchain-specific). Value mappings:
  u8 uniform      : top 8 bits of the 64-bit output
  float uniform   : (z >> 11) * 2^-53 in [0, 1)
  normal          : Box-Muller on two uniforms
Every draw is identified by (seed, stream) so independent draws never overlap.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix(z: np.ndarray) -> np.ndarray:
    z = z.copy()
    z ^= z >> np.uint64(30)
    z *= _M1
    z ^= z >> np.uint64(27)
    z *= _M2
    z ^= z >> np.uint64(31)
    return z


def u64(seed: int, stream: int, n: int) -> np.ndarray:
    """n 64-bit outputs of stream `stream` of generator `seed`."""
    with np.errstate(over="ignore"):
        base = _mix(np.array([(seed * 0x100000001B3 + stream * 0xD1B54A32D192ED03) & (2**64 - 1)], np.uint64))[0]
        ctr = np.arange(1, n + 1, dtype=np.uint64) * _GOLDEN + base
        return _mix(ctr)


def uniform(seed: int, stream: int, n: int) -> np.ndarray:
    return (u64(seed, stream, n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed: int, stream: int, n: int) -> np.ndarray:
    u1 = 1.0 - uniform(seed, 2 * stream + 1000003, n)  # (0, 1]
    u2 = uniform(seed, 2 * stream + 1000004, n)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)


def uniform_u8(shape, seed: int, stream: int = 0) -> np.ndarray:
    n = int(np.prod(shape))
    return (u64(seed, stream, n) >> np.uint64(56)).astype(np.uint8).reshape(shape)


def uniform_f32(shape, seed: int, stream: int = 0) -> np.ndarray:
    return uniform(seed, stream, int(np.prod(shape))).astype(np.float32).reshape(shape)


def squeeze(img: np.ndarray) -> np.ndarray:
    return img[..., 0] if img.ndim == 3 and img.shape[2] == 1 else img


# --------------------------------------------------------------------- configs

def c1_image() -> np.ndarray:
    """1920x1080x3 uint8, i.i.d. uniform [0, 255], seed 1."""
    return uniform_u8((1080, 1920, 3), seed=1)


def sinusoid_image(h: int = 4320, w: int = 7680, seed: int = 2) -> np.ndarray:
    """float32 in [0, 1]: 0.5 + sum of 16 plane waves (amplitudes summing to
    0.5, frequencies log-uniform in [1/2048, 1/4] cycles/px, uniform angles and
    phases) + N(0, 0.05) noise, clamped. Each wave is evaluated separably:
    sin(a x + b y + p) = sin(a x + p) cos(b y) + cos(a x + p) sin(b y)."""
    k = 16
    wts = uniform(seed, 1, k)
    amp = 0.5 * wts / wts.sum()
    freq = np.exp(np.log(1 / 2048) + (np.log(1 / 4) - np.log(1 / 2048)) * uniform(seed, 2, k))
    theta = 2 * math.pi * uniform(seed, 3, k)
    phase = 2 * math.pi * uniform(seed, 4, k)
    x = np.arange(w, dtype=np.float64)
    y = np.arange(h, dtype=np.float64)
    img = np.full((h, w), 0.5)
    for i in range(k):
        a = 2 * math.pi * freq[i] * math.cos(theta[i])
        b = 2 * math.pi * freq[i] * math.sin(theta[i])
        sx, cx = np.sin(a * x + phase[i]), np.cos(a * x + phase[i])
        sy, cy = np.sin(b * y), np.cos(b * y)
        img += amp[i] * (np.outer(cy, sx) + np.outer(sy, cx))
    img += 0.05 * normal(seed, 5, h * w).reshape(h, w)
    return np.clip(img, 0.0, 1.0).astype(np.float32)


def polygon_image(h: int = 2160, w: int = 3840, cn: int = 3, seed: int = 3, regions: int = 64,
                  noise_std: float = 8.0, sp_density: float = 0.05) -> np.ndarray:
    """uint8: Voronoi cells of `regions` seeded points with uniform colours,
    plus N(0, noise_std) noise, saturated, then salt-and-pepper on a
    `sp_density` fraction of pixels (half 0, half 255, all channels)."""
    from scipy.spatial import cKDTree

    pts = np.stack([uniform(seed, 1, regions) * w, uniform(seed, 2, regions) * h], 1)
    colors = uniform(seed, 3, regions * cn).reshape(regions, cn) * 255.0
    tree = cKDTree(pts)
    label = np.empty((h, w), np.int32)
    xs = np.arange(w) + 0.5
    step = max(1, (1 << 20) // w)
    for y0 in range(0, h, step):
        ys = np.arange(y0, min(h, y0 + step)) + 0.5
        gx, gy = np.meshgrid(xs, ys)
        _, idx = tree.query(np.stack([gx.ravel(), gy.ravel()], 1))
        label[y0:y0 + len(ys)] = idx.reshape(len(ys), w)
    img = colors[label] + noise_std * normal(seed, 4, h * w * cn).reshape(h, w, cn)
    img = np.clip(np.floor(img + 0.5), 0, 255).astype(np.uint8)
    sel = uniform(seed, 5, h * w).reshape(h, w)
    img[sel < sp_density / 2] = 0
    img[(sel >= sp_density / 2) & (sel < sp_density)] = 255
    return squeeze(img)


def c4_images(n: int = 4096):
    """(u8 uniform, f32 uniform [0, 1)) n x n images, seed 4."""
    return uniform_u8((n, n), seed=4, stream=0), uniform_f32((n, n), seed=4, stream=1)


def c4_kernel(k: int = 31, seed: int = 5):
    """k x k N(0, 1) weights normalised to sum 1 in FP64, then float32.
    Returns (kernel, raw sum)."""
    raw = normal(seed, 0, k * k).reshape(k, k)
    s = raw.sum()
    return (raw / s).astype(np.float32), float(s)


# --------------------------------------------------------------------- serving

@dataclass
class Request:
    rid: int
    arrival: float  # seconds
    height: int
    width: int
    channels: int
    op: str  # "gaussian" | "median5" | "bilateral9"
    sigma: float


def serving_stream(n: int = 20000, rate: float = 1000.0, seed: int = 6, min_side: int = 64,
                   max_side: int = 4096) -> list[Request]:
    """Poisson arrivals at `rate` requests/s; W and H independently
    log-uniform in [min_side, max_side]; channels uniform in {1, 3, 4}; ops
    50% gaussian (sigma ~ U[0.5, 8]), 25% median 5x5, 25% bilateral d=9."""
    gaps = -np.log(1.0 - uniform(seed, 1, n)) / rate
    t = np.cumsum(gaps)
    lo, hi = math.log(min_side), math.log(max_side + 1)
    hs = np.exp(lo + (hi - lo) * uniform(seed, 2, n)).astype(np.int64).clip(min_side, max_side)
    ws = np.exp(lo + (hi - lo) * uniform(seed, 3, n)).astype(np.int64).clip(min_side, max_side)
    cns = np.array([1, 3, 4])[(uniform(seed, 4, n) * 3).astype(np.int64).clip(0, 2)]
    opu = uniform(seed, 5, n)
    sig = 0.5 + 7.5 * uniform(seed, 6, n)
    out = []
    for i in range(n):
        op = "gaussian" if opu[i] < 0.5 else ("median5" if opu[i] < 0.75 else "bilateral9")
        out.append(Request(i, float(t[i]), int(hs[i]), int(ws[i]), int(cns[i]), op, float(sig[i])))
    return out


def request_image(req: Request, seed: int = 6) -> np.ndarray:
    return squeeze(uniform_u8((req.height, req.width, req.channels), seed=seed, stream=100 + req.rid))


def circles_scene(h: int = 1080, w: int = 1920, n: int = 24, seed: int = 7, noise: float = 8.0) -> np.ndarray:
    """Filled discs (radius 12 .. min(h, w) / 8 + 12, grey levels 90 .. 250) on a
    40 background plus N(0, noise) noise, uint8: a houghCircles workload."""
    u = uniform(seed, 0, 4 * n)
    yy, xx = np.mgrid[0:h, 0:w]
    img = np.full((h, w), 40.0)
    for i in range(n):
        rad = 12 + u[4 * i] * (min(h, w) / 8)
        cx = rad + u[4 * i + 1] * (w - 2 * rad)
        cy = rad + u[4 * i + 2] * (h - 2 * rad)
        img[(xx - cx) ** 2 + (yy - cy) ** 2 <= rad ** 2] = 90 + 160 * u[4 * i + 3]
    img += normal(seed, 1, h * w).reshape(h, w) * noise
    return np.clip(np.rint(img), 0, 255).astype(np.uint8)
