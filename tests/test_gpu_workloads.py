"""GPU tests at the workloads' real sizes, and regressions for the runtime's
host bookkeeping:

- filter2D with the dense 31 x 31 kernel on 4096 x 4096 images: float32 in
  DEFAULT precision bit-identical to the reference (SHA-256 of the
  reference's output), and the split-precision MFMA paths (BF16X3, FP16X2)
  on uint8 and float32 inputs within the stated bound of the FP64 oracle
  (filter.hpp:159-194 evaluated in FP64), before and after saturation;
- the serving stream at its real sizes (up to 4096 x 4096, 1/3/4 channels)
  through the server under a 256 MB budget: every result byte-equal to the
  single-image call, a subset against the oracle, and the memory the server
  holds within its budget;
- batches run back to back on two streams; in-place calls on strided device
  views that end at their allocation's end.
"""
from __future__ import annotations

import ctypes as C
import hashlib
import json

import numpy as np
import pytest

from tests.test_gpu_parity import _mfma_bound, _mixed_jobs, _oracle_result

pytestmark = pytest.mark.gpu

FULL = json.loads(open(__file__.replace("test_gpu_workloads.py", "golden/fullsize.json")).read())


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def _log(msg):
    print(msg, flush=True)


# ----------------------------------------------------------------------- filter2D 31 x 31 on 4096^2
@pytest.fixture(scope="module")
def c4(oracle):
    from imageprocessingflex_amd import workloads as wl
    u8, f32 = wl.c4_images()
    k, ksum = wl.c4_kernel()
    oracle.set_num_threads(0)
    want = {"u8": oracle.filter2d(u8, k, border=3, out_dtype=np.float64),
            "f32": oracle.filter2d(f32, k, border=3, out_dtype=np.float64)}
    return {"u8": u8, "f32": f32, "k": k, "ksum": ksum, "want": want}


def test_full_c4_filter2d_f32_default_bit_exact(ipf, c4):
    _log("c4 float32 4096^2 filter2D DEFAULT")
    out = ipf.filter2d(c4["f32"], c4["k"])
    assert _sha(out) == FULL["c4_filter2d31_reflect101_f32"]
    _log(f"c4 f32 fix-ups {ipf.last_fixups()}")


def _saturate_u8(v):
    # saturate_cast<uint8_t>(double) (core.hpp:67-103): round half away from zero, clamp
    r = np.where(v < 0, v - 0.5, v + 0.5)
    return np.clip(np.trunc(r), 0, 255).astype(np.uint8)


@pytest.mark.parametrize("precision", [3, 4], ids=["bf16x3", "fp16x2"])
@pytest.mark.parametrize("kind", ["u8", "f32"])
def test_full_c4_mfma_error_vs_fp64_oracle(ipf, c4, precision, kind):
    img, k, want = c4[kind], c4["k"], c4["want"][kind]
    _log(f"c4 MFMA {kind} precision {precision} 4096^2")
    got = ipf.filter2d(img, k, 3, out_dtype=np.float32, precision=precision).astype(np.float64)
    err = np.abs(got - want)
    bound = _mfma_bound(k, float(np.abs(img).max()), precision, kind == "f32") + 2.0 ** -24 * np.abs(want).max()
    rms = float(np.sqrt(np.mean(err ** 2)))
    _log(f"  pre-saturation max {err.max():.3e} rms {rms:.3e} bound {bound:.3e}")
    assert err.max() <= bound
    if kind == "u8":
        # uint8 output: the reference rounds the FP64 sum once; the MFMA sum may
        # land on the other side of a rounding boundary, by one unit at most
        out = ipf.filter2d(img, k, 3, precision=precision)
        ref = _saturate_u8(want)
        d = np.abs(out.astype(np.int64) - ref.astype(np.int64))
        _log(f"  post-saturation off-by-one {int((d > 0).sum())} of {d.size}, max {int(d.max())}")
        assert d.max() <= 1
        # only samples within the bound of a rounding boundary may differ
        frac = np.abs((want + 0.5) - np.floor(want + 0.5))
        near = (np.minimum(frac, 1.0 - frac) <= bound) | (want < 0.5 + bound) | (want > 254.5 - bound)
        assert not np.any((d > 0) & ~near)
    else:
        out = ipf.filter2d(img, k, 3, precision=precision)
        assert out.dtype == np.float32
        assert np.array_equal(out, got.astype(np.float32))


# ----------------------------------------------------------------------- serving stream at real sizes
def _request_kwargs(req):
    if req.op == "gaussian":
        return "gaussian", dict(sigma=req.sigma, border=3)
    if req.op == "median5":
        return "median", dict(ksize=5, border=1)
    return "bilateral", dict(d=9, sigma_color=75.0, sigma_space=75.0, border=3)


def _single(ipf, img, op, kw):
    if op == "gaussian":
        return ipf.gaussian_blur(img, kw["sigma"], border=kw["border"])
    if op == "median":
        return ipf.median_blur(img, kw["ksize"], kw["border"])
    return ipf.bilateral_filter(img, kw["d"], kw["sigma_color"], kw["sigma_space"], kw["border"])


def test_serving_stream_real_sizes_same_alone_and_in_the_mix(ipf, oracle):
    from imageprocessingflex_amd import workloads as wl
    reqs = wl.serving_stream(n=40)
    # plus the largest requests the stream allows: 4096 x 4096 with 1, 3 and 4 channels
    extra = [wl.Request(40 + i, 0.0, 4096, 4096, cn, op, 7.9)
             for i, (cn, op) in enumerate(((4, "gaussian"), (3, "median5"), (1, "bilateral9")))]
    reqs = reqs[:40] + extra
    imgs = [wl.request_image(r) for r in reqs]
    srv = ipf.Server(256 << 20, max_batch=32, graph=True)
    outs = []
    for r, img in zip(reqs, imgs):
        op, kw = _request_kwargs(r)
        out = np.empty_like(img)
        outs.append(out)
        srv.submit(img, out, op, **kw)
    srv.drain()
    done = srv.poll()
    assert sorted(i for i, _ in done) == list(range(len(reqs)))
    held, peak = srv.memory()
    _log(f"serving: {len(reqs)} requests, held {held} B, peak in use {peak} B, stats {srv.stats()}")
    assert held <= 256 << 20 and 0 < peak <= held
    srv.close()
    px = 0
    for i, (r, img, out) in enumerate(zip(reqs, imgs, outs)):
        op, kw = _request_kwargs(r)
        assert np.array_equal(out, _single(ipf, img, op, kw)), (i, r)
        # against the oracle: every request up to 1.5 Mpx and the three largest
        if img.shape[0] * img.shape[1] <= 1_500_000 or i >= 40:
            assert np.array_equal(out, _oracle_result(oracle, img, op, kw)), (i, r)
            px += img.shape[0] * img.shape[1]
    _log(f"serving: {px} px checked against the oracle")


def test_server_memory_within_budget_under_many_small_requests(ipf, oracle):
    # 400 host requests of up to 257 x 300 x 4 against a 3 MB budget: the
    # server holds one 3 MB block and never places more than that
    rng = np.random.default_rng(77)
    budget = 3 << 20
    srv = ipf.Server(budget, max_batch=64, graph=True)
    items, outs = [], []
    for i in range(400):
        h, w = int(rng.integers(1, 258)), int(rng.integers(1, 301))
        cn = int(rng.choice([1, 3, 4]))
        img = rng.integers(0, 256, (h, w, cn) if cn > 1 else (h, w), dtype=np.uint8)
        op = ("gaussian", dict(sigma=1.5, border=i % 5)) if i % 2 else ("median", dict(ksize=3, border=1 + i % 4))
        out = np.empty_like(img)
        items.append((img, *op))
        outs.append(out)
        srv.submit(img, out, op[0], **op[1])
    srv.drain()
    assert len(srv.poll(1024)) == 400
    held, peak = srv.memory()
    assert held == budget and 0 < peak <= budget
    batches, biggest = srv.stats()
    _log(f"small requests: {batches} batches, largest {biggest}, peak {peak} of {held}")
    srv.close()
    for i, ((img, op, kw), out) in enumerate(zip(items, outs)):
        assert np.array_equal(out, _oracle_result(oracle, img, op, kw)), i


# ----------------------------------------------------------------------- batches on two streams
def test_batch_runs_back_to_back_on_two_streams(ipf, oracle):
    rng = np.random.default_rng(43)
    first, second = _mixed_jobs(rng, 22), _mixed_jobs(rng, 22)[::-1]
    s1, s2 = ipf.Stream(), ipf.Stream()
    b = ipf.Batch(32, graph=True)
    runs = []
    for items, s in ((first, s1), (second, s2), (first, s2), (second, s1)):
        srcs = [ipf.DeviceImage.from_array(img, s) for img, _, _ in items]
        dsts = [ipf.DeviceImage.like(d) for d in srcs]
        s.synchronize()  # uploads done; the runs below are not ordered by the host
        runs.append((items, srcs, dsts, s))
    for items, srcs, dsts, s in runs:
        b.run([ipf.job(d, o, op, **kw) for d, o, (_, op, kw) in zip(srcs, dsts, items)], stream=s)
    s1.synchronize()
    s2.synchronize()
    for r, (items, _, dsts, s) in enumerate(runs):
        for i, ((img, op, kw), o) in enumerate(zip(items, dsts)):
            want = _oracle_result(oracle, img, op, kw)
            assert np.array_equal(o.download(s).reshape(want.shape), want), (r, i, op, kw)
    b.close()


# ----------------------------------------------------------------------- in-place strided views
@pytest.mark.parametrize("op", ["gaussian", "median", "box"])
def test_in_place_view_ending_at_allocation_end(ipf, oracle, op):
    # a region of interest: the bottom 41 rows of a 120 x 200 x 3 parent, columns
    # [30, 180), filtered in place through a descriptor with the parent's stride;
    # its last row ends 60 bytes before the parent's last row does
    from imageprocessingflex_amd import _lib
    rng = np.random.default_rng(91)
    parent = rng.integers(0, 256, (120, 200, 3), dtype=np.uint8)
    d = ipf.DeviceImage.from_array(parent)
    y0, x0, h, w, cn = 79, 30, 41, 150, 3
    stride = 200 * cn
    base = d.ptr.value + y0 * stride + x0 * cn
    view = _lib.IpfgImage(C.c_void_p(base), w, h, cn, 0, stride, 1)
    roi = parent[y0:y0 + h, x0:x0 + w].copy()
    lib = _lib.load()
    if op == "gaussian":
        _lib.check(lib.ipfg_gaussian_blur(C.byref(view), C.byref(view), 1.7, 0.0, 0, 3, 0, None))
        want = oracle.gaussian_blur(roi, 1.7)
    elif op == "median":
        _lib.check(lib.ipfg_median(C.byref(view), C.byref(view), 5, 1, None))
        want = oracle.median(roi, 5, 1)
    else:
        _lib.check(lib.ipfg_box_filter(C.byref(view), C.byref(view), 5, 3, 1, 2, 0.0, 0, None))
        want = oracle.box_filter(roi, 5, 3, True, 2)
    ipf.synchronize()
    got = d.download()
    assert np.array_equal(got[y0:y0 + h, x0:x0 + w], want)
    outside = np.ones(parent.shape[:2], bool)
    outside[y0:y0 + h, x0:x0 + w] = False
    assert np.array_equal(got[outside], parent[outside])
