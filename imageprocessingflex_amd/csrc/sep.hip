// Separable correlation (ipf::sepFilter2D / gaussianBlur) for gfx950.
//
// Reference: include/ipf/filter.hpp:34-90 (detail::sepFilterImpl), 196-203,
// 261-267. The reference pads each source row through border tables, runs the
// horizontal taps into a full-height FP64 buffer of H+kh-1 rows, then the
// vertical taps (starting from delta), and rounds once with saturate_cast.
//
// Fast path (one launch, no intermediate in HBM): a workgroup walks down a
// strip of TX pixels (64 for one channel, 32 otherwise) x SH rows, 16 rows per
// step. Each step stages 16 border-padded input rows into LDS (the loads of
// the next step are issued before this step computes), runs the horizontal
// taps into a ring of rows in LDS and, once kh - 1 rows of history are in the
// ring, the vertical taps for 16 output rows. Both passes keep a sliding
// register window (P = 4 outputs per thread horizontally, M = 4 vertically)
// so every LDS value feeds several FMAs, and prefetch weights as wave-uniform
// scalars. Taps are applied in the reference's index order.
//
// Arithmetic: Acc = double by default. MI355X issues FP64 FMA at the rate of
// (non-packed) FP32 FMA, and the FP64 result differs from the reference's
// separate multiply/add chain by a tiny bound, so every output -- integer or
// float32 -- whose value is not within that bound of a rounding boundary is
// certified bit-identical; the few others are recomputed in the reference's
// exact order by the last workgroup of the same launch. Acc = float is the
// FP32 option (no certification).
//
// Exact path: the reference's two passes in FP64 with separate multiply and
// add (the library is built with -ffp-contract=off), one thread per sample.
#include "sep1.hpp"

namespace ipfg {
namespace {

using sepk::SepArgs;

// ------------------------------------------------------------------ exact
__global__ __launch_bounds__(kBlock) void sep_exact_h(DevImage src, double* __restrict__ tmp,
                                                      const double* __restrict__ kx, int kw, int ax, int ay, int kh,
                                                      int border, double bval) {
    const int W = src.w, H = src.h, cn = src.cn;
    const size_t rowLen = static_cast<size_t>(W) * cn;
    const size_t total = static_cast<size_t>(H + kh - 1) * rowLen;
    const size_t es = dtype_size(src.dt);
    for (size_t idx = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; idx < total;
         idx += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const int py = static_cast<int>(idx / rowLen);
        const int q = static_cast<int>(idx - static_cast<size_t>(py) * rowLen);
        const int x = q / cn, c = q - x * cn;
        const int sy = border_index(py - ay, H, border);
        double t = 0.0;
        for (int i = 0; i < kw; ++i) {
            const int sx = border_index(x + i - ax, W, border);
            const double v = (sy < 0 || sx < 0)
                                 ? bval
                                 : load_sample(src.data, static_cast<size_t>(sy) * src.stride + (static_cast<size_t>(sx) * cn + c) * es, src.dt);
            t = t + kx[i] * v;
        }
        tmp[idx] = t;
    }
}

__global__ __launch_bounds__(kBlock) void sep_exact_v(DevImage dst, const double* __restrict__ tmp,
                                                      const double* __restrict__ ky, int kh, double delta) {
    const int W = dst.w, H = dst.h, cn = dst.cn;
    const size_t rowLen = static_cast<size_t>(W) * cn;
    const size_t total = static_cast<size_t>(H) * rowLen;
    const size_t es = dtype_size(dst.dt);
    for (size_t idx = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; idx < total;
         idx += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const int y = static_cast<int>(idx / rowLen);
        const size_t q = idx - static_cast<size_t>(y) * rowLen;
        double acc = delta;
        for (int j = 0; j < kh; ++j) acc = acc + ky[j] * tmp[static_cast<size_t>(y + j) * rowLen + q];
        store_sample(dst.data, static_cast<size_t>(y) * dst.stride + q * es, dst.dt, acc);
    }
}

__global__ void zero_counter_kernel(unsigned long long* c, unsigned int* d) {
    *c = 0ull;
    *d = 0u;
}

int sep_grid(size_t n) {
    const size_t b = (n + kBlock - 1) / kBlock;
    return static_cast<int>(b < 8192 ? (b ? b : 1) : 8192);
}


}  // namespace

using namespace sepk;

// The one-channel float32 certified kernel (sep1.hpp): kw = kh odd in
// [3, 49], centred anchors, float-aligned rows.
static bool sep1_fits(const DevImage& src, const DevImage& dst, int kw, int kh, int ax, int ay) {
    return src.cn == 1 && src.dt == IPFG_F32 && dst.dt == IPFG_F32 && kw == kh && kw % 2 == 1 && kw >= 3 &&
           kw <= 49 && ax == kw / 2 && ay == kh / 2 && src.stride % 4 == 0 && dst.stride % 4 == 0 && src.w > 0 &&
           src.h > 0 && static_cast<size_t>(src.h) * src.stride < (size_t(1) << 31) &&
           static_cast<size_t>(dst.h) * dst.stride < (size_t(1) << 31);
}

static hipError_t sep1_launch(const SepArgs& a, hipStream_t s) {
    if (!sep1_fits(a.src, a.dst, a.kw, a.kh, a.ax, a.ay)) return hipErrorNotSupported;
    if (a.kw <= 13) return sep1::launch_k3_13(a, s);
    if (a.kw <= 25) return sep1::launch_k15_25(a, s);
    return a.kw <= 37 ? sep1::launch_k27_37(a, s) : sep1::launch_k39_49(a, s);
}

static bool sep_planar(int cn, int src_dt, int dst_dt) {
    return (cn == 1 || cn == 3 || cn == 4) &&
           ((src_dt == IPFG_U8 && (dst_dt == IPFG_U8 || dst_dt == IPFG_F32)) || (src_dt == IPFG_F32 && dst_dt == IPFG_F32));
}

bool sep_fast_fits(int cn, int kw, int kh, int src_dt, int dst_dt, bool cert) {
    const size_t acc = cert ? sizeof(double) : sizeof(float);
    if (sep_planar(cn, src_dt, dst_dt)) {
        const size_t hr = (cert && !plane_ring_is_float(dst_dt)) ? sizeof(double) : sizeof(float);
        return plane_lds_bytes(cn, kw, kh, hr) <= 156 * 1024 && (PTX + kw - 1) * cn <= 64 * 8;
    }
    return sep_lds_bytes(32, cn, kw, kh, acc) <= 156 * 1024 && (32 + kw - 1) * cn <= 64 * 8;
}

hipError_t launch_zero_counter(unsigned long long* count, unsigned int* done, hipStream_t s) {
    hipLaunchKernelGGL(zero_counter_kernel, dim3(1), dim3(1), 0, s, count, done);
    return hipGetLastError();
}

std::vector<int> sep_sh_candidates(const SepLaunch& l) {
    const int h = l.src.h;
    std::vector<int> c;
    if (l.mode == MODE_CERTIFY && sep1_fits(l.src, l.dst, l.kw, l.kh, l.ax, l.ay)) {
        // the one-channel float32 kernel: 16-row steps, then 8-row steps;
        // short strips give small images enough workgroups to fill the CUs
        for (int v : {16, 32, 64, 128, 256, 512})
            if (v < 2 * h) c.push_back(v);
        for (int v : {8, 16, 32, 64, 128, 256})
            if (v < 2 * h) c.push_back(sep1::kHalfStep + v);
        return c;
    }
    for (int v : {16, 32, 48, 64, 96, 128, 192, 256, 384, 512})
        if (v < 2 * h) c.push_back(v);
    if (l.mode == MODE_CERTIFY && tile_u8_fits(l.src, l.dst, l.kw, l.kh, l.ax, l.ay, l.unit != 0)) {
        c.push_back(kTileSH + 16);  // the one-pass tile kernel, 16- and 32-row tiles
        c.push_back(kTileSH + 32);
    }
    return c;
}

size_t sep_exact_tmp_bytes(int w, int h, int cn, int kh) {
    return sizeof(double) * static_cast<size_t>(h + kh - 1) * static_cast<size_t>(w) * cn;
}

hipError_t launch_sep(const SepLaunch& l, hipStream_t s) {
    const size_t rowLen = static_cast<size_t>(l.src.w) * l.src.cn;
    if (l.mode == MODE_EXACT) {
        const size_t nh = static_cast<size_t>(l.src.h + l.kh - 1) * rowLen;
        hipLaunchKernelGGL(sep_exact_h, dim3(sep_grid(nh)), dim3(kBlock), 0, s, l.src, l.tmp, l.kxd, l.kw, l.ax,
                           l.ay, l.kh, l.border, l.bval);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(sep_exact_v, dim3(sep_grid(static_cast<size_t>(l.src.h) * rowLen)), dim3(kBlock), 0, s,
                           l.dst, l.tmp, l.kyd, l.kh, l.delta);
        return hipGetLastError();
    }
    SepArgs a{};
    a.src = l.src;
    a.dst = l.dst;
    const bool cert = l.mode == MODE_CERTIFY;
    a.kx = cert ? static_cast<const void*>(l.kxd) : static_cast<const void*>(l.kxf);
    a.ky = cert ? static_cast<const void*>(l.kyd) : static_cast<const void*>(l.kyf);
    a.kxd = l.kxd;
    a.kyd = l.kyd;
    a.kw = l.kw;
    a.kh = l.kh;
    a.ax = l.ax;
    a.ay = l.ay;
    a.border = l.border;
    a.delta = l.delta;
    a.bval = l.bval;
    a.bcoef = l.bcoef;
    a.bdelta = l.bdelta;
    a.xbound = l.xbound;
    a.fix = l.fix;
    a.SH = l.sh;
    a.oscale = l.oscale != 0.0 ? l.oscale : 1.0;
    a.unit = l.unit;
    a.kabs = l.kabs;
    int tile_rows = a.SH >= kTileSH ? (a.SH == kTileSH ? 32 : a.SH - kTileSH) : 0;
    if (tile_rows) a.SH = 0;  // the strip kernels' own height when the tile kernel does not take the call
    // Without a measured shape the tile kernel takes the calls it fits: it was
    // faster than the strip kernels on most measured uint8 shapes, with
    // 16-row tiles up to a megapixel and 32-row tiles above.
    if (l.sh <= 0) tile_rows = static_cast<long long>(l.src.w) * l.src.h <= (1LL << 20) ? 16 : 32;
    const int cn = l.src.cn;
    const bool planar = sep_planar(cn, l.src.dt, l.dst.dt);
    if (planar && l.src.dt == IPFG_U8) {
        if (cert && cn == 1 && l.dst.dt == IPFG_F32) {
            const hipError_t e = plane_u8f_k1(a, s);
            if (e != hipErrorNotSupported) return e;
            (void)hipGetLastError();
        }
        if (cert && plane_ring_is_float(l.dst.dt)) {
            // the ring holds the FP64 horizontal sums rounded to float
            a.bcoef += 1.01 * 5.9604644775390625e-08 * l.kabs;
            if (l.dst.dt == IPFG_U8 && tile_rows) {
                const hipError_t e = tile_u8(a, tile_rows, s);
                if (e != hipErrorNotSupported) return e;
                (void)hipGetLastError();
            }
            if (l.dst.dt == IPFG_U8) {
                const hipError_t e = cn == 1 ? plane_u8_k1(a, s) : cn == 3 ? plane_u8_k3(a, s) : plane_u8_k4(a, s);
                if (e != hipErrorNotSupported) return e;
                (void)hipGetLastError();
            }
        }
        return plane_u8(a, cert, s);
    }
    if (planar) {  // float32 -> float32
        if (cert && cn == 1) {
            hipError_t e = sep1_launch(a, s);
            if (e != hipErrorNotSupported) return e;
            (void)hipGetLastError();
            e = plane_f32_k1(a, s);
            if (e != hipErrorNotSupported) return e;
            (void)hipGetLastError();
        }
        return plane_f32(a, cert, s);
    }
    return generic_any(a, cert, s);
}

}  // namespace ipfg
