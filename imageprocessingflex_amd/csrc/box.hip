// Box filter (ipf::boxFilter / blur) for gfx950.
//
// Reference: include/ipf/filter.hpp:205-259. The reference forms horizontal
// window sums with a running sum per row in FP64, then vertical running sums
// per column, and rounds saturate_cast(sum * (1 / (kw * kh))).
//
// Fast path: a workgroup tile of TX pixels x TH rows; input rows (tile + halo)
// are staged through LDS per channel plane; each thread forms P horizontal
// window sums with a running sum seeded by a direct sum, the vertical pass
// forms M column sums the same way. Accumulation is FP64, so for integer
// images (and integer border values) every sum is an exact integer and the
// result is bit-identical to the reference. For float images the sums differ
// from the reference's left-to-right running sums only by FP64 rounding.
//
// Exact path: the reference's running sums in its order, one thread per row
// (horizontal) and per column (vertical).
#include "kernels.hpp"

namespace ipfg {
namespace {

constexpr int P = 8;     // horizontal outputs per thread
constexpr int M = 8;     // vertical outputs per thread
constexpr int CHR = 8;   // staged input rows per chunk
constexpr int TX = 64;   // tile width in pixels

struct BoxTile {
    int HR, HP, XP;
};

__host__ __device__ inline BoxTile box_tile(int cn, int TH, int kw, int kh) {
    BoxTile t;
    t.HR = TH + kh - 1;
    t.HP = TX * cn + 1;
    t.XP = (TX + kw - 1) | 1;
    return t;
}

__host__ inline size_t box_lds_bytes(int cn, int TH, int kw, int kh) {
    const BoxTile t = box_tile(cn, TH, kw, kh);
    return sizeof(double) * (static_cast<size_t>(t.HR) * t.HP + static_cast<size_t>(cn) * CHR * t.XP);
}

template <class Ti>
__global__ __launch_bounds__(kBlock) void box_fast_kernel(const Ti* __restrict__ src, size_t sstride, int W, int H,
                                                          int cn, void* __restrict__ dst, size_t dstride, int dt_out,
                                                          int kw, int kh, int border, double bval, double scale,
                                                          int TH) {
    extern __shared__ double dsm[];
    const BoxTile T = box_tile(cn, TH, kw, kh);
    double* h_tile = dsm;
    double* in_tile = dsm + static_cast<size_t>(T.HR) * T.HP;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int ax = kw / 2, ay = kh / 2;
    const int x0 = blockIdx.x * TX, y0 = blockIdx.y * TH;
    const int XW = TX + kw - 1;
    const int rowSamples = XW * cn;
    const bool inside_x = (x0 - ax >= 0) && (x0 - ax + XW <= W);
    const int plane = CHR * T.XP;
    constexpr int G = TX / P;

    for (int r0 = 0; r0 < T.HR; r0 += CHR) {
        const int nr = min(CHR, T.HR - r0);
        for (int r = wave; r < nr; r += kBlock / 64) {
            const int gy = y0 - ay + r0 + r;
            const int sy = (gy >= 0 && gy < H) ? gy : border_index(gy, H, border);
            double* trow = in_tile + r * T.XP;
            if (sy < 0) {
                for (int s = lane; s < rowSamples; s += 64) {
                    const int px = s / cn, c = s - px * cn;
                    trow[c * plane + px] = bval;
                }
                continue;
            }
            const Ti* srow = reinterpret_cast<const Ti*>(reinterpret_cast<const char*>(src) + static_cast<size_t>(sy) * sstride);
            for (int s = lane; s < rowSamples; s += 64) {
                const int px = s / cn, c = s - px * cn;
                double v;
                if (inside_x) {
                    v = static_cast<double>(srow[static_cast<size_t>(x0 - ax) * cn + s]);
                } else {
                    const int gx = x0 - ax + px;
                    const int sx = (gx >= 0 && gx < W) ? gx : border_index(gx, W, border);
                    v = sx < 0 ? bval : static_cast<double>(srow[static_cast<size_t>(sx) * cn + c]);
                }
                trow[c * plane + px] = v;
            }
        }
        __syncthreads();
        const int ntask = CHR * cn * G;
        for (int t = tid; t < ntask; t += kBlock) {
            const int r = t % CHR;
            const int rest = t / CHR;
            const int g = rest % G;
            const int c = rest / G;
            if (r >= nr) continue;
            const double* base = in_tile + c * plane + r * T.XP + g * P;
            double s = 0.0;
            for (int i = 0; i < kw; ++i) s += base[i];
            double* hrow = h_tile + (r0 + r) * T.HP + c;
            hrow[(g * P) * cn] = s;
#pragma unroll
            for (int m = 1; m < P; ++m) {
                s += base[m + kw - 1] - base[m - 1];
                hrow[(g * P + m) * cn] = s;
            }
        }
        __syncthreads();
    }
    const int ncol = TX * cn;
    const int nv = TH / M;
    for (int t = tid; t < nv * ncol; t += kBlock) {
        const int col = t % ncol;
        const int v = t / ncol;
        const double* base = h_tile + (v * M) * T.HP + col;
        const int px = col / cn, c = col - px * cn;
        const int x = x0 + px;
        double s = 0.0;
        for (int j = 0; j < kh; ++j) s += base[j * T.HP];
        if (x >= W) continue;
        const size_t q = static_cast<size_t>(x) * cn + c;
        const size_t es = dtype_size(dt_out);
#pragma unroll
        for (int m = 0; m < M; ++m) {
            const int y = y0 + v * M + m;
            if (y >= H) break;
            if (m > 0) s += base[(m + kh - 1) * T.HP] - base[(m - 1) * T.HP];
            store_sample(dst, static_cast<size_t>(y) * dstride + q * es, dt_out, s * scale);
        }
    }
}

// ------------------------------------------------------------------ exact
// Horizontal running sums for one (padded row, channel) per thread, in the
// reference's order (filter.hpp:220-237).
__global__ __launch_bounds__(kBlock) void box_exact_h(DevImage src, double* __restrict__ tmp, int kw, int kh,
                                                      int border, double bval) {
    const int W = src.w, H = src.h, cn = src.cn;
    const int ax = kw / 2, ay = kh / 2;
    const int PH = H + kh - 1;
    const size_t es = dtype_size(src.dt);
    const size_t rowLen = static_cast<size_t>(W) * cn;
    const long long n = static_cast<long long>(PH) * cn;
    for (long long idx = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; idx < n;
         idx += static_cast<long long>(gridDim.x) * blockDim.x) {
        const int py = static_cast<int>(idx / cn), c = static_cast<int>(idx - static_cast<long long>(py) * cn);
        const int sy = border_index(py - ay, H, border);
        auto val = [&](int px) -> double {
            const int sx = border_index(px - ax, W, border);
            if (sy < 0 || sx < 0) return bval;
            return load_sample(src.data, static_cast<size_t>(sy) * src.stride + (static_cast<size_t>(sx) * cn + c) * es,
                               src.dt);
        };
        double* out = tmp + static_cast<size_t>(py) * rowLen + c;
        double s = 0.0;
        for (int i = 0; i < kw; ++i) s = s + val(i);
        out[0] = s;
        for (int x = 1; x < W; ++x) {
            const double step = val(x + kw - 1) - val(x - 1);
            s = s + step;
            out[static_cast<size_t>(x) * cn] = s;
        }
    }
}

// Vertical running sums per column in the reference's order (filter.hpp:240-251).
__global__ __launch_bounds__(kBlock) void box_exact_v(DevImage dst, const double* __restrict__ tmp, int kh,
                                                      double scale) {
    const int W = dst.w, H = dst.h, cn = dst.cn;
    const size_t rowLen = static_cast<size_t>(W) * cn;
    const size_t es = dtype_size(dst.dt);
    for (size_t q = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; q < rowLen;
         q += static_cast<size_t>(gridDim.x) * blockDim.x) {
        double col = 0.0;
        for (int j = 0; j < kh; ++j) col = col + tmp[static_cast<size_t>(j) * rowLen + q];
        for (int y = 0; y < H; ++y) {
            store_sample(dst.data, static_cast<size_t>(y) * dst.stride + q * es, dst.dt, col * scale);
            if (y + 1 < H) {
                const double step = tmp[static_cast<size_t>(y + kh) * rowLen + q] - tmp[static_cast<size_t>(y) * rowLen + q];
                col = col + step;
            }
        }
    }
}

int grid_for(size_t n) {
    const size_t b = (n + kBlock - 1) / kBlock;
    return static_cast<int>(b < 8192 ? (b ? b : 1) : 8192);
}

template <class Ti>
hipError_t box_fast_launch(const BoxLaunch& a, double scale, hipStream_t s) {
    const int cn = a.src.cn;
    int TH = ((2 * (a.kh - 1) + M - 1) / M) * M;
    TH = TH < 32 ? 32 : (TH > 128 ? 128 : TH);
    const long long gx = (a.src.w + TX - 1) / TX;
    while (TH > 16 && gx * ((a.src.h + TH - 1) / TH) < 2048) TH -= M;
    while (TH > M && box_lds_bytes(cn, TH, a.kw, a.kh) > 64 * 1024) TH -= M;
    const size_t lds = box_lds_bytes(cn, TH, a.kw, a.kh);
    if (lds > 64 * 1024) return hipErrorInvalidConfiguration;
    const dim3 grid(static_cast<unsigned>(gx), static_cast<unsigned>((a.src.h + TH - 1) / TH));
    hipLaunchKernelGGL(box_fast_kernel<Ti>, grid, dim3(kBlock), lds, s, static_cast<const Ti*>(a.src.data),
                       a.src.stride, a.src.w, a.src.h, cn, a.dst.data, a.dst.stride, a.dst.dt, a.kw, a.kh, a.border,
                       a.bval, scale, TH);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_box(const BoxLaunch& a, hipStream_t s) {
    const double scale = a.normalize ? 1.0 / (static_cast<double>(a.kw) * a.kh) : 1.0;
    if (a.mode == MODE_EXACT) {
        const size_t rows = static_cast<size_t>(a.src.h + a.kh - 1) * a.src.cn;
        hipLaunchKernelGGL(box_exact_h, dim3(grid_for(rows)), dim3(kBlock), 0, s, a.src, a.tmp, a.kw, a.kh, a.border,
                           a.bval);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(box_exact_v, dim3(grid_for(static_cast<size_t>(a.src.w) * a.src.cn)), dim3(kBlock), 0, s,
                           a.dst, a.tmp, a.kh, scale);
        return hipGetLastError();
    }
    switch (a.src.dt) {
        case IPFG_U8: return box_fast_launch<uint8_t>(a, scale, s);
        case IPFG_U16: return box_fast_launch<uint16_t>(a, scale, s);
        case IPFG_I16: return box_fast_launch<int16_t>(a, scale, s);
        case IPFG_I32: return box_fast_launch<int32_t>(a, scale, s);
        case IPFG_F32: return box_fast_launch<float>(a, scale, s);
        case IPFG_F64: return box_fast_launch<double>(a, scale, s);
    }
    return hipErrorInvalidValue;
}

}  // namespace ipfg
