// Separable correlation (ipf::sepFilter2D / gaussianBlur) for gfx950.
//
// Reference: include/ipf/filter.hpp:34-90 (detail::sepFilterImpl), 196-203,
// 261-267. The reference pads each source row through border tables, runs the
// horizontal taps into a full-height FP64 buffer of H+kh-1 rows, then the
// vertical taps, and rounds once with saturate_cast.
//
// Fast path (one launch, no intermediate in HBM): each workgroup owns a tile
// of TX pixels x TH rows. Input rows (tile + halo) are staged through LDS in
// chunks of 16 rows, de-interleaved into one plane per channel; the horizontal
// pass writes the FP32 row results for all TH + kh - 1 rows into an LDS tile;
// the vertical pass reads columns of that tile. Both passes keep a sliding
// window in registers (P = 4 outputs per thread horizontally, M = 8
// vertically), so each LDS value feeds several FMAs. Taps are applied in the
// reference's index order. For integer outputs the FP32 value of every sample
// is checked against its error bound; samples that might round differently
// from the reference's FP64 value are queued and recomputed exactly.
//
// Exact path: the reference's two passes in FP64 with separate multiply and
// add (the library is built with -ffp-contract=off), one thread per sample.
#include "kernels.hpp"

namespace ipfg {
namespace {

constexpr int P = 4;     // horizontal outputs per thread
constexpr int M = 8;     // vertical outputs per thread
constexpr int CHR = 16;  // staged input rows per chunk

template <int CN>
constexpr int tile_x() {
    return CN == 1 ? 64 : 32;
}

struct SepTile {
    int TH, HR, HP, XP;
};

__host__ __device__ inline SepTile sep_tile(int TX, int cn, int TH, int kw, int kh) {
    SepTile t;
    t.TH = TH;
    t.HR = TH + kh - 1;
    t.HP = TX * cn + 1;
    t.XP = (TX + kw) | 1;
    return t;
}

__host__ inline size_t sep_lds_bytes(int TX, int cn, int TH, int kw, int kh) {
    const SepTile t = sep_tile(TX, cn, TH, kw, kh);
    return sizeof(float) * (static_cast<size_t>(t.HR + 1) * t.HP + static_cast<size_t>(cn) * CHR * t.XP);
}

template <class Ti, int CN>
__global__ __launch_bounds__(kBlock) void sep_fast_kernel(
    const Ti* __restrict__ src, size_t sstride, int W, int H, int cn_rt, void* __restrict__ dst, size_t dstride,
    int dt_out, const float* __restrict__ kx, int kw, int ax, const float* __restrict__ ky, int kh, int ay, int border,
    float delta, float bval, int TH, int certify, double bound, double lo, double hi, FixList fix) {
    extern __shared__ float smem[];
    constexpr int TX = tile_x<CN>();
    constexpr int G = TX / P;
    const int cn = CN > 0 ? CN : cn_rt;
    const SepTile T = sep_tile(TX, cn, TH, kw, kh);
    float* h_tile = smem;
    float* in_tile = smem + (T.HR + 1) * T.HP;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int x0 = blockIdx.x * TX, y0 = blockIdx.y * TH;
    const int XW = TX + kw - 1;
    const int rowSamples = XW * cn;
    const bool inside_x = (x0 - ax >= 0) && (x0 - ax + XW <= W);
    const int plane = CHR * T.XP;

    for (int r0 = 0; r0 < T.HR; r0 += CHR) {
        const int nr = min(CHR, T.HR - r0);
        // ---- stage input rows [r0, r0 + nr) of the padded tile, planar per channel
        for (int r = wave; r < nr; r += kBlock / 64) {
            const int gy = y0 - ay + r0 + r;
            const int sy = (gy >= 0 && gy < H) ? gy : border_index(gy, H, border);
            float* trow = in_tile + r * T.XP;
            if (sy < 0) {
                for (int s = lane; s < rowSamples; s += 64) {
                    const int px = s / cn, c = s - px * cn;
                    trow[c * plane + px] = bval;
                }
                continue;
            }
            const Ti* srow = reinterpret_cast<const Ti*>(reinterpret_cast<const char*>(src) + static_cast<size_t>(sy) * sstride);
            if (inside_x) {
                const Ti* p = srow + static_cast<size_t>(x0 - ax) * cn;
                for (int s = lane; s < rowSamples; s += 64) {
                    const int px = s / cn, c = s - px * cn;
                    trow[c * plane + px] = static_cast<float>(p[s]);
                }
            } else {
                for (int s = lane; s < rowSamples; s += 64) {
                    const int px = s / cn, c = s - px * cn;
                    const int gx = x0 - ax + px;
                    const int sx = (gx >= 0 && gx < W) ? gx : border_index(gx, W, border);
                    trow[c * plane + px] = sx < 0 ? bval : static_cast<float>(srow[static_cast<size_t>(sx) * cn + c]);
                }
            }
        }
        __syncthreads();
        // ---- horizontal taps for the staged rows (reference order: i ascending from 0)
        const int ntask = CHR * cn * G;
        for (int t = tid; t < ntask; t += kBlock) {
            const int r = t % CHR;
            const int rest = t / CHR;
            const int g = rest % G;
            const int c = rest / G;
            if (r >= nr) continue;
            const float* base = in_tile + c * plane + r * T.XP + g * P;
            float acc[P], win[P];
#pragma unroll
            for (int m = 0; m < P; ++m) {
                acc[m] = 0.f;
                win[m] = base[m];
            }
            int i0 = 0;
            for (; i0 + P <= kw; i0 += P) {
#pragma unroll
                for (int ii = 0; ii < P; ++ii) {
                    const float k = kx[i0 + ii];
#pragma unroll
                    for (int m = 0; m < P; ++m) acc[m] = fmaf(k, win[(ii + m) % P], acc[m]);
                    win[ii] = base[i0 + ii + P];
                }
            }
#pragma unroll
            for (int ii = 0; ii < P; ++ii) {
                if (i0 + ii < kw) {
                    const float k = kx[i0 + ii];
#pragma unroll
                    for (int m = 0; m < P; ++m) acc[m] = fmaf(k, win[(ii + m) % P], acc[m]);
                    win[ii] = base[i0 + ii + P];
                }
            }
            float* hrow = h_tile + (r0 + r) * T.HP + c;
#pragma unroll
            for (int m = 0; m < P; ++m) hrow[(g * P + m) * cn] = acc[m];
        }
        __syncthreads();
    }

    // ---- vertical taps (reference order: acc = delta, j ascending)
    const int ncol = TX * cn;
    const int nv = TH / M;
    for (int t = tid; t < nv * ncol; t += kBlock) {
        const int col = t % ncol;
        const int v = t / ncol;
        const float* base = h_tile + (v * M) * T.HP + col;
        float acc[M], win[M];
#pragma unroll
        for (int m = 0; m < M; ++m) {
            acc[m] = delta;
            win[m] = base[m * T.HP];
        }
        int j0 = 0;
        for (; j0 + M <= kh; j0 += M) {
#pragma unroll
            for (int jj = 0; jj < M; ++jj) {
                const float k = ky[j0 + jj];
#pragma unroll
                for (int m = 0; m < M; ++m) acc[m] = fmaf(k, win[(jj + m) % M], acc[m]);
                win[jj] = base[(j0 + jj + M) * T.HP];
            }
        }
#pragma unroll
        for (int jj = 0; jj < M; ++jj) {
            if (j0 + jj < kh) {
                const float k = ky[j0 + jj];
#pragma unroll
                for (int m = 0; m < M; ++m) acc[m] = fmaf(k, win[(jj + m) % M], acc[m]);
                win[jj] = base[(j0 + jj + M) * T.HP];
            }
        }
        const int px = col / cn, c = col - px * cn;
        const int x = x0 + px;
        if (x >= W) continue;
        const size_t q = static_cast<size_t>(x) * cn + c;
#pragma unroll
        for (int m = 0; m < M; ++m) {
            const int y = y0 + v * M + m;
            if (y >= H) break;
            char* drow = static_cast<char*>(dst) + static_cast<size_t>(y) * dstride;
            switch (dt_out) {
                case IPFG_U8: reinterpret_cast<uint8_t*>(drow)[q] = saturate_from_double<uint8_t>(acc[m]); break;
                case IPFG_U16: reinterpret_cast<uint16_t*>(drow)[q] = saturate_from_double<uint16_t>(acc[m]); break;
                case IPFG_I16: reinterpret_cast<int16_t*>(drow)[q] = saturate_from_double<int16_t>(acc[m]); break;
                case IPFG_I32: reinterpret_cast<int32_t*>(drow)[q] = saturate_from_double<int32_t>(acc[m]); break;
                case IPFG_F32: reinterpret_cast<float*>(drow)[q] = acc[m]; break;
                case IPFG_F64: reinterpret_cast<double*>(drow)[q] = acc[m]; break;
            }
            if (certify && needs_exact(acc[m], bound, lo, hi))
                fix_push(fix, static_cast<unsigned long long>(y) * W * cn + q);
        }
    }
}

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

// One output sample in the reference's FP64 operation order.
__device__ inline double sep_exact_sample(const DevImage& src, int x, int y, int c, const double* kx, int kw,
                                          const double* ky, int kh, int ax, int ay, int border, double delta,
                                          double bval) {
    const size_t es = dtype_size(src.dt);
    double acc = delta;
    for (int j = 0; j < kh; ++j) {
        const int sy = border_index(y + j - ay, src.h, border);
        double t = 0.0;
        for (int i = 0; i < kw; ++i) {
            const int sx = border_index(x + i - ax, src.w, border);
            const double v =
                (sy < 0 || sx < 0)
                    ? bval
                    : load_sample(src.data, static_cast<size_t>(sy) * src.stride + (static_cast<size_t>(sx) * src.cn + c) * es, src.dt);
            t = t + kx[i] * v;
        }
        acc = acc + ky[j] * t;
    }
    return acc;
}

__global__ __launch_bounds__(kBlock) void sep_fix_kernel(DevImage src, DevImage dst, const double* __restrict__ kx,
                                                         int kw, const double* __restrict__ ky, int kh, int ax, int ay,
                                                         int border, double delta, double bval, FixList fix) {
    const unsigned long long n = *fix.count;
    const size_t rowLen = static_cast<size_t>(dst.w) * dst.cn;
    const bool all = n > fix.capacity;
    const unsigned long long N = all ? static_cast<unsigned long long>(dst.h) * rowLen : n;
    const size_t es = dtype_size(dst.dt);
    for (unsigned long long k = blockIdx.x * static_cast<unsigned long long>(blockDim.x) + threadIdx.x; k < N;
         k += static_cast<unsigned long long>(gridDim.x) * blockDim.x) {
        const unsigned long long s = all ? k : fix.index[k];
        const int y = static_cast<int>(s / rowLen);
        const size_t q = s - static_cast<size_t>(y) * rowLen;
        const int x = static_cast<int>(q / dst.cn), c = static_cast<int>(q - static_cast<size_t>(x) * dst.cn);
        const double v = sep_exact_sample(src, x, y, c, kx, kw, ky, kh, ax, ay, border, delta, bval);
        store_sample(dst.data, static_cast<size_t>(y) * dst.stride + q * es, dst.dt, v);
    }
}

__global__ void zero_counter_kernel(unsigned long long* c) { *c = 0ull; }

int grid_for(size_t n) {
    const size_t b = (n + kBlock - 1) / kBlock;
    return static_cast<int>(b < 8192 ? (b ? b : 1) : 8192);
}

template <class Ti, int CN>
hipError_t sep_fast_launch(const SepLaunch& a, hipStream_t s) {
    constexpr int TX = tile_x<CN>();
    const int cn = a.src.cn;
    // Tall tiles amortise the vertical halo: (TH + kh - 1) / TH <= 1.25 where possible.
    int TH = ((4 * (a.kh - 1) + M - 1) / M) * M;
    TH = TH < 32 ? 32 : (TH > 256 ? 256 : TH);
    const long long gx = (a.src.w + TX - 1) / TX;
    while (TH > 16 && gx * ((a.src.h + TH - 1) / TH) < 2048) TH -= M;
    while (TH > M && sep_lds_bytes(TX, cn, TH, a.kw, a.kh) > 64 * 1024) TH -= M;
    const size_t lds = sep_lds_bytes(TX, cn, TH, a.kw, a.kh);
    if (lds > 64 * 1024) return hipErrorInvalidConfiguration;
    const dim3 grid(static_cast<unsigned>(gx), static_cast<unsigned>((a.src.h + TH - 1) / TH));
    const int certify = a.mode == MODE_CERTIFY;
    hipLaunchKernelGGL((sep_fast_kernel<Ti, CN>), grid, dim3(kBlock), lds, s, static_cast<const Ti*>(a.src.data),
                       a.src.stride, a.src.w, a.src.h, cn, a.dst.data, a.dst.stride, a.dst.dt, a.kxf, a.kw, a.ax,
                       a.kyf, a.kh, a.ay, a.border, static_cast<float>(a.delta), static_cast<float>(a.bval), TH,
                       certify, a.bound, dtype_lo(a.dst.dt), dtype_hi(a.dst.dt), a.fix);
    return hipGetLastError();
}

template <class Ti>
hipError_t sep_fast_cn(const SepLaunch& a, hipStream_t s) {
    switch (a.src.cn) {
        case 1: return sep_fast_launch<Ti, 1>(a, s);
        case 3: return sep_fast_launch<Ti, 3>(a, s);
        case 4: return sep_fast_launch<Ti, 4>(a, s);
        default: return sep_fast_launch<Ti, 0>(a, s);
    }
}

}  // namespace

hipError_t launch_zero_counter(unsigned long long* count, hipStream_t s) {
    hipLaunchKernelGGL(zero_counter_kernel, dim3(1), dim3(1), 0, s, count);
    return hipGetLastError();
}

size_t sep_exact_tmp_bytes(int w, int h, int cn, int kh) {
    return sizeof(double) * static_cast<size_t>(h + kh - 1) * static_cast<size_t>(w) * cn;
}

hipError_t launch_sep(const SepLaunch& a, hipStream_t s) {
    const size_t rowLen = static_cast<size_t>(a.src.w) * a.src.cn;
    if (a.mode == MODE_EXACT) {
        const size_t nh = static_cast<size_t>(a.src.h + a.kh - 1) * rowLen;
        hipLaunchKernelGGL(sep_exact_h, dim3(grid_for(nh)), dim3(kBlock), 0, s, a.src, a.tmp, a.kxd, a.kw, a.ax,
                           a.ay, a.kh, a.border, a.bval);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(sep_exact_v, dim3(grid_for(static_cast<size_t>(a.src.h) * rowLen)), dim3(kBlock), 0, s,
                           a.dst, a.tmp, a.kyd, a.kh, a.delta);
        return hipGetLastError();
    }
    hipError_t e;
    if (a.mode == MODE_CERTIFY) {
        e = launch_zero_counter(a.fix.count, s);
        if (e != hipSuccess) return e;
    }
    switch (a.src.dt) {
        case IPFG_U8: e = sep_fast_cn<uint8_t>(a, s); break;
        case IPFG_U16: e = sep_fast_cn<uint16_t>(a, s); break;
        case IPFG_I16: e = sep_fast_cn<int16_t>(a, s); break;
        case IPFG_F32: e = sep_fast_cn<float>(a, s); break;
        default: return hipErrorInvalidValue;
    }
    if (e != hipSuccess || a.mode != MODE_CERTIFY) return e;
    hipLaunchKernelGGL(sep_fix_kernel, dim3(kFixBlocks), dim3(kBlock), 0, s, a.src, a.dst, a.kxd, a.kw, a.kyd, a.kh,
                       a.ax, a.ay, a.border, a.delta, a.bval, a.fix);
    return hipGetLastError();
}

}  // namespace ipfg
