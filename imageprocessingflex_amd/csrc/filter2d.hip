// filter2D (generic 2-D correlation) for gfx950.
//
// Reference: include/ipf/filter.hpp:157-194. Only kernel weights that compare
// unequal to zero take part, in row-major order; the sum starts at delta and is
// rounded once with saturate_cast.
//
// Fast path (FP32): a workgroup tile of TX pixels x TH rows with the input
// tile (plus kernel halo) staged in LDS, one plane per channel. Each thread
// produces P = 4 horizontally adjacent outputs, walking the kernel rows and
// sliding a register window along each row, so every LDS value feeds P FMAs.
// Weights are read as wave-uniform scalars. Zero weights are skipped when the
// kernel has any (a uniform branch), matching the reference's tap list for
// non-finite inputs. Integer outputs are certified with a per-sample bound
// from the accumulated |w| * |v|; doubtful samples are recomputed in FP64.
//
// Exact path: the reference's tap list in FP64, one thread per sample.
#include "kernels.hpp"

namespace ipfg {
namespace {

constexpr int P = 4;
constexpr int TX = 64;

__host__ __device__ inline int f2d_xp(int kw) { return (TX + kw - 1 + P) | 1; }

__host__ inline size_t f2d_lds_bytes(int cn, int TH, int kw, int kh) {
    return sizeof(float) * static_cast<size_t>(cn) * (TH + kh - 1) * f2d_xp(kw);
}

template <class Ti, bool HZ, bool CERT>
__global__ __launch_bounds__(kBlock) void f2d_fast_kernel(const Ti* __restrict__ src, size_t sstride, int W, int H,
                                                          int cn, void* __restrict__ dst, size_t dstride, int dt_out,
                                                          const float* __restrict__ K, int kw, int kh, int ax, int ay,
                                                          int border, float delta, float bval, int TH,
                                                          double bscale, double lo, double hi, FixList fix) {
    extern __shared__ float smem[];
    const int XP = f2d_xp(kw);
    const int HR = TH + kh - 1;
    const int plane = HR * XP;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int x0 = blockIdx.x * TX, y0 = blockIdx.y * TH;
    const int XW = TX + kw - 1;
    const int rowSamples = XW * cn;
    const bool inside_x = (x0 - ax >= 0) && (x0 - ax + XW <= W);
    for (int r = wave; r < HR; r += kBlock / 64) {
        const int gy = y0 - ay + r;
        const int sy = (gy >= 0 && gy < H) ? gy : border_index(gy, H, border);
        float* trow = smem + r * XP;
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
            float v;
            if (inside_x) {
                v = static_cast<float>(srow[static_cast<size_t>(x0 - ax) * cn + s]);
            } else {
                const int gx = x0 - ax + px;
                const int sx = (gx >= 0 && gx < W) ? gx : border_index(gx, W, border);
                v = sx < 0 ? bval : static_cast<float>(srow[static_cast<size_t>(sx) * cn + c]);
            }
            trow[c * plane + px] = v;
        }
    }
    __syncthreads();
    constexpr int G = TX / P;
    const int ntask = TH * G * cn;
    for (int t = tid; t < ntask; t += kBlock) {
        const int g = t % G;
        const int rest = t / G;
        const int r = rest % TH;
        const int c = rest / TH;
        float acc[P], aab[P];
#pragma unroll
        for (int m = 0; m < P; ++m) {
            acc[m] = delta;
            aab[m] = 0.f;
        }
        for (int j = 0; j < kh; ++j) {
            const float* base = smem + c * plane + (r + j) * XP + g * P;
            const float* krow = K + j * kw;
            float win[P];
#pragma unroll
            for (int m = 0; m < P; ++m) win[m] = base[m];
            int i0 = 0;
            for (; i0 + P <= kw; i0 += P) {
#pragma unroll
                for (int ii = 0; ii < P; ++ii) {
                    const float k = krow[i0 + ii];
                    if (!HZ || k != 0.f) {
#pragma unroll
                        for (int m = 0; m < P; ++m) {
                            acc[m] = fmaf(k, win[(ii + m) % P], acc[m]);
                            if (CERT) aab[m] = fmaf(fabsf(k), fabsf(win[(ii + m) % P]), aab[m]);
                        }
                    }
                    win[ii] = base[i0 + ii + P];
                }
            }
#pragma unroll
            for (int ii = 0; ii < P; ++ii) {
                if (i0 + ii < kw) {
                    const float k = krow[i0 + ii];
                    if (!HZ || k != 0.f) {
#pragma unroll
                        for (int m = 0; m < P; ++m) {
                            acc[m] = fmaf(k, win[(ii + m) % P], acc[m]);
                            if (CERT) aab[m] = fmaf(fabsf(k), fabsf(win[(ii + m) % P]), aab[m]);
                        }
                    }
                    win[ii] = base[i0 + ii + P];
                }
            }
        }
        const int y = y0 + r;
        if (y >= H) continue;
        char* drow = static_cast<char*>(dst) + static_cast<size_t>(y) * dstride;
#pragma unroll
        for (int m = 0; m < P; ++m) {
            const int x = x0 + g * P + m;
            if (x >= W) break;
            const size_t q = static_cast<size_t>(x) * cn + c;
            switch (dt_out) {
                case IPFG_U8: reinterpret_cast<uint8_t*>(drow)[q] = saturate_from_double<uint8_t>(acc[m]); break;
                case IPFG_U16: reinterpret_cast<uint16_t*>(drow)[q] = saturate_from_double<uint16_t>(acc[m]); break;
                case IPFG_I16: reinterpret_cast<int16_t*>(drow)[q] = saturate_from_double<int16_t>(acc[m]); break;
                case IPFG_I32: reinterpret_cast<int32_t*>(drow)[q] = saturate_from_double<int32_t>(acc[m]); break;
                case IPFG_F32: reinterpret_cast<float*>(drow)[q] = acc[m]; break;
                case IPFG_F64: reinterpret_cast<double*>(drow)[q] = acc[m]; break;
            }
            if (CERT) {
                // |fp32 sum - FP64 reference| <= bscale * (sum of |w||v| + |delta|)
                const double bound = bscale * (static_cast<double>(aab[m]) + fabs(static_cast<double>(delta)));
                if (needs_exact(acc[m], bound, lo, hi)) fix_push(fix, static_cast<unsigned long long>(y) * W * cn + q);
            }
        }
    }
}

// ------------------------------------------------------------------ exact
__device__ inline double f2d_exact_sample(const DevImage& src, int x, int y, int c, const int* taps,
                                          const double* tw, int ntaps, int ax, int ay, int border, double delta,
                                          double bval) {
    const size_t es = dtype_size(src.dt);
    double s = delta;
    for (int t = 0; t < ntaps; ++t) {
        const int sx = border_index(x + taps[2 * t] - ax, src.w, border);
        const int sy = border_index(y + taps[2 * t + 1] - ay, src.h, border);
        const double v = (sx < 0 || sy < 0)
                             ? bval
                             : load_sample(src.data, static_cast<size_t>(sy) * src.stride + (static_cast<size_t>(sx) * src.cn + c) * es, src.dt);
        s = s + tw[t] * v;
    }
    return s;
}

__global__ __launch_bounds__(kBlock) void f2d_exact_kernel(DevImage src, DevImage dst, const int* __restrict__ taps,
                                                           const double* __restrict__ tw, int ntaps, int ax, int ay,
                                                           int border, double delta, double bval) {
    const size_t rowLen = static_cast<size_t>(dst.w) * dst.cn;
    const size_t total = rowLen * dst.h;
    const size_t es = dtype_size(dst.dt);
    for (size_t idx = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; idx < total;
         idx += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const int y = static_cast<int>(idx / rowLen);
        const size_t q = idx - static_cast<size_t>(y) * rowLen;
        const int x = static_cast<int>(q / dst.cn), c = static_cast<int>(q - static_cast<size_t>(x) * dst.cn);
        const double v = f2d_exact_sample(src, x, y, c, taps, tw, ntaps, ax, ay, border, delta, bval);
        store_sample(dst.data, static_cast<size_t>(y) * dst.stride + q * es, dst.dt, v);
    }
}

__global__ __launch_bounds__(kBlock) void f2d_fix_kernel(DevImage src, DevImage dst, const int* __restrict__ taps,
                                                         const double* __restrict__ tw, int ntaps, int ax, int ay,
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
        const double v = f2d_exact_sample(src, x, y, c, taps, tw, ntaps, ax, ay, border, delta, bval);
        store_sample(dst.data, static_cast<size_t>(y) * dst.stride + q * es, dst.dt, v);
    }
}

int grid_for(size_t n) {
    const size_t b = (n + kBlock - 1) / kBlock;
    return static_cast<int>(b < 16384 ? (b ? b : 1) : 16384);
}

template <class Ti, bool HZ, bool CERT>
hipError_t f2d_fast_launch(const F2dLaunch& a, hipStream_t s) {
    const int cn = a.src.cn;
    // rows per tile: at least 4 tasks per thread of work, LDS within 64 KB
    int TH = 16;
    while (TH < 64 && static_cast<long long>(TH + a.kh - 1) * 4 > static_cast<long long>(TH) * 5) TH *= 2;
    while (TH > 1 && f2d_lds_bytes(cn, TH, a.kw, a.kh) > 64 * 1024) TH /= 2;
    const size_t lds = f2d_lds_bytes(cn, TH, a.kw, a.kh);
    if (lds > 64 * 1024) return hipErrorInvalidConfiguration;
    const dim3 grid(static_cast<unsigned>((a.src.w + TX - 1) / TX), static_cast<unsigned>((a.src.h + TH - 1) / TH));
    // per-sample bound factor: (taps + 4) unit roundoffs, 1% and FP64 slack
    const double bscale = 1.01 * (a.ntaps + 5) / 16777216.0 + 1e-9;
    hipLaunchKernelGGL((f2d_fast_kernel<Ti, HZ, CERT>), grid, dim3(kBlock), lds, s,
                       static_cast<const Ti*>(a.src.data), a.src.stride, a.src.w, a.src.h, cn, a.dst.data,
                       a.dst.stride, a.dst.dt, a.kf, a.kw, a.kh, a.ax, a.ay, a.border, static_cast<float>(a.delta),
                       static_cast<float>(a.bval), TH, bscale, dtype_lo(a.dst.dt), dtype_hi(a.dst.dt), a.fix);
    return hipGetLastError();
}

template <class Ti>
hipError_t f2d_fast_dispatch(const F2dLaunch& a, hipStream_t s) {
    const bool cert = a.mode == MODE_CERTIFY;
    if (a.has_zero) return cert ? f2d_fast_launch<Ti, true, true>(a, s) : f2d_fast_launch<Ti, true, false>(a, s);
    return cert ? f2d_fast_launch<Ti, false, true>(a, s) : f2d_fast_launch<Ti, false, false>(a, s);
}

}  // namespace

hipError_t launch_filter2d_mfma(const F2dLaunch& a, hipStream_t s);

hipError_t launch_filter2d(const F2dLaunch& a, hipStream_t s) {
    const size_t total = static_cast<size_t>(a.src.w) * a.src.h * a.src.cn;
    if (a.mode == MODE_EXACT) {
        hipLaunchKernelGGL(f2d_exact_kernel, dim3(grid_for(total)), dim3(kBlock), 0, s, a.src, a.dst, a.taps,
                           a.tapw, a.ntaps, a.ax, a.ay, a.border, a.delta, a.bval);
        return hipGetLastError();
    }
    hipError_t e;
    if (a.mode == MODE_CERTIFY) {
        e = launch_zero_counter(a.fix.count, s);
        if (e != hipSuccess) return e;
    }
    switch (a.src.dt) {
        case IPFG_U8: e = f2d_fast_dispatch<uint8_t>(a, s); break;
        case IPFG_U16: e = f2d_fast_dispatch<uint16_t>(a, s); break;
        case IPFG_I16: e = f2d_fast_dispatch<int16_t>(a, s); break;
        case IPFG_F32: e = f2d_fast_dispatch<float>(a, s); break;
        default: return hipErrorInvalidValue;
    }
    if (e != hipSuccess || a.mode != MODE_CERTIFY) return e;
    hipLaunchKernelGGL(f2d_fix_kernel, dim3(kFixBlocks), dim3(kBlock), 0, s, a.src, a.dst, a.taps, a.tapw, a.ntaps,
                       a.ax, a.ay, a.border, a.delta, a.bval, a.fix);
    return hipGetLastError();
}

}  // namespace ipfg
