// filter2D (generic 2-D correlation) for gfx950.
//
// Reference: include/ipf/filter.hpp:157-194. Only kernel weights that compare
// unequal to zero take part, in row-major order; the sum starts at delta and is
// rounded once with saturate_cast.
//
// Fast path (FP64 certified, or FP32): a workgroup tile of TX pixels x TH
// rows with the input tile (plus kernel halo) staged in LDS, one plane per
// channel. Each thread produces P = 8 (kernels 8 or more wide) or 4
// horizontally adjacent outputs, walking the kernel rows and sliding a P-value
// register window along each row: taps go in blocks of 4, each block reads
// the next 4 samples with one 16-byte LDS load, so every LDS value feeds P
// FMAs and one conversion. The window is a ring whose origin moves by 4 per
// block (for P = 8 two phases, O = 0 and 4). Weights
// are read as wave-uniform scalars. Zero weights are skipped when the kernel
// has any (a uniform branch), matching the reference's tap list for
// non-finite inputs. Certified launches bound each output's distance to the
// reference's value; doubtful samples are recomputed in FP64 in the
// reference's order by the same launch.
//
// Exact path: the reference's tap list in FP64, one thread per sample.
#include "kernels.hpp"

namespace ipfg {
namespace {

constexpr int TX = 64;

// LDS row pitch (floats): a multiple of 4 for 16-byte loads, 4 more than a
// multiple of 8 to spread rows over the banks, and room for the last block's
// load (up to TX + kw + 10 samples from the row start).
__host__ __device__ inline int f2d_xp(int kw) { return (TX + kw + 11 + 7) / 8 * 8 + 4; }

__host__ inline size_t f2d_lds_bytes(int cn, int TH, int kw, int kh) {
    return sizeof(float) * static_cast<size_t>(cn) * (TH + kh - 1) * f2d_xp(kw);
}

// One output sample in the reference's order (filter.hpp:185-190: s = delta,
// then s += w * v over the non-zero taps in row-major order), computed by a
// whole wave: each lane forms the products of 64 taps at a time and the sum
// runs over broadcast values. All 64 lanes must call it; all return the result.
template <class Ti>
__device__ inline double f2d_exact_wave(const DevImage& src, int x, int y, int c, const int* __restrict__ taps,
                                        const double* __restrict__ tw, int ntaps, int ax, int ay, int border,
                                        double delta, double bval) {
    const int lane = lane_id();
    double s = delta;
    for (int t0 = 0; t0 < ntaps; t0 += 64) {
        const int t = t0 + lane;
        double p = 0.0;
        if (t < ntaps) {
            const int sx = border_index(x + taps[2 * t] - ax, src.w, border);
            const int sy = border_index(y + taps[2 * t + 1] - ay, src.h, border);
            const double v = (sx < 0 || sy < 0)
                                 ? bval
                                 : static_cast<double>(*reinterpret_cast<const Ti*>(
                                       static_cast<const char*>(src.data) + static_cast<size_t>(sy) * src.stride +
                                       (static_cast<size_t>(sx) * src.cn + c) * sizeof(Ti)));
            p = tw[t] * v;
        }
        const int n = min(64, ntaps - t0);
        for (int k = 0; k < n; ++k) s = s + bcast_d(p, k);
    }
    return s;
}

struct F2dArgs {
    DevImage src, dst;
    const void* K;       // kw * kh weights of type Acc, row-major
    const int* taps;     // non-zero taps (dx, dy) for exact recomputation
    const double* tapw;
    int ntaps, kw, kh, ax, ay, border, TH;
    double delta, bval;
    double bcoef, bdelta, xbound;  // certification bound = bcoef * X + bdelta
    FixList fix;
};

template <class Acc>
__device__ __forceinline__ Acc f2d_fma(Acc a, Acc b, Acc c) {
    if constexpr (sizeof(Acc) == 8) return fma(a, b, c);
    else return fmaf(a, b, c);
}

// Taps i0 .. i0 + nt - 1 (nt <= 4) of one kernel row for P outputs. On entry
// win[(O + q) % P] holds sample i0 + q of the window (q < P); each tap then
// retires its oldest sample for sample i0 + P + ii, which the following taps
// address by the same formula.
template <int P, int O, bool HZ, class Acc>
__device__ __forceinline__ void f2d_block4(const float* __restrict__ base, int i0, int nt, const Acc* __restrict__ krow,
                                           Acc (&acc)[P], Acc (&win)[P]) {
    const float4 f = *reinterpret_cast<const float4*>(base + i0 + P);
    const float nx[4] = {f.x, f.y, f.z, f.w};
#pragma unroll
    for (int ii = 0; ii < 4; ++ii) {
        if (ii < nt) {
            const Acc k = krow[i0 + ii];
            if (!HZ || k != Acc(0)) {
#pragma unroll
                for (int m = 0; m < P; ++m) acc[m] = f2d_fma(k, win[(O + ii + m) % P], acc[m]);
            }
            win[(O + ii) % P] = static_cast<Acc>(nx[ii]);
        }
    }
}

// One kernel row: P = 8 alternates the two window phases, P = 4 has one.
template <int P, bool HZ, class Acc>
__device__ __forceinline__ void f2d_row(const float* __restrict__ base, int kw, const Acc* __restrict__ krow,
                                        Acc (&acc)[P]) {
    Acc win[P];
#pragma unroll
    for (int q = 0; q < P; q += 4) {
        const float4 f = *reinterpret_cast<const float4*>(base + q);
        win[q] = f.x;
        win[q + 1] = f.y;
        win[q + 2] = f.z;
        win[q + 3] = f.w;
    }
    int i0 = 0;
    if constexpr (P == 8) {
        for (; i0 + 8 <= kw; i0 += 8) {
            f2d_block4<P, 0, HZ>(base, i0, 4, krow, acc, win);
            f2d_block4<P, 4, HZ>(base, i0 + 4, 4, krow, acc, win);
        }
        if (i0 + 4 <= kw) {
            f2d_block4<P, 0, HZ>(base, i0, 4, krow, acc, win);
            if (i0 + 4 < kw) f2d_block4<P, 4, HZ>(base, i0 + 4, kw - i0 - 4, krow, acc, win);
        } else if (i0 < kw) {
            f2d_block4<P, 0, HZ>(base, i0, kw - i0, krow, acc, win);
        }
    } else {
        for (; i0 + 4 <= kw; i0 += 4) f2d_block4<P, 0, HZ>(base, i0, 4, krow, acc, win);
        if (i0 < kw) f2d_block4<P, 0, HZ>(base, i0, kw - i0, krow, acc, win);
    }
}

// KS > 0: the kernel is KS x KS (the row's tap loop is unrolled and its
// weights are loaded ahead); 0: sizes at run time.
template <class Ti, class Acc, bool HZ, bool CERT, int P, int KS>
__global__ __launch_bounds__(kBlock) void f2d_fast_kernel(F2dArgs a, const Acc* __restrict__ K) {
    extern __shared__ float smem[];
    __shared__ unsigned int tile_max_bits;
    const int W = a.src.w, H = a.src.h, cn = a.src.cn;
    const int kw = KS ? KS : a.kw, kh = KS ? KS : a.kh, TH = a.TH;
    const int XP = f2d_xp(kw);
    const int HR = TH + kh - 1;
    const int plane = HR * XP;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int x0 = blockIdx.x * TX, y0 = blockIdx.y * TH;
    const int XW = TX + kw - 1;
    const int rowSamples = XW * cn;
    const bool inside_x = (x0 - a.ax >= 0) && (x0 - a.ax + XW <= W);
    const float bval = static_cast<float>(a.bval);
    const bool measure = CERT && a.xbound < 0;
    const Ti* __restrict__ src = static_cast<const Ti*>(a.src.data);
    if (tid == 0) tile_max_bits = 0u;
    float my_max = 0.f;
    __syncthreads();
    for (int r = wave; r < HR; r += kBlock / 64) {
        const int gy = y0 - a.ay + r;
        const int sy = (gy >= 0 && gy < H) ? gy : border_index(gy, H, a.border);
        float* trow = smem + r * XP;
        if (sy < 0) {
            for (int s = lane; s < rowSamples; s += 64) {
                const int px = s / cn, c = s - px * cn;
                trow[c * plane + px] = bval;
            }
            if (measure) my_max = fmaxf(my_max, fabsf(bval));
            continue;
        }
        const Ti* srow = reinterpret_cast<const Ti*>(reinterpret_cast<const char*>(src) + static_cast<size_t>(sy) * a.src.stride);
        for (int s = lane; s < rowSamples; s += 64) {
            const int px = s / cn, c = s - px * cn;
            float v;
            if (inside_x) {
                v = static_cast<float>(srow[static_cast<size_t>(x0 - a.ax) * cn + s]);
            } else {
                const int gx = x0 - a.ax + px;
                const int sx = (gx >= 0 && gx < W) ? gx : border_index(gx, W, a.border);
                v = sx < 0 ? bval : static_cast<float>(srow[static_cast<size_t>(sx) * cn + c]);
            }
            if (measure) my_max = fmaxf(my_max, (v == v) ? fabsf(v) : INFINITY);
            trow[c * plane + px] = v;
        }
    }
    if (measure) wave_max_into(&tile_max_bits, my_max);
    __syncthreads();
    double bound = 0.0;
    if (CERT) bound = a.bcoef * (measure ? static_cast<double>(__uint_as_float(tile_max_bits)) : a.xbound) + a.bdelta;
    constexpr int G = TX / P;
    const int ntask = TH * G * cn;
    const int ntask64 = (ntask + 63) & ~63;  // whole waves stay together for the fix-ups
    // wave-uniform loop over groups of 64 tasks (the trip count and the tap
    // loops below stay scalar)
    for (int t0 = __builtin_amdgcn_readfirstlane(wave) * 64; t0 < ntask64; t0 += kBlock) {
        const int t = t0 + lane;
        // every lane walks the same taps (so the weights stay wave-uniform
        // scalars); lanes past the last task filter row 0 of plane 0 and
        // store nothing
        const bool act = t < ntask;
        const int g = t % G;
        const int rest = act ? t / G : 0;
        const int r = rest % TH;
        const int c = rest / TH;
        Acc acc[P];
#pragma unroll
        for (int m = 0; m < P; ++m) acc[m] = static_cast<Acc>(a.delta);
        for (int j = 0; j < kh; ++j)
            f2d_row<P, HZ>(smem + c * plane + (r + j) * XP + g * P, kw, K + j * kw, acc);
        const int y = y0 + r;
        unsigned int dmask = 0u;
        if (act && y < H) {
            char* drow = static_cast<char*>(a.dst.data) + static_cast<size_t>(y) * a.dst.stride;
#pragma unroll
            for (int m = 0; m < P; ++m) {
                const int x = x0 + g * P + m;
                if (x >= W) break;
                const size_t q = static_cast<size_t>(x) * cn + c;
                const double v = static_cast<double>(acc[m]);
                bool doubt = false;
                switch (a.dst.dt) {
                    case IPFG_U8:
                        reinterpret_cast<uint8_t*>(drow)[q] = saturate_from_double<uint8_t>(v);
                        if (CERT) doubt = needs_exact(v, bound, 0.0, 255.0);
                        break;
                    case IPFG_U16:
                        reinterpret_cast<uint16_t*>(drow)[q] = saturate_from_double<uint16_t>(v);
                        if (CERT) doubt = needs_exact(v, bound, 0.0, 65535.0);
                        break;
                    case IPFG_I16:
                        reinterpret_cast<int16_t*>(drow)[q] = saturate_from_double<int16_t>(v);
                        if (CERT) doubt = needs_exact(v, bound, -32768.0, 32767.0);
                        break;
                    case IPFG_I32:
                        reinterpret_cast<int32_t*>(drow)[q] = saturate_from_double<int32_t>(v);
                        if (CERT) doubt = needs_exact(v, bound, -2147483648.0, 2147483647.0);
                        break;
                    case IPFG_F32:
                        reinterpret_cast<float*>(drow)[q] = static_cast<float>(v);
                        if (CERT) doubt = needs_exact_f32(v, bound);
                        break;
                    default:
                        reinterpret_cast<double*>(drow)[q] = v;
                        break;
                }
                if (CERT && doubt) dmask |= 1u << m;
            }
        }
        if (CERT) {
            const size_t es = dtype_size(a.dst.dt);
            wave_fix(a.fix, dmask, [&](int l, int m) {
                const int xl = x0 + bcast_i(g, l) * P + m, yl = bcast_i(y, l), cl = bcast_i(c, l);
                const double e = f2d_exact_wave<Ti>(a.src, xl, yl, cl, a.taps, a.tapw, a.ntaps, a.ax, a.ay, a.border,
                                                    a.delta, a.bval);
                if (lane_id() == l)
                    store_sample(a.dst.data, static_cast<size_t>(yl) * a.dst.stride + (static_cast<size_t>(xl) * cn + cl) * es,
                                 a.dst.dt, e);
            });
        }
    }
    if (CERT) fix_finish(a.fix);
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

int grid_for(size_t n) {
    const size_t b = (n + kBlock - 1) / kBlock;
    return static_cast<int>(b < 16384 ? (b ? b : 1) : 16384);
}

template <class Ti, class Acc, bool HZ, bool CERT, int P, int KS>
hipError_t f2d_fast_launch(const F2dLaunch& l, hipStream_t s) {
    const int cn = l.src.cn;
    // rows per tile: halo overhead (TH + kh - 1) / TH <= 1.25 up to 64 rows; LDS within 64 KB
    int TH = 16;
    if (static_cast<long long>(l.kw) * l.kh >= 64) TH = 32;  // arithmetic-bound: workgroups per CU over halo rows
    else
        while (TH < 64 && static_cast<long long>(TH + l.kh - 1) * 4 > static_cast<long long>(TH) * 5) TH *= 2;
    while (TH > 1 && f2d_lds_bytes(cn, TH, l.kw, l.kh) > 64 * 1024) TH /= 2;
    const size_t lds = f2d_lds_bytes(cn, TH, l.kw, l.kh);
    if (lds > 64 * 1024) return hipErrorInvalidConfiguration;
    F2dArgs a{};
    a.src = l.src;
    a.dst = l.dst;
    a.K = sizeof(Acc) == 8 ? static_cast<const void*>(l.kd) : static_cast<const void*>(l.kf);
    a.taps = l.taps;
    a.tapw = l.tapw;
    a.ntaps = l.ntaps;
    a.kw = l.kw;
    a.kh = l.kh;
    a.ax = l.ax;
    a.ay = l.ay;
    a.border = l.border;
    a.TH = TH;
    a.delta = l.delta;
    a.bval = l.bval;
    a.bcoef = l.bcoef;
    a.bdelta = l.bdelta;
    a.xbound = l.xbound;
    a.fix = l.fix;
    const dim3 grid(static_cast<unsigned>((l.src.w + TX - 1) / TX), static_cast<unsigned>((l.src.h + TH - 1) / TH));
    hipLaunchKernelGGL((f2d_fast_kernel<Ti, Acc, HZ, CERT, P, KS>), grid, dim3(kBlock), lds, s, a,
                       static_cast<const Acc*>(a.K));
    return hipGetLastError();
}

template <class Ti, int P>
hipError_t f2d_fast_dispatch_p(const F2dLaunch& a, hipStream_t s) {
    if (a.mode == MODE_CERTIFY) {
        if (P == 8 && !a.has_zero && a.kw == 31 && a.kh == 31)  // dense 31 x 31
            return f2d_fast_launch<Ti, double, false, true, P, 31>(a, s);
        return a.has_zero ? f2d_fast_launch<Ti, double, true, true, P, 0>(a, s)
                          : f2d_fast_launch<Ti, double, false, true, P, 0>(a, s);
    }
    return a.has_zero ? f2d_fast_launch<Ti, float, true, false, P, 0>(a, s)
                      : f2d_fast_launch<Ti, float, false, false, P, 0>(a, s);
}

template <class Ti>
hipError_t f2d_fast_dispatch(const F2dLaunch& a, hipStream_t s) {
    return a.kw >= 8 ? f2d_fast_dispatch_p<Ti, 8>(a, s) : f2d_fast_dispatch_p<Ti, 4>(a, s);
}

}  // namespace

hipError_t launch_filter2d(const F2dLaunch& a, hipStream_t s) {
    const size_t total = static_cast<size_t>(a.src.w) * a.src.h * a.src.cn;
    if (a.mode == MODE_EXACT) {
        hipLaunchKernelGGL(f2d_exact_kernel, dim3(grid_for(total)), dim3(kBlock), 0, s, a.src, a.dst, a.taps,
                           a.tapw, a.ntaps, a.ax, a.ay, a.border, a.delta, a.bval);
        return hipGetLastError();
    }
    switch (a.src.dt) {
        case IPFG_U8: return f2d_fast_dispatch<uint8_t>(a, s);
        case IPFG_U16: return f2d_fast_dispatch<uint16_t>(a, s);
        case IPFG_I16: return f2d_fast_dispatch<int16_t>(a, s);
        case IPFG_F32: return f2d_fast_dispatch<float>(a, s);
    }
    return hipErrorInvalidValue;
}

}  // namespace ipfg
