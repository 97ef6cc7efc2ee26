// One-channel float32 separable filter (gaussianBlur / sepFilter2D / box with
// kw = kh = K odd, centred anchor), certified FP64: the kernel of the 8K
// float32 roofline workload. Reference: include/ipf/filter.hpp:34-90 (rows
// filtered with kx into an FP64 buffer, then columns with ky from delta),
// 205-253 (boxFilter), 261-267 (gaussianBlur).
//
// A workgroup (256 threads) walks down a strip of TX = 128 output columns
// and SH rows, CH padded rows per step (CH = 8 or 16):
//   stage      the CH rows x (S0 + 128 + K - 1) input samples of the step go
//              from registers (loaded two steps ahead: whole-wave rows, one
//              16-byte load per lane, scalar row addresses) to LDS as float,
//              shifted so that every horizontal window starts 16-byte aligned;
//   horizontal thread t filters P = CH / 2 consecutive columns of one staged
//              row (row t % CH: the 8 lanes an LDS cycle serves read 8
//              different rows, 4 banks apart) and writes P FP64 sums into a
//              ring of R = (K0 + 1) * CH rows;
//   vertical   thread t owns column t % 128 and M = CH / 2 consecutive output
//              rows, reads its M + K - 1 ring values and accumulates from
//              delta in the reference's tap order.
// Weights are wave-uniform scalar loads. Each output is rounded to float
// from both ends of its certification interval [r - b, r + b]: when the two
// floats agree the output is the reference's (it is stored as the lower one);
// otherwise the wave recomputes it in the reference's exact order
// (sepk::sep_exact_wave) in the same launch. b covers both FMA chains'
// rounding (bcoef * X + bdelta, X = running maximum |input| the workgroup
// staged) plus the rounding of r -/+ b itself.
#pragma once

#include "sep_kernels.hpp"

namespace ipfg {
namespace sep1 {

using sepk::SepArgs;

constexpr int TX = 128;  // output columns per workgroup
constexpr int kThreads = 256;
// A strip-height value at or above this selects the 8-row step variant (the
// value minus kHalfStep is the strip height).
constexpr int kHalfStep = 1 << 20;

constexpr int gcd(int a, int b) { return b == 0 ? a : gcd(b, a % b); }
constexpr int lcm(int a, int b) { return a / gcd(a, b) * b; }

template <int K, int CH>
struct Geo {
    static constexpr int P = CH / 2;                       // horizontal outputs per thread
    static constexpr int M = CH / 2;                       // vertical outputs per thread
    static constexpr int HG = TX / P;                      // horizontal tasks per staged row
    static_assert(CH * HG == kThreads && (CH / M) * TX == kThreads, "one task per thread and pass");
    static constexpr int A = K / 2;                        // anchor (centred)
    static constexpr int K0 = (K - 1 + CH - 1) / CH;       // steps before the first output rows
    static constexpr int NR = K0 + 1;                      // steps the ring holds
    static constexpr int R = NR * CH;                      // ring rows
    static constexpr int RP = TX + 2;                      // ring pitch (doubles): rows 4 banks apart
    static constexpr int S0 = (4 - A % 4) % 4;             // staged samples before column x0 - A
    static constexpr int PAD = (4 - S0) % 4;               // LDS slots before staged sample 0
    static constexpr int WB = S0 + PAD;                    // LDS index of column x0 - A (0 or 4)
    static constexpr int LR = (S0 + TX + K - 1 + 3) / 4;   // 16-byte segments per staged row
    static constexpr int NC = (P + K - 1 + 3) / 4;         // 16-byte LDS reads per horizontal task
    static constexpr int NEED = P * (HG - 1) + WB + 4 * NC;  // floats of a row the tasks read
    static constexpr int SPMIN = (PAD + 4 * LR > NEED ? PAD + 4 * LR : NEED);
    static constexpr int SP = ((SPMIN + 31) / 32) * 32 + 4;  // stage pitch (floats): 4 mod 32
    static constexpr size_t ring_bytes = sizeof(double) * R * RP;
    static constexpr size_t lds_bytes = ring_bytes + sizeof(float) * CH * SP;
    static_assert(LR <= 64, "a staged row is one 16-byte load per lane");
    static constexpr int D = 2;                            // chunks whose loads are in flight
    static constexpr int U = lcm(D, NR);                   // steps of one unrolled loop turn
};

// f(integral_constant<int, 0>), ..., f(integral_constant<int, U - 1>).
template <int U, int I = 0, class F>
__device__ __forceinline__ void unroll_steps(F&& f) {
    if constexpr (I < U) {
        f(std::integral_constant<int, I>{});
        unroll_steps<U, I + 1>(f);
    }
}

__device__ __forceinline__ float max3abs(float a, float b, float m) {
    return __builtin_fmaxf(__builtin_fmaxf(__builtin_fabsf(a), __builtin_fabsf(b)), m);
}

// One strip. HIN: every staged column lies inside the image (the interior
// strips); otherwise columns go through border_index. Loads are
// unconditional and never branch on data (rows outside the image load a
// valid row; Constant border values are substituted when the chunk is
// written to LDS), so the loads of a chunk stay in flight for two steps
// without register copies.
template <int K, int CH, bool UNIT, bool HIN>
__device__ __forceinline__ void sep1_strip(const SepArgs& a, const double* __restrict__ kx,
                                           const double* __restrict__ ky, int x0, int y0, int sh,
                                           unsigned char* __restrict__ smem, float* __restrict__ wave_max) {
    using G = Geo<K, CH>;
    constexpr int P = G::P, M = G::M;
    constexpr int RPW = CH / 4;  // staged rows per wave
    double* __restrict__ ring = reinterpret_cast<double*>(smem);
    float* __restrict__ stage = reinterpret_cast<float*>(smem + G::ring_bytes);
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int W = a.src.w, H = a.src.h;
    const int rows_out = min(sh, H - y0);
    const int nchunks = (rows_out + K - 1 + CH - 1) / CH;
    const int nsteps = G::K0 + (rows_out + CH - 1) / CH;
    const int gxs = x0 - G::A - G::S0;  // image column of staged sample 0 (a multiple of 4)
    const bool cst = a.border == IPFG_BORDER_CONSTANT;
    const float bvalf = static_cast<float>(a.bval);
    // Buffer resources over the source and destination (the launcher checks
    // both fit 2^31 bytes): row offsets go in scalar registers, lane offsets
    // stay fixed, so no per-row address arithmetic runs on the vector ALU.
    const int sstride = static_cast<int>(a.src.stride);
    const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
        a.src.data, 0, static_cast<int>(static_cast<size_t>(H) * a.src.stride), 0x00020000);

    // ---- staging: wave w loads rows w, w + 4, ... of a chunk; lane l loads
    // segment min(l, LR - 1) (lanes past the row re-read its last segment)
    const int seg = min(lane, G::LR - 1);
    int sxe[4] = {0, 0, 0, 0};  // edge strips: image column of each sample (-1: Constant border)
    if constexpr (!HIN) {
#pragma unroll
        for (int q = 0; q < 4; ++q) sxe[q] = border_index(gxs + 4 * seg + q, W, a.border);
    }
    auto src_row = [&](int gy) {  // byte offset of padded row gy (wave-uniform); Constant: row 0
        const int sy = (gy >= 0 && gy < H) ? gy : border_index(gy, H, a.border);
        return max(sy, 0) * sstride;
    };
    const int voff = (gxs + 4 * seg) * 4;  // interior strips: this lane's 16 bytes of a row
    auto load = [&](float4 (&reg)[RPW], int c) {
#pragma unroll
        for (int i = 0; i < RPW; ++i) {
            const int ro = src_row(y0 - G::A + CH * c + wave + 4 * i);
            if constexpr (HIN) {
                const auto v = __builtin_amdgcn_raw_buffer_load_b128(rsrc, voff, ro, 0);
                reg[i] = make_float4(__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]),
                                     __uint_as_float(v[3]));
            } else {
                float e[4];
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    e[q] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rsrc, max(sxe[q], 0) * 4, ro, 0));
                reg[i] = make_float4(e[0], e[1], e[2], e[3]);
            }
        }
    };
    float my_max = 0.f;
    auto store = [&](const float4 (&reg)[RPW], int c) {
#pragma unroll
        for (int i = 0; i < RPW; ++i) {
            float4 v = reg[i];
            if (cst) {
                const int gy = y0 - G::A + CH * c + wave + 4 * i;
                if (gy < 0 || gy >= H) v = make_float4(bvalf, bvalf, bvalf, bvalf);
                if constexpr (!HIN) {
                    if (sxe[0] < 0) v.x = bvalf;
                    if (sxe[1] < 0) v.y = bvalf;
                    if (sxe[2] < 0) v.z = bvalf;
                    if (sxe[3] < 0) v.w = bvalf;
                }
            }
            float* p = stage + (wave + 4 * i) * G::SP + G::PAD + 4 * seg;
            if constexpr (G::PAD == 0) {
                *reinterpret_cast<float4*>(p) = v;
            } else if constexpr (G::PAD == 2) {
                reinterpret_cast<float2*>(p)[0] = make_float2(v.x, v.y);
                reinterpret_cast<float2*>(p)[1] = make_float2(v.z, v.w);
            } else {
                p[0] = v.x;
                p[1] = v.y;
                p[2] = v.z;
                p[3] = v.w;
            }
            my_max = max3abs(v.x, v.y, my_max);
            my_max = max3abs(v.z, v.w, my_max);
        }
    };

    // ---- tasks
    const int hrow = tid & (CH - 1), hgrp = tid / CH;  // horizontal: staged row, column group
    const int vcol = tid & (TX - 1);                    // vertical: column
    const int vhalf = __builtin_amdgcn_readfirstlane(tid >> 7);  // vertical: which M of the step's CH rows
    const int xg = x0 + vcol;
    float* __restrict__ dst = static_cast<float*>(a.dst.data);
    const size_t dstride = a.dst.stride / sizeof(float);
    const __amdgpu_buffer_rsrc_t wsrc = __builtin_amdgcn_make_buffer_rsrc(
        a.dst.data, 0, static_cast<int>(static_cast<size_t>(H) * a.dst.stride), 0x00020000);
    const double oscale = a.oscale;
    unsigned int xmax_bits = 0u;

    // D register sets rotate: the loads of chunk k + D are issued when chunk
    // k has been stored, so each chunk's loads have D steps to land.
    // PH = k mod U (U a multiple of the ring period NR and of D): the ring
    // rows of a step are compile-time constants.
    auto step = [&](auto phc, float4 (&reg)[RPW], const int k) {
        constexpr int PH = decltype(phc)::value;
        constexpr int slot = PH % G::NR;
        if (k < nchunks) {
            store(reg, k);
            // wave maximum of what this wave staged so far (only grows); the
            // values are non-negative floats, which order like their bits
            const unsigned int m = wave_max_u32(__float_as_uint(my_max));
            if (lane == 0) wave_max[wave] = __uint_as_float(m);
        }
        load(reg, k + G::D);  // past the last chunk: valid rows, never stored
        __syncthreads();
        // ---- horizontal: row hrow of chunk k, columns [P hgrp, P hgrp + P)
        if (k < nchunks) {
            const float* wp = stage + hrow * G::SP + P * hgrp + G::WB;  // 16-byte aligned
            float w[4 * G::NC];
#pragma unroll
            for (int c = 0; c < G::NC; ++c) {
                const float4 f = *reinterpret_cast<const float4*>(wp + 4 * c);
                w[4 * c] = f.x;
                w[4 * c + 1] = f.y;
                w[4 * c + 2] = f.z;
                w[4 * c + 3] = f.w;
            }
            double acc[P];
            if constexpr (UNIT) {
                double s = 0.0;
#pragma unroll
                for (int i = 0; i < K; ++i) s += static_cast<double>(w[i]);
                acc[0] = s;
#pragma unroll
                for (int m = 1; m < P; ++m) {
                    s += static_cast<double>(w[m + K - 1]) - static_cast<double>(w[m - 1]);
                    acc[m] = s;
                }
            } else {
#pragma unroll
                for (int m = 0; m < P; ++m) acc[m] = 0.0;
#pragma unroll
                for (int i = 0; i < K; ++i) {
                    const double ki = kx[i];
#pragma unroll
                    for (int m = 0; m < P; ++m) acc[m] = fma(ki, static_cast<double>(w[m + i]), acc[m]);
                }
            }
            double* hp = ring + (slot * CH + hrow) * G::RP + P * hgrp;
#pragma unroll
            for (int m = 0; m < P; m += 2) *reinterpret_cast<double2*>(hp + m) = make_double2(acc[m], acc[m + 1]);
        }
        // certification bound of this step (X only grows: valid for every row used so far)
        if (k >= G::K0) {
            const float4 wm = *reinterpret_cast<const float4*>(wave_max);
            const float xm = __builtin_fmaxf(__builtin_fmaxf(wm.x, wm.y), __builtin_fmaxf(wm.z, wm.w));
            xmax_bits = __float_as_uint(xm);
        }
        __syncthreads();
        // ---- vertical: output rows [CH (k - K0) + M vhalf, + M)
        const int ob = CH * (k - G::K0) + M * vhalf;
        if (k < G::K0 || ob >= rows_out) return;
        const double X = static_cast<double>(__uint_as_float(xmax_bits));
        const double rmax = a.kabs * X + fabs(a.delta);
        const double bb = a.bcoef * X + a.bdelta + 4.5e-16 * rmax;  // non-finite X: every output doubtful
        // ring row of the task's first tap: CH ((k - K0) mod NR) + M vhalf,
        // known at compile time for each (PH, vhalf)
        constexpr int base = CH * (((PH - G::K0) % G::NR + G::NR) % G::NR);
        double h[M + K - 1];
        auto read_window = [&](auto vhc) {
            constexpr int s0 = base + M * decltype(vhc)::value;
#pragma unroll
            for (int j = 0; j < M + K - 1; ++j) h[j] = ring[((s0 + j) % G::R) * G::RP + vcol];
        };
        if (vhalf == 0) read_window(std::integral_constant<int, 0>{});
        else read_window(std::integral_constant<int, 1>{});
        double acc[M];
        if constexpr (UNIT) {
            double s = a.delta;
#pragma unroll
            for (int j = 0; j < K; ++j) s += h[j];
            acc[0] = s;
#pragma unroll
            for (int m = 1; m < M; ++m) {
                s += h[m + K - 1] - h[m - 1];
                acc[m] = s;
            }
        } else {
#pragma unroll
            for (int m = 0; m < M; ++m) acc[m] = a.delta;
#pragma unroll
            for (int j = 0; j < K; ++j) {
                const double kj = ky[j];
#pragma unroll
                for (int m = 0; m < M; ++m) acc[m] = fma(kj, h[m + j], acc[m]);
            }
        }
        const int yb = y0 + ob;
        float* op = dst + static_cast<size_t>(yb) * dstride + xg;
        const int orow = yb * static_cast<int>(a.dst.stride);  // byte offset of row yb (fits 2^31)
        // Results and doubt flags of all M rows first (independent chains),
        // then the stores: whole tiles store unconditionally.
        float out[M];
        unsigned long long bal[M];
        unsigned long long any = 0ull;
#pragma unroll
        for (int m = 0; m < M; ++m) {
            const double r = UNIT ? acc[m] * oscale : acc[m];
            out[m] = static_cast<float>(r - bb);
            const float fhi = static_cast<float>(r + bb);
            // a NaN result stays NaN whichever way it is computed: no flag needed
            bal[m] = __ballot(__float_as_uint(out[m]) != __float_as_uint(fhi));
            any |= bal[m];
        }
        const int nrow = min(M, rows_out - ob);
        if (x0 + TX <= W && nrow == M) {
#pragma unroll
            for (int m = 0; m < M; ++m)
                __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(out[m]), wsrc, xg * 4,
                                                      orow + m * static_cast<int>(a.dst.stride), 0);
        } else {
            const unsigned long long inside = __ballot(xg < W);
            any = 0ull;
#pragma unroll
            for (int m = 0; m < M; ++m) {
                bal[m] = m < nrow ? bal[m] & inside : 0ull;
                any |= bal[m];
                if (m < nrow && xg < W) op[static_cast<size_t>(m) * dstride] = out[m];
            }
        }
        if (any) {  // rare: recompute the doubtful outputs in the reference's order, one wave at a time
            unsigned int total = 0;
#pragma unroll 1
            for (int m = 0; m < M; ++m) {
                unsigned long long b = bal[m];
                total += __popcll(b);
                while (b) {
                    const int l = __builtin_ctzll(b);
                    b &= b - 1ull;
                    const int xlane = bcast_i(xg, l);
                    const int yl = yb + m;
                    const double e = sepk::sep_exact_wave<float>(a.src, xlane, yl, 0, a.kxd, K, a.kyd, K, G::A, G::A,
                                                                 a.border, a.delta, a.bval) *
                                     (UNIT ? oscale : 1.0);
                    if (lane == l) dst[static_cast<size_t>(yl) * dstride + xlane] = static_cast<float>(e);
                }
            }
            if (lane == 0) atomicAdd(a.fix.count, static_cast<unsigned long long>(total));
        }
    };
    float4 regs[G::D][RPW];
    unroll_steps<G::D>([&](auto c) { load(regs[decltype(c)::value], decltype(c)::value); });
    for (int k = 0; k < nsteps; k += G::U) {
        unroll_steps<G::U>([&](auto phc) {
            constexpr int PH = decltype(phc)::value;
            if (k + PH < nsteps) step(phc, regs[PH % G::D], k + PH);
        });
    }
}

template <int K, int CH, bool UNIT>
__global__ __launch_bounds__(kThreads) void sep1_kernel(SepArgs a, const double* __restrict__ kx,
                                                         const double* __restrict__ ky, int nbx, int nby, int sh) {
    using G = Geo<K, CH>;
    extern __shared__ __align__(16) unsigned char sep1_smem[];
    __shared__ float wave_max[4];
    // Tile order: workgroups are dispatched round robin over the 8 XCDs;
    // XCD j takes tiles [j * n / 8, (j + 1) * n / 8), rows of neighbouring
    // strips, so the strips sharing column halos share an L2.
    const int nblk = nbx * nby;
    int bid = blockIdx.x;
    if ((nblk & 7) == 0) bid = (bid & 7) * (nblk >> 3) + (bid >> 3);
    const int bx = bid % nbx, by = bid / nbx;
    const int x0 = bx * TX, y0 = by * sh;
    const int gxs = x0 - G::A - G::S0;
    if (gxs >= 0 && gxs + 4 * G::LR <= a.src.w)
        sep1_strip<K, CH, UNIT, true>(a, kx, ky, x0, y0, sh, sep1_smem, wave_max);
    else
        sep1_strip<K, CH, UNIT, false>(a, kx, ky, x0, y0, sh, sep1_smem, wave_max);
    fix_finish(a.fix);
}

template <int K, int CH, bool UNIT>
hipError_t sep1_launch_kc(const SepArgs& a, int sh, hipStream_t s) {
    using G = Geo<K, CH>;
    auto kern = sep1_kernel<K, CH, UNIT>;
    static int per_cu = 0;  // workgroups a CU holds (occupancy of this instantiation)
    if (per_cu == 0) {
        if (G::lds_bytes > 64 * 1024) {
            const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                                     static_cast<int>(G::lds_bytes));
            if (e != hipSuccess) return e;
        }
        int n = 0;
        const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, reinterpret_cast<const void*>(kern),
                                                                          kThreads, G::lds_bytes);
        if (e != hipSuccess) return e;
        per_cu = n > 0 ? n : 1;
    }
    const int nbx = (a.src.w + TX - 1) / TX;
    if (sh <= 0) {  // tall strips (few halo rows), one round of workgroups over the CUs
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
        const long long slots = static_cast<long long>(cus) * per_cu;
        long long nby = slots / nbx;
        if (nby < 1) nby = 1;
        sh = static_cast<int>((a.src.h + nby - 1) / nby);
    }
    sh = (sh + CH - 1) / CH * CH;
    const int nby = (a.src.h + sh - 1) / sh;
    hipLaunchKernelGGL(kern, dim3(static_cast<unsigned>(nbx * nby)), dim3(kThreads), G::lds_bytes, s, a,
                       static_cast<const double*>(a.kx), static_cast<const double*>(a.ky), nbx, nby, sh);
    return hipGetLastError();
}

// a.SH: strip height of the 16-row-step variant, or kHalfStep + the strip
// height of the 8-row-step variant; <= 0 picks the variant by window size.
template <int K, bool UNIT>
hipError_t sep1_launch_k(const SepArgs& a, hipStream_t s) {
    const bool half = a.SH >= kHalfStep || (a.SH <= 0 && K <= 31);  // 16-row steps pay off for long windows
    const int sh = a.SH >= kHalfStep ? a.SH - kHalfStep : a.SH;
    return half ? sep1_launch_kc<K, 8, UNIT>(a, sh, s) : sep1_launch_kc<K, 16, UNIT>(a, sh, s);
}

// Dispatch over the instantiated sizes; hipErrorNotSupported for calls this
// kernel does not take.
// Four translation units share the sizes so that they compile in parallel.
hipError_t launch_k3_13(const SepArgs& a, hipStream_t s);   // K = 3 .. 13
hipError_t launch_k15_25(const SepArgs& a, hipStream_t s);  // K = 15 .. 25
hipError_t launch_k27_37(const SepArgs& a, hipStream_t s);  // K = 27 .. 37
hipError_t launch_k39_49(const SepArgs& a, hipStream_t s);  // K = 39 .. 49

template <int KLO, int KHI>
hipError_t launch_range(const SepArgs& a, hipStream_t s) {
    if constexpr (KLO > KHI) {
        return hipErrorNotSupported;
    } else {
        if (a.kw == KLO) return a.unit ? sep1_launch_k<KLO, true>(a, s) : sep1_launch_k<KLO, false>(a, s);
        return launch_range<KLO + 2, KHI>(a, s);
    }
}

}  // namespace sep1
}  // namespace ipfg
