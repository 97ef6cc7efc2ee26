// filter2D as an implicit GEMM on the MFMA units (precision modes BF16X3 and
// FP16X2), for large kernels.
//
// Reference semantics: include/ipf/filter.hpp:157-194 (correlation with the
// kernel, anchor, border handling, delta, saturate_cast). These modes trade
// the reference's FP64 arithmetic for matrix-core throughput: the result
// carries a small error against the FP64 value, which the tests measure and
// bound (max and RMS against the FP64 oracle). Zero weights, and the zero
// padding of the kernels below (the Toeplitz band; the rows kernel's 32 x 32
// weight table), take part as zeros: a non-finite input (which the reference
// would skip under a zero tap) can reach outputs up to 16 columns (Toeplitz)
// or 32 - kw columns and 32 - kh rows (rows kernel) outside its window.
// Finite inputs are unaffected.
//
// Two formulations: kernels up to 32 x 31 run the rows-on-M kernel (further
// below), larger ones the Toeplitz blocks described here.
//
// Toeplitz blocks. For one output row block of 16 rows and 16 columns, kernel row
// j contributes a 1-D correlation of input rows with K[j][.]; with the input
// columns split into chunks of 16, each (row j, chunk c) pair is a
// 16 x 16 x 16 product In[16 rows][16 cols] * T[16 cols][16 outputs] with the
// banded (Toeplitz) matrix T[p][n] = K[j][16 c + p - n]. Chunks are flattened
// (q = j * NC + c) and paired into the K = 32 dimension of
// v_mfma_f32_16x16x32_{f16,bf16}: lanes 0-31 carry chunk 2s, lanes 32-63 chunk
// 2s + 1 of pair s. T never depends on the tile position, so the host lays out
// every pair's B fragment once, in register order (one 16-byte load per lane
// per pair and weight part); A fragments are 16-byte LDS reads from the
// staged input tile.
//
// Split precision. Weights are split on the host into parts that sum to the
// FP64 weight to 22 (fp16, two parts) or 24 (bf16, three parts) bits; inputs
// that the format cannot hold exactly (float images; uint8 is exact in both)
// are split into hi + lo on the device. Products kept: FP16X2: x_hi w_hi +
// x_hi w_lo + x_lo w_hi; BF16X3: uint8: x (w0 + w1 + w2), float:
// x0 w0 + x0 w1 + x1 w0. Everything accumulates in FP32 in the MFMA
// accumulators. fp16 operands are scaled by powers of two (weights by the
// host, float inputs per tile from the tile's max |x|) so that they stay in
// fp16's normal range; the scale is removed exactly in the epilogue.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>

#include "kernels.hpp"

namespace ipfg {
namespace {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float float4v __attribute__((ext_vector_type(4)));


constexpr int MT = 1;              // 16-row tiles per wave (weight fragments are shared by MT x NT tiles)
constexpr int TYo = 4 * 16 * MT;   // output rows per workgroup
constexpr int TXo = 64;            // output columns per workgroup (4 tiles of 16)
constexpr int NT = TXo / 16;
constexpr int PD = 4;              // weight-fragment prefetch depth (pairs)

__host__ __device__ inline int mf_nc(int kw) { return (kw + 15 + 15) / 16; }          // chunks per kernel row
__host__ __device__ inline int mf_cols(int kw) { return TXo + 16 * (mf_nc(kw) - 1); }  // staged columns
__host__ __device__ inline int mf_pitch(int kw) { return mf_cols(kw) + 8; }             // halves, 16-byte rows, skewed banks
__host__ __device__ inline int mf_rows(int kh) { return TYo + kh - 1; }

template <bool BF>
struct Fmt;
template <> struct Fmt<false> {
    using T = _Float16;
    using V = half8;
    __device__ static float4v mfma(V a, V b, float4v c) { return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0); }
};
template <> struct Fmt<true> {
    using T = __bf16;
    using V = bf16x8;
    __device__ static float4v mfma(V a, V b, float4v c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }
};

struct MfArgs {
    DevImage src, dst;
    const void* bfrag;   // [parts][pairs][64 lanes][8] weight fragments
    int pairs, kw, kh, ax, ay, border;
    double delta, bval;
    double wscale;       // weights were multiplied by this power of two
};

// BF: bf16 (else fp16). WP: weight parts. XS: split the input into hi + lo.
template <class Ti, bool BF, int WP, bool XS>
__global__ __launch_bounds__(kBlock) void f2d_mfma_kernel(MfArgs a) {
    using F = Fmt<BF>;
    using T = typename F::T;
    using V = typename F::V;
    extern __shared__ __align__(16) unsigned char mf_smem[];
    __shared__ unsigned int tile_max_bits;
    const int W = a.src.w, H = a.src.h, cn = a.src.cn;
    const int kw = a.kw, kh = a.kh;
    const int NC = mf_nc(kw), COLS = mf_cols(kw), PITCH = mf_pitch(kw), ROWS = mf_rows(kh);
    T* hi = reinterpret_cast<T*>(mf_smem);
    T* lo = hi + ROWS * PITCH;  // used when XS
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int x0 = blockIdx.x * TXo, y0 = blockIdx.y * TYo, c = blockIdx.z;
    const int gx0 = x0 - a.ax, gy0 = y0 - a.ay;
    const Ti* __restrict__ src = static_cast<const Ti*>(a.src.data);
    const bool inside = gx0 >= 0 && gx0 + COLS <= W && gy0 >= 0 && gy0 + ROWS <= H;

    // ---- stage the input tile (border-mapped), optionally measuring max |x|.
    // Each lane converts two adjacent columns and writes them as one 32-bit
    // LDS word; the row is wave-uniform, so its address is computed once.
    auto fetch2 = [&](int r, int col, float& v0, float& v1) {
        if (inside) {
            const Ti* rp = reinterpret_cast<const Ti*>(reinterpret_cast<const char*>(src) +
                                                       static_cast<size_t>(gy0 + r) * a.src.stride) + c;
            const int o = (gx0 + col) * cn;
            v0 = static_cast<float>(rp[o]);
            v1 = static_cast<float>(rp[o + cn]);
            return;
        }
        const int sy = border_index(gy0 + r, H, a.border);
        const Ti* rp = reinterpret_cast<const Ti*>(reinterpret_cast<const char*>(src) +
                                                   static_cast<size_t>(sy < 0 ? 0 : sy) * a.src.stride) + c;
        const int sx0 = border_index(gx0 + col, W, a.border), sx1 = border_index(gx0 + col + 1, W, a.border);
        v0 = (sy < 0 || sx0 < 0) ? static_cast<float>(a.bval) : static_cast<float>(rp[static_cast<size_t>(sx0) * cn]);
        v1 = (sy < 0 || sx1 < 0) ? static_cast<float>(a.bval) : static_cast<float>(rp[static_cast<size_t>(sx1) * cn]);
    };
    const int COLS2 = COLS / 2;  // COLS is a multiple of 16
    float xscale = 1.0f;
    if constexpr (XS && !BF) {
        // fp16: scale the tile so that max |x| lands in [2^13, 2^14)
        if (tid == 0) tile_max_bits = 0u;
        __syncthreads();
        float m = 0.f;
        for (int r = wave; r < ROWS; r += kBlock / 64)
            for (int c2 = lane; c2 < COLS2; c2 += 64) {
                float v0, v1;
                fetch2(r, 2 * c2, v0, v1);
                m = fmaxf(m, fmaxf(fabsf(v0), fabsf(v1)));
            }
        wave_max_into(&tile_max_bits, m);
        __syncthreads();
        const float mx = __uint_as_float(tile_max_bits);
        if (mx > 0.f && mx < INFINITY) {
            int ex;
            frexpf(mx, &ex);  // mx in [2^(ex-1), 2^ex)
            xscale = ldexpf(1.0f, 14 - ex);
        }
    }
    typedef T T2 __attribute__((ext_vector_type(2)));
    for (int r = wave; r < ROWS; r += kBlock / 64) {
        for (int c2 = lane; c2 < COLS2; c2 += 64) {
            float v0, v1;
            fetch2(r, 2 * c2, v0, v1);
            v0 *= xscale;
            v1 *= xscale;
            const T h0 = static_cast<T>(v0), h1 = static_cast<T>(v1);
            *reinterpret_cast<T2*>(hi + r * PITCH + 2 * c2) = T2{h0, h1};
            if constexpr (XS)
                *reinterpret_cast<T2*>(lo + r * PITCH + 2 * c2) =
                    T2{static_cast<T>(v0 - static_cast<float>(h0)), static_cast<T>(v1 - static_cast<float>(h1))};
        }
    }
    __syncthreads();

    // ---- MFMA main loop over chunk pairs
    float4v acc[MT][NT];
#pragma unroll
    for (int u = 0; u < MT; ++u)
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[u][t] = float4v{0.f, 0.f, 0.f, 0.f};
    const int half = lane >> 5;                      // which chunk of the pair
    const int m = lane & 15;                         // A row within the 16-row tile
    const int kofs = 8 * ((lane & 31) >> 4);         // 0 or 8: first of the lane's 8 k values
    const int lane_base = (wave * 16 * MT + m) * PITCH + kofs;
    const uint4* __restrict__ bf = static_cast<const uint4*>(a.bfrag);
    const int nq = kh * NC;
    auto load_b = [&](V (&b)[WP], int s) {
#pragma unroll
        for (int p = 0; p < WP; ++p) b[p] = __builtin_bit_cast(V, bf[(static_cast<size_t>(p) * a.pairs + s) * 64 + lane]);
    };
    // chunk q = 2 s + half of this lane: kernel row j, column chunk cc (advanced incrementally).
    // Every chunk of a pair is live except the second one of the last pair when
    // kh * NC is odd: that lane half reads chunk 0 of its own row instead (its
    // weight fragment is zero, and the data stays inside the staged tile).
    int q = half, j = half / NC, cc = half - (half / NC) * NC;
    // a ring of PD register sets of weight fragments: pairs s + 1 .. s + PD - 1
    // are in flight while pair s is multiplied (the pair count is a multiple
    // of PD; the padding pairs have zero weights)
    V bring[PD][WP];
#pragma unroll
    for (int r = 0; r < PD; ++r) load_b(bring[r], r);
    for (int s0 = 0; s0 < a.pairs; s0 += PD) {
#pragma unroll
        for (int r = 0; r < PD; ++r) {
            const int base = lane_base + (q < nq ? j * PITCH + 16 * cc : 0);
#pragma unroll
            for (int u = 0; u < MT; ++u) {
                V x[NT], xl[NT];
#pragma unroll
                for (int t = 0; t < NT; ++t) {
                    x[t] = *reinterpret_cast<const V*>(hi + base + 16 * u * PITCH + 16 * t);
                    if constexpr (XS) xl[t] = *reinterpret_cast<const V*>(lo + base + 16 * u * PITCH + 16 * t);
                }
                // parts outer, tiles inner: consecutive MFMAs never feed each other
#pragma unroll
                for (int p = 0; p < WP; ++p)
#pragma unroll
                    for (int t = 0; t < NT; ++t) acc[u][t] = F::mfma(x[t], bring[r][p], acc[u][t]);
                if constexpr (XS) {
#pragma unroll
                    for (int t = 0; t < NT; ++t) acc[u][t] = F::mfma(xl[t], bring[r][0], acc[u][t]);
                }
            }
            if (s0 + r + PD < a.pairs) load_b(bring[r], s0 + r + PD);
            q += 2;
            cc += 2;
            while (cc >= NC) {
                cc -= NC;
                ++j;
            }
        }
    }

    // ---- epilogue: lane holds column lane % 16, rows 4 * (lane / 16) + i of each tile
    const double unscale = 1.0 / (static_cast<double>(xscale) * a.wscale);
    const int es = static_cast<int>(dtype_size(a.dst.dt));
#pragma unroll
    for (int u = 0; u < MT; ++u) {
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int x = x0 + 16 * t + (lane & 15);
            if (x >= W) continue;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int y = y0 + (wave * MT + u) * 16 + 4 * (lane >> 4) + i;
                if (y >= H) break;
                const double v = static_cast<double>(acc[u][t][i]) * unscale + a.delta;
                store_sample(a.dst.data, static_cast<size_t>(y) * a.dst.stride + (static_cast<size_t>(x) * cn + c) * es,
                             a.dst.dt, v);
            }
        }
    }
}

template <class Ti, bool BF, int WP, bool XS>
hipError_t mf_launch(const MfArgs& a, hipStream_t s) {
    const size_t lds = static_cast<size_t>(mf_rows(a.kh)) * mf_pitch(a.kw) * 2 * (XS ? 2 : 1);
    if (lds > 156 * 1024) return hipErrorInvalidConfiguration;
    auto kern = f2d_mfma_kernel<Ti, BF, WP, XS>;
    if (lds > 64 * 1024) {
        static size_t raised = 0;
        if (lds > raised) {
            const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                                     hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
            if (e != hipSuccess) return e;
            raised = lds;
        }
    }
    const dim3 grid(static_cast<unsigned>((a.src.w + TXo - 1) / TXo), static_cast<unsigned>((a.src.h + TYo - 1) / TYo),
                    static_cast<unsigned>(a.src.cn));
    hipLaunchKernelGGL(kern, grid, dim3(kBlock), lds, s, a);
    return hipGetLastError();
}

template <class Ti, bool XS>
hipError_t mf_by_prec(const MfArgs& a, int precision, hipStream_t s) {
    if (precision == IPFG_PREC_FP16X2) return mf_launch<Ti, false, 2, XS>(a, s);
    return mf_launch<Ti, true, 3, XS>(a, s);
}

// ---------------------------------------------------------------------------
// Rows-on-M formulation (kernels up to 32 x 31). One wave owns a strip of 128
// output columns and walks down its rows, one staged input row per step. For
// input row R and a tile s (s = 0..7) of the 16 output columns x0 + s + 8 n,
// one MFMA pair forms C[M][n] += sum_k A[M][k] B[k][n] with
//   B[k][n] = S_R[s + 8 n + k]       (the row, 32 taps per output column),
//   A[M][k] = K[(ph - M) mod 32][k]  (a rotation of the kernel rows),
// M = 0..31 over two 16-row halves and ph = step mod 32. Accumulator row M
// thereby collects output row y with kernel row j = (ph - M) mod 32 at every
// step: one output row per accumulator row, j advancing with the steps. At
// j = 31 (a zero kernel row) the row is complete: its 16 x 8 = 128 values
// (8 consecutive columns per lane) are stored and the row starts over at zero.
// The A fragments come from one 32 x 40 table of the scaled weights per part
// in LDS (a lane reads kernel row (ph - M) mod 32, taps 8 (l / 16) ..+7, one
// step ahead); the data fragments are 16-byte LDS reads from 8 copies of the
// staged row, copy s shifted left by s, so that every lane's 8 taps start
// 16-byte aligned. Executed work: 32 x 32 products per
// 31 x 31 taps (1.066) times (rows + 31) / rows per strip, times the parts.
constexpr int RW_TX = 128;               // output columns per wave
constexpr int RW_SPAN = 160;             // staged samples per row (128 + 31 taps, rounded up)
constexpr int RW_PAD = 8;                // copy s keeps sample c at c - s + RW_PAD (no bounds tests)
constexpr int RW_CP = RW_SPAN + RW_PAD;  // elements per shifted copy (pitch, a multiple of 8)
constexpr int RW_ROWB = 8 * RW_CP;       // elements of one staged row (8 copies)
constexpr int RW_AHEAD = 4;              // rows in flight from global memory
constexpr int RW_WK = 40;                // weight table pitch (32 taps + 8: rows 80 bytes apart, no bank conflicts)

struct RwArgs {
    DevImage src, dst;
    const uint4* wt;     // [parts][32 kernel rows][RW_WK] weight table (zero outside kh x kw)
    int kw, kh, ax, ay, border, sh;
    double delta, bval, wscale;
};

// Stores outputs x and x + 1 of row y (columns past the image are skipped).
template <class To>
__device__ __forceinline__ void rw_store2(const DevImage& d, int y, int x, int ch, float v0, float v1,
                                          double unscale, double delta) {
    To* row = reinterpret_cast<To*>(static_cast<char*>(d.data) + static_cast<size_t>(y) * d.stride);
    const To o0 = saturate_from_double<To>(static_cast<double>(v0) * unscale + delta);
    const To o1 = saturate_from_double<To>(static_cast<double>(v1) * unscale + delta);
    struct alignas(2 * sizeof(To)) Pair {
        To a, b;
    };
    if (d.cn == 1 && x + 1 < d.w && (reinterpret_cast<uintptr_t>(row) & (2 * sizeof(To) - 1)) == 0) {
        *reinterpret_cast<Pair*>(row + x) = Pair{o0, o1};
    } else {
        if (x < d.w) row[static_cast<size_t>(x) * d.cn + ch] = o0;
        if (x + 1 < d.w) row[static_cast<size_t>(x + 1) * d.cn + ch] = o1;
    }
}

// Waves per SIMD the rows kernels are compiled for (registers) and launched
// at (one full round of workgroups): 3, or 2 for bf16 with split inputs
// (whose LDS traffic gains nothing from a third wave) and for double inputs
// (which would spill at 3).
template <class Ti, bool BF, bool XS>
constexpr int rw_occupancy() {
    return ((BF && XS) || sizeof(Ti) == 8) ? 2 : 3;
}

template <class Ti, bool BF, int WP, bool XS>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(rw_occupancy<Ti, BF, XS>(), rw_occupancy<Ti, BF, XS>())))
void f2d_mfma_rows_kernel(RwArgs a) {
    using F = Fmt<BF>;
    using T = typename F::T;
    using V = typename F::V;
    // per wave: two staged-row buffers (hi and, with XS, lo) and a parking
    // row for the finished sums; waves never touch each other's, so after the
    // weight table is in place they synchronise with themselves only
    __shared__ __align__(16) T stage[4][2][XS ? 2 : 1][RW_ROWB];
    __shared__ __align__(16) T wtab[WP][32][RW_WK];
    __shared__ __align__(16) float park[4][RW_TX];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    {
        constexpr int n16 = WP * 32 * RW_WK * 2 / 16;
        uint4* w = reinterpret_cast<uint4*>(&wtab[0][0][0]);
        for (int i = tid; i < n16; i += kBlock) w[i] = a.wt[i];
    }
    const int W = a.src.w, H = a.src.h, cn = a.src.cn, ch = blockIdx.z;
    const int x0 = (blockIdx.x * 4 + wave) * RW_TX;
    const int y0 = blockIdx.y * a.sh;
    const int rows = min(a.sh, H - y0);
    const int nsteps = rows + 31;
    const int cx0 = x0 - a.ax;  // image column of staged sample 0
    const char* __restrict__ src = static_cast<const char*>(a.src.data);
    // lane l stages samples l, l + 64, l + 128 (< RW_SPAN) of every row: their
    // byte offsets in a row are fixed for the strip. Loads always read inside
    // the image (offset 0 stands in for a Constant-border column, row 0 for a
    // Constant-border row) and the border value replaces them when the row is
    // staged, so no load sits under a branch and the ring keeps them in flight.
    long long coff[3];
    bool cok[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const int sx = border_index(cx0 + lane + 64 * i, W, a.border);
        cok[i] = sx >= 0;
        coff[i] = cok[i] ? (static_cast<long long>(sx) * cn + ch) * static_cast<long long>(sizeof(Ti)) : 0;
    }
    const float bvalf = static_cast<float>(a.bval);
    auto fetch = [&](int q, Ti (&v)[3]) -> bool {
        const int sy = border_index(y0 - a.ay + q, H, a.border);  // wave-uniform
        const char* rp = src + static_cast<size_t>(sy < 0 ? 0 : sy) * a.src.stride;
#pragma unroll
        for (int i = 0; i < 3; ++i) v[i] = *reinterpret_cast<const Ti*>(rp + coff[i]);
        return sy >= 0;
    };
    auto value = [&](const Ti (&v)[3], bool row_ok, int i) {
        return (row_ok && cok[i]) ? static_cast<float>(v[i]) : bvalf;
    };
    // fp16 keeps float inputs in its normal range: one power-of-two scale per
    // strip, max |x| of the strip's staged samples in [2^13, 2^14)
    float xscale = 1.0f;
    if constexpr (XS && !BF) {
        float m = 0.f;
        Ti pv[RW_AHEAD][3];
        bool pok[RW_AHEAD];
#pragma unroll
        for (int r = 0; r < RW_AHEAD; ++r) pok[r] = fetch(r, pv[r]);
        for (int q0 = 0; q0 < nsteps; q0 += RW_AHEAD) {
#pragma unroll
            for (int E = 0; E < RW_AHEAD; ++E) {
                if (q0 + E < nsteps) {
                    m = fmaxf(m, fmaxf(fabsf(value(pv[E], pok[E], 0)),
                                       fmaxf(fabsf(value(pv[E], pok[E], 1)),
                                             lane + 128 < RW_SPAN ? fabsf(value(pv[E], pok[E], 2)) : 0.f)));
                    pok[E] = fetch(q0 + E + RW_AHEAD, pv[E]);
                }
            }
        }
        const float mx = __uint_as_float(wave_max_u32(__float_as_uint(m)));
        if (mx > 0.f && mx < INFINITY) xscale = ldexpf(1.0f, 13 - ilogbf(mx));  // mx in [2^e, 2^(e+1))
    }
    // rows in flight: ring of RW_AHEAD raw sample sets and their row flags
    Ti ring[RW_AHEAD][3];
    bool rok[RW_AHEAD];
    auto store_row = [&](const Ti (&v)[3], bool row_ok, int buf) {
        T* hi = stage[wave][buf][0] + RW_PAD;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const int c = lane + 64 * i;
            if (i == 2 && c >= RW_SPAN) break;
            const float x = value(v, row_ok, i) * xscale;
            const T h = static_cast<T>(x);
#pragma unroll
            for (int sft = 0; sft < 8; ++sft) hi[sft * RW_CP + c - sft] = h;
            if constexpr (XS) {
                const T l = static_cast<T>(x - static_cast<float>(h));
                T* lo = stage[wave][buf][1] + RW_PAD;
#pragma unroll
                for (int sft = 0; sft < 8; ++sft) lo[sft * RW_CP + c - sft] = l;
            }
        }
    };
    auto wave_sync = [] {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };
    float4v acc[8][2];
#pragma unroll
    for (int t = 0; t < 8; ++t) acc[t][0] = acc[t][1] = float4v{0.f, 0.f, 0.f, 0.f};
    const int n = lane & 15, g = lane >> 4;
    const int boff = RW_PAD + 8 * n + 8 * g;  // element offset of this lane's 8 taps in a copy
    const double unscale = 1.0 / (static_cast<double>(xscale) * a.wscale);
#pragma unroll
    for (int r = 0; r < RW_AHEAD; ++r) rok[r] = fetch(r, ring[r]);
    store_row(ring[0], rok[0], 0);
    rok[0] = fetch(RW_AHEAD, ring[0]);
    __syncthreads();  // weight table and every wave's first row
    // A fragments of the next step (read one step ahead): accumulator row
    // M = 16 mh + n takes kernel row (ph - M) mod 32
    V wnext[2][WP];
    auto load_weights = [&](int ph) {
#pragma unroll
        for (int mh = 0; mh < 2; ++mh)
#pragma unroll
            for (int p = 0; p < WP; ++p)
                wnext[mh][p] = *reinterpret_cast<const V*>(&wtab[p][(ph - 16 * mh - n) & 31][8 * g]);
    };
    load_weights(0);
    // steps in groups of 4 (= RW_AHEAD): the accumulator element that
    // finishes and the register set of the prefetch ring are compile-time
    // indices after unrolling; which half finishes is uniform
    for (int q0 = 0; q0 < nsteps; q0 += 4) {
#pragma unroll
        for (int E = 0; E < 4; ++E) {
            const int q = q0 + E;
            if (q < nsteps) {
                const int buf = q & 1, ph = q & 31;
                V wa[2][WP];
#pragma unroll
                for (int mh = 0; mh < 2; ++mh)
#pragma unroll
                    for (int p = 0; p < WP; ++p) wa[mh][p] = wnext[mh][p];
                V xb[8], xl[8];
#pragma unroll
                for (int t = 0; t < 8; ++t) {
                    xb[t] = *reinterpret_cast<const V*>(&stage[wave][buf][0][t * RW_CP + boff]);
                    if constexpr (XS) xl[t] = *reinterpret_cast<const V*>(&stage[wave][buf][1][t * RW_CP + boff]);
                }
                if constexpr (XS) {
#pragma unroll
                    for (int p = 0; p < WP; ++p)
#pragma unroll
                        for (int t = 0; t < 8; ++t)
#pragma unroll
                            for (int mh = 0; mh < 2; ++mh) acc[t][mh] = F::mfma(wa[mh][p], xb[t], acc[t][mh]);
#pragma unroll
                    for (int t = 0; t < 8; ++t)
#pragma unroll
                        for (int mh = 0; mh < 2; ++mh) acc[t][mh] = F::mfma(wa[mh][0], xl[t], acc[t][mh]);
                } else {
#pragma unroll
                    for (int p = 0; p < WP; ++p)
#pragma unroll
                        for (int t = 0; t < 8; ++t) acc[t][0] = F::mfma(wa[0][p], xb[t], acc[t][0]);
                    // rows M >= 16 finish (as warm-up rows, discarded) at step
                    // M - 1 >= 15 and start over: their earlier products are unused
                    if (q >= 16) {
#pragma unroll
                        for (int p = 0; p < WP; ++p)
#pragma unroll
                            for (int t = 0; t < 8; ++t) acc[t][1] = F::mfma(wa[1][p], xb[t], acc[t][1]);
                    }
                }
                // the next row into the other buffer (read two steps ago), the
                // row after the prefetched ones into the freed register set
                store_row(ring[(E + 1) & 3], rok[(E + 1) & 3], buf ^ 1);
                load_weights(ph + 1);
                rok[(E + 1) & 3] = fetch(q + 1 + RW_AHEAD, ring[(E + 1) & 3]);
                // accumulator row (ph + 1) mod 32 has seen kernel rows 0..31:
                // output row y0 + q - 31, held by lane group gf (16 lanes x 8
                // columns); they park it, then every lane converts two columns
                const int ef = (E + 1) & 3;
                const int Mf = (ph + 1) & 31, gf = (Mf & 15) >> 2;
                const bool upper = (Mf >> 4) != 0, mine = g == gf;
                float v[8];
#pragma unroll
                for (int t = 0; t < 8; ++t) {
                    const float v0 = acc[t][0][ef], v1 = acc[t][1][ef];
                    v[t] = upper ? v1 : v0;
                    acc[t][0][ef] = (mine && !upper) ? 0.f : v0;
                    acc[t][1][ef] = (mine && upper) ? 0.f : v1;
                }
                if (q >= 31) {
                    if (mine) {
                        float4* pk = reinterpret_cast<float4*>(&park[wave][8 * n]);
                        pk[0] = make_float4(v[0], v[1], v[2], v[3]);
                        pk[1] = make_float4(v[4], v[5], v[6], v[7]);
                    }
                    wave_sync();
                    const float2 pv = *reinterpret_cast<const float2*>(&park[wave][2 * lane]);
                    const int y = y0 + q - 31, x = x0 + 2 * lane;
                    switch (a.dst.dt) {
                        case IPFG_U8: rw_store2<uint8_t>(a.dst, y, x, ch, pv.x, pv.y, unscale, a.delta); break;
                        case IPFG_U16: rw_store2<uint16_t>(a.dst, y, x, ch, pv.x, pv.y, unscale, a.delta); break;
                        case IPFG_I16: rw_store2<int16_t>(a.dst, y, x, ch, pv.x, pv.y, unscale, a.delta); break;
                        case IPFG_I32: rw_store2<int32_t>(a.dst, y, x, ch, pv.x, pv.y, unscale, a.delta); break;
                        case IPFG_F32: rw_store2<float>(a.dst, y, x, ch, pv.x, pv.y, unscale, a.delta); break;
                        default: rw_store2<double>(a.dst, y, x, ch, pv.x, pv.y, unscale, a.delta); break;
                    }
                }
                wave_sync();
            }
        }
    }
}

// Compute units of the current device (cached per device; 256 if the query fails).
int device_cus() {
    static std::atomic<int> cache[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    int n = cache[dev].load(std::memory_order_relaxed);
    if (n == 0) {
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
        cache[dev].store(n, std::memory_order_relaxed);
    }
    return n;
}

template <class Ti, bool BF, int WP, bool XS>
hipError_t rw_launch(const RwArgs& a0, hipStream_t s) {
    RwArgs a = a0;
    const int strips = (a.src.w + RW_TX - 1) / RW_TX;
    const int gx = (strips + 3) / 4;
    // one round of workgroups: rw_occupancy per CU (one wave per SIMD each);
    // strips at least 32 rows tall (each strip spends 31 steps warming up)
    const long long nby = std::max(1LL, static_cast<long long>(device_cus()) * rw_occupancy<Ti, BF, XS>() / gx);
    int sh = static_cast<int>((a.src.h + nby - 1) / nby);
    sh = std::max(sh, 32);
    a.sh = sh;
    const dim3 grid(static_cast<unsigned>(gx), static_cast<unsigned>((a.src.h + sh - 1) / sh),
                    static_cast<unsigned>(a.src.cn));
    hipLaunchKernelGGL((f2d_mfma_rows_kernel<Ti, BF, WP, XS>), grid, dim3(kBlock), 0, s, a);
    return hipGetLastError();
}

template <class Ti, bool XS>
hipError_t rw_by_prec(const RwArgs& a, int precision, hipStream_t s) {
    if (precision == IPFG_PREC_FP16X2) return rw_launch<Ti, false, 2, XS>(a, s);
    return rw_launch<Ti, true, 3, XS>(a, s);
}

bool rows_fit(int kw, int kh) { return kw <= 32 && kh <= 31; }

}  // namespace

size_t f2d_mfma_lds_bytes(int kw, int kh, bool split_input) {
    return static_cast<size_t>(mf_rows(kh)) * mf_pitch(kw) * 2 * (split_input ? 2 : 1);
}

int f2d_mfma_pairs(int kw, int kh) { return (((kh * mf_nc(kw) + 1) / 2 + PD - 1) / PD) * PD; }

namespace {
// The parts (hi, lo[, lo2]) of a scaled weight in the operand format, each
// rounded from what the earlier parts left.
void split_parts(double w, bool bf, int parts, uint16_t* bits) {
    double rest = w;
    for (int p = 0; p < parts; ++p) {
        double got;
        if (bf) {
            const __bf16 h = static_cast<__bf16>(static_cast<float>(rest));
            std::memcpy(&bits[p], &h, 2);
            got = static_cast<double>(static_cast<float>(h));
        } else {
            const _Float16 h = static_cast<_Float16>(static_cast<float>(rest));
            std::memcpy(&bits[p], &h, 2);
            got = static_cast<double>(static_cast<float>(h));
        }
        rest -= got;
    }
}
}  // namespace

// Weight fragments in register order.
// Rows-on-M kernels (kw <= 32, kh <= 31): the scaled weights as a table
// [part][32 kernel rows][RW_WK], zero outside the kernel.
// Otherwise: for part p, pair s, lane l, element e: chunk q = 2 s + (l >= 32),
// row j = q / NC, chunk column cc = q % NC, k = 8 ((l % 32) / 16) + e,
// n = l % 16, tap i = 16 cc + k - n.
void f2d_mfma_fragments(const double* kd, int kw, int kh, int precision, std::vector<uint16_t>& out,
                        double& wscale) {
    const bool bf = precision != IPFG_PREC_FP16X2;
    const int parts = bf ? 3 : 2;
    double mx = 0;
    for (int i = 0; i < kw * kh; ++i) mx = std::max(mx, std::fabs(kd[i]));
    wscale = 1.0;
    if (!bf && mx > 0) {  // fp16: max |w| * wscale in [2^13, 2^14)
        int ex;
        std::frexp(mx, &ex);
        wscale = std::ldexp(1.0, 14 - ex);
    }
    uint16_t bits[3];
    if (rows_fit(kw, kh)) {
        out.assign(static_cast<size_t>(parts) * 32 * RW_WK, 0);
        for (int j = 0; j < kh; ++j)
            for (int k = 0; k < kw; ++k) {
                split_parts(kd[static_cast<size_t>(j) * kw + k] * wscale, bf, parts, bits);
                for (int p = 0; p < parts; ++p) out[(static_cast<size_t>(p) * 32 + j) * RW_WK + k] = bits[p];
            }
        return;
    }
    const int NC = mf_nc(kw), pairs = f2d_mfma_pairs(kw, kh);
    out.assign(static_cast<size_t>(parts) * pairs * 64 * 8, 0);
    for (int s = 0; s < pairs; ++s)
        for (int l = 0; l < 64; ++l)
            for (int e = 0; e < 8; ++e) {
                const int q = 2 * s + (l >= 32 ? 1 : 0);
                const int j = q / NC, cc = q % NC;
                const int k = 8 * ((l % 32) / 16) + e, n = l % 16;
                const int i = 16 * cc + k - n;
                double w = 0.0;
                if (j < kh && i >= 0 && i < kw) w = kd[static_cast<size_t>(j) * kw + i] * wscale;
                split_parts(w, bf, parts, bits);
                for (int p = 0; p < parts; ++p) out[((static_cast<size_t>(p) * pairs + s) * 64 + l) * 8 + e] = bits[p];
            }
}

hipError_t launch_filter2d_mfma(const F2dLaunch& l, hipStream_t s) {
    if (rows_fit(l.kw, l.kh)) {
        RwArgs r{};
        r.src = l.src;
        r.dst = l.dst;
        r.wt = static_cast<const uint4*>(l.mfma_w);
        r.kw = l.kw;
        r.kh = l.kh;
        r.ax = l.ax;
        r.ay = l.ay;
        r.border = l.border;
        r.delta = l.delta;
        r.bval = l.bval;
        r.wscale = l.mfma_wscale;
        switch (l.src.dt) {
            case IPFG_U8: return rw_by_prec<uint8_t, false>(r, l.precision, s);
            case IPFG_U16: return rw_by_prec<uint16_t, true>(r, l.precision, s);
            case IPFG_I16: return rw_by_prec<int16_t, true>(r, l.precision, s);
            case IPFG_I32: return rw_by_prec<int32_t, true>(r, l.precision, s);
            case IPFG_F32: return rw_by_prec<float, true>(r, l.precision, s);
            case IPFG_F64: return rw_by_prec<double, true>(r, l.precision, s);
        }
        return hipErrorInvalidValue;
    }
    MfArgs a{};
    a.src = l.src;
    a.dst = l.dst;
    a.bfrag = l.mfma_w;
    a.pairs = f2d_mfma_pairs(l.kw, l.kh);
    a.kw = l.kw;
    a.kh = l.kh;
    a.ax = l.ax;
    a.ay = l.ay;
    a.border = l.border;
    a.delta = l.delta;
    a.bval = l.bval;
    a.wscale = l.mfma_wscale;
    switch (l.src.dt) {
        case IPFG_U8: return mf_by_prec<uint8_t, false>(a, l.precision, s);
        case IPFG_U16: return mf_by_prec<uint16_t, true>(a, l.precision, s);
        case IPFG_I16: return mf_by_prec<int16_t, true>(a, l.precision, s);
        case IPFG_I32: return mf_by_prec<int32_t, true>(a, l.precision, s);
        case IPFG_F32: return mf_by_prec<float, true>(a, l.precision, s);
        case IPFG_F64: return mf_by_prec<double, true>(a, l.precision, s);
    }
    return hipErrorInvalidValue;
}

}  // namespace ipfg
