// filter2D as an implicit GEMM on the MFMA units (precision modes BF16X3 and
// FP16X2), for large kernels.
//
// Reference semantics: include/ipf/filter.hpp:157-194 (correlation with the
// kernel, anchor, border handling, delta, saturate_cast). These modes trade
// the reference's FP64 arithmetic for matrix-core throughput: the result
// carries a small error against the FP64 value, which the tests measure and
// bound (max and RMS against the FP64 oracle). Zero weights, and the zero
// band of the Toeplitz blocks below, take part as zeros: a non-finite input
// (which the reference would skip under a zero tap) can reach outputs up to
// 16 columns outside its window. Finite inputs are unaffected.
//
// Formulation. For one output row block of 16 rows and 16 columns, kernel row
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

}  // namespace

size_t f2d_mfma_lds_bytes(int kw, int kh, bool split_input) {
    return static_cast<size_t>(mf_rows(kh)) * mf_pitch(kw) * 2 * (split_input ? 2 : 1);
}

int f2d_mfma_pairs(int kw, int kh) { return (((kh * mf_nc(kw) + 1) / 2 + PD - 1) / PD) * PD; }

// Weight fragments in register order: for part p, pair s, lane l, element e:
// chunk q = 2 s + (l >= 32), row j = q / NC, chunk column cc = q % NC,
// k = 8 ((l % 32) / 16) + e, n = l % 16, tap i = 16 cc + k - n.
void f2d_mfma_fragments(const double* kd, int kw, int kh, int precision, std::vector<uint16_t>& out,
                        double& wscale) {
    const int NC = mf_nc(kw), pairs = f2d_mfma_pairs(kw, kh);
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
                double rest = w;
                for (int p = 0; p < parts; ++p) {
                    uint16_t bits;
                    double got;
                    if (bf) {
                        const __bf16 h = static_cast<__bf16>(static_cast<float>(rest));
                        std::memcpy(&bits, &h, 2);
                        got = static_cast<double>(static_cast<float>(h));
                    } else {
                        const _Float16 h = static_cast<_Float16>(static_cast<float>(rest));
                        std::memcpy(&bits, &h, 2);
                        got = static_cast<double>(static_cast<float>(h));
                    }
                    out[((static_cast<size_t>(p) * pairs + s) * 64 + l) * 8 + e] = bits;
                    rest -= got;
                }
            }
}

hipError_t launch_filter2d_mfma(const F2dLaunch& l, hipStream_t s) {
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
