// unsharpMask on uint8 images in one launch (filter.hpp:472-483).
//
// The reference blurs in FP64 (gaussianBlur<double>, Reflect101), then sets
// out = saturate(|v - b| > threshold ? v + amount (v - b) : v). Here each
// thread blurs 4 consecutive flat samples (one dword of an interleaved row)
// in 4 consecutive rows with FP32 separable sums held in registers: per input
// row, the horizontal taps read a window of 1 + 2 DH dwords, and the vertical
// taps fold that row into the outputs it reaches. FP32 chains of ks taps with
// weights rounded to float stay within `bnd` of the reference's blur; every
// diff v - b within that bound of v - b32 is then tried (threshold branch and
// rounding of v + amount diff): when all give the same byte, it is the
// reference's byte. The rest are recomputed by the whole wave in the
// reference's order (sep_exact_wave). Workgroups past the interior compute
// the frame (rows within ks / 2 of the top or bottom, the first and last
// samples of a row) one sample per thread in the reference's order.
#include "sep_kernels.hpp"

namespace ipfg {
namespace {

constexpr int kUnsharpMaxK = 13;
constexpr int kUnsharpRows = 4;

struct UnsharpArgs {
    DevImage src, dst;
    const double* kd;  // ks FP64 weights (+ 16 zeros), device
    float kf[kUnsharpMaxK];
    double amount, threshold, bnd;
    int L, glo, GI, HI, nbi, nblocks, lside, rside;
    long long nframe;
};

template <int N>
__device__ __forceinline__ float ubyte(const unsigned (&w)[N], int q) {
    return static_cast<float>((w[q >> 2] >> (8 * (q & 3))) & 0xffu);
}

// The reference's result from its exact blur b.
__device__ __forceinline__ uint8_t unsharp_out(double v, double b, double amount, double threshold) {
    const double diff = v - b;
    return saturate_from_double<uint8_t>(fabs(diff) > threshold ? v + amount * diff : v);
}

// The byte every blur within bnd of b32 gives, or -1 when they differ.
__device__ __forceinline__ int unsharp_sure(double v, float b32, double amount, double threshold, double bnd) {
    const double d = v - static_cast<double>(b32);
    const double w = bnd + 2.3e-16 * fabs(d);
    const double lo = d - w, hi = d + w;
    const double mina = (lo <= 0.0 && hi >= 0.0) ? 0.0 : fmin(fabs(lo), fabs(hi));
    const double maxa = fmax(fabs(lo), fabs(hi));
    const int keep = saturate_from_double<uint8_t>(v);
    if (!(maxa > threshold)) return keep;  // the sample is kept whatever the blur
    const double r1 = v + amount * lo, r2 = v + amount * hi;
    const double eps = 4.5e-16 * (fabs(v) + fabs(amount) * maxa) + 1e-300;
    const int o1 = saturate_from_double<uint8_t>(fmin(r1, r2) - eps);
    const int o2 = saturate_from_double<uint8_t>(fmax(r1, r2) + eps);
    if (o1 != o2) return -1;
    if (!(mina > threshold) && o1 != keep) return -1;
    return o1;
}

// The reference's blur at (x, y, c) by one thread (frame samples).
template <int KS>
__device__ double unsharp_blur_seq(const DevImage& src, int x, int y, int c, const double* __restrict__ k) {
    constexpr int R = KS / 2;
    double acc = 0.0;
    for (int j = 0; j < KS; ++j) {
        const int sy = border_index(y + j - R, src.h, IPFG_BORDER_REFLECT101);
        const uint8_t* row = static_cast<const uint8_t*>(src.data) + static_cast<size_t>(sy) * src.stride;
        double t = 0.0;
        for (int i = 0; i < KS; ++i) {
            const int sx = border_index(x + i - R, src.w, IPFG_BORDER_REFLECT101);
            t = t + k[i] * static_cast<double>(row[static_cast<size_t>(sx) * src.cn + c]);
        }
        acc = acc + k[j] * t;
    }
    return acc;
}

template <int KS, int CN>
__global__ __launch_bounds__(kBlock) void unsharp_u8_kernel(UnsharpArgs a) {
    constexpr int R = KS / 2;
    constexpr int DH = (R * CN + 3) / 4;
    constexpr int ND = 1 + 2 * DH;
    constexpr int NG = kUnsharpRows;
    constexpr int NR = NG + KS - 1;
    const int b = blockIdx.x;
    if (b >= a.nbi) {  // frame: one sample per thread, the reference's order
        const long long idx = static_cast<long long>(b - a.nbi) * kBlock + threadIdx.x;
        if (idx >= a.nframe) return;
        const int H = a.src.h;
        const long long full_top = static_cast<long long>(a.HI > 0 ? R : H) * a.L;
        const long long full = a.HI > 0 ? full_top + static_cast<long long>(R) * a.L : full_top;
        int y, s;
        if (idx < full_top) {
            y = static_cast<int>(idx / a.L);
            s = static_cast<int>(idx - static_cast<long long>(y) * a.L);
        } else if (idx < full) {
            const long long k = idx - full_top;
            const int rr = static_cast<int>(k / a.L);
            y = H - R + rr;
            s = static_cast<int>(k - static_cast<long long>(rr) * a.L);
        } else {
            const int sc = a.lside + a.rside;
            const long long k = idx - full;
            const int rr = static_cast<int>(k / sc);
            const int t = static_cast<int>(k - static_cast<long long>(rr) * sc);
            y = R + rr;
            s = t < a.lside ? t : a.L - a.rside + (t - a.lside);
        }
        const int px = s / CN, c = s - px * CN;
        const uint8_t* srow = static_cast<const uint8_t*>(a.src.data) + static_cast<size_t>(y) * a.src.stride;
        const double v = static_cast<double>(srow[s]);
        uint8_t* drow = static_cast<uint8_t*>(a.dst.data) + static_cast<size_t>(y) * a.dst.stride;
        drow[s] = unsharp_out(v, unsharp_blur_seq<KS>(a.src, px, y, c, a.kd), a.amount, a.threshold);
        return;
    }
    // consecutive workgroups go to the 8 XCDs in turn: give each XCD a
    // contiguous band of rows
    const int blk = (b & 7) * (a.nbi >> 3) + (b >> 3);
    if (blk >= a.nblocks) return;  // whole workgroups: the wave-wide fix-ups below need every lane
    const unsigned nrb = (static_cast<unsigned>(a.HI) + NG - 1) / NG;
    const unsigned n = static_cast<unsigned>(a.GI) * nrb;
    const unsigned idx0 = static_cast<unsigned>(blk) * kBlock + threadIdx.x;
    const bool act = idx0 < n;
    const unsigned idx = act ? idx0 : n - 1u;  // idle lanes redo the last group and store nothing
    const unsigned rb = idx / static_cast<unsigned>(a.GI);
    const int s0 = (static_cast<int>(idx - rb * static_cast<unsigned>(a.GI)) + a.glo) * 4;
    const int r0 = static_cast<int>(rb) * NG;  // output row R + r0 + k reads input rows r0 + k .. r0 + k + KS - 1
    const int H = a.src.h;
    float acc[NG][4];
#pragma unroll
    for (int k = 0; k < NG; ++k)
#pragma unroll
        for (int m = 0; m < 4; ++m) acc[k][m] = 0.f;
    unsigned cw[NG] = {};  // the thread's own dword of each output row
    // input rows one at a time (a wave-uniform loop), so only one row's window
    // is live; other waves cover the load latency
#pragma unroll 2
    for (int t = 0; t < NR; ++t) {
        const int rr = min(r0 + t, H - 1);  // rows past the image feed only rows that are not stored
        const unsigned* p =
            reinterpret_cast<const unsigned*>(static_cast<const uint8_t*>(a.src.data) + static_cast<size_t>(rr) * a.src.stride + s0) - DH;
        unsigned w[ND];
#pragma unroll
        for (int d = 0; d < ND; ++d) w[d] = p[d];
#pragma unroll
        for (int k = 0; k < NG; ++k)
            if (t == k + R) cw[k] = w[DH];
        float h[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            h[m] = 0.f;
#pragma unroll
            for (int i = 0; i < KS; ++i) h[m] = fmaf(a.kf[i], ubyte(w, 4 * DH + m + (i - R) * CN), h[m]);
        }
#pragma unroll
        for (int k = 0; k < NG; ++k) {
            const int j = t - k;
            if (j >= 0 && j < KS) {
                const float kj = a.kf[j];
#pragma unroll
                for (int m = 0; m < 4; ++m) acc[k][m] = fmaf(kj, h[m], acc[k][m]);
            }
        }
    }
    unsigned doubt = 0u;  // bit 4 k + m
#pragma unroll
    for (int k = 0; k < NG; ++k) {
        unsigned o = 0u;
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const double v = static_cast<double>((cw[k] >> (8 * m)) & 0xffu);
            const int r = unsharp_sure(v, acc[k][m], a.amount, a.threshold, a.bnd);
            if (r < 0) doubt |= 1u << (4 * k + m);
            o |= static_cast<unsigned>(r < 0 ? 0 : r) << (8 * m);
        }
        if (act && r0 + k < a.HI) {
            uint8_t* drow = static_cast<uint8_t*>(a.dst.data) + static_cast<size_t>(R + r0 + k) * a.dst.stride;
            *reinterpret_cast<unsigned*>(drow + s0) = o;
        }
    }
    if (!act) doubt = 0u;
    for (int k = 0; k < NG; ++k)
        if (r0 + k >= a.HI) doubt &= ~(0xfu << (4 * k));
    unsigned long long bal = __ballot(doubt != 0u);
    while (bal) {
        const int l = __builtin_ctzll(bal);
        bal &= bal - 1ull;
        unsigned mm = static_cast<unsigned>(bcast_i(static_cast<int>(doubt), l));
        const int ls0 = bcast_i(s0, l), lr0 = bcast_i(r0, l);
        while (mm) {
            const int bit = __builtin_ctz(mm);
            mm &= mm - 1u;
            const int k = bit >> 2, m = bit & 3;
            const int sidx = ls0 + m, px = sidx / CN, c = sidx - px * CN, y = R + lr0 + k;
            const double bx = sepk::sep_exact_wave<uint8_t>(a.src, px, y, c, a.kd, KS, a.kd, KS, R, R,
                                                            IPFG_BORDER_REFLECT101, 0.0, 0.0);
            if (lane_id() == l) {
                uint8_t* drow = static_cast<uint8_t*>(a.dst.data) + static_cast<size_t>(y) * a.dst.stride;
                const uint8_t* srow = static_cast<const uint8_t*>(a.src.data) + static_cast<size_t>(y) * a.src.stride;
                drow[sidx] = unsharp_out(static_cast<double>(srow[sidx]), bx, a.amount, a.threshold);
            }
        }
    }
}

template <int KS>
hipError_t unsharp_by_cn(const UnsharpArgs& a, int cn, dim3 grid, hipStream_t s) {
    switch (cn) {
        case 1: hipLaunchKernelGGL((unsharp_u8_kernel<KS, 1>), grid, dim3(kBlock), 0, s, a); break;
        case 3: hipLaunchKernelGGL((unsharp_u8_kernel<KS, 3>), grid, dim3(kBlock), 0, s, a); break;
        case 4: hipLaunchKernelGGL((unsharp_u8_kernel<KS, 4>), grid, dim3(kBlock), 0, s, a); break;
        default: return hipErrorNotSupported;
    }
    return hipGetLastError();
}

}  // namespace

hipError_t launch_unsharp_u8(const DevImage& src, const DevImage& dst, int ks, const double* k_host,
                             const double* k_dev, double amount, double threshold, double bnd, hipStream_t s) {
    const int cn = src.cn;
    if (src.dt != IPFG_U8 || dst.dt != IPFG_U8) return hipErrorNotSupported;
    if (ks < 3 || ks > kUnsharpMaxK || ks % 2 == 0) return hipErrorNotSupported;
    if (cn != 1 && cn != 3 && cn != 4) return hipErrorNotSupported;
    if ((reinterpret_cast<uintptr_t>(src.data) & 15u) || (src.stride & 15u) ||
        (reinterpret_cast<uintptr_t>(dst.data) & 15u) || (dst.stride & 15u))
        return hipErrorNotSupported;
    const long long L = static_cast<long long>(src.w) * cn;
    if (L <= 0 || src.h <= 0 || L > (1LL << 30)) return hipErrorNotSupported;
    UnsharpArgs a{};
    a.src = src;
    a.dst = dst;
    a.kd = k_dev;
    for (int i = 0; i < ks; ++i) a.kf[i] = static_cast<float>(k_host[i]);
    a.amount = amount;
    a.threshold = threshold;
    a.bnd = bnd;
    a.L = static_cast<int>(L);
    const int R = ks / 2;
    const int dh = (R * cn + 3) / 4;
    // interior groups g: 4 g >= 4 dh and 4 g + 4 + 4 dh <= L; rows R .. H - R - 1
    a.glo = dh;
    const int ghi = (a.L - 4 - 4 * dh) >= 0 ? (a.L - 4 - 4 * dh) / 4 : -1;
    a.GI = ghi >= a.glo ? ghi - a.glo + 1 : 0;
    a.HI = src.h > 2 * R ? src.h - 2 * R : 0;
    if (a.GI == 0 || a.HI == 0) {
        a.GI = a.HI = 0;
        a.lside = a.rside = 0;
        a.nframe = static_cast<long long>(src.h) * a.L;
    } else {
        a.lside = 4 * a.glo;
        a.rside = a.L - 4 * (ghi + 1);
        a.nframe = static_cast<long long>(2 * R) * a.L + static_cast<long long>(a.HI) * (a.lside + a.rside);
    }
    const long long nint = static_cast<long long>(a.GI) * ((a.HI + kUnsharpRows - 1) / kUnsharpRows);
    const long long nb = (nint + kBlock - 1) / kBlock;
    const long long nbf = (a.nframe + kBlock - 1) / kBlock;
    if (nint > (1LL << 31) - 1 || nb + nbf + 8 > (1LL << 31) - 1) return hipErrorNotSupported;
    a.nblocks = static_cast<int>(nb);
    a.nbi = static_cast<int>((nb + 7) / 8 * 8);
    const dim3 grid(static_cast<unsigned>(a.nbi + nbf));
    switch (ks) {
        case 3: return unsharp_by_cn<3>(a, cn, grid, s);
        case 5: return unsharp_by_cn<5>(a, cn, grid, s);
        case 7: return unsharp_by_cn<7>(a, cn, grid, s);
        case 9: return unsharp_by_cn<9>(a, cn, grid, s);
        case 11: return unsharp_by_cn<11>(a, cn, grid, s);
        default: return unsharp_by_cn<13>(a, cn, grid, s);
    }
}

}  // namespace ipfg
