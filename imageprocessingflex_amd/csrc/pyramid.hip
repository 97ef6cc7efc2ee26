// Gaussian pyramid steps (ipf::pyrDown / ipf::pyrUp) for gfx950.
//
// Reference: include/ipf/geometry.hpp:501-524. pyrDown filters with the 5-tap
// binomial kernel {1, 4, 6, 4, 1} / 16 (sepFilter2D<double>, Reflect101) and
// keeps the samples at even (x, y), rounded once with saturate_cast; pyrUp
// filters the zero-inserted image of twice the size (source sample (x, y) at
// (2x, 2y), zeros elsewhere) with the kernel / 8.
//
// Each thread produces two horizontally adjacent outputs of one channel in
// R = 4 consecutive rows, loading the samples the pair shares once and
// computing the horizontal sums of the rows the R outputs share once.
// The arithmetic is the reference's, operation for operation: every
// horizontal sum starts at +0.0 and adds w * v in tap order, every vertical
// sum starts at delta = +0.0 and adds w * h in row order, FP64 with separate
// multiply and add. So every output is bit-identical for every pixel type.
//
// pyrUp leaves out the taps that meet an inserted zero and the rows that are
// all zeros: w * 0.0 is +0.0, and s + (+0.0) == s for every s the sums can
// hold (a sum that starts at +0.0 is never -0.0 under round-to-nearest: a
// zero sum of nonzero terms is +0.0, and +0.0 + -0.0 is +0.0).
//
// Per 8 outputs pyrDown loads 77 samples and runs 150 multiply-adds, pyrUp
// loads 12 and runs 44: in integers for 8- and 16-bit images (exact, see
// pyr_int_path), otherwise in FP64 as separate multiply and add. Loads address a
// wave-uniform row base plus a 32-bit per-lane column offset.
#include "kernels.hpp"

namespace ipfg {
namespace {

constexpr int kPyrRows = 4;  // outputs per thread (one column)

__device__ __forceinline__ double pyr_madd(double s, double w, double v) { return __dadd_rn(s, __dmul_rn(w, v)); }

// Sample at byte offset off of a row (the row base is wave-uniform, the
// offset a 32-bit per-lane value).
template <class T>
__device__ __forceinline__ double pyr_ld(const char* row, int off) {
    return static_cast<double>(*reinterpret_cast<const T*>(row + off));
}

// 8- and 16-bit integer images: with the dyadic weights every FP64 sum of the
// reference is exact, so its result is the integer sum S (weights 1, 4, 6, 4,
// 1 along both axes) over 2^sh (256 for pyrDown, 64 for pyrUp), rounded once
// half away from zero; |S| < 2^25 fits an int. The values stay inside the
// input range (the weights of every output sum to 1), so nothing saturates.
template <class T>
constexpr bool pyr_int_path() {
    return std::is_integral<T>::value && sizeof(T) <= 2;
}

template <class T, int SH>
__device__ __forceinline__ void pyr_store_int(const DevImage& im, int q, int y, int S) {
    constexpr int half = 1 << (SH - 1);
    const int r = S >= 0 ? (S + half) >> SH : -((half - S) >> SH);
    *reinterpret_cast<T*>(static_cast<char*>(im.data) + static_cast<size_t>(y) * im.stride + static_cast<size_t>(q) * sizeof(T)) =
        static_cast<T>(r);
}

template <class T>
__device__ __forceinline__ int pyr_ldi(const char* row, int off) {
    return static_cast<int>(*reinterpret_cast<const T*>(row + off));
}

template <class T>
__device__ __forceinline__ void pyr_store(const DevImage& im, int q, int y, double v) {
    *reinterpret_cast<T*>(static_cast<char*>(im.data) + static_cast<size_t>(y) * im.stride + static_cast<size_t>(q) * sizeof(T)) =
        saturate_from_double<T>(v);
}

// pyrDown: output (x, y) is the blurred sample (2x, 2y). Blurred row 2y sums
// source rows border(2y + j - 2), j = 0..4. A thread produces the column pair
// x = 2 xp, 2 xp + 1 (source columns 4 xp - 2 .. 4 xp + 4, shared) of channel
// c for R rows, which use the 2R + 3 padded rows starting at 2 y0.
template <class T>
__global__ __launch_bounds__(kBlock) void pyr_down_kernel(DevImage src, DevImage dst) {
    const double k[5] = {1 / 16.0, 4 / 16.0, 6 / 16.0, 4 / 16.0, 1 / 16.0};
    const int cn = src.cn;
    const int npair = (dst.w + 1) / 2;
    const long long qq = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    if (qq >= static_cast<long long>(npair) * cn) return;
    const int xp = static_cast<int>(qq) / cn, c = static_cast<int>(qq) - xp * cn;
    const bool two = 2 * xp + 1 < dst.w;
    const int sz = static_cast<int>(sizeof(T));
    int off[7];
#pragma unroll
    for (int t = 0; t < 7; ++t) off[t] = (border_index(4 * xp - 2 + t, src.w, IPFG_BORDER_REFLECT101) * cn + c) * sz;
    const int q0 = (2 * xp) * cn + c, q1 = q0 + cn;  // output sample columns
    const char* base = static_cast<const char*>(src.data);
    for (int y0 = blockIdx.y * kPyrRows; y0 < dst.h; y0 += gridDim.y * kPyrRows) {
        const int ny = min(kPyrRows, dst.h - y0);
        if constexpr (pyr_int_path<T>()) {
            constexpr int b[5] = {1, 4, 6, 4, 1};
            int h0[2 * kPyrRows + 3], h1[2 * kPyrRows + 3];
#pragma unroll
            for (int p = 0; p < 2 * kPyrRows + 3; ++p) {
                if (p < 2 * ny + 3) {
                    const int sy = border_index(2 * y0 + p - 2, src.h, IPFG_BORDER_REFLECT101);
                    const char* row = base + static_cast<size_t>(sy) * src.stride;
                    int v[7];
#pragma unroll
                    for (int t = 0; t < 7; ++t) v[t] = pyr_ldi<T>(row, off[t]);
                    int t0 = 0, t1 = 0;
#pragma unroll
                    for (int i = 0; i < 5; ++i) {
                        t0 += b[i] * v[i];
                        t1 += b[i] * v[i + 2];
                    }
                    h0[p] = t0;
                    h1[p] = t1;
                }
            }
#pragma unroll
            for (int m = 0; m < kPyrRows; ++m) {
                if (m < ny) {
                    int S0 = 0, S1 = 0;
#pragma unroll
                    for (int j = 0; j < 5; ++j) {
                        S0 += b[j] * h0[2 * m + j];
                        S1 += b[j] * h1[2 * m + j];
                    }
                    pyr_store_int<T, 8>(dst, q0, y0 + m, S0);
                    if (two) pyr_store_int<T, 8>(dst, q1, y0 + m, S1);
                }
            }
            continue;
        }
        double h0[2 * kPyrRows + 3], h1[2 * kPyrRows + 3];
#pragma unroll
        for (int p = 0; p < 2 * kPyrRows + 3; ++p) {
            if (p < 2 * ny + 3) {
                const int sy = border_index(2 * y0 + p - 2, src.h, IPFG_BORDER_REFLECT101);
                const char* row = base + static_cast<size_t>(sy) * src.stride;
                double v[7];
#pragma unroll
                for (int t = 0; t < 7; ++t) v[t] = pyr_ld<T>(row, off[t]);
                double t0 = 0.0, t1 = 0.0;
#pragma unroll
                for (int i = 0; i < 5; ++i) {
                    t0 = pyr_madd(t0, k[i], v[i]);
                    t1 = pyr_madd(t1, k[i], v[i + 2]);
                }
                h0[p] = t0;
                h1[p] = t1;
            }
        }
#pragma unroll
        for (int m = 0; m < kPyrRows; ++m) {
            if (m < ny) {
                double a0 = 0.0, a1 = 0.0;
#pragma unroll
                for (int j = 0; j < 5; ++j) {
                    a0 = pyr_madd(a0, k[j], h0[2 * m + j]);
                    a1 = pyr_madd(a1, k[j], h1[2 * m + j]);
                }
                pyr_store<T>(dst, q0, y0 + m, a0);
                if (two) pyr_store<T>(dst, q1, y0 + m, a1);
            }
        }
    }
}

// pyrUp: output (X, Y) filters the zero-inserted image U (2w x 2h) over
// padded rows Y .. Y + 4 (row border(p - 2) of U). U(2x, 2y) = src(x, y) and
// U is zero at odd rows and columns; Reflect101 on an even length keeps the
// parity of a position, so the taps that meet a nonzero sample are the i
// with i = X (mod 2) and the rows j = Y (mod 2). A thread produces the column
// pair X = 2x, 2x + 1 of channel c: the even column sums source columns
// a = border(2x - 2) / 2, x, b = border(2x + 2) / 2 with weights 1, 6, 1 (/ 8),
// the odd column x and b with 4, 4 (/ 8). Its R = 4 rows (y0 even) use the 4
// even padded rows y0, y0 + 2, y0 + 4, y0 + 6.
template <class T>
__global__ __launch_bounds__(kBlock) void pyr_up_kernel(DevImage src, DevImage dst) {
    static_assert(kPyrRows == 4, "the row pattern below assumes 4 rows per thread");
    const double k[5] = {1 / 8.0, 4 / 8.0, 6 / 8.0, 4 / 8.0, 1 / 8.0};
    const int cn = src.cn;
    const long long qq = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    if (qq >= static_cast<long long>(src.w) * cn) return;
    const int x = static_cast<int>(qq) / cn, c = static_cast<int>(qq) - x * cn;
    const int sz = static_cast<int>(sizeof(T));
    const int oa = ((border_index(2 * x - 2, dst.w, IPFG_BORDER_REFLECT101) >> 1) * cn + c) * sz;
    const int ox = (x * cn + c) * sz;
    const int ob = ((border_index(2 * x + 2, dst.w, IPFG_BORDER_REFLECT101) >> 1) * cn + c) * sz;
    const int q0 = (2 * x) * cn + c, q1 = q0 + cn;
    const char* base = static_cast<const char*>(src.data);
    for (int y0 = blockIdx.y * kPyrRows; y0 < dst.h; y0 += gridDim.y * kPyrRows) {
        const int ny = min(kPyrRows, dst.h - y0);
        if constexpr (pyr_int_path<T>()) {
            int he[4], ho[4];  // even / odd column sums of the 4 even padded rows
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int r = border_index(y0 + 2 * e - 2, dst.h, IPFG_BORDER_REFLECT101);
                const char* row = base + static_cast<size_t>(r >> 1) * src.stride;
                const int va = pyr_ldi<T>(row, oa), vx = pyr_ldi<T>(row, ox), vb = pyr_ldi<T>(row, ob);
                he[e] = va + 6 * vx + vb;
                ho[e] = 4 * vx + 4 * vb;
            }
#pragma unroll
            for (int m = 0; m < kPyrRows; ++m) {
                if (m < ny) {
                    const int Se = m % 2 == 0 ? he[m / 2] + 6 * he[m / 2 + 1] + he[m / 2 + 2]
                                              : 4 * he[(m + 1) / 2] + 4 * he[(m + 1) / 2 + 1];
                    const int So = m % 2 == 0 ? ho[m / 2] + 6 * ho[m / 2 + 1] + ho[m / 2 + 2]
                                              : 4 * ho[(m + 1) / 2] + 4 * ho[(m + 1) / 2 + 1];
                    pyr_store_int<T, 6>(dst, q0, y0 + m, Se);
                    pyr_store_int<T, 6>(dst, q1, y0 + m, So);
                }
            }
            continue;
        }
        double he[4], ho[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int r = border_index(y0 + 2 * e - 2, dst.h, IPFG_BORDER_REFLECT101);  // even
            const char* row = base + static_cast<size_t>(r >> 1) * src.stride;
            const double va = pyr_ld<T>(row, oa), vx = pyr_ld<T>(row, ox), vb = pyr_ld<T>(row, ob);
            double t = pyr_madd(0.0, k[0], va);
            t = pyr_madd(t, k[2], vx);
            he[e] = pyr_madd(t, k[4], vb);
            ho[e] = pyr_madd(pyr_madd(0.0, k[1], vx), k[3], vb);
        }
#pragma unroll
        for (int m = 0; m < kPyrRows; ++m) {
            if (m < ny) {
                double ae, ao;
                if (m % 2 == 0) {
                    ae = pyr_madd(pyr_madd(pyr_madd(0.0, k[0], he[m / 2]), k[2], he[m / 2 + 1]), k[4], he[m / 2 + 2]);
                    ao = pyr_madd(pyr_madd(pyr_madd(0.0, k[0], ho[m / 2]), k[2], ho[m / 2 + 1]), k[4], ho[m / 2 + 2]);
                } else {
                    ae = pyr_madd(pyr_madd(0.0, k[1], he[(m + 1) / 2]), k[3], he[(m + 1) / 2 + 1]);
                    ao = pyr_madd(pyr_madd(0.0, k[1], ho[(m + 1) / 2]), k[3], ho[(m + 1) / 2 + 1]);
                }
                pyr_store<T>(dst, q0, y0 + m, ae);
                pyr_store<T>(dst, q1, y0 + m, ao);
            }
        }
    }
}

template <class T>
hipError_t launch_pyr_t(const DevImage& src, const DevImage& dst, bool up, hipStream_t s) {
    // one thread per column pair: pyrDown (dst.w + 1) / 2 pairs, pyrUp src.w
    const long long cols = static_cast<long long>(up ? src.w : (dst.w + 1) / 2) * dst.cn;
    const long long gx = (cols + kBlock - 1) / kBlock;
    const long long gy = std::min<long long>((dst.h + kPyrRows - 1) / kPyrRows, 65535);
    if (gx > 0x7fffffffLL) return hipErrorInvalidConfiguration;
    const dim3 grid(static_cast<unsigned>(gx), static_cast<unsigned>(gy));
    if (up) hipLaunchKernelGGL(pyr_up_kernel<T>, grid, dim3(kBlock), 0, s, src, dst);
    else hipLaunchKernelGGL(pyr_down_kernel<T>, grid, dim3(kBlock), 0, s, src, dst);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_pyr(const DevImage& src, const DevImage& dst, bool up, hipStream_t s) {
    if (src.dt != dst.dt || src.cn != dst.cn) return hipErrorInvalidValue;
    if (up ? (dst.w != 2 * src.w || dst.h != 2 * src.h) : (dst.w != (src.w + 1) / 2 || dst.h != (src.h + 1) / 2))
        return hipErrorInvalidValue;
    if (static_cast<long long>(dst.w) * dst.h == 0) return hipSuccess;
    // per-lane column offsets are 32-bit
    if (static_cast<long long>(src.w) * src.cn * static_cast<long long>(dtype_size(src.dt)) > 0x7fffffffLL)
        return hipErrorInvalidValue;
    switch (src.dt) {
        case IPFG_U8: return launch_pyr_t<uint8_t>(src, dst, up, s);
        case IPFG_U16: return launch_pyr_t<uint16_t>(src, dst, up, s);
        case IPFG_I16: return launch_pyr_t<int16_t>(src, dst, up, s);
        case IPFG_I32: return launch_pyr_t<int32_t>(src, dst, up, s);
        case IPFG_F32: return launch_pyr_t<float>(src, dst, up, s);
        case IPFG_F64: return launch_pyr_t<double>(src, dst, up, s);
    }
    return hipErrorInvalidValue;
}

}  // namespace ipfg
