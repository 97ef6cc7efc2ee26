// filter2DFFT for gfx950 (reference: include/ipf/fft.hpp:21-98, 163-216).
//
// The reference correlates through FP64 radix-2 FFTs: per channel, the padded
// image (real part) and the kernel (imaginary part) share one complex
// pow2 x pow2 buffer; one forward 2-D FFT (rows, then columns); the two real
// spectra are separated and multiplied; one inverse 2-D FFT; the real part,
// scaled, is the correlation. Every butterfly of a radix-2 decimation-in-time
// transform depends only on its two inputs and its twiddle, so a GPU schedule
// that runs the same butterflies stage after stage, with the same twiddle
// values (computed on the host with the same libm calls) and the same complex
// arithmetic, gives the reference's numbers bit for bit. That is the design:
//
//  fft_pack_kernel      builds the complex buffer already in bit-reversed
//                       order along both axes (the permutation every 1-D
//                       transform starts with), padding by the border mode
//  fft_stages_kernel    runs up to 7 consecutive stages of the 1-D transforms
//                       of every row or every column: a workgroup stages a
//                       tile of 2048 complex values (groups of 2^stages
//                       elements that only meet each other in those stages,
//                       16+ groups side by side so every global access is a
//                       run of >= 256 contiguous bytes) in LDS, applies the
//                       stages there and writes the tile back in place
//  fft_separate_kernel  separates the image and kernel spectra and forms
//                       F_img conj(F_ker), written bit-reversed for the
//                       inverse transform
//  fft_extract_kernel   real part x 1 / (fw fh) + delta, saturated
//
// A 2^13-point axis takes two passes (7 + 6 stages); the buffers stay in HBM
// between passes. Complex products are (ac - bd, ad + bc) with separate
// multiplies and adds, as GCC evaluates std::complex<double> products of
// finite values.
#include <algorithm>

#include "fft.hpp"

namespace ipfg {
namespace {

constexpr int kTile = 2048;  // complex doubles per workgroup tile (32 KB of LDS)

__device__ __forceinline__ unsigned bitrev(unsigned v, int bits) {
    return bits ? __brev(v) >> (32 - bits) : 0u;
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// The complex buffer of one channel: element (x, y) of the padded image in the
// real part, of the kernel in the imaginary part, zero elsewhere; stored at
// (bitrev(x), bitrev(y)). One thread per stored element.
template <class T>
__global__ __launch_bounds__(kBlock) void fft_pack_kernel(DevImage src, int c, int kw, int kh, int border, double bval,
                                                          const float* __restrict__ K, double2* __restrict__ A,
                                                          int lw, int lh) {
    const int fw = 1 << lw;
    const int X = blockIdx.x * kBlock + threadIdx.x;
    const int Y = blockIdx.y;
    if (X >= fw) return;
    const int x = static_cast<int>(bitrev(static_cast<unsigned>(X), lw));
    const int y = static_cast<int>(bitrev(static_cast<unsigned>(Y), lh));
    const int ax = kw / 2, ay = kh / 2;
    const int PW = src.w + kw - 1, PH = src.h + kh - 1;
    double re = 0.0, im = 0.0;
    if (x < PW && y < PH) {
        const int sx = border_index(x - ax, src.w, border), sy = border_index(y - ay, src.h, border);
        re = (sx < 0 || sy < 0)
                 ? bval
                 : static_cast<double>(reinterpret_cast<const T*>(static_cast<const char*>(src.data) +
                                                                  static_cast<size_t>(sy) * src.stride)[static_cast<size_t>(sx) * src.cn + c]);
    }
    if (x < kw && y < kh) im = static_cast<double>(K[static_cast<size_t>(y) * kw + x]);
    A[static_cast<size_t>(Y) * fw + X] = make_double2(re, im);
}

// Stages s0 + 1 .. s0 + S of the radix-2 transforms of length n = 2^L along
// one axis. Transform b's element i sits at data[b * ts + i * es]; rows have
// es = 1 and ts = n, columns es = n_rows and ts = 1, count = 2^lc transforms.
// The elements that meet in these stages form groups of G = 2^S with spacing
// D = 2^s0: i = hi 2^(s0+S) + m D + lo. Groups are numbered so that
// neighbours are neighbours in memory (rows: by b, hi, lo; columns: by hi, lo,
// b) and a tile holds TE = min(kTile, count n) values, ng = TE / G
// consecutive groups. Stage j (butterfly span 2^(j-1)) pairs m with m + h,
// h = 2^(j-1-s0), and takes the twiddle of index (m mod 2h) D + lo from the
// table of length 2^j (tw + 2^(j-1) - 1). In LDS, group g's element m sits at
// g G + (m XOR (g mod G)), so that neither the tile loads (consecutive
// groups) nor the butterflies (consecutive m) meet in one bank.
__global__ __launch_bounds__(kBlock) void fft_stages_kernel(double2* __restrict__ data, int L, int s0, int S, int lc,
                                                            int rows, const double2* __restrict__ tw) {
    __shared__ double2 t[kTile];
    const int lte = min(11, L + lc);
    const int TE = 1 << lte, G = 1 << S, lng = lte - S, D = 1 << s0;
    const size_t first = static_cast<size_t>(blockIdx.x) << lng;  // the tile's first group
    const int nrow = 1 << L;
    // address of element m of the tile's group g, and the group's lo
    auto locate = [&](int g, int* lo_out) -> size_t {
        const size_t gid = first + g;
        size_t b, pos;
        if (rows) {
            b = gid >> (L - S);
            pos = gid & ((static_cast<size_t>(1) << (L - S)) - 1);
        } else {
            b = gid & ((static_cast<size_t>(1) << lc) - 1);
            pos = gid >> lc;
        }
        const size_t hi = pos >> s0;
        const int lo = static_cast<int>(pos & (D - 1));
        *lo_out = lo;
        const size_t i0 = (hi << (s0 + S)) + lo;  // element m adds m D
        return rows ? b * nrow + i0 : b + i0 * (static_cast<size_t>(1) << lc);
    };
    const size_t mstride = rows ? static_cast<size_t>(D) : static_cast<size_t>(D) << lc;
    const bool grouped_runs = rows && D == 1;  // each group is contiguous: walk m fastest
    for (int e = threadIdx.x; e < TE; e += kBlock) {
        int g, m, lo;
        if (grouped_runs) {
            g = e >> S;
            m = e & (G - 1);
        } else {
            m = e >> lng;
            g = e & ((1 << lng) - 1);
        }
        t[g * G + (m ^ (g & (G - 1)))] = data[locate(g, &lo) + m * mstride];
    }
    __syncthreads();
    for (int st = 0; st < S; ++st) {
        const int h = 1 << st;
        const double2* w = tw + ((1 << (s0 + st)) - 1);  // stage j = s0 + st + 1
        for (int q = threadIdx.x; q < TE / 2; q += kBlock) {
            const int g = q >> (S - 1);
            const int r = q & ((G >> 1) - 1);
            const int m = ((r >> st) << (st + 1)) | (r & (h - 1));
            int lo;
            (void)locate(g, &lo);
            const double2 wk = w[(m & (2 * h - 1)) * D + lo];
            const int i0 = g * G + (m ^ (g & (G - 1))), i1 = i0 ^ h;  // m has bit st clear
            const double2 u = t[i0];
            const double2 v = cmul(t[i1], wk);
            t[i0] = make_double2(u.x + v.x, u.y + v.y);
            t[i1] = make_double2(u.x - v.x, u.y - v.y);
        }
        __syncthreads();
    }
    for (int e = threadIdx.x; e < TE; e += kBlock) {
        int g, m, lo;
        if (grouped_runs) {
            g = e >> S;
            m = e & (G - 1);
        } else {
            m = e >> lng;
            g = e & ((1 << lng) - 1);
        }
        data[locate(g, &lo) + m * mstride] = t[g * G + (m ^ (g & (G - 1)))];
    }
}

// B(v, u) = F_img(v, u) conj(F_ker(v, u)) with F_img = (Z(k) + conj Z(-k)) / 2
// and F_ker = (Z(k) - conj Z(-k)) (-i / 2), written at (bitrev(u), bitrev(v)).
// One thread per stored element.
__global__ __launch_bounds__(kBlock) void fft_separate_kernel(const double2* __restrict__ A, double2* __restrict__ B,
                                                              int lw, int lh) {
    const int fw = 1 << lw, fh = 1 << lh;
    const int U = blockIdx.x * kBlock + threadIdx.x;
    const int V = blockIdx.y;
    if (U >= fw) return;
    const int u = static_cast<int>(bitrev(static_cast<unsigned>(U), lw));
    const int v = static_cast<int>(bitrev(static_cast<unsigned>(V), lh));
    const double2 z = A[static_cast<size_t>(v) * fw + u];
    const double2 mz = A[static_cast<size_t>((fh - v) & (fh - 1)) * fw + ((fw - u) & (fw - 1))];
    const double2 zm = make_double2(mz.x, -mz.y);  // conj
    const double2 fi = make_double2((z.x + zm.x) * 0.5, (z.y + zm.y) * 0.5);
    const double2 d = make_double2(z.x - zm.x, z.y - zm.y);
    const double2 fk = cmul(d, make_double2(0.0, -0.5));
    B[static_cast<size_t>(V) * fw + U] = cmul(fi, make_double2(fk.x, -fk.y));
}

template <class To>
__global__ __launch_bounds__(kBlock) void fft_extract_kernel(const double2* __restrict__ B, int lw, int c, double s,
                                                             double delta, DevImage dst) {
    const int x = blockIdx.x * kBlock + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= dst.w) return;
    const double r = B[(static_cast<size_t>(y) << lw) + x].x * s + delta;
    reinterpret_cast<To*>(static_cast<char*>(dst.data) + static_cast<size_t>(y) * dst.stride)[static_cast<size_t>(x) * dst.cn + c] =
        saturate_from_double<To>(r);
}

int ilog2(int n) {
    int l = 0;
    while ((1 << l) < n) ++l;
    return l;
}

// All stages of the 2^lc transforms of length 2^L along one axis (rows or
// columns), in passes of <= 7 stages.
hipError_t run_axis(double2* data, int L, int lc, bool rows, const double2* tw, hipStream_t s) {
    const int passes = (L + 6) / 7;
    const int lte = std::min(11, L + lc);
    const unsigned tiles = static_cast<unsigned>((static_cast<size_t>(1) << (L + lc)) >> lte);
    int s0 = 0;
    for (int p = 0; p < passes; ++p) {
        const int S = (L - s0) / (passes - p);  // split the remaining stages evenly
        hipLaunchKernelGGL(fft_stages_kernel, dim3(tiles), dim3(kBlock), 0, s, data, L, s0, S, lc, rows ? 1 : 0, tw);
        s0 += S;
    }
    return hipGetLastError();
}

template <class F>
hipError_t by_type(int dt, F f) {
    switch (dt) {
        case IPFG_U8: return f(static_cast<uint8_t*>(nullptr));
        case IPFG_U16: return f(static_cast<uint16_t*>(nullptr));
        case IPFG_I16: return f(static_cast<int16_t*>(nullptr));
        case IPFG_I32: return f(static_cast<int32_t*>(nullptr));
        case IPFG_F32: return f(static_cast<float*>(nullptr));
        case IPFG_F64: return f(static_cast<double*>(nullptr));
    }
    return hipErrorInvalidValue;
}

}  // namespace

bool fft_shape_ok(int lw, int lh) {
    // grid rows of the element-wise launches (<= 65535) and 4 GiB per buffer
    return lw >= 0 && lh >= 0 && lw <= 20 && lh <= 15 && lw + lh <= 28;
}

hipError_t launch_filter2d_fft(const FftLaunch& a, hipStream_t s) {
    const int fw = 1 << a.lw, fh = 1 << a.lh;
    const dim3 grid(static_cast<unsigned>((fw + kBlock - 1) / kBlock), static_cast<unsigned>(fh));
    const double scale = 1.0 / (static_cast<double>(fw) * fh);
    for (int c = 0; c < a.src.cn; ++c) {
        hipError_t e = by_type(a.src.dt, [&](auto* tag) {
            using T = std::remove_pointer_t<decltype(tag)>;
            hipLaunchKernelGGL(fft_pack_kernel<T>, grid, dim3(kBlock), 0, s, a.src, c, a.kw, a.kh, a.border, a.bval,
                               a.kernel, a.buf0, a.lw, a.lh);
            return hipGetLastError();
        });
        if (e != hipSuccess) return e;
        // forward: every row (length fw), then every column (length fh)
        if ((e = run_axis(a.buf0, a.lw, a.lh, true, a.tw_fwd_w, s)) != hipSuccess) return e;
        if ((e = run_axis(a.buf0, a.lh, a.lw, false, a.tw_fwd_h, s)) != hipSuccess) return e;
        hipLaunchKernelGGL(fft_separate_kernel, grid, dim3(kBlock), 0, s, a.buf0, a.buf1, a.lw, a.lh);
        if ((e = run_axis(a.buf1, a.lw, a.lh, true, a.tw_inv_w, s)) != hipSuccess) return e;
        if ((e = run_axis(a.buf1, a.lh, a.lw, false, a.tw_inv_h, s)) != hipSuccess) return e;
        const dim3 og(static_cast<unsigned>((a.dst.w + kBlock - 1) / kBlock), static_cast<unsigned>(a.dst.h));
        e = by_type(a.dst.dt, [&](auto* tag) {
            using To = std::remove_pointer_t<decltype(tag)>;
            hipLaunchKernelGGL(fft_extract_kernel<To>, og, dim3(kBlock), 0, s, a.buf1, a.lw, c, scale, a.delta, a.dst);
            return hipGetLastError();
        });
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace ipfg
