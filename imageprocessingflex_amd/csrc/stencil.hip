// Small dense stencils (3x3, 5x5, centred anchor) on interleaved uint8
// images whose sums are exact in FP32, and the 3x3 median of uint8 images.
//
// Used for linear filters on uint8 images when every weight is a dyadic
// rational m * 2^-f and 255 * sum|m| (+ |delta| in units of 2^-f) stays below
// 2^23: each product and partial sum is then an exact FP32 value, so any
// summation order gives the reference's FP64 result bit for bit (sobel,
// scharr, prewitt, laplacian, sharpen, emboss and integer-valued user
// kernels; filter.hpp:157-203, 393-496).
//
// Layout: a row of W * cn interleaved samples is a flat array in which the
// horizontal neighbours of a sample sit cn samples apart. A thread owns 4
// consecutive flat samples (one dword of the row) in 8 consecutive rows and
// loads, per input row, its dword and DH halo dwords on each side, so a
// wave's loads and stores are contiguous and nothing passes through LDS. The workgroups after the
// interior ones compute the frame the interior leaves out (rows whose window
// leaves the image, the first and the last samples of each row) one sample
// per thread through border_index, in the same launch.
#include "kernels.hpp"

namespace ipfg {
namespace {

constexpr int kStencilMax = 25;

struct StencilArgs {
    const uint8_t* src;
    char* dst;
    size_t sstride, dstride;
    int W, H, L, ay, border;
    int glo, GI, HI;   // interior: groups glo .. glo + GI - 1 of rows ay .. ay + HI - 1
    int nbi;           // workgroups of the interior part (a multiple of 8)
    int nblocks;       // of which carry interior groups
    int ntop, nbot;    // frame: whole rows at the top and bottom
    int lside, rside;  // frame: the first lside and the last rside samples of each interior row
    long long nframe;  // frame samples
    float bval, delta;
    float k[kStencilMax];   // row-major kh x kw weights
    float kx[5], ky[5];     // separable launches: k = ky (x) kx
};

// Saturated uint8 of an exact float (round half away from zero): on (0, 255)
// the sum v + 0.5 is exact, so truncation rounds like the reference.
__device__ __forceinline__ unsigned sat_u8(float v) {
    return v <= 0.f ? 0u : v >= 255.f ? 255u : static_cast<unsigned>(v + 0.5f);
}

// Sample q (compile-time after unrolling) of the dword window as float.
template <int N>
__device__ __forceinline__ float win_byte(const unsigned (&w)[N], int q) {
    return static_cast<float>((w[q >> 2] >> (8 * (q & 3))) & 0xffu);
}

// Row y and flat sample s of frame sample idx: the whole top rows, the whole
// bottom rows, then the first lside and last rside samples of each interior row.
__device__ __forceinline__ void frame_pos(const StencilArgs& a, long long idx, int& y, int& s) {
    const long long full_top = static_cast<long long>(a.ntop) * a.L;
    const long long full = full_top + static_cast<long long>(a.nbot) * a.L;
    if (idx < full_top) {
        y = static_cast<int>(idx / a.L);
        s = static_cast<int>(idx - static_cast<long long>(y) * a.L);
    } else if (idx < full) {
        const long long k = idx - full_top;
        const int rr = static_cast<int>(k / a.L);
        y = a.H - a.nbot + rr;
        s = static_cast<int>(k - static_cast<long long>(rr) * a.L);
    } else {
        const int sc = a.lside + a.rside;
        const long long k = idx - full;
        const int rr = static_cast<int>(k / sc);
        const int t = static_cast<int>(k - static_cast<long long>(rr) * sc);
        y = a.ay + rr;
        s = t < a.lside ? t : a.L - a.rside + (t - a.lside);
    }
}

template <class To, int KW, int KH, int CN>
__device__ __forceinline__ void frame_sample(const StencilArgs& a, long long idx) {
    int y, s;
    frame_pos(a, idx, y, s);
    const int px = s / CN, c = s - px * CN;
    float v = a.delta;
#pragma unroll
    for (int j = 0; j < KH; ++j) {
        const int sy = border_index(y - a.ay + j, a.H, a.border);
        const uint8_t* srow = a.src + static_cast<size_t>(sy < 0 ? 0 : sy) * a.sstride;
#pragma unroll
        for (int i = 0; i < KW; ++i) {
            const int sx = border_index(px + i - KW / 2, a.W, a.border);
            const float x = (sy < 0 || sx < 0) ? a.bval : static_cast<float>(srow[static_cast<size_t>(sx) * CN + c]);
            v = fmaf(a.k[j * KW + i], x, v);
        }
    }
    char* drow = a.dst + static_cast<size_t>(y) * a.dstride;
    if constexpr (sizeof(To) == 1) reinterpret_cast<uint8_t*>(drow)[s] = static_cast<uint8_t>(sat_u8(v));
    else reinterpret_cast<float*>(drow)[s] = v;
}

// Output rows per thread: a thread computes one group of 4 samples in NG
// consecutive rows, so the NG + KH - 1 input rows it loads (and converts) are
// shared by all of them.
constexpr int kStencilRows = 8;

// Workgroups [0, nbi) compute interior groups of 4 samples (one dword of the
// row) from windows of 1 + 2 DH dwords per input row; consecutive lanes take
// consecutive groups of a row, so every load and store instruction of a wave
// covers contiguous dwords. The remaining workgroups compute the frame one
// sample per thread.
template <class To, int KW, int KH, int CN, bool SEP>
__global__ __launch_bounds__(kBlock) void stencil_u8_kernel(StencilArgs a) {
    constexpr int AX = KW / 2;
    constexpr int DH = (AX * CN + 3) / 4;  // halo dwords on each side
    constexpr int ND = 1 + 2 * DH;
    constexpr int NG = kStencilRows;
    constexpr int NR = NG + KH - 1;
    const int b = blockIdx.x;
    if (b >= a.nbi) {
        const long long idx = static_cast<long long>(b - a.nbi) * kBlock + threadIdx.x;
        if (idx < a.nframe) frame_sample<To, KW, KH, CN>(a, idx);
        return;
    }
    // consecutive workgroups go to the 8 XCDs in turn: give each XCD a
    // contiguous run of row blocks so rows shared by neighbours meet in its L2
    const int blk = (b & 7) * (a.nbi >> 3) + (b >> 3);
    if (blk >= a.nblocks) return;
    const unsigned nrb = (static_cast<unsigned>(a.HI) + NG - 1) / NG;
    const unsigned idx = static_cast<unsigned>(blk) * kBlock + threadIdx.x;  // GI * nrb < 2^31
    if (idx >= static_cast<unsigned>(a.GI) * nrb) return;
    const unsigned rb = idx / static_cast<unsigned>(a.GI);
    const int s0 = (static_cast<int>(idx - rb * static_cast<unsigned>(a.GI)) + a.glo) * 4;
    const int r0 = static_cast<int>(rb) * NG;  // first output row - ay = first input row
    unsigned w[NR][ND];
#pragma unroll
    for (int t = 0; t < NR; ++t) {
        const int rr = min(r0 + t, a.H - 1);  // rows past the image feed only rows that are not stored
        const unsigned* p = reinterpret_cast<const unsigned*>(a.src + static_cast<size_t>(rr) * a.sstride + s0) - DH;
#pragma unroll
        for (int d = 0; d < ND; ++d) w[t][d] = p[d];
    }
    // separable kernels: the horizontal sums of each input row, shared by the
    // output rows it reaches
    float h[SEP ? NR : 1][4];
    if constexpr (SEP) {
#pragma unroll
        for (int t = 0; t < NR; ++t)
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                h[t][m] = 0.f;
#pragma unroll
                for (int i = 0; i < KW; ++i) h[t][m] = fmaf(a.kx[i], win_byte(w[t], 4 * DH + m + (i - AX) * CN), h[t][m]);
            }
    }
#pragma unroll
    for (int k = 0; k < NG; ++k) {
        float acc[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) acc[m] = a.delta;
        if constexpr (SEP) {
#pragma unroll
            for (int j = 0; j < KH; ++j)
#pragma unroll
                for (int m = 0; m < 4; ++m) acc[m] = fmaf(a.ky[j], h[k + j][m], acc[m]);
        } else {
#pragma unroll
            for (int j = 0; j < KH; ++j)
#pragma unroll
                for (int i = 0; i < KW; ++i) {
                    const float kk = a.k[j * KW + i];
#pragma unroll
                    for (int m = 0; m < 4; ++m)
                        acc[m] = fmaf(kk, win_byte(w[k + j], 4 * DH + m + (i - AX) * CN), acc[m]);
                }
        }
        if (r0 + k >= a.HI) break;
        char* drow = a.dst + static_cast<size_t>(a.ay + r0 + k) * a.dstride;
        if constexpr (sizeof(To) == 1)
            *reinterpret_cast<unsigned*>(drow + s0) = sat_u8(acc[0]) | (sat_u8(acc[1]) << 8) |
                                                      (sat_u8(acc[2]) << 16) | (sat_u8(acc[3]) << 24);
        else
            *reinterpret_cast<float4*>(drow + static_cast<size_t>(s0) * sizeof(float)) =
                make_float4(acc[0], acc[1], acc[2], acc[3]);
    }
}

// Interior and frame of a KW x KH window over a (fields L, H, ay set) and the
// grid: interior groups g with 4 g >= 4 dh and 4 g + 4 + 4 dh <= L, on rows
// whose window stays inside the image; the frame is everything else.
hipError_t stencil_geometry(StencilArgs& a, int kw, int kh, int cn, dim3& grid) {
    const int ah = (kw / 2) * cn;
    const int dh = (ah + 3) / 4;
    a.glo = dh;
    const int ghi = (a.L - 4 - 4 * dh) >= 0 ? (a.L - 4 - 4 * dh) / 4 : -1;
    a.GI = ghi >= a.glo ? ghi - a.glo + 1 : 0;
    a.HI = a.H >= kh ? a.H - kh + 1 : 0;
    if (a.GI == 0 || a.HI == 0) {
        a.GI = a.HI = 0;
        a.ntop = a.H;
        a.nbot = 0;
        a.lside = a.rside = 0;
    } else {
        a.ntop = a.ay;
        a.nbot = a.H - a.HI - a.ay;
        a.lside = 4 * a.glo;
        a.rside = a.L - 4 * (ghi + 1);
    }
    const long long nint = static_cast<long long>(a.GI) * ((a.HI + kStencilRows - 1) / kStencilRows);
    const long long nb = (nint + kBlock - 1) / kBlock;
    a.nframe = static_cast<long long>(a.ntop + a.nbot) * a.L + static_cast<long long>(a.HI) * (a.lside + a.rside);
    const long long nbf = (a.nframe + kBlock - 1) / kBlock;
    if (nint > (1LL << 31) - 1 || nb + nbf + 8 > (1LL << 31) - 1) return hipErrorNotSupported;
    a.nblocks = static_cast<int>(nb);
    a.nbi = static_cast<int>((nb + 7) / 8 * 8);
    grid = dim3(static_cast<unsigned>(a.nbi + nbf));
    return hipSuccess;
}

template <class To, int KW, int KH, bool SEP>
hipError_t stencil_by_cn(StencilArgs a, int cn, hipStream_t s) {
    dim3 grid;
    const hipError_t e = stencil_geometry(a, KW, KH, cn, grid);
    if (e != hipSuccess) return e;
    switch (cn) {
        case 1: hipLaunchKernelGGL((stencil_u8_kernel<To, KW, KH, 1, SEP>), grid, dim3(kBlock), 0, s, a); break;
        case 3: hipLaunchKernelGGL((stencil_u8_kernel<To, KW, KH, 3, SEP>), grid, dim3(kBlock), 0, s, a); break;
        case 4: hipLaunchKernelGGL((stencil_u8_kernel<To, KW, KH, 4, SEP>), grid, dim3(kBlock), 0, s, a); break;
        default: return hipErrorNotSupported;
    }
    return hipGetLastError();
}

// ------------------------------------------------------------ 3 x 3 median
// medianBlur(k = 3) on uint8 images (filter.hpp:302-332) in the stencil's
// layout. The median of 9 is med3(max of the three row minima, med3 of the row
// medians, min of the three row maxima) over the window's rows, each row's
// triple sorted once and shared by the three output rows it reaches.
__device__ __forceinline__ void sort3u(unsigned& x, unsigned& y, unsigned& z) {
    const unsigned a = min(x, y), b = max(x, y);
    const unsigned c = min(b, z);
    z = max(b, z);
    x = min(a, c);
    y = max(a, c);
}
__device__ __forceinline__ unsigned med3u(unsigned x, unsigned y, unsigned z) {
    return max(min(x, y), min(max(x, y), z));
}
template <int N>
__device__ __forceinline__ unsigned win_u8(const unsigned (&w)[N], int q) {
    return (w[q >> 2] >> (8 * (q & 3))) & 0xffu;
}

template <int CN>
__device__ __forceinline__ void median3_frame_sample(const StencilArgs& a, long long idx) {
    int y, s;
    frame_pos(a, idx, y, s);
    const int px = s / CN, c = s - px * CN;
    unsigned lo[3], md[3], hi[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const int sy = border_index(y - 1 + j, a.H, a.border);  // never Constant: sy >= 0
        const uint8_t* srow = a.src + static_cast<size_t>(sy) * a.sstride;
        unsigned v[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) v[i] = srow[static_cast<size_t>(border_index(px + i - 1, a.W, a.border)) * CN + c];
        sort3u(v[0], v[1], v[2]);
        lo[j] = v[0];
        md[j] = v[1];
        hi[j] = v[2];
    }
    const unsigned m = med3u(max(max(lo[0], lo[1]), lo[2]), med3u(md[0], md[1], md[2]), min(min(hi[0], hi[1]), hi[2]));
    reinterpret_cast<uint8_t*>(a.dst + static_cast<size_t>(y) * a.dstride)[s] = static_cast<uint8_t>(m);
}

template <int CN>
__global__ __launch_bounds__(kBlock) void median3_u8_kernel(StencilArgs a) {
    constexpr int DH = (CN + 3) / 4;  // halo dwords on each side
    constexpr int ND = 1 + 2 * DH;
    constexpr int NG = kStencilRows;
    constexpr int NR = NG + 2;
    const int b = blockIdx.x;
    if (b >= a.nbi) {
        const long long idx = static_cast<long long>(b - a.nbi) * kBlock + threadIdx.x;
        if (idx < a.nframe) median3_frame_sample<CN>(a, idx);
        return;
    }
    const int blk = (b & 7) * (a.nbi >> 3) + (b >> 3);
    if (blk >= a.nblocks) return;
    const unsigned nrb = (static_cast<unsigned>(a.HI) + NG - 1) / NG;
    const unsigned idx = static_cast<unsigned>(blk) * kBlock + threadIdx.x;
    if (idx >= static_cast<unsigned>(a.GI) * nrb) return;
    const unsigned rb = idx / static_cast<unsigned>(a.GI);
    const int s0 = (static_cast<int>(idx - rb * static_cast<unsigned>(a.GI)) + a.glo) * 4;
    const int r0 = static_cast<int>(rb) * NG;
    unsigned w[NR][ND];
#pragma unroll
    for (int t = 0; t < NR; ++t) {
        const int rr = min(r0 + t, a.H - 1);  // rows past the image feed only rows that are not stored
        const unsigned* p = reinterpret_cast<const unsigned*>(a.src + static_cast<size_t>(rr) * a.sstride + s0) - DH;
#pragma unroll
        for (int d = 0; d < ND; ++d) w[t][d] = p[d];
    }
    // sorted horizontal triples of the last three input rows (a ring)
    unsigned lo[3][4], md[3][4], hi[3][4];
    auto sort_row = [&](int t, int slot) {
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            unsigned x = win_u8(w[t], 4 * DH + m - CN), y = win_u8(w[t], 4 * DH + m), z = win_u8(w[t], 4 * DH + m + CN);
            sort3u(x, y, z);
            lo[slot][m] = x;
            md[slot][m] = y;
            hi[slot][m] = z;
        }
    };
    sort_row(0, 0);
    sort_row(1, 1);
#pragma unroll
    for (int k = 0; k < NG; ++k) {
        sort_row(k + 2, (k + 2) % 3);
        if (r0 + k >= a.HI) break;
        unsigned out = 0u;
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const unsigned p = max(max(lo[0][m], lo[1][m]), lo[2][m]);
            const unsigned q = med3u(md[0][m], md[1][m], md[2][m]);
            const unsigned r = min(min(hi[0][m], hi[1][m]), hi[2][m]);
            out |= med3u(p, q, r) << (8 * m);
        }
        *reinterpret_cast<unsigned*>(a.dst + static_cast<size_t>(a.ay + r0 + k) * a.dstride + s0) = out;
    }
}

bool aligned16(const void* p, size_t stride) {
    return (reinterpret_cast<uintptr_t>(p) & 15u) == 0 && (stride & 15u) == 0;
}

}  // namespace

hipError_t launch_median3_u8(const DevImage& src, const DevImage& dst, int border, hipStream_t s) {
    const int cn = src.cn;
    if (src.dt != IPFG_U8 || dst.dt != IPFG_U8 || border == IPFG_BORDER_CONSTANT) return hipErrorNotSupported;
    if (cn != 1 && cn != 3 && cn != 4) return hipErrorNotSupported;
    if (!aligned16(src.data, src.stride) || !aligned16(dst.data, dst.stride)) return hipErrorNotSupported;
    const long long L = static_cast<long long>(src.w) * cn;
    if (L <= 0 || src.h <= 0 || L > (1LL << 30)) return hipErrorNotSupported;
    StencilArgs a{};
    a.src = static_cast<const uint8_t*>(src.data);
    a.dst = static_cast<char*>(dst.data);
    a.sstride = src.stride;
    a.dstride = dst.stride;
    a.W = src.w;
    a.H = src.h;
    a.L = static_cast<int>(L);
    a.ay = 1;
    a.border = border;
    dim3 grid;
    const hipError_t e = stencil_geometry(a, 3, 3, cn, grid);
    if (e != hipSuccess) return e;
    switch (cn) {
        case 1: hipLaunchKernelGGL(median3_u8_kernel<1>, grid, dim3(kBlock), 0, s, a); break;
        case 3: hipLaunchKernelGGL(median3_u8_kernel<3>, grid, dim3(kBlock), 0, s, a); break;
        default: hipLaunchKernelGGL(median3_u8_kernel<4>, grid, dim3(kBlock), 0, s, a); break;
    }
    return hipGetLastError();
}

hipError_t launch_stencil_u8(const DevImage& src, const DevImage& dst, const float* k, int kw, int kh, int ay,
                             int border, double delta, double bval, hipStream_t s, const float* kx, const float* ky) {
    const int cn = src.cn;
    if (src.dt != IPFG_U8 || (dst.dt != IPFG_U8 && dst.dt != IPFG_F32)) return hipErrorNotSupported;
    if (!((kw == 3 && kh == 3) || (kw == 5 && kh == 5))) return hipErrorNotSupported;
    if (cn != 1 && cn != 3 && cn != 4) return hipErrorNotSupported;
    if (!aligned16(src.data, src.stride) || !aligned16(dst.data, dst.stride)) return hipErrorNotSupported;
    if (border == IPFG_BORDER_CONSTANT && !(bval == static_cast<double>(static_cast<float>(bval))))
        return hipErrorNotSupported;
    const long long L = static_cast<long long>(src.w) * cn;
    if (L <= 0 || src.h <= 0 || L > (1LL << 30)) return hipErrorNotSupported;
    StencilArgs a{};
    a.src = static_cast<const uint8_t*>(src.data);
    a.dst = static_cast<char*>(dst.data);
    a.sstride = src.stride;
    a.dstride = dst.stride;
    a.W = src.w;
    a.H = src.h;
    a.L = static_cast<int>(L);
    a.ay = ay;
    a.border = border;
    a.bval = static_cast<float>(bval);
    a.delta = static_cast<float>(delta);
    for (int i = 0; i < kw * kh; ++i) a.k[i] = k[i];
    const bool sep = kx != nullptr && ky != nullptr;
    if (sep) {
        for (int i = 0; i < kw; ++i) a.kx[i] = kx[i];
        for (int j = 0; j < kh; ++j) a.ky[j] = ky[j];
    }
    const bool f32 = dst.dt == IPFG_F32;
    if (kw == 3) {
        if (sep) return f32 ? stencil_by_cn<float, 3, 3, true>(a, cn, s) : stencil_by_cn<uint8_t, 3, 3, true>(a, cn, s);
        return f32 ? stencil_by_cn<float, 3, 3, false>(a, cn, s) : stencil_by_cn<uint8_t, 3, 3, false>(a, cn, s);
    }
    if (sep) return f32 ? stencil_by_cn<float, 5, 5, true>(a, cn, s) : stencil_by_cn<uint8_t, 5, 5, true>(a, cn, s);
    return f32 ? stencil_by_cn<float, 5, 5, false>(a, cn, s) : stencil_by_cn<uint8_t, 5, 5, false>(a, cn, s);
}

}  // namespace ipfg
