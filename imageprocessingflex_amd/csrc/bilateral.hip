// bilateralFilter for gfx950.
//
// Reference: include/ipf/filter.hpp:334-387. Circular taps i^2 + j^2 <= r^2
// (j outer, i inner) with FP64 space weights; the range weight of the L1
// colour distance comes from an FP64 table for integer images; Constant
// border skips taps outside the image; out = saturate_cast(acc / wsum).
//
// Fast path (uint8, 1..4 channels): pixels are packed into one 32-bit word in
// an LDS tile with an r-pixel halo, so the L1 distance over all channels is a
// single v_sad_u8; the range table (<= 1021 entries) and the tap list (LDS
// offset, dx, dy, space weight) sit in LDS, so every lane of a wave walks the
// same taps with broadcast reads. One thread per output pixel. Certified
// launches check every channel's FP64 result against its error bound;
// doubtful pixels are recomputed in FP64 in the reference's order.
//
// Exact path: the reference loop in FP64 per pixel, for every pixel type.
#include "batch.hpp"
#include "kernels.hpp"

namespace ipfg {
namespace {

constexpr int TX = 64;

__host__ __device__ inline int bil_xp(int r) { return (TX + 2 * r) | 1; }

// Tap list entry in LDS: tile offset dy * XP + dx, and dx, dy for the
// Constant border's bounds test.
struct BilTap {
    int off, dx, dy, pad;
};

// Largest tap count of a radius-r disc is (2r + 1)^2.
__host__ __device__ inline int bil_max_taps(int r) { return (2 * r + 1) * (2 * r + 1); }

__host__ inline size_t bil_lds_bytes(int TH, int r, int lut_len, size_t acc_size) {
    const size_t tile = sizeof(uint32_t) * ((static_cast<size_t>(TH + 2 * r) * bil_xp(r) + 3) / 4 * 4);
    const size_t lut = (acc_size * lut_len + 15) / 16 * 16;
    return tile + lut + (sizeof(BilTap) + acc_size) * bil_max_taps(r) + 16;
}

// One pixel, every channel, in the reference's FP64 order (filter.hpp:366-383:
// for each tap, w = space * range(L1 distance); wsum += w; acc[c] += w * p[c];
// out = acc / wsum; Constant border skips taps), computed by a whole wave:
// each lane forms the weights and products of 64 taps at a time and the sums
// run over broadcast values. All 64 lanes must call it; all return the results.
template <int CN>
__device__ inline void bil_exact_wave(const DevImage& src, int x, int y, const int* __restrict__ taps,
                                      const double* __restrict__ spaced, int ntaps, const double* __restrict__ lutd,
                                      int lut_len, int border, double cc, double (&out)[CN]) {
    const int lane = lane_id();
    const uint8_t* p0 = static_cast<const uint8_t*>(src.data) + static_cast<size_t>(y) * src.stride + static_cast<size_t>(x) * CN;
    double wsum = 0.0, acc[CN];
#pragma unroll
    for (int c = 0; c < CN; ++c) acc[c] = 0.0;
    for (int k0 = 0; k0 < ntaps; k0 += 64) {
        const int k = k0 + lane;
        bool valid = false;
        double w = 0.0, wp[CN];
#pragma unroll
        for (int c = 0; c < CN; ++c) wp[c] = 0.0;
        if (k < ntaps) {
            const int sx = border_index(x + taps[2 * k], src.w, border);
            const int sy = border_index(y + taps[2 * k + 1], src.h, border);
            if (sx >= 0 && sy >= 0) {
                valid = true;
                const uint8_t* p = static_cast<const uint8_t*>(src.data) + static_cast<size_t>(sy) * src.stride +
                                   static_cast<size_t>(sx) * CN;
                double dist = 0.0;
#pragma unroll
                for (int c = 0; c < CN; ++c) dist = dist + fabs(static_cast<double>(p[c]) - static_cast<double>(p0[c]));
                const double rw = dist < static_cast<double>(lut_len) ? lutd[static_cast<size_t>(dist)] : exp(dist * dist * cc);
                w = spaced[k] * rw;
#pragma unroll
                for (int c = 0; c < CN; ++c) wp[c] = w * static_cast<double>(p[c]);
            }
        }
        const unsigned long long vb = __ballot(valid);
        const int n = min(64, ntaps - k0);
        for (int kk = 0; kk < n; ++kk) {
            if (!((vb >> kk) & 1ull)) continue;
            wsum = wsum + bcast_d(w, kk);
#pragma unroll
            for (int c = 0; c < CN; ++c) acc[c] = acc[c] + bcast_d(wp[c], kk);
        }
    }
#pragma unroll
    for (int c = 0; c < CN; ++c) out[c] = acc[c] / wsum;
}

__device__ inline double bil_exact_channel(const DevImage& src, int x, int y, int c, const int* taps,
                                           const double* spaced, int ntaps, const double* lutd, int lut_len, int r,
                                           int border, double cc);

struct BilArgs {
    DevImage src, dst;
    const int* taps;
    const void* space;    // ntaps space weights of type Acc
    const void* lut;      // lut_len range weights of type Acc
    const double* spaced; // FP64 tables for exact recomputation
    const double* lutd;
    int ntaps, lut_len, r, border, TH;
    double cc;
    double bound;         // certification bound on |result - reference|
    FixList fix;
};

// One TX x TH output tile (bx, by); the workgroup's LDS is bsm. Certified
// launches recompute doubtful pixels in place (wave_fix); the caller runs
// fix_finish once per workgroup.
template <int CN, bool CONST_BORDER, class Acc, bool CERT>
__device__ __forceinline__ void bilateral_u8_tile(const BilArgs& a, int bx, int by, uint32_t* bsm) {
    const int r = a.r, TH = a.TH;
    const int W = a.src.w, H = a.src.h;
    const int XP = bil_xp(r);
    const int HR = TH + 2 * r;
    uint32_t* tile = bsm;
    Acc* lut = reinterpret_cast<Acc*>(bsm + ((HR * XP + 3) / 4) * 4);
    BilTap* ltap = reinterpret_cast<BilTap*>(reinterpret_cast<char*>(lut) + (sizeof(Acc) * a.lut_len + 15) / 16 * 16);
    Acc* lsw = reinterpret_cast<Acc*>(ltap + a.ntaps);
    const Acc* __restrict__ lutg = static_cast<const Acc*>(a.lut);
    const Acc* __restrict__ space = static_cast<const Acc*>(a.space);
    const uint8_t* __restrict__ src = static_cast<const uint8_t*>(a.src.data);
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform: scalar loop bounds
    const int x0 = bx * TX, y0 = by * TH;
    for (int i = tid; i < a.lut_len; i += kBlock) lut[i] = lutg[i];
    for (int i = tid; i < a.ntaps; i += kBlock) {
        const int dx = a.taps[2 * i], dy = a.taps[2 * i + 1];
        ltap[i] = BilTap{dy * XP + dx, dx, dy, 0};
        lsw[i] = space[i];
    }
    // Staging: each wave takes rows wave, wave + 4, ...; per row every lane
    // loads all its pixels (px = lane + 64 j, clamped to the row) before it
    // packs and writes them, so a row costs one memory round trip.
    const int XW = TX + 2 * r;
    constexpr int PJ = 2;  // pixels per lane per pass: rows up to 128 pixels take one pass
    for (int rr = wave; rr < HR; rr += kBlock / 64) {
        const int gy = y0 - r + rr;
        const int sy = (gy >= 0 && gy < H) ? gy : border_index(gy, H, a.border);
        uint32_t* trow = tile + rr * XP;
        const uint8_t* srow = src + static_cast<size_t>(sy < 0 ? 0 : sy) * a.src.stride;
        for (int p0x = 0; p0x < XW; p0x += 64 * PJ) {
            uint8_t b[PJ][CN];
            bool out[PJ];
#pragma unroll
            for (int j = 0; j < PJ; ++j) {
                const int px = min(p0x + lane + 64 * j, XW - 1);
                const int gx = x0 - r + px;
                const int sx = (gx >= 0 && gx < W) ? gx : border_index(gx, W, a.border);
                out[j] = sy < 0 || sx < 0;
                const uint8_t* p = srow + static_cast<size_t>(sx < 0 ? 0 : sx) * CN;
#pragma unroll
                for (int c = 0; c < CN; ++c) b[j][c] = p[c];
            }
#pragma unroll
            for (int j = 0; j < PJ; ++j) {
                const int px = p0x + lane + 64 * j;
                if (px < XW) {
                    uint32_t v = 0;
#pragma unroll
                    for (int c = 0; c < CN; ++c) v |= static_cast<uint32_t>(b[j][c]) << (8 * c);
                    trow[px] = out[j] ? 0u : v;
                }
            }
        }
    }
    __syncthreads();
    // wave-uniform loops: every lane walks the same taps (lanes past the
    // image filter a position inside the tile and store nothing)
    const int ntaps = a.ntaps;
    for (int t0 = wave * 64; t0 < TX * TH; t0 += kBlock) {
        const int t = t0 + lane;
        const int lx = t % TX, ly = t / TX;
        const int x = x0 + lx, y = y0 + ly;
        const bool act = x < W && y < H;
        const uint32_t* cen = tile + (ly + r) * XP + (lx + r);
        const uint32_t p0 = *cen;
        Acc wsum = Acc(0);
        Acc acc[CN];
#pragma unroll
        for (int c = 0; c < CN; ++c) acc[c] = Acc(0);
        for (int k = 0; k < ntaps; ++k) {
            const BilTap tp = ltap[k];
            const uint32_t p = cen[tp.off];
            const uint32_t dist = __builtin_amdgcn_sad_u8(p, p0, 0u);
            Acc w = lsw[k] * lut[dist];
            if (CONST_BORDER) {  // a tap outside the image takes no part (weight 0 adds exactly nothing)
                if (static_cast<unsigned>(x + tp.dx) >= static_cast<unsigned>(W) ||
                    static_cast<unsigned>(y + tp.dy) >= static_cast<unsigned>(H))
                    w = Acc(0);
            }
            wsum += w;
#pragma unroll
            for (int c = 0; c < CN; ++c) {
                const Acc pc = static_cast<Acc>((p >> (8 * c)) & 255u);
                if constexpr (sizeof(Acc) == 8) acc[c] = fma(w, pc, acc[c]);
                else acc[c] = fmaf(w, pc, acc[c]);
            }
        }
        unsigned int dmask = 0u;
        if (act) {
            uint8_t* o = static_cast<uint8_t*>(a.dst.data) + static_cast<size_t>(y) * a.dst.stride + static_cast<size_t>(x) * CN;
            bool doubt = false;
#pragma unroll
            for (int c = 0; c < CN; ++c) {
                const double v = static_cast<double>(acc[c] / wsum);
                o[c] = saturate_from_double<uint8_t>(v);
                if (CERT) doubt |= needs_exact(v, a.bound, 0.0, 255.0);
            }
            dmask = doubt ? 1u : 0u;
        }
        if (CERT) {
            wave_fix(a.fix, dmask, [&](int l, int) {
                const int xl = bcast_i(x, l), yl = bcast_i(y, l);
                double e[CN];
                bil_exact_wave<CN>(a.src, xl, yl, a.taps, a.spaced, a.ntaps, a.lutd, a.lut_len, a.border, a.cc, e);
                if (lane_id() == l) {
                    uint8_t* o = static_cast<uint8_t*>(a.dst.data) + static_cast<size_t>(yl) * a.dst.stride +
                                 static_cast<size_t>(xl) * CN;
#pragma unroll
                    for (int c = 0; c < CN; ++c) o[c] = saturate_from_double<uint8_t>(e[c]);
                }
            });
        }
    }
}

template <int CN, bool CONST_BORDER, class Acc, bool CERT>
__global__ __launch_bounds__(kBlock) void bilateral_u8_kernel(BilArgs a) {
    extern __shared__ __align__(16) uint32_t bsm[];
    bilateral_u8_tile<CN, CONST_BORDER, Acc, CERT>(a, blockIdx.x, blockIdx.y, bsm);
    if (CERT) fix_finish(a.fix);
}

// One pixel, one channel, in the reference's FP64 order. The range weight of
// an integer image comes from the table while the distance is inside it.
__device__ inline double bil_exact_channel(const DevImage& src, int x, int y, int c, const int* taps,
                                           const double* spaced, int ntaps, const double* lutd, int lut_len, int r,
                                           int border, double cc) {
    const size_t es = dtype_size(src.dt);
    const int cn = src.cn;
    const char* p0 = static_cast<const char*>(src.data) + static_cast<size_t>(y) * src.stride + static_cast<size_t>(x) * cn * es;
    double wsum = 0.0, acc = 0.0;
    for (int k = 0; k < ntaps; ++k) {
        const int sx = border_index(x + taps[2 * k], src.w, border);
        const int sy = border_index(y + taps[2 * k + 1], src.h, border);
        if (sx < 0 || sy < 0) continue;
        const char* p = static_cast<const char*>(src.data) + static_cast<size_t>(sy) * src.stride + static_cast<size_t>(sx) * cn * es;
        double dist = 0.0;
        for (int ch = 0; ch < cn; ++ch)
            dist = dist + fabs(load_sample(p, ch * es, src.dt) - load_sample(p0, ch * es, src.dt));
        double rw;
        if (lutd != nullptr && dist < static_cast<double>(lut_len))
            rw = lutd[static_cast<size_t>(dist)];
        else
            rw = exp(dist * dist * cc);
        const double w = spaced[k] * rw;
        wsum = wsum + w;
        acc = acc + w * load_sample(p, c * es, src.dt);
    }
    (void)r;
    return acc / wsum;
}

__global__ __launch_bounds__(kBlock) void bilateral_exact_kernel(DevImage src, DevImage dst,
                                                                 const int* __restrict__ taps,
                                                                 const double* __restrict__ spaced, int ntaps,
                                                                 const double* __restrict__ lutd, int lut_len, int r,
                                                                 int border, double cc) {
    const size_t es = dtype_size(dst.dt);
    const unsigned long long N = static_cast<unsigned long long>(dst.w) * dst.h;
    for (unsigned long long pix = blockIdx.x * static_cast<unsigned long long>(blockDim.x) + threadIdx.x; pix < N;
         pix += static_cast<unsigned long long>(gridDim.x) * blockDim.x) {
        const int y = static_cast<int>(pix / dst.w);
        const int x = static_cast<int>(pix - static_cast<unsigned long long>(y) * dst.w);
        for (int c = 0; c < dst.cn; ++c) {
            const double v = bil_exact_channel(src, x, y, c, taps, spaced, ntaps, lutd, lut_len, r, border, cc);
            store_sample(dst.data, static_cast<size_t>(y) * dst.stride + (static_cast<size_t>(x) * dst.cn + c) * es,
                         dst.dt, v);
        }
    }
}

int grid_for(size_t n) {
    const size_t b = (n + kBlock - 1) / kBlock;
    return static_cast<int>(b < 16384 ? (b ? b : 1) : 16384);
}

template <int CN, bool CB, class Acc, bool CERT>
hipError_t bil_fast_launch(const BilateralLaunch& l, hipStream_t s) {
    int TH = 32;
    while (TH > 4 && bil_lds_bytes(TH, l.r, l.lut_len, sizeof(Acc)) > 64 * 1024) TH /= 2;
    const size_t lds = bil_lds_bytes(TH, l.r, l.lut_len, sizeof(Acc));
    BilArgs a{};
    a.src = l.src;
    a.dst = l.dst;
    a.taps = l.taps;
    a.space = sizeof(Acc) == 8 ? static_cast<const void*>(l.spaced) : static_cast<const void*>(l.spacef);
    a.lut = sizeof(Acc) == 8 ? static_cast<const void*>(l.lutd) : static_cast<const void*>(l.lutf);
    a.spaced = l.spaced;
    a.lutd = l.lutd;
    a.ntaps = l.ntaps;
    a.lut_len = l.lut_len;
    a.r = l.r;
    a.border = l.border;
    a.TH = TH;
    a.cc = l.cc;
    a.bound = l.bound;
    a.fix = l.fix;
    const dim3 grid(static_cast<unsigned>((l.src.w + TX - 1) / TX), static_cast<unsigned>((l.src.h + TH - 1) / TH));
    hipLaunchKernelGGL((bilateral_u8_kernel<CN, CB, Acc, CERT>), grid, dim3(kBlock), lds, s, a);
    return hipGetLastError();
}

template <int CN, bool CB>
hipError_t bil_fast_mode(const BilateralLaunch& a, hipStream_t s) {
    if (a.mode == MODE_CERTIFY) return bil_fast_launch<CN, CB, double, true>(a, s);
    return bil_fast_launch<CN, CB, float, false>(a, s);
}

template <bool CB>
hipError_t bil_fast_cn(const BilateralLaunch& a, hipStream_t s) {
    switch (a.src.cn) {
        case 1: return bil_fast_mode<1, CB>(a, s);
        case 2: return bil_fast_mode<2, CB>(a, s);
        case 3: return bil_fast_mode<3, CB>(a, s);
        case 4: return bil_fast_mode<4, CB>(a, s);
    }
    return hipErrorInvalidValue;
}

}  // namespace

// True when the uint8 fast kernel can run this call (tile fits LDS).
bool bilateral_fast_ok(int dt, int cn, int r, int lut_len) {
    return dt == IPFG_U8 && cn >= 1 && cn <= 4 && lut_len > 0 && bil_lds_bytes(4, r, lut_len, 8) <= 64 * 1024;
}

hipError_t launch_bilateral(const BilateralLaunch& a, hipStream_t s) {
    const size_t npix = static_cast<size_t>(a.src.w) * a.src.h;
    if (a.mode == MODE_EXACT) {
        hipLaunchKernelGGL(bilateral_exact_kernel, dim3(grid_for(npix)), dim3(kBlock), 0, s, a.src, a.dst, a.taps,
                           a.spaced, a.ntaps, a.lutd, a.lut_len, a.r, a.border, a.cc);
        return hipGetLastError();
    }
    return a.border == IPFG_BORDER_CONSTANT ? bil_fast_cn<true>(a, s) : bil_fast_cn<false>(a, s);
}

// ------------------------------------------------------------------ batched
namespace {
template <int CN>
__device__ __forceinline__ void bil_batch_tile(const BilArgs& a, int tx, int ty, uint32_t* bsm) {
    if (a.border == IPFG_BORDER_CONSTANT) bilateral_u8_tile<CN, true, double, true>(a, tx, ty, bsm);
    else bilateral_u8_tile<CN, false, double, true>(a, tx, ty, bsm);
}

__global__ __launch_bounds__(kBlock) void bilateral_batch_kernel(const BatchClassHeader* __restrict__ hdr,
                                                                 const BatchJob* __restrict__ jobs, BatchQueue* q,
                                                                 FixList fix) {
    extern __shared__ __align__(16) uint32_t bsm[];
    batch_loop(hdr, jobs, q, [&](const BatchJob& J, int tx, int ty) {
        BilArgs a{};
        a.src = J.src;
        a.dst = J.dst;
        a.taps = J.taps;
        a.space = J.spaced;
        a.lut = J.lutd;
        a.spaced = J.spaced;
        a.lutd = J.lutd;
        a.ntaps = J.ntaps;
        a.lut_len = J.lut_len;
        a.r = J.r;
        a.border = J.border;
        a.TH = J.rows;
        a.cc = J.cc;
        a.bound = J.bound;
        a.fix = fix;
        switch (J.src.cn) {
            case 1: bil_batch_tile<1>(a, tx, ty, bsm); break;
            case 3: bil_batch_tile<3>(a, tx, ty, bsm); break;
            default: bil_batch_tile<4>(a, tx, ty, bsm); break;
        }
    });
    fix_finish(fix);
}
}  // namespace

hipError_t launch_bilateral_batch(const BatchClassHeader* hdr, const BatchJob* jobs, BatchQueue* q, FixList fix,
                                  int grid, size_t lds, hipStream_t s) {
    hipLaunchKernelGGL(bilateral_batch_kernel, dim3(grid), dim3(kBlock), lds, s, hdr, jobs, q, fix);
    return hipGetLastError();
}

int bilateral_batch_occupancy(size_t lds) {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, reinterpret_cast<const void*>(bilateral_batch_kernel), kBlock,
                                                     lds) != hipSuccess)
        return 1;
    return n > 0 ? n : 1;
}

// Largest radius and range table the batched bilateral filter takes.
size_t bilateral_batch_lds() { return bil_lds_bytes(kBatchTileRows, 5, 255 * 4 + 1, sizeof(double)); }

}  // namespace ipfg
