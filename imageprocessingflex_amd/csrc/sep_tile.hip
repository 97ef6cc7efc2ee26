// Certified uint8 -> uint8 separable filter in one pass per tile: gaussianBlur /
// sepFilter2D with kw = kh = K odd in [3, 15] and a centred anchor, on 1, 3 or
// 4 interleaved channels. Reference: include/ipf/filter.hpp:34-90 (rows
// filtered with kx into an FP64 buffer, then columns with ky from delta) and
// 261-267 (gaussianBlur).
//
// The strip kernels walk a workgroup down a column of the image in pipelined
// steps; on an image of a few megapixels the strips are short (a 1080-row
// image gives 16-row strips two steps long) and every workgroup spends most
// of its life waiting for its first rows. Here the image is cut into tiles of
// TXP pixels x TY rows (TY = 16 or 32), and one round of workgroups walks
// them, each tile in three phases:
//   stage      the (TY + K - 1) rows x (TXP + K - 1) pixels the tile reads go
//              to LDS as float, interleaved as in the image (interior tiles:
//              whole-wave rows of aligned dwords, loaded while the previous
//              tile's vertical pass runs; tiles at the left or right edge: one
//              sample per item through border_index). Rows come through
//              border_index too; Constant rows and columns take the border value.
//   horizontal task (row, group of Q pixels): all channels of Q consecutive
//              pixels from 16-byte LDS reads, K FP64 FMAs per sample, the sums
//              rounded to float into a second LDS tile.
//   vertical   task (sample column, group of M rows): M + K - 1 sums, K FP64
//              FMAs per output from delta, saturate_cast to uint8.
// Each output is certified as in the strip kernels: bound = bcoef * X +
// bdelta with X = max |input| (255, or |border value|), bcoef widened by the
// caller for the float-rounded sums; a doubtful output is recomputed by its
// wave in the reference's exact order (sep_exact_wave).
#include "sep_kernels.hpp"

namespace ipfg {
namespace sepk {
namespace {

constexpr int kMaxK = 15;
constexpr int kQ = 4;   // pixels per horizontal task
constexpr int kM = 8;   // rows per vertical task

template <int CN, int K, int TY>
struct TileGeo {
    static constexpr int A = K / 2;
    static constexpr int TXP = CN == 1 ? 64 : CN == 3 ? 32 : 16;  // output pixels per tile row
    static constexpr int TXS = TXP * CN;                             // output samples per tile row
    static constexpr int NR = TY + K - 1;                            // staged rows
    static constexpr int G = TXP / kQ;                               // horizontal tasks per row
    static constexpr int NH = NR * G;                                // horizontal tasks
    static constexpr int NV = TXS * (TY / kM);                       // vertical tasks
    static constexpr int XS = (TXP + K - 1) * CN;                    // staged samples per row
    // The first staged byte of a row, (x0 - A) CN with x0 a multiple of 4,
    // lies O bytes past a 4-byte boundary; staged sample 0 sits at LDS float
    // OFF (a multiple of 4, so every task's reads are 16-byte aligned) and the
    // aligned dword d of the row at float PADL + 4 d.
    static constexpr int O = (4 - (A * CN) % 4) % 4;
    static constexpr int OFF = O ? 4 : 0;
    static constexpr int PADL = OFF - O;
    static constexpr int ND = (O + XS + 3) / 4;                      // dwords per staged row
    static constexpr int NSPAN4 = ((kQ + K - 1) * CN + 3) / 4;       // float4 reads per horizontal task
    static constexpr int SPMIN1 = PADL + 4 * ND;
    static constexpr int SPMIN2 = OFF + (TXP - kQ) * CN + 4 * NSPAN4;
    static constexpr int SP = ((SPMIN1 > SPMIN2 ? SPMIN1 : SPMIN2) + 3) / 4 * 4;  // staged row pitch (floats)
    static constexpr int SPS = TXS + 4;                              // sum row pitch (floats)
    static constexpr size_t lds_bytes = sizeof(float) * NR * (SP + SPS);
    static_assert(TXP % kQ == 0 && (kQ * CN) % 4 == 0, "a task's sums are whole float4s");
    static constexpr int NRW = (NR + 3) / 4;                         // staged rows per wave
    static_assert(TY % kM == 0 && NV % 64 == 0, "vertical tasks fill whole waves (the fix-up is wave-collective)");
    static_assert(ND <= 64, "a staged row is one dword load per lane");
    static_assert(lds_bytes <= 64 * 1024, "fits the default LDS limit");
};

__device__ __forceinline__ float ubyte_f(unsigned int v, int q) {
    return static_cast<float>((v >> (8 * q)) & 0xffu);
}

// needs_exact(r, bound, 0, 255) as one comparison, so that the compiler keeps
// a task's M accumulation chains together and needs no mask register per
// test. thr = 0.5 - bound with bound < 0.25 (the caller recomputes every
// output otherwise). r is doubtful when r + 0.5 lies within bound of an
// integer (|fract(r + 0.5) - 0.5| >= thr) unless r < 0 or r > 255: a jump
// within bound of such an r lies at or below -0.5 or at or above 255.5, so
// both ends of [r - bound, r + bound] saturate alike. Each difference below
// has the exact sign of the comparison it stands for; a NaN r gives true.
__device__ __forceinline__ bool doubt_u8(double r, double thr) {
    const double t = r + 0.5;
    const double fr = t - floor(t);
    return !(fmax(thr - fabs(fr - 0.5), fmax(-r, r - 255.0)) > 0.0);
}

// saturate_cast<uint8_t>(double) without branches: v + 0.5 truncated and
// clamped to [0, 255]. Below zero the reference's v - 0.5 and this v + 0.5
// both end at 0; a NaN sum is dropped by the maximum (IEEE maxNum) and gives 0.
__device__ __forceinline__ unsigned char sat_u8(double v) {
    return static_cast<unsigned char>(static_cast<unsigned int>(__builtin_fmin(__builtin_fmax(v + 0.5, 0.0), 255.0)));
}

template <int CN, int K, int TY>
__global__ __launch_bounds__(kBlock) void tile_u8_kernel(SepArgs a, const double* __restrict__ kx,
                                                         const double* __restrict__ ky, int nbx, int ntiles) {
    using T = TileGeo<CN, K, TY>;
    extern __shared__ __align__(16) unsigned char tile_smem[];
    // The weights go through LDS and reach the FMAs as broadcast VGPR
    // operands: held in scalar registers across both passes they spill.
    __shared__ double wts[2 * kMaxK];
    float* __restrict__ stage = reinterpret_cast<float*>(tile_smem);
    float* __restrict__ hs = stage + T::NR * T::SP;
    const int tid = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    const int W = a.src.w, H = a.src.h;
    const unsigned char* __restrict__ src = static_cast<const unsigned char*>(a.src.data);
    const size_t sstride = a.src.stride;
    const float bvalf = static_cast<float>(a.bval);
    const double bound = a.bcoef * a.xbound + a.bdelta;
    const double thr = 0.5 - bound;
    const bool wide = !(bound < 0.25);  // also a NaN bound: recompute every output
    unsigned char* __restrict__ dst = static_cast<unsigned char*>(a.dst.data);
    const size_t dstride = a.dst.stride;
    if (tid < 2 * K) wts[tid < K ? tid : kMaxK + tid - K] = tid < K ? kx[tid] : ky[tid - K];

    // Interior tiles (every staged column inside the image) are staged from
    // whole-wave rows: wave w loads rows w, w + 4, ... of the tile (the row's
    // address is wave-uniform), lane d dword d of the row (lanes past the row
    // re-read its last dword). The loads of the next tile are issued before
    // this tile's vertical pass and land while it runs.
    const int d = min(lane, T::ND - 1);
    unsigned int v[T::NRW];
    auto tile_x0 = [&](int t) { return (t % nbx) * T::TXP; };
    auto tile_y0 = [&](int t) { return (t / nbx) * TY; };
    auto aligned0 = [&](int t) { return (tile_x0(t) - T::A) * CN - T::O; };  // 4-byte aligned
    auto interior = [&](int t) { return aligned0(t) >= 0 && aligned0(t) + 4 * T::ND <= W * CN; };
    auto load = [&](int t) {
        const int y0 = tile_y0(t), a0 = aligned0(t);
#pragma unroll
        for (int u = 0; u < T::NRW; ++u) {
            const int r = min(wave + 4 * u, T::NR - 1);
            const int sy = border_index(y0 - T::A + r, H, a.border);
            v[u] = *reinterpret_cast<const unsigned int*>(src + static_cast<size_t>(max(sy, 0)) * sstride + a0 + 4 * d);
        }
    };

    int tile = blockIdx.x;
    if (tile < ntiles && interior(tile)) load(tile);
    for (; tile < ntiles; tile += gridDim.x) {
        const int x0 = tile_x0(tile), y0 = tile_y0(tile);
        // ---- stage
        if (interior(tile)) {
#pragma unroll
            for (int u = 0; u < T::NRW; ++u) {
                const int r = wave + 4 * u;
                if (r >= T::NR) break;
                float4 f = make_float4(ubyte_f(v[u], 0), ubyte_f(v[u], 1), ubyte_f(v[u], 2), ubyte_f(v[u], 3));
                if (border_index(y0 - T::A + r, H, a.border) < 0) f = make_float4(bvalf, bvalf, bvalf, bvalf);
                float* p = stage + r * T::SP + T::PADL + 4 * d;
                if (lane < T::ND) {
                    if constexpr (T::PADL == 0) {
                        *reinterpret_cast<float4*>(p) = f;
                    } else if constexpr (T::PADL == 2) {
                        reinterpret_cast<float2*>(p)[0] = make_float2(f.x, f.y);
                        reinterpret_cast<float2*>(p)[1] = make_float2(f.z, f.w);
                    } else {
                        p[0] = f.x;
                        p[1] = f.y;
                        p[2] = f.z;
                        p[3] = f.w;
                    }
                }
            }
        } else {  // tiles at the left or right edge: one sample per item
            for (int t = tid; t < T::NR * T::XS; t += kBlock) {
                const int r = t / T::XS, q = t - r * T::XS;
                const int px = q / CN, c = q - px * CN;
                const int sy = border_index(y0 - T::A + r, H, a.border);
                const int sx = border_index(x0 - T::A + px, W, a.border);
                float f = bvalf;
                if (sy >= 0 && sx >= 0)
                    f = static_cast<float>(src[static_cast<size_t>(sy) * sstride + static_cast<size_t>(sx) * CN + c]);
                stage[r * T::SP + T::OFF + q] = f;
            }
        }
        __syncthreads();

        // ---- horizontal: row r, pixels [Q g, Q g + Q), every channel. The
        // tasks go to the threads in reverse order: the waves left with a
        // partial last round here take the extra round of the vertical pass,
        // so every SIMD carries the same share of a tile.
        for (int t = kBlock - 1 - tid; t < T::NH; t += kBlock) {
            const int r = t / T::G, g = t - r * T::G;
            const float* sp = stage + r * T::SP + T::OFF + g * kQ * CN;
            float f[4 * T::NSPAN4];
#pragma unroll
            for (int u = 0; u < T::NSPAN4; ++u) {
                const float4 q4 = reinterpret_cast<const float4*>(sp)[u];
                f[4 * u] = q4.x;
                f[4 * u + 1] = q4.y;
                f[4 * u + 2] = q4.z;
                f[4 * u + 3] = q4.w;
            }
            float out[kQ * CN];
#pragma unroll
            for (int c = 0; c < CN; ++c) {
                double acc[kQ];
#pragma unroll
                for (int m = 0; m < kQ; ++m) acc[m] = 0.0;
#pragma unroll
                for (int i = 0; i < K; ++i) {
                    const double ki = wts[i];
#pragma unroll
                    for (int m = 0; m < kQ; ++m) acc[m] = fma(ki, static_cast<double>(f[(m + i) * CN + c]), acc[m]);
                }
#pragma unroll
                for (int m = 0; m < kQ; ++m) out[m * CN + c] = static_cast<float>(acc[m]);
            }
            float4* hp = reinterpret_cast<float4*>(hs + r * T::SPS + g * kQ * CN);
#pragma unroll
            for (int u = 0; u < kQ * CN / 4; ++u) hp[u] = make_float4(out[4 * u], out[4 * u + 1], out[4 * u + 2], out[4 * u + 3]);
        }
        __syncthreads();
        // the next tile's rows are in flight during this tile's vertical pass
        if (tile + static_cast<int>(gridDim.x) < ntiles && interior(tile + gridDim.x)) load(tile + gridDim.x);

        // ---- vertical: sample column s, rows [M v, M v + M)
        for (int t = tid; t < T::NV; t += kBlock) {  // NV is a multiple of 64: whole waves iterate together
            const int vg = t / T::TXS, s = t - vg * T::TXS;
            const float* hp = hs + vg * kM * T::SPS + s;
            double win[kM + K - 1];
#pragma unroll
            for (int j = 0; j < kM + K - 1; ++j) win[j] = static_cast<double>(hp[j * T::SPS]);
            double acc[kM];
#pragma unroll
            for (int m = 0; m < kM; ++m) acc[m] = a.delta;
#pragma unroll
            for (int j = 0; j < K; ++j) {
                const double kj = wts[kMaxK + j];
#pragma unroll
                for (int m = 0; m < kM; ++m) acc[m] = fma(kj, win[m + j], acc[m]);
            }
            const int sx = x0 * CN + s;  // sample column in the row
            const int yb = y0 + vg * kM;
            // Results and doubt flags of all M rows first (independent chains),
            // then the stores: whole tasks store unconditionally.
            unsigned char out[kM];
            unsigned int dmask = 0u;
#pragma unroll
            for (int m = 0; m < kM; ++m) {
                const double rv = acc[m];
                out[m] = sat_u8(rv);
                dmask |= doubt_u8(rv, thr) ? 1u << m : 0u;
            }
            if (wide) dmask = (1u << kM) - 1u;
            unsigned char* op = dst + static_cast<size_t>(yb) * dstride + sx;
            if (sx < W * CN && yb + kM <= H) {
#pragma unroll
                for (int m = 0; m < kM; ++m) op[m * dstride] = out[m];
            } else {
                const int nrow = sx < W * CN ? min(kM, H - yb) : 0;
                dmask &= (1u << max(nrow, 0)) - 1u;
#pragma unroll
                for (int m = 0; m < kM; ++m)
                    if (m < nrow) op[m * dstride] = out[m];
            }
            wave_fix(a.fix, dmask, [&](int l, int m) {
                const int sl = bcast_i(sx, l);
                const int yl = bcast_i(yb, l) + m;
                const int xl = sl / CN, cl = sl - xl * CN;
                const double e = sep_exact_wave<uint8_t>(a.src, xl, yl, cl, a.kxd, K, a.kyd, K, T::A, T::A, a.border,
                                                         a.delta, a.bval);
                if (lane_id() == l) dst[static_cast<size_t>(yl) * dstride + sl] = saturate_from_double<uint8_t>(e);
            });
        }
    }
    fix_finish(a.fix);
}

// One round of workgroups (as many as the CUs hold) walks the tiles in
// row-major order, each workgroup every gridDim-th tile.
template <int CN, int K, int TY>
hipError_t tile_launch_ck(const SepArgs& a, hipStream_t s) {
    using T = TileGeo<CN, K, TY>;
    auto kern = tile_u8_kernel<CN, K, TY>;
    static int per_cu = 0;  // workgroups a CU holds (occupancy of this instantiation)
    if (per_cu == 0) {
        int n = 0;
        const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, reinterpret_cast<const void*>(kern),
                                                                          kBlock, T::lds_bytes);
        if (e != hipSuccess) return e;
        per_cu = n > 0 ? n : 1;
    }
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        cus = 256;
    const int nbx = (a.src.w + T::TXP - 1) / T::TXP;
    const long long nt = static_cast<long long>(nbx) * ((a.src.h + TY - 1) / TY);
    if (nt > (1LL << 30)) return hipErrorNotSupported;
    const int ntiles = static_cast<int>(nt);
    const int grid = static_cast<int>(std::min<long long>(nt, static_cast<long long>(cus) * per_cu));
    hipLaunchKernelGGL(kern, dim3(static_cast<unsigned>(grid)), dim3(kBlock), T::lds_bytes, s, a,
                       static_cast<const double*>(a.kx), static_cast<const double*>(a.ky), nbx, ntiles);
    return hipGetLastError();
}

template <int CN, int TY, int K = 3>
hipError_t tile_launch_c(const SepArgs& a, hipStream_t s) {
    if constexpr (K > kMaxK) {
        return hipErrorNotSupported;
    } else {
        if (a.kw == K) return tile_launch_ck<CN, K, TY>(a, s);
        return tile_launch_c<CN, TY, K + 2>(a, s);
    }
}

template <int TY>
hipError_t tile_launch_ty(const SepArgs& a, hipStream_t s) {
    switch (a.src.cn) {
        case 1: return tile_launch_c<1, TY>(a, s);
        case 3: return tile_launch_c<3, TY>(a, s);
        default: return tile_launch_c<4, TY>(a, s);
    }
}

}  // namespace

bool tile_u8_fits(const DevImage& src, const DevImage& dst, int kw, int kh, int ax, int ay, bool unit) {
    const int cn = src.cn;
    return src.dt == IPFG_U8 && dst.dt == IPFG_U8 && (cn == 1 || cn == 3 || cn == 4) && dst.cn == cn &&
           kw == kh && kw % 2 == 1 && kw >= 3 && kw <= kMaxK && ax == kw / 2 && ay == kh / 2 && !unit &&
           src.w > 0 && src.h > 0 && reinterpret_cast<uintptr_t>(src.data) % 4 == 0 &&
           src.stride % 4 == 0 && static_cast<long long>(src.w) * cn < (1LL << 30);
}

hipError_t tile_u8(const SepArgs& a, int rows, hipStream_t s) {
    if (!tile_u8_fits(a.src, a.dst, a.kw, a.kh, a.ax, a.ay, a.unit != 0) || !(a.xbound >= 0.0) || a.oscale != 1.0)
        return hipErrorNotSupported;
    if (rows == 16) return tile_launch_ty<16>(a, s);
    if (rows == 32) return tile_launch_ty<32>(a, s);
    return hipErrorNotSupported;
}

}  // namespace sepk
}  // namespace ipfg
