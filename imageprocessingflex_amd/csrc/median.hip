// medianBlur for gfx950: exact k x k median (rank k*k/2) per channel.
//
// Reference: include/ipf/filter.hpp:269-332 (Huang's sliding histogram for
// uint8 with k >= 5, nth_element otherwise; both give the exact order
// statistic, so any exact selection is bit-identical).
//
// Samples are mapped to order-preserving unsigned keys and staged in LDS (one
// plane per channel, border-mapped halo). Each thread produces P = 4
// horizontally adjacent outputs:
//  - k = 3: the three rows of each window column are sorted once (min/med/max)
//    and shared by the three outputs that use the column; the median is
//    med3(max of column minima, med3 of column medians, min of column maxima).
//  - k = 5: each of the P + 4 columns is sorted once (9 comparators) and
//    shared; the median of the 5 sorted columns comes from a pruned odd-even
//    merge network (82 comparators, generated and verified by
//    gen_median_networks.py).
//  - any other k: bitwise selection -- the largest key t with
//    #(window keys < t) <= rank is built MSB first by counting in LDS.
#include "batch.hpp"
#include "kernels.hpp"

namespace ipfg {
namespace {

template <class K>
__device__ __forceinline__ K kmin(K a, K b) {
    return a < b ? a : b;
}
template <class K>
__device__ __forceinline__ K kmax(K a, K b) {
    return a < b ? b : a;
}

// Two 16-bit keys in one register: the comparators of two independent
// windows run as one v_pk_min_u16 / v_pk_max_u16 each.
typedef unsigned short KeyPair __attribute__((ext_vector_type(2)));
template <>
__device__ __forceinline__ KeyPair kmin<KeyPair>(KeyPair a, KeyPair b) {
    return __builtin_elementwise_min(a, b);
}
template <>
__device__ __forceinline__ KeyPair kmax<KeyPair>(KeyPair a, KeyPair b) {
    return __builtin_elementwise_max(a, b);
}

}  // namespace
}  // namespace ipfg

namespace ipfg {
namespace {
#include "median_networks.hpp"

// Order-preserving key of a sample.
template <class T>
struct KeyOf;
template <> struct KeyOf<uint8_t> {
    using K = uint32_t;
    static constexpr int bits = 8;
    __device__ static K to(uint8_t v) { return v; }
    __device__ static uint8_t from(K k) { return static_cast<uint8_t>(k); }
};
template <> struct KeyOf<uint16_t> {
    using K = uint32_t;
    static constexpr int bits = 16;
    __device__ static K to(uint16_t v) { return v; }
    __device__ static uint16_t from(K k) { return static_cast<uint16_t>(k); }
};
template <> struct KeyOf<int16_t> {
    using K = uint32_t;
    static constexpr int bits = 16;
    __device__ static K to(int16_t v) { return static_cast<K>(static_cast<int>(v) + 32768); }
    __device__ static int16_t from(K k) { return static_cast<int16_t>(static_cast<int>(k) - 32768); }
};
template <> struct KeyOf<int32_t> {
    using K = uint32_t;
    static constexpr int bits = 32;
    __device__ static K to(int32_t v) { return static_cast<K>(v) ^ 0x80000000u; }
    __device__ static int32_t from(K k) { return static_cast<int32_t>(k ^ 0x80000000u); }
};
template <> struct KeyOf<float> {
    using K = uint32_t;
    static constexpr int bits = 32;
    __device__ static K to(float v) {
        const K b = __float_as_uint(v);
        return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
    }
    __device__ static float from(K k) { return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k); }
};
template <> struct KeyOf<double> {
    using K = unsigned long long;
    static constexpr int bits = 64;
    __device__ static K to(double v) {
        const K b = static_cast<K>(__double_as_longlong(v));
        return (b >> 63) ? ~b : (b | (1ull << 63));
    }
    __device__ static double from(K k) {
        const K b = (k >> 63) ? (k & ~(1ull << 63)) : ~k;
        return __longlong_as_double(static_cast<long long>(b));
    }
};

constexpr int P = 4;
constexpr int TX = 64;

__host__ __device__ inline int med_xp(int k) { return (TX + k - 1) | 1; }

template <class T>
__host__ inline size_t med_lds_bytes(int cn, int TH, int k) {
    using K = typename KeyOf<T>::K;
    return sizeof(K) * static_cast<size_t>(cn) * (TH + k - 1) * med_xp(k);
}

template <class K>
__device__ __forceinline__ K med3(K a, K b, K c) {
    return kmax(kmin(a, b), kmin(kmax(a, b), c));
}

// One TX x TH output tile (bx, by) of a median; the workgroup's LDS is smem.
template <class T, int KS>
__device__ __forceinline__ void median_tile(const T* __restrict__ src, size_t sstride, int W, int H, int cn,
                                            T* __restrict__ dst, size_t dstride, int k, int border, int TH, int bx,
                                            int by, unsigned char* smem_raw) {
    using KT = KeyOf<T>;
    using K = typename KT::K;
    K* tile = reinterpret_cast<K*>(smem_raw);
    const int ksz = KS > 0 ? KS : k;
    const int r = ksz / 2;
    const int XP = med_xp(ksz);
    const int HR = TH + ksz - 1;
    const int plane = HR * XP;
    const int tid = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    const int x0 = bx * TX, y0 = by * TH;
    const int XW = TX + ksz - 1;
    const int rowSamples = XW * cn;
    const bool inside_x = (x0 - r >= 0) && (x0 - r + XW <= W);
    // Staging: each wave takes rows wave, wave + 4, ...; per row, every lane
    // issues its loads (samples lane + 64 j, j < SJ, clamped to the row)
    // before it writes any, so a row costs one memory round trip.
    constexpr int SJ = 5;  // (64 + 4) samples x 4 channels fit in one pass
    int px_l[SJ], c_l[SJ];
#pragma unroll
    for (int j = 0; j < SJ; ++j) {
        px_l[j] = (lane + 64 * j) / cn;
        c_l[j] = lane + 64 * j - px_l[j] * cn;
    }
    for (int rr = wave; rr < HR; rr += kBlock / 64) {
        const int gy = y0 - r + rr;
        const int sy = (gy >= 0 && gy < H) ? gy : border_index(gy, H, border);
        const T* srow = reinterpret_cast<const T*>(reinterpret_cast<const char*>(src) + static_cast<size_t>(sy) * sstride);
        K* trow = tile + rr * XP;
        for (int s0 = 0; s0 < rowSamples; s0 += 64 * SJ) {
            T v[SJ];
#pragma unroll
            for (int j = 0; j < SJ; ++j) {
                const int s = min(s0 + lane + 64 * j, rowSamples - 1);
                if (inside_x) {
                    v[j] = srow[static_cast<size_t>(x0 - r) * cn + s];
                } else {
                    const int px = s0 == 0 && s == lane + 64 * j ? px_l[j] : s / cn;
                    const int c = s - px * cn;
                    const int gx = x0 - r + px;
                    const int sx = (gx >= 0 && gx < W) ? gx : border_index(gx, W, border);
                    v[j] = srow[static_cast<size_t>(sx) * cn + c];
                }
            }
#pragma unroll
            for (int j = 0; j < SJ; ++j) {
                const int s = s0 + lane + 64 * j;
                if (s < rowSamples) {
                    const int px = s0 == 0 ? px_l[j] : s / cn;
                    const int c = s0 == 0 ? c_l[j] : s - px * cn;
                    trow[c * plane + px] = KT::to(v[j]);
                }
            }
        }
    }
    __syncthreads();

    if constexpr ((KS == 3 || KS == 5) && KT::bits <= 16) {
        // Rows row and row + TH / 2 of the tile share every comparator (keys of
        // at most 16 bits packed in one register).
        constexpr int G = TX / P;
        static_assert(kBlock % G == 0, "a thread keeps its column group");
        const int HT = TH / 2;  // TH is an even power of two here
        const int ntask = HT * G * cn;
        const int g = tid % G;
        int row = (tid / G) % HT, c = (tid / G) / HT;
        for (int t = tid; t < ntask; t += kBlock, row += kBlock / G) {
            while (row >= HT) {
                row -= HT;
                ++c;
            }
            const K* base = tile + c * plane + row * XP + g * P;
            const K* base2 = base + HT * XP;
            auto ld = [&](int off) {
                KeyPair v;
                v.x = static_cast<unsigned short>(base[off]);
                v.y = static_cast<unsigned short>(base2[off]);
                return v;
            };
            KeyPair res[P];
            if constexpr (KS == 3) {
                KeyPair lo[P + 2], md[P + 2], hi[P + 2];
#pragma unroll
                for (int cc = 0; cc < P + 2; ++cc) {
                    const KeyPair a = ld(cc), b = ld(XP + cc), d = ld(2 * XP + cc);
                    const KeyPair mn = kmin(a, b), mx = kmax(a, b);
                    lo[cc] = kmin(mn, d);
                    hi[cc] = kmax(mx, d);
                    md[cc] = kmax(mn, kmin(mx, d));
                }
#pragma unroll
                for (int m = 0; m < P; ++m) {
                    const KeyPair a = kmax(kmax(lo[m], lo[m + 1]), lo[m + 2]);
                    const KeyPair b = med3(md[m], md[m + 1], md[m + 2]);
                    const KeyPair d = kmin(kmin(hi[m], hi[m + 1]), hi[m + 2]);
                    res[m] = med3(a, b, d);
                }
            } else {
                KeyPair col[P + 4][5];
#pragma unroll
                for (int cc = 0; cc < P + 4; ++cc) {
#pragma unroll
                    for (int j = 0; j < 5; ++j) col[cc][j] = ld(j * XP + cc);
                    sort5(col[cc]);
                }
#pragma unroll
                for (int m = 0; m < P; ++m) {
                    KeyPair win[25];
#pragma unroll
                    for (int cc = 0; cc < 5; ++cc)
#pragma unroll
                        for (int j = 0; j < 5; ++j) win[cc * 5 + j] = col[m + cc][j];
                    res[m] = med25_cols(win);
                }
            }
#pragma unroll
            for (int half = 0; half < 2; ++half) {
                const int y = y0 + row + half * HT;
                if (y >= H) continue;
                T* drow = reinterpret_cast<T*>(reinterpret_cast<char*>(dst) + static_cast<size_t>(y) * dstride);
#pragma unroll
                for (int m = 0; m < P; ++m) {
                    const int x = x0 + g * P + m;
                    if (x >= W) break;
                    drow[static_cast<size_t>(x) * cn + c] = KT::from(half ? res[m].y : res[m].x);
                }
            }
        }
    } else if constexpr (KS == 3 || KS == 5) {
        constexpr int G = TX / P;
        static_assert(kBlock % G == 0, "a thread keeps its column group");
        const int ntask = TH * G * cn;
        // task t = (c * TH + row) * G + g: g stays tid % G, (c, row) advance
        // by kBlock / G rows per step (no division per task)
        const int g = tid % G;
        int row = (tid / G) % TH, c = (tid / G) / TH;
        for (int t = tid; t < ntask; t += kBlock, row += kBlock / G) {
            while (row >= TH) {
                row -= TH;
                ++c;
            }
            const K* base = tile + c * plane + row * XP + g * P;
            K res[P];
            if constexpr (KS == 3) {
                K lo[P + 2], md[P + 2], hi[P + 2];
#pragma unroll
                for (int cc = 0; cc < P + 2; ++cc) {
                    const K a = base[cc], b = base[XP + cc], d = base[2 * XP + cc];
                    const K mn = kmin(a, b), mx = kmax(a, b);
                    lo[cc] = kmin(mn, d);
                    hi[cc] = kmax(mx, d);
                    md[cc] = kmax(mn, kmin(mx, d));
                }
#pragma unroll
                for (int m = 0; m < P; ++m) {
                    const K a = kmax(kmax(lo[m], lo[m + 1]), lo[m + 2]);
                    const K b = med3(md[m], md[m + 1], md[m + 2]);
                    const K d = kmin(kmin(hi[m], hi[m + 1]), hi[m + 2]);
                    res[m] = med3(a, b, d);
                }
            } else {
                K col[P + 4][5];
#pragma unroll
                for (int cc = 0; cc < P + 4; ++cc) {
#pragma unroll
                    for (int j = 0; j < 5; ++j) col[cc][j] = base[j * XP + cc];
                    sort5(col[cc]);
                }
#pragma unroll
                for (int m = 0; m < P; ++m) {
                    K win[25];
#pragma unroll
                    for (int cc = 0; cc < 5; ++cc)
#pragma unroll
                        for (int j = 0; j < 5; ++j) win[cc * 5 + j] = col[m + cc][j];
                    res[m] = med25_cols(win);
                }
            }
            const int y = y0 + row;
            if (y >= H) continue;
            T* drow = reinterpret_cast<T*>(reinterpret_cast<char*>(dst) + static_cast<size_t>(y) * dstride);
#pragma unroll
            for (int m = 0; m < P; ++m) {
                const int x = x0 + g * P + m;
                if (x >= W) break;
                drow[static_cast<size_t>(x) * cn + c] = KT::from(res[m]);
            }
        }
    } else {
        const int rank = ksz * ksz / 2;
        const int ntask = TH * TX * cn;
        for (int t = tid; t < ntask; t += kBlock) {
            const int px = t % TX;
            const int rest = t / TX;
            const int row = rest % TH;
            const int c = rest / TH;
            const int x = x0 + px, y = y0 + row;
            if (x >= W || y >= H) continue;
            const K* base = tile + c * plane + row * XP + px;
            K prefix = 0;
            for (int b = KT::bits - 1; b >= 0; --b) {
                const K cand = prefix | (static_cast<K>(1) << b);
                int below = 0;
                for (int j = 0; j < ksz; ++j)
                    for (int i = 0; i < ksz; ++i) below += base[j * XP + i] < cand;
                if (below <= rank) prefix = cand;
            }
            T* drow = reinterpret_cast<T*>(reinterpret_cast<char*>(dst) + static_cast<size_t>(y) * dstride);
            drow[static_cast<size_t>(x) * cn + c] = KT::from(prefix);
        }
    }
}

template <class T, int KS>
__global__ __launch_bounds__(kBlock) void median_kernel(const T* __restrict__ src, size_t sstride, int W, int H,
                                                        int cn, T* __restrict__ dst, size_t dstride, int k,
                                                        int border, int TH) {
    extern __shared__ unsigned char smem_raw[];
    median_tile<T, KS>(src, sstride, W, H, cn, dst, dstride, k, border, TH, blockIdx.x, blockIdx.y, smem_raw);
}

// Windows too large for an LDS tile: the same bitwise selection reading the
// source through the border tables directly, one thread per sample.
template <class T>
__global__ __launch_bounds__(kBlock) void median_global_kernel(const T* __restrict__ src, size_t sstride, int W,
                                                               int H, int cn, T* __restrict__ dst, size_t dstride,
                                                               int k, int border) {
    using KT = KeyOf<T>;
    using K = typename KT::K;
    const int r = k / 2, rank = k * k / 2;
    const size_t total = static_cast<size_t>(W) * H * cn;
    for (size_t idx = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; idx < total;
         idx += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const int y = static_cast<int>(idx / (static_cast<size_t>(W) * cn));
        const size_t q = idx - static_cast<size_t>(y) * W * cn;
        const int x = static_cast<int>(q / cn), c = static_cast<int>(q - static_cast<size_t>(x) * cn);
        K prefix = 0;
        for (int b = KT::bits - 1; b >= 0; --b) {
            const K cand = prefix | (static_cast<K>(1) << b);
            long long below = 0;
            for (int j = 0; j < k; ++j) {
                const int sy = border_index(y + j - r, H, border);
                const T* srow = reinterpret_cast<const T*>(reinterpret_cast<const char*>(src) + static_cast<size_t>(sy) * sstride);
                for (int i = 0; i < k; ++i) {
                    const int sx = border_index(x + i - r, W, border);
                    below += KT::to(srow[static_cast<size_t>(sx) * cn + c]) < cand;
                }
            }
            if (below <= rank) prefix = cand;
        }
        T* drow = reinterpret_cast<T*>(reinterpret_cast<char*>(dst) + static_cast<size_t>(y) * dstride);
        drow[q] = KT::from(prefix);
    }
}

template <class T, int KS>
hipError_t median_launch(const MedianLaunch& a, hipStream_t s) {
    const int cn = a.src.cn;
    int TH = 32;
    while (TH > 2 && med_lds_bytes<T>(cn, TH, a.k) > 64 * 1024) TH /= 2;  // even: the 16-bit path pairs rows
    const size_t lds = med_lds_bytes<T>(cn, TH, a.k);
    if (lds > 64 * 1024) {
        const size_t total = static_cast<size_t>(a.src.w) * a.src.h * cn;
        const size_t blocks = (total + kBlock - 1) / kBlock;
        hipLaunchKernelGGL(median_global_kernel<T>, dim3(static_cast<unsigned>(blocks < 16384 ? blocks : 16384)),
                           dim3(kBlock), 0, s, static_cast<const T*>(a.src.data), a.src.stride, a.src.w, a.src.h, cn,
                           static_cast<T*>(a.dst.data), a.dst.stride, a.k, a.border);
        return hipGetLastError();
    }
    const dim3 grid(static_cast<unsigned>((a.src.w + TX - 1) / TX), static_cast<unsigned>((a.src.h + TH - 1) / TH));
    hipLaunchKernelGGL((median_kernel<T, KS>), grid, dim3(kBlock), lds, s, static_cast<const T*>(a.src.data),
                       a.src.stride, a.src.w, a.src.h, cn, static_cast<T*>(a.dst.data), a.dst.stride, a.k, a.border,
                       TH);
    return hipGetLastError();
}

template <class T>
hipError_t median_dispatch(const MedianLaunch& a, hipStream_t s) {
    if (a.k == 3) return median_launch<T, 3>(a, s);
    if (a.k == 5) return median_launch<T, 5>(a, s);
    return median_launch<T, 0>(a, s);
}

}  // namespace

hipError_t launch_median(const MedianLaunch& a, hipStream_t s) {
    switch (a.src.dt) {
        case IPFG_U8: return median_dispatch<uint8_t>(a, s);
        case IPFG_U16: return median_dispatch<uint16_t>(a, s);
        case IPFG_I16: return median_dispatch<int16_t>(a, s);
        case IPFG_I32: return median_dispatch<int32_t>(a, s);
        case IPFG_F32: return median_dispatch<float>(a, s);
        case IPFG_F64: return median_dispatch<double>(a, s);
    }
    return hipErrorInvalidValue;
}

// ------------------------------------------------------------------ batched
namespace {
__global__ __launch_bounds__(kBlock) void median_batch_kernel(const BatchClassHeader* __restrict__ hdr,
                                                              const BatchJob* __restrict__ jobs, BatchQueue* q) {
    extern __shared__ unsigned char smem_raw[];
    batch_loop(hdr, jobs, q, [&](const BatchJob& J, int tx, int ty) {
        const uint8_t* src = static_cast<const uint8_t*>(J.src.data);
        uint8_t* dst = static_cast<uint8_t*>(J.dst.data);
        if (J.k == 5)
            median_tile<uint8_t, 5>(src, J.src.stride, J.src.w, J.src.h, J.src.cn, dst, J.dst.stride, 5, J.border, J.rows,
                                    tx, ty, smem_raw);
        else
            median_tile<uint8_t, 3>(src, J.src.stride, J.src.w, J.src.h, J.src.cn, dst, J.dst.stride, 3, J.border, J.rows,
                                    tx, ty, smem_raw);
    });
}
}  // namespace

hipError_t launch_median_batch(const BatchClassHeader* hdr, const BatchJob* jobs, BatchQueue* q, int grid, size_t lds,
                               hipStream_t s) {
    hipLaunchKernelGGL(median_batch_kernel, dim3(grid), dim3(kBlock), lds, s, hdr, jobs, q);
    return hipGetLastError();
}

int median_batch_occupancy(size_t lds) {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, reinterpret_cast<const void*>(median_batch_kernel), kBlock,
                                                     lds) != hipSuccess)
        return 1;
    return n > 0 ? n : 1;
}

size_t median_batch_lds() { return med_lds_bytes<uint8_t>(4, kBatchTileRows, 5); }

}  // namespace ipfg
