// Feature callers of the filters for gfx950: canny, cannyAuto, cornerHarris,
// cornerMinEigenVal (reference: include/ipf/features.hpp:26-160).
//
// The derivatives come from the separable kernels (sobel, certified to the
// reference's bits); what is here is the per-pixel work after them, with every
// intermediate in device memory:
//
//  canny      canny_nms_kernel     strongest channel, quantised-direction
//                                  non-maximum suppression, weak/strong state
//             canny_merge_kernel   8-connected components of the candidates
//                                  (union-find with CAS hooking and pointer
//                                  jumping, after ECL-CC: Jaiganesh and
//                                  Burtscher, HPDC 2018)
//             canny_flag_kernel    roots that hold a strong candidate
//             canny_out_kernel     255 where the pixel's component is flagged
//  cannyAuto  to_gray_kernel, then a radix select of the magnitude's rank
//             7n/10 (select_*), whose value stays on the device and sets the
//             thresholds canny_nms_kernel reads
//  corners    tensor_rows_kernel   products of the derivatives and their
//                                  horizontal running sums (the reference's
//                                  boxFilter order), one wave per row
//             tensor_cols_kernel   vertical running sums and the Harris or
//                                  minimum-eigenvalue response, one thread per
//                                  column
//
// Hysteresis keeps the weak candidates connected to a strong one through other
// candidates; that set is the union of the candidate components that contain
// a strong pixel, so labelling the components gives the reference's result
// whatever order the reference's queue visits pixels in.
#include <algorithm>

#include "kernels.hpp"

namespace ipfg {
namespace {

__device__ __forceinline__ const float* frow(const DevImage& im, int y) {
    return reinterpret_cast<const float*>(static_cast<const char*>(im.data) + static_cast<size_t>(y) * im.stride);
}

// canny's per-pixel magnitude (features.hpp:35-47): the channel with the
// largest L1 (or float L2) magnitude; ties keep the first channel. With no
// channel above -1 (all NaN) the magnitude is -1 and the direction (0, 0).
__device__ __forceinline__ float best_mag(const DevImage& gx, const DevImage& gy, int x, int y, bool l2, float* sx,
                                          float* sy) {
    const float* ax = frow(gx, y) + static_cast<size_t>(x) * gx.cn;
    const float* ay = frow(gy, y) + static_cast<size_t>(x) * gy.cn;
    float best = -1.0f, bx = 0.0f, by = 0.0f;
    for (int c = 0; c < gx.cn; ++c) {
        const float a = ax[c], b = ay[c];
        const float m = l2 ? __builtin_sqrtf(a * a + b * b) : fabsf(a) + fabsf(b);
        if (m > best) {
            best = m;
            bx = a;
            by = b;
        }
    }
    *sx = bx;
    *sy = by;
    return best;
}

__device__ __forceinline__ float mag_at(const DevImage& gx, const DevImage& gy, int x, int y, bool l2) {
    if (x < 0 || y < 0 || x >= gx.w || y >= gx.h) return 0.0f;
    float a, b;
    return best_mag(gx, gy, x, y, l2, &a, &b);
}

// Non-maximum suppression (features.hpp:50-69). state: 0 none, 1 weak, 2
// strong; parent: the pixel's own index for candidates (union-find roots).
// When `sel` is set, the thresholds are cannyAuto's (features.hpp:98-100):
// high = max(1e-6, the selected magnitude), low = 0.4 high.
__global__ __launch_bounds__(kBlock) void canny_nms_kernel(DevImage gx, DevImage gy, double low, double high, int l2,
                                                           const unsigned int* __restrict__ sel,
                                                           uint8_t* __restrict__ state, int* __restrict__ parent) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= gx.w || y >= gx.h) return;
    if (sel != nullptr) {
        high = fmax(1e-6, static_cast<double>(__uint_as_float(*sel)));
        low = 0.4 * high;
    }
    const size_t p = static_cast<size_t>(y) * gx.w + x;
    float dx, dy;
    const float m = best_mag(gx, gy, x, y, l2 != 0, &dx, &dy);
    uint8_t st = 0;
    if (static_cast<double>(m) > low) {
        const float tan22 = 0.41421356f;
        const float ax = fabsf(dx), ay = fabsf(dy);
        bool peak;
        if (ay <= ax * tan22) {
            peak = m > mag_at(gx, gy, x - 1, y, l2) && m >= mag_at(gx, gy, x + 1, y, l2);
        } else if (ay * tan22 >= ax) {
            peak = m > mag_at(gx, gy, x, y - 1, l2) && m >= mag_at(gx, gy, x, y + 1, l2);
        } else {
            const int s = (dx < 0) != (dy < 0) ? -1 : 1;
            peak = m > mag_at(gx, gy, x - s, y - 1, l2) && m > mag_at(gx, gy, x + s, y + 1, l2);
        }
        if (peak) st = static_cast<double>(m) > high ? 2 : 1;
    }
    state[p] = st;
    parent[p] = st ? static_cast<int>(p) : -1;
}

// Union-find over the candidates. Every parent index is below its child's
// (roots point at themselves) and only roots are changed by CAS, so the plain
// pointer-jumping stores of `root` only ever replace a parent by one of its
// ancestors; agent-scope atomics keep the view of other XCDs' updates current.
__device__ __forceinline__ int ld(const int* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int root(int* par, int v) {
    int cur = ld(par + v);
    if (cur == v) return v;
    int prev = v, next;
    while (cur > (next = ld(par + cur))) {
        __hip_atomic_store(par + prev, next, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        prev = cur;
        cur = next;
    }
    return cur;
}

__device__ __forceinline__ void unite(int* par, int a, int b) {
    a = root(par, a);
    b = root(par, b);
    while (a != b) {
        if (a < b) {  // hang b's tree under a (only if b is still a root)
            int expect = b;
            if (__hip_atomic_compare_exchange_strong(par + b, &expect, a, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT))
                return;
            b = root(par, expect);
        } else {
            int expect = a;
            if (__hip_atomic_compare_exchange_strong(par + a, &expect, b, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT))
                return;
            a = root(par, expect);
        }
    }
}

// Joins each candidate with its candidate neighbours above and to the left
// (the other four of the 8 are joined from their side).
__global__ __launch_bounds__(kBlock) void canny_merge_kernel(const uint8_t* __restrict__ state, int* parent, int W,
                                                             int H) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= W || y >= H) return;
    const int p = y * W + x;
    if (!state[p]) return;
    if (x > 0 && state[p - 1]) unite(parent, p, p - 1);
    if (y > 0) {
        const int q = p - W;
        if (x > 0 && state[q - 1]) unite(parent, p, q - 1);
        if (state[q]) unite(parent, p, q);
        if (x + 1 < W && state[q + 1]) unite(parent, p, q + 1);
    }
}

// Marks the roots of components that hold a strong candidate (and shortens
// the candidates' paths; another thread's pointer jump may still leave a
// candidate on an intermediate node, so readers walk to the root).
__global__ __launch_bounds__(kBlock) void canny_flag_kernel(const uint8_t* __restrict__ state, int* parent,
                                                            uint8_t* __restrict__ flag, int n) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= n) return;
    const uint8_t st = state[p];
    if (!st) return;
    const int r = root(parent, p);
    __hip_atomic_store(parent + p, r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (st == 2) flag[r] = 1;
}

__global__ __launch_bounds__(kBlock) void canny_out_kernel(const uint8_t* __restrict__ state,
                                                           const int* __restrict__ parent,
                                                           const uint8_t* __restrict__ flag, DevImage dst) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= dst.w || y >= dst.h) return;
    const size_t p = static_cast<size_t>(y) * dst.w + x;
    bool edge = false;
    if (state[p]) {  // the forest is final here; walk (read only) to the root
        int r = parent[p];
        for (int up = parent[r]; up != r; up = parent[r]) r = up;
        edge = flag[r] != 0;
    }
    (static_cast<uint8_t*>(dst.data) + static_cast<size_t>(y) * dst.stride)[x] = edge ? 255 : 0;
}

__global__ __launch_bounds__(kBlock) void fill_u8_kernel(uint8_t* __restrict__ p, size_t n) {
    for (size_t i = blockIdx.x * static_cast<size_t>(kBlock) + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * kBlock)
        p[i] = 0;
}

// toGray (color.hpp:51-81): channel 0 of a 1- or 2-channel image; Luma 601 of
// the first three channels in FP64 (0.299 r + 0.587 g + 0.114 b, left to
// right) with saturate_cast for 3 and 4 channels.
template <class T>
__global__ __launch_bounds__(kBlock) void to_gray_kernel(DevImage src, DevImage dst) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= src.w || y >= src.h) return;
    const T* s = reinterpret_cast<const T*>(static_cast<const char*>(src.data) + static_cast<size_t>(y) * src.stride) +
                 static_cast<size_t>(x) * src.cn;
    T* d = reinterpret_cast<T*>(static_cast<char*>(dst.data) + static_cast<size_t>(y) * dst.stride) + x;
    if (src.cn <= 2) {
        *d = s[0];
        return;
    }
    const double r = static_cast<double>(s[0]), g = static_cast<double>(s[1]), b = static_cast<double>(s[2]);
    *d = saturate_from_double<T>(0.299 * r + 0.587 * g + 0.114 * b);
}

// ------------------------------------------------------------ radix select
// The element of rank k (0-based, ascending) of n non-negative floats, whose
// bit patterns order like their values: three passes over the digits 31-21,
// 20-10 and 9-0, each counting the elements that match the digits chosen so
// far. sel[0] accumulates the chosen bits, sel[1] the mask of chosen bits,
// sel[2..3] the remaining rank (64-bit).
constexpr int kSelBins = 2048;

__global__ __launch_bounds__(kBlock) void select_init_kernel(unsigned int* __restrict__ sel,
                                                             unsigned int* __restrict__ hist, unsigned long long k) {
    for (int i = threadIdx.x; i < kSelBins; i += kBlock) hist[i] = 0u;
    if (threadIdx.x == 0) {
        sel[0] = 0u;
        sel[1] = 0u;
        *reinterpret_cast<unsigned long long*>(sel + 2) = k;
    }
}

__global__ __launch_bounds__(kBlock) void select_hist_kernel(const unsigned int* __restrict__ v, size_t n,
                                                             const unsigned int* __restrict__ sel,
                                                             unsigned int* __restrict__ hist, int shift) {
    __shared__ unsigned int h[kSelBins];
    for (int i = threadIdx.x; i < kSelBins; i += kBlock) h[i] = 0u;
    __syncthreads();
    const unsigned int want = sel[0], mask = sel[1];
    for (size_t i = blockIdx.x * static_cast<size_t>(kBlock) + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * kBlock) {
        const unsigned int b = v[i];
        if ((b & mask) == want) atomicAdd(&h[(b >> shift) & (kSelBins - 1)], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kSelBins; i += kBlock)
        if (h[i]) atomicAdd(&hist[i], h[i]);
}

// One workgroup: finds the bin holding the remaining rank, fixes its digit and
// clears the histogram for the next pass.
__global__ __launch_bounds__(kBlock) void select_pick_kernel(unsigned int* __restrict__ sel,
                                                             unsigned int* __restrict__ hist, int shift, int bits) {
    constexpr int per = kSelBins / kBlock;  // 8 bins per thread
    __shared__ unsigned long long part[kBlock];
    unsigned long long c[per], sum = 0;
    for (int j = 0; j < per; ++j) {
        c[j] = hist[threadIdx.x * per + j];
        sum += c[j];
    }
    part[threadIdx.x] = sum;
    __syncthreads();
    // inclusive scan of the per-thread sums (Hillis-Steele)
    for (int off = 1; off < kBlock; off <<= 1) {
        const unsigned long long add = threadIdx.x >= off ? part[threadIdx.x - off] : 0ull;
        __syncthreads();
        part[threadIdx.x] += add;
        __syncthreads();
    }
    const unsigned long long k = *reinterpret_cast<const unsigned long long*>(sel + 2);
    unsigned long long below = threadIdx.x ? part[threadIdx.x - 1] : 0ull;
    __syncthreads();
    if (k >= below && k < part[threadIdx.x]) {  // exactly one thread holds the rank
        for (int j = 0; j < per; ++j) {
            if (k < below + c[j]) {
                const unsigned int digit = static_cast<unsigned int>(threadIdx.x * per + j);
                sel[0] |= digit << shift;
                sel[1] |= ((1u << bits) - 1u) << shift;
                *reinterpret_cast<unsigned long long*>(sel + 2) = k - below;
                break;
            }
            below += c[j];
        }
    }
    for (int j = 0; j < per; ++j) hist[threadIdx.x * per + j] = 0u;
}

// ------------------------------------------------------------ corners
// Structure tensor (features.hpp:107-126) and the responses (128-153). The
// reference forms xx = dx dx, xy = dx dy, yy = dy dy in float, then
// boxFilter(., block, block, false) with Reflect101: horizontal running sums
// s += v(x + kw - 1) - v(x - 1) seeded by the sum of the first window, then
// vertical running sums of those rows, each in FP64 and in that order, and
// static_cast<float>(sum * 1.0). The passes below keep that order exactly.
// rows: 3 planes of H x W doubles (xx, xy, yy horizontal sums of image row y;
// padded rows reuse the row they reflect).
//
// Rows: one wave per image row. Lanes load 64 consecutive window positions
// (coalesced) and form the entering-minus-leaving products in parallel; the
// running sums then pass through the lanes in order (lane i's term broadcast
// with v_readlane, added, kept by lane i), the reference's sequential
// additions exactly. Columns: one thread per column, a dependent chain whose
// loads two register buffers of kTensorChunk rows keep in flight.
constexpr int kTensorChunk = 8;

__global__ __launch_bounds__(kBlock) void tensor_rows_kernel(DevImage dx, DevImage dy, int kw,
                                                             double* __restrict__ rows) {
    const int lane = threadIdx.x & 63;
    const int y = blockIdx.x * (blockDim.x / 64) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int W = dx.w, H = dx.h;
    if (y >= H) return;  // whole waves
    const float* __restrict__ a = frow(dx, y);
    const float* __restrict__ b = frow(dy, y);
    const size_t plane = static_cast<size_t>(W) * H;
    double* __restrict__ rxx = rows + static_cast<size_t>(y) * W;
    double* __restrict__ rxy = rxx + plane;
    double* __restrict__ ryy = rxy + plane;
    const int ax = kw / 2;
    auto src_x = [&](int px) { return border_index(px - ax, W, IPFG_BORDER_REFLECT101); };
    auto prod = [](float u, float v, double& xx, double& xy, double& yy) {
        xx = static_cast<double>(u * u);
        xy = static_cast<double>(u * v);
        yy = static_cast<double>(v * v);
    };
    // the first window, summed in order (every lane holds the sums)
    double sxx = 0.0, sxy = 0.0, syy = 0.0;
    for (int i = 0; i < kw; ++i) {
        const int sx = src_x(i);
        double p, q, r;
        prod(a[sx], b[sx], p, q, r);
        sxx += p;
        sxy += q;
        syy += r;
    }
    if (lane == 0) {
        rxx[0] = sxx;
        rxy[0] = sxy;
        ryy[0] = syy;
    }
    // step x (1 <= x < W) adds window position x + kw - 1 and drops x - 1;
    // the next chunk's samples are loaded before this chunk's chain runs
    float ui, vi, uo, vo;
    auto load = [&](int x0) {
        const int x = min(x0 + lane, W - 1);  // lanes past the row load a valid sample (unused)
        const int si = src_x(x + kw - 1), so = src_x(x - 1);
        ui = a[si];
        vi = b[si];
        uo = a[so];
        vo = b[so];
    };
    load(1);
    for (int x0 = 1; x0 < W; x0 += 64) {
        double p1, q1, r1, p0, q0, r0;
        prod(ui, vi, p1, q1, r1);
        prod(uo, vo, p0, q0, r0);
        const double dxx = p1 - p0, dxy = q1 - q0, dyy = r1 - r0;
        if (x0 + 64 < W) load(x0 + 64);
        const int n = min(64, W - x0);
        double oxx = 0.0, oxy = 0.0, oyy = 0.0;
#pragma unroll 8
        for (int i = 0; i < n; ++i) {
            sxx += bcast_d(dxx, i);
            sxy += bcast_d(dxy, i);
            syy += bcast_d(dyy, i);
            if (lane == i) {
                oxx = sxx;
                oxy = sxy;
                oyy = syy;
            }
        }
        const int x = x0 + lane;
        if (x < W) {
            rxx[x] = oxx;
            rxy[x] = oxy;
            ryy[x] = oyy;
        }
    }
}

// harris != 0: A C - B^2 - k (A + C)^2 (features.hpp:128-139); else the
// minimum eigenvalue of [[A/2, B], [B, C/2]] form (141-153).
__global__ __launch_bounds__(kBlock) void tensor_cols_kernel(const double* __restrict__ rows, int W, int H, int kh,
                                                             int harris, double k, DevImage dst) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= W) return;
    const size_t plane = static_cast<size_t>(W) * H;
    const int ay = kh / 2;
    auto row = [&](int py) { return rows + static_cast<size_t>(border_index(py - ay, H, IPFG_BORDER_REFLECT101)) * W + x; };
    double ca = 0.0, cb = 0.0, cc = 0.0;
    for (int j = 0; j < kh; ++j) {
        const double* r = row(j);
        ca += r[0];
        cb += r[plane];
        cc += r[2 * plane];
    }
    // after output row y the window adds row y + kh and drops row y
    struct Chunk {
        double ia[kTensorChunk], ib[kTensorChunk], ic[kTensorChunk];
        double oa[kTensorChunk], ob[kTensorChunk], oc[kTensorChunk];
    };
    auto load = [&](Chunk& c, int y0) {  // steps past the last row load valid rows (unused)
#pragma unroll
        for (int j = 0; j < kTensorChunk; ++j) {
            const int y = min(y0 + j, H - 1);
            const double* in = row(y + kh);
            const double* out = row(y);
            c.ia[j] = in[0];
            c.ib[j] = in[plane];
            c.ic[j] = in[2 * plane];
            c.oa[j] = out[0];
            c.ob[j] = out[plane];
            c.oc[j] = out[2 * plane];
        }
    };
    auto run = [&](const Chunk& c, int y0) {
#pragma unroll
        for (int j = 0; j < kTensorChunk; ++j) {
            const int y = y0 + j;
            if (y < H) {
                const float fa = static_cast<float>(ca * 1.0), fb = static_cast<float>(cb * 1.0),
                            fc = static_cast<float>(cc * 1.0);
                float out;
                if (harris) {
                    const double A = fa, B = fb, C = fc;
                    out = static_cast<float>(A * C - B * B - k * (A + C) * (A + C));
                } else {
                    const double A = fa * 0.5, B = fb, C = fc * 0.5;
                    out = static_cast<float>((A + C) - __builtin_sqrt((A - C) * (A - C) + B * B));
                }
                reinterpret_cast<float*>(static_cast<char*>(dst.data) + static_cast<size_t>(y) * dst.stride)[x] = out;
                if (y + 1 < H) {
                    ca += c.ia[j] - c.oa[j];
                    cb += c.ib[j] - c.ob[j];
                    cc += c.ic[j] - c.oc[j];
                }
            }
        }
    };
    Chunk qa, qb;
    load(qa, 0);
    for (int y0 = 0; y0 < H; y0 += 2 * kTensorChunk) {
        load(qb, y0 + kTensorChunk);
        run(qa, y0);
        load(qa, y0 + 2 * kTensorChunk);
        run(qb, y0 + kTensorChunk);
    }
}

dim3 pixel_grid(int w, int h) {
    return dim3(static_cast<unsigned>((w + 63) / 64), static_cast<unsigned>((h + 3) / 4));
}

}  // namespace

hipError_t launch_canny(const DevImage& gx, const DevImage& gy, double low, double high, bool l2,
                        const unsigned int* sel, const CannyScratch& sc, const DevImage& dst, hipStream_t s) {
    const int W = gx.w, H = gx.h;
    const size_t n = static_cast<size_t>(W) * H;
    if (n == 0) return hipSuccess;
    if (n > static_cast<size_t>(INT32_MAX)) return hipErrorNotSupported;  // int labels
    const dim3 g = pixel_grid(W, H);
    hipLaunchKernelGGL(canny_nms_kernel, g, dim3(kBlock), 0, s, gx, gy, low, high, l2 ? 1 : 0, sel, sc.state,
                       sc.parent);
    hipLaunchKernelGGL(fill_u8_kernel, dim3(static_cast<unsigned>(std::min<size_t>((n + kBlock - 1) / kBlock, 4096))),
                       dim3(kBlock), 0, s, sc.flag, n);
    hipLaunchKernelGGL(canny_merge_kernel, g, dim3(kBlock), 0, s, sc.state, sc.parent, W, H);
    const unsigned lin = static_cast<unsigned>((n + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(canny_flag_kernel, dim3(lin), dim3(kBlock), 0, s, sc.state, sc.parent, sc.flag,
                       static_cast<int>(n));
    hipLaunchKernelGGL(canny_out_kernel, g, dim3(kBlock), 0, s, sc.state, sc.parent, sc.flag, dst);
    return hipGetLastError();
}

hipError_t launch_to_gray(const DevImage& src, const DevImage& dst, hipStream_t s) {
    if (static_cast<long long>(src.w) * src.h == 0) return hipSuccess;
    const dim3 g = pixel_grid(src.w, src.h);
    switch (src.dt) {
        case IPFG_U8: hipLaunchKernelGGL(to_gray_kernel<uint8_t>, g, dim3(kBlock), 0, s, src, dst); break;
        case IPFG_U16: hipLaunchKernelGGL(to_gray_kernel<uint16_t>, g, dim3(kBlock), 0, s, src, dst); break;
        case IPFG_I16: hipLaunchKernelGGL(to_gray_kernel<int16_t>, g, dim3(kBlock), 0, s, src, dst); break;
        case IPFG_I32: hipLaunchKernelGGL(to_gray_kernel<int32_t>, g, dim3(kBlock), 0, s, src, dst); break;
        case IPFG_F32: hipLaunchKernelGGL(to_gray_kernel<float>, g, dim3(kBlock), 0, s, src, dst); break;
        case IPFG_F64: hipLaunchKernelGGL(to_gray_kernel<double>, g, dim3(kBlock), 0, s, src, dst); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

size_t select_scratch_bytes() { return (4 + kSelBins) * sizeof(unsigned int); }

hipError_t launch_select_rank(const float* v, size_t n, unsigned long long k, unsigned int* scratch, hipStream_t s) {
    if (n == 0 || k >= n) return hipErrorInvalidValue;
    unsigned int* sel = scratch;
    unsigned int* hist = scratch + 4;
    hipLaunchKernelGGL(select_init_kernel, dim3(1), dim3(kBlock), 0, s, sel, hist, k);
    const unsigned grid = static_cast<unsigned>(std::min<size_t>((n + kBlock - 1) / kBlock, 2048));
    const int shifts[3] = {21, 10, 0}, bits[3] = {11, 11, 10};
    for (int pass = 0; pass < 3; ++pass) {
        hipLaunchKernelGGL(select_hist_kernel, dim3(grid), dim3(kBlock), 0, s,
                           reinterpret_cast<const unsigned int*>(v), n, sel, hist, shifts[pass]);
        hipLaunchKernelGGL(select_pick_kernel, dim3(1), dim3(kBlock), 0, s, sel, hist, shifts[pass], bits[pass]);
    }
    return hipGetLastError();
}

hipError_t launch_structure_response(const DevImage& dx, const DevImage& dy, int block, bool harris, double k,
                                     double* rows, const DevImage& dst, hipStream_t s) {
    const int W = dx.w, H = dx.h;
    if (static_cast<long long>(W) * H == 0) return hipSuccess;
    // one image row per one-wave workgroup (spread evenly over the CUs)
    hipLaunchKernelGGL(tensor_rows_kernel, dim3(static_cast<unsigned>(H)), dim3(64), 0, s, dx, dy, block, rows);
    // one wave per workgroup: the few column waves spread over as many CUs
    hipLaunchKernelGGL(tensor_cols_kernel, dim3(static_cast<unsigned>((W + 63) / 64)), dim3(64), 0, s, rows, W, H, block,
                       harris ? 1 : 0, k, dst);
    return hipGetLastError();
}

}  // namespace ipfg
