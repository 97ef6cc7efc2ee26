// houghCircles for gfx950 (reference: include/ipf/features.hpp:373-504).
//
// The reference runs canny and a Gaussian-smoothed Sobel gradient (the filter
// kernels, on the device here), then two loops that dominate its time: every
// edge pixel votes for centres at every radius, and every accumulator peak
// scans every edge pixel several times (radius histogram, two least-squares
// fits, angular coverage). Here:
//
//  hough_count_kernel / hough_scan_kernel / hough_write_kernel
//      ordered compaction (chunks of 1024 pixels, wave ballots): the edge
//      points in row-major order, as the reference's vector holds them, and
//      later the accumulator peaks in row-major order
//  hough_vote_kernel     one thread per (point, side, radius); votes are
//                        float additions of 1 (exact integers, any order)
//  hough_box_kernel      3 x 3 box sums of the votes (Constant border)
//  hough_refine_kernel   one workgroup per peak: centroid, radius histogram
//                        (LDS atomics), the two Kasa fits (points tested in
//                        parallel, ballots in LDS, the sums accumulated by one
//                        lane in point order so they round as the reference's),
//                        coverage sectors
//
// Sorting the peaks, the minimum-distance pass over them and the final
// de-duplication are short host loops over the candidates, as in the
// reference. Distances use hypot_dd (below): its results match glibc's hypot
// except in rare last-bit cases, which only matter when a rounded distance or
// alignment lands exactly on a decision threshold.
#include "hough.hpp"

namespace ipfg {
namespace {

constexpr int kChunk = 1024;
constexpr double kPiH = 3.14159265358979323846;
constexpr double kAlign = 0.92;  // features.hpp:425

// sqrt(x^2 + y^2) from the exact square sum (double-double through FMA) and
// one correction step of the square root.
__device__ __forceinline__ double hypot_dd(double x, double y) {
    x = fabs(x);
    y = fabs(y);
    const double big = fmax(x, y), small = fmin(x, y);
    if (!(big < 1e150) || (small > 0.0 && small < 1e-150)) return hypot(x, y);  // inf, NaN, extreme ranges
    const double xx = x * x, yy = y * y;
    const double exx = fma(x, x, -xx), eyy = fma(y, y, -yy);
    const double sh = xx + yy;
    const double bb = sh - xx;
    const double err = (xx - (sh - bb)) + (yy - bb);
    const double sl = err + (exx + eyy);
    const double r = sqrt(sh);
    if (r == 0.0) return 0.0;
    const double res = fma(-r, r, sh) + sl;
    return r + res / (2.0 * r);
}

__device__ __forceinline__ double radial(const HoughPoint& e, double cx, double cy, double* dist) {
    const double ux = e.x - cx, uy = e.y - cy;
    *dist = hypot_dd(ux, uy);
    return *dist > 1e-9 ? fabs(ux * e.nx + uy * e.ny) / *dist : 0.0;
}

__device__ __forceinline__ float at_f(const DevImage& im, int x, int y) {
    return reinterpret_cast<const float*>(static_cast<const char*>(im.data) + static_cast<size_t>(y) * im.stride)[x];
}

// mode 0: edge pixel with a usable gradient; mode 1: accumulator peak.
struct Pick {
    int mode;
    DevImage edges, dx, dy;
    const float* sm;
    float thr;
    int w, h;
    __device__ bool operator()(long long p) const {
        if (p >= static_cast<long long>(w) * h) return false;
        const int y = static_cast<int>(p / w), x = static_cast<int>(p - static_cast<long long>(y) * w);
        if (mode == 0) {
            if (!(static_cast<const uint8_t*>(edges.data) + static_cast<size_t>(y) * edges.stride)[x]) return false;
            return !(hypot_dd(at_f(dx, x, y), at_f(dy, x, y)) < 1e-9);
        }
        const float v = sm[p];
        if (!(v > thr)) return false;
        for (int j = max(0, y - 1); j <= min(h - 1, y + 1); ++j)
            for (int i = max(0, x - 1); i <= min(w - 1, x + 1); ++i)
                if (v < sm[static_cast<size_t>(j) * w + i]) return false;
        return true;
    }
};

__global__ __launch_bounds__(kBlock) void hough_count_kernel(Pick pick, unsigned* __restrict__ counts) {
    __shared__ unsigned n;
    if (threadIdx.x == 0) n = 0;
    __syncthreads();
    unsigned mine = 0;
    for (int k = 0; k < kChunk / kBlock; ++k)
        mine += pick(static_cast<long long>(blockIdx.x) * kChunk + k * kBlock + threadIdx.x) ? 1u : 0u;
    atomicAdd(&n, mine);
    __syncthreads();
    if (threadIdx.x == 0) counts[blockIdx.x] = n;
}

// Exclusive scan of the chunk counts (one workgroup); total receives the sum.
__global__ __launch_bounds__(kBlock) void hough_scan_kernel(const unsigned* __restrict__ counts,
                                                            unsigned* __restrict__ offsets, int n,
                                                            unsigned* __restrict__ total) {
    __shared__ unsigned part[kBlock];
    unsigned carry = 0;
    for (int base = 0; base < n; base += kBlock) {
        const int i = base + threadIdx.x;
        const unsigned v = i < n ? counts[i] : 0u;
        part[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < kBlock; off <<= 1) {
            const unsigned add = threadIdx.x >= off ? part[threadIdx.x - off] : 0u;
            __syncthreads();
            part[threadIdx.x] += add;
            __syncthreads();
        }
        if (i < n) offsets[i] = carry + part[threadIdx.x] - v;
        carry += part[kBlock - 1];
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

// Writes the picked items of each chunk in order: points (mode 0: x, y, the
// unit gradient and the gradient with its length) or peaks (mode 1: x, y and
// the box sum).
__global__ __launch_bounds__(kBlock) void hough_write_kernel(Pick pick, const unsigned* __restrict__ offsets,
                                                             HoughPoint* __restrict__ pts, double* __restrict__ grad,
                                                             int* __restrict__ cand_xy, float* __restrict__ cand_v) {
    __shared__ unsigned wave_n[kBlock / 64];
    unsigned base = offsets[blockIdx.x];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int k = 0; k < kChunk / kBlock; ++k) {
        const long long p = static_cast<long long>(blockIdx.x) * kChunk + k * kBlock + threadIdx.x;
        const bool take = pick(p);
        const unsigned long long bal = __ballot(take);
        if (lane == 0) wave_n[wave] = static_cast<unsigned>(__popcll(bal));
        __syncthreads();
        unsigned before = 0, all = 0;
        for (int w = 0; w < kBlock / 64; ++w) {
            before += w < wave ? wave_n[w] : 0u;
            all += wave_n[w];
        }
        const unsigned idx = base + before + static_cast<unsigned>(__popcll(bal & ((1ull << lane) - 1ull)));
        if (take) {
            const int y = static_cast<int>(p / pick.w), x = static_cast<int>(p - static_cast<long long>(y) * pick.w);
            if (pick.mode == 0) {
                const double gx = at_f(pick.dx, x, y), gy = at_f(pick.dy, x, y), g = hypot_dd(gx, gy);
                pts[idx] = HoughPoint{static_cast<double>(x), static_cast<double>(y), gx / g, gy / g};
                grad[3 * static_cast<size_t>(idx)] = gx;
                grad[3 * static_cast<size_t>(idx) + 1] = gy;
                grad[3 * static_cast<size_t>(idx) + 2] = g;
            } else {
                cand_xy[2 * static_cast<size_t>(idx)] = x;
                cand_xy[2 * static_cast<size_t>(idx) + 1] = y;
                cand_v[idx] = pick.sm[p];
            }
        }
        base += all;
        __syncthreads();
    }
}

// Each point votes at lround(p + s r g / |g|) for s = -1, 1 and every radius
// (features.hpp:408-413); grid-stride over (point, side, radius).
__global__ __launch_bounds__(kBlock) void hough_vote_kernel(const HoughPoint* __restrict__ pts,
                                                            const double* __restrict__ grad,
                                                            const unsigned* __restrict__ total, int min_r, int nr,
                                                            float* __restrict__ acc, int w, int h) {
    const unsigned long long n = static_cast<unsigned long long>(*total) * 2ull * nr;
    for (unsigned long long i = blockIdx.x * static_cast<unsigned long long>(kBlock) + threadIdx.x; i < n;
         i += static_cast<unsigned long long>(gridDim.x) * kBlock) {
        const unsigned long long p = i / (2ull * nr);
        const int rem = static_cast<int>(i - p * 2ull * nr);
        const int sgn = rem < nr ? -1 : 1;
        const int r = min_r + (rem < nr ? rem : rem - nr);
        const double gx = grad[3 * p], gy = grad[3 * p + 1], g = grad[3 * p + 2];
        const double x = pts[p].x, y = pts[p].y;
        const long cx = lround(x + static_cast<double>(sgn * r) * gx / g);
        const long cy = lround(y + static_cast<double>(sgn * r) * gy / g);
        if (cx >= 0 && cy >= 0 && cx < w && cy < h) atomicAdd(&acc[static_cast<size_t>(cy) * w + cx], 1.0f);
    }
}

__global__ __launch_bounds__(kBlock) void hough_box_kernel(const float* __restrict__ acc, float* __restrict__ sm,
                                                           int w, int h) {
    const int x = blockIdx.x * kBlock + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    double s = 0.0;
    for (int j = -1; j <= 1; ++j)
        for (int i = -1; i <= 1; ++i) {
            const int sx = x + i, sy = y + j;
            if (sx >= 0 && sy >= 0 && sx < w && sy < h) s += acc[static_cast<size_t>(sy) * w + sx];
        }
    sm[static_cast<size_t>(y) * w + x] = static_cast<float>(s);
}

__global__ __launch_bounds__(kBlock) void zero_f32_kernel(float* __restrict__ p, size_t n) {
    for (size_t i = blockIdx.x * static_cast<size_t>(kBlock) + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * kBlock)
        p[i] = 0.0f;
}

// First index in [0, n) whose point lies on row y or below (the points are
// in row-major order, so their rows never decrease).
__device__ unsigned first_row_at(const HoughPoint* __restrict__ pts, unsigned n, double y) {
    unsigned lo = 0, hi = n;
    while (lo < hi) {
        const unsigned mid = lo + (hi - lo) / 2;
        if (pts[mid].y < y) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// One workgroup per peak (features.hpp:419-488). Every pass takes only points
// within (a radius of) the current centre, so it scans the rows
// [cy - reach, cy + reach] with reach >= that radius + 3: a contiguous range
// of the row-major points holding every point the pass can take, in the same
// order.
__global__ __launch_bounds__(kBlock) void hough_refine_kernel(const HoughPoint* __restrict__ pts, unsigned npts,
                                                              const float* __restrict__ acc, int w, int h,
                                                              const int* __restrict__ peaks, int min_r, int max_r,
                                                              double min_coverage, HoughFit* __restrict__ out) {
    extern __shared__ int hist[];  // max_r + 2
    __shared__ double s_cx, s_cy, s_radius;
    __shared__ int s_best, s_stop, s_support;
    __shared__ unsigned s_lo, s_hi;
    __shared__ unsigned long long s_mask[kBlock / 64];
    __shared__ int s_sector[36];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int px = peaks[2 * blockIdx.x], py = peaks[2 * blockIdx.x + 1];
    if (tid == 0) {  // vote-weighted centroid of the 5 x 5 accumulator window
        double sw = 0, scx = 0, scy = 0;
        for (int j = -2; j <= 2; ++j)
            for (int i = -2; i <= 2; ++i) {
                const int x = px + i, y = py + j;
                if (x < 0 || y < 0 || x >= w || y >= h) continue;
                const double v = acc[static_cast<size_t>(y) * w + x];
                sw += v;
                scx += v * x;
                scy += v * y;
            }
        s_cx = sw > 0 ? scx / sw : px;
        s_cy = sw > 0 ? scy / sw : py;
    }
    for (int i = tid; i < max_r + 2; i += kBlock) hist[i] = 0;
    __syncthreads();
    double cx = s_cx, cy = s_cy;
    // rows [cy - reach, cy + reach] as a point range [lo, hi); every point if
    // the centre is not finite (the distance tests then take NaN paths)
    unsigned lo = 0, hi = npts;
    auto window = [&](double reach) {
        if (tid == 0) {
            const bool finite = isfinite(cy) && isfinite(reach);
            s_lo = finite ? first_row_at(pts, npts, cy - reach) : 0u;
            s_hi = finite ? first_row_at(pts, npts, nextafter(cy + reach, INFINITY)) : npts;
        }
        __syncthreads();
        lo = s_lo;
        hi = s_hi;
        __syncthreads();  // every thread has its copy before the words are rewritten
    };
    window(static_cast<double>(max_r) + 3.0);  // lround(d) <= max_r
    for (unsigned i = lo + tid; i < hi; i += kBlock) {  // radially aligned support per radius
        double d;
        if (radial(pts[i], cx, cy, &d) < kAlign) continue;
        const long r = lround(d);
        if (r >= min_r && r <= max_r) atomicAdd(&hist[r], 1);
    }
    __syncthreads();
    if (tid == 0) {
        int best = 0;
        double best_score = 0;
        for (int r = min_r; r <= max_r; ++r) {
            const double score = (hist[r] + 0.5 * (hist[r - 1] + hist[r + 1])) / (2 * kPiH * r);
            if (score > best_score) {
                best_score = score;
                best = r;
            }
        }
        s_best = best;
        s_radius = best;
        s_stop = 0;
    }
    __syncthreads();
    if (s_best == 0) {
        if (tid == 0) out[blockIdx.x] = HoughFit{cx, cy, 0.0, 0, 0};
        return;
    }
    // two algebraic (Kasa) fits on the supporting points, summed in point order
    for (int iter = 0; iter < 2; ++iter) {
        const double radius = s_radius;
        double sx = 0, sy = 0, sxx = 0, syy = 0, sxy = 0, sz = 0, sxz = 0, syz = 0, cnt = 0;
        window(radius + 3.0);  // |d - radius| <= 2
        for (unsigned base = lo; base < hi; base += kBlock) {
            const unsigned i = base + tid;
            bool take = false;
            if (i < hi) {
                double d;
                take = !(radial(pts[i], cx, cy, &d) < kAlign) && !(fabs(d - radius) > 2.0);
            }
            const unsigned long long bal = __ballot(take);
            if (lane == 0) s_mask[wave] = bal;
            __syncthreads();
            if (tid == 0) {
                for (int wv = 0; wv < kBlock / 64; ++wv) {
                    unsigned long long m = s_mask[wv];
                    while (m) {
                        const int b = __builtin_ctzll(m);
                        m &= m - 1ull;
                        const HoughPoint e = pts[base + wv * 64 + b];
                        const double x = e.x - cx, y = e.y - cy, z = x * x + y * y;
                        sx += x;
                        sy += y;
                        sxx += x * x;
                        syy += y * y;
                        sxy += x * y;
                        sz += z;
                        sxz += x * z;
                        syz += y * z;
                        cnt += 1;
                    }
                }
            }
            __syncthreads();
        }
        if (tid == 0) {
            bool stop = cnt < 3;
            if (!stop) {
                const double m[9] = {sxx, sxy, sx, sxy, syy, sy, sx, sy, cnt}, rhs[3] = {-sxz, -syz, -sz};
                auto det3 = [](const double* t) {
                    return t[0] * (t[4] * t[8] - t[5] * t[7]) - t[1] * (t[3] * t[8] - t[5] * t[6]) +
                           t[2] * (t[3] * t[7] - t[4] * t[6]);
                };
                const double det = det3(m);
                if (fabs(det) < 1e-12) {
                    stop = true;
                } else {
                    double sol[3];
                    for (int col = 0; col < 3; ++col) {
                        double t[9];
                        for (int q = 0; q < 9; ++q) t[q] = m[q];
                        for (int r = 0; r < 3; ++r) t[r * 3 + col] = rhs[r];
                        sol[col] = det3(t) / det;
                    }
                    const double ox = -sol[0] / 2, oy = -sol[1] / 2, r2 = ox * ox + oy * oy - sol[2];
                    if (r2 <= 0) {
                        stop = true;
                    } else {
                        s_cx = cx + ox;
                        s_cy = cy + oy;
                        s_radius = sqrt(r2);
                    }
                }
            }
            s_stop = stop ? 1 : 0;
        }
        __syncthreads();
        cx = s_cx;
        cy = s_cy;
        if (s_stop) break;
    }
    const double radius = s_radius;
    if (radius < min_r - 1 || radius > max_r + 1) {
        if (tid == 0) out[blockIdx.x] = HoughFit{cx, cy, radius, 0, 0};
        return;
    }
    // angular coverage of the supporting points (36 sectors)
    if (tid < 36) s_sector[tid] = 0;
    if (tid == 0) s_support = 0;
    __syncthreads();
    int mine = 0;
    window(radius + 3.0);  // |d - radius| <= 1.5
    for (unsigned i = lo + tid; i < hi; i += kBlock) {
        const HoughPoint e = pts[i];
        double d;
        if (radial(e, cx, cy, &d) < kAlign || fabs(d - radius) > 1.5) continue;
        ++mine;
        const double a = atan2(e.y - cy, e.x - cx) + kPiH;
        s_sector[min(35, static_cast<int>(a / (2 * kPiH) * 36))] = 1;
    }
    atomicAdd(&s_support, mine);
    __syncthreads();
    if (tid == 0) {
        int covered = 0;
        for (int k = 0; k < 36; ++k) covered += s_sector[k];
        out[blockIdx.x] = HoughFit{cx, cy, radius, s_support, covered < min_coverage * 36 ? 0 : 1};
    }
}

}  // namespace

size_t hough_chunks(int w, int h) { return (static_cast<size_t>(w) * h + kChunk - 1) / kChunk; }

hipError_t launch_hough_votes(const DevImage& edges, const DevImage& dx, const DevImage& dy, int min_r, int max_r,
                              float threshold, const HoughScratch& sc, hipStream_t s) {
    const int w = edges.w, h = edges.h;
    const size_t n = static_cast<size_t>(w) * h;
    const unsigned chunks = static_cast<unsigned>(hough_chunks(w, h));
    unsigned* offsets = sc.counts + chunks;
    Pick pick{0, edges, dx, dy, sc.sm, threshold, w, h};
    hipLaunchKernelGGL(hough_count_kernel, dim3(chunks), dim3(kBlock), 0, s, pick, sc.counts);
    hipLaunchKernelGGL(hough_scan_kernel, dim3(1), dim3(kBlock), 0, s, sc.counts, offsets, static_cast<int>(chunks),
                       sc.total);
    hipLaunchKernelGGL(hough_write_kernel, dim3(chunks), dim3(kBlock), 0, s, pick, offsets, sc.pts, sc.grad,
                       sc.cand_xy, sc.cand_v);
    const unsigned fill = static_cast<unsigned>(std::min<size_t>((n + kBlock - 1) / kBlock, 8192));
    hipLaunchKernelGGL(zero_f32_kernel, dim3(fill), dim3(kBlock), 0, s, sc.acc, n);
    hipLaunchKernelGGL(hough_vote_kernel, dim3(16384), dim3(kBlock), 0, s, sc.pts, sc.grad, sc.total, min_r,
                       max_r - min_r + 1, sc.acc, w, h);
    hipLaunchKernelGGL(hough_box_kernel, dim3(static_cast<unsigned>((w + kBlock - 1) / kBlock), static_cast<unsigned>(h)),
                       dim3(kBlock), 0, s, sc.acc, sc.sm, w, h);
    pick.mode = 1;
    hipLaunchKernelGGL(hough_count_kernel, dim3(chunks), dim3(kBlock), 0, s, pick, sc.counts);
    hipLaunchKernelGGL(hough_scan_kernel, dim3(1), dim3(kBlock), 0, s, sc.counts, offsets, static_cast<int>(chunks),
                       sc.total + 1);
    hipLaunchKernelGGL(hough_write_kernel, dim3(chunks), dim3(kBlock), 0, s, pick, offsets, sc.pts, sc.grad,
                       sc.cand_xy, sc.cand_v);
    return hipGetLastError();
}

hipError_t launch_hough_refine(const HoughPoint* pts, unsigned npts, const float* acc, int w, int h, const int* peaks,
                               int n, int min_r, int max_r, double min_coverage, HoughFit* out, hipStream_t s) {
    if (n == 0) return hipSuccess;
    if (max_r > kHoughMaxRadius) return hipErrorNotSupported;
    const size_t lds = sizeof(int) * static_cast<size_t>(max_r + 2);
    hipLaunchKernelGGL(hough_refine_kernel, dim3(static_cast<unsigned>(n)), dim3(kBlock), lds, s, pts, npts, acc, w, h,
                       peaks, min_r, max_r, min_coverage, out);
    return hipGetLastError();
}

}  // namespace ipfg
