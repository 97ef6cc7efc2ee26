// Structural similarity (ipf::ssim) for gfx950.
//
// Reference: include/ipf/metrics.hpp:46-78. Both images are converted to
// double, the products x*x, y*y and x*y formed per sample, and the five
// images x, y, xx, yy, xy each blurred with gaussianBlur(1.5, 1.5, 11,
// Reflect101) (sepFilter2D<double>: a horizontal pass into a double buffer,
// every sum starting at +0.0 and adding w * v in tap order, then a vertical
// pass starting at delta = +0.0 and adding w * h in row order). Each map
// sample is float(((2 ux uy + C1)(2 cxy + C2)) / ((ux^2 + uy^2 + C1)(vx + vy
// + C2))) and the result is the mean of the map over the valid region.
//
// One launch forms the map: a thread owns one sample column and kSsimRows
// output rows, streams the kSsimRows + 10 padded rows it needs, computes the
// five horizontal sums of each once and adds them to the vertical sums of
// the outputs that use the row (so each vertical sum still runs in row
// order). The arithmetic is the reference's operation for operation (FP64,
// separate multiply and add, IEEE division), so the map is bit-identical.
// The same launch sums the map over the valid region per workgroup; a second
// one-workgroup launch adds the partial sums in a fixed order. The mean thus
// differs from the reference's row-major running sum only by summation order.
#include "kernels.hpp"

namespace ipfg {
namespace {

constexpr int kSsimRows = 8;   // output rows per thread
constexpr int kSsimTaps = 11;  // window of the reference (radius 5)
constexpr int kSsimR = kSsimTaps / 2;
constexpr int kSumBlock = 1024;

template <class T>
__device__ __forceinline__ double ssim_ld(const char* row, int off) {
    return static_cast<double>(*reinterpret_cast<const T*>(row + off));
}

__device__ __forceinline__ double madd(double s, double w, double v) { return __dadd_rn(s, __dmul_rn(w, v)); }

// Sum of v over the workgroup (kBlock lanes), valid in thread 0.
__device__ double block_sum(double v) {
    __shared__ double part[kBlock / 64];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) part[wv] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < kBlock / 64; ++i) s += part[i];
    return s;
}

template <class T>
__global__ __launch_bounds__(kBlock) void ssim_map_kernel(DevImage a, DevImage b, SsimTaps k, float* map,
                                                          size_t map_stride, SsimRoi roi, double* partial) {
    const int cn = a.cn;
    const int row_samples = a.w * cn;
    const int q = static_cast<int>(blockIdx.x) * kBlock + static_cast<int>(threadIdx.x);
    const bool live = q < row_samples;
    const int qc = live ? q : 0;
    const int x = qc / cn, c = qc - x * cn;
    const int sz = static_cast<int>(sizeof(T));
    int off[kSsimTaps];
#pragma unroll
    for (int t = 0; t < kSsimTaps; ++t)
        off[t] = (border_index(x - kSsimR + t, a.w, IPFG_BORDER_REFLECT101) * cn + c) * sz;
    const bool in_cols = live && x >= roi.x0 && x < roi.x1;
    const char* abase = static_cast<const char*>(a.data);
    const char* bbase = static_cast<const char*>(b.data);
    double part = 0.0;
    for (int y0 = blockIdx.y * kSsimRows; y0 < a.h; y0 += gridDim.y * kSsimRows) {
        double mx[kSsimRows], my[kSsimRows], sxx[kSsimRows], syy[kSsimRows], sxy[kSsimRows];
#pragma unroll
        for (int m = 0; m < kSsimRows; ++m) mx[m] = my[m] = sxx[m] = syy[m] = sxy[m] = 0.0;
        if (live) {
#pragma unroll
            for (int p = 0; p < kSsimRows + kSsimTaps - 1; ++p) {
                const int sy = border_index(y0 + p - kSsimR, a.h, IPFG_BORDER_REFLECT101);
                const char* ra = abase + static_cast<size_t>(sy) * a.stride;
                const char* rb = bbase + static_cast<size_t>(sy) * b.stride;
                double hx = 0.0, hy = 0.0, hxx = 0.0, hyy = 0.0, hxy = 0.0;
#pragma unroll
                for (int t = 0; t < kSsimTaps; ++t) {
                    const double vx = ssim_ld<T>(ra, off[t]), vy = ssim_ld<T>(rb, off[t]);
                    hx = madd(hx, k.w[t], vx);
                    hy = madd(hy, k.w[t], vy);
                    hxx = madd(hxx, k.w[t], __dmul_rn(vx, vx));
                    hyy = madd(hyy, k.w[t], __dmul_rn(vy, vy));
                    hxy = madd(hxy, k.w[t], __dmul_rn(vx, vy));
                }
#pragma unroll
                for (int m = 0; m < kSsimRows; ++m) {
                    const int j = p - m;  // tap of padded row p in output row m
                    if (j >= 0 && j < kSsimTaps) {
                        mx[m] = madd(mx[m], k.w[j], hx);
                        my[m] = madd(my[m], k.w[j], hy);
                        sxx[m] = madd(sxx[m], k.w[j], hxx);
                        syy[m] = madd(syy[m], k.w[j], hyy);
                        sxy[m] = madd(sxy[m], k.w[j], hxy);
                    }
                }
            }
#pragma unroll
            for (int m = 0; m < kSsimRows; ++m) {
                const int y = y0 + m;
                if (y < a.h) {
                    const double ux = mx[m], uy = my[m];
                    const double vx = __dsub_rn(sxx[m], __dmul_rn(ux, ux));
                    const double vy = __dsub_rn(syy[m], __dmul_rn(uy, uy));
                    const double cxy = __dsub_rn(sxy[m], __dmul_rn(ux, uy));
                    const double num = __dmul_rn(__dadd_rn(__dmul_rn(__dmul_rn(2.0, ux), uy), k.c1),
                                                 __dadd_rn(__dmul_rn(2.0, cxy), k.c2));
                    const double den =
                        __dmul_rn(__dadd_rn(__dadd_rn(__dmul_rn(ux, ux), __dmul_rn(uy, uy)), k.c1),
                                  __dadd_rn(__dadd_rn(vx, vy), k.c2));
                    const float v = static_cast<float>(num / den);
                    if (map) map[static_cast<size_t>(y) * map_stride + q] = v;
                    if (in_cols && y >= roi.y0 && y < roi.y1) part += static_cast<double>(v);
                }
            }
        }
    }
    const double s = block_sum(part);
    if (threadIdx.x == 0) partial[static_cast<size_t>(blockIdx.y) * gridDim.x + blockIdx.x] = s;
}

// out[0] = sum of partial[0..n) in a fixed order (strided lanes, then lanes in order).
__global__ __launch_bounds__(kSumBlock) void ssim_sum_kernel(const double* partial, long long n, double* out) {
    __shared__ double lanes[kSumBlock];
    double s = 0.0;
    for (long long i = threadIdx.x; i < n; i += kSumBlock) s += partial[i];
    lanes[threadIdx.x] = s;
    __syncthreads();
    for (int w = kSumBlock / 2; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) lanes[threadIdx.x] += lanes[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = lanes[0];
}

template <class T>
hipError_t launch_ssim_t(const DevImage& a, const DevImage& b, const SsimTaps& k, float* map, size_t map_stride,
                         const SsimRoi& roi, double* partial, double* sum, hipStream_t s) {
    const long long gx = (static_cast<long long>(a.w) * a.cn + kBlock - 1) / kBlock;
    const long long gy = std::min<long long>((a.h + kSsimRows - 1) / kSsimRows, 4096);
    const dim3 grid(static_cast<unsigned>(gx), static_cast<unsigned>(gy));
    hipLaunchKernelGGL(ssim_map_kernel<T>, grid, dim3(kBlock), 0, s, a, b, k, map, map_stride, roi, partial);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(ssim_sum_kernel, dim3(1), dim3(kSumBlock), 0, s, partial, gx * gy, sum);
    return hipGetLastError();
}

}  // namespace

long long ssim_partials(const DevImage& a) {
    if (static_cast<long long>(a.w) * a.h == 0) return 0;
    const long long gx = (static_cast<long long>(a.w) * a.cn + kBlock - 1) / kBlock;
    return gx * std::min<long long>((a.h + kSsimRows - 1) / kSsimRows, 4096);
}

hipError_t launch_ssim(const DevImage& a, const DevImage& b, const SsimTaps& k, float* map, size_t map_stride,
                       const SsimRoi& roi, double* partial, double* sum, hipStream_t s) {
    if (a.dt != b.dt || a.cn != b.cn || a.w != b.w || a.h != b.h) return hipErrorInvalidValue;
    if (static_cast<long long>(a.w) * a.h == 0) return hipSuccess;
    // per-lane column offsets and sample indices are 32-bit
    if (static_cast<long long>(a.w) * a.cn * static_cast<long long>(dtype_size(a.dt)) > 0x7fffffffLL)
        return hipErrorInvalidValue;
    switch (a.dt) {
        case IPFG_U8: return launch_ssim_t<uint8_t>(a, b, k, map, map_stride, roi, partial, sum, s);
        case IPFG_U16: return launch_ssim_t<uint16_t>(a, b, k, map, map_stride, roi, partial, sum, s);
        case IPFG_I16: return launch_ssim_t<int16_t>(a, b, k, map, map_stride, roi, partial, sum, s);
        case IPFG_I32: return launch_ssim_t<int32_t>(a, b, k, map, map_stride, roi, partial, sum, s);
        case IPFG_F32: return launch_ssim_t<float>(a, b, k, map, map_stride, roi, partial, sum, s);
        case IPFG_F64: return launch_ssim_t<double>(a, b, k, map, map_stride, roi, partial, sum, s);
    }
    return hipErrorInvalidValue;
}

}  // namespace ipfg
