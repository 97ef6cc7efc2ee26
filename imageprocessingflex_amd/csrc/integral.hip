// integralImage for gfx950 (reference: include/ipf/filter.hpp:522-538).
//
// The reference fills the (W + 1) x (H + 1) double table row by row: a
// running sum of the row (left to right) added to the entry above,
// I(x + 1, y + 1) = I(x + 1, y) + rowSum(x, y). Both running sums are
// sequential in FP64, so the bits of a float image's table depend on that
// order; integer images give exact integer sums in any order as long as they
// stay below 2^53.
//
//  integral_rows_kernel  one thread per (row, channel): the row's running
//                        sums, stored in the table's own row y + 1 (the
//                        column pass adds the row above in place)
//  integral_cols_kernel  one thread per (column, channel): walks down the
//                        column adding each entry to the one below, in the
//                        reference's row order; consecutive threads take
//                        consecutive samples, so every row step is one
//                        contiguous access per wave
// Row 0 and column 0 are zero.
#include "kernels.hpp"

namespace ipfg {
namespace {

template <class T>
__global__ __launch_bounds__(kBlock) void integral_rows_kernel(DevImage src, double* __restrict__ I, size_t istride,
                                                               int squared) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    const int cn = src.cn;
    if (t >= src.h * cn) return;
    const int y = t / cn, c = t - y * cn;
    const T* s = reinterpret_cast<const T*>(static_cast<const char*>(src.data) + static_cast<size_t>(y) * src.stride);
    double* row = I + static_cast<size_t>(y + 1) * istride;
    row[c] = 0.0;
    double sum = 0.0;
    for (int x = 0; x < src.w; ++x) {
        const double v = static_cast<double>(s[static_cast<size_t>(x) * cn + c]);
        sum += squared ? v * v : v;
        row[static_cast<size_t>(x + 1) * cn + c] = sum;
    }
}

__global__ __launch_bounds__(kBlock) void integral_cols_kernel(double* __restrict__ I, size_t istride, int W, int H,
                                                               int cn) {
    const int q = blockIdx.x * kBlock + threadIdx.x;  // sample index along a table row
    const int n = (W + 1) * cn;
    if (q >= n) return;
    I[q] = 0.0;  // row 0
    double above = 0.0;
    for (int y = 1; y <= H; ++y) {
        double* p = I + static_cast<size_t>(y) * istride + q;
        above = above + *p;  // I(x + 1, y) + rowSum: the entry above plus this row's running sum
        *p = above;
    }
}

}  // namespace

hipError_t launch_integral(const DevImage& src, double* I, size_t istride_samples, bool squared, hipStream_t s) {
    const int rows = src.h * src.cn;
    if (rows > 0 && src.w > 0)
        switch (src.dt) {
            case IPFG_U8:
                hipLaunchKernelGGL(integral_rows_kernel<uint8_t>, dim3((rows + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                                   src, I, istride_samples, squared ? 1 : 0);
                break;
            case IPFG_U16:
                hipLaunchKernelGGL(integral_rows_kernel<uint16_t>, dim3((rows + kBlock - 1) / kBlock), dim3(kBlock), 0,
                                   s, src, I, istride_samples, squared ? 1 : 0);
                break;
            case IPFG_I16:
                hipLaunchKernelGGL(integral_rows_kernel<int16_t>, dim3((rows + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                                   src, I, istride_samples, squared ? 1 : 0);
                break;
            case IPFG_I32:
                hipLaunchKernelGGL(integral_rows_kernel<int32_t>, dim3((rows + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                                   src, I, istride_samples, squared ? 1 : 0);
                break;
            case IPFG_F32:
                hipLaunchKernelGGL(integral_rows_kernel<float>, dim3((rows + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                                   src, I, istride_samples, squared ? 1 : 0);
                break;
            case IPFG_F64:
                hipLaunchKernelGGL(integral_rows_kernel<double>, dim3((rows + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                                   src, I, istride_samples, squared ? 1 : 0);
                break;
            default: return hipErrorInvalidValue;
        }
    const int n = (src.w + 1) * src.cn;
    hipLaunchKernelGGL(integral_cols_kernel, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, I, istride_samples,
                       src.w, src.h, src.cn);
    return hipGetLastError();
}

}  // namespace ipfg
