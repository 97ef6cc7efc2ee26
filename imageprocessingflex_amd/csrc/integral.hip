// integralImage for gfx950 (reference: include/ipf/filter.hpp:522-538).
//
// The reference fills the (W + 1) x (H + 1) double table row by row: a
// running sum of the row (left to right) added to the entry above,
// I(x + 1, y + 1) = I(x + 1, y) + rowSum(x, y). Both running sums are
// sequential in FP64, so the bits of a float image's table depend on that
// order; integer images give exact integer sums in any order as long as they
// stay below 2^53.
//
//  integral_rows_kernel  one wave per (row, channel): the row's running
//                        sums (the chain carried through the lanes), stored
//                        in the table's own row y + 1 (the column pass adds
//                        the row above in place)
//  integral_cols_kernel  one thread per (column, channel): walks down the
//                        column adding each entry to the one below, in the
//                        reference's row order; consecutive threads take
//                        consecutive samples, so every row step is one
//                        contiguous access per wave
// Row 0 and column 0 are zero.
#include "kernels.hpp"

namespace ipfg {
namespace {

// Rows: one wave per (row, channel). Lanes load 64 consecutive samples and
// square them in parallel; the running sum then passes through the lanes in
// order (lane i's term broadcast with v_readlane, added, kept by lane i), the
// reference's sequential additions exactly, with coalesced loads and stores.
template <class T>
__global__ __launch_bounds__(kBlock) void integral_rows_kernel(DevImage src, double* __restrict__ I, size_t istride,
                                                               int squared) {
    const int lane = threadIdx.x & 63;
    const int t = blockIdx.x * (blockDim.x / 64) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int cn = src.cn, W = src.w;
    if (t >= src.h * cn) return;  // whole waves
    const int y = t / cn, c = t - y * cn;
    const T* __restrict__ s =
        reinterpret_cast<const T*>(static_cast<const char*>(src.data) + static_cast<size_t>(y) * src.stride) + c;
    double* __restrict__ row = I + static_cast<size_t>(y + 1) * istride + c;
    if (lane == 0) row[0] = 0.0;
    // the next chunk's samples are loaded before this chunk's chain runs
    // (lanes past the row load its last sample, unused)
    T v = s[static_cast<size_t>(min(lane, W - 1)) * cn];
    double sum = 0.0;
    for (int x0 = 0; x0 < W; x0 += 64) {
        const double d = static_cast<double>(v);
        const double term = squared ? d * d : d;
        if (x0 + 64 < W) v = s[static_cast<size_t>(min(x0 + 64 + lane, W - 1)) * cn];
        const int n = min(64, W - x0);
        double out = 0.0;
#pragma unroll 8
        for (int i = 0; i < n; ++i) {
            sum += bcast_d(term, i);
            if (lane == i) out = sum;
        }
        if (x0 + lane < W) row[static_cast<size_t>(x0 + lane + 1) * cn] = out;
    }
}

// Columns: one thread per table column, a dependent chain; two register
// buffers of kChunk rows alternate, so the next chunk's loads are in flight
// while the current one is summed.
constexpr int kChunk = 16;

__global__ __launch_bounds__(kBlock) void integral_cols_kernel(double* __restrict__ I, size_t istride, int W, int H,
                                                               int cn) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;  // sample index along a table row
    const int n = (W + 1) * cn;
    if (q >= n) return;
    double* __restrict__ col = I + q;
    col[0] = 0.0;  // row 0
    // Rows 1 .. H hold the row running sums; each becomes the entry above
    // plus itself, top to bottom. A chunk of rows is read before any row at
    // or after it is written (positions past H re-read row H and are unused).
    auto load = [&](double (&v)[kChunk], int y0) {
#pragma unroll
        for (int j = 0; j < kChunk; ++j) v[j] = col[static_cast<size_t>(min(y0 + j, H)) * istride];
    };
    double above = 0.0;
    auto add = [&](const double (&v)[kChunk], int y0) {
#pragma unroll
        for (int j = 0; j < kChunk; ++j) {
            if (y0 + j <= H) {
                above = above + v[j];  // I(x + 1, y) + rowSum: the entry above plus this row's running sum
                col[static_cast<size_t>(y0 + j) * istride] = above;
            }
        }
    };
    double a[kChunk], b[kChunk];
    load(a, 1);
    for (int y0 = 1; y0 <= H; y0 += 2 * kChunk) {
        load(b, y0 + kChunk);
        add(a, y0);
        load(a, y0 + 2 * kChunk);
        add(b, y0 + kChunk);
    }
}

}  // namespace

hipError_t launch_integral(const DevImage& src, double* I, size_t istride_samples, bool squared, hipStream_t s) {
    const int rows = src.h * src.cn;
    const dim3 rgrid(static_cast<unsigned>(rows));  // one (row, channel) per one-wave workgroup
    if (rows > 0 && src.w > 0)
        switch (src.dt) {
            case IPFG_U8:
                hipLaunchKernelGGL(integral_rows_kernel<uint8_t>, rgrid, dim3(64), 0, s,
                                   src, I, istride_samples, squared ? 1 : 0);
                break;
            case IPFG_U16:
                hipLaunchKernelGGL(integral_rows_kernel<uint16_t>, rgrid, dim3(64), 0,
                                   s, src, I, istride_samples, squared ? 1 : 0);
                break;
            case IPFG_I16:
                hipLaunchKernelGGL(integral_rows_kernel<int16_t>, rgrid, dim3(64), 0, s,
                                   src, I, istride_samples, squared ? 1 : 0);
                break;
            case IPFG_I32:
                hipLaunchKernelGGL(integral_rows_kernel<int32_t>, rgrid, dim3(64), 0, s,
                                   src, I, istride_samples, squared ? 1 : 0);
                break;
            case IPFG_F32:
                hipLaunchKernelGGL(integral_rows_kernel<float>, rgrid, dim3(64), 0, s,
                                   src, I, istride_samples, squared ? 1 : 0);
                break;
            case IPFG_F64:
                hipLaunchKernelGGL(integral_rows_kernel<double>, rgrid, dim3(64), 0, s,
                                   src, I, istride_samples, squared ? 1 : 0);
                break;
            default: return hipErrorInvalidValue;
        }
    const int n = (src.w + 1) * src.cn;
    // one wave per workgroup: the few column waves spread over as many CUs
    hipLaunchKernelGGL(integral_cols_kernel, dim3((n + 63) / 64), dim3(64), 0, s, I, istride_samples,
                       src.w, src.h, src.cn);
    return hipGetLastError();
}

}  // namespace ipfg
