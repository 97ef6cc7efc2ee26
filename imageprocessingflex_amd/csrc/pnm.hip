// Netpbm / PFM raster unpacking for gfx950 (reference: include/ipf/codec/pnm.hpp:58-107,
// 143-176). The host parses the header and copies the raster bytes to the
// device untouched; one thread per output sample decodes its bits, bytes or
// words there: P4 bit rows, maxval rescaling (sample * max / maxval + 0.5 in
// FP64, truncated, after clamping to maxval), 16-bit big-endian words, PFM's
// bottom-up rows in either byte order.
#include "pnm.hpp"

namespace ipfg {
namespace {

template <class T>
__global__ __launch_bounds__(kBlock) void pnm_unpack_kernel(const uint8_t* __restrict__ raw, int layout, int maxval,
                                                            DevImage dst) {
    const size_t row = static_cast<size_t>(dst.w) * dst.cn;
    const size_t i = blockIdx.x * static_cast<size_t>(kBlock) + threadIdx.x;
    if (i >= row * dst.h) return;
    const size_t y = i / row, q = i - y * row;
    T* out = reinterpret_cast<T*>(static_cast<char*>(dst.data) + y * dst.stride) + q;
    if constexpr (std::is_same<T, float>::value) {
        const uint8_t* b = raw + ((static_cast<size_t>(dst.h) - 1 - y) * row + q) * 4;
        const unsigned u = layout == PFM_LE ? (b[0] | (b[1] << 8) | (b[2] << 16) | (static_cast<unsigned>(b[3]) << 24))
                                            : ((static_cast<unsigned>(b[0]) << 24) | (b[1] << 16) | (b[2] << 8) | b[3]);
        *out = __uint_as_float(u);
    } else {
        const double tmax = static_cast<double>(std::numeric_limits<T>::max());
        if (layout == PNM_BITS) {
            const size_t rb = (static_cast<size_t>(dst.w) + 7) / 8;
            const int bit = (raw[y * rb + (q >> 3)] >> (7 - (q & 7))) & 1;
            *out = bit ? T(0) : static_cast<T>(tmax);
            return;
        }
        long long v = layout == PNM_BYTES ? raw[i] : ((raw[2 * i] << 8) | raw[2 * i + 1]);
        if (v > maxval) v = maxval;
        *out = static_cast<T>((static_cast<double>(v) * tmax) / static_cast<double>(maxval) + 0.5);
    }
}

}  // namespace

hipError_t launch_pnm_unpack(const uint8_t* raw, int layout, int maxval, const DevImage& dst, hipStream_t s) {
    const size_t n = static_cast<size_t>(dst.w) * dst.h * dst.cn;
    if (n == 0) return hipSuccess;
    const dim3 grid(static_cast<unsigned>((n + kBlock - 1) / kBlock));
    if (layout == PFM_LE || layout == PFM_BE) {
        if (dst.dt != IPFG_F32) return hipErrorInvalidValue;
        hipLaunchKernelGGL(pnm_unpack_kernel<float>, grid, dim3(kBlock), 0, s, raw, layout, maxval, dst);
    } else if (dst.dt == IPFG_U8) {
        hipLaunchKernelGGL(pnm_unpack_kernel<uint8_t>, grid, dim3(kBlock), 0, s, raw, layout, maxval, dst);
    } else if (dst.dt == IPFG_U16) {
        hipLaunchKernelGGL(pnm_unpack_kernel<uint16_t>, grid, dim3(kBlock), 0, s, raw, layout, maxval, dst);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace ipfg
