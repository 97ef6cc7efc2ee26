// Instantiates the planar separable kernels for uint8 images with run-time
// kernel sizes (uint8 -> uint8 and uint8 -> float32 outputs).
#include "sep_kernels.hpp"

namespace ipfg {
namespace sepk {

template <class To, class Acc, bool CERT>
static hipError_t by_cn(const SepArgs& a, hipStream_t s) {
    switch (a.src.cn) {
        case 1: return plane_launch<uint8_t, To, Acc, 1, CERT, 0>(a, s);
        case 3: return plane_launch<uint8_t, To, Acc, 3, CERT, 0>(a, s);
        case 4: return plane_launch<uint8_t, To, Acc, 4, CERT, 0>(a, s);
    }
    return hipErrorInvalidValue;
}

hipError_t plane_u8(const SepArgs& a, bool cert, hipStream_t s) {
    if (a.dst.dt == IPFG_U8) return cert ? by_cn<uint8_t, double, true>(a, s) : by_cn<uint8_t, float, false>(a, s);
    return cert ? by_cn<float, double, true>(a, s) : by_cn<float, float, false>(a, s);
}

}  // namespace sepk
}  // namespace ipfg
