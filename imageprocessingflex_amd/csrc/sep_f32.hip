// Instantiates the planar separable kernels for float32 images and outputs
// with run-time kernel sizes.
#include "sep_kernels.hpp"

namespace ipfg {
namespace sepk {

template <class Acc, bool CERT>
static hipError_t by_cn(const SepArgs& a, hipStream_t s) {
    switch (a.src.cn) {
        case 1: return plane_launch<float, float, Acc, 1, CERT, 0>(a, s);
        case 3: return plane_launch<float, float, Acc, 3, CERT, 0>(a, s);
        case 4: return plane_launch<float, float, Acc, 4, CERT, 0>(a, s);
    }
    return hipErrorInvalidValue;
}

hipError_t plane_f32(const SepArgs& a, bool cert, hipStream_t s) {
    return cert ? by_cn<double, true>(a, s) : by_cn<float, false>(a, s);
}

}  // namespace sepk
}  // namespace ipfg
