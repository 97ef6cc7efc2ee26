// Instantiates the interleaved separable kernel (any channel count, output
// type chosen at run time) for uint8, uint16, int16 and float32 images.
#include "sep_kernels.hpp"

namespace ipfg {
namespace sepk {

template <class Ti>
static hipError_t by_mode(const SepArgs& a, bool cert, hipStream_t s) {
    return cert ? sep_fast_launch<Ti, double, 0, true>(a, s) : sep_fast_launch<Ti, float, 0, false>(a, s);
}

hipError_t generic_any(const SepArgs& a, bool cert, hipStream_t s) {
    switch (a.src.dt) {
        case IPFG_U8: return by_mode<uint8_t>(a, cert, s);
        case IPFG_U16: return by_mode<uint16_t>(a, cert, s);
        case IPFG_I16: return by_mode<int16_t>(a, cert, s);
        case IPFG_F32: return by_mode<float>(a, cert, s);
    }
    return hipErrorInvalidValue;
}

}  // namespace sepk
}  // namespace ipfg
