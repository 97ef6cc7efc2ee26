// Instantiates the size-specialised planar separable kernels for certified
// uint8 -> uint8 calls on 4-channel images (kw = kh, odd, 3..49).
#include "sep_kernels.hpp"

namespace ipfg {
namespace sepk {

hipError_t plane_u8_k4(const SepArgs& a, hipStream_t s) {
    if (a.src.cn != 4 || a.dst.dt != IPFG_U8 || a.kw != a.kh) return hipErrorNotSupported;
    return plane_launch_k<uint8_t, uint8_t, double, 4, true>(a, s);
}

}  // namespace sepk
}  // namespace ipfg
