// Instantiates the size-specialised planar separable kernels for certified
// float32 -> float32 calls on one-channel images (kw = kh, odd, 3..49).
#include "sep_kernels.hpp"

namespace ipfg {
namespace sepk {

hipError_t plane_f32_k1(const SepArgs& a, hipStream_t s) {
    if (a.src.cn != 1 || a.dst.dt != IPFG_F32 || a.kw != a.kh) return hipErrorNotSupported;
    return plane_launch_k<float, float, double, 1, true>(a, s);
}

}  // namespace sepk
}  // namespace ipfg
