// Instantiates the size-specialised planar separable kernels for certified
// uint8 -> float32 calls on 1-channel images with kw = kh in {3, 5, 7}: the
// scaled derivative filters of the feature callers (cornerHarris,
// cornerMinEigenVal) and small Gaussians.
#include "sep_kernels.hpp"

namespace ipfg {
namespace sepk {

hipError_t plane_u8f_k1(const SepArgs& a, hipStream_t s) {
    if (a.src.cn != 1 || a.dst.dt != IPFG_F32 || a.kw != a.kh) return hipErrorNotSupported;
    switch (a.kw) {
        case 3: return plane_launch<uint8_t, float, double, 1, true, 3>(a, s);
        case 5: return plane_launch<uint8_t, float, double, 1, true, 5>(a, s);
        case 7: return plane_launch<uint8_t, float, double, 1, true, 7>(a, s);
    }
    return hipErrorNotSupported;
}

}  // namespace sepk
}  // namespace ipfg
