// Instantiates the one-channel float32 certified separable kernel for
// K = 3, 5, ..., 13 (gaussian windows up to sigma 2; box windows).
#include "sep1.hpp"

namespace ipfg {
namespace sep1 {

hipError_t launch_k3_13(const SepArgs& a, hipStream_t s) { return launch_range<3, 13>(a, s); }

}  // namespace sep1
}  // namespace ipfg
