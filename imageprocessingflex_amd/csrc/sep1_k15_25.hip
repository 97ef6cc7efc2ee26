// Instantiates the one-channel float32 certified separable kernel for
// K = 15, 17, ..., 25 (gaussian windows up to sigma 4; box windows).
#include "sep1.hpp"

namespace ipfg {
namespace sep1 {

hipError_t launch_k15_25(const SepArgs& a, hipStream_t s) { return launch_range<15, 25>(a, s); }

}  // namespace sep1
}  // namespace ipfg
