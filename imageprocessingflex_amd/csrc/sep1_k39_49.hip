// Instantiates the one-channel float32 certified separable kernel for
// K = 39, 41, ..., 49 (gaussian windows up to sigma 8).
#include "sep1.hpp"

namespace ipfg {
namespace sep1 {

hipError_t launch_k39_49(const SepArgs& a, hipStream_t s) { return launch_range<39, 49>(a, s); }

}  // namespace sep1
}  // namespace ipfg
