// Instantiates the one-channel float32 certified separable kernel for
// K = 27, 29, ..., 37 (gaussian windows up to sigma 6).
#include "sep1.hpp"

namespace ipfg {
namespace sep1 {

hipError_t launch_k27_37(const SepArgs& a, hipStream_t s) { return launch_range<27, 37>(a, s); }

}  // namespace sep1
}  // namespace ipfg
