// Instantiates the one-channel float32 certified separable kernel for
// K = 27, 29, ..., 49 (gaussian windows for sigma up to 8).
#include "sep1.hpp"

namespace ipfg {
namespace sep1 {

hipError_t launch_large(const SepArgs& a, hipStream_t s) { return launch_range<27, 49>(a, s); }

}  // namespace sep1
}  // namespace ipfg
