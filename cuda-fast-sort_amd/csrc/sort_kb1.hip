// Instantiates the onesweep op sequence for 1-byte keys (all value widths).
#include "rsort_launch_impl.hpp"

namespace rsa {
hipError_t launch_core_kb1(const SortJob& j, hipStream_t s) { return launch_core<1>(j, s); }
}  // namespace rsa
