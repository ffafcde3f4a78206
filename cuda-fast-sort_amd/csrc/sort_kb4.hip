// Instantiates the onesweep op sequence for 4-byte keys (all value widths).
#include "rsort_launch_impl.hpp"

namespace rsa {
hipError_t launch_core_kb4(const SortJob& j, hipStream_t s) { return launch_core<4>(j, s); }
}  // namespace rsa
