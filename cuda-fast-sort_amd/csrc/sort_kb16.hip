// Instantiates the onesweep op sequence for 16-byte keys (all value widths).
#include "rsort_launch_impl.hpp"

namespace rsa {
hipError_t launch_core_kb16(const SortJob& j, hipStream_t s) { return launch_core<16>(j, s); }
}  // namespace rsa
