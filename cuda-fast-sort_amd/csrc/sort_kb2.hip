// Instantiates the onesweep op sequence for 2-byte keys (all value widths).
#include "rsort_launch_impl.hpp"

namespace rsa {
hipError_t launch_core_kb2(const SortJob& j, hipStream_t s) { return launch_core<2>(j, s); }
}  // namespace rsa
