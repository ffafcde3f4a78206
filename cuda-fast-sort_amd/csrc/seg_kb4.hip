// Instantiates the segmented sort for 4-byte keys (all value widths).
#include "rsort_seg_launch_impl.hpp"

namespace rsa {
hipError_t launch_seg_kb4(const SegJob& j, hipStream_t s) { return launch_seg<4>(j, s); }
hipError_t launch_req_kb4(const ReqJob& j, hipStream_t s) { return launch_req<4>(j, s); }
}  // namespace rsa
