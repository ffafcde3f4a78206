// Instantiates the segmented sort for 16-byte keys (all value widths).
#include "rsort_seg_launch_impl.hpp"

namespace rsa {
hipError_t launch_seg_kb16(const SegJob& j, hipStream_t s) { return launch_seg<16>(j, s); }
hipError_t launch_req_kb16(const ReqJob& j, hipStream_t s) { return launch_req<16>(j, s); }
}  // namespace rsa
