// Instantiates the segmented sort for 2-byte keys (all value widths).
#include "rsort_seg_launch_impl.hpp"

namespace rsa {
hipError_t launch_seg_kb2(const SegJob& j, hipStream_t s) { return launch_seg<2>(j, s); }
hipError_t launch_req_kb2(const ReqJob& j, hipStream_t s) { return launch_req<2>(j, s); }
}  // namespace rsa
