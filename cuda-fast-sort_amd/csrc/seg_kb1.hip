// Instantiates the segmented sort for 1-byte keys (all value widths).
#include "rsort_seg_launch_impl.hpp"

namespace rsa {
hipError_t launch_seg_kb1(const SegJob& j, hipStream_t s) { return launch_seg<1>(j, s); }
hipError_t launch_req_kb1(const ReqJob& j, hipStream_t s) { return launch_req<1>(j, s); }
}  // namespace rsa
