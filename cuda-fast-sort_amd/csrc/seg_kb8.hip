// Instantiates the segmented sort for 8-byte keys (all value widths).
#include "rsort_seg_launch_impl.hpp"

namespace rsa {
hipError_t launch_seg_kb8(const SegJob& j, hipStream_t s) { return launch_seg<8>(j, s); }
hipError_t launch_req_kb8(const ReqJob& j, hipStream_t s) { return launch_req<8>(j, s); }
}  // namespace rsa
