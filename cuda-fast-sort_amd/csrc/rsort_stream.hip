// Bandwidth ceiling of the device for the sort's access pattern: streaming
// kernels that only read, only write, or copy, in 16-byte units per lane, a
// grid of `blocks_per_cu` 256-thread workgroups per CU striding over the
// buffer with four units in flight per lane.  A measurement helper for the
// benchmark (the sort's passes read and write every byte once, so a copy of
// the same bytes is the practical ceiling of a pass).  No reference
// counterpart.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/rsort_c.h"

namespace rsa {
namespace {

constexpr int kUnroll = 4;

// MODE 0: read (an xor of everything read is stored by lanes whose value
// matches an unlikely pattern, so the loads cannot be dropped); 1: write a
// constant; 2: copy.
template <int MODE>
__global__ __launch_bounds__(256) void k_stream(uint4* __restrict__ dst, const uint4* __restrict__ src,
                                                uint64_t units) {
    const uint64_t stride = (uint64_t)gridDim.x * 256u;
    uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    uint4 acc = make_uint4(0u, 0u, 0u, 0u);
    const uint4 fill = make_uint4(0x01234567u, 0x89ABCDEFu, 0xFEDCBA98u, 0x76543210u);
    for (; i + (kUnroll - 1) * stride < units; i += kUnroll * stride) {
        uint4 q[kUnroll];
        if constexpr (MODE != 1) {
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) q[u] = src[i + u * stride];
        }
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            if constexpr (MODE == 0) {
                acc.x ^= q[u].x;
                acc.y ^= q[u].y;
                acc.z ^= q[u].z;
                acc.w ^= q[u].w;
            } else if constexpr (MODE == 1) {
                dst[i + u * stride] = fill;
            } else {
                dst[i + u * stride] = q[u];
            }
        }
    }
    for (; i < units; i += stride) {
        if constexpr (MODE == 0) {
            const uint4 q = src[i];
            acc.x ^= q.x;
            acc.y ^= q.y;
            acc.z ^= q.z;
            acc.w ^= q.w;
        } else if constexpr (MODE == 1) {
            dst[i] = fill;
        } else {
            dst[i] = src[i];
        }
    }
    if constexpr (MODE == 0) {
        if ((acc.x ^ acc.y ^ acc.z ^ acc.w) == 0x9E3779B9u && acc.x == 0x7F4A7C15u) dst[threadIdx.x] = acc;
    }
}

}  // namespace
}  // namespace rsa

extern "C" rsort_status rsort_stream_bandwidth(uint32_t mode, void* dst, const void* src, uint64_t bytes,
                                               uint32_t blocks_per_cu, void* stream) {
    if (mode > 2 || (bytes & 15u) || (((uintptr_t)dst | (uintptr_t)src) & 15u) || blocks_per_cu < 1 ||
        blocks_per_cu > 32)
        return RSORT_ERROR_INVALID_ARGUMENT;
    if ((mode != 0 && !dst) || (mode != 1 && !src) || (mode == 0 && !dst)) return RSORT_ERROR_INVALID_ARGUMENT;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
        return RSORT_ERROR_DEVICE;
    const dim3 grid((uint32_t)cus * blocks_per_cu), block(256);
    const uint64_t units = bytes / 16u;
    hipStream_t s = (hipStream_t)stream;
    if (mode == 0)
        hipLaunchKernelGGL(rsa::k_stream<0>, grid, block, 0, s, (uint4*)dst, (const uint4*)src, units);
    else if (mode == 1)
        hipLaunchKernelGGL(rsa::k_stream<1>, grid, block, 0, s, (uint4*)dst, (const uint4*)src, units);
    else
        hipLaunchKernelGGL(rsa::k_stream<2>, grid, block, 0, s, (uint4*)dst, (const uint4*)src, units);
    return hipGetLastError() == hipSuccess ? RSORT_SUCCESS : RSORT_ERROR_DEVICE;
}
