// Bandwidth ceiling of the device for the sort's access pattern: streaming
// kernels that only read, only write, or copy, in 16-byte units per lane.  A
// measurement helper for the benchmark (each pass of the sort reads and
// writes every byte once, so a copy of the same bytes is the practical
// ceiling of a pass).  No reference counterpart.
//
// Two access forms, each with plain or non-temporal loads and stores:
//   strided  every lane strides over the buffer by the grid's width, four
//            units in flight;
//   chunked  each workgroup takes contiguous chunks of 256 x 8 units (32 KiB),
//            chunk c = workgroup + k * grid, eight units in flight per lane.
// Measured on MI355X (copy of 2 GiB): chunked with non-temporal accesses
// 5.4-5.5 TB/s, chunked plain 5.2-5.3, strided 4.2-4.7 (and a one-unit
// strided copy up to 5.5 at 4 workgroups per CU).
//
// A third form, the scatter copy, has the write pattern of one LSD pass: the
// buffer is tiles of 16384 4-byte keys (64 KiB), read in order, and tile t is
// written as 256 runs of 64 keys, run d at position t * 64 of bucket d's
// region (256 regions of tiles * 64 keys, one after another).  It is the pass
// of a sort whose 256 digits are equally frequent in every tile, an exact
// permutation with no ranking: the ceiling of a pass's own access pattern.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/rsort_c.h"

namespace rsa {
namespace {

typedef unsigned int v4u __attribute__((ext_vector_type(4)));

template <bool NT>
__device__ __forceinline__ uint4 ld16(const uint4* p) {
    if constexpr (NT) {
        const v4u v = __builtin_nontemporal_load((const v4u*)p);
        return make_uint4(v.x, v.y, v.z, v.w);
    } else {
        return *p;
    }
}

template <bool NT>
__device__ __forceinline__ void st16(uint4* p, const uint4& q) {
    if constexpr (NT) {
        const v4u v = {q.x, q.y, q.z, q.w};
        __builtin_nontemporal_store(v, (v4u*)p);
    } else {
        *p = q;
    }
}

// OP 0: read (the xor of everything read is stored only if it hits an
// unlikely pattern, so the loads cannot be dropped); 1: write a constant;
// 2: copy.  Unit i of the buffer is touched by exactly one lane.
template <int OP, bool NT, bool CHUNKED>
__global__ __launch_bounds__(256) void k_stream(uint4* __restrict__ dst, const uint4* __restrict__ src,
                                                uint64_t units) {
    constexpr int U = CHUNKED ? 8 : 4;
    uint4 acc = make_uint4(0u, 0u, 0u, 0u);
    const uint4 fill = make_uint4(0x01234567u, 0x89ABCDEFu, 0xFEDCBA98u, 0x76543210u);
    auto body = [&](uint64_t i0, uint64_t step) {
        uint4 q[U];
        if constexpr (OP != 1) {
#pragma unroll
            for (int u = 0; u < U; ++u) q[u] = ld16<NT>(src + i0 + (uint64_t)u * step);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if constexpr (OP == 0) {
                acc.x ^= q[u].x;
                acc.y ^= q[u].y;
                acc.z ^= q[u].z;
                acc.w ^= q[u].w;
            } else {
                st16<NT>(dst + i0 + (uint64_t)u * step, OP == 1 ? fill : q[u]);
            }
        }
    };
    uint64_t tail0;
    if constexpr (CHUNKED) {
        const uint64_t chunk = 256u * U;
        const uint64_t nch = units / chunk;
        for (uint64_t c = blockIdx.x; c < nch; c += gridDim.x) body(c * chunk + threadIdx.x, 256u);
        tail0 = nch * chunk;
    } else {
        const uint64_t stride = (uint64_t)gridDim.x * 256u;
        uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x;
        for (; i + (U - 1) * stride < units; i += U * stride) body(i, stride);
        tail0 = i - ((uint64_t)blockIdx.x * 256u + threadIdx.x);  // this lane's units from the round it stopped at
    }
    for (uint64_t i = tail0 + (uint64_t)blockIdx.x * 256u + threadIdx.x; i < units; i += (uint64_t)gridDim.x * 256u) {
        if constexpr (OP == 0) {
            const uint4 q = src[i];
            acc.x ^= q.x;
            acc.y ^= q.y;
            acc.z ^= q.z;
            acc.w ^= q.w;
        } else {
            dst[i] = OP == 1 ? fill : src[i];
        }
    }
    if constexpr (OP == 0) {
        if ((acc.x ^ acc.y ^ acc.z ^ acc.w) == 0x9E3779B9u && acc.x == 0x7F4A7C15u) dst[threadIdx.x] = acc;
    }
}

// Scatter copy: one 256-thread workgroup per tile (grid-stride over tiles),
// 16 units of 16 bytes per lane, all loads in flight before the stores.
// Sixteen consecutive lanes write one run (256 contiguous bytes).
template <bool NT>
__global__ __launch_bounds__(256) void k_scatter(uint4* __restrict__ dst, const uint4* __restrict__ src,
                                                 uint64_t tiles) {
    constexpr int U = 16;
    const uint64_t region = tiles * 16u;  // 16-byte units per bucket region
    for (uint64_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        const uint4* s = src + t * 4096u;
        uint4 q[U];
#pragma unroll
        for (int k = 0; k < U; ++k) q[k] = ld16<NT>(s + k * 256 + threadIdx.x);
#pragma unroll
        for (int k = 0; k < U; ++k) {
            const uint32_t u = (uint32_t)k * 256u + threadIdx.x;
            st16<NT>(dst + (uint64_t)(u >> 4) * region + t * 16u + (u & 15u), q[k]);
        }
    }
}

template <int OP>
void launch_op(uint32_t form, dim3 grid, hipStream_t s, uint4* dst, const uint4* src, uint64_t units) {
    switch (form) {
        case 0: hipLaunchKernelGGL((k_stream<OP, false, false>), grid, dim3(256), 0, s, dst, src, units); break;
        case 1: hipLaunchKernelGGL((k_stream<OP, true, false>), grid, dim3(256), 0, s, dst, src, units); break;
        case 2: hipLaunchKernelGGL((k_stream<OP, false, true>), grid, dim3(256), 0, s, dst, src, units); break;
        default: hipLaunchKernelGGL((k_stream<OP, true, true>), grid, dim3(256), 0, s, dst, src, units); break;
    }
}

}  // namespace
}  // namespace rsa

extern "C" rsort_status rsort_stream_bandwidth(uint32_t mode, void* dst, const void* src, uint64_t bytes,
                                               uint32_t blocks_per_cu, void* stream) {
    const uint32_t op = mode & 3u, form = (mode >> 2) & 3u;
    if ((mode & ~15u) || (bytes & 15u) || (((uintptr_t)dst | (uintptr_t)src) & 15u) || blocks_per_cu < 1 ||
        blocks_per_cu > 32)
        return RSORT_ERROR_INVALID_ARGUMENT;
    if (op == 3 && ((bytes & 65535u) || (form & 2u) || dst == src)) return RSORT_ERROR_INVALID_ARGUMENT;
    if (!dst || (op != 1 && !src)) return RSORT_ERROR_INVALID_ARGUMENT;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
        return RSORT_ERROR_DEVICE;
    const dim3 grid((uint32_t)cus * blocks_per_cu);
    const uint64_t units = bytes / 16u;
    hipStream_t s = (hipStream_t)stream;
    if (op == 0) rsa::launch_op<0>(form, grid, s, (uint4*)dst, (const uint4*)src, units);
    else if (op == 1) rsa::launch_op<1>(form, grid, s, (uint4*)dst, (const uint4*)src, units);
    else if (op == 2) rsa::launch_op<2>(form, grid, s, (uint4*)dst, (const uint4*)src, units);
    else if (form & 1u)
        hipLaunchKernelGGL((rsa::k_scatter<true>), grid, dim3(256), 0, s, (uint4*)dst, (const uint4*)src, bytes >> 16);
    else
        hipLaunchKernelGGL((rsa::k_scatter<false>), grid, dim3(256), 0, s, (uint4*)dst, (const uint4*)src, bytes >> 16);
    return hipGetLastError() == hipSuccess ? RSORT_SUCCESS : RSORT_ERROR_DEVICE;
}
