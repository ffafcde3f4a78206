// Device-side input generator and result checks, for sizes where a CPU check
// is too slow and for runs on the reference's own benchmark inputs.
//
// Reference counterparts (restated, same semantics):
//   mix32 / set_seed_radix / init_keys   include/rsort/detail/utils.cuh:227-317, 435-437
//   check_sorted                         include/rsort/detail/utils.cuh:341-365
//   check_pairings                       include/rsort/detail/utils.cuh:377-432
//   per-pass digit histogram             include/rsort/detail/histogram.cuh:88-211
// The checks count violations into a caller-zeroed device counter instead of
// returning a flag, so one launch covers any n and the caller reads one word.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/rsort_c.h"
#include "rsort_kernels.hpp"

namespace rsa {

namespace {

// Hash of a 32-bit word (the reference's mix32).
__device__ __forceinline__ uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// init_keys: mode 1 random, 2 byte_skip (every other byte fixed to 0x55 / 0x00
// so half the passes see one digit), 3 ascending index.
template <int KB>
__global__ void k_init_keys(typename KeyTraits<KB>::Mem* __restrict__ out, uint64_t n, uint32_t mode, uint32_t seed) {
    using Reg = typename KeyTraits<KB>::Reg;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        Reg x;
        if constexpr (KB <= 4) {
            uint32_t v = mode == 3 ? (uint32_t)i : mix32((uint32_t)(i ^ (uint64_t)seed));
            if (mode == 2) v = (v & 0xFF00FF00u) | 0x00550055u;
            if constexpr (KB < 4) v &= (1u << (8 * KB)) - 1u;
            x = (Reg)v;
        } else if constexpr (KB == 8) {
            uint64_t v;
            if (mode == 3) {
                v = i;
            } else {
                v = ((uint64_t)(mix32((uint32_t)i) ^ seed) << 32) | (uint64_t)(mix32((uint32_t)(i + 0x9e3779b9u)) ^ seed);
                if (mode == 2) v = (v & 0xFF00FF00FF00FF00ull) | 0x0055005500550055ull;
            }
            x = v;
        } else {
            U128 v;
            if (mode == 3) {
                v.lo = i;
                v.hi = 0;
            } else {
                v.hi = ((uint64_t)(mix32((uint32_t)i) ^ seed) << 32) | (uint64_t)(mix32((uint32_t)(i + 0x9e3779b9u)) ^ seed);
                v.lo = ((uint64_t)(mix32((uint32_t)(i + 0x3c6ef372u)) ^ seed) << 32) |
                       (uint64_t)(mix32((uint32_t)(i + 0xdaa66d2bu)) ^ seed);
                if (mode == 2) {
                    v.hi = (v.hi & 0xFF00FF00FF00FF00ull) | 0x0055005500550055ull;
                    v.lo = (v.lo & 0xFF00FF00FF00FF00ull) | 0x0055005500550055ull;
                }
            }
            x = v;
        }
        out[i] = key_to_mem<KB>(x);
    }
}

// a precedes b in the sort order: strict "a > b" (ascending) on the key values;
// float and double compare as floating point (a NaN never violates), the
// others (integers, half, long double) on their order-preserving words.
template <int KB>
__device__ __forceinline__ bool out_of_order(typename KeyTraits<KB>::Reg a, typename KeyTraits<KB>::Reg b,
                                             uint32_t kind, bool desc, const TwMasks& tw) {
    if constexpr (KB == 4) {
        if (kind == kKindFloat) {
            const float fa = __uint_as_float(a), fb = __uint_as_float(b);
            return desc ? (fa < fb) : (fa > fb);
        }
    }
    if constexpr (KB == 8) {
        if (kind == kKindFloat) {
            const double fa = __longlong_as_double((long long)a), fb = __longlong_as_double((long long)b);
            return desc ? (fa < fb) : (fa > fb);
        }
    }
    const auto ta = tw_in<KB>(a, tw), tb = tw_in<KB>(b, tw);  // tw carries no descending complement
    if constexpr (KB == 16) {
        const bool gt = ta.hi != tb.hi ? ta.hi > tb.hi : ta.lo > tb.lo;
        const bool lt = ta.hi != tb.hi ? ta.hi < tb.hi : ta.lo < tb.lo;
        return desc ? lt : gt;
    } else {
        return desc ? (ta < tb) : (ta > tb);
    }
}

__device__ __forceinline__ void count_violations(bool bad, unsigned long long* counter) {
    const uint64_t m = __ballot(bad);
    if (m && (threadIdx.x & 63u) == 0u) atomicAdd(counter, (unsigned long long)__builtin_popcountll(m));
}

template <int KB>
__global__ void k_check_sorted(const typename KeyTraits<KB>::Mem* __restrict__ keys, uint64_t n, uint32_t kind,
                               uint32_t desc, TwMasks tw, unsigned long long* bad) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i0 = (uint64_t)blockIdx.x * blockDim.x; i0 + 1 < n; i0 += stride) {
        const uint64_t i = i0 + threadIdx.x;
        bool v = false;
        if (i + 1 < n) v = out_of_order<KB>(key_to_reg<KB>(keys[i]), key_to_reg<KB>(keys[i + 1]), kind, desc != 0, tw);
        count_violations(v, bad);
    }
}

template <int KB, int VB>
__global__ void k_check_pairings(const typename KeyTraits<KB>::Mem* __restrict__ ko,
                                 const typename KeyTraits<KB>::Mem* __restrict__ ks, const void* __restrict__ vs,
                                 uint64_t n, unsigned long long* bad) {
    using V = std::conditional_t<VB == 8, uint64_t, uint32_t>;
    const V* vals = (const V*)vs;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i0 = (uint64_t)blockIdx.x * blockDim.x; i0 < n; i0 += stride) {
        const uint64_t i = i0 + threadIdx.x;
        bool v = false;
        if (i < n) {
            const uint64_t x = vals[i];
            const auto k = key_to_reg<KB>(ks[i]);
            if (x >= n) {
                v = true;
            } else {
                const auto o = key_to_reg<KB>(ko[x]);
                if constexpr (KB == 16) {
                    v = o.lo != k.lo || o.hi != k.hi;
                } else {
                    v = o != k;
                }
            }
            if (!v && i + 1 < n) {
                const auto k2 = key_to_reg<KB>(ks[i + 1]);
                bool eq;
                if constexpr (KB == 16) {
                    eq = k2.lo == k.lo && k2.hi == k.hi;
                } else {
                    eq = k2 == k;
                }
                v = eq && x > (uint64_t)vals[i + 1];  // equal keys keep their input order
            }
        }
        count_violations(v, bad);
    }
}

// Per-pass digit counts of the twiddled keys: out[p*256 + d].
template <int KB>
__global__ __launch_bounds__(256) void k_digit_counts(const typename KeyTraits<KB>::Mem* __restrict__ keys,
                                                      uint64_t n, TwMasks tw, unsigned long long* out) {
    __shared__ uint32_t H[KB * 256];
    for (uint32_t k = threadIdx.x; k < KB * 256u; k += 256u) H[k] = 0u;
    __syncthreads();
    for (uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256u) {
        const auto t = tw_in<KB>(key_to_reg<KB>(keys[i]), tw);
#pragma unroll
        for (int p = 0; p < KB; ++p) atomicAdd(&H[p * 256 + key_digit(t, 8u * (p & 3), (uint32_t)p >> 2)], 1u);
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < KB * 256u; k += 256u)
        if (H[k]) atomicAdd(out + k, (unsigned long long)H[k]);
}

uint32_t grid_of(uint64_t n, uint32_t per) {
    uint64_t g = (n + per - 1) / per;
    if (g > 8192) g = 8192;
    return g < 1 ? 1u : (uint32_t)g;
}

bool key_width_ok(uint32_t kb) { return kb == 1 || kb == 2 || kb == 4 || kb == 8 || kb == 16; }

}  // namespace

}  // namespace rsa

using namespace rsa;

extern "C" {

rsort_status rsort_init_keys(uint32_t key_bytes, uint32_t mode, uint32_t seed, void* keys, uint64_t n, void* stream) {
    if (!key_width_ok(key_bytes) || mode < 1 || mode > 3) return RSORT_ERROR_INVALID_ARGUMENT;
    if (n == 0) return RSORT_SUCCESS;
    if (!keys || (uintptr_t)keys % key_bytes) return RSORT_ERROR_INVALID_ARGUMENT;
    hipStream_t s = (hipStream_t)stream;
    const dim3 g(grid_of(n, 256)), b(256);
    switch (key_bytes) {
        case 1: hipLaunchKernelGGL(k_init_keys<1>, g, b, 0, s, (uint8_t*)keys, n, mode, seed); break;
        case 2: hipLaunchKernelGGL(k_init_keys<2>, g, b, 0, s, (uint16_t*)keys, n, mode, seed); break;
        case 4: hipLaunchKernelGGL(k_init_keys<4>, g, b, 0, s, (uint32_t*)keys, n, mode, seed); break;
        case 8: hipLaunchKernelGGL(k_init_keys<8>, g, b, 0, s, (uint64_t*)keys, n, mode, seed); break;
        default: hipLaunchKernelGGL(k_init_keys<16>, g, b, 0, s, (U128*)keys, n, mode, seed); break;
    }
    return hipGetLastError() == hipSuccess ? RSORT_SUCCESS : RSORT_ERROR_DEVICE;
}

uint32_t rsort_set_seed_radix(int32_t index) { return (uint32_t)((index + 81) * 17); }

rsort_status rsort_check_sorted(const rsort_desc* desc, const void* keys, uint64_t n,
                                unsigned long long* d_violations, void* stream) {
    if (!desc || !key_width_ok(desc->key_bytes) || desc->key_kind > RSORT_KEY_FLOAT || !d_violations)
        return RSORT_ERROR_INVALID_ARGUMENT;
    if (n < 2) return RSORT_SUCCESS;
    if (!keys) return RSORT_ERROR_INVALID_ARGUMENT;
    const TwMasks tw = make_tw_masks(desc->key_kind, desc->key_bytes, false);
    hipStream_t s = (hipStream_t)stream;
    const dim3 g(grid_of(n, 256)), b(256);
    const uint32_t k = desc->key_kind, d = desc->descending ? 1u : 0u;
    switch (desc->key_bytes) {
        case 1: hipLaunchKernelGGL(k_check_sorted<1>, g, b, 0, s, (const uint8_t*)keys, n, k, d, tw, d_violations); break;
        case 2: hipLaunchKernelGGL(k_check_sorted<2>, g, b, 0, s, (const uint16_t*)keys, n, k, d, tw, d_violations); break;
        case 4: hipLaunchKernelGGL(k_check_sorted<4>, g, b, 0, s, (const uint32_t*)keys, n, k, d, tw, d_violations); break;
        case 8: hipLaunchKernelGGL(k_check_sorted<8>, g, b, 0, s, (const uint64_t*)keys, n, k, d, tw, d_violations); break;
        default: hipLaunchKernelGGL(k_check_sorted<16>, g, b, 0, s, (const U128*)keys, n, k, d, tw, d_violations); break;
    }
    return hipGetLastError() == hipSuccess ? RSORT_SUCCESS : RSORT_ERROR_DEVICE;
}

rsort_status rsort_check_pairings(uint32_t key_bytes, uint32_t value_bytes, const void* keys_orig,
                                  const void* keys_sorted, const void* vals_sorted, uint64_t n,
                                  unsigned long long* d_violations, void* stream) {
    if (!key_width_ok(key_bytes) || (value_bytes != 4 && value_bytes != 8) || !d_violations)
        return RSORT_ERROR_INVALID_ARGUMENT;
    if (n == 0) return RSORT_SUCCESS;
    if (!keys_orig || !keys_sorted || !vals_sorted) return RSORT_ERROR_INVALID_ARGUMENT;
    hipStream_t s = (hipStream_t)stream;
    const dim3 g(grid_of(n, 256)), b(256);
#define RSA_PAIR(KB, KT)                                                                                     \
    do {                                                                                                      \
        if (value_bytes == 8)                                                                                 \
            hipLaunchKernelGGL((k_check_pairings<KB, 8>), g, b, 0, s, (const KT*)keys_orig,                   \
                               (const KT*)keys_sorted, vals_sorted, n, d_violations);                        \
        else                                                                                                  \
            hipLaunchKernelGGL((k_check_pairings<KB, 4>), g, b, 0, s, (const KT*)keys_orig,                   \
                               (const KT*)keys_sorted, vals_sorted, n, d_violations);                        \
    } while (0)
    switch (key_bytes) {
        case 1: RSA_PAIR(1, uint8_t); break;
        case 2: RSA_PAIR(2, uint16_t); break;
        case 4: RSA_PAIR(4, uint32_t); break;
        case 8: RSA_PAIR(8, uint64_t); break;
        default: RSA_PAIR(16, U128); break;
    }
#undef RSA_PAIR
    return hipGetLastError() == hipSuccess ? RSORT_SUCCESS : RSORT_ERROR_DEVICE;
}

rsort_status rsort_digit_counts(const rsort_desc* desc, const void* keys, uint64_t n, unsigned long long* d_counts,
                                void* stream) {
    if (!desc || !key_width_ok(desc->key_bytes) || desc->key_kind > RSORT_KEY_FLOAT || !d_counts)
        return RSORT_ERROR_INVALID_ARGUMENT;
    if (n == 0) return RSORT_SUCCESS;
    if (!keys) return RSORT_ERROR_INVALID_ARGUMENT;
    const TwMasks tw = make_tw_masks(desc->key_kind, desc->key_bytes, desc->descending != 0);
    hipStream_t s = (hipStream_t)stream;
    const dim3 g(grid_of(n, 256u * 16u)), b(256);
    switch (desc->key_bytes) {
        case 1: hipLaunchKernelGGL(k_digit_counts<1>, g, b, 0, s, (const uint8_t*)keys, n, tw, d_counts); break;
        case 2: hipLaunchKernelGGL(k_digit_counts<2>, g, b, 0, s, (const uint16_t*)keys, n, tw, d_counts); break;
        case 4: hipLaunchKernelGGL(k_digit_counts<4>, g, b, 0, s, (const uint32_t*)keys, n, tw, d_counts); break;
        case 8: hipLaunchKernelGGL(k_digit_counts<8>, g, b, 0, s, (const uint64_t*)keys, n, tw, d_counts); break;
        default: hipLaunchKernelGGL(k_digit_counts<16>, g, b, 0, s, (const U128*)keys, n, tw, d_counts); break;
    }
    return hipGetLastError() == hipSuccess ? RSORT_SUCCESS : RSORT_ERROR_DEVICE;
}

}  // extern "C"
