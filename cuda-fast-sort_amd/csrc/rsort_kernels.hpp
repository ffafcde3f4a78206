// gfx950 kernels of the onesweep LSD radix sort (8-bit digits).
//
//   k_histogram  - one HBM read counts every digit of every pass into
//                  wave-private LDS counters, folds them into global counts and
//                  (last workgroup) turns them into exclusive digit offsets and
//                  the uniform-pass skip mask.
//                  Reference: include/rsort/detail/histogram.cuh:88-211.
//   k_onesweep   - one stable counting-sort pass: load tile, rank in LDS,
//                  decoupled lookback on a ticketed tile order, LDS-staged
//                  coalesced scatter of keys then values.
//                  Reference: include/rsort/detail/radix_kernel.cuh:118-294,
//                  ranker.cuh:62-170, lookback.cuh:252-303, scatter.cuh:159-202.
//   k_fill_zero, k_copy_back, k_iota, k_gather - small helpers of the op
//                  sequence (radix_kernel.cuh:377-386, 468-473).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "rsort_protocol.hpp"
#include "rsort_tuning.hpp"

namespace rsa {

// ---------------------------------------------------------------------------
// Key and value storage traits
// ---------------------------------------------------------------------------
struct alignas(16) U128 {
    uint64_t lo, hi;
};

template <int KB>
struct KeyTraits;

template <>
struct KeyTraits<1> {
    using Mem = uint8_t;
    using Reg = uint32_t;
};
template <>
struct KeyTraits<2> {
    using Mem = uint16_t;
    using Reg = uint32_t;
};
template <>
struct KeyTraits<4> {
    using Mem = uint32_t;
    using Reg = uint32_t;
};
template <>
struct KeyTraits<8> {
    using Mem = uint64_t;
    using Reg = uint64_t;
};
template <>
struct KeyTraits<16> {
    using Mem = U128;
    using Reg = U128;
};

template <int VB>
struct ValTraits;
template <>
struct ValTraits<0> {
    using Mem = uint8_t;  // unused placeholder
};
template <>
struct ValTraits<1> {
    using Mem = uint8_t;
};
template <>
struct ValTraits<2> {
    using Mem = uint16_t;
};
template <>
struct ValTraits<4> {
    using Mem = uint32_t;
};
template <>
struct ValTraits<8> {
    using Mem = uint64_t;
};
template <>
struct ValTraits<16> {
    using Mem = U128;
};

// ---------------------------------------------------------------------------
// Register-level key operations
// ---------------------------------------------------------------------------
template <int KB>
__device__ __forceinline__ uint32_t key_filler32() {
    if constexpr (KB >= 4) {
        return 0xFFFFFFFFu;
    } else {
        return (1u << (8 * KB)) - 1u;
    }
}

template <int KB>
__device__ __forceinline__ typename KeyTraits<KB>::Reg key_filler() {
    if constexpr (KB <= 4) {
        return key_filler32<KB>();
    } else if constexpr (KB == 8) {
        return ~0ull;
    } else {
        return U128{~0ull, ~0ull};
    }
}

// Twiddle in: order-preserving map to unsigned (radix_traits::twiddle_in,
// tuning.cuh:132-148), with the descending complement folded in.
template <int KB>
__device__ __forceinline__ typename KeyTraits<KB>::Reg tw_in(typename KeyTraits<KB>::Reg x, const TwMasks& t) {
    if constexpr (KB <= 4) {
        constexpr int SH = 32 - 8 * KB;
        const uint32_t neg = (uint32_t)(((int32_t)(x << SH)) >> 31);
        return x ^ ((neg & (uint32_t)t.fmask) | (uint32_t)t.smask) ^ (uint32_t)t.dmask;
    } else if constexpr (KB == 8) {
        const uint64_t neg = (uint64_t)(((int64_t)x) >> 63);
        return x ^ ((neg & t.fmask) | t.smask) ^ t.dmask;
    } else {
        const uint64_t neg = (uint64_t)(((int64_t)x.hi) >> 63);
        U128 r;
        r.lo = x.lo ^ (neg & t.fmask) ^ t.dmask;
        r.hi = x.hi ^ ((neg & t.fmask) | t.smask) ^ t.dmask;
        return r;
    }
}

// Twiddle out: inverse of tw_in (radix_traits::twiddle_out, tuning.cuh:151-165).
template <int KB>
__device__ __forceinline__ typename KeyTraits<KB>::Reg tw_out(typename KeyTraits<KB>::Reg x, const TwMasks& t) {
    if constexpr (KB <= 4) {
        constexpr int SH = 32 - 8 * KB;
        const uint32_t u = x ^ (uint32_t)t.dmask;
        const uint32_t neg = (uint32_t)(((int32_t)(u << SH)) >> 31);
        return u ^ ((~neg & (uint32_t)t.fmask) | (uint32_t)t.smask);
    } else if constexpr (KB == 8) {
        const uint64_t u = x ^ t.dmask;
        const uint64_t neg = (uint64_t)(((int64_t)u) >> 63);
        return u ^ ((~neg & t.fmask) | t.smask);
    } else {
        U128 u{x.lo ^ t.dmask, x.hi ^ t.dmask};
        const uint64_t neg = (uint64_t)(((int64_t)u.hi) >> 63);
        u.lo ^= (~neg & t.fmask);
        u.hi ^= ((~neg & t.fmask) | t.smask);
        return u;
    }
}

// Digit of a pass: byte (8*pass) of the key; `sh` = 8*(pass&3) selects the
// byte in a dword, `wsel` = pass>>2 the dword (both wave-uniform).
__device__ __forceinline__ uint32_t key_digit(uint32_t x, uint32_t sh, uint32_t) {
    return __builtin_amdgcn_ubfe(x, sh, 8);
}
__device__ __forceinline__ uint32_t key_digit(uint64_t x, uint32_t sh, uint32_t wsel) {
    const uint32_t w = wsel ? (uint32_t)(x >> 32) : (uint32_t)x;
    return __builtin_amdgcn_ubfe(w, sh, 8);
}
// The dword is picked by a mask merge, not a select: a select between the
// two halves of a key held in a per-thread array is folded by the compiler
// into a load from a computed address, which moves the whole key array to
// scratch memory.  The masks pass through readfirstlane, which the optimiser
// does not see through.
__device__ __forceinline__ uint32_t key_digit(const U128& x, uint32_t sh, uint32_t wsel) {
    const uint32_t mh = __builtin_amdgcn_readfirstlane(0u - ((wsel >> 1) & 1u));
    const uint32_t mw = __builtin_amdgcn_readfirstlane(0u - (wsel & 1u));
    const uint64_t m64 = ((uint64_t)mh << 32) | mh;
    const uint64_t h = (x.hi & m64) | (x.lo & ~m64);
    const uint32_t w = ((uint32_t)(h >> 32) & mw) | ((uint32_t)h & ~mw);
    return __builtin_amdgcn_ubfe(w, sh, 8);
}

template <int KB>
__device__ __forceinline__ typename KeyTraits<KB>::Reg key_to_reg(typename KeyTraits<KB>::Mem m) {
    return (typename KeyTraits<KB>::Reg)m;
}
template <>
__device__ __forceinline__ U128 key_to_reg<16>(U128 m) {
    return m;
}
template <int KB>
__device__ __forceinline__ typename KeyTraits<KB>::Mem key_to_mem(typename KeyTraits<KB>::Reg r) {
    return (typename KeyTraits<KB>::Mem)r;
}
template <>
__device__ __forceinline__ U128 key_to_mem<16>(U128 r) {
    return r;
}

// ---------------------------------------------------------------------------
// Wave / LDS helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t lane_id() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

// Number of set bits of a 64-lane mask below the calling lane.
__device__ __forceinline__ uint32_t mask_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

template <typename T>
__device__ __forceinline__ T wave_incl_scan(T x) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const T y = __shfl_up(x, (unsigned)o);
        if (lane >= (uint32_t)o) x += y;
    }
    return x;
}

// Compiler-only barrier: keeps the per-item LDS sequence of the ranker in
// program order (the LDS executes one wave's instructions in order).
#define RSA_COMPILER_FENCE() __atomic_signal_fence(__ATOMIC_SEQ_CST)
#ifndef RSA_RANK_PIN
#define RSA_RANK_PIN 1
#endif

// Phase timing (debug builds with -DRSA_PHASE_TIMING): thread 0 of each
// workgroup adds the wall-clock ticks (100 MHz) of each phase to a slot of
// the debug buffer; 64 copies of the slots spread the atomics.
#ifdef RSA_PHASE_TIMING
struct PhaseTimer {
    unsigned long long* dbg;
    uint64_t t, t0;
};
#define RSA_TIMER_START(dbgbase)                                                                      \
    PhaseTimer rsa_pt{(dbgbase) ? (dbgbase) + (blockIdx.x & 63u) * 32u : nullptr, (uint64_t)wall_clock64(), 0}; \
    rsa_pt.t0 = rsa_pt.t;
#define RSA_MARK(slot)                                                           \
    do {                                                                         \
        if (threadIdx.x == 0 && rsa_pt.dbg) {                                    \
            const uint64_t n_ = wall_clock64();                                  \
            atomicAdd(rsa_pt.dbg + (slot), (unsigned long long)(n_ - rsa_pt.t)); \
            rsa_pt.t = n_;                                                       \
        }                                                                        \
    } while (0)
#define RSA_COUNT(slot, v)                                                       \
    do {                                                                         \
        if (rsa_pt.dbg) atomicAdd(rsa_pt.dbg + (slot), (unsigned long long)(v)); \
    } while (0)
#define RSA_TOTAL(slot)                                                                          \
    do {                                                                                         \
        if (threadIdx.x == 0 && rsa_pt.dbg) {                                                    \
            atomicAdd(rsa_pt.dbg + (slot), (unsigned long long)(wall_clock64() - rsa_pt.t0));    \
            atomicAdd(rsa_pt.dbg + (slot) + 1, 1ull);                                            \
        }                                                                                        \
    } while (0)
#else
struct PhaseTimer {};
#define RSA_TIMER_START(dbgbase) \
    PhaseTimer rsa_pt{};         \
    (void)rsa_pt;
#define RSA_MARK(slot) \
    do {               \
    } while (0)
#define RSA_COUNT(slot, v) \
    do {                   \
    } while (0)
#define RSA_TOTAL(slot) \
    do {                \
    } while (0)
#endif

// Loads of data read once per launch (a pass's input keys and values, the
// histogram's keys): plain, or with the non-temporal hint.  Measured on
// MI355X at 2^28: the histogram's keys non-temporal 0.241 -> 0.182 ms (u32
// keys; u32/u32 sorts +5 %, u32 keys +1.7 %, u64 Zipf +1.3 %); the pass
// kernel's inputs non-temporal were slower (u32 keys pass 0.529 -> 0.540 ms,
// u64 +3 %), so RSA_NT_IN stays 0.
#ifndef RSA_NT_IN
#define RSA_NT_IN 0
#endif
#ifndef RSA_NT_HIST
#define RSA_NT_HIST 1
#endif
typedef unsigned int rsa_v4u __attribute__((ext_vector_type(4)));

template <bool NT, typename T>
__device__ __forceinline__ T ld_once(const T* p) {
    if constexpr (!NT) {
        return *p;
    } else if constexpr (std::is_same_v<T, U128>) {
        U128 r;
        r.lo = __builtin_nontemporal_load(&p->lo);
        r.hi = __builtin_nontemporal_load(&p->hi);
        return r;
    } else if constexpr (std::is_same_v<T, uint4>) {
        const rsa_v4u v = __builtin_nontemporal_load((const rsa_v4u*)p);
        return make_uint4(v.x, v.y, v.z, v.w);
    } else {
        return __builtin_nontemporal_load(p);
    }
}

template <typename T>
__device__ __forceinline__ T lds_ld(T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}
template <typename T>
__device__ __forceinline__ void lds_st(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}

// Stable rank of each of a wave's K items among the wave's items of the same
// digit that precede it (item order: (i, lane), i.e. item i of lane l comes
// after items 0..i-1 of every lane and after item i of lanes 0..l-1).
//   M  - the wave's 256 peer-mask words, zero on entry, zero on exit
//   C  - the wave's 256 digit counters, zero on entry, the wave's digit
//        totals on exit
//   dr - rank << 8 | digit per item
// A wave whose item i shares one digit takes the fast path (rank = lane);
// otherwise the lanes of a digit find each other by or-ing their lane bit
// into the digit's mask word, and the lowest such lane advances the counter.
// The LDS executes one wave's operations in order, so the sequence needs only
// compiler fences (reference: ranker.cuh:62-170 match_early_counts, which
// builds the peer mask from 8 ballots instead).
template <int KB, int K, typename CT>
__device__ __forceinline__ void rank_wave(const typename KeyTraits<KB>::Reg (&key)[K], uint32_t sh, uint32_t wsel,
                                          uint64_t* M, CT* C, uint32_t lane, uint32_t (&dr)[K]) {
#pragma unroll
    for (int i = 0; i < K; ++i) {
        const uint32_t d = key_digit(key[i], sh, wsel);
        const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
        uint32_t rank;
        if (__all(d == d0)) {
            const uint32_t c = lds_ld(C + d0);
            rank = c + lane;
#if RSA_RANK_PIN
            asm volatile("" : "+v"(rank));
#endif
            RSA_COMPILER_FENCE();
            if (lane == 0) lds_st(C + d0, (CT)(c + 64u));
        } else {
            // each half-wave ors and clears its own 32-bit half of the mask
            // word (half the LDS bytes of 64-bit ors and clears); the read
            // is the whole 64-bit word
            uint32_t* half = (uint32_t*)(M + d) + (lane >> 5);
            __hip_atomic_fetch_or(half, 1u << (lane & 31u), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
            RSA_COMPILER_FENCE();
            const uint64_t m = lds_ld(M + d);
            RSA_COMPILER_FENCE();
            lds_st(half, 0u);
            RSA_COMPILER_FENCE();
            const uint32_t c = lds_ld(C + d);
            const uint32_t below = mask_below(m);
            rank = c + below;
#if RSA_RANK_PIN
            // the rank is taken before the counter store is issued: otherwise
            // the compiler computes it after the store and waits for the
            // store's completion too, a second LDS round trip per item
            asm volatile("" : "+v"(rank));
#endif
            RSA_COMPILER_FENCE();
            if (below == 0u) lds_st(C + d, (CT)(c + (uint32_t)__builtin_popcountll(m)));
        }
        RSA_COMPILER_FENCE();
        dr[i] = (rank << 8) | d;
    }
}

// ---------------------------------------------------------------------------
// Kernel parameter blocks
// ---------------------------------------------------------------------------
struct HistParams {
    const void* keys;      // caller's keys (original representation)
    uint32_t* partials;    // [grid][passes][256] per-workgroup counts
    void* counts;          // [passes][256] Off, zeroed by k_fill_zero
    void* gp;              // [passes][256] Off exclusive digit offsets (out)
    Ctrl* ctrl;
    uint64_t n;
    uint64_t head;         // keys before the first 16-byte aligned group
    uint32_t passes;
    uint32_t allow_skip;
    uint32_t direct;       // fold by device atomics from every workgroup (small grids)
    uint4* self_clear;     // no fill kernel: the histogram clears this region (control header, counts,
    uint64_t self_clear_units;  // tickets, lookback) itself, in 16-byte units
    uint64_t head_units;   // its first part, header and counts (16-byte units)
    uint64_t clear_id;     // several workgroups clearing: this call's id (nonzero), see k_histogram
    uint32_t early_claim;  // workgroup 0 claims the clearing when it starts (0: only where counts are needed)
    uint32_t* xcd_lock;    // the device's per-XCD ticket lock word (null: never taken)
    uint32_t want_xcd;     // the passes can run on per-XCD tickets (see k_onesweep)
    TwMasks tw;
    unsigned long long* dbg;  // phase-timing slots (RSA_PHASE_TIMING builds), else null
};

struct PassParams {
    void* keys_a;          // caller's buffer (pass input when executed-pass index is even)
    void* keys_b;          // workspace ping-pong buffer
    void* vals_a;
    void* vals_b;
    const void* gp;        // [passes][256] exclusive digit offsets
    void* lookback;        // [tiles][256] words
    void* lb_groups;       // [passes][groups][256] group words (kLbGroup tiles each), zeroed per sort
    uint32_t* tickets_xcd; // this pass's kTicketParts per-XCD tile ticket counters, zeroed per sort
    uint32_t* xcd_lock;    // the device's per-XCD ticket lock word (released by the last walking pass)
    uint32_t xcd_ok;       // per-XCD tickets allowed for this kernel (the grid is whole rounds of them)
    Ctrl* ctrl;
    uint64_t n;
    uint32_t num_tiles;
    uint32_t pass;
    uint32_t passes;
    TwMasks tw;
    unsigned long long* dbg;  // phase-timing slots (RSA_PHASE_TIMING builds), else null
};

// ---------------------------------------------------------------------------
// Helpers
// ---------------------------------------------------------------------------
static __global__ void k_fill_zero(uint32_t* __restrict__ p, uint32_t words) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < words; i += gridDim.x * blockDim.x) p[i] = 0u;
}

// Clears the lookback array (16-byte units) once per sort.
static __global__ void k_fill_zero16(uint4* __restrict__ p, uint64_t units) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < units; i += stride)
        p[i] = make_uint4(0u, 0u, 0u, 0u);
}

// Odd number of executed passes leaves the result in the workspace buffer:
// copy it back (radix_kernel.cuh:468-473), decided on the device because
// uniform passes are skipped on the device.
// Copies `elems` elements of E; in 16-byte units when both pointers are
// 16-byte aligned (`vec`), element by element otherwise.
template <typename E>
__device__ __forceinline__ void copy_elems(E* __restrict__ dst, const E* __restrict__ src, uint64_t elems, bool vec,
                                           uint64_t t0, uint64_t stride) {
    uint64_t done = 0;
    if (vec) {
        constexpr uint64_t PER = 16u / sizeof(E);
        const uint64_t units = elems / PER;
        const uint4* s4 = (const uint4*)src;
        uint4* d4 = (uint4*)dst;
        for (uint64_t i = t0; i < units; i += stride) d4[i] = s4[i];
        done = units * PER;
    }
    for (uint64_t i = done + t0; i < elems; i += stride) dst[i] = src[i];
}

// Keys (and values) back into the caller's buffers, in one launch, when the
// executed pass count is odd; the grid is bounded.
template <typename KE, typename VE>
__global__ __launch_bounds__(256) void k_copy_back(const Ctrl* __restrict__ ctrl, uint32_t passes,
                                                   KE* __restrict__ dk, const KE* __restrict__ sk,
                                                   VE* __restrict__ dv, const VE* __restrict__ sv, uint64_t elems,
                                                   uint32_t vec_k, uint32_t vec_v) {
    if ((__builtin_popcount(pass_run_mask(ctrl->skip_mask, passes)) & 1u) == 0u) return;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t t0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    copy_elems(dk, sk, elems, vec_k != 0u, t0, stride);
    if (dv) copy_elems(dv, sv, elems, vec_v != 0u, t0, stride);
}

template <typename E>
__global__ void k_copy(E* __restrict__ dst, const E* __restrict__ src, uint64_t elems) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < elems; i += stride) dst[i] = src[i];
}

static __global__ void k_iota(uint32_t* __restrict__ p, uint64_t n) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = (uint32_t)i;
}

// Keys of records: dst key i (kel elements of E) = the kel elements at
// element offset `off` of record i (rel elements of E).
template <typename E>
__global__ void k_extract_keys(E* __restrict__ dst, const E* __restrict__ rec, uint64_t n, uint32_t rel,
                               uint32_t off, uint32_t kel) {
    const uint64_t total = n * kel;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += stride) {
        const uint64_t i = t / kel;
        const uint32_t c = (uint32_t)(t - i * kel);
        dst[t] = rec[i * rel + off + c];
    }
}

// dst record i = src record idx[i]; records of `rec` elements of type E.
template <typename E>
__global__ void k_gather(E* __restrict__ dst, const E* __restrict__ src, const uint32_t* __restrict__ idx,
                         uint64_t n, uint32_t rec) {
    const uint64_t total = n * rec;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += stride) {
        const uint64_t i = t / rec;
        const uint32_t c = (uint32_t)(t - i * rec);
        dst[t] = src[(uint64_t)idx[i] * rec + c];
    }
}

// ---------------------------------------------------------------------------
// Global histogram of all passes (one read).
// ---------------------------------------------------------------------------
template <int KB>
struct HistVec {
    static constexpr int V = 16 / KB;  // keys per lane per 16-byte load
    using Reg = typename KeyTraits<KB>::Reg;
    __device__ __forceinline__ static void unpack(const uint4& q, Reg (&k)[V]) {
        if constexpr (KB == 1) {
            const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
            for (int j = 0; j < 16; ++j) k[j] = (w[j >> 2] >> (8 * (j & 3))) & 0xFFu;
        } else if constexpr (KB == 2) {
            const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
            for (int j = 0; j < 8; ++j) k[j] = (w[j >> 1] >> (16 * (j & 1))) & 0xFFFFu;
        } else if constexpr (KB == 4) {
            k[0] = q.x;
            k[1] = q.y;
            k[2] = q.z;
            k[3] = q.w;
        } else if constexpr (KB == 8) {
            k[0] = (uint64_t)q.x | ((uint64_t)q.y << 32);
            k[1] = (uint64_t)q.z | ((uint64_t)q.w << 32);
        } else {
            k[0].lo = (uint64_t)q.x | ((uint64_t)q.y << 32);
            k[0].hi = (uint64_t)q.z | ((uint64_t)q.w << 32);
        }
    }
};

// byte p (compile-time) of a register key
template <int P>
__device__ __forceinline__ uint32_t reg_byte(uint32_t x) {
    return __builtin_amdgcn_ubfe(x, 8 * (P & 3), 8);
}
template <int P>
__device__ __forceinline__ uint32_t reg_byte(uint64_t x) {
    return __builtin_amdgcn_ubfe((uint32_t)(x >> (32 * (P >> 2))), 8 * (P & 3), 8);
}
template <int P>
__device__ __forceinline__ uint32_t reg_byte(const U128& x) {
    const uint64_t h = (P >= 8) ? x.hi : x.lo;
    return __builtin_amdgcn_ubfe((uint32_t)(h >> (32 * ((P >> 2) & 1))), 8 * (P & 3), 8);
}

__device__ __forceinline__ uint32_t rfl(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ uint64_t rfl(uint64_t x) {
    return (uint64_t)rfl((uint32_t)x) | ((uint64_t)rfl((uint32_t)(x >> 32)) << 32);
}
__device__ __forceinline__ U128 rfl(const U128& x) { return U128{rfl(x.lo), rfl(x.hi)}; }
__device__ __forceinline__ uint32_t reg_or(uint32_t a, uint32_t b) { return a | b; }
__device__ __forceinline__ uint64_t reg_or(uint64_t a, uint64_t b) { return a | b; }
__device__ __forceinline__ U128 reg_or(const U128& a, const U128& b) { return U128{a.lo | b.lo, a.hi | b.hi}; }
__device__ __forceinline__ uint32_t reg_and(uint32_t a, uint32_t b) { return a & b; }
__device__ __forceinline__ uint64_t reg_and(uint64_t a, uint64_t b) { return a & b; }
__device__ __forceinline__ U128 reg_and(const U128& a, const U128& b) { return U128{a.lo & b.lo, a.hi & b.hi}; }
__device__ __forceinline__ uint32_t reg_xor(uint32_t a, uint32_t b) { return a ^ b; }
__device__ __forceinline__ uint64_t reg_xor(uint64_t a, uint64_t b) { return a ^ b; }
__device__ __forceinline__ U128 reg_xor(const U128& a, const U128& b) { return U128{a.lo ^ b.lo, a.hi ^ b.hi}; }

// Count pass P of V keys held by each lane; counter (P, digit) is word
// (P * 256 + digit) * STRIDE from H.  If the whole wave's group has one digit
// in this pass, a single lane adds the group size (avoids 64-way LDS atomic
// serialisation on uniform bytes).
template <int KB, int P, int V, int STRIDE>
__device__ __forceinline__ void hist_count_pass(uint32_t* H, const typename KeyTraits<KB>::Reg (&k)[V],
                                                const typename KeyTraits<KB>::Reg& v,
                                                const typename KeyTraits<KB>::Reg& o0) {
    if constexpr (P < KB) {
        if (!__any(reg_byte<P>(v) != 0u)) {
            if (lane_id() == 0) atomicAdd(H + (P * 256 + reg_byte<P>(o0)) * STRIDE, (uint32_t)(64 * V));
        } else {
#pragma unroll
            for (int j = 0; j < V; ++j) atomicAdd(H + (P * 256 + reg_byte<P>(k[j])) * STRIDE, 1u);
        }
        hist_count_pass<KB, P + 1, V, STRIDE>(H, k, v, o0);
    }
}

template <int KB, int P, int STRIDE>
__device__ __forceinline__ void hist_count_one(uint32_t* H, const typename KeyTraits<KB>::Reg& k) {
    if constexpr (P < KB) {
        atomicAdd(H + (P * 256 + reg_byte<P>(k)) * STRIDE, 1u);
        hist_count_one<KB, P + 1, STRIDE>(H, k);
    }
}

// The histogram's clearing protocol without a fill kernel (k_histogram):
// thread 0 claims the clearing of the control header and the counts for call
// `id` if the word holds a stale value (neither id nor id - 1), by a
// compare-and-swap to id - 1.  `seen` is a value read from the word.
__device__ __forceinline__ bool hist_claim_clear(uint64_t* word, uint64_t id, uint64_t seen) {
    while (seen != id && seen != id - 1u) {
        if (__hip_atomic_compare_exchange_strong(word, &seen, id - 1u, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT))
            return true;
        // `seen` now holds the word's value: id - 1 or id (only this call's
        // workgroups write it while the call runs)
    }
    return false;
}

// Every thread of the claiming workgroup: coherent zeros over the header (but
// its clear_id word) and the counts, then, once they have completed, thread 0
// publishes the call's id.
__device__ __forceinline__ void hist_clear_head(const HistParams& a, uint32_t tid) {
    uint32_t* w32 = (uint32_t*)a.self_clear;
    const uint32_t words = (uint32_t)(a.head_units * 4u);
    constexpr uint32_t cw = (uint32_t)(offsetof(Ctrl, clear_id) / 4u);
    for (uint32_t i = tid; i < words; i += kHistThreads)
        if (i != cw && i != cw + 1u) __hip_atomic_store(w32 + i, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (tid == 0) __hip_atomic_store(&a.ctrl->clear_id, a.clear_id, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int KB, bool WIDE>
__global__ __launch_bounds__(kHistThreads) void k_histogram(HistParams a) {
    using Reg = typename KeyTraits<KB>::Reg;
    using Mem = typename KeyTraits<KB>::Mem;
    using Off = std::conditional_t<WIDE, uint64_t, uint32_t>;
    constexpr int P = KB;
    constexpr int WAVES = kHistThreads / 64;
    constexpr int V = HistVec<KB>::V;
    constexpr uint32_t G = 64u * V;  // keys per wave group
    constexpr int UNROLL = 4;  // 16-byte loads per lane per batch
#if RSA_HIST_SLOTS
    // Lane slots: counter (pass, digit) has SLOTS copies side by side and lane
    // l adds to copy l mod SLOTS, so the 32 lanes an LDS atomic serves at once
    // hit distinct banks whatever their digits (32 slots), or at most
    // 32 / SLOTS of them share one.
    constexpr int SLOTS = hist_slots(KB);
    constexpr int STRIDE = SLOTS;
    constexpr uint32_t HWORDS = (uint32_t)P * 256u * SLOTS;
#else
    // One copy of the counters per wave (pairs of waves for wide keys).
    constexpr int COPIES = hist_copies(KB);
    constexpr int WPC = WAVES / COPIES;
    constexpr int STRIDE = 1;
    constexpr uint32_t HWORDS = (uint32_t)COPIES * P * 256u;
#endif

    __shared__ uint32_t s_hist[HWORDS];
    __shared__ Off s_wsum[4];
    __shared__ uint32_t s_last;
    __shared__ uint32_t s_skip;
    __shared__ uint32_t s_mine;

    const uint32_t tid = threadIdx.x;
    const uint32_t w = tid >> 6;
    const uint32_t lane = tid & 63u;
    RSA_TIMER_START(a.dbg);

    for (uint32_t i = tid; i < HWORDS; i += kHistThreads) s_hist[i] = 0u;
    if (tid == 0) s_skip = 0u;
    if (a.self_clear_units && gridDim.x == 1u) {
        // the only workgroup: nothing else reads the control header, counts,
        // tickets or lookback words before the passes, so it clears them here.
        // The header goes through thread 0, which later stores the skip mask
        // into it.
        constexpr uint32_t hdr = (uint32_t)(sizeof(Ctrl) / 16u);
        if (tid == 0)
            for (uint32_t u = 0; u < hdr; ++u) a.self_clear[u] = make_uint4(0u, 0u, 0u, 0u);
        for (uint64_t u = hdr + tid; u < a.self_clear_units; u += kHistThreads)
            a.self_clear[u] = make_uint4(0u, 0u, 0u, 0u);
    } else if (a.self_clear_units) {
        // Several workgroups and no fill kernel.  The header word clear_id
        // holds this call's id (odd, unique per process) once the header and
        // the counts are cleared for this call, id - 1 while a workgroup clears
        // them, and anything else (a stale value) before.  One running
        // workgroup claims the clearing by a compare-and-swap from a stale
        // value to id - 1: workgroup 0 at its start, or, if it has not got
        // there yet, whichever workgroup first reaches the point where it
        // needs the counts (below).  The claimer clears the header (but
        // clear_id) and the counts with coherent stores, waits for them, then
        // stores id.  Every workgroup waits to read id before its first access
        // to the header or the counts; it only ever waits on a claimer that
        // was running when it claimed and waits on nothing, so the wait ends
        // whatever the dispatch order and whatever else runs on the device.
        // The ticket and lookback words, read only by the pass kernels, are
        // cleared by every workgroup with plain stores.
        if (blockIdx.x == 0 && a.early_claim) {
            if (tid == 0) s_mine = hist_claim_clear(&a.ctrl->clear_id, a.clear_id,
                                                    __hip_atomic_load(&a.ctrl->clear_id, __ATOMIC_RELAXED,
                                                                      __HIP_MEMORY_SCOPE_AGENT))
                                       ? 1u
                                       : 0u;
            __syncthreads();
            if (s_mine) hist_clear_head(a, tid);
        }
        for (uint64_t u = a.head_units + (uint64_t)blockIdx.x * kHistThreads + tid; u < a.self_clear_units;
             u += (uint64_t)gridDim.x * kHistThreads)
            a.self_clear[u] = make_uint4(0u, 0u, 0u, 0u);
    }
    __syncthreads();

#if RSA_HIST_SLOTS
    uint32_t* H = s_hist + (lane & (SLOTS - 1));
#else
    uint32_t* H = s_hist + (w / WPC) * (P * 256);
#endif
    // this workgroup's total of counter i = pass * 256 + digit
    auto wg_total = [&](uint32_t i) {
        uint32_t s = 0;
#if RSA_HIST_SLOTS
#pragma unroll
        for (int c = 0; c < SLOTS; ++c) s += s_hist[i * SLOTS + ((c + i) & (SLOTS - 1))];  // rotated: no conflicts
#else
#pragma unroll
        for (int c = 0; c < COPIES; ++c) s += s_hist[c * P * 256 + i];
#endif
        return s;
    };
    const Mem* keys = (const Mem*)a.keys;
    const bool twid = a.tw.active != 0u;

    const uint64_t body = (a.n - a.head) / G;  // full wave groups
    const uint64_t g0 = body * blockIdx.x / gridDim.x;
    const uint64_t g1 = body * (blockIdx.x + 1) / gridDim.x;
    const uint4* base = (const uint4*)(keys + a.head);

    // Software-pipelined: the next batch of UNROLL loads is in flight while
    // the current one is counted.
    auto count_group = [&](const uint4& q) {
        Reg k[V];
        HistVec<KB>::unpack(q, k);
        if (twid) {
#pragma unroll
            for (int j = 0; j < V; ++j) k[j] = tw_in<KB>(k[j], a.tw);
        }
        Reg o = k[0], an = k[0];
#pragma unroll
        for (int j = 1; j < V; ++j) {
            o = reg_or(o, k[j]);
            an = reg_and(an, k[j]);
        }
        const Reg o0 = rfl(o);
        const Reg v = reg_or(reg_xor(o, an), reg_xor(o, o0));
        hist_count_pass<KB, 0, V, STRIDE>(H, k, v, o0);
    };
    constexpr uint64_t STEP = (uint64_t)UNROLL * WAVES;
    uint64_t g = g0 + w;
    if (g + (UNROLL - 1) * WAVES < g1) {
        // two register batches alternate: one is counted while the other loads
        uint4 qa[UNROLL], qb[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) qa[u] = ld_once<RSA_NT_HIST != 0>(base + (g + (uint64_t)u * WAVES) * 64 + lane);
        while (true) {
            uint64_t gn = g + STEP;
            bool more = gn + (UNROLL - 1) * WAVES < g1;
            if (more) {
#pragma unroll
                for (int u = 0; u < UNROLL; ++u) qb[u] = ld_once<RSA_NT_HIST != 0>(base + (gn + (uint64_t)u * WAVES) * 64 + lane);
            }
#pragma unroll
            for (int u = 0; u < UNROLL; ++u) count_group(qa[u]);
            g = gn;
            if (!more) break;
            gn = g + STEP;
            more = gn + (UNROLL - 1) * WAVES < g1;
            if (more) {
#pragma unroll
                for (int u = 0; u < UNROLL; ++u) qa[u] = ld_once<RSA_NT_HIST != 0>(base + (gn + (uint64_t)u * WAVES) * 64 + lane);
            }
#pragma unroll
            for (int u = 0; u < UNROLL; ++u) count_group(qb[u]);
            g = gn;
            if (!more) break;
        }
    }
    for (; g < g1; g += WAVES) count_group(base[g * 64 + lane]);
    // head (before the aligned body) and tail (after it): last workgroup
    if (blockIdx.x == gridDim.x - 1) {
        const uint64_t tail0 = a.head + body * G;
        const uint64_t extra = a.head + (a.n - tail0);
        for (uint64_t t = tid; t < extra; t += kHistThreads) {
            const uint64_t idx = t < a.head ? t : tail0 + (t - a.head);
            Reg k = key_to_reg<KB>(keys[idx]);
            if (twid) k = tw_in<KB>(k, a.tw);
            hist_count_one<KB, 0, STRIDE>(H, k);
        }
    }
    __syncthreads();
    RSA_MARK(16);
    if (a.clear_id) {
        // the header and the counts must be cleared for this call (see above):
        // read the id, or claim the clearing if nobody has, or wait for the
        // running claimer
        if (tid == 0) {
            uint32_t mine = 0u;
            while (true) {
                const uint64_t cw = __hip_atomic_load(&a.ctrl->clear_id, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (cw == a.clear_id) break;
                if (cw != a.clear_id - 1u && hist_claim_clear(&a.ctrl->clear_id, a.clear_id, cw)) {
                    mine = 1u;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            s_mine = mine;
        }
        __syncthreads();
        if (s_mine) hist_clear_head(a, tid);
    }

    Off* counts = (Off*)a.counts;
    const uint32_t groups = (gridDim.x + kHistGroup - 1) / kHistGroup;
    bool from_lds = false;  // the finalising workgroup holds the totals in LDS
    if (gridDim.x == 1u) {
        // the only workgroup holds every count: totals into LDS, no fold
        constexpr int PER = (P * 256 + kHistThreads - 1) / kHistThreads;
        uint32_t t[PER];
#pragma unroll
        for (int r = 0; r < PER; ++r) {
            const uint32_t i = tid + (uint32_t)r * kHistThreads;
            t[r] = i < (uint32_t)(P * 256) ? wg_total(i) : 0u;
        }
        __syncthreads();  // every slot read before the totals overwrite them
#pragma unroll
        for (int r = 0; r < PER; ++r) {
            const uint32_t i = tid + (uint32_t)r * kHistThreads;
            if (i < (uint32_t)(P * 256)) s_hist[i] = t[r];
        }
        from_lds = true;
        __syncthreads();
        RSA_MARK(17);
    } else if (a.direct) {
        // Small grids: every workgroup adds its counts to the global counts
        // with device atomics, and the last arrival finalises (three coherent
        // round trips fewer than the two-level fold below).
        for (uint32_t i = tid; i < (uint32_t)(P * 256); i += kHistThreads) {
            const uint32_t s = wg_total(i);
            if (s) atomicAdd(counts + i, (Off)s);
        }
        __builtin_amdgcn_s_waitcnt(0);  // device atomics on the counts have completed
        __syncthreads();
        if (tid == 0) {
            const uint32_t prev =
                __hip_atomic_fetch_add(&a.ctrl->hist_done, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            s_last = (prev == gridDim.x - 1u) ? 1u : 0u;
        }
        __syncthreads();
        RSA_MARK(17);
        if (!s_last) {
            RSA_TOTAL(20);
            return;
        }
    } else {
        // Two-level fold: every workgroup stores its counts; the last arrival of
        // each group of kHistGroup workgroups sums the group's rows and adds them
        // to the global counts (kHistGroup x fewer device atomics); the last group
        // to finish finalises.
        {
            uint32_t* row = a.partials + (size_t)blockIdx.x * (P * 256);
            for (uint32_t i = tid; i < (uint32_t)(P * 256); i += kHistThreads)
                __hip_atomic_store(row + i, wg_total(i), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        const uint32_t group = blockIdx.x / kHistGroup;
        const uint32_t grp0 = group * kHistGroup;
        const uint32_t gsize = (gridDim.x - grp0) < (uint32_t)kHistGroup ? (gridDim.x - grp0) : (uint32_t)kHistGroup;
        // The rows were written with coherent (write-through) atomic stores, so
        // waiting for their completion orders them before the arrival count; an
        // agent-scope release here would write back the whole XCD L2 and, with
        // every workgroup of the XCD arriving together, serialise on it.
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (tid == 0) {
            const uint32_t prev =
                __hip_atomic_fetch_add(&a.ctrl->group_done[group], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            s_last = (prev == gsize - 1u) ? 1u : 0u;
        }
        __syncthreads();
        RSA_MARK(17);
        if (!s_last) {
            RSA_TOTAL(20);
            return;
        }
        // all of a counter's rows are loaded at once (one coherent round trip,
        // not one per row: at small n this fold is the kernel's critical path).
        // With a single group its last arrival holds the totals itself: they go
        // to LDS and straight into the offsets below, with no device atomics and
        // no second round trip.
        for (uint32_t i = tid; i < (uint32_t)(P * 256); i += kHistThreads) {
            const uint32_t* col = a.partials + (size_t)grp0 * (P * 256) + i;
            uint32_t v[kHistGroup];
#pragma unroll
            for (int m = 0; m < kHistGroup; ++m)
                v[m] = (uint32_t)m < gsize
                           ? __hip_atomic_load(col + (size_t)m * (P * 256), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                           : 0u;
            uint32_t s = 0;
#pragma unroll
            for (int m = 0; m < kHistGroup; ++m) s += v[m];
            if (groups == 1u) s_hist[i] = s;
            else if (s) atomicAdd(counts + i, (Off)s);
        }
        from_lds = groups == 1u;
        __builtin_amdgcn_s_waitcnt(0);  // device atomics on the counts have completed
        __syncthreads();
        if (groups > 1u) {
            if (tid == 0) {
                const uint32_t prev =
                    __hip_atomic_fetch_add(&a.ctrl->hist_done, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                s_last = (prev == groups - 1u) ? 1u : 0u;
            }
            __syncthreads();
            if (!s_last) {
                RSA_TOTAL(20);
                return;
            }
        }
    }

    // last workgroup: exclusive digit offsets per pass + uniform-pass mask
    // (counts are read with coherent atomic loads)
    Off* gp = (Off*)a.gp;
    Off cp[P];  // every pass's count of digit tid, loaded at once
    if (tid < 256) {
#pragma unroll
        for (int p = 0; p < P; ++p)
            cp[p] = from_lds ? (Off)s_hist[p * 256 + tid]
                             : __hip_atomic_load(counts + p * 256 + tid, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
#pragma unroll
    for (int p = 0; p < P; ++p) {
        Off c = 0, incl = 0;
        if (tid < 256) {
            c = cp[p];
            if ((uint64_t)c == a.n) atomicOr(&s_skip, 1u << p);
            incl = wave_incl_scan<Off>(c);
            if (lane == 63) s_wsum[w] = incl;
        }
        __syncthreads();
        if (tid < 256) {
            Off off = 0;
            for (uint32_t ww = 0; ww < w; ++ww) off += s_wsum[ww];
            gp[p * 256 + tid] = off + incl - c;
        }
        __syncthreads();
    }
    if (tid == 0) {
        const uint32_t skip = a.allow_skip ? s_skip : 0u;
        a.ctrl->skip_mask = skip;
        // Per-XCD tile tickets can make a tile wait on a tile that no running
        // workgroup holds yet (k_onesweep); of the sorts in flight on the
        // device, only the one holding this lock uses them, so at most one
        // pass kernel at a time depends on dispatch order.  Taken here, once
        // per sort, when some pass walks the lookback; released by the last
        // tile of the last such pass.
        uint32_t mode = 0u;
        if (a.want_xcd && a.xcd_lock) {
            const uint32_t run = pass_run_mask(skip, (uint32_t)P);
            if (run & ~skip) {
                uint32_t expect = 0u;
                mode = __hip_atomic_compare_exchange_strong(a.xcd_lock, &expect, 1u, __ATOMIC_RELAXED,
                                                            __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                           ? 1u
                           : 0u;
            }
        }
        a.ctrl->xcd_mode = mode;
    }
    RSA_MARK(18);
    RSA_TOTAL(20);
}

// ---------------------------------------------------------------------------
// Onesweep pass.
// ---------------------------------------------------------------------------
#ifndef RSA_LB_W0
#define RSA_LB_W0 8
#endif
#ifndef RSA_LB_WMAX
#define RSA_LB_WMAX 8
#endif
// Groups read per round by the grouped walk (two words each).
#ifndef RSA_LB_GW
#define RSA_LB_GW 4
#endif
// Where the pass kernel loads its values: 0 after the lookback (overlapping
// the key scatter), 1 with the keys, 2 after the key staging (overlapping
// the lookback; measured fastest).  Kernels whose LDS allows two workgroups
// per CU always use 0: their registers are budgeted for it.
#ifndef RSA_SCAN_BATCH
#define RSA_SCAN_BATCH 2  // column words per LDS round trip in the pass tile's digit scan (4 spills at 64 VGPRs)
#endif
#ifndef RSA_SCAN_COL_REGS
#define RSA_SCAN_COL_REGS 1
#endif
#ifndef RSA_VAL_AT
#define RSA_VAL_AT 2
#endif


// Lookback words are read and written with agent-scope relaxed atomics
// (coherent across the XCD L2s without cache maintenance); no other data
// travels through them, so no fence is needed.
template <typename W>
__device__ __forceinline__ W lb_load(const W* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename W>
__device__ __forceinline__ void lb_store(W* p, W v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// LDS of one onesweep tile.  The staging buffer `big` holds the per-wave
// peer-mask tables while ranking, then the tile's keys, then its values.
template <int KB, int VB, bool WIDE, int THREADS, int K>
struct OnesweepLds {
    using KM = typename KeyTraits<KB>::Mem;
    using VM = typename ValTraits<VB>::Mem;
    using Off = std::conditional_t<WIDE, uint64_t, uint32_t>;
    static constexpr int WAVES = THREADS / 64;
    static constexpr uint32_t TILE = (uint32_t)THREADS * K;
    static constexpr size_t KSTAGE = (size_t)TILE * sizeof(KM);
    static constexpr size_t VSTAGE = VB ? (size_t)TILE * sizeof(VM) : 0;
    static constexpr size_t MATCH = (size_t)WAVES * 256 * sizeof(uint64_t);
    static constexpr size_t BIG =
        KSTAGE > VSTAGE ? (KSTAGE > MATCH ? KSTAGE : MATCH) : (VSTAGE > MATCH ? VSTAGE : MATCH);
    static_assert(THREADS % 64 == 0 && THREADS >= 256, "workgroup must hold 256 digit owners");
    static_assert(TILE <= 65535u * 64u, "tile too large");

    // Per-wave digit counters, then tile-local digit starts: both < TILE,
    // so 16 bits suffice for tiles below 64K keys (4 KiB less LDS per
    // 8-wave tile, which lets a fourth 8192-key tile fit in a CU's 160 KiB).
    using Cnt = std::conditional_t<(TILE < 65536u), uint16_t, uint32_t>;

    __attribute__((aligned(16))) uint8_t big[BIG];
    __attribute__((aligned(16))) Cnt cnt[WAVES * 256];
    Off base[256];
    uint32_t wsum[4];
    uint32_t tile;
    uint32_t early[256];  // tile digit counts taken before the rank (early aggregate)
#ifdef RSA_PHASE_TIMING
    uint32_t dbg[2];
#endif
};

// Where one executed pass reads and writes, and what it applies.
template <int KB, int VB>
struct PassIo {
    const typename KeyTraits<KB>::Mem* __restrict__ kin;
    typename KeyTraits<KB>::Mem* __restrict__ kout;
    const typename ValTraits<VB>::Mem* __restrict__ vin;
    typename ValTraits<VB>::Mem* __restrict__ vout;
    uint32_t sh, wsel;  // digit position (key_digit)
    uint32_t epoch;     // lookback epoch of the pass
    bool first, last;   // apply twiddle-in on load / twiddle-out on store
};

// One tile of a pass: its lookback row, whether it heads its lookback chain
// (publishes an inclusive prefix at once), its first key, and the end of the
// key range it belongs to (the tile holds keys tbase .. min(tbase+TILE, kend)).
struct TileGeom {
    uint32_t lb_tile;
    bool head;
    uint64_t tbase;
    uint64_t kend;
};

// Take a tile ticket and clear the per-wave rank tables.  Returns the ticket
// (every thread); the tables are ready for onesweep_tile.
template <int KB, int VB, bool WIDE, int THREADS, int K>
__device__ __forceinline__ uint32_t onesweep_begin(OnesweepLds<KB, VB, WIDE, THREADS, K>& L, uint32_t* ticket) {
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    const uint32_t w = tid >> 6;
    if (tid == 0) L.tile = atomicAdd(ticket, 1u);
    {
        // each wave clears its own match table and digit counters
        uint4* m = (uint4*)(L.big + (size_t)w * 256 * sizeof(uint64_t));
        m[lane] = make_uint4(0u, 0u, 0u, 0u);
        m[lane + 64] = make_uint4(0u, 0u, 0u, 0u);
        constexpr uint32_t cnt_units = 256u * sizeof(L.cnt[0]) / 16u;
        if (lane < cnt_units) ((uint4*)(L.cnt + w * 256))[lane] = make_uint4(0u, 0u, 0u, 0u);
    }
    if (tid < 256) L.early[tid] = 0u;
    __syncthreads();
    return L.tile;
}

// How pass `pass` runs (pass_run_mask): whether it runs, whether it is a plain
// copy, its executed-pass index e (buffer parity: even reads the caller's
// buffer), its lookback epoch (parity among the passes that walk the lookback:
// a copy pass writes no lookback words, so it must not flip the epoch), and
// whether it is the first or last executed pass (twiddle in / out).
struct PassRun {
    bool runs, copy, first, last;
    bool last_walk;  // no later pass walks the lookback
    uint32_t e, epoch;
};

__device__ __forceinline__ PassRun pass_run(uint32_t skip_mask, uint32_t passes, uint32_t pass) {
    const uint32_t run = pass_run_mask(skip_mask, passes);
    const uint32_t below = (1u << pass) - 1u;
    PassRun r;
    r.runs = (run >> pass) & 1u;
    r.copy = (skip_mask >> pass) & 1u;
    r.e = (uint32_t)__builtin_popcount(run & below);
    r.epoch = (uint32_t)__builtin_popcount(run & ~skip_mask & below) & 1u;
    r.first = r.e == 0u;
    r.last = (run >> (pass + 1u)) == 0u;
    r.last_walk = ((run & ~skip_mask) >> (pass + 1u)) == 0u;
    return r;
}

// A pass whose digit is uniform, run only to keep the executed pass count
// even: the stable sort by a constant digit is the identity, so keys
// [lo, min(hi, n)) are copied (twiddled in or out when it is the first or last
// executed pass) and values with them, `stride` threads apart.
template <int KB, int VB>
__device__ __forceinline__ void copy_pass_range(const PassIo<KB, VB>& io, const TwMasks& tw, uint64_t lo, uint64_t hi,
                                                uint64_t n, uint64_t stride) {
    const uint64_t end = hi < n ? hi : n;
    for (uint64_t i = lo + threadIdx.x; i < end; i += stride) {
        auto k = key_to_reg<KB>(io.kin[i]);
        if (io.first) k = tw_in<KB>(k, tw);
        if (io.last) k = tw_out<KB>(k, tw);
        io.kout[i] = key_to_mem<KB>(k);
        if constexpr (VB > 0) io.vout[i] = io.vin[i];
    }
}

// One LSD pass over the whole array: one workgroup per tile, tiles taken in
// ticket order so every tile's predecessors are already running.  SEGWALK:
// the lookback walk in one-round segments (passes of few tiles).  EARLY: the
// tile aggregate is published from digit counts taken before the rank.
template <int KB, int VB, bool WIDE, int THREADS, int K, bool SEGWALK, bool EARLY>
__global__ __launch_bounds__(THREADS, pass_min_waves(KB, VB, K)) void k_onesweep(PassParams a) {
    using KM = typename KeyTraits<KB>::Mem;
    using KR = typename KeyTraits<KB>::Reg;
    using VM = typename ValTraits<VB>::Mem;
    using Off = std::conditional_t<WIDE, uint64_t, uint32_t>;
    constexpr int WAVES = THREADS / 64;
    constexpr uint32_t TILE = (uint32_t)THREADS * K;
    constexpr int LB_W0 = RSA_LB_W0;
    constexpr int LB_WMAX = RSA_LB_WMAX;
    __shared__ OnesweepLds<KB, VB, WIDE, THREADS, K> L;

    // the two header words every pass reads first, in one load (the second,
    // xcd_mode, picks the ticket counter below)
    const uint2 hdr = *(const uint2*)&a.ctrl->skip_mask;
    // every kernel argument the prologue and the tile need, loaded together
    // with that word: otherwise the compiler loads the argument lines one
    // miss after another, ahead of the ticket
    asm volatile("" ::"s"(a.keys_a), "s"(a.keys_b), "s"(a.vals_a), "s"(a.vals_b), "s"(a.gp), "s"(a.lookback),
                 "s"(a.lb_groups), "s"(a.tickets_xcd), "s"(a.xcd_ok), "s"(a.n), "s"(a.num_tiles), "s"(a.tw.fmask),
                 "s"(a.tw.active));
    const PassRun pr = pass_run(hdr.x, a.passes, a.pass);
    if (!pr.runs) return;  // uniform digit: pass is the identity
    PassIo<KB, VB> io;
    io.kin = (const KM*)((pr.e & 1u) ? a.keys_b : a.keys_a);
    io.kout = (KM*)((pr.e & 1u) ? a.keys_a : a.keys_b);
    io.vin = (const VM*)((pr.e & 1u) ? a.vals_b : a.vals_a);
    io.vout = (VM*)((pr.e & 1u) ? a.vals_a : a.vals_b);
    io.sh = 8u * (a.pass & 3u);
    io.wsel = a.pass >> 2;
    io.epoch = pr.epoch;
    io.first = pr.first && a.tw.active;
    io.last = pr.last && a.tw.active;
    if (pr.copy) {
        copy_pass_range<KB, VB>(io, a.tw, (uint64_t)blockIdx.x * TILE, (uint64_t)(blockIdx.x + 1) * TILE, a.n,
                                THREADS);
        return;
    }

    RSA_TIMER_START(a.dbg);
#ifdef RSA_PHASE_TIMING
    if (threadIdx.x < 2) L.dbg[threadIdx.x] = 0u;
#endif
    // Ticket k of counter `part` (of `parts`) is tile (k / C) * (parts * C)
    // + part * C + k % C.  One counter with C = 1 is the global ticket: tiles
    // are handed out in increasing order, so a tile only ever waits on tiles
    // already held by running workgroups, whatever else runs on the device.
    // With a counter per XCD (workgroup b is placed on XCD b mod 8, so each
    // XCD works through runs of C neighbouring tiles, whose digit runs share
    // partial cache lines) each counter still hands out its tiles in
    // increasing order, but a tile's predecessors from other counters may not
    // be handed out yet.  That needs the next parts * C workgroups in
    // dispatch order to be resident together: the host allows it only where
    // the kernel alone can hold twice that (xcd_ok), and of the sorts in
    // flight on the device only the one holding the device's lock uses it
    // (ctrl->xcd_mode, taken by the histogram), so no two pass kernels can
    // each fill XCDs with workgroups waiting on the other's undispatched
    // tiles.  The grid is whole rounds of parts * C tiles when xcd_ok:
    // workgroups past the last tile end here.
    const bool xcd = a.xcd_ok && hdr.y;
    const uint32_t parts = xcd ? kTicketParts : 1u;
    const uint32_t chunk = xcd ? tile_chunk(KB) : 1u;
    const uint32_t part = blockIdx.x % parts;
    const uint32_t tk = onesweep_begin(L, xcd ? a.tickets_xcd + part : &a.ctrl->ticket[a.pass]);
    const uint32_t tile = ticket_tile(tk, part, parts, chunk);
    if (tile >= a.num_tiles) return;
    TileGeom g;
    g.lb_tile = tile;
    g.head = tile == 0u;
    g.tbase = (uint64_t)tile * TILE;
    g.kend = a.n;
    const Off* gp_row = (const Off*)a.gp + a.pass * 256u;
    Off* lb = (Off*)a.lookback;
    constexpr bool LB_GROUPS = true;
    Off* lbg = (Off*)a.lb_groups + (size_t)a.pass * ((a.num_tiles + kLbGroup - 1) / kLbGroup) * 256u;
    constexpr bool LB_SEGWALK = SEGWALK;
    constexpr bool EARLY_AGG = EARLY;
    const TwMasks& tw = a.tw;
    const uint32_t tile_tid = threadIdx.x;
#include "rsort_onesweep_tile.inc"
    // The last tile of the last pass that walks: every tile of the pass has
    // been handed out, so the sort no longer depends on dispatch order.
    if (xcd && pr.last_walk && tile + 1u == a.num_tiles && threadIdx.x == 0)
        __hip_atomic_store(a.xcd_lock, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    RSA_TOTAL(10);
}

}  // namespace rsa
