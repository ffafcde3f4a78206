// Persistent chunked onesweep pass for gfx950.
//
// One workgroup of THREADS (16 waves) per CU loops over chunks of the array,
// taking chunk numbers from the pass's ticket counter.  A chunk is SUBS
// sub-tiles of SUB = THREADS*IPT keys (96 KiB of keys per chunk):
//
//   1. rank every sub-tile in registers (per-wave LDS peer masks + per
//      (sub-tile, wave, digit) u16 counters),
//   2. publish the chunk's digit aggregate and scan the sub-tile totals (the
//      16-bit totals of a digit are packed in one 64-bit word, so one
//      256-wide scan yields every sub-tile's digit starts),
//   3. stage every sub-tile at its chunk-local rank in LDS -- the key and
//      rank registers are then free for the next chunk's loads and for the
//      lookback window,
//   4. run one decoupled lookback per chunk (the lookback traffic per key
//      falls with the square of the unit size),
//   5. write the staged sub-tiles as coalesced digit runs; values are loaded
//      during 4-5 and follow through the same LDS buffer.
// The next chunk's ticket is taken and its keys are loaded right after the
// ranking, so those loads are in flight during steps 2-5.
//
// Reference: the per-tile onesweep_byte kernel (include/rsort/detail/
// radix_kernel.cuh:118-294) with its ranker (ranker.cuh:62-170), lookback
// (lookback.cuh:252-303) and staged scatter (scatter.cuh:159-202); the chunk
// structure and prefetch are this implementation's.
#pragma once

#include "rsort_kernels.hpp"

namespace rsa {

// Lookback windows of the chunked pass (first round, maximum).
#ifndef RSA_PLB_W0
#define RSA_PLB_W0 8
#endif
#ifndef RSA_PLB_WMAX
#define RSA_PLB_WMAX 32
#endif

// Keys of chunk c into r (wave-striped per sub-tile; fillers past n).  The
// chunk and wave offsets form a wave-uniform base pointer, so each load needs
// only the lane offset in a vector register.
template <int KB, int IPT, int SUBS, uint32_t SUB>
__device__ __forceinline__ void load_chunk_keys(typename KeyTraits<KB>::Reg (&r)[IPT * SUBS],
                                                const typename KeyTraits<KB>::Mem* __restrict__ kin, uint32_t c,
                                                uint32_t w, uint32_t lane, uint64_t n) {
    constexpr uint64_t CHUNK = (uint64_t)SUB * SUBS;
    const uint64_t wbase = (uint64_t)c * CHUNK + (uint64_t)w * 64u * IPT;
    const typename KeyTraits<KB>::Mem* pw = kin + wbase;
    if ((uint64_t)c * CHUNK + CHUNK <= n) {
#pragma unroll
        for (int s = 0; s < SUBS; ++s)
#pragma unroll
            for (int i = 0; i < IPT; ++i) r[s * IPT + i] = key_to_reg<KB>(pw[(size_t)s * SUB + (size_t)i * 64u + lane]);
    } else {
#pragma unroll
        for (int s = 0; s < SUBS; ++s)
#pragma unroll
            for (int i = 0; i < IPT; ++i) {
                const uint64_t idx = wbase + (uint64_t)s * SUB + (uint64_t)i * 64u + lane;
                r[s * IPT + i] = idx < n ? key_to_reg<KB>(kin[idx]) : key_filler<KB>();
            }
    }
}

template <int KB, int VB, bool WIDE, int THREADS, int IPT, int SUBS>
__global__ __launch_bounds__(THREADS) void k_onesweep_p(PassParams a) {
    using KM = typename KeyTraits<KB>::Mem;
    using KR = typename KeyTraits<KB>::Reg;
    using VM = typename ValTraits<VB>::Mem;
    using Off = std::conditional_t<WIDE, uint64_t, uint32_t>;
    constexpr int WAVES = THREADS / 64;
    constexpr uint32_t SUB = (uint32_t)THREADS * IPT;
    constexpr uint32_t CHUNK = SUB * SUBS;
    constexpr int K = IPT * SUBS;
    constexpr size_t STAGE = (size_t)SUB * (sizeof(KM) > (VB ? sizeof(VM) : 0) ? sizeof(KM) : sizeof(VM));
    constexpr size_t MATCH = (size_t)WAVES * 256 * sizeof(uint64_t);
    constexpr size_t BIG = (size_t)SUBS * STAGE > MATCH ? (size_t)SUBS * STAGE : MATCH;
    static_assert(SUBS >= 1 && SUBS <= 4, "sub-tile totals are packed four to a 64-bit word");
    static_assert(BIG == (size_t)SUBS * STAGE || MATCH > (size_t)SUBS * STAGE, "staging holds every sub-tile");
    static_assert(SUB <= 65535u, "sub-tile offsets are 16-bit");
    static_assert(THREADS >= SUBS * 256, "one thread per (sub-tile, digit)");
    static_assert(THREADS % 256 == 0, "digit owners fill whole waves");

    __shared__ __attribute__((aligned(16))) uint8_t s_big[BIG];
    __shared__ __attribute__((aligned(16))) uint16_t s_off[SUBS * WAVES * 256];
    __shared__ uint32_t s_tot[SUBS * 256];
    __shared__ Off s_base[SUBS * 256];
    __shared__ uint64_t s_wsum[4];
    __shared__ uint32_t s_ticket;
#ifdef RSA_PHASE_TIMING
    __shared__ uint32_t s_dbg[2];
#endif

    const PassRun pr = pass_run(a.ctrl->skip_mask, a.passes, a.pass);
    if (!pr.runs) return;  // uniform digit: pass is the identity
    const uint32_t e = pr.e;
    const uint32_t epoch = pr.epoch;
    const bool first = pr.first && a.tw.active;
    const bool last = pr.last && a.tw.active;
    const KM* __restrict__ kin = (const KM*)((e & 1u) ? a.keys_b : a.keys_a);
    KM* __restrict__ kout = (KM*)((e & 1u) ? a.keys_a : a.keys_b);
    const VM* __restrict__ vin = (const VM*)((e & 1u) ? a.vals_b : a.vals_a);
    VM* __restrict__ vout = (VM*)((e & 1u) ? a.vals_a : a.vals_b);
    if (pr.copy) {
        PassIo<KB, VB> io{kin, kout, vin, vout, 0u, 0u, epoch, first, last};
        copy_pass_range<KB, VB>(io, a.tw, (uint64_t)blockIdx.x * THREADS, a.n, a.n, (uint64_t)gridDim.x * THREADS);
        return;
    }
    const uint32_t nchunks = a.num_tiles;
    const uint64_t n = a.n;
    Off* lb = (Off*)a.lookback;
    const Off* gp = (const Off*)a.gp + (size_t)a.pass * 256u;

    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    const uint32_t w = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform
    const uint32_t sh = 8u * (a.pass & 3u);
    const uint32_t wsel = a.pass >> 2;
    RSA_TIMER_START(a.dbg);

    // ---- persistent loop over chunks; the next chunk's keys load into nxt ----
    if (tid == 0) s_ticket = atomicAdd(&a.ctrl->ticket[a.pass], 1u);
    __syncthreads();
    uint32_t c = __builtin_amdgcn_readfirstlane(s_ticket);
    __syncthreads();  // every thread has read the ticket before it is reused
    if (c >= nchunks) return;
    KR cur[K], nxt[K];
    load_chunk_keys<KB, IPT, SUBS, SUB>(cur, kin, c, w, lane, n);
    while (true) {
        // Opaque copies of the thread ids: without them the compiler hoists
        // every per-item address out of the loop and spills them.
        uint32_t tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        const uint32_t lane = tid & 63u;
        const uint64_t lanebit = 1ull << lane;
        const uint64_t cb = (uint64_t)c * CHUNK;
        const uint32_t valid = (cb + CHUNK <= n) ? CHUNK : (uint32_t)(n - cb);
        if (first) {
            // the first executed pass reads the caller's keys: twiddle valid keys
            const uint64_t base = cb + (uint64_t)w * 64u * IPT + lane;
#pragma unroll
            for (int s = 0; s < SUBS; ++s)
#pragma unroll
                for (int i = 0; i < IPT; ++i)
                    if (base + (uint64_t)s * SUB + (uint64_t)i * 64u < n)
                        cur[s * IPT + i] = tw_in<KB>(cur[s * IPT + i], a.tw);
        }
        // clear the counters and this wave's peer-mask table
        {
            uint4* o = (uint4*)s_off;
            constexpr uint32_t OFF_UNITS = SUBS * WAVES * 256 * 2 / 16;
            for (uint32_t u = tid; u < OFF_UNITS; u += THREADS) o[u] = make_uint4(0u, 0u, 0u, 0u);
            uint4* m = (uint4*)(s_big + (size_t)w * 256 * sizeof(uint64_t));
            m[lane] = make_uint4(0u, 0u, 0u, 0u);
            m[lane + 64] = make_uint4(0u, 0u, 0u, 0u);
        }
        if (tid == 0) s_ticket = atomicAdd(&a.ctrl->ticket[a.pass], 1u);
        __syncthreads();
        RSA_MARK(0);

        // ---- rank: per (sub-tile, wave) stable rank among same-digit items ----
        uint64_t* M = (uint64_t*)s_big + (size_t)w * 256;
        uint32_t dr[K];  // rank << 8 | digit
#pragma unroll
        for (int s = 0; s < SUBS; ++s) {
            uint16_t* C = s_off + ((size_t)s * WAVES + w) * 256;
#pragma unroll
            for (int i = 0; i < IPT; ++i) {
                const uint32_t d = key_digit(cur[s * IPT + i], sh, wsel);
                const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
                uint32_t rank;
                if (__all(d == d0)) {
                    const uint32_t cnt = lds_ld(C + d0);
                    rank = cnt + lane;
                    RSA_COMPILER_FENCE();
                    if (lane == 0) lds_st(C + d0, (uint16_t)(cnt + 64u));
                } else {
                    __hip_atomic_fetch_or(M + d, lanebit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
                    RSA_COMPILER_FENCE();
                    const uint64_t m = lds_ld(M + d);
                    RSA_COMPILER_FENCE();
                    lds_st(M + d, (uint64_t)0);
                    RSA_COMPILER_FENCE();
                    const uint32_t cnt = lds_ld(C + d);
                    const uint32_t below = mask_below(m);
                    rank = cnt + below;
                    RSA_COMPILER_FENCE();
                    if (below == 0u) lds_st(C + d, (uint16_t)(cnt + (uint32_t)__builtin_popcountll(m)));
                }
                RSA_COMPILER_FENCE();
                dr[s * IPT + i] = (rank << 8) | d;
            }
        }
        __syncthreads();
        RSA_MARK(2);

        // the next chunk: its keys load while this one is scanned and scattered
        const uint32_t cn = __builtin_amdgcn_readfirstlane(s_ticket);
        if (cn < nchunks) load_chunk_keys<KB, IPT, SUBS, SUB>(nxt, kin, cn, w, lane, n);

        // ---- per (sub-tile, digit): prefix over waves, sub-tile totals ----
        if (tid < SUBS * 256u) {
            const uint32_t s = tid >> 8, d = tid & 255u;
            uint16_t* col = s_off + (size_t)s * WAVES * 256 + d;
            uint32_t r = 0;
#pragma unroll
            for (int ww = 0; ww < WAVES; ++ww) {
                const uint32_t v = col[ww * 256];
                col[ww * 256] = (uint16_t)r;
                r += v;
            }
            s_tot[s * 256 + d] = r;
        }
        __syncthreads();

        // ---- per digit: chunk aggregate publish + packed scan of sub-tile totals ----
        uint32_t tot[SUBS];
        uint64_t packed = 0, incl = 0;
        uint32_t pub = 0;
        Off gbase = 0;
        if (tid < 256u) {
            gbase = gp[tid];
            uint32_t sum = 0;
#pragma unroll
            for (int s = 0; s < SUBS; ++s) {
                tot[s] = s_tot[s * 256 + tid];
                sum += tot[s];
                packed |= (uint64_t)tot[s] << (16 * s);
            }
            pub = sum - (tid == 255u ? (CHUNK - valid) : 0u);  // fillers are not keys
            lb_store(lb + (size_t)c * kRadix + tid, lb_pack<Off>(c == 0u ? kLbPrefix : kLbAggregate, epoch, (Off)pub));
            incl = wave_incl_scan<uint64_t>(packed);
            if (lane == 63u) s_wsum[w] = incl;
        }
        __syncthreads();
        uint64_t excl = 0;
        if (tid < 256u) {
            uint64_t woff = 0;
            for (uint32_t ww = 0; ww < w; ++ww) woff += s_wsum[ww];
            excl = woff + incl - packed;  // field s: start of digit tid in sub-tile s
#pragma unroll
            for (int s = 0; s < SUBS; ++s) s_tot[s * 256 + tid] = (uint32_t)(excl >> (16 * s)) & 0xFFFFu;
        }
        __syncthreads();
        if (tid < SUBS * 256u) {
            const uint32_t s = tid >> 8, d = tid & 255u;
            const uint16_t start = (uint16_t)s_tot[s * 256 + d];
            uint16_t* col = s_off + (size_t)s * WAVES * 256 + d;
#pragma unroll
            for (int ww = 0; ww < WAVES; ++ww) col[ww * 256] = (uint16_t)(col[ww * 256] + start);
        }
        __syncthreads();
        RSA_MARK(3);

        // ---- stage every sub-tile at its rank (region s = sub-tile s) ----
        // (key-value sorts keep each item's 16-bit staging position, two per
        // register, to stage its value later)
        uint32_t pp[VB > 0 ? (K + 1) / 2 : 1];
#pragma unroll
        for (int s = 0; s < SUBS; ++s) {
            KM* SK = (KM*)(s_big + (size_t)s * STAGE);
            const uint16_t* C = s_off + ((size_t)s * WAVES + w) * 256;
#pragma unroll
            for (int i = 0; i < IPT; ++i) {
                const uint32_t x = dr[s * IPT + i];
                const uint32_t pos = (uint32_t)C[x & 255u] + (x >> 8);
                SK[pos] = key_to_mem<KB>(cur[s * IPT + i]);
                if constexpr (VB > 0) {
                    const int k = s * IPT + i;
                    if (k & 1) pp[k >> 1] |= pos << 16;
                    else pp[k >> 1] = pos;
                }
            }
        }
        RSA_MARK(5);

        // ---- chunk lookback (digit owners) ----
        if (tid < 256u) {
            Off prefix = 0;
            if (c > 0u) {
                uint32_t steps = 0, spins = 0;
                const Off* colp = lb + tid;
                prefix = lookback_walk<Off, RSA_PLB_W0, RSA_PLB_WMAX>(
                    (int64_t)c, epoch, [colp](int64_t t) { return lb_load(colp + t * (int64_t)kRadix); },
                    []() { __builtin_amdgcn_s_sleep(1); }, steps, spins);
                lb_store(lb + (size_t)c * kRadix + tid, lb_pack<Off>(kLbPrefix, epoch, prefix + (Off)pub));
#ifdef RSA_PHASE_TIMING
                RSA_COUNT(8, steps);
                RSA_COUNT(9, spins);
#endif
            }
            Off cum = 0;
#pragma unroll
            for (int s = 0; s < SUBS; ++s) {
                const Off start = (Off)((excl >> (16 * s)) & 0xFFFFu);
                s_base[s * 256 + tid] = gbase + prefix + cum - start;
                cum += (Off)tot[s];
            }
        }
        __syncthreads();
        RSA_MARK(4);

        // values are needed only after their keys are out: load them now (not
        // before the lookback, whose window needs the registers)
        VM vv[VB > 0 ? K : 1];
        if constexpr (VB > 0) {
            const uint64_t wbase = cb + (uint64_t)w * 64u * IPT;
            const VM* pw = vin + wbase;
            const bool vfull = cb + CHUNK <= n;
#pragma unroll
            for (int s = 0; s < SUBS; ++s)
#pragma unroll
                for (int i = 0; i < IPT; ++i) {
                    const size_t o = (size_t)s * SUB + (size_t)i * 64u + lane;
                    vv[s * IPT + i] = (vfull || wbase + o < n) ? pw[o] : VM{};
                }
        }

        // ---- write the staged sub-tiles as coalesced digit runs; a
        //      key-value sort then stages that sub-tile's values in the same
        //      region and writes them to the same places ----
#pragma unroll
        for (int s = 0; s < SUBS; ++s) {
            const KM* SK = (const KM*)(s_big + (size_t)s * STAGE);
            const uint32_t sv = valid > (uint32_t)s * SUB ? valid - (uint32_t)s * SUB : 0u;
            const uint32_t vs = sv < SUB ? sv : SUB;
            const Off* B = s_base + s * 256;
            Off gi[VB > 0 ? IPT : 1];
#pragma unroll
            for (int j = 0; j < IPT; ++j) {
                const uint32_t pos = (uint32_t)j * THREADS + tid;
                if (pos < vs) {
                    KR k = key_to_reg<KB>(SK[pos]);
                    const Off g = B[key_digit(k, sh, wsel)] + (Off)pos;
                    if (last) k = tw_out<KB>(k, a.tw);
                    kout[g] = key_to_mem<KB>(k);
                    if constexpr (VB > 0) gi[j] = g;
                }
            }
            if constexpr (VB > 0) {
                __syncthreads();
                VM* SV = (VM*)(s_big + (size_t)s * STAGE);
#pragma unroll
                for (int i = 0; i < IPT; ++i) {
                    const int k = s * IPT + i;
                    SV[(pp[k >> 1] >> ((k & 1) * 16)) & 0xFFFFu] = vv[k];
                }
                __syncthreads();
#pragma unroll
                for (int j = 0; j < IPT; ++j) {
                    const uint32_t pos = (uint32_t)j * THREADS + tid;
                    if (pos < vs) vout[gi[j]] = SV[pos];
                }
            }
        }
        __syncthreads();  // the next chunk clears the LDS this chunk read
        RSA_MARK(6);
        RSA_TOTAL(10);
        if (cn >= nchunks) break;
        c = cn;
#pragma unroll
        for (int k = 0; k < K; ++k) cur[k] = nxt[k];
    }
}

}  // namespace rsa
