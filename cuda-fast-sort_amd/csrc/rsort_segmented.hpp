// gfx950 kernels of the segmented sort: every segment [off[i], off[i+1]) of
// the key (and value) array is sorted stably and independently, in place.
// The reference has no segmented entry point; this is the batched form of its
// sort (the same stable LSD order inside each segment).
//
//   k_seg_count / k_seg_scan / k_seg_fill
//       classify every segment by length into an LDS bin (0..kSegBins-1) or
//       the onesweep category (kSegBins), and build per-category queues of segment ids
//       (in segment order) plus a descriptor and a tile range per long
//       segment -- three launches, no host round trip.
//   k_seg_lds<bin>
//       persistent: one workgroup takes the next segment of its bin from a
//       ticket (taking the following ticket while it sorts), loads the
//       segment into registers, runs every LSD pass through LDS
//       (rank in registers, exchange through one LDS staging buffer) and
//       writes it back: two HBM touches per key.
//   k_seg_large_init / k_seg_hist / k_seg_gp / k_seg_onesweep / k_seg_copy_back
//       long segments: tile table, per-segment digit counts -> offsets, then
//       the onesweep passes over all long segments' tiles at once; each
//       segment's first tile heads its own lookback chain.
#pragma once

#include "rsort_kernels.hpp"

namespace rsa {

// Descriptor of one long segment (queue order).
struct alignas(16) SegDesc {
    uint64_t begin, end;
    uint32_t first_tile, tiles;
    uint64_t seg;  // segment index
};

struct SegParams {
    void* keys;            // caller's arrays; segments index into them
    void* vals;
    void* tmp_keys;        // ping-pong buffers of the long segments
    void* tmp_vals;
    const void* offsets;   // num_segments + 1 offsets, 32- or 64-bit
    SegCtrl* ctrl;
    uint32_t* chunk;       // [chunks][kSegRow] category counts, then exclusive bases
    uint32_t* queue;       // [num_segments] segment ids grouped by category
    SegDesc* desc;         // [max_large] long segments
    uint32_t* tile_seg;    // [max_tiles] long-segment index of each tile
    void* seg_gp;          // [max_large][passes][256] Off: digit counts, then global offsets
    void* lookback;        // [max_tiles][256] Off
    uint64_t num_segments;
    uint64_t n;            // key array extent: offsets are clamped to [0, n]
    uint64_t max_large;    // capacity of desc / seg_gp
    uint64_t max_tiles;    // capacity of tile_seg / lookback
    uint32_t off64;        // offsets are 64-bit
    uint32_t passes;
    uint32_t tile;         // onesweep tile (keys)
    uint32_t pass;         // onesweep pass of this launch
    uint32_t cap[kSegBins];  // LDS bin capacities (keys)
    uint32_t pad;
    TwMasks tw;
    unsigned long long* dbg;  // phase-timing slots (RSA_PHASE_TIMING builds), else null
    const SegReq* reqs;       // direct form: request list (may be host-mapped memory)
};

// Segment s as a clamped, non-negative range [b, e) of [0, n).
__device__ __forceinline__ void seg_range(const SegParams& a, uint64_t s, uint64_t& b, uint64_t& e) {
    uint64_t lo, hi;
    if (a.off64) {
        lo = ((const uint64_t*)a.offsets)[s];
        hi = ((const uint64_t*)a.offsets)[s + 1];
    } else {
        lo = ((const uint32_t*)a.offsets)[s];
        hi = ((const uint32_t*)a.offsets)[s + 1];
    }
    lo = lo < a.n ? lo : a.n;
    hi = hi < lo ? lo : (hi < a.n ? hi : a.n);
    b = lo;
    e = hi;
}

__device__ __forceinline__ uint32_t seg_category(const SegParams& a, uint64_t m) {
    if (m <= 1u) return kSegCatNone;
#pragma unroll
    for (int b = 0; b < kSegBins; ++b)
        if (m <= a.cap[b]) return (uint32_t)b;
    return kSegCatLarge;
}

// ---------------------------------------------------------------------------
// Classification.  chunk[c*kSegRow + k]: k = 0..kSegCatLarge segments of
// category k in chunk c, k = kSegCatTiles onesweep tiles of chunk c's long
// segments.
// ---------------------------------------------------------------------------
static __global__ __launch_bounds__(kSegChunk) void k_seg_count(SegParams a) {
    __shared__ uint32_t s_cnt[kSegRow];
    const uint32_t tid = threadIdx.x;
    if (tid < kSegRow) s_cnt[tid] = 0u;
    __syncthreads();
    const uint64_t s = (uint64_t)blockIdx.x * kSegChunk + tid;
    uint32_t cat = kSegCatNone, tiles = 0;
    if (s < a.num_segments) {
        uint64_t b, e;
        seg_range(a, s, b, e);
        cat = seg_category(a, e - b);
        if (cat == kSegCatLarge) tiles = (uint32_t)((e - b + a.tile - 1u) / a.tile);
    }
#pragma unroll
    for (uint32_t c = 0; c <= kSegCatLarge; ++c) {
        const uint64_t m = __ballot(cat == c);
        if ((tid & 63u) == 0u && m) atomicAdd(&s_cnt[c], (uint32_t)__builtin_popcountll(m));
    }
    uint32_t t = tiles;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o);
    if ((tid & 63u) == 0u && t) atomicAdd(&s_cnt[kSegCatTiles], t);
    __syncthreads();
    if (tid < kSegRow) a.chunk[blockIdx.x * kSegRow + tid] = s_cnt[tid];
}

// One workgroup: wave k scans column k of the chunk table (exclusive), then
// the category totals become the queue layout.
constexpr uint32_t kSegScanThreads = 64u * (kSegCatTiles + 1u);
static __global__ __launch_bounds__(kSegScanThreads) void k_seg_scan(SegParams a, uint32_t chunks) {
    __shared__ uint32_t s_tot[kSegRow];
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t c = threadIdx.x >> 6;
    if (c <= kSegCatTiles) {
        uint32_t carry = 0;
        for (uint32_t k0 = 0; k0 < chunks; k0 += 64u) {
            const uint32_t k = k0 + lane;
            const uint32_t v = k < chunks ? a.chunk[k * kSegRow + c] : 0u;
            const uint32_t incl = wave_incl_scan<uint32_t>(v);
            if (k < chunks) a.chunk[k * kSegRow + c] = carry + incl - v;
            carry += __shfl(incl, 63);
        }
        if (lane == 0) s_tot[c] = carry;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t base = 0;
        for (uint32_t cc = 0; cc <= kSegCatLarge; ++cc) {
            a.ctrl->cat_count[cc] = s_tot[cc];
            a.ctrl->cat_base[cc] = base;
            base += s_tot[cc];
        }
        a.ctrl->large_tiles = s_tot[kSegCatTiles];
        // Disjoint segments never exceed the bounds the workspace was sized
        // for; overlapping ones can: then the long segments are left unsorted
        // (flagged) rather than written past the tables.
        if (s_tot[kSegCatLarge] > a.max_large || s_tot[kSegCatTiles] > a.max_tiles) {
            a.ctrl->overflow = 1u;
            a.ctrl->cat_count[kSegCatLarge] = 0u;
            a.ctrl->large_tiles = 0u;
        }
    }
}

// Queue slots (stable: segment order inside each category) and the long
// segments' descriptors.
static __global__ __launch_bounds__(kSegChunk) void k_seg_fill(SegParams a) {
    constexpr int WAVES = kSegChunk / 64;
    __shared__ uint32_t s_wc[WAVES][kSegRow];
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    const uint32_t w = tid >> 6;
    const uint64_t s = (uint64_t)blockIdx.x * kSegChunk + tid;
    uint32_t cat = kSegCatNone, tiles = 0;
    uint64_t b = 0, e = 0;
    if (s < a.num_segments) {
        seg_range(a, s, b, e);
        cat = seg_category(a, e - b);
        if (cat == kSegCatLarge) tiles = (uint32_t)((e - b + a.tile - 1u) / a.tile);
    }
    uint32_t rank = 0;
#pragma unroll
    for (uint32_t c = 0; c <= kSegCatLarge; ++c) {
        const uint64_t m = __ballot(cat == c);
        if (cat == c) rank = mask_below(m);
        if (lane == 0) s_wc[w][c] = (uint32_t)__builtin_popcountll(m);
    }
    const uint32_t tincl = wave_incl_scan<uint32_t>(tiles);
    if (lane == 63u) s_wc[w][kSegCatTiles] = tincl;
    __syncthreads();
    if (cat <= kSegCatLarge) {
        uint32_t pre = 0, tpre = 0;
        for (uint32_t ww = 0; ww < w; ++ww) {
            pre += s_wc[ww][cat];
            tpre += s_wc[ww][kSegCatTiles];
        }
        const uint32_t pos = a.chunk[blockIdx.x * kSegRow + cat] + pre + rank;
        a.queue[a.ctrl->cat_base[cat] + pos] = (uint32_t)s;
        if (cat == kSegCatLarge && pos < a.max_large) {
            SegDesc d;
            d.begin = b;
            d.end = e;
            d.first_tile = a.chunk[blockIdx.x * kSegRow + kSegCatTiles] + tpre + tincl - tiles;
            d.tiles = tiles;
            d.seg = s;
            a.desc[pos] = d;
        }
    }
}

// ---------------------------------------------------------------------------
// LDS bins: one workgroup sorts one segment of at most THREADS*K keys.
// ---------------------------------------------------------------------------
// DIRECT: workgroup b sorts request reqs[bin + b] (grid = the bin's request
// count), or, without a request list, the one array keys/vals of n keys
// count); otherwise a persistent workgroup takes segments of bin `bin` by ticket.
template <int KB, int VB, int THREADS, int K, bool DIRECT>
__global__ __launch_bounds__(THREADS) void k_seg_lds(SegParams a, uint32_t bin) {
    using KM = typename KeyTraits<KB>::Mem;
    using KR = typename KeyTraits<KB>::Reg;
    using VM = typename ValTraits<VB>::Mem;
    constexpr int WAVES = THREADS / 64;
    constexpr uint32_t TILE = (uint32_t)THREADS * K;
    constexpr size_t KSTAGE = (size_t)TILE * sizeof(KM);
    constexpr size_t VSTAGE = VB ? (size_t)TILE * sizeof(VM) : 0;
    constexpr size_t MATCH = (size_t)WAVES * 256 * sizeof(uint64_t);
#ifndef RSA_SEG_PAIR_XCHG
#define RSA_SEG_PAIR_XCHG 1
#endif
    // 4-byte keys with 4-byte values: one 64-bit exchange moves both, with
    // half the exchange instructions and barriers (twice the staging LDS).
    // Bins of more threads keep keys and values apart: their staging LDS then
    // leaves room for other bins' workgroups on the CU (measured on MI355X,
    // the 100k-segment batch: paired up to 1024 threads 102.3 GKeys/s, up to
    // 384 103.3-104.2, up to 256 104.5-105.6).
#ifndef RSA_SEG_PAIR_MAX_THREADS
#define RSA_SEG_PAIR_MAX_THREADS 256
#endif
    constexpr bool PAIR = RSA_SEG_PAIR_XCHG && KB == 4 && VB == 4 && THREADS <= RSA_SEG_PAIR_MAX_THREADS;
    constexpr size_t BIG0 = KSTAGE > VSTAGE ? (KSTAGE > MATCH ? KSTAGE : MATCH) : (VSTAGE > MATCH ? VSTAGE : MATCH);
    constexpr size_t BIG = PAIR ? (KSTAGE + VSTAGE > MATCH ? KSTAGE + VSTAGE : MATCH) : BIG0;
    // digits owned per thread in the digit scan, and how many threads own digits
    constexpr int R = THREADS >= 256 ? 1 : 256 / THREADS;
    constexpr uint32_t OWNERS = THREADS >= 256 ? 256u : (uint32_t)THREADS;
    static_assert(THREADS % 64 == 0 && THREADS <= 1024, "workgroup of whole waves");

    __shared__ __attribute__((aligned(16))) uint8_t s_big[BIG];
    // Per-wave digit counters, then the waves' digit starts: both below TILE,
    // so 16 bits hold them for bins under 64K keys (half the LDS bytes of the
    // rank's counter reads and writes and of the exchange's offset reads;
    // 100k-segment batch +1.4 % on MI355X)
#ifndef RSA_SEG_CNT16
#define RSA_SEG_CNT16 1
#endif
    using SCnt = std::conditional_t<(RSA_SEG_CNT16 && TILE < 65536u), uint16_t, uint32_t>;
    __shared__ __attribute__((aligned(16))) SCnt s_cnt[WAVES * 256];
    __shared__ uint32_t s_wsum[4];
    __shared__ uint32_t s_item[2];  // by segment parity: this segment's ticket, the next one's
    __shared__ uint32_t s_uniform[2];  // by pass parity: reset two passes after its last read

    const uint32_t count = DIRECT ? 0u : a.ctrl->cat_count[bin];
    const uint32_t* __restrict__ q = DIRECT ? nullptr : a.queue + a.ctrl->cat_base[bin];
    RSA_TIMER_START(a.dbg);
    uint32_t next = 0;  // thread 0: the ticket of the workgroup's next segment
    if constexpr (!DIRECT) {
        if (threadIdx.x == 0) s_item[0] = atomicAdd(&a.ctrl->bin_ticket[bin], 1u);
        __syncthreads();
    }

    for (uint32_t round = 0;; ++round) {
        // the thread index is re-read opaquely each segment, so that the
        // compiler keeps no per-item indexes live across segments
        uint32_t tid = threadIdx.x;
        if constexpr (!DIRECT) asm volatile("" : "+v"(tid));
        const uint32_t lane = tid & 63u;
        const uint32_t w = tid >> 6;
        const uint32_t wl = w * 64u * K + lane;  // tile-local index of item 0
        KM* __restrict__ keys;
        VM* __restrict__ vals;
        uint64_t sb;
        uint32_t m;
        if constexpr (DIRECT) {
            if (round) break;
            if (a.reqs) {
                const SegReq r = a.reqs[bin + blockIdx.x];
                keys = (KM*)r.keys;
                vals = (VM*)r.vals;
                m = (uint32_t)r.n;
            } else {
                // no request list: one array of a.n keys (a plain sort that fits the bin)
                keys = (KM*)a.keys;
                vals = (VM*)a.vals;
                m = (uint32_t)a.n;
            }
            sb = 0;
        } else {
            // uniform: the segment's range and base pointers live in scalar registers
            const uint32_t item = __builtin_amdgcn_readfirstlane(s_item[round & 1u]);
            if (item >= count) break;
            uint64_t se;
            seg_range(a, q[item], sb, se);
            m = (uint32_t)(se - sb);
            keys = (KM*)a.keys + sb;
            vals = VB > 0 ? (VM*)a.vals + sb : nullptr;
        }
        RSA_MARK(22);

        // unconditional loads (slots past the end re-read the last key) so
        // that all of them are in flight at once; fillers are chosen after
        KR key[K];
        VM val[VB > 0 ? K : 1];
#pragma unroll
        for (int i = 0; i < K; ++i) {
            const uint32_t idx = wl + (uint32_t)i * 64u;
            const uint32_t at = idx < m ? idx : m - 1u;
            key[i] = key_to_reg<KB>(keys[at]);
            if constexpr (VB > 0) val[i] = vals[at];
        }
#pragma unroll
        for (int i = 0; i < K; ++i) {
            const uint32_t idx = wl + (uint32_t)i * 64u;
            key[i] = idx < m ? tw_in<KB>(key[i], a.tw) : key_filler<KB>();
        }
#ifdef RSA_PHASE_TIMING
        __builtin_amdgcn_s_waitcnt(0);
#endif
        RSA_MARK(23);
        // the next segment's ticket: its round trip overlaps this segment's passes
        if (!DIRECT && tid == 0) next = atomicAdd(&a.ctrl->bin_ticket[bin], 1u);

        for (uint32_t p = 0; p < (uint32_t)KB; ++p) {
            // the previous exchange has been read back (barrier at its end)
            {
                uint64_t* M = (uint64_t*)s_big + (size_t)w * 256;
                for (uint32_t k = lane; k < 256u; k += 64u) {
                    M[k] = 0ull;
                    s_cnt[w * 256 + k] = (SCnt)0u;
                }
            }
            RSA_COMPILER_FENCE();  // own-wave tables: the LDS keeps a wave's order
            if (tid == 0) s_uniform[p & 1u] = 0u;
            uint32_t dr[K];
            rank_wave<KB, K>(key, 8u * (p & 3u), p >> 2, (uint64_t*)s_big + (size_t)w * 256, s_cnt + w * 256, lane,
                             dr);
            __syncthreads();
            RSA_MARK(24);

            // digit scan: per digit, exclusive over waves; then exclusive over
            // digits; s_cnt[w][d] becomes the start of wave w's run of digit d
            // the column's exclusive prefix stays in registers until the base is known
            uint32_t tot[R], col[R][WAVES], incl = 0, local = 0;
            if (tid < OWNERS) {
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const uint32_t d = tid * R + r;
                    uint32_t s = 0;
#pragma unroll
                    for (int ww = 0; ww < WAVES; ++ww) {
                        col[r][ww] = s;
                        s += s_cnt[ww * 256 + d];
                    }
                    tot[r] = s;
                    local += s;
                    if (s == TILE) s_uniform[p & 1u] = 1u;
                }
                incl = wave_incl_scan<uint32_t>(local);
                if (lane == 63u) s_wsum[w] = incl;
            }
            __syncthreads();
            if (tid < OWNERS) {
                uint32_t base = incl - local;
                for (uint32_t ww = 0; ww < w; ++ww) base += s_wsum[ww];
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const uint32_t d = tid * R + r;
#pragma unroll
                    for (int ww = 0; ww < WAVES; ++ww) s_cnt[ww * 256 + d] = (SCnt)(base + col[r][ww]);
                    base += tot[r];
                }
            }
            __syncthreads();
            RSA_MARK(25);
            if (s_uniform[p & 1u]) continue;  // one digit: this pass keeps the order
            if constexpr (PAIR) {
                uint64_t* SP = (uint64_t*)s_big;
#pragma unroll
                for (int i = 0; i < K; ++i) {
                    const uint32_t at = s_cnt[w * 256 + (dr[i] & 255u)] + (dr[i] >> 8);
                    SP[at] = (uint64_t)(uint32_t)key_to_mem<KB>(key[i]) | ((uint64_t)(uint32_t)val[i] << 32);
                }
                __syncthreads();
#pragma unroll
                for (int i = 0; i < K; ++i) {
                    const uint64_t x = SP[wl + (uint32_t)i * 64u];
                    key[i] = key_to_reg<KB>((KM)(uint32_t)x);
                    val[i] = (VM)(uint32_t)(x >> 32);
                }
                __syncthreads();
                continue;
            }
            KM* SK = (KM*)s_big;
#pragma unroll
            for (int i = 0; i < K; ++i) {
                const uint32_t at = s_cnt[w * 256 + (dr[i] & 255u)] + (dr[i] >> 8);
                SK[at] = key_to_mem<KB>(key[i]);
                if constexpr (VB > 0) dr[i] = at;  // the value goes to the same slot
            }
            __syncthreads();
#pragma unroll
            for (int i = 0; i < K; ++i) key[i] = key_to_reg<KB>(SK[wl + (uint32_t)i * 64u]);
            RSA_MARK(26);
            if constexpr (VB > 0) {
                __syncthreads();
                VM* SV = (VM*)s_big;
#pragma unroll
                for (int i = 0; i < K; ++i) SV[dr[i]] = val[i];
                __syncthreads();
#pragma unroll
                for (int i = 0; i < K; ++i) val[i] = SV[wl + (uint32_t)i * 64u];
            }
            __syncthreads();
            RSA_MARK(27);
        }

#pragma unroll
        for (int i = 0; i < K; ++i) {
            const uint32_t idx = wl + (uint32_t)i * 64u;
            if (idx < m) {
                keys[idx] = key_to_mem<KB>(tw_out<KB>(key[i], a.tw));
                if constexpr (VB > 0) vals[idx] = val[i];
            }
        }
        if (!DIRECT && tid == 0) s_item[(round + 1u) & 1u] = next;
        // s_item and the LDS tables are reused by the next segment
        __syncthreads();
        RSA_MARK(28);
        if (tid == 0) RSA_COUNT(29, 1u);
    }
    RSA_TOTAL(30);
}

// ---------------------------------------------------------------------------
// Long segments.
// ---------------------------------------------------------------------------
// Tile table and zeroed digit counts, one workgroup per long segment.
template <bool WIDE>
__global__ __launch_bounds__(256) void k_seg_large_init(SegParams a) {
    using Off = std::conditional_t<WIDE, uint64_t, uint32_t>;
    const uint32_t nl = a.ctrl->cat_count[kSegCatLarge];
    for (uint32_t li = blockIdx.x; li < nl; li += gridDim.x) {
        const SegDesc d = a.desc[li];
        for (uint32_t k = threadIdx.x; k < d.tiles; k += 256u) a.tile_seg[d.first_tile + k] = li;
        Off* cnt = (Off*)a.seg_gp + (size_t)li * a.passes * 256u;
        for (uint32_t k = threadIdx.x; k < a.passes * 256u; k += 256u) cnt[k] = 0;
    }
}

// Digit counts of every pass per long segment.  Each workgroup takes a
// contiguous range of tiles, counts in LDS and flushes with device atomics
// when its range moves to the next segment.  Also zeroes the tiles' lookback
// rows.
template <int KB, bool WIDE>
__global__ __launch_bounds__(512) void k_seg_hist(SegParams a) {
    using KM = typename KeyTraits<KB>::Mem;
    using Off = std::conditional_t<WIDE, uint64_t, uint32_t>;
    constexpr int P = KB;
    __shared__ uint32_t H[P * 256];
    const uint32_t tid = threadIdx.x;
    const uint32_t nt = a.ctrl->large_tiles;
    const uint32_t per = (nt + gridDim.x - 1u) / gridDim.x;
    const uint32_t t0 = blockIdx.x * per;
    const uint32_t t1 = t0 + per < nt ? t0 + per : nt;
    if (t0 >= t1) return;
    const KM* __restrict__ keys = (const KM*)a.keys;
    Off* lb = (Off*)a.lookback;

    for (uint32_t k = tid; k < P * 256u; k += 512u) H[k] = 0u;
    uint32_t li = a.tile_seg[t0];
    SegDesc d = a.desc[li];
    __syncthreads();
    for (uint32_t t = t0; t < t1; ++t) {
        if (t >= d.first_tile + d.tiles) {
            __syncthreads();
            Off* cnt = (Off*)a.seg_gp + (size_t)li * P * 256u;
            for (uint32_t k = tid; k < P * 256u; k += 512u) {
                if (H[k]) atomicAdd(cnt + k, (Off)H[k]);
                H[k] = 0u;
            }
            ++li;  // tiles of consecutive long segments are consecutive
            d = a.desc[li];
            __syncthreads();
        }
        if (tid < 256u) lb[(size_t)t * 256u + tid] = 0;
        const uint64_t tb = d.begin + (uint64_t)(t - d.first_tile) * a.tile;
        const uint64_t te = tb + a.tile < d.end ? tb + a.tile : d.end;
        for (uint64_t i = tb + tid; i < te; i += 512u) {
            const auto k = tw_in<KB>(key_to_reg<KB>(keys[i]), a.tw);
#pragma unroll
            for (int p = 0; p < P; ++p) atomicAdd(&H[p * 256 + key_digit(k, 8u * (p & 3), (uint32_t)p >> 2)], 1u);
        }
    }
    __syncthreads();
    Off* cnt = (Off*)a.seg_gp + (size_t)li * P * 256u;
    for (uint32_t k = tid; k < P * 256u; k += 512u)
        if (H[k]) atomicAdd(cnt + k, (Off)H[k]);
}

// Counts -> global exclusive digit offsets (segment begin included), in place.
template <bool WIDE>
__global__ __launch_bounds__(256) void k_seg_gp(SegParams a) {
    using Off = std::conditional_t<WIDE, uint64_t, uint32_t>;
    __shared__ Off s_wsum[4];
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    const uint32_t w = tid >> 6;
    const uint32_t nl = a.ctrl->cat_count[kSegCatLarge];
    for (uint32_t li = blockIdx.x; li < nl; li += gridDim.x) {
        const uint64_t begin = a.desc[li].begin;
        Off* g = (Off*)a.seg_gp + (size_t)li * a.passes * 256u;
        for (uint32_t p = 0; p < a.passes; ++p) {
            const Off c = g[p * 256u + tid];
            const Off incl = wave_incl_scan<Off>(c);
            if (lane == 63u) s_wsum[w] = incl;
            __syncthreads();
            Off off = (Off)begin + incl - c;
            for (uint32_t ww = 0; ww < w; ++ww) off += s_wsum[ww];
            g[p * 256u + tid] = off;
            __syncthreads();
        }
    }
}

// One onesweep pass over the tiles of every long segment.  Persistent: each
// workgroup takes tiles by ticket until none is left; a segment's first tile
// publishes an inclusive prefix at once, so no lookback crosses segments.
template <int KB, int VB, bool WIDE, int THREADS, int K>
__global__ __launch_bounds__(THREADS) void k_seg_onesweep(SegParams a) {
    using KM = typename KeyTraits<KB>::Mem;
    using KR = typename KeyTraits<KB>::Reg;
    using VM = typename ValTraits<VB>::Mem;
    using Off = std::conditional_t<WIDE, uint64_t, uint32_t>;
    constexpr int WAVES = THREADS / 64;
    constexpr uint32_t TILE = (uint32_t)THREADS * K;
    constexpr int LB_W0 = RSA_LB_W0;
    constexpr int LB_WMAX = RSA_LB_WMAX;
    constexpr bool LB_GROUPS = false;  // each segment heads its own chain
    Off* lbg = nullptr;
    constexpr bool LB_SEGWALK = false;
    constexpr bool EARLY_AGG = false;
    __shared__ OnesweepLds<KB, VB, WIDE, THREADS, K> L;

    const uint32_t pass = a.pass;
    const bool odd = (pass & 1u) != 0u;
    PassIo<KB, VB> io;
    io.kin = (const KM*)(odd ? a.tmp_keys : a.keys);
    io.kout = (KM*)(odd ? a.keys : a.tmp_keys);
    io.vin = (const VM*)(odd ? a.tmp_vals : a.vals);
    io.vout = (VM*)(odd ? a.vals : a.tmp_vals);
    io.sh = 8u * (pass & 3u);
    io.wsel = pass >> 2;
    io.epoch = pass & 1u;
    io.first = pass == 0u && a.tw.active;
    io.last = pass + 1u == a.passes && a.tw.active;
    Off* lb = (Off*)a.lookback;
    const TwMasks& tw = a.tw;
    const uint32_t nt = a.ctrl->large_tiles;
    unsigned long long* const no_dbg = nullptr;
    (void)no_dbg;
    RSA_TIMER_START(no_dbg);
#ifdef RSA_PHASE_TIMING
    if (threadIdx.x < 2) L.dbg[threadIdx.x] = 0u;
#endif

    for (;;) {
        const uint32_t t = onesweep_begin(L, &a.ctrl->ticket[pass]);
        if (t >= nt) break;
        const uint32_t li = a.tile_seg[t];
        const SegDesc sd = a.desc[li];
        TileGeom g;
        g.lb_tile = t;
        g.head = t == sd.first_tile;
        g.tbase = sd.begin + (uint64_t)(t - sd.first_tile) * TILE;
        g.kend = sd.end;
        const Off* gp_row = (const Off*)a.seg_gp + ((size_t)li * a.passes + pass) * 256u;
        uint32_t tile_tid = threadIdx.x;
        asm volatile("" : "+v"(tile_tid));
#include "rsort_onesweep_tile.inc"
        __syncthreads();  // the next tile reuses the LDS
    }
}

// Odd pass count: the long segments' results are in the ping-pong buffers.
template <typename E>
__global__ void k_seg_copy_back(SegParams a, E* __restrict__ dst, const E* __restrict__ src) {
    const uint32_t nl = a.ctrl->cat_count[kSegCatLarge];
    for (uint32_t li = blockIdx.y; li < nl; li += gridDim.y) {
        const SegDesc d = a.desc[li];
        for (uint64_t i = d.begin + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < d.end;
             i += (uint64_t)gridDim.x * blockDim.x)
            dst[i] = src[i];
    }
}

}  // namespace rsa
