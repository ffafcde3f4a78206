// Host-side op sequence of one sort (the counterpart of
// onesweep_byte_sort_enqueue, include/rsort/detail/radix_kernel.cuh:336-474):
//   clear control block -> histogram (+ lookback clear, offsets, skip mask)
//   -> one onesweep launch per pass -> device-decided copy-back.
// Every launch of a plain sort goes on the caller's stream; the segmented
// sort forks onto the library's side streams after its classification and
// joins back before it returns.  Nothing here allocates.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rsort_protocol.hpp"
#include "rsort_tuning.hpp"

namespace rsa {

struct SortJob {
    uint32_t tile_cand;   // tile_cand index of the pass kernel
    uint32_t hist_bpc;    // histogram workgroups per CU
    uint32_t kb;          // key bytes 1,2,4,8,16
    uint32_t vb;          // value bytes as seen by the kernels: 0,1,2,4,8,16
    bool wide;            // 64-bit offsets and lookback words
    bool allow_skip;      // skip passes whose digit is uniform
    bool lb_wide;         // one-round segmented lookback walk (passes of few tiles)
    bool early_agg;       // tile aggregates from digit counts taken before the rank
    bool xcd_tickets;     // per-XCD tile tickets allowed (see pass_variant for when they are used)
    uint32_t* xcd_lock;   // the device's per-XCD ticket lock word (device address; null: global ticket only)
    uint32_t hist_direct_max;  // histogram grids up to this fold by direct device atomics
    uint64_t clear_id;    // nonzero: the histogram may clear the control region itself (unique per call)
    bool late_claim;      // test setting: no workgroup claims that clearing before it needs the counts
    TwMasks tw;
    void* keys;           // caller buffer (in/out)
    void* vals;           // caller buffer (in/out) or index buffer
    void* tmp_keys;       // workspace ping-pong
    void* tmp_vals;
    Ctrl* ctrl;           // control header, followed by counts then gp
    void* counts;
    void* gp;
    uint32_t* hist_partials;    // [hist grid][passes][256]
    void* lookback;
    uint64_t lb_bytes;    // ticket counters, tile words and group words (cleared together)
    void* lb_groups;      // group words: [passes][groups][256]
    uint32_t* part_tickets;  // per-pass, per-XCD tile ticket counters (start of the lookback region)
    uint64_t n;
    uint64_t num_tiles;
    uint32_t passes;
    int num_cus;
    unsigned long long* dbg;  // phase-timing slots (RSA_PHASE_TIMING builds)
    hipEvent_t* events;   // optional: one event recorded before and after each launch
    int max_events;
    int* n_events;
};

// One segmented sort (no reference counterpart): classification, LDS bins,
// then the onesweep passes over the long segments.
struct SegJob {
    uint32_t kb;          // key bytes 1,2,4,8,16
    uint32_t vb;          // value bytes as seen by the kernels: 0,1,2,4,8,16
    bool wide;            // 64-bit offsets and lookback words
    TwMasks tw;
    void* keys;
    void* vals;
    void* tmp_keys;       // ping-pong buffers (long segments)
    void* tmp_vals;
    const void* offsets;  // num_segments + 1 entries
    uint32_t off64;       // 64-bit offsets
    uint64_t num_segments;
    uint64_t n;
    SegCtrl* ctrl;
    uint32_t* chunk;
    uint32_t* queue;
    void* desc;           // SegDesc[max_large]
    uint32_t* tile_seg;
    void* seg_gp;
    void* lookback;
    uint32_t chunks;      // ceil(num_segments / kSegChunk)
    uint64_t max_large;   // bound on long segments
    uint64_t max_tiles;   // bound on their tiles
    uint32_t lds_threshold;   // longest segment sorted in LDS (bins above it are not used)
    int num_cus;
    unsigned long long* dbg;  // phase-timing slots (RSA_PHASE_TIMING builds)
    hipEvent_t* events;
    int max_events;
    int* n_events;
    // Side streams (null: one stream): the LDS bins and the long segments
    // sort disjoint segments, so after the classification they run on the
    // caller's stream and three side streams at once, forked and joined with
    // these events (a graph capture records the same fork and join).
    hipStream_t side[3];
    hipEvent_t fork;
    hipEvent_t join[3];
};

// Requests sorted in place by the LDS bins, one workgroup each (the serving
// loop's direct form, and plain sorts small enough for one bin): reqs =
// SegReq list grouped by bin, count[b] of bin b starting at first[b].  Every
// request fits its bin (n <= bin capacity).
struct ReqJob {
    uint32_t kb, vb;
    TwMasks tw;
    const void* reqs;     // null: one request, keys/vals/n below (count 1 in its bin)
    uint32_t first[kSegBins];
    uint32_t count[kSegBins];
    void* keys;
    void* vals;
    uint64_t n;
};

// One launcher per key width, each in its own translation unit.
hipError_t launch_core_kb1(const SortJob& j, hipStream_t s);
hipError_t launch_core_kb2(const SortJob& j, hipStream_t s);
hipError_t launch_core_kb4(const SortJob& j, hipStream_t s);
hipError_t launch_core_kb8(const SortJob& j, hipStream_t s);
hipError_t launch_core_kb16(const SortJob& j, hipStream_t s);

hipError_t launch_req_kb1(const ReqJob& j, hipStream_t s);
hipError_t launch_req_kb2(const ReqJob& j, hipStream_t s);
hipError_t launch_req_kb4(const ReqJob& j, hipStream_t s);
hipError_t launch_req_kb8(const ReqJob& j, hipStream_t s);
hipError_t launch_req_kb16(const ReqJob& j, hipStream_t s);
hipError_t launch_seg_kb1(const SegJob& j, hipStream_t s);
hipError_t launch_seg_kb2(const SegJob& j, hipStream_t s);
hipError_t launch_seg_kb4(const SegJob& j, hipStream_t s);
hipError_t launch_seg_kb8(const SegJob& j, hipStream_t s);
hipError_t launch_seg_kb16(const SegJob& j, hipStream_t s);

// Helpers living in the host translation unit.
hipError_t launch_iota(uint32_t* p, uint64_t n, int num_cus, hipStream_t s);
hipError_t launch_gather(void* dst, const void* src, const uint32_t* idx, uint64_t n, uint32_t rec_bytes,
                         int num_cus, hipStream_t s);
hipError_t launch_copy(void* dst, const void* src, uint64_t bytes, int num_cus, hipStream_t s);
hipError_t launch_extract(void* dst, const void* rec, uint64_t n, uint32_t record_bytes, uint32_t key_offset,
                          uint32_t key_bytes, int num_cus, hipStream_t s);

// Resident workgroups of a kernel per CU (queried once per kernel; every
// device this library runs on is gfx950).
template <typename F>
int blocks_per_cu(F kernel, int threads) {
    int b = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, kernel, threads, 0) != hipSuccess || b < 1) b = 1;
    return b;
}

inline uint32_t grid_for(uint64_t work, uint32_t per_block, int num_cus, uint32_t blocks_per_cu) {
    uint64_t g = (work + per_block - 1) / per_block;
    const uint64_t cap = (uint64_t)(num_cus > 0 ? num_cus : 256) * blocks_per_cu;
    if (g > cap) g = cap;
    if (g < 1) g = 1;
    return (uint32_t)g;
}

template <typename Job>
inline void record_event(const Job& j, hipStream_t s) {
    if (j.events && j.n_events && *j.n_events < j.max_events) {
        (void)hipEventRecord(j.events[*j.n_events], s);
        ++*j.n_events;
    }
}

}  // namespace rsa
