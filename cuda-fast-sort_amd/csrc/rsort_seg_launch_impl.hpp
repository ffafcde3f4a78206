// Launcher of the segmented sort, instantiated once per key width (seg_kb*.hip).
#pragma once

#include "rsort_launch.hpp"
#include "rsort_segmented.hpp"

namespace rsa {

template <int KB, int VB, int BIN>
hipError_t launch_seg_bin(const SegParams& a, const SegJob& j, hipStream_t s) {
    constexpr int THREADS = (int)seg_bin_threads(KB, VB, BIN);
    constexpr int K = (int)seg_bin_items(KB, VB, BIN);
    static const int occ = blocks_per_cu(k_seg_lds<KB, VB, THREADS, K, false>, THREADS);
    uint64_t grid = (uint64_t)j.num_cus * (uint64_t)occ;
    if (grid > j.num_segments) grid = j.num_segments;
    hipLaunchKernelGGL((k_seg_lds<KB, VB, THREADS, K, false>), dim3((uint32_t)grid), dim3(THREADS), 0, s, a,
                       (uint32_t)BIN);
    return hipGetLastError();
}

// Direct form: count[b] requests of bin b at reqs[first[b] ..], one
// workgroup each.
template <int KB, int VB, int BIN>
void launch_req_bin(const SegParams& a, const uint32_t* first, const uint32_t* count, hipStream_t s) {
    constexpr int THREADS = (int)seg_bin_threads(KB, VB, BIN);
    constexpr int K = (int)seg_bin_items(KB, VB, BIN);
    if (count[BIN])
        hipLaunchKernelGGL((k_seg_lds<KB, VB, THREADS, K, true>), dim3(count[BIN]), dim3(THREADS), 0, s, a,
                           first[BIN]);
}

template <int KB, int VB>
hipError_t launch_req_v(const ReqJob& j, hipStream_t s) {
    SegParams a{};
    a.reqs = (const SegReq*)j.reqs;
    a.keys = j.keys;
    a.vals = j.vals;
    a.n = j.n;
    a.tw = j.tw;
    static_assert(kSegBins == 9, "one launch per bin below");
    launch_req_bin<KB, VB, 0>(a, j.first, j.count, s);
    launch_req_bin<KB, VB, 1>(a, j.first, j.count, s);
    launch_req_bin<KB, VB, 2>(a, j.first, j.count, s);
    launch_req_bin<KB, VB, 3>(a, j.first, j.count, s);
    launch_req_bin<KB, VB, 4>(a, j.first, j.count, s);
    launch_req_bin<KB, VB, 5>(a, j.first, j.count, s);
    launch_req_bin<KB, VB, 6>(a, j.first, j.count, s);
    launch_req_bin<KB, VB, 7>(a, j.first, j.count, s);
    launch_req_bin<KB, VB, 8>(a, j.first, j.count, s);
    return hipGetLastError();
}

template <int KB>
hipError_t launch_req(const ReqJob& j, hipStream_t s) {
    switch (j.vb) {
        case 0: return launch_req_v<KB, 0>(j, s);
        case 1: return launch_req_v<KB, 1>(j, s);
        case 2: return launch_req_v<KB, 2>(j, s);
        case 4: return launch_req_v<KB, 4>(j, s);
        case 8: return launch_req_v<KB, 8>(j, s);
        case 16: return launch_req_v<KB, 16>(j, s);
        default: return hipErrorInvalidValue;
    }
}

template <int KB, int VB, bool WIDE>
hipError_t launch_seg_long(SegParams a, const SegJob& j, hipStream_t s);

template <int KB, int VB, bool WIDE>
hipError_t launch_seg_t(const SegJob& j, hipStream_t s) {
    constexpr TileCfg C = tile_cfg(KB, VB);
    constexpr uint32_t P = KB;

    SegParams a{};
    a.keys = j.keys;
    a.vals = j.vals;
    a.tmp_keys = j.tmp_keys;
    a.tmp_vals = j.tmp_vals;
    a.offsets = j.offsets;
    a.ctrl = j.ctrl;
    a.chunk = j.chunk;
    a.queue = j.queue;
    a.desc = (SegDesc*)j.desc;
    a.tile_seg = j.tile_seg;
    a.seg_gp = j.seg_gp;
    a.lookback = j.lookback;
    a.num_segments = j.num_segments;
    a.n = j.n;
    a.max_large = j.max_large;
    a.max_tiles = j.max_tiles;
    a.off64 = j.off64;
    a.passes = P;
    a.tile = C.threads * C.items;
    // bins whose capacity exceeds the LDS threshold are not used: their
    // segments go to the onesweep path
    for (int b = 0; b < kSegBins; ++b) a.cap[b] = seg_cap(KB, VB, b) <= j.lds_threshold ? seg_cap(KB, VB, b) : 0u;
    a.tw = j.tw;
    a.dbg = j.dbg;

    // 1. classification and queues
    record_event(j, s);
    hipLaunchKernelGGL(k_fill_zero, dim3(1), dim3(64), 0, s, (uint32_t*)j.ctrl, (uint32_t)(sizeof(SegCtrl) / 4u));
    hipLaunchKernelGGL(k_seg_count, dim3(j.chunks), dim3(kSegChunk), 0, s, a);
    hipLaunchKernelGGL(k_seg_scan, dim3(1), dim3(kSegScanThreads), 0, s, a, j.chunks);
    hipLaunchKernelGGL(k_seg_fill, dim3(j.chunks), dim3(kSegChunk), 0, s, a);
    record_event(j, s);

    // 2. LDS bins and long segments; with side streams they (disjoint
    //    segments) run at once, so that one kernel's tail and spare CU
    //    resources are filled by another's workgroups
    const bool fork = j.side[0] && !j.events;
    hipStream_t sb[3] = {s, s, s};
    hipError_t e = hipSuccess;
    if (fork) {
        e = hipEventRecord(j.fork, s);
        for (int i = 0; i < 3 && e == hipSuccess; ++i) {
            e = hipStreamWaitEvent(j.side[i], j.fork, 0);
            sb[i] = j.side[i];
        }
        if (e != hipSuccess) return e;
    }
    // Four streams, the longest-running bins first on separate streams, the
    // long segments behind the short bin 7 (measured spans when run side by
    // side: bin 5 ~2.0 ms, bin 8 ~1.1 ms, bin 4 ~0.9 ms, bins 0-3, 6 and 7
    // 0.25-0.6 ms, the long segments ~0.5 ms).  Measured on MI355X, 100k
    // lognormal segments, u32/u32 (GKeys/s): nine bins with 3,0 | 5,8 | 4,2
    // | 7,6,1,long 100.4; 7 | 5,long | 4,3,1 | 0,2,8,6 102.3, 103.3 with the
    // large bins exchanging keys and values separately; this one 105.0
    // (105.4-105.9 in a later same-box A/B, where bin 6 behind bin 5 on the
    // first side stream measured 105.1-105.6).
    // (Five bins of capacity 1024 << b: one stream 92.0; 4,3 | 2 | 1,0 |
    // long 94.2; 4,1 | 3 | 2,0 | long 96.9; 4,0 | 3 | 2 | 1,long 100.3; the
    // long segments first on their own stream 95-100.)
    static_assert(kSegBins == 9, "one launch per bin below");
    e = launch_seg_bin<KB, VB, 7>(a, j, s);
    if (e == hipSuccess) e = launch_seg_bin<KB, VB, 5>(a, j, sb[0]);
    if (e == hipSuccess) e = launch_seg_bin<KB, VB, 4>(a, j, sb[1]);
    if (e == hipSuccess) e = launch_seg_bin<KB, VB, 0>(a, j, sb[2]);
    if (e == hipSuccess) e = launch_seg_bin<KB, VB, 3>(a, j, sb[1]);
    if (e == hipSuccess) e = launch_seg_bin<KB, VB, 2>(a, j, sb[2]);
    if (e == hipSuccess) e = launch_seg_bin<KB, VB, 8>(a, j, sb[2]);
    if (e == hipSuccess) e = launch_seg_bin<KB, VB, 1>(a, j, sb[1]);
    record_event(j, s);
    if (e == hipSuccess) e = launch_seg_long<KB, VB, WIDE>(a, j, s);
    if (e == hipSuccess) e = launch_seg_bin<KB, VB, 6>(a, j, s);
    if (fork) {
        for (int i = 0; i < 3 && e == hipSuccess; ++i) {
            e = hipEventRecord(j.join[i], j.side[i]);
            if (e == hipSuccess) e = hipStreamWaitEvent(s, j.join[i], 0);
        }
    }
    return e;
}

// 3. long segments: tile table, digit offsets, onesweep passes
template <int KB, int VB, bool WIDE>
hipError_t launch_seg_long(SegParams a, const SegJob& j, hipStream_t s) {
    using KM = typename KeyTraits<KB>::Mem;
    using VM = typename ValTraits<VB>::Mem;
    constexpr TileCfg C = tile_cfg(KB, VB);
    constexpr uint32_t P = KB;
    if (j.max_large == 0) return hipGetLastError();
    const uint32_t gl = grid_for(j.max_large, 1, j.num_cus, 4);
    hipLaunchKernelGGL((k_seg_large_init<WIDE>), dim3(gl), dim3(256), 0, s, a);
    hipLaunchKernelGGL((k_seg_hist<KB, WIDE>), dim3(grid_for(j.max_tiles, 1, j.num_cus, 2)), dim3(512), 0, s, a);
    hipLaunchKernelGGL((k_seg_gp<WIDE>), dim3(gl), dim3(256), 0, s, a);
    record_event(j, s);
    static const int occ = blocks_per_cu(k_seg_onesweep<KB, VB, WIDE, C.threads, C.items>, C.threads);
    const uint32_t go = grid_for(j.max_tiles, 1, j.num_cus, (uint32_t)occ);
    for (uint32_t p = 0; p < P; ++p) {
        a.pass = p;
        hipLaunchKernelGGL((k_seg_onesweep<KB, VB, WIDE, C.threads, C.items>), dim3(go), dim3(C.threads), 0, s, a);
        record_event(j, s);
    }
    if (P & 1u) {
        const dim3 g(8, grid_for(j.max_large, 1, j.num_cus, 4));
        hipLaunchKernelGGL((k_seg_copy_back<KM>), g, dim3(256), 0, s, a, (KM*)j.keys, (const KM*)j.tmp_keys);
        if constexpr (VB > 0)
            hipLaunchKernelGGL((k_seg_copy_back<VM>), g, dim3(256), 0, s, a, (VM*)j.vals, (const VM*)j.tmp_vals);
        record_event(j, s);
    }
    return hipGetLastError();
}

template <int KB, int VB>
hipError_t launch_seg_v(const SegJob& j, hipStream_t s) {
    return j.wide ? launch_seg_t<KB, VB, true>(j, s) : launch_seg_t<KB, VB, false>(j, s);
}

template <int KB>
hipError_t launch_seg(const SegJob& j, hipStream_t s) {
    switch (j.vb) {
        case 0: return launch_seg_v<KB, 0>(j, s);
        case 1: return launch_seg_v<KB, 1>(j, s);
        case 2: return launch_seg_v<KB, 2>(j, s);
        case 4: return launch_seg_v<KB, 4>(j, s);
        case 8: return launch_seg_v<KB, 8>(j, s);
        case 16: return launch_seg_v<KB, 16>(j, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace rsa
