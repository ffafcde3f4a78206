// Launcher template instantiated once per key width (sort_kb*.hip).
#pragma once

#include "rsort_kernels.hpp"
#include "rsort_launch.hpp"

namespace rsa {

// The pass kernel variant of a sort, chosen once before the histogram (which
// takes the device's per-XCD ticket lock only when the variant may use it).
struct PassKernel {
    const void* fn;
    uint32_t threads;
    uint32_t grid;
    bool xcd_ok;
};

// Per-XCD tile tickets make a workgroup wait, in the worst case, on tiles
// held by the next kTicketParts * C workgroups in dispatch order, which must
// then be able to be resident at once: they are allowed only on a whole
// 8-XCD device (256 CUs) where the kernel's resident workgroups number at
// least twice that, and used only by the sort that holds the device's lock
// (k_histogram); otherwise the single global ticket, which never waits on a
// tile not yet handed out.
template <int KB, int VB, bool WIDE, int THREADS, int K, bool SEG, bool EARLY>
PassKernel pass_variant(const SortJob& j) {
    constexpr uint32_t C = tile_chunk(KB);
    static const int occ = blocks_per_cu(k_onesweep<KB, VB, WIDE, THREADS, K, SEG, EARLY>, THREADS);
    PassKernel pk;
    pk.fn = (const void*)k_onesweep<KB, VB, WIDE, THREADS, K, SEG, EARLY>;
    pk.threads = THREADS;
    pk.xcd_ok = C > 0 && j.xcd_tickets && j.xcd_lock && j.num_cus >= 256 &&
                (uint64_t)occ * (uint64_t)j.num_cus >= 2ull * kTicketParts * C;
    // whole rounds of the per-XCD ticket counters (extra workgroups end at once)
    const uint64_t round = pk.xcd_ok ? (uint64_t)kTicketParts * C : 1u;
    pk.grid = (uint32_t)((j.num_tiles + round - 1) / round * round);
    return pk;
}

// Passes of few tiles take the variant with the segmented lookback walk,
// large u32/u32 passes the one that publishes tile aggregates before the rank.
template <int KB, int VB, bool WIDE, int THREADS, int K>
PassKernel pass_variant_of(const SortJob& j) {
    if constexpr (!WIDE && THREADS >= 512) {
        if (j.lb_wide) return pass_variant<KB, VB, WIDE, THREADS, K, true, false>(j);
    }
    if constexpr (KB == 4 && VB == 4 && !WIDE) {
        if (j.early_agg) return pass_variant<KB, VB, WIDE, THREADS, K, false, true>(j);
    }
    return pass_variant<KB, VB, WIDE, THREADS, K, false, false>(j);
}

// The tile candidate the tuning table chose (tile_cand).
template <int KB, int VB, bool WIDE>
PassKernel choose_pass_kernel(const SortJob& j) {
    constexpr TileCfg C = tile_cfg(KB, VB);
    if constexpr (tile_cands_built(KB, VB)) {
        constexpr TileCfg C1 = tile_cand(KB, VB, 1), C2 = tile_cand(KB, VB, 2), C3 = tile_cand(KB, VB, 3);
        if (j.tile_cand == 1) return pass_variant_of<KB, VB, WIDE, C1.threads, C1.items>(j);
        if (j.tile_cand == 2) return pass_variant_of<KB, VB, WIDE, C2.threads, C2.items>(j);
        if (j.tile_cand == 3) return pass_variant_of<KB, VB, WIDE, C3.threads, C3.items>(j);
    }
    return pass_variant_of<KB, VB, WIDE, C.threads, C.items>(j);
}

template <int KB, int VB, bool WIDE>
hipError_t launch_core_t(const SortJob& j, hipStream_t s) {
    using Off = std::conditional_t<WIDE, uint64_t, uint32_t>;
    using KM = typename KeyTraits<KB>::Mem;
    using VM = typename ValTraits<VB>::Mem;
    constexpr uint32_t P = KB;

    record_event(j, s);
    // 1. control header, global digit counts, tickets and lookback words
    //    (flag 0: never ready in any epoch), contiguous in the workspace: one
    //    fill, or none when the histogram clears them itself (a single
    //    workgroup, or, given a call id, a region of at most
    //    kSelfClearMaxBytes; the fill stays for larger regions, where its
    //    cost is lost in the sort, and for captured launches, whose replays
    //    would repeat one id)
    const uint64_t head_units = (uint64_t)((sizeof(Ctrl) + (size_t)P * kRadix * sizeof(Off)) / 16u);
    const uint64_t clear_units = head_units + (uint64_t)(j.lb_bytes / 16u);
    const uint32_t hg = grid_for(j.n, hist_keys_per_block_iter(KB), j.num_cus, j.hist_bpc ? j.hist_bpc : 1u);
    const uint32_t hgrid = hg < kHistMaxGrid ? hg : kHistMaxGrid;
    const bool self_clear = hgrid == 1u || (j.clear_id != 0u && clear_units * 16u <= kSelfClearMaxBytes);
    if (!self_clear)
        hipLaunchKernelGGL(k_fill_zero16, dim3(grid_for(clear_units, 256u * 8u, j.num_cus, 4)), dim3(256), 0, s,
                           (uint4*)j.ctrl, clear_units);
    record_event(j, s);

    // 2. histogram of all passes, offsets, skip mask
    HistParams h{};
    h.keys = j.keys;
    h.partials = j.hist_partials;
    h.counts = j.counts;
    h.gp = j.gp;
    h.ctrl = j.ctrl;
    h.n = j.n;
    {
        const uintptr_t addr = (uintptr_t)j.keys;
        uint64_t head = ((16u - (addr & 15u)) & 15u) / KB;
        if (head > j.n) head = j.n;
        h.head = head;
    }
    h.passes = P;
    h.allow_skip = j.allow_skip ? 1u : 0u;
    h.direct = (hgrid > (uint32_t)kHistGroup && hgrid <= j.hist_direct_max) ? 1u : 0u;
    h.self_clear = self_clear ? (uint4*)j.ctrl : nullptr;
    h.self_clear_units = self_clear ? clear_units : 0u;
    h.head_units = head_units;
    h.clear_id = self_clear && hgrid > 1u ? j.clear_id : 0u;
    h.early_claim = j.late_claim ? 0u : 1u;
    const PassKernel pk = choose_pass_kernel<KB, VB, WIDE>(j);
    h.xcd_lock = pk.xcd_ok ? j.xcd_lock : nullptr;
    h.want_xcd = pk.xcd_ok ? 1u : 0u;
    h.tw = j.tw;
    h.dbg = j.dbg;
    hipLaunchKernelGGL((k_histogram<KB, WIDE>), dim3(hgrid), dim3(kHistThreads), 0, s, h);
    record_event(j, s);

    // 3. passes
    PassParams pp{};
    pp.keys_a = j.keys;
    pp.keys_b = j.tmp_keys;
    pp.vals_a = j.vals;
    pp.vals_b = j.tmp_vals;
    pp.gp = j.gp;
    pp.lookback = j.lookback;
    pp.lb_groups = j.lb_groups;
    pp.ctrl = j.ctrl;
    pp.n = j.n;
    pp.num_tiles = (uint32_t)j.num_tiles;
    pp.passes = P;
    pp.tw = j.tw;
    pp.dbg = j.dbg;
    pp.tickets_xcd = j.part_tickets;
    pp.xcd_lock = j.xcd_lock;
    pp.xcd_ok = pk.xcd_ok ? 1u : 0u;
    for (uint32_t p = 0; p < P; ++p) {
        pp.pass = p;
        pp.tickets_xcd = j.part_tickets + p * kTicketParts;
        void* args[] = {&pp};
        const hipError_t e = hipLaunchKernel(pk.fn, dim3(pk.grid), dim3(pk.threads), args, 0, s);
        if (e != hipSuccess) return e;
        record_event(j, s);
    }

    // 4. copy back when the executed pass count is odd (decided on the device;
    //    with skipped passes the count is made even, so only an odd pass count
    //    can leave the result in the workspace)
    if (P & 1u) {
        using VE = std::conditional_t<(VB > 0), VM, uint8_t>;
        const uint64_t bytes = j.n * (uint64_t)(KB > VB ? KB : VB);
        const uint32_t g = grid_for(bytes / 16u + 1u, 256u * 8u, j.num_cus, 4);
        const uint32_t vk = (((uintptr_t)j.keys | (uintptr_t)j.tmp_keys) & 15u) == 0u;
        const uint32_t vv = VB > 0 && (((uintptr_t)j.vals | (uintptr_t)j.tmp_vals) & 15u) == 0u;
        hipLaunchKernelGGL((k_copy_back<KM, VE>), dim3(g), dim3(256), 0, s, (const Ctrl*)j.ctrl, P, (KM*)j.keys,
                           (const KM*)j.tmp_keys, VB > 0 ? (VE*)j.vals : (VE*)nullptr,
                           VB > 0 ? (const VE*)j.tmp_vals : (const VE*)nullptr, j.n, vk, vv);
        record_event(j, s);
    }
    return hipGetLastError();
}

template <int KB, int VB>
hipError_t launch_core_v(const SortJob& j, hipStream_t s) {
    return j.wide ? launch_core_t<KB, VB, true>(j, s) : launch_core_t<KB, VB, false>(j, s);
}

template <int KB>
hipError_t launch_core(const SortJob& j, hipStream_t s) {
    switch (j.vb) {
        case 0: return launch_core_v<KB, 0>(j, s);
        case 1: return launch_core_v<KB, 1>(j, s);
        case 2: return launch_core_v<KB, 2>(j, s);
        case 4: return launch_core_v<KB, 4>(j, s);
        case 8: return launch_core_v<KB, 8>(j, s);
        case 16: return launch_core_v<KB, 16>(j, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace rsa
