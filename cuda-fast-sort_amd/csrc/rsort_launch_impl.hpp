// Launcher template instantiated once per key width (sort_kb*.hip).
#pragma once

#include "rsort_kernels.hpp"
#include "rsort_launch.hpp"
#include "rsort_onesweep_p.hpp"

namespace rsa {

// One pass launch of one kernel variant.  The per-XCD tile tickets make a
// workgroup wait, in the worst case, on tiles held by the next
// kTicketParts * C workgroups in dispatch order, which must then be able to
// be resident at once: they are used only on a whole 8-XCD device (256 CUs)
// where the kernel's resident workgroups number at least twice that;
// otherwise the single global ticket, which never waits on a tile not yet
// handed out.
template <int KB, int VB, bool WIDE, int THREADS, int K, bool SEG, bool EARLY>
void launch_pass_variant(const SortJob& j, PassParams pp, hipStream_t s) {
    constexpr uint32_t C = tile_chunk(KB);
    static const int occ = blocks_per_cu(k_onesweep<KB, VB, WIDE, THREADS, K, SEG, EARLY>, THREADS);
    const bool xcd = C > 0 && j.xcd_tickets && j.num_cus >= 256 &&
                     (uint64_t)occ * (uint64_t)j.num_cus >= 2ull * kTicketParts * C;
    pp.tickets = xcd ? j.part_tickets + pp.pass * kTicketParts : &j.ctrl->ticket[pp.pass];
    pp.tparts = xcd ? kTicketParts : 1u;
    pp.tchunk = xcd ? C : 1u;
    // whole rounds of the per-XCD ticket counters (extra workgroups end at once)
    const uint64_t round = xcd ? (uint64_t)kTicketParts * C : 1u;
    const uint32_t grid = (uint32_t)((j.num_tiles + round - 1) / round * round);
    hipLaunchKernelGGL((k_onesweep<KB, VB, WIDE, THREADS, K, SEG, EARLY>), dim3(grid), dim3(THREADS), 0, s, pp);
}

// One pass of the tile kernel; passes of few tiles take the variant with the
// segmented lookback walk, large u32/u32 passes the one that publishes tile
// aggregates before the rank.
template <int KB, int VB, bool WIDE, int THREADS, int K>
void launch_onesweep(const SortJob& j, const PassParams& pp, hipStream_t s) {
    if constexpr (!WIDE && THREADS >= 512) {
        if (j.lb_wide) return launch_pass_variant<KB, VB, WIDE, THREADS, K, true, false>(j, pp, s);
    }
    if constexpr (KB == 4 && VB == 4 && !WIDE) {
        if (j.early_agg) return launch_pass_variant<KB, VB, WIDE, THREADS, K, false, true>(j, pp, s);
    }
    launch_pass_variant<KB, VB, WIDE, THREADS, K, false, false>(j, pp, s);
}

template <int KB, int VB, bool WIDE>
hipError_t launch_core_t(const SortJob& j, hipStream_t s) {
    using Off = std::conditional_t<WIDE, uint64_t, uint32_t>;
    using KM = typename KeyTraits<KB>::Mem;
    using VM = typename ValTraits<VB>::Mem;
    constexpr TileCfg C = tile_cfg(KB, VB);
    constexpr uint32_t P = KB;

    record_event(j, s);
    // 1. control header, global digit counts, tickets and lookback words
    //    (flag 0: never ready in any epoch), contiguous in the workspace: one
    //    fill, or none when the histogram is a single workgroup (it clears
    //    them itself)
    const uint64_t clear_units =
        (uint64_t)((sizeof(Ctrl) + (size_t)P * kRadix * sizeof(Off) + j.lb_bytes) / 16u);
    const uint32_t hg = grid_for(j.n, hist_keys_per_block_iter(KB), j.num_cus, j.hist_bpc ? j.hist_bpc : 1u);
    const uint32_t hgrid = hg < kHistMaxGrid ? hg : kHistMaxGrid;
    if (hgrid > 1u)
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
    h.self_clear = hgrid == 1u ? (uint4*)j.ctrl : nullptr;
    h.self_clear_units = hgrid == 1u ? clear_units : 0u;
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
    constexpr ChunkCfg CC = chunk_cfg(KB, VB);
    for (uint32_t p = 0; p < P; ++p) {
        pp.pass = p;
        if (j.pass_kernel == kPassChunk) {
            // persistent: one workgroup per CU (its LDS allows no second)
            const uint32_t grid = j.num_tiles < (uint64_t)j.num_cus ? (uint32_t)j.num_tiles : (uint32_t)j.num_cus;
            hipLaunchKernelGGL((k_onesweep_p<KB, VB, WIDE, CC.threads, CC.ipt, CC.subs>), dim3(grid),
                               dim3(CC.threads), 0, s, pp);
        } else if (tile_cands_built(KB, VB) && j.tile_cand == 1) {
            constexpr TileCfg C1 = tile_cand(KB, VB, 1);
            launch_onesweep<KB, VB, WIDE, C1.threads, C1.items>(j, pp, s);
        } else if (tile_cands_built(KB, VB) && j.tile_cand == 2) {
            constexpr TileCfg C2 = tile_cand(KB, VB, 2);
            launch_onesweep<KB, VB, WIDE, C2.threads, C2.items>(j, pp, s);
        } else if (tile_cands_built(KB, VB) && j.tile_cand == 3) {
            constexpr TileCfg C3 = tile_cand(KB, VB, 3);
            launch_onesweep<KB, VB, WIDE, C3.threads, C3.items>(j, pp, s);
        } else {
            launch_onesweep<KB, VB, WIDE, C.threads, C.items>(j, pp, s);
        }
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
