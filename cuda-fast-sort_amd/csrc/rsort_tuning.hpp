// Launch geometry for gfx950.  The reference keeps an sm_86 table of
// threads x items per thread by key/value width
// (include/rsort/detail/tuning.cuh:138-213); this is the wave64 / 160 KiB-LDS
// counterpart.  A tile of THREADS*ITEMS keys is one workgroup of the pass
// kernel; the staging buffer holds one tile of keys (then of values), so the
// LDS footprint is THREADS*ITEMS*max(key,value bytes) + 1 KiB per wave.
#pragma once

#include <stdint.h>

#include "rsort_protocol.hpp"

namespace rsa {

// Histogram counter layout: 1 = lane slots (bank-conflict-free LDS atomics,
// one 1024-thread workgroup per CU), 0 = one counter copy per wave (512
// threads, several workgroups per CU).
#ifndef RSA_HIST_SLOTS
#define RSA_HIST_SLOTS 1
#endif
constexpr int kHistThreads = RSA_HIST_SLOTS ? 1024 : 512;
// Histogram workgroups fold their counts in groups of this many (the last
// arrival of a group sums the group's rows) before touching global atomics.
constexpr int kHistGroup = 32;
// Upper bound of the histogram grid (32 groups, the arrival counters that fit
// in the control header).
constexpr uint32_t kHistMaxGrid = 32 * kHistGroup;

// The histogram kernel clears the control region (header, counts, tickets,
// lookback words) itself, with no fill kernel, up to this size (a u32 sort
// of up to 2^25 keys: 1.5 KiB per 16K-key tile).  Above it, the plain stores
// cost the histogram more than the fill kernel costs the sort.
constexpr uint64_t kSelfClearMaxBytes = 4ull << 20;
// Histogram grids up to this size fold by device atomics from every
// workgroup (fewer coherent round trips; the two-level fold above it keeps
// the device atomics few).
constexpr uint32_t kHistDirectMaxGrid = 256;

// Tile tickets of the pass kernel, by XCD: workgroup b takes its tickets
// from counter b mod kTicketParts (the round-robin order in which workgroups
// are placed on the 8 XCDs), and each counter hands out runs of
// tile_chunk() consecutive tiles, so that neighbouring tiles -- whose digit
// runs share the partial cache lines at their ends -- are written through
// the same L2.  Measured on MI355X with the tuned tiles: runs of 16 tiles;
// runs of 8 were 1-2 % slower for u32/u32 and 3 % for u64, runs of 32 and 2
// slower still.  RSA_XCD_CHUNK sets the run length; 0 = one global ticket
// counter.
#ifndef RSA_XCD_CHUNK
#define RSA_XCD_CHUNK 16
#endif
constexpr uint32_t kTicketParts = 8;
RSA_HD constexpr uint32_t tile_chunk(int) { return (uint32_t)RSA_XCD_CHUNK; }
// bytes of the per-pass ticket counters at the start of the lookback region
constexpr size_t kPartTicketBytes = 16 * kTicketParts * 4;  // 16 passes at most

// u32/u32 passes of at least this many tiles publish each tile's aggregate
// from digit counts taken before the rank (one LDS atomic per key more; the
// successors then wait less).  Measured on MI355X: u32/u32 2^28 -2.9 %, but
// 2^24 +4 %, and u32 keys, u64/u64 slower.
constexpr uint64_t kEarlyAggMinTiles = 8192;

// Passes of at most this many tiles use the one-round segmented lookback
// walk (all of their tiles are resident at once on 256 CUs).
constexpr uint64_t kLbWideMaxTiles = 512;

// Histogram grid for n keys of key_bytes: each workgroup gets at least 4
// loads of 16 bytes per lane.
RSA_HD constexpr uint32_t hist_keys_per_block_iter(int key_bytes) {
    return (uint32_t)(kHistThreads / 64) * 64u * (16u / (uint32_t)key_bytes) * 4u;
}

// LDS copies of the per-pass digit counters in the histogram kernel: one per
// wave while it fits in 64 KiB, otherwise shared by pairs of waves.
RSA_HD constexpr int hist_copies(int key_bytes) {
    return (key_bytes * 1024 * (kHistThreads / 64) <= 64 * 1024) ? (kHistThreads / 64)
                                                                   : (64 * 1024) / (key_bytes * 1024);
}

// Lane slots per histogram counter: 32 (one per LDS bank) while the
// counters fit in 128 KiB, fewer for 8- and 16-byte keys.
RSA_HD constexpr int hist_slots(int key_bytes) { return key_bytes <= 4 ? 32 : 128 / key_bytes; }

// Histogram workgroups per CU: as many as the LDS counters allow, up to 4.
RSA_HD constexpr uint32_t hist_blocks_per_cu(int key_bytes) {
    const uint32_t lds = RSA_HIST_SLOTS ? (uint32_t)hist_slots(key_bytes) * (uint32_t)key_bytes * 1024u
                                        : (uint32_t)hist_copies(key_bytes) * (uint32_t)key_bytes * 1024u;
    const uint32_t b = (160u * 1024u - 1024u) / lds;
    return b < 1u ? 1u : (b > 4u ? 4u : b);
}

struct TileCfg {
    uint32_t threads;
    uint32_t items;
};

// Tile of the pass kernel per key / value width: threads x items per thread.
// Measured on MI355X: fewer, larger tiles win -- every tile adds a column of
// 256 lookback words that its successors must read -- as long as the
// workgroups resident per CU still cover the latency.  Whole-call GKeys/s
// against the earlier 512-thread tiles: u32 keys 2^28 1024x16 (two
// workgroups per CU) 98 vs 82; u32/u32 1024x20 (one per CU) 56 vs 45; u64
// 2^27 1024x8 28.8 vs 24.9; u64/u64 1024x10 15.8 vs 12.7; u128 2^26 1024x8
// 5.9 vs 5.6; u128/u64 1024x6 4.45 vs 3.61.
#ifndef RSA_TILE_K4_V0
#define RSA_TILE_K4_V0 1024, 16
#endif
#ifndef RSA_TILE_K4_V4
#define RSA_TILE_K4_V4 1024, 20
#endif
#ifndef RSA_TILE_K4_V8
#define RSA_TILE_K4_V8 1024, 12
#endif
#ifndef RSA_TILE_K4_V16
#define RSA_TILE_K4_V16 1024, 8
#endif
#ifndef RSA_TILE_K8_V0
#define RSA_TILE_K8_V0 1024, 8
#endif
#ifndef RSA_TILE_K8_V8
#define RSA_TILE_K8_V8 1024, 10
#endif
#ifndef RSA_TILE_K8_V16
#define RSA_TILE_K8_V16 1024, 8
#endif
#ifndef RSA_TILE_K16_V0
#define RSA_TILE_K16_V0 1024, 8
#endif
#ifndef RSA_TILE_K16_VX
#define RSA_TILE_K16_VX 1024, 6
#endif
#define RSA_TILE_CFG(...) TileCfg{__VA_ARGS__}

// key_bytes in {1,2,4,8,16}; val_bytes in {0,1,2,4,8,16}
RSA_HD constexpr TileCfg tile_cfg(int key_bytes, int val_bytes) {
    const int wide = key_bytes > val_bytes ? key_bytes : val_bytes;
    if (key_bytes <= 4) {
        if (val_bytes == 0) return RSA_TILE_CFG(RSA_TILE_K4_V0);
        if (wide <= 4) return RSA_TILE_CFG(RSA_TILE_K4_V4);
        if (wide == 8) return RSA_TILE_CFG(RSA_TILE_K4_V8);
        return RSA_TILE_CFG(RSA_TILE_K4_V16);
    }
    if (key_bytes == 8) {
        if (val_bytes == 0) return RSA_TILE_CFG(RSA_TILE_K8_V0);
        if (wide == 8) return RSA_TILE_CFG(RSA_TILE_K8_V8);
        return RSA_TILE_CFG(RSA_TILE_K8_V16);
    }
    // 16-byte keys
    if (val_bytes == 0) return RSA_TILE_CFG(RSA_TILE_K16_V0);
    return RSA_TILE_CFG(RSA_TILE_K16_VX);
}

// Minimum waves per SIMD the pass kernel is compiled for: 8 (at most 64
// VGPRs) where the tile fits that without spilling, so that two 1024-thread
// (or four 512-thread) workgroups share a CU; 1 (no bound) elsewhere.
RSA_HD constexpr int pass_min_waves(int key_bytes, int val_bytes, int items) {
    return (key_bytes <= 4 && (val_bytes == 0 || val_bytes == 4) && items <= 16) ||
                   (key_bytes == 8 && (val_bytes == 0 || val_bytes == 8) && items <= 8)
               ? 8
               : 1;
}

// Tile candidates of the pass kernel the autotuner chooses from.  Candidate
// 0 is tile_cfg; 1 has fewer items per thread (four fewer from 12 items up,
// else two fewer), 2 half the threads, 3 half the items.  They are built for
// the widths sorted most -- 4- and 8-byte keys alone or with 4- or 8-byte
// values, 16-byte keys alone or with 8-byte values -- and fall back to
// candidate 0 elsewhere.
constexpr int kTileCands = 4;

RSA_HD constexpr bool tile_cands_built(int key_bytes, int val_bytes) {
    return ((key_bytes == 4 || key_bytes == 8) && (val_bytes == 0 || val_bytes == 4 || val_bytes == 8)) ||
           (key_bytes == 16 && (val_bytes == 0 || val_bytes == 8));
}

RSA_HD constexpr TileCfg tile_cand(int key_bytes, int val_bytes, int cand) {
    const TileCfg d = tile_cfg(key_bytes, val_bytes);
    if (!tile_cands_built(key_bytes, val_bytes) || cand <= 0 || cand >= kTileCands) return d;
    if (cand == 1) return TileCfg{d.threads, d.items >= 12u ? d.items - 4u : d.items - 2u};
    if (cand == 2) return TileCfg{d.threads / 2u, d.items};
    return TileCfg{d.threads, d.items / 2u};
}

// ---------------------------------------------------------------------------
// Segmented sort.  A segment of at most seg_cap(kb, vb, kSegBins-1) keys is
// sorted by one workgroup entirely in LDS; bin b takes the segments too long
// for bin b-1.  Longer segments go through the onesweep passes with
// per-segment digit offsets.  The item count keeps one bin's keys (or values)
// at <= 64 KiB of LDS.
// ---------------------------------------------------------------------------
constexpr int kSegBins = 9;
constexpr uint32_t kSegCatLarge = kSegBins;      // category of the onesweep segments
constexpr uint32_t kSegCatTiles = kSegBins + 1;  // chunk-table column: the long segments' tiles
constexpr uint32_t kSegRow = 16;                 // chunk-table columns per classification chunk
constexpr uint32_t kSegCatNone = kSegRow - 1;    // 0 or 1 key: nothing to do
constexpr uint32_t kSegChunk = 1024;             // segments per classification workgroup
static_assert(kSegCatTiles < kSegCatNone, "categories and the tile column fit one chunk row");

RSA_HD constexpr uint32_t seg_items(int key_bytes, int val_bytes) {
    const int wide = key_bytes > val_bytes ? key_bytes : val_bytes;
    return wide <= 4 ? 16u : (wide == 8 ? 8u : 4u);
}

// Bin b: threads, and items per thread as a fraction of seg_items (in
// quarters).  Capacities grow by 1.33x or 1.5x from bin to bin (1024, 1536,
// 2048, 3072, 4096, 6144, 8192, 12288, 16384 keys for 4-byte keys and
// values), so a segment fills at least 2/3 of its workgroup's slots: the
// bins are bound by LDS throughput, and filler slots cost LDS work too.
// Measured on MI355X (100k lognormal segments, u32/u32): five bins of
// capacity 1024 << b, ~70 % full, 100.3 GKeys/s; these nine, see the design
// notes.  Bin 0 runs two waves at half the items: its one-wave segments were
// bound by the latency of each wave's rank chain.  (Twice the threads at half
// the items in the larger bins, which are bound by LDS throughput, was
// slower.)
RSA_HD constexpr uint32_t seg_bin_threads(int, int, int bin) {
    return bin <= 2 ? 128u : bin <= 4 ? 256u : bin == 5 ? 384u : bin == 6 ? 512u : bin == 7 ? 768u : 1024u;
}
RSA_HD constexpr uint32_t seg_bin_quarters(int bin) { return bin == 0 ? 2u : (bin == 1 || bin == 3) ? 3u : 4u; }
RSA_HD constexpr uint32_t seg_bin_items(int key_bytes, int val_bytes, int bin) {
    return seg_items(key_bytes, val_bytes) * seg_bin_quarters(bin) / 4u;
}

RSA_HD constexpr uint32_t seg_cap(int key_bytes, int val_bytes, int bin) {
    return seg_bin_items(key_bytes, val_bytes, bin) * seg_bin_threads(key_bytes, val_bytes, bin);
}

}  // namespace rsa
