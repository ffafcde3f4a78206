// Shared host/device definitions of the onesweep protocol: the control block
// that lives in the caller's workspace, the lookback word layout, and the
// key-twiddle masks.  Included by the HIP kernels, by the host dispatcher and
// by the CPU-side protocol simulation test, so it must compile with plain g++.
//
// Reference counterparts:
//   lookback word layout    include/rsort/detail/lookback.cuh:31-58 (flags + epoch + value)
//   epoch validity rule     README.md:117 ("matching epoch and no flag = not yet published")
//   twiddles                include/rsort/detail/tuning.cuh:297-397 (radix_traits)
#pragma once

#include <stddef.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define RSA_HD __host__ __device__ __forceinline__
#else
#define RSA_HD inline
#endif

namespace rsa {

constexpr uint32_t kRadix = 256;
constexpr uint32_t kMaxPasses = 16;

// ---------------------------------------------------------------------------
// Control block.  256 bytes at the start of the control region; it is cleared
// by the init kernel at the start of every sort, so stale workspace contents
// never leak into a new sort.
// ---------------------------------------------------------------------------
struct Ctrl {
    uint32_t ticket[kMaxPasses];  // per-pass tile tickets (dynamic tile order)
    // skip_mask and xcd_mode side by side (8-byte aligned): every pass
    // workgroup reads them first, in one scalar load
    uint32_t skip_mask;           // bit p set: every key has the same digit in pass p
    uint32_t xcd_mode;            // 1: this sort holds the device's per-XCD ticket lock
    uint32_t hist_done;           // histogram workgroup groups finished
    uint32_t pad;
    uint64_t clear_id;            // id of the call whose histogram cleared this header (no fill kernel)
    uint32_t group_done[32];      // histogram workgroups finished per group
    uint32_t reserved[10];
};
static_assert(sizeof(Ctrl) == 256, "control header must stay 256 bytes");
static_assert(offsetof(Ctrl, skip_mask) % 8 == 0 && offsetof(Ctrl, xcd_mode) == offsetof(Ctrl, skip_mask) + 4,
              "skip_mask and xcd_mode form one aligned 8-byte word");

// Passes that run.  A pass whose digit is the same for every key is the
// identity and is skipped, except that when the passes left are odd in
// number the lowest skipped pass runs as a plain copy: an even count of
// executed passes leaves the result in the caller's buffer, with no copy-back
// launch.  Only an odd count with nothing skipped (one-byte keys) ends in the
// workspace buffer.  Bit p set: pass p runs.
RSA_HD uint32_t pass_run_mask(uint32_t skip_mask, uint32_t passes) {
    const uint32_t pmask = passes >= 32 ? ~0u : ((1u << passes) - 1u);
    const uint32_t active = ~skip_mask & pmask;
    const uint32_t idle = pmask & ~active;
    if ((__builtin_popcount(active) & 1u) && idle) return active | (idle & (0u - idle));
    return active;
}

// Control block of the segmented sort (512 bytes, cleared per sort).
// Categories 0..kSegBins-1 are the LDS bins, kSegBins the onesweep segments.
struct SegCtrl {
    uint32_t ticket[kMaxPasses];  // onesweep pass tickets over the long segments' tiles
    uint32_t bin_ticket[16];      // work tickets of the LDS bins
    uint32_t cat_count[16];       // segments per category
    uint32_t cat_base[16];        // first queue slot of each category
    uint32_t large_tiles;         // onesweep tiles of all long segments
    uint32_t overflow;            // 1: segments overlap beyond the workspace bounds
    uint32_t reserved[62];
};
static_assert(sizeof(SegCtrl) == 512, "segmented control header must stay 512 bytes");

// One request of the direct (pointer-list) form of the segmented LDS bins:
// its own key and value arrays, sorted in place.
struct SegReq {
    uint64_t keys;  // device address
    uint64_t vals;  // device address (0 without values)
    uint64_t n;
    uint64_t pad;
};

// ---------------------------------------------------------------------------
// Lookback words.  One word per (tile, digit).  Flag, epoch and value share
// one machine word so a single atomic load observes a consistent triple.
//   32-bit: [31:30] flag  [29] epoch  [28:0] value   (n <= 2^29 - 1)
//   64-bit: [63:62] flag  [61] epoch  [60:0] value
// flag 0 = not published, 1 = tile aggregate only, 2 = inclusive prefix.
// The epoch is the parity of the executed-pass index, so a word left over
// from the previous executed pass never reads as valid; the whole array is
// zeroed once per sort (flag 0 is invalid for every epoch).
// ---------------------------------------------------------------------------
constexpr uint32_t kLbInvalid = 0;
constexpr uint32_t kLbAggregate = 1;
constexpr uint32_t kLbPrefix = 2;

template <typename W>
struct LbLayout;

template <>
struct LbLayout<uint32_t> {
    static constexpr int kFlagShift = 30;
    static constexpr int kEpochShift = 29;
    static constexpr uint32_t kValueMask = (1u << 29) - 1u;
};

template <>
struct LbLayout<uint64_t> {
    static constexpr int kFlagShift = 62;
    static constexpr int kEpochShift = 61;
    static constexpr uint64_t kValueMask = (1ull << 61) - 1ull;
};

template <typename W>
RSA_HD W lb_pack(uint32_t flag, uint32_t epoch, W value) {
    using L = LbLayout<W>;
    return (W(flag) << L::kFlagShift) | (W(epoch & 1u) << L::kEpochShift) | (value & L::kValueMask);
}

template <typename W>
RSA_HD uint32_t lb_flag(W w) {
    return uint32_t(w >> LbLayout<W>::kFlagShift) & 3u;
}

template <typename W>
RSA_HD uint32_t lb_epoch(W w) {
    return uint32_t(w >> LbLayout<W>::kEpochShift) & 1u;
}

template <typename W>
RSA_HD W lb_value(W w) {
    return w & LbLayout<W>::kValueMask;
}

// A word is usable by a pass of epoch `epoch` iff it carries a flag AND the
// matching epoch.  (The reference tests only the epoch bits,
// lookback.cuh:200-207; requiring a flag as well removes its tag-word hazard.)
template <typename W>
RSA_HD bool lb_ready(W w, uint32_t epoch) {
    return lb_flag(w) != kLbInvalid && lb_epoch(w) == (epoch & 1u);
}

// Largest n the 32-bit word can describe: an inclusive prefix may equal n.
constexpr uint64_t kNarrowMaxN = (1ull << 29) - 1ull;

// Decoupled lookback for one digit of `tile` (tile >= 1): sum the published
// values of the preceding tiles, walking back until an inclusive prefix.
// Each round issues the loads of `win` predecessors at once (independent
// loads, one round trip) and consumes them in order.  The window starts at
// W0 words and doubles up to WMAX while every word read is an aggregate, so
// short walks (the steady state) read few words and the long walks at the
// start of a pass (every resident tile starts at once) take few rounds; a
// not-ready word ends the round and is polled again after `backoff()`.
// Tile 0 always publishes an inclusive prefix, so the walk never passes it.
// `load(t)` returns the word of tile t.  Shared by the pass kernel and the
// host-thread protocol simulation.
template <typename W, int W0, int WMAX, typename Load, typename Backoff>
RSA_HD W lookback_walk(int64_t tile, uint32_t epoch, Load&& load, Backoff&& backoff, uint32_t& steps,
                       uint32_t& spins) {
    W acc = 0;
    int64_t j = tile - 1;  // next predecessor to consume
    int win = W0;
    while (true) {
        W word[WMAX];
#pragma unroll
        for (int k = 0; k < WMAX; ++k) word[k] = (k < win && (j - k) >= 0) ? load(j - k) : (W)0;
        int used = win;
        bool done = false;
#pragma unroll
        for (int k = 0; k < WMAX; ++k) {
            if (k < win && used == win && !done) {
                if (!lb_ready(word[k], epoch)) {
                    used = k;
                } else {
                    acc += lb_value(word[k]);
                    if (lb_flag(word[k]) == kLbPrefix) {
                        done = true;
                        used = k + 1;
                    }
                }
            }
        }
        steps += (uint32_t)used;
        if (done) return acc;
        j -= used;
        if (used < win) {
            ++spins;
            backoff();
        } else {
            win = win * 2 < WMAX ? win * 2 : WMAX;
        }
    }
}

// ---------------------------------------------------------------------------
// Group aggregates (plain sorts): tiles are also summed per group of
// kLbGroup consecutive tiles.  Each tile adds (1 << count shift) | aggregate
// to its group's word for every digit with a device-scope atomic when it
// publishes its aggregate, so a group word with count == kLbGroup holds the
// group's total.  One group word per (pass, group, digit), zeroed per sort.
// A walk that reaches a complete group adds its total in one read instead of
// reading its kLbGroup tile words, and a group whose last tile has published
// an inclusive prefix ends the walk.
// ---------------------------------------------------------------------------
#ifndef RSA_LB_GROUP
#define RSA_LB_GROUP 8
#endif
constexpr int kLbGroup = RSA_LB_GROUP;
static_assert(kLbGroup >= 2 && kLbGroup <= 63, "the group count field has 6 bits");

template <typename W>
struct GrpLayout {
    static constexpr int kCountShift = (int)sizeof(W) * 8 - 6;
    static constexpr W kSumMask = (W(1) << kCountShift) - W(1);
};

template <typename W>
RSA_HD W grp_add(W aggregate) {
    return (W(1) << GrpLayout<W>::kCountShift) | aggregate;
}

template <typename W>
RSA_HD bool grp_complete(W w) {
    return (uint32_t)(w >> GrpLayout<W>::kCountShift) == (uint32_t)kLbGroup;
}

template <typename W>
RSA_HD W grp_sum(W w) {
    return w & GrpLayout<W>::kSumMask;
}

// Adds the tile words of tiles hi, hi-1, ..., lo (ready ones; a not-ready
// word is polled again after backoff()) until an inclusive prefix is met.
// Returns true when one was met (acc then holds the finished prefix).
template <typename W, int WIN, typename Load, typename Backoff>
RSA_HD bool lookback_range(int64_t hi, int64_t lo, uint32_t epoch, Load&& load, Backoff&& backoff, W& acc,
                           uint32_t& steps, uint32_t& spins) {
    int64_t j = hi;  // next tile to consume
    while (j >= lo) {
        W word[WIN];
#pragma unroll
        for (int k = 0; k < WIN; ++k) word[k] = (j - k) >= lo ? load(j - k) : (W)0;
        int used = 0;
        bool stop = false;
#pragma unroll
        for (int k = 0; k < WIN; ++k) {
            if (!stop && (j - k) >= lo) {
                if (!lb_ready(word[k], epoch)) {
                    stop = true;
                } else {
                    acc += lb_value(word[k]);
                    used = k + 1;
                    if (lb_flag(word[k]) == kLbPrefix) {
                        steps += (uint32_t)used;
                        return true;
                    }
                }
            }
        }
        steps += (uint32_t)used;
        j -= used;
        if (stop) {
            ++spins;
            backoff();
        }
    }
    return false;
}

// The grouped walk from a given point: the tile words hi, hi-1, ..., lo
// (none when hi < lo), then whole groups h, h-1, ..., GW per round (the last
// tile's word and the group word of each), falling back to the tile words of
// a group that is not complete yet.  Adds to acc and returns the finished
// prefix.
template <typename W, int WIN, int GW, typename LoadTile, typename LoadGroup, typename Backoff>
RSA_HD W lookback_walk_groups_from(W acc, int64_t hi, int64_t lo, int64_t h, uint32_t epoch, LoadTile&& load_tile,
                                   LoadGroup&& load_group, Backoff&& backoff, uint32_t& steps, uint32_t& spins) {
    if (lookback_range<W, WIN>(hi, lo, epoch, load_tile, backoff, acc, steps, spins)) return acc;
    while (h >= 0) {
        W last[GW], grp[GW];
#pragma unroll
        for (int k = 0; k < GW; ++k) {
            const bool in = (h - k) >= 0;
            last[k] = in ? load_tile((h - k) * kLbGroup + (kLbGroup - 1)) : (W)0;
            grp[k] = in ? load_group(h - k) : (W)0;
        }
        int used = 0;
        bool fallback = false;
#pragma unroll
        for (int k = 0; k < GW; ++k) {
            if (!fallback && (h - k) >= 0) {
                if (lb_ready(last[k], epoch) && lb_flag(last[k]) == kLbPrefix) {
                    steps += (uint32_t)(2 * (k + 1));
                    return acc + lb_value(last[k]);
                }
                if (grp_complete(grp[k])) {
                    acc += grp_sum(grp[k]);
                    used = k + 1;
                } else {
                    fallback = true;
                }
            }
        }
        steps += (uint32_t)(2 * used);
        h -= used;
        if (fallback) {
            // group h is not complete yet: its tile words, polling as needed
            if (lookback_range<W, WIN>(h * kLbGroup + (kLbGroup - 1), h * kLbGroup, epoch, load_tile, backoff, acc,
                                       steps, spins))
                return acc;
            --h;
        }
    }
    return acc;
}

// Exclusive prefix of `tile` over the earlier tiles with group aggregates:
// the earlier tiles of its own group at tile level, then whole groups.  Tile
// 0 publishes an inclusive prefix and group 0 is complete once its tiles have
// published, so every walk ends.
template <typename W, int WIN, int GW, typename LoadTile, typename LoadGroup, typename Backoff>
RSA_HD W lookback_walk_grouped(int64_t tile, uint32_t epoch, LoadTile&& load_tile, LoadGroup&& load_group,
                               Backoff&& backoff, uint32_t& steps, uint32_t& spins) {
    const int64_t g = tile / kLbGroup;
    return lookback_walk_groups_from<W, WIN, GW>((W)0, tile - 1, g * kLbGroup, g - 1, epoch, load_tile, load_group,
                                                 backoff, steps, spins);
}

// ---------------------------------------------------------------------------
// Wide walk (plain sorts with few tiles, where every walk runs back to the
// first tiles while they are still in flight): the walk of one tile is cut
// into segments that separate threads scan in one round of loads, nearest
// first.  Segment 0 is the earlier tiles of the tile's own group; segment
// s >= 1 is the groups g-1-(s-1)*GW down to g-s*GW.  A segment gives its
// partial sum and how it ended:
//   kSegOn    passed through every item;
//   kSegEnd   met an inclusive prefix, or ran past group 0 (the sum is then
//             complete);
//   kSegStop  met a not-ready tile word or an incomplete group at item k,
//             which is stored as k << 2 (items before it are in the sum).
// lookback_wide_finish adds the segments in order and resumes the ordinary
// walk where one stopped, so the prefix is the one the ordinary walk finds.
// ---------------------------------------------------------------------------
constexpr uint32_t kSegOn = 0u, kSegEnd = 1u, kSegStop = 2u;

template <typename W, typename LoadTile>
RSA_HD uint32_t lb_segment_tiles(int64_t tile, uint32_t epoch, LoadTile&& load_tile, W& part, uint32_t& steps) {
    constexpr int N = kLbGroup - 1;  // earlier tiles of a group
    const int64_t lo = (tile / kLbGroup) * kLbGroup;
    W word[N];
#pragma unroll
    for (int k = 0; k < N; ++k) word[k] = (tile - 1 - k) >= lo ? load_tile(tile - 1 - k) : (W)0;
    part = 0;
#pragma unroll
    for (int k = 0; k < N; ++k) {
        if ((tile - 1 - k) < lo) return kSegOn;
        ++steps;
        if (!lb_ready(word[k], epoch)) return kSegStop | ((uint32_t)k << 2);
        part += lb_value(word[k]);
        if (lb_flag(word[k]) == kLbPrefix) return kSegEnd;
    }
    return kSegOn;
}

template <typename W, int GW, typename LoadTile, typename LoadGroup>
RSA_HD uint32_t lb_segment_groups(int64_t h, uint32_t epoch, LoadTile&& load_tile, LoadGroup&& load_group, W& part,
                                  uint32_t& steps) {
    W last[GW], grp[GW];
#pragma unroll
    for (int k = 0; k < GW; ++k) {
        const bool in = (h - k) >= 0;
        last[k] = in ? load_tile((h - k) * kLbGroup + (kLbGroup - 1)) : (W)0;
        grp[k] = in ? load_group(h - k) : (W)0;
    }
    part = 0;
#pragma unroll
    for (int k = 0; k < GW; ++k) {
        if ((h - k) < 0) return kSegEnd;
        steps += 2u;
        if (lb_ready(last[k], epoch) && lb_flag(last[k]) == kLbPrefix) {
            part += lb_value(last[k]);
            return kSegEnd;
        }
        if (!grp_complete(grp[k])) return kSegStop | ((uint32_t)k << 2);
        part += grp_sum(grp[k]);
    }
    return kSegOn;
}

// Prefix of `tile` from the results of segments 0..NSEG-1 (part, st).
template <typename W, int WIN, int GW, int NSEG, typename LoadTile, typename LoadGroup, typename Backoff>
RSA_HD W lookback_wide_finish(int64_t tile, uint32_t epoch, const W (&part)[NSEG], const uint32_t (&st)[NSEG],
                              LoadTile&& load_tile, LoadGroup&& load_group, Backoff&& backoff, uint32_t& steps,
                              uint32_t& spins) {
    const int64_t g = tile / kLbGroup;
    W acc = 0;
#pragma unroll
    for (int s = 0; s < NSEG; ++s) {
        acc += part[s];
        const uint32_t how = st[s] & 3u;
        const int64_t k = (int64_t)(st[s] >> 2);
        if (how == kSegEnd) return acc;
        if (how == kSegStop) {
            if (s == 0)  // not-ready tile word tile-1-k of the own group
                return lookback_walk_groups_from<W, WIN, GW>(acc, tile - 1 - k, g * kLbGroup, g - 1, epoch,
                                                             load_tile, load_group, backoff, steps, spins);
            const int64_t hs = g - 1 - (int64_t)(s - 1) * GW - k;  // incomplete group
            return lookback_walk_groups_from<W, WIN, GW>(acc, hs * kLbGroup + (kLbGroup - 1), hs * kLbGroup, hs - 1,
                                                         epoch, load_tile, load_group, backoff, steps, spins);
        }
    }
    return lookback_walk_groups_from<W, WIN, GW>(acc, -1, 0, g - 1 - (int64_t)(NSEG - 1) * GW, epoch, load_tile,
                                                 load_group, backoff, steps, spins);
}

// ---------------------------------------------------------------------------
// Tile tickets of a pass.  Ticket k of counter `part` (of `parts`, each
// handing out runs of `chunk` consecutive tiles) is this tile; one counter
// with chunk 1 is the global ticket (tile = k).
// ---------------------------------------------------------------------------
RSA_HD uint32_t ticket_tile(uint32_t k, uint32_t part, uint32_t parts, uint32_t chunk) {
    return (k / chunk) * (parts * chunk) + part * chunk + k % chunk;
}

// ---------------------------------------------------------------------------
// Key twiddle masks (runtime, wave-uniform).  For a key of W bits held in an
// unsigned word u:
//   in : t = u ^ ((sext(msb(u)) & fmask) | smask) ^ dmask
//   out: u = t' ^ ((~sext(msb(t')) & fmask) | smask),  t' = t ^ dmask
// with fmask = ones(W) for floating keys, smask = sign bit for signed and
// floating keys, dmask = ones(W) for descending order.  For 128-bit keys the
// masks apply per 64-bit half and smask only to the high half.
// ---------------------------------------------------------------------------
enum KeyKind : uint32_t { kKindUnsigned = 0, kKindSigned = 1, kKindFloat = 2 };

struct TwMasks {
    uint64_t fmask;  // per 64-bit half
    uint64_t smask;  // high half only
    uint64_t dmask;  // per 64-bit half
    uint32_t active; // nonzero when the twiddle is not the identity
    uint32_t pad;
};

RSA_HD TwMasks make_tw_masks(uint32_t kind, uint32_t key_bytes, bool descending) {
    const uint32_t bits = key_bytes >= 8 ? 64u : key_bytes * 8u;
    const uint64_t ones = bits == 64 ? ~0ull : ((1ull << bits) - 1ull);
    const uint64_t sign = 1ull << (bits - 1u);
    TwMasks m{};
    m.fmask = (kind == kKindFloat) ? ones : 0ull;
    m.smask = (kind == kKindSigned || kind == kKindFloat) ? sign : 0ull;
    m.dmask = descending ? ones : 0ull;
    m.active = (m.fmask | m.smask | m.dmask) != 0ull;
    return m;
}

}  // namespace rsa
