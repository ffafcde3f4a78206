// CPU simulation of the tile-ticket + decoupled lookback protocol of the
// onesweep pass (rsort_protocol.hpp word layout, k_onesweep's publish /
// lookback order).  Host threads play workgroups: each takes a ticket, publishes
// its aggregate, walks back until an inclusive prefix, publishes its own
// prefix.  Several executed passes reuse one lookback array with the epoch
// parity and no clearing in between; words start zeroed once per sort.  The
// walk is rsa::lookback_walk, the function the kernel runs, with first/maximum
// windows of 1/1, 1/4, 3/8 and 4/32 words, and rsa::lookback_walk_grouped
// (window -1, -4): tiles also add their aggregate to a per-pass group word
// with an atomic, and the walk skips complete groups; window -5 is the
// segmented walk (rsa::lb_segment_tiles / lb_segment_groups, then
// rsa::lookback_wide_finish).
// Exit status 0 when every tile's prefix equals the sum of earlier tiles.
#include <stdint.h>
#include <stdio.h>

#include <atomic>
#include <random>
#include <thread>
#include <vector>

#include "../cuda-fast-sort_amd/csrc/rsort_protocol.hpp"

template <typename W>
static int run(uint32_t tiles, uint32_t digits, uint32_t passes, uint32_t threads, uint32_t seed, int window) {
    std::vector<std::atomic<W>> lb((size_t)tiles * digits);
    for (auto& w : lb) w.store(0, std::memory_order_relaxed);  // cleared once per sort
    std::mt19937 rng(seed);
    int errors = 0;
    for (uint32_t e = 0; e < passes; ++e) {
        const uint32_t epoch = e & 1u;
        std::vector<uint32_t> count((size_t)tiles * digits);
        for (auto& c : count) c = rng() % 300u;
        std::vector<W> expect((size_t)tiles * digits);
        for (uint32_t d = 0; d < digits; ++d) {
            W acc = 0;
            for (uint32_t t = 0; t < tiles; ++t) {
                expect[(size_t)t * digits + d] = acc;
                acc += count[(size_t)t * digits + d];
            }
        }
        const uint32_t groups = (tiles + rsa::kLbGroup - 1) / rsa::kLbGroup;
        std::vector<std::atomic<W>> grp((size_t)groups * digits);  // one array per pass, zeroed
        for (auto& w : grp) w.store(0, std::memory_order_relaxed);
        std::atomic<uint32_t> ticket{0};
        std::atomic<int> bad{0};
        auto worker = [&](uint32_t id) {
            std::mt19937 r(seed * 7919u + id + e * 131u);
            for (;;) {
                const uint32_t tile = ticket.fetch_add(1, std::memory_order_relaxed);
                if (tile >= tiles) return;
                if (r() % 8 == 0) std::this_thread::yield();
                for (uint32_t d = 0; d < digits; ++d) {
                    const W c = count[(size_t)tile * digits + d];
                    lb[(size_t)tile * digits + d].store(
                        rsa::lb_pack<W>(tile == 0 ? rsa::kLbPrefix : rsa::kLbAggregate, epoch, c),
                        std::memory_order_relaxed);
                    if (window < 0) {
                        if (r() % 16 == 0) std::this_thread::yield();
                        grp[(size_t)(tile / rsa::kLbGroup) * digits + d].fetch_add(rsa::grp_add<W>(c),
                                                                                   std::memory_order_relaxed);
                    }
                }
                if (r() % 8 == 0) std::this_thread::yield();
                for (uint32_t d = 0; d < digits; ++d) {
                    W prefix = 0;
                    if (tile > 0) {
                        uint32_t steps = 0, spins = 0;
                        auto load = [&](int64_t t) { return lb[(size_t)t * digits + d].load(std::memory_order_relaxed); };
                        auto backoff = [] { std::this_thread::yield(); };
                        auto gload = [&](int64_t h) { return grp[(size_t)h * digits + d].load(std::memory_order_relaxed); };
                        if (window == -5) {
                            // segmented walk: the segments are scanned one after
                            // another here (by separate threads on the device)
                            W part[4];
                            uint32_t st[4];
                            st[0] = rsa::lb_segment_tiles<W>(tile, epoch, load, part[0], steps);
                            for (int sg = 1; sg < 4; ++sg) {
                                if (r() % 4 == 0) std::this_thread::yield();
                                st[sg] = rsa::lb_segment_groups<W, 4>((int64_t)(tile / rsa::kLbGroup) - 1 - (sg - 1) * 4,
                                                                      epoch, load, gload, part[sg], steps);
                            }
                            prefix = rsa::lookback_wide_finish<W, 8, 4>(tile, epoch, part, st, load, gload, backoff,
                                                                       steps, spins);
                        } else
                        prefix = (window == -1)  ? rsa::lookback_walk_grouped<W, 2, 1>(tile, epoch, load, gload, backoff, steps, spins)
                                 : (window == -4) ? rsa::lookback_walk_grouped<W, 8, 4>(tile, epoch, load, gload, backoff, steps, spins)
                                 : (window == 1) ? rsa::lookback_walk<W, 1, 1>(tile, epoch, load, backoff, steps, spins)
                                 : (window == 4) ? rsa::lookback_walk<W, 1, 4>(tile, epoch, load, backoff, steps, spins)
                                 : (window == 8) ? rsa::lookback_walk<W, 3, 8>(tile, epoch, load, backoff, steps, spins)
                                                 : rsa::lookback_walk<W, 4, 32>(tile, epoch, load, backoff, steps, spins);
                        lb[(size_t)tile * digits + d].store(
                            rsa::lb_pack<W>(rsa::kLbPrefix, epoch, prefix + count[(size_t)tile * digits + d]),
                            std::memory_order_relaxed);
                    }
                    if (prefix != expect[(size_t)tile * digits + d]) bad.fetch_add(1);
                }
            }
        };
        std::vector<std::thread> th;
        for (uint32_t i = 0; i < threads; ++i) th.emplace_back(worker, i);
        for (auto& t : th) t.join();
        errors += bad.load();
    }
    return errors;
}

// Fixed lookback states, one thread: every tile word ready (aggregate, or an
// inclusive prefix with probability p_incl; tile 0 always inclusive) and each
// group complete with probability p_complete (else one tile short).  Every
// walk of every tile must give the exclusive prefix.
template <typename W>
static int fixed_states(uint32_t tiles, double p_incl, double p_complete, uint32_t seed) {
    std::mt19937 rng(seed);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    const uint32_t epoch = seed & 1u;
    std::vector<W> cnt(tiles), pre(tiles), lb(tiles);
    W acc = 0;
    for (uint32_t t = 0; t < tiles; ++t) {
        cnt[t] = rng() % 1000u;
        pre[t] = acc;
        acc += cnt[t];
        const bool incl = t == 0 || u(rng) < p_incl;
        lb[t] = rsa::lb_pack<W>(incl ? rsa::kLbPrefix : rsa::kLbAggregate, epoch, incl ? pre[t] + cnt[t] : cnt[t]);
    }
    const uint32_t groups = (tiles + rsa::kLbGroup - 1) / rsa::kLbGroup;
    std::vector<W> grp(groups, 0);
    for (uint32_t h = 0; h < groups; ++h) {
        const bool complete = u(rng) < p_complete;
        for (uint32_t t = h * rsa::kLbGroup; t < (h + 1) * rsa::kLbGroup && t < tiles; ++t)
            if (complete || t + 1 < (h + 1) * rsa::kLbGroup) grp[h] += rsa::grp_add<W>(cnt[t]);
    }
    auto load = [&](int64_t t) { return lb[t]; };
    auto gload = [&](int64_t h) { return grp[h]; };
    auto backoff = [] {};
    int errors = 0;
    for (uint32_t t = 1; t < tiles; ++t) {
        uint32_t steps = 0, spins = 0;
        if (rsa::lookback_walk_grouped<W, 8, 4>(t, epoch, load, gload, backoff, steps, spins) != pre[t]) ++errors;
        W part[4];
        uint32_t st[4];
        st[0] = rsa::lb_segment_tiles<W>(t, epoch, load, part[0], steps);
        for (int sg = 1; sg < 4; ++sg)
            st[sg] = rsa::lb_segment_groups<W, 4>((int64_t)(t / rsa::kLbGroup) - 1 - (sg - 1) * 4, epoch, load, gload,
                                                 part[sg], steps);
        if (rsa::lookback_wide_finish<W, 8, 4>(t, epoch, part, st, load, gload, backoff, steps, spins) != pre[t])
            ++errors;
    }
    return errors;
}

int main() {
    int errors = 0;
    // segments passing through, ending and stopping at every position
    for (uint32_t seed = 1; seed <= 4; ++seed) {
        errors += fixed_states<uint32_t>(400, 0.0, 1.0, seed);
        errors += fixed_states<uint32_t>(400, 0.02, 0.9, seed);
        errors += fixed_states<uint64_t>(300, 0.1, 0.5, seed);
        errors += fixed_states<uint32_t>(301, 0.0, 0.0, seed);
    }
    for (uint32_t seed = 1; seed <= 3; ++seed) {
        for (int window : {1, 4, 8, 32, -1, -4, -5}) {
            errors += run<uint32_t>(300, 32, 5, 8, seed, window);
            errors += run<uint64_t>(200, 16, 4, 6, seed + 100, window);
        }
    }
    // longer chains of groups under more threads
    errors += run<uint32_t>(6000, 8, 3, 16, 7, -4);
    errors += run<uint64_t>(3000, 4, 3, 12, 8, -1);
    errors += run<uint32_t>(6000, 8, 3, 16, 9, -5);
    errors += run<uint32_t>(90, 64, 6, 64, 10, -5);  // few tiles, all in flight at once
    // the packed word keeps flag, epoch and the largest narrow value apart
    const uint32_t w = rsa::lb_pack<uint32_t>(rsa::kLbPrefix, 1, (uint32_t)rsa::kNarrowMaxN);
    if (rsa::lb_flag(w) != rsa::kLbPrefix || rsa::lb_epoch(w) != 1 || rsa::lb_value(w) != rsa::kNarrowMaxN) ++errors;
    if (rsa::lb_ready(0u, 0) || rsa::lb_ready(0u, 1)) ++errors;  // zeroed words are never ready
    if (rsa::lb_ready(rsa::lb_pack<uint32_t>(rsa::kLbAggregate, 0, 5), 1)) ++errors;  // stale epoch
    // a group word is complete only after exactly kLbGroup additions
    uint32_t gw = 0;
    for (int i = 0; i < rsa::kLbGroup; ++i) {
        if (rsa::grp_complete(gw)) ++errors;
        gw += rsa::grp_add<uint32_t>(20480u);
    }
    if (!rsa::grp_complete(gw) || rsa::grp_sum(gw) != 20480u * rsa::kLbGroup) ++errors;
    uint64_t gw64 = 0;
    for (int i = 0; i < rsa::kLbGroup; ++i) gw64 += rsa::grp_add<uint64_t>(1ull << 40);
    if (!rsa::grp_complete(gw64) || rsa::grp_sum(gw64) != ((uint64_t)rsa::kLbGroup << 40)) ++errors;
    // executed passes: every non-uniform pass runs; an odd count gains the
    // lowest uniform pass (run as a copy), so the count is even unless no pass
    // is uniform
    for (uint32_t passes : {1u, 2u, 4u, 8u, 16u}) {
        const uint32_t pmask = (1u << passes) - 1u;
        for (uint32_t skip = 0; skip <= pmask; skip += (passes == 16u ? 7u : 1u)) {
            const uint32_t active = ~skip & pmask, run = rsa::pass_run_mask(skip, passes);
            const uint32_t extra = run & ~active;
            if ((run & active) != active || (run & ~pmask) != 0u) ++errors;
            if (__builtin_popcount(run) & 1u) {
                if (active != pmask) ++errors;
            }
            if (extra && (extra != (skip & (0u - skip)) || !(__builtin_popcount(active) & 1u))) ++errors;
        }
    }
    printf("lookback simulation: %d errors\n", errors);
    return errors == 0 ? 0 : 1;
}
