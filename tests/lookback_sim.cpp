// CPU simulation of the tile-ticket + decoupled lookback protocol of the
// onesweep pass (rsort_protocol.hpp word layout, k_onesweep's publish /
// lookback order).  Host threads play workgroups: each takes a ticket, publishes
// its aggregate, walks back until an inclusive prefix, publishes its own
// prefix.  Several executed passes reuse one lookback array with the epoch
// parity and no clearing in between; words start zeroed once per sort.  The
// walk is rsa::lookback_walk, the function the kernel runs, with first/maximum
// windows of 1/1, 1/4, 3/8 and 4/32 words, and rsa::lookback_walk_grouped
// (window < 0 here): tiles also add their aggregate to a per-pass group word
// with an atomic, and the walk skips complete groups.
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

int main() {
    int errors = 0;
    for (uint32_t seed = 1; seed <= 3; ++seed) {
        for (int window : {1, 4, 8, 32, -1, -4}) {
            errors += run<uint32_t>(300, 32, 5, 8, seed, window);
            errors += run<uint64_t>(200, 16, 4, 6, seed + 100, window);
        }
    }
    // longer chains of groups under more threads
    errors += run<uint32_t>(6000, 8, 3, 16, 7, -4);
    errors += run<uint64_t>(3000, 4, 3, 12, 8, -1);
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
    printf("lookback simulation: %d errors\n", errors);
    return errors == 0 ? 0 : 1;
}
