// CPU model of pass kernels sharing one device: 8 XCDs with a fixed number of
// workgroup slots each; workgroup b of a kernel is placed on XCD b mod 8 and
// each XCD dispatches its share of a kernel in index order, choosing freely
// between the kernels that have workgroups left for it.  A dispatched
// workgroup takes its tile ticket at once (rsa::ticket_tile: the per-XCD
// counters, runs of 16 tiles, or the global ticket) and then waits, holding
// its slot, until every lower tile of its kernel has been handed out (its
// lookback walk needs their aggregates); then it finishes and frees the slot.
//
// Three cases, each over many schedules of 2-4 overlapping kernels:
//   xcd     every kernel on per-XCD tickets: an adversarial schedule (XCDs
//           0-3 favour one kernel, 4-7 the other) deadlocks -- the hazard;
//   lock    the rule the library follows: a kernel uses per-XCD tickets only
//           if it takes the device's lock when it starts (else the global
//           ticket) and frees it when its last tile finishes: never deadlocks;
//   global  every kernel on the global ticket: never deadlocks.
// Two more put a histogram kernel that clears the control header itself
// (k_histogram without a fill kernel) beside a pass kernel holding the lock:
//   hist-wait   every histogram workgroup waits until workgroup 0 has cleared
//               (workgroup 0 goes first only on its own XCD): the adversarial
//               schedule (XCD 0 favours the pass kernel) deadlocks;
//   hist-claim  the library's protocol: workgroup 0 claims the clearing when
//               it starts, and a workgroup that needs the counts before anyone
//               has claimed claims it itself; a claimer is running and waits on
//               nothing, so no histogram workgroup waits for long: never
//               deadlocks.
// Prints the deadlock counts; exit status 0 when they are as stated.
#include <stdint.h>
#include <stdio.h>

#include <random>
#include <vector>

#include "../cuda-fast-sort_amd/csrc/rsort_protocol.hpp"

namespace {

constexpr uint32_t kXcds = 8;
constexpr uint32_t kParts = 8;
constexpr uint32_t kChunk = 16;

struct Kernel {
    uint32_t tiles = 0, grid = 0;
    bool hist = false;       // a histogram kernel (no tickets): workgroup b is "tile" b
    bool cleared = false;    // its control header is cleared for this call
    uint64_t arrive = 0;  // step at which it is launched
    bool started = false, xcd = false, holds_lock = false;
    uint32_t counter[kParts] = {0};
    std::vector<uint8_t> claimed;
    uint32_t lowest_unclaimed = 0;
    uint32_t next_wg[kXcds] = {0};  // next workgroup index to dispatch on each XCD
    uint32_t finished = 0;          // tiles finished
    bool done() const { return finished == tiles; }
    bool has_work(uint32_t x) const { return started && next_wg[x] < grid; }
};

struct Slot {
    int k = -1;  // kernel, -1 = free
    uint32_t tile = 0;
};

enum Mode { kAllXcd, kLock, kAllGlobal, kHistWait, kHistClaim };

// Returns true when the schedule deadlocks.
bool simulate(Mode mode, uint32_t slots_per_xcd, std::vector<Kernel> ks, bool adversarial, uint32_t seed) {
    std::mt19937 rng(seed);
    std::vector<Slot> slots(kXcds * slots_per_xcd);
    bool lock = false;
    uint64_t step = 0;
    for (auto& k : ks) {
        k.claimed.assign(k.tiles, 0);
        for (uint32_t x = 0; x < kXcds; ++x) k.next_wg[x] = x;
    }
    auto start_due = [&]() {
        for (auto& k : ks) {
            if (k.started || step < k.arrive) continue;
            k.started = true;
            if (k.hist) continue;
            if (mode == kAllXcd) k.xcd = true;
            else if (mode != kAllGlobal && !lock) k.xcd = k.holds_lock = lock = true;
        }
    };
    auto take_ticket = [&](Kernel& k, uint32_t b) -> uint32_t {
        const uint32_t parts = k.xcd ? kParts : 1u, chunk = k.xcd ? kChunk : 1u;
        const uint32_t part = b % parts;
        const uint32_t t = rsa::ticket_tile(k.counter[part]++, part, parts, chunk);
        if (t < k.tiles) {
            k.claimed[t] = 1;
            while (k.lowest_unclaimed < k.tiles && k.claimed[k.lowest_unclaimed]) ++k.lowest_unclaimed;
        }
        return t;
    };
    auto finish_tile = [&](Kernel& k, uint32_t t) {
        ++k.finished;
        if (t + 1 == k.tiles && k.holds_lock) k.holds_lock = lock = false;  // the last tile frees the lock
    };
    for (;;) {
        start_due();
        bool all_done = true;
        uint64_t next_arrival = UINT64_MAX;
        for (auto& k : ks) {
            all_done = all_done && k.started && k.done();
            if (!k.started && k.arrive < next_arrival) next_arrival = k.arrive;
        }
        if (all_done) return false;
        // candidate events: finish a workgroup whose lower tiles are all handed out,
        // or dispatch on an XCD with a free slot
        std::vector<uint32_t> runnable, dispatchable;
        for (uint32_t i = 0; i < slots.size(); ++i) {
            if (slots[i].k < 0) continue;
            const Kernel& k = ks[slots[i].k];
            // a histogram workgroup needs the cleared header; under the claim
            // protocol it clears it itself if nobody has claimed the clearing
            const bool ready = k.hist ? (k.cleared || mode == kHistClaim) : slots[i].tile < k.lowest_unclaimed;
            if (ready) runnable.push_back(i);
        }
        for (uint32_t x = 0; x < kXcds; ++x) {
            bool free = false, work = false;
            for (uint32_t s = 0; s < slots_per_xcd; ++s) free = free || slots[x * slots_per_xcd + s].k < 0;
            for (auto& k : ks) work = work || k.has_work(x);
            if (free && work) dispatchable.push_back(x);
        }
        if (runnable.empty() && dispatchable.empty()) {
            if (next_arrival == UINT64_MAX) return true;  // nothing can move: deadlock
            step = next_arrival;
            continue;
        }
        ++step;
        // the adversarial schedule dispatches greedily; otherwise a coin
        const bool do_dispatch =
            !dispatchable.empty() && (runnable.empty() || adversarial || (rng() & 1u));
        if (!do_dispatch) {
            const uint32_t i = runnable[rng() % runnable.size()];
            Kernel& k = ks[slots[i].k];
            if (k.hist) {
                k.cleared = true;  // it read the id, or claimed and cleared
                ++k.finished;
            } else {
                finish_tile(k, slots[i].tile);
            }
            slots[i].k = -1;
            continue;
        }
        const uint32_t x = dispatchable[rng() % dispatchable.size()];
        std::vector<int> cand;
        for (int i = 0; i < (int)ks.size(); ++i)
            if (ks[i].has_work(x)) cand.push_back(i);
        int ki = cand[rng() % cand.size()];
        if (adversarial) {
            // XCDs 0-3 favour kernel 0, XCDs 4-7 kernel 1
            const int fav = x < kXcds / 2 ? 0 : 1;
            for (int c : cand)
                if (c == fav) ki = c;
        }
        Kernel& k = ks[ki];
        const uint32_t b = k.next_wg[x];
        k.next_wg[x] += kXcds;
        if (k.hist) {
            if (b == 0) k.cleared = true;  // workgroup 0 clears when it starts
            for (uint32_t s = 0; s < slots_per_xcd; ++s) {
                Slot& sl = slots[x * slots_per_xcd + s];
                if (sl.k < 0) {
                    sl.k = ki;
                    sl.tile = b;
                    break;
                }
            }
            continue;
        }
        const uint32_t t = take_ticket(k, b);
        if (t >= k.tiles) continue;  // past the last tile: ends at once
        for (uint32_t s = 0; s < slots_per_xcd; ++s) {
            Slot& sl = slots[x * slots_per_xcd + s];
            if (sl.k < 0) {
                sl.k = ki;
                sl.tile = t;
                break;
            }
        }
    }
}

std::vector<Kernel> make_kernels(std::mt19937& rng, uint32_t count, bool same_start) {
    std::vector<Kernel> ks(count);
    for (auto& k : ks) {
        k.tiles = 200u + rng() % 1300u;
        const uint32_t round = kParts * kChunk;
        k.grid = (k.tiles + round - 1) / round * round;  // whole rounds, as launched
        k.arrive = same_start ? 0u : rng() % 3000u;
    }
    return ks;
}

}  // namespace

// A pass kernel and a histogram kernel of 16-256 workgroups that clears its
// own header (its grid is at most one workgroup per CU: 32 per XCD).
std::vector<Kernel> make_pass_and_hist(std::mt19937& rng, bool same_start) {
    std::vector<Kernel> ks = make_kernels(rng, 2u, same_start);
    ks[1].hist = true;
    ks[1].grid = ks[1].tiles = 16u + rng() % 241u;
    return ks;
}

int main() {
    int errors = 0;
    const char* names[] = {"xcd", "lock", "global", "hist-wait", "hist-claim"};
    for (Mode m : {kAllXcd, kLock, kAllGlobal, kHistWait, kHistClaim}) {
        uint32_t dead_adv = 0, dead_rand = 0, runs_adv = 0, runs_rand = 0;
        for (uint32_t seed = 1; seed <= 400; ++seed) {
            std::mt19937 rng(seed);
            const uint32_t spx = 4u + rng() % 13u;  // 4..16 slots per XCD
            const bool adversarial = seed % 2 == 0;
            const bool hist = m == kHistWait || m == kHistClaim;
            auto ks = hist ? make_pass_and_hist(rng, adversarial)
                           : make_kernels(rng, adversarial ? 2u : 2u + rng() % 3u, adversarial);
            const bool dead = simulate(m, spx, ks, adversarial, seed * 977u);
            (adversarial ? dead_adv : dead_rand) += dead;
            (adversarial ? runs_adv : runs_rand) += 1;
        }
        printf("%-6s deadlocks: adversarial %u/%u, random %u/%u\n", names[m], dead_adv, runs_adv, dead_rand,
               runs_rand);
        const bool hazard = m == kAllXcd || m == kHistWait;
        if (hazard && dead_adv == 0) ++errors;                 // the hazard must show
        if (!hazard && dead_adv + dead_rand != 0) ++errors;    // the rule must remove it
    }
    printf("dispatch simulation: %d errors\n", errors);
    return errors == 0 ? 0 : 1;
}
