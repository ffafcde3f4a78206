// Command-line benchmark and validation driver of the sort, with the
// reference driver's options (benchmark/bench_parser.cpp:11-106,
// benchmark/radix_gpu.cu): --n --iterations --warmup --warmup_ms --descending
// --validation --help.
//
//   benchmark (default): uint32 keys from the reference generator (random
//     mode, seed of run i = set_seed_radix(i)), regenerated before every
//     iteration outside the timed region; warm-up runs until both --warmup
//     iterations and --warmup_ms have passed; each timed iteration is bracketed
//     by hipEvents around rsort::sort (the reference times the call the same
//     way, benchmark/benchmark_sort.cuh:623-634).  After the last iteration the
//     result is checked on the device: order, per-pass digit histogram.
//   --validation: every key type x value type (none, u32, u8, u16, u64, u128) x
//     input mode (random, byte_skip, ascending) x order x size (10, 100, 1000
//     and 2^(b-10) .. 2^b keys, b = floor(log2(--n))), checked on the device
//     (order, digit histograms, and for 4- and 8-byte values the pairing with
//     the original index and stability) -- the reference's matrix
//     (benchmark/validate_sort.cuh:36-236), 6552 cases.
//
// Build: cuda-fast-sort_amd/build.py (g++ against librsort_amd.so).
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/rsort_c.h"

namespace {

struct Options {
    uint64_t n = 1ull << 27;
    uint32_t iterations = 30;
    uint32_t warmup = 10;
    uint32_t warmup_ms = 250;
    bool descending = false;
    bool validation = false;
};

void usage(const char* exe) {
    printf("Usage: %s [options]\n\n"
           "  --n <value>            number of keys (validation: largest size)\n"
           "  --iterations <value>   timed iterations\n"
           "  --warmup <value>       warm-up iterations\n"
           "  --warmup_ms <value>    minimum warm-up time in ms\n"
           "  --descending           sort descending\n"
           "  --validation           run the validation matrix\n"
           "  --help                 this text\n",
           exe);
}

bool parse_u64(const char* s, uint64_t lim, uint64_t* out) {
    if (!s || !*s) return false;
    char* end = nullptr;
    const unsigned long long v = strtoull(s, &end, 10);
    if (end == s || *end || v > lim) return false;
    *out = v;
    return true;
}

// 0 ok, 1 error, 2 help
int parse(int argc, char** argv, Options* o) {
    for (int i = 1; i < argc; ++i) {
        const char* a = argv[i];
        uint64_t v = 0;
        auto value = [&](uint64_t lim) {
            if (i + 1 >= argc || !parse_u64(argv[i + 1], lim, &v)) {
                printf("Error: missing or invalid value after %s\n", a);
                return false;
            }
            ++i;
            return true;
        };
        if (!strcmp(a, "--help")) return 2;
        if (!strcmp(a, "--descending")) {
            o->descending = true;
        } else if (!strcmp(a, "--validation")) {
            o->validation = true;
        } else if (!strcmp(a, "--n")) {
            if (!value(~0ull)) return 1;
            o->n = v;
        } else if (!strcmp(a, "--iterations")) {
            if (!value(0xFFFFFFFFull)) return 1;
            o->iterations = (uint32_t)v;
        } else if (!strcmp(a, "--warmup")) {
            if (!value(0xFFFFFFFFull)) return 1;
            o->warmup = (uint32_t)v;
        } else if (!strcmp(a, "--warmup_ms")) {
            if (!value(0xFFFFFFFFull)) return 1;
            o->warmup_ms = (uint32_t)v;
        } else {
            printf("Error: unknown argument: %s\n", a);
            return 1;
        }
    }
    if (o->n == 0) {
        printf("Error: --n must be at least 1\n");
        return 1;
    }
    return 0;
}

#define HIP_OK(x)                                                                        \
    do {                                                                                 \
        const hipError_t e_ = (x);                                                       \
        if (e_ != hipSuccess) {                                                          \
            fprintf(stderr, "%s:%d: %s\n", __FILE__, __LINE__, hipGetErrorString(e_));   \
            exit(3);                                                                     \
        }                                                                                \
    } while (0)

// Device buffers of one case, grown on demand and reused.
struct Buffers {
    void *keys = nullptr, *vals = nullptr, *orig = nullptr, *ws = nullptr;
    unsigned long long* counters = nullptr;  // [0] violations, [1..] two digit histograms
    size_t keys_b = 0, orig_b = 0, vals_b = 0, ws_b = 0;
    static constexpr size_t kCounterWords = 1 + 2 * 16 * 256;

    static void grow(void** p, size_t* have, size_t need) {
        if (need <= *have) return;
        if (*p) HIP_OK(hipFree(*p));
        HIP_OK(hipMalloc(p, need));
        *have = need;
    }
    void ensure(size_t key_b, size_t val_b, size_t ws_need) {
        grow(&keys, &keys_b, key_b);
        grow(&orig, &orig_b, key_b);
        grow(&vals, &vals_b, val_b ? val_b : 16);
        grow(&ws, &ws_b, ws_need ? ws_need : 16);
        if (!counters) HIP_OK(hipMalloc((void**)&counters, kCounterWords * 8));
    }
};

struct Checks {
    bool order = false, hist = false, pairs = true;
};

// Device checks of a finished sort: order, digit histogram (before in
// counters[1..], after in counters[1+P*256..]), pairings for 4/8-byte values.
Checks check_result(const rsort_desc& d, Buffers& b, uint64_t n, hipStream_t s) {
    Checks c;
    const size_t hist_words = (size_t)d.key_bytes * 256u;
    HIP_OK(hipMemsetAsync(b.counters, 0, 8, s));
    HIP_OK(hipMemsetAsync(b.counters + 1 + hist_words, 0, hist_words * 8, s));
    if (rsort_check_sorted(&d, b.keys, n, b.counters, s) != RSORT_SUCCESS) return c;
    if (rsort_digit_counts(&d, b.keys, n, b.counters + 1 + hist_words, s) != RSORT_SUCCESS) return c;
    std::vector<unsigned long long> h(1 + 2 * hist_words);
    HIP_OK(hipMemcpyAsync(h.data(), b.counters, h.size() * 8, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    c.order = h[0] == 0;
    c.hist = std::equal(h.begin() + 1, h.begin() + 1 + (long)hist_words, h.begin() + 1 + (long)hist_words);
    if (d.value_bytes == 8 || (d.value_bytes == 4 && n < 0xFFFFFFFFull)) {
        c.pairs = false;
        HIP_OK(hipMemsetAsync(b.counters, 0, 8, s));
        if (rsort_check_pairings(d.key_bytes, d.value_bytes, b.orig, b.keys, b.vals, n, b.counters, s) !=
            RSORT_SUCCESS)
            return c;
        unsigned long long bad = 1;
        HIP_OK(hipMemcpyAsync(&bad, b.counters, 8, hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
        c.pairs = bad == 0;
    }
    return c;
}

// Generates keys (and index values), records the digit histogram of the input.
bool prepare(const rsort_desc& d, Buffers& b, uint64_t n, uint32_t mode, uint32_t seed, hipStream_t s) {
    const size_t hist_words = (size_t)d.key_bytes * 256u;
    if (rsort_init_keys(d.key_bytes, mode, seed, b.keys, n, s) != RSORT_SUCCESS) return false;
    if (d.value_bytes && rsort_init_keys(d.value_bytes, 3, 0, b.vals, n, s) != RSORT_SUCCESS) return false;
    HIP_OK(hipMemcpyAsync(b.orig, b.keys, n * d.key_bytes, hipMemcpyDeviceToDevice, s));
    HIP_OK(hipMemsetAsync(b.counters + 1, 0, hist_words * 8, s));
    return rsort_digit_counts(&d, b.keys, n, b.counters + 1, s) == RSORT_SUCCESS;
}

const char* mode_name(uint32_t m) { return m == 1 ? "random" : m == 2 ? "byte_skip" : "ascending"; }

int benchmark(const Options& o) {
    hipStream_t s;
    HIP_OK(hipStreamCreate(&s));
    rsort_desc d{};
    d.key_kind = RSORT_KEY_UNSIGNED;
    d.key_bytes = 4;
    d.descending = o.descending;
    size_t tb = 0;
    if (rsort_sort(&d, nullptr, nullptr, nullptr, &tb, o.n, nullptr) != RSORT_SUCCESS) {
        printf("size query failed: %s\n", rsort_last_error());
        return 1;
    }
    Buffers b;
    b.ensure(o.n * 4, 0, tb);
    hipEvent_t e0, e1;
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
    uint32_t seed_i = 0, warm = 0;
    const auto t0 = std::chrono::steady_clock::now();
    auto run = [&](float* ms) {
        if (!prepare(d, b, o.n, 1, rsort_set_seed_radix((int32_t)seed_i++), s)) return false;
        HIP_OK(hipEventRecord(e0, s));
        size_t t = tb;
        if (rsort_sort(&d, b.keys, nullptr, b.ws, &t, o.n, s) != RSORT_SUCCESS) return false;
        HIP_OK(hipEventRecord(e1, s));
        HIP_OK(hipEventSynchronize(e1));
        if (ms) HIP_OK(hipEventElapsedTime(ms, e0, e1));
        return true;
    };
    while (warm < o.warmup ||
           std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() < o.warmup_ms) {
        if (!run(nullptr)) {
            printf("sort failed: %s\n", rsort_last_error());
            return 1;
        }
        ++warm;
    }
    std::vector<float> ms(o.iterations);
    for (auto& m : ms)
        if (!run(&m)) {
            printf("sort failed: %s\n", rsort_last_error());
            return 1;
        }
    const Checks c = ms.empty() ? Checks{} : check_result(d, b, o.n, s);
    double mean = 0, var = 0;
    for (float m : ms) mean += m;
    mean /= std::max<size_t>(1, ms.size());
    for (float m : ms) var += (m - mean) * (m - mean);
    const double sd_us = ms.size() > 1 ? std::sqrt(var / (ms.size() - 1)) * 1e3 : 0.0;
    rsort_plan_info p{};
    (void)rsort_plan(&d, o.n, &p);
    hipDeviceProp_t prop;
    HIP_OK(hipGetDeviceProperties(&prop, 0));
    printf("\nGPU        %s (%s)\n", prop.name[0] ? prop.name : "device 0", prop.gcnArchName);
    printf("n          %llu uint32 keys, %s input, %s\n", (unsigned long long)o.n, mode_name(1),
           o.descending ? "descending" : "ascending");
    printf("tile       %u threads x %u keys, %llu tiles, %u passes, %s lookback words\n", p.threads,
           p.items_per_thread, (unsigned long long)p.num_tiles, p.passes, p.wide ? "64-bit" : "32-bit");
    printf("workspace  %.3f MB\n", (double)tb / (1024.0 * 1024.0));
    printf("time       %.4f ms (+- %.3f us) over %u iterations after %u warm-up runs\n", mean, sd_us,
           (unsigned)ms.size(), warm);
    printf("rate       %.3f GKeys/s, %.3f ps per key\n", o.n / (mean * 1e-3) / 1e9, mean * 1e9 / o.n);
    printf("checks     sorted %s, digit histograms %s\n", c.order ? "yes" : "NO", c.hist ? "yes" : "NO");
    return c.order && c.hist ? 0 : 1;
}

struct TypeEntry {
    const char* name;
    uint32_t kind, bytes;
};

int validation(const Options& o) {
    static const TypeEntry keys[] = {
        {"uint32_t", 0, 4}, {"uint8_t", 0, 1},  {"uint16_t", 0, 2}, {"uint64_t", 0, 8}, {"unsigned __int128", 0, 16},
        {"int16_t", 1, 2},  {"int8_t", 1, 1},   {"int32_t", 1, 4},  {"int64_t", 1, 8},  {"__int128", 1, 16},
        {"float", 2, 4},    {"double", 2, 8},   {"long double", 2, 16}};
    static const TypeEntry vals[] = {{"none", 0, 0},     {"uint32_t", 0, 4}, {"uint8_t", 0, 1},
                                     {"uint16_t", 0, 2}, {"uint64_t", 0, 8}, {"unsigned __int128", 0, 16}};
    const uint32_t bit_end = o.n >= 2 ? (uint32_t)std::floor(std::log2((double)o.n)) : 1u;
    const uint32_t bit_start = bit_end >= 10 ? bit_end - 10 : 0;
    std::vector<uint64_t> sizes = {10, 100, 1000};
    for (uint32_t b = bit_start; b <= bit_end; ++b) sizes.push_back(1ull << b);
    hipStream_t s;
    HIP_OK(hipStreamCreate(&s));
    Buffers buf;
    uint32_t tests = 0, passed = 0, seed_i = 0;
    printf("validation: %zu key types x %zu value types x 3 modes x 2 orders x %zu sizes (10 .. 2^%u)\n",
           sizeof(keys) / sizeof(keys[0]), sizeof(vals) / sizeof(vals[0]), sizes.size(), bit_end);
    fflush(stdout);
    for (uint32_t mode = 1; mode <= 3; ++mode) {
        for (uint32_t desc = 0; desc < 2; ++desc) {
            for (const auto& k : keys) {
                for (const auto& v : vals) {
                    for (uint64_t n : sizes) {
                        rsort_desc d{};
                        d.key_kind = k.kind;
                        d.key_bytes = k.bytes;
                        d.value_bytes = v.bytes;
                        d.descending = desc;
                        size_t tb = 0;
                        bool ok = rsort_sort(&d, nullptr, nullptr, nullptr, &tb, n, nullptr) == RSORT_SUCCESS;
                        if (ok) {
                            buf.ensure(n * k.bytes, n * v.bytes, tb);
                            ok = prepare(d, buf, n, mode, rsort_set_seed_radix((int32_t)seed_i++), s);
                        }
                        size_t t = tb;
                        ok = ok && rsort_sort(&d, buf.keys, v.bytes ? buf.vals : nullptr, buf.ws, &t, n, s) ==
                                       RSORT_SUCCESS;
                        Checks c;
                        if (ok) c = check_result(d, buf, n, s);
                        ok = ok && c.order && c.hist && c.pairs;
                        ++tests;
                        passed += ok ? 1 : 0;
                        if (!ok)
                            printf("FAILED: n = %llu, %s, %s input, key %s, value %s [order %d, histogram %d, "
                                   "pairs %d] %s\n",
                                   (unsigned long long)n, desc ? "DESC" : "ASC", mode_name(mode), k.name, v.name,
                                   c.order, c.hist, c.pairs, rsort_last_error());
                    }
                }
                printf("  %s input, %s, key %-18s: %u/%u passed so far\n", mode_name(mode), desc ? "DESC" : "ASC",
                       k.name, passed, tests);
                fflush(stdout);
            }
        }
    }
    printf("\nDone... passed %u/%u tests.\n", passed, tests);
    return passed == tests ? 0 : 1;
}

}  // namespace

int main(int argc, char** argv) {
    Options o;
    const int r = parse(argc, argv, &o);
    if (r) {
        if (r == 1) printf("\n");
        usage(argv[0]);
        return r == 2 ? 0 : 1;
    }
    return o.validation ? validation(o) : benchmark(o);
}
