// Batched segmented sort against one plain sort call per segment, from host
// C++ over the C ABI (the caller a service would be).
//
//   seg_percall <sizes.bin> [--calls M] [--replays R] [--seed S]
//
// sizes.bin holds the segment lengths as little-endian uint64.  Keys are u32
// from the reference generator (rsort_init_keys random mode, seed S), values
// u32 = index.  Prints one JSON line:
//   batched:  R replays of one rsort_segmented_sort over every segment with
//             RSORT_FLAG_GRAPH (the first, direct call is a warm-up), timed
//             with hipEvents around the replays;
//   per_call: one rsort_sort (u32/u32 pairs) per segment for the first M
//             segments, back to back on one stream with one workspace sized
//             for the longest, timed with hipEvents around the whole loop
//             (and the host clock);
// plus the keys/s of each and their ratio.  Both results are checked on the
// device: every segment in order, every value pointing at its key.
#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/rsort_c.h"

namespace {

#define CHECK_HIP(x)                                                             \
    do {                                                                         \
        const hipError_t e_ = (x);                                               \
        if (e_ != hipSuccess) {                                                  \
            fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));              \
            exit(2);                                                             \
        }                                                                        \
    } while (0)
#define CHECK_RS(x)                                                              \
    do {                                                                         \
        const rsort_status s_ = (x);                                             \
        if (s_ != RSORT_SUCCESS) {                                               \
            fprintf(stderr, "%s: %s (%s)\n", #x, rsort_status_string(s_), rsort_last_error()); \
            exit(2);                                                             \
        }                                                                        \
    } while (0)

// Every segment sorted and every value the index of its key: checked per
// segment with the library's device checks (order; value range, key
// identity, stability), one violation counter.
uint64_t check(const std::vector<uint64_t>& off, size_t segs, const uint32_t* d_orig, const uint32_t* d_keys,
               const uint32_t* d_vals, unsigned long long* d_bad, hipStream_t s) {
    const rsort_desc d{RSORT_KEY_UNSIGNED, 4, 4, 0, 0, 0};
    CHECK_HIP(hipMemsetAsync(d_bad, 0, 8, s));
    for (size_t i = 0; i < segs; ++i) {
        const uint64_t b = off[i], m = off[i + 1] - off[i];
        CHECK_RS(rsort_check_sorted(&d, d_keys + b, m, d_bad, s));
    }
    // values are global indices: check the pairing against the whole original array
    CHECK_RS(rsort_check_pairings(4, 4, d_orig, d_keys, d_vals, off[segs], d_bad, s));
    unsigned long long bad = 0;
    CHECK_HIP(hipMemcpyAsync(&bad, d_bad, 8, hipMemcpyDeviceToHost, s));
    CHECK_HIP(hipStreamSynchronize(s));
    return bad;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        fprintf(stderr, "usage: %s sizes.bin [--calls M] [--replays R] [--seed S]\n", argv[0]);
        return 1;
    }
    size_t calls = 20000;
    int replays = 10;
    uint32_t seed = 7;
    for (int i = 2; i + 1 < argc; i += 2) {
        if (!strcmp(argv[i], "--calls")) calls = strtoull(argv[i + 1], nullptr, 10);
        else if (!strcmp(argv[i], "--replays")) replays = atoi(argv[i + 1]);
        else if (!strcmp(argv[i], "--seed")) seed = (uint32_t)strtoul(argv[i + 1], nullptr, 10);
    }
    FILE* f = fopen(argv[1], "rb");
    if (!f) {
        perror(argv[1]);
        return 1;
    }
    std::vector<uint64_t> sizes;
    uint64_t v;
    while (fread(&v, 8, 1, f) == 1) sizes.push_back(v);
    fclose(f);
    const size_t segs = sizes.size();
    if (segs == 0 || replays < 1) return 1;
    if (calls > segs) calls = segs;
    std::vector<uint64_t> off(segs + 1, 0);
    uint64_t longest = 0;
    for (size_t i = 0; i < segs; ++i) {
        off[i + 1] = off[i] + sizes[i];
        if (i < calls && sizes[i] > longest) longest = sizes[i];
    }
    const uint64_t n = off[segs];

    hipStream_t s;
    CHECK_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    uint32_t *d_orig, *d_keys, *d_vals;
    uint64_t* d_off;
    unsigned long long* d_bad;
    CHECK_HIP(hipMalloc(&d_orig, n * 4));
    CHECK_HIP(hipMalloc(&d_keys, n * 4));
    CHECK_HIP(hipMalloc(&d_vals, n * 4));
    CHECK_HIP(hipMalloc(&d_off, (segs + 1) * 8));
    CHECK_HIP(hipMalloc(&d_bad, 8));
    CHECK_HIP(hipMemcpy(d_off, off.data(), (segs + 1) * 8, hipMemcpyHostToDevice));
    CHECK_RS(rsort_init_keys(4, 1, seed, d_orig, n, s));
    std::vector<uint32_t> idx(n);
    for (uint64_t i = 0; i < n; ++i) idx[i] = (uint32_t)i;
    auto reset = [&]() {
        CHECK_HIP(hipMemcpyAsync(d_keys, d_orig, n * 4, hipMemcpyDeviceToDevice, s));
        CHECK_HIP(hipMemcpyAsync(d_vals, idx.data(), n * 4, hipMemcpyHostToDevice, s));
        CHECK_HIP(hipStreamSynchronize(s));
    };
    hipEvent_t e0, e1;
    CHECK_HIP(hipEventCreate(&e0));
    CHECK_HIP(hipEventCreate(&e1));

    // batched: one segmented sort of every segment, replayed from its graph
    const rsort_desc bd{RSORT_KEY_UNSIGNED, 4, 4, 0, RSORT_FLAG_GRAPH, 0};
    size_t tb = 0;
    CHECK_RS(rsort_segmented_sort(&bd, nullptr, nullptr, n, nullptr, 8, segs, nullptr, &tb, s));
    void* ws;
    CHECK_HIP(hipMalloc(&ws, tb));
    reset();
    CHECK_RS(rsort_segmented_sort(&bd, d_keys, d_vals, n, d_off, 8, segs, ws, &tb, s));  // direct + capture
    CHECK_HIP(hipStreamSynchronize(s));
    float batched_ms = 0.f;
    for (int r = 0; r < replays; ++r) {
        reset();
        CHECK_HIP(hipEventRecord(e0, s));
        CHECK_RS(rsort_segmented_sort(&bd, d_keys, d_vals, n, d_off, 8, segs, ws, &tb, s));
        CHECK_HIP(hipEventRecord(e1, s));
        CHECK_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        CHECK_HIP(hipEventElapsedTime(&ms, e0, e1));
        batched_ms += ms;
    }
    batched_ms /= (float)replays;
    const uint64_t bad_batched = check(off, segs, d_orig, d_keys, d_vals, d_bad, s);
    CHECK_HIP(hipFree(ws));

    // per call: one rsort_sort per segment, back to back
    const rsort_desc pd{RSORT_KEY_UNSIGNED, 4, 4, 0, 0, 0};
    size_t ptb = 0;
    CHECK_RS(rsort_sort(&pd, nullptr, nullptr, nullptr, &ptb, longest, s));
    CHECK_HIP(hipMalloc(&ws, ptb));
    reset();
    for (size_t i = 0; i < 50 && i < calls; ++i) {  // warm-up: first-use costs of the kernels
        size_t t = ptb;
        CHECK_RS(rsort_sort(&pd, d_keys + off[i], d_vals + off[i], ws, &t, sizes[i], s));
    }
    reset();
    const auto t0 = std::chrono::steady_clock::now();
    CHECK_HIP(hipEventRecord(e0, s));
    for (size_t i = 0; i < calls; ++i) {
        size_t t = ptb;
        CHECK_RS(rsort_sort(&pd, d_keys + off[i], d_vals + off[i], ws, &t, sizes[i], s));
    }
    CHECK_HIP(hipEventRecord(e1, s));
    CHECK_HIP(hipEventSynchronize(e1));
    const double wall_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    float per_call_ms = 0.f;
    CHECK_HIP(hipEventElapsedTime(&per_call_ms, e0, e1));
    const uint64_t bad_per_call = check(off, calls, d_orig, d_keys, d_vals, d_bad, s);
    const uint64_t call_keys = off[calls];

    const double batched_gks = (double)n / (batched_ms * 1e-3) / 1e9;
    const double per_call_gks = (double)call_keys / (per_call_ms * 1e-3) / 1e9;
    printf("{\"segments\": %zu, \"keys\": %llu, \"batched\": {\"ms\": %.4f, \"gkeys_per_s\": %.3f, \"replays\": %d, "
           "\"violations\": %llu}, \"per_call\": {\"calls\": %zu, \"keys\": %llu, \"ms\": %.3f, \"wall_ms\": %.3f, "
           "\"us_per_call\": %.2f, \"gkeys_per_s\": %.5f, \"violations\": %llu}, \"batched_over_per_call\": %.1f}\n",
           segs, (unsigned long long)n, batched_ms, batched_gks, replays, (unsigned long long)bad_batched, calls,
           (unsigned long long)call_keys, per_call_ms, wall_s * 1e3, per_call_ms * 1e3 / (double)calls, per_call_gks,
           (unsigned long long)bad_per_call, batched_gks / per_call_gks);
    CHECK_HIP(hipFree(ws));
    CHECK_HIP(hipFree(d_orig));
    CHECK_HIP(hipFree(d_keys));
    CHECK_HIP(hipFree(d_vals));
    CHECK_HIP(hipFree(d_off));
    CHECK_HIP(hipFree(d_bad));
    return (bad_batched || bad_per_call) ? 3 : 0;
}
