// Drop-in check of <rsort/rsort.hpp>: the four templates, called exactly as
// the reference's examples call them (examples/example_basic.cu:27-43:
// size query with a null workspace, hipMalloc, sort), for several key and
// value types; each result is compared with std::stable_sort on the host.
// Exit status 0 when every case matches.
#include <hip/hip_runtime_api.h>
#include <rsort/rsort.hpp>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>

#define CHECK_HIP(x)                                                                           \
    do {                                                                                       \
        hipError_t e_ = (x);                                                                   \
        if (e_ != hipSuccess) {                                                                \
            fprintf(stderr, "%s:%d %s\n", __FILE__, __LINE__, hipGetErrorString(e_));          \
            exit(2);                                                                           \
        }                                                                                      \
    } while (0)

static int failures = 0;

// order used by the host check: numeric for integers and floats (no NaN in
// the inputs); for 16-byte long double the sign-magnitude 128-bit pattern
template <typename K>
static bool less_key(const K& a, const K& b) {
    if constexpr (std::is_same_v<K, long double>) {
        unsigned __int128 x, y;
        memcpy(&x, &a, 16);
        memcpy(&y, &b, 16);
        const unsigned __int128 s = (unsigned __int128)1 << 127;
        const bool nx = x & s, ny = y & s;
        if (nx != ny) return nx;
        return nx ? (x & ~s) > (y & ~s) : x < y;
    } else {
        return a < b;
    }
}

template <typename K>
static K make_key(std::mt19937_64& g) {
    if constexpr (rsort::detail::is_float16<K>::value) {
        // multiples of 1/8 in [-125, 125]: exact in half precision, many duplicates
        return (K)(((double)(int64_t)(g() % 2001) - 1000.0) / 8.0);
    } else if constexpr (std::is_floating_point_v<K>) {
        K v;
        memset(&v, 0, sizeof(K));
        v = (K)((double)(int64_t)(g() % 2001) - 1000.0) / (K)8;  // duplicates and +-values
        if constexpr (std::is_same_v<K, long double>) {
            unsigned char b[16];
            memcpy(b, &v, 16);
            memset(b + 10, 0, 6);  // x87 padding bytes are unspecified: clear them
            memcpy(&v, b, 16);
        }
        return v;
    } else {
        return (K)(g() % 5000) * (K)(sizeof(K) > 1 ? 977 : 1);
    }
}

template <typename K, typename V, bool Desc, bool Pairs>
static void run_case(const char* name, size_t n) {
    std::mt19937_64 g(n * 131 + sizeof(K) * 7 + Desc + 2 * Pairs);
    std::vector<K> hk(n);
    std::vector<V> hv(n);
    for (size_t i = 0; i < n; ++i) {
        hk[i] = make_key<K>(g);
        hv[i] = (V)i;
    }
    K* dk = nullptr;
    V* dv = nullptr;
    uint8_t* ws = nullptr;
    size_t temp_bytes = 0;
    CHECK_HIP(hipMalloc(&dk, n * sizeof(K)));
    CHECK_HIP(hipMalloc(&dv, n * sizeof(V)));
    CHECK_HIP(hipMemcpy(dk, hk.data(), n * sizeof(K), hipMemcpyHostToDevice));
    CHECK_HIP(hipMemcpy(dv, hv.data(), n * sizeof(V), hipMemcpyHostToDevice));
    // size query, then the sort
    for (int phase = 0; phase < 2; ++phase) {
        if constexpr (Pairs) {
            if constexpr (Desc) rsort::sort_pairs_descending(dk, dv, &temp_bytes, ws, n);
            else rsort::sort_pairs(dk, dv, &temp_bytes, ws, n);
        } else {
            if constexpr (Desc) rsort::sort_descending(dk, &temp_bytes, ws, n);
            else rsort::sort(dk, &temp_bytes, ws, n);
        }
        if (rsort::last_status() != RSORT_SUCCESS) {
            printf("FAIL %s: %s\n", name, rsort::last_error());
            ++failures;
            return;
        }
        if (phase == 0) CHECK_HIP(hipMalloc(&ws, temp_bytes));
    }
    CHECK_HIP(hipDeviceSynchronize());
    std::vector<K> rk(n);
    std::vector<V> rv(n);
    CHECK_HIP(hipMemcpy(rk.data(), dk, n * sizeof(K), hipMemcpyDeviceToHost));
    CHECK_HIP(hipMemcpy(rv.data(), dv, n * sizeof(V), hipMemcpyDeviceToHost));
    std::vector<size_t> idx(n);
    std::iota(idx.begin(), idx.end(), (size_t)0);
    std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) {
        return Desc ? less_key(hk[b], hk[a]) : less_key(hk[a], hk[b]);
    });
    bool ok = true;
    for (size_t i = 0; i < n && ok; ++i) {
        ok = memcmp(&rk[i], &hk[idx[i]], sizeof(K)) == 0;
        if (Pairs) ok = ok && rv[i] == (V)idx[i];
        else ok = ok && rv[i] == (V)i;  // values untouched by key-only sorts
    }
    printf("%s %s n=%zu\n", ok ? "ok  " : "FAIL", name, n);
    failures += ok ? 0 : 1;
    CHECK_HIP(hipFree(ws));
    CHECK_HIP(hipFree(dv));
    CHECK_HIP(hipFree(dk));
}

// Array-of-structs sort (rsort::sort_records): a 24-byte record sorted by a
// float member (ascending) or by a uint16_t member (descending); each record
// must equal the std::stable_sort of the host records by that member.
struct Particle {
    double x;
    float y;
    uint16_t tag;
    uint32_t id;
};

template <bool ByTag>
static void run_records(const char* name, size_t n) {
    std::mt19937_64 g(n * 17 + ByTag);
    std::vector<Particle> h(n);
    for (size_t i = 0; i < n; ++i) {
        h[i].x = (double)(g() % 1000);
        h[i].y = (float)((int)(g() % 2001) - 1000) * 0.5f;
        h[i].tag = (uint16_t)(g() % 300);
        h[i].id = (uint32_t)i;
    }
    Particle* d = nullptr;
    uint8_t* ws = nullptr;
    size_t temp_bytes = 0;
    CHECK_HIP(hipMalloc(&d, n * sizeof(Particle)));
    CHECK_HIP(hipMemcpy(d, h.data(), n * sizeof(Particle), hipMemcpyHostToDevice));
    for (int phase = 0; phase < 2; ++phase) {
        if constexpr (ByTag) rsort::sort_records_descending(d, &Particle::tag, &temp_bytes, ws, n);
        else rsort::sort_records(d, &Particle::y, &temp_bytes, ws, n);
        if (rsort::last_status() != RSORT_SUCCESS) {
            printf("FAIL %s: %s\n", name, rsort::last_error());
            ++failures;
            return;
        }
        if (phase == 0) CHECK_HIP(hipMalloc(&ws, temp_bytes));
    }
    CHECK_HIP(hipDeviceSynchronize());
    std::vector<Particle> r(n);
    CHECK_HIP(hipMemcpy(r.data(), d, n * sizeof(Particle), hipMemcpyDeviceToHost));
    std::stable_sort(h.begin(), h.end(), [](const Particle& a, const Particle& b) {
        return ByTag ? a.tag > b.tag : a.y < b.y;
    });
    const bool ok = memcmp(r.data(), h.data(), n * sizeof(Particle)) == 0;
    printf("%s %s n=%zu\n", ok ? "ok  " : "FAIL", name, n);
    failures += ok ? 0 : 1;
    CHECK_HIP(hipFree(ws));
    CHECK_HIP(hipFree(d));
}

int main() {
#if defined(__FLT16_MAX__)
    printf("compiler has _Float16: half keys tested\n");
#else
    printf("compiler without _Float16 in C++: half keys through the C ABI only\n");
#endif
    for (size_t n : {(size_t)1000, (size_t)123457}) {
        run_case<uint32_t, uint32_t, false, false>("sort<uint32_t>", n);
        run_case<int32_t, uint32_t, true, false>("sort_descending<int32_t>", n);
        run_case<uint32_t, uint32_t, false, true>("sort_pairs<uint32_t,uint32_t>", n);
        run_case<int64_t, uint64_t, true, true>("sort_pairs_descending<int64_t,uint64_t>", n);
        run_case<float, uint32_t, false, true>("sort_pairs<float,uint32_t>", n);
        run_case<double, uint64_t, true, true>("sort_pairs_descending<double,uint64_t>", n);
        run_case<int8_t, uint16_t, false, true>("sort_pairs<int8_t,uint16_t>", n);
        run_case<uint16_t, uint8_t, true, true>("sort_pairs_descending<uint16_t,uint8_t>", n);
        run_case<unsigned __int128, uint64_t, false, true>("sort_pairs<unsigned __int128,uint64_t>", n);
        run_case<__int128, uint32_t, true, true>("sort_pairs_descending<__int128,uint32_t>", n);
        run_case<long double, uint32_t, false, true>("sort_pairs<long double,uint32_t>", n);
#if defined(__FLT16_MAX__)
        // half keys through the templates (compilers with _Float16 in C++)
        run_case<_Float16, uint32_t, false, true>("sort_pairs<_Float16,uint32_t>", n);
        run_case<_Float16, uint32_t, true, false>("sort_descending<_Float16>", n);
#endif
    }
    for (size_t n : {(size_t)1000, (size_t)200003}) {
        run_records<false>("sort_records<Particle, &Particle::y>", n);
        run_records<true>("sort_records_descending<Particle, &Particle::tag>", n);
    }
    printf("%d failures\n", failures);
    return failures ? 1 : 0;
}
