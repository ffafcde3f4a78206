// CPU oracle for the rsort onesweep path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: it may be loaded by
// tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and by
// nothing else.  The product (librsort_amd.so) never links or calls it.
//
// It restates the reference's semantics on the CPU:
//   twiddle_in / twiddle_out   include/rsort/detail/tuning.cuh:132-165
//   long double as u128 word   include/rsort/detail/radix_kernel.cuh:351-356,
//                              include/rsort/detail/tuning.cuh:232-293
//   stable LSD sort result     include/rsort/detail/radix_kernel.cuh:336-474 (stable
//                              counting passes over bytes 0..P-1 = std::stable_sort
//                              on the twiddled key; descending = ascending on ~key)
//   one LSD pass              include/rsort/detail/radix_kernel.cuh:118-294
//   digit histogram           include/rsort/detail/histogram.cuh:88-211
//   mix32 / init_keys          include/rsort/detail/utils.cuh:227-317
//   set_seed_radix             include/rsort/detail/utils.cuh:435-437
//   check_sorted              include/rsort/detail/utils.cuh:341-365
//   check_pairings            include/rsort/detail/utils.cuh:377-432
// plus the workload generators of the benchmark configurations, written from
// raw std::mt19937 outputs (the std:: distributions are implementation-defined).
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <execution>
#include <numeric>
#include <random>
#include <thread>
#include <utility>
#include <vector>

typedef unsigned __int128 u128;

enum { KIND_UNSIGNED = 0, KIND_SIGNED = 1, KIND_FLOAT = 2 };

static inline u128 load_word(const uint8_t* p, int kb) {
    u128 x = 0;
    memcpy(&x, p, (size_t)kb);  // little-endian host
    return x;
}

static inline void store_word(uint8_t* p, u128 x, int kb) { memcpy(p, &x, (size_t)kb); }

static inline u128 ones_of(int kb) { return kb == 16 ? ~(u128)0 : (((u128)1 << (8 * kb)) - 1); }

// radix_traits<T>::twiddle_in<Descending> (tuning.cuh:132-148): floating keys
// flip all bits when negative, else the sign bit; signed keys flip the sign
// bit; descending complements the result.
static inline u128 twiddle_in(u128 x, int kind, int kb, int desc) {
    const u128 ones = ones_of(kb);
    const u128 sign = (u128)1 << (8 * kb - 1);
    if (kind == KIND_FLOAT) {
        x = (x & sign) ? (~x & ones) : (x ^ sign);
    } else if (kind == KIND_SIGNED) {
        x ^= sign;
    }
    if (desc) x = ~x & ones;
    return x;
}

// radix_traits<T>::twiddle_out<Descending> (tuning.cuh:151-165)
static inline u128 twiddle_out(u128 x, int kind, int kb, int desc) {
    const u128 ones = ones_of(kb);
    const u128 sign = (u128)1 << (8 * kb - 1);
    if (desc) x = ~x & ones;
    if (kind == KIND_FLOAT) {
        x = (x & sign) ? (x ^ sign) : (~x & ones);
    } else if (kind == KIND_SIGNED) {
        x ^= sign;
    }
    return x;
}

template <typename W>
static void stable_permutation(const uint8_t* keys, uint64_t n, int kind, int kb, int desc, std::vector<uint64_t>& idx) {
    std::vector<W> tw(n);
    for (uint64_t i = 0; i < n; ++i) tw[i] = (W)twiddle_in(load_word(keys + i * kb, kb), kind, kb, desc);
    idx.resize(n);
    std::iota(idx.begin(), idx.end(), (uint64_t)0);
    std::stable_sort(idx.begin(), idx.end(), [&](uint64_t a, uint64_t b) { return tw[a] < tw[b]; });
}

static void apply_permutation(uint8_t* data, int width, const std::vector<uint64_t>& idx) {
    if (!data || width <= 0) return;
    const uint64_t n = idx.size();
    std::vector<uint8_t> src(data, data + n * (uint64_t)width);
    for (uint64_t i = 0; i < n; ++i) memcpy(data + i * width, src.data() + idx[i] * width, (size_t)width);
}

static bool valid_desc(int kind, int kb) {
    if (kind < 0 || kind > 2) return false;
    if (!(kb == 1 || kb == 2 || kb == 4 || kb == 8 || kb == 16)) return false;
    if (kind == KIND_FLOAT && kb == 1) return false;
    return true;
}

extern "C" {

int oracle_version(void) { return 1; }

// Twiddle n keys of kb bytes into their unsigned radix words (kb bytes each).
int oracle_twiddle_in(int kind, int kb, int desc, const void* keys, void* out, uint64_t n) {
    if (!valid_desc(kind, kb)) return -1;
    const uint8_t* s = (const uint8_t*)keys;
    uint8_t* d = (uint8_t*)out;
    for (uint64_t i = 0; i < n; ++i) store_word(d + i * kb, twiddle_in(load_word(s + i * kb, kb), kind, kb, desc), kb);
    return 0;
}

int oracle_twiddle_out(int kind, int kb, int desc, const void* words, void* out, uint64_t n) {
    if (!valid_desc(kind, kb)) return -1;
    const uint8_t* s = (const uint8_t*)words;
    uint8_t* d = (uint8_t*)out;
    for (uint64_t i = 0; i < n; ++i) store_word(d + i * kb, twiddle_out(load_word(s + i * kb, kb), kind, kb, desc), kb);
    return 0;
}

// The result rsort::sort / sort_pairs (and _descending) must produce: keys and
// values permuted in place by std::stable_sort on the twiddled key.
int oracle_sort(int kind, int kb, int vb, int desc, void* keys, void* vals, uint64_t n) {
    if (!valid_desc(kind, kb) || vb < 0) return -1;
    std::vector<uint64_t> idx;
    if (kb <= 8) {
        stable_permutation<uint64_t>((const uint8_t*)keys, n, kind, kb, desc, idx);
    } else {
        stable_permutation<u128>((const uint8_t*)keys, n, kind, kb, desc, idx);
    }
    apply_permutation((uint8_t*)keys, kb, idx);
    apply_permutation((uint8_t*)vals, vb, idx);
    return 0;
}

// Segmented sort (no reference counterpart): every segment, clamped to
// [min(off[i], n), min(max(off[i+1], begin), n)), sorted as oracle_sort sorts
// a whole array; keys outside every segment stay where they are.
int oracle_segmented_sort(int kind, int kb, int vb, int desc, void* keys, void* vals, uint64_t n,
                          const uint64_t* offsets, uint64_t num_segments) {
    if (!valid_desc(kind, kb) || vb < 0) return -1;
    for (uint64_t s = 0; s < num_segments; ++s) {
        uint64_t b = offsets[s] < n ? offsets[s] : n;
        uint64_t e = offsets[s + 1] < b ? b : (offsets[s + 1] < n ? offsets[s + 1] : n);
        if (e - b < 2) continue;
        const int rc = oracle_sort(kind, kb, vb, desc, (uint8_t*)keys + b * kb, vals ? (uint8_t*)vals + b * vb : nullptr,
                                   e - b);
        if (rc != 0) return rc;
    }
    return 0;
}

// Stable permutation only: out[i] = original index of the i-th sorted key.
int oracle_sort_permutation(int kind, int kb, int desc, const void* keys, uint64_t n, uint64_t* out) {
    if (!valid_desc(kind, kb)) return -1;
    std::vector<uint64_t> idx;
    if (kb <= 8) {
        stable_permutation<uint64_t>((const uint8_t*)keys, n, kind, kb, desc, idx);
    } else {
        stable_permutation<u128>((const uint8_t*)keys, n, kind, kb, desc, idx);
    }
    memcpy(out, idx.data(), n * sizeof(uint64_t));
    return 0;
}

// One LSD pass (radix_kernel.cuh:118-294): stable counting sort of twiddled
// words by byte `pass`, carrying vb-byte values.  Independent of std::stable_sort.
int oracle_lsd_pass(int kb, int vb, int pass, const void* words_in, const void* vals_in, void* words_out,
                    void* vals_out, uint64_t n) {
    if (pass < 0 || pass >= kb) return -1;
    const uint8_t* ki = (const uint8_t*)words_in;
    const uint8_t* vi = (const uint8_t*)vals_in;
    uint8_t* ko = (uint8_t*)words_out;
    uint8_t* vo = (uint8_t*)vals_out;
    uint64_t count[256] = {0};
    for (uint64_t i = 0; i < n; ++i) count[ki[i * kb + pass]]++;
    uint64_t start[256];
    uint64_t run = 0;
    for (int d = 0; d < 256; ++d) {
        start[d] = run;
        run += count[d];
    }
    for (uint64_t i = 0; i < n; ++i) {
        const uint64_t o = start[ki[i * kb + pass]]++;
        memcpy(ko + o * kb, ki + i * kb, (size_t)kb);
        if (vb > 0) memcpy(vo + o * vb, vi + i * vb, (size_t)vb);
    }
    return 0;
}

// Digit counts of every pass on twiddled keys (histogram.cuh:88-211 counts the
// same digits; the reference stores their exclusive scan): out[p*256 + d].
int oracle_digit_counts(int kind, int kb, int desc, const void* keys, uint64_t n, uint64_t* out) {
    if (!valid_desc(kind, kb)) return -1;
    memset(out, 0, sizeof(uint64_t) * 256 * (size_t)kb);
    const uint8_t* s = (const uint8_t*)keys;
    for (uint64_t i = 0; i < n; ++i) {
        const u128 t = twiddle_in(load_word(s + i * kb, kb), kind, kb, desc);
        for (int p = 0; p < kb; ++p) out[p * 256 + (int)((t >> (8 * p)) & 0xFF)]++;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Reference input generator (utils.cuh:227-317, 435-437)
// ---------------------------------------------------------------------------
uint32_t oracle_mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

int oracle_set_seed_radix(int index) { return (index + 81) * 17; }

// mode: 1 random, 2 byte_skip, 3 asc (Array_Modes, utils.cuh:208-214).
// Writes the key bit patterns (kb bytes each).
int oracle_init_keys(int kb, int mode, uint32_t seed, void* out, uint64_t n) {
    if (!(kb == 1 || kb == 2 || kb == 4 || kb == 8 || kb == 16)) return -1;
    if (mode < 1 || mode > 3) return -1;
    uint8_t* d = (uint8_t*)out;
    const u128 ones = ones_of(kb);
    for (uint64_t i = 0; i < n; ++i) {
        u128 x;
        if (mode != 3) {
            if (kb <= 4) {
                x = oracle_mix32((uint32_t)(i ^ (uint64_t)seed));
            } else if (kb == 8) {
                x = ((u128)(oracle_mix32((uint32_t)(i + 0x00000000u)) ^ seed) << 32) |
                    (u128)(oracle_mix32((uint32_t)(i + 0x9e3779b9u)) ^ seed);
            } else {
                x = ((u128)(oracle_mix32((uint32_t)(i + 0x00000000u)) ^ seed) << 96) |
                    ((u128)(oracle_mix32((uint32_t)(i + 0x9e3779b9u)) ^ seed) << 64) |
                    ((u128)(oracle_mix32((uint32_t)(i + 0x3c6ef372u)) ^ seed) << 32) |
                    (u128)(oracle_mix32((uint32_t)(i + 0xdaa66d2bu)) ^ seed);
            }
            x &= ones;
            if (mode == 2) {
                const u128 keep = ((u128)0xFF00FF00FF00FF00ull << 64) | (u128)0xFF00FF00FF00FF00ull;
                const u128 fill = ((u128)0x0055005500550055ull << 64) | (u128)0x0055005500550055ull;
                x = ((x & keep) | fill) & ones;
            }
        } else {
            x = (u128)i & ones;
        }
        store_word(d + i * kb, x, kb);
    }
    return 0;
}

// check_sorted (utils.cuh:341-365): adjacent pair order on the key values;
// comparisons with NaN are false so NaNs never fail; 16-byte float keys
// (long double) are compared on their twiddled word.  Returns 1 when sorted.
int oracle_check_sorted(int kind, int kb, int desc, const void* keys, uint64_t n) {
    if (!valid_desc(kind, kb)) return -1;
    const uint8_t* s = (const uint8_t*)keys;
    for (uint64_t i = 0; i + 1 < n; ++i) {
        const u128 a = load_word(s + i * kb, kb), b = load_word(s + (i + 1) * kb, kb);
        bool bad = false;
        if (kind == KIND_FLOAT && kb == 4) {
            float fa, fb;
            uint32_t ua = (uint32_t)a, ub = (uint32_t)b;
            memcpy(&fa, &ua, 4);
            memcpy(&fb, &ub, 4);
            bad = desc ? (fa < fb) : (fa > fb);
        } else if (kind == KIND_FLOAT && kb == 8) {
            double fa, fb;
            uint64_t ua = (uint64_t)a, ub = (uint64_t)b;
            memcpy(&fa, &ua, 8);
            memcpy(&fb, &ub, 8);
            bad = desc ? (fa < fb) : (fa > fb);
        } else {
            // integers (and half / long double): order of the twiddled word
            const u128 ta = twiddle_in(a, kind, kb, 0), tb = twiddle_in(b, kind, kb, 0);
            bad = desc ? (ta < tb) : (ta > tb);
        }
        if (bad) return 0;
    }
    return 1;
}

// check_pairings (utils.cuh:377-432) for values = original index (vb = 4 or 8):
// value in range, key bit-identical to the original key at that index, and
// equal keys carry increasing values (stability).  Returns 1 when all hold.
int oracle_check_pairings(int kb, int vb, const void* keys_orig, const void* keys_sorted, const void* vals_sorted,
                          uint64_t n) {
    if (!(vb == 4 || vb == 8)) return -1;
    const uint8_t* ko = (const uint8_t*)keys_orig;
    const uint8_t* ks = (const uint8_t*)keys_sorted;
    const uint8_t* vs = (const uint8_t*)vals_sorted;
    auto val = [&](uint64_t i) -> uint64_t {
        uint64_t v = 0;
        memcpy(&v, vs + i * vb, (size_t)vb);
        return v;
    };
    for (uint64_t i = 0; i < n; ++i) {
        const uint64_t v = val(i);
        if (v >= n) return 0;
        if (memcmp(ks + i * kb, ko + v * kb, (size_t)kb) != 0) return 0;
        if (i + 1 < n && memcmp(ks + i * kb, ks + (i + 1) * kb, (size_t)kb) == 0 && v > val(i + 1)) return 0;
    }
    return 1;
}

// ---------------------------------------------------------------------------
// Workload generators of the benchmark configurations (raw std::mt19937).
// ---------------------------------------------------------------------------
void oracle_mt19937(uint32_t seed, uint32_t* out, uint64_t n) {
    std::mt19937 g(seed);
    for (uint64_t i = 0; i < n; ++i) out[i] = (uint32_t)g();
}

static inline double u01_53(std::mt19937& g) {  // [0,1) with 53 bits (genrand_res53)
    const uint32_t a = (uint32_t)g() >> 5, b = (uint32_t)g() >> 6;
    return (a * 67108864.0 + b) * (1.0 / 9007199254740992.0);
}

static inline double normal_bm(std::mt19937& g) {  // Box-Muller, cosine branch
    const double u1 = 1.0 - u01_53(g);            // (0, 1]
    const double u2 = u01_53(g);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586476925 * u2);
}

// 64-bit bijective finaliser (splitmix64 output function).
static inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

uint64_t oracle_mix64(uint64_t z) { return mix64(z); }

// uint64 keys: low 16 bits uniform, upper 48 bits zero.
void oracle_gen_u64_low16(uint32_t seed, uint64_t* out, uint64_t n) {
    std::mt19937 g(seed);
    for (uint64_t i = 0; i < n; ++i) out[i] = (uint64_t)((uint32_t)g() & 0xFFFFu);
}

// uint64 keys: rank ~ Zipf(s) over `ranks` values (inverse CDF), key = mix64(rank).
void oracle_gen_zipf_u64(uint32_t seed, double s, uint64_t ranks, uint64_t* out, uint64_t n) {
    std::vector<double> cdf(ranks);
    double acc = 0.0;
    for (uint64_t k = 0; k < ranks; ++k) {
        acc += std::exp(-s * std::log((double)(k + 1)));
        cdf[k] = acc;
    }
    for (auto& c : cdf) c /= acc;
    std::mt19937 g(seed);
    for (uint64_t i = 0; i < n; ++i) {
        const double u = u01_53(g);
        uint64_t r = (uint64_t)(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
        if (r >= ranks) r = ranks - 1;
        out[i] = mix64(r + 1);
    }
}

// float32 ~ N(0,1), 1% replaced by +-0.0 (half each), 0.1% by +-inf.
void oracle_gen_f32_normal_specials(uint32_t seed, float* out, uint64_t n) {
    std::mt19937 g(seed);
    for (uint64_t i = 0; i < n; ++i) {
        const uint32_t r = (uint32_t)g();
        const double u = r * (1.0 / 4294967296.0);
        if (u < 0.001) {
            out[i] = (r & 1u) ? -INFINITY : INFINITY;
        } else if (u < 0.011) {
            out[i] = (r & 1u) ? -0.0f : 0.0f;
        } else {
            out[i] = (float)normal_bm(g);
        }
    }
}

// 128-bit keys {lo, hi}: hi uniform over 2^20 values, lo uniform over 2^64.
void oracle_gen_u128_pairs(uint32_t seed, uint64_t* out_lohi, uint64_t n) {
    std::mt19937 g(seed);
    for (uint64_t i = 0; i < n; ++i) {
        const uint64_t hi = (uint64_t)((uint32_t)g() & 0xFFFFFu);
        const uint64_t a = (uint32_t)g(), b = (uint32_t)g();
        out_lohi[2 * i] = (a << 32) | b;
        out_lohi[2 * i + 1] = hi;
    }
}

// Segment sizes ~ round(median * exp(sigma * N(0,1))) clipped to [lo, hi].
void oracle_gen_lognormal_sizes(uint32_t seed, double median, double sigma, uint64_t lo, uint64_t hi,
                                uint64_t* out, uint64_t count) {
    std::mt19937 g(seed);
    for (uint64_t i = 0; i < count; ++i) {
        double v = std::floor(median * std::exp(sigma * normal_bm(g)) + 0.5);
        if (v < (double)lo) v = (double)lo;
        if (v > (double)hi) v = (double)hi;
        out[i] = (uint64_t)v;
    }
}

// ---------------------------------------------------------------------------
// CPU baselines (timed on the host cores, outside generation and copies).
// ---------------------------------------------------------------------------
static double median_of(std::vector<double> v) {
    std::sort(v.begin(), v.end());
    return v.empty() ? 0.0 : v[v.size() / 2];
}

// std::stable_sort of (key, value) records by key, single thread.
// vals == NULL times key-only.  Returns the median seconds over reps.
double oracle_time_stable_sort_u32(const uint32_t* keys, const uint32_t* vals, uint64_t n, int reps) {
    std::vector<double> t;
    for (int r = 0; r < reps; ++r) {
        if (vals) {
            std::vector<std::pair<uint32_t, uint32_t>> rec(n);
            for (uint64_t i = 0; i < n; ++i) rec[i] = {keys[i], vals[i]};
            const auto t0 = std::chrono::steady_clock::now();
            std::stable_sort(rec.begin(), rec.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
            t.push_back(std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
        } else {
            std::vector<uint32_t> k(keys, keys + n);
            const auto t0 = std::chrono::steady_clock::now();
            std::stable_sort(k.begin(), k.end());
            t.push_back(std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
        }
    }
    return median_of(t);
}

// std::sort(std::execution::par_unseq) of the same records (not stable).
double oracle_time_sort_par_u32(const uint32_t* keys, const uint32_t* vals, uint64_t n, int reps) {
    std::vector<double> t;
    for (int r = 0; r < reps; ++r) {
        if (vals) {
            std::vector<std::pair<uint32_t, uint32_t>> rec(n);
            for (uint64_t i = 0; i < n; ++i) rec[i] = {keys[i], vals[i]};
            const auto t0 = std::chrono::steady_clock::now();
            std::sort(std::execution::par_unseq, rec.begin(), rec.end(),
                      [](const auto& a, const auto& b) { return a.first < b.first; });
            t.push_back(std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
        } else {
            std::vector<uint32_t> k(keys, keys + n);
            const auto t0 = std::chrono::steady_clock::now();
            std::sort(std::execution::par_unseq, k.begin(), k.end());
            t.push_back(std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
        }
    }
    return median_of(t);
}

// Stable sort over `threads` host threads: each thread std::stable_sorts one
// of `threads` contiguous runs, then runs are merged pairwise in rounds, the
// merges of a round in parallel (std::merge takes from the left run on equal
// keys, so the result is the same as one std::stable_sort).
extern "C++" {
template <typename R, typename Less>
static void par_stable_sort(std::vector<R>& a, int threads, Less less) {
    const size_t n = a.size();
    const size_t T = threads < 1 ? 1 : (size_t)threads;
    std::vector<size_t> b(T + 1);
    for (size_t i = 0; i <= T; ++i) b[i] = n * i / T;
    {
        std::vector<std::thread> th;
        for (size_t i = 0; i < T; ++i)
            th.emplace_back([&, i] { std::stable_sort(a.begin() + b[i], a.begin() + b[i + 1], less); });
        for (auto& t : th) t.join();
    }
    std::vector<R> tmp(n);
    std::vector<R>* src = &a;
    std::vector<R>* dst = &tmp;
    for (size_t width = 1; width < T; width *= 2) {
        std::vector<std::thread> th;
        for (size_t i = 0; i < T; i += 2 * width) {
            const size_t lo = b[i], mid = b[std::min(i + width, T)], hi = b[std::min(i + 2 * width, T)];
            th.emplace_back([=] {
                std::merge(src->begin() + lo, src->begin() + mid, src->begin() + mid, src->begin() + hi,
                           dst->begin() + lo, less);
            });
        }
        for (auto& t : th) t.join();
        std::swap(src, dst);
    }
    if (src != &a) a.swap(tmp);
}
}  // extern "C++"

// The multi-threaded stable sort of (key, value) records by key (vals NULL:
// keys only); results to out_keys / out_vals.
int oracle_stable_sort_par_u32(const uint32_t* keys, const uint32_t* vals, uint64_t n, int threads,
                               uint32_t* out_keys, uint32_t* out_vals) {
    if (vals) {
        std::vector<std::pair<uint32_t, uint32_t>> rec(n);
        for (uint64_t i = 0; i < n; ++i) rec[i] = {keys[i], vals[i]};
        par_stable_sort(rec, threads, [](const auto& x, const auto& y) { return x.first < y.first; });
        for (uint64_t i = 0; i < n; ++i) {
            out_keys[i] = rec[i].first;
            out_vals[i] = rec[i].second;
        }
    } else {
        std::vector<uint32_t> k(keys, keys + n);
        par_stable_sort(k, threads, [](uint32_t x, uint32_t y) { return x < y; });
        std::copy(k.begin(), k.end(), out_keys);
    }
    return 0;
}

// Median seconds of the multi-threaded stable sort over reps.
double oracle_time_stable_sort_par_u32(const uint32_t* keys, const uint32_t* vals, uint64_t n, int threads,
                                       int reps) {
    std::vector<double> t;
    for (int r = 0; r < reps; ++r) {
        if (vals) {
            std::vector<std::pair<uint32_t, uint32_t>> rec(n);
            for (uint64_t i = 0; i < n; ++i) rec[i] = {keys[i], vals[i]};
            const auto t0 = std::chrono::steady_clock::now();
            par_stable_sort(rec, threads, [](const auto& x, const auto& y) { return x.first < y.first; });
            t.push_back(std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
        } else {
            std::vector<uint32_t> k(keys, keys + n);
            const auto t0 = std::chrono::steady_clock::now();
            par_stable_sort(k, threads, [](uint32_t x, uint32_t y) { return x < y; });
            t.push_back(std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
        }
    }
    return median_of(t);
}

// 1 when libstdc++'s parallel algorithms have the TBB backend, 0 when they
// run serially.
int oracle_pstl_parallel_backend(void) {
#if defined(_PSTL_PAR_BACKEND_TBB) || __has_include(<tbb/tbb.h>)
    return 1;
#else
    return 0;
#endif
}

}  // extern "C"
