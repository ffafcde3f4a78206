// rsort.hpp - drop-in C++ API of the MI355X onesweep radix sort.
//
// Same four templates, argument order and meaning as the reference
// (AwardOfSky/CUDA-Fast-Sort include/rsort/rsort.cuh:13-92), with hipStream_t
// in place of cudaStream_t.  Each forwards to the C ABI in <rsort_c.h>:
//
//   size_t temp_bytes = 0;
//   rsort::sort(d_keys, &temp_bytes, nullptr, n);      // size query, no device work
//   hipMalloc(&d_workspace, temp_bytes);
//   rsort::sort(d_keys, &temp_bytes, d_workspace, n);  // in-place, stream-ordered
//
// Differences from the reference, all in the caller's favour:
//   * the sort runs on `stream` (the reference ignores it and captures a graph
//     on the legacy stream, radix_kernel.cuh:583-600);
//   * errors are recorded instead of ignored: rsort::last_status() and
//     rsort::last_error() report the outcome of the calling thread's last call;
//   * n == 0 is a no-op (undefined in the reference).
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>
#include <type_traits>

#include "../rsort_c.h"

namespace rsort {

namespace detail {

// Placeholder value type of key-only sorts (the reference's no_value_t).
struct no_value_t {};

// IEEE half keys (_Float16), where the compiler has the type in C++ (clang,
// hipcc's host compiler; gcc 12 and later on x86-64).  The reference takes
// arithmetic types only (tuning.cuh:301-304); half keys are an addition, also
// reachable through the C ABI (RSORT_KEY_FLOAT, 2 bytes) with any compiler.
template <typename T>
struct is_float16 : std::false_type {};
#if defined(__FLT16_MAX__)
template <>
struct is_float16<_Float16> : std::true_type {};
#endif

// Key kind / width of Key_T (radix_traits, include/rsort/detail/tuning.cuh:297-397).
template <typename T>
struct key_desc {
    static constexpr bool is_half = is_float16<T>::value;
    static constexpr bool is_ld = std::is_same_v<T, long double> && sizeof(long double) == 16;
#if defined(__SIZEOF_INT128__)
    static constexpr bool is_u128 = std::is_same_v<T, unsigned __int128>;
    static constexpr bool is_i128 = std::is_same_v<T, __int128>;
#else
    static constexpr bool is_u128 = false;
    static constexpr bool is_i128 = false;
#endif
    static_assert(std::is_arithmetic_v<T> || is_u128 || is_i128 || is_half, "Radix type must be a numeric.");
    static_assert(sizeof(T) <= 16, "Radix type too large.");
    static_assert(!std::is_floating_point_v<T> || sizeof(T) != 1, "no 1-byte floating keys");
    static constexpr uint32_t bytes = (uint32_t)sizeof(T);
    static constexpr uint32_t kind = (std::is_floating_point_v<T> || is_ld || is_half) ? RSORT_KEY_FLOAT
                                     : ((std::is_signed_v<T> || is_i128) ? RSORT_KEY_SIGNED : RSORT_KEY_UNSIGNED);
};

inline rsort_status& last_status_ref() {
    static thread_local rsort_status s = RSORT_SUCCESS;
    return s;
}

template <bool Descending, typename Key_T, typename Value_T>
inline void call(Key_T* keys, Value_T* values, size_t* temp_bytes, uint8_t* workspace, size_t n, hipStream_t stream) {
    static_assert(sizeof(long double) == 16 || !std::is_same_v<Key_T, long double> || sizeof(long double) == 8,
                  "Sorting long doubles expects 64 or 128 bits.");
    rsort_desc d{};
    d.key_kind = key_desc<Key_T>::kind;
    d.key_bytes = key_desc<Key_T>::bytes;
    d.value_bytes = std::is_same_v<Value_T, no_value_t> ? 0u : (uint32_t)sizeof(Value_T);
    d.descending = Descending ? 1u : 0u;
    last_status_ref() = rsort_sort(&d, static_cast<void*>(keys),
                                   std::is_same_v<Value_T, no_value_t> ? nullptr : static_cast<void*>(values),
                                   static_cast<void*>(workspace), temp_bytes, (uint64_t)n,
                                   static_cast<void*>(stream));
}

// Byte offset of a data member, from its address in a record-sized buffer
// (no Record object is constructed).
template <typename Record, typename Key_T>
inline uint32_t member_offset(Key_T Record::*key) {
    alignas(Record) static unsigned char buf[sizeof(Record)];
    const Record* r = reinterpret_cast<const Record*>(buf);
    return (uint32_t)(reinterpret_cast<const unsigned char*>(&(r->*key)) - buf);
}

template <bool Descending, typename Record, typename Key_T>
inline void call_records(Record* records, Key_T Record::*key, size_t* temp_bytes, uint8_t* workspace, size_t n,
                         hipStream_t stream) {
    static_assert(std::is_trivially_copyable<Record>::value, "records are moved as bytes");
    rsort_desc d{};
    d.key_kind = key_desc<Key_T>::kind;
    d.key_bytes = key_desc<Key_T>::bytes;
    d.value_bytes = 0;
    d.descending = Descending ? 1u : 0u;
    last_status_ref() = rsort_sort_records(&d, static_cast<void*>(records), (uint64_t)n, (uint32_t)sizeof(Record),
                                           member_offset(key), static_cast<void*>(workspace), temp_bytes,
                                           static_cast<void*>(stream));
}

}  // namespace detail

// Outcome of the calling thread's last rsort call.
inline rsort_status last_status() { return detail::last_status_ref(); }
inline const char* last_error() { return rsort_last_error(); }

template <typename Key_T, typename Len_T>
inline void sort(Key_T* d_inout, size_t* temp_bytes, uint8_t* d_workspace, Len_T n, hipStream_t stream = nullptr) {
    static_assert(std::is_integral_v<Len_T>, "Type of N (Len_T) must be an integral type");
    detail::call<false, Key_T, detail::no_value_t>(d_inout, nullptr, temp_bytes, d_workspace, (size_t)n, stream);
}

template <typename Key_T, typename Len_T>
inline void sort_descending(Key_T* d_inout, size_t* temp_bytes, uint8_t* d_workspace, Len_T n,
                            hipStream_t stream = nullptr) {
    static_assert(std::is_integral_v<Len_T>, "Type of N (Len_T) must be an integral type");
    detail::call<true, Key_T, detail::no_value_t>(d_inout, nullptr, temp_bytes, d_workspace, (size_t)n, stream);
}

template <typename Key_T, typename Len_T, typename Value_T>
inline void sort_pairs(Key_T* d_inout, Value_T* d_inout_values, size_t* temp_bytes, uint8_t* d_workspace, Len_T n,
                       hipStream_t stream = nullptr) {
    static_assert(std::is_integral_v<Len_T>, "Type of N (Len_T) must be an integral type");
    detail::call<false, Key_T, Value_T>(d_inout, d_inout_values, temp_bytes, d_workspace, (size_t)n, stream);
}

template <typename Key_T, typename Len_T, typename Value_T>
inline void sort_pairs_descending(Key_T* d_inout, Value_T* d_inout_values, size_t* temp_bytes, uint8_t* d_workspace,
                                  Len_T n, hipStream_t stream = nullptr) {
    static_assert(std::is_integral_v<Len_T>, "Type of N (Len_T) must be an integral type");
    detail::call<true, Key_T, Value_T>(d_inout, d_inout_values, temp_bytes, d_workspace, (size_t)n, stream);
}

// Array-of-structs sort (no counterpart in the reference's API; it plans
// one, rsort.cuh:95-102): sorts d_records stably by the member `key`, e.g.
//   struct Particle { float x, y; uint32_t cell; };
//   rsort::sort_records(d_particles, &Particle::cell, &temp_bytes, d_ws, n);
// Record must be trivially copyable; the member's type must be a key type.
template <typename Record, typename Key_T, typename Len_T>
inline void sort_records(Record* d_records, Key_T Record::*key, size_t* temp_bytes, uint8_t* d_workspace, Len_T n,
                         hipStream_t stream = nullptr) {
    detail::call_records<false>(d_records, key, temp_bytes, d_workspace, (size_t)n, stream);
}

template <typename Record, typename Key_T, typename Len_T>
inline void sort_records_descending(Record* d_records, Key_T Record::*key, size_t* temp_bytes, uint8_t* d_workspace,
                                    Len_T n, hipStream_t stream = nullptr) {
    detail::call_records<true>(d_records, key, temp_bytes, d_workspace, (size_t)n, stream);
}

}  // namespace rsort
