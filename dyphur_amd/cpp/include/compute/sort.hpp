#pragma once
// sort_by_key / sort (compute/include/compute/sort.hpp:29,89): ascending,
// stable LSD radix sort on the device; asynchronous.  The reference requires
// n to be a power of two (bitonic network); any n is accepted here.  Key and
// value types: u64 or u32 keys with u32 values; u32 or u64 key-only.
#include "stream.hpp"

#include <cstdint>
#include <type_traits>

namespace dyphur {

template <typename Key, typename Value>
void sort_by_key(Stream &s, Key *keys, Value *values, size_t n) {
    static_assert(std::is_same_v<Value, uint32_t>, "sort_by_key: values must be uint32_t");
    if (n <= 1) return;
    if constexpr (std::is_same_v<Key, uint64_t>)
        check(dy_sort_by_key_u64_u32(s.handle(), keys, values, n), "sort_by_key");
    else if constexpr (std::is_same_v<Key, uint32_t>)
        check(dy_sort_by_key_u32_u32(s.handle(), keys, values, n), "sort_by_key");
    else
        static_assert(sizeof(Key) == 0, "sort_by_key: keys must be uint32_t or uint64_t");
}

template <typename Key>
void sort(Stream &s, Key *keys, size_t n) {
    if (n <= 1) return;
    if constexpr (std::is_same_v<Key, uint64_t>)
        check(dy_sort_u64(s.handle(), keys, n), "sort");
    else if constexpr (std::is_same_v<Key, uint32_t>)
        check(dy_sort_u32(s.handle(), keys, n), "sort");
    else
        static_assert(sizeof(Key) == 0, "sort: keys must be uint32_t or uint64_t");
}

}  // namespace dyphur
