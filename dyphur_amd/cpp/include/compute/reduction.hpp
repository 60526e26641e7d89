#pragma once
// reduce_sum (compute/include/compute/reduction.hpp:22-62): deterministic sum,
// workgroup trees of 256 in a fixed order then a sequential host merge of the
// partial sums; synchronous.  float only.
#include "stream.hpp"

#include <type_traits>

namespace dyphur {

template <typename T>
T reduce_sum(Stream &s, const T *d_data, size_t n) {
    static_assert(std::is_same_v<T, float>, "reduce_sum: float only");
    if (n == 0) return T{0};
    float out = 0.f;
    check(dy_reduce_sum_f32(s.handle(), d_data, n, &out), "reduce_sum");
    return out;
}

}  // namespace dyphur
