#pragma once
// parallel_for (compute/include/compute/kernel.hpp:11-15): runs kernel(i) for
// i in [0, n) on the stream; the callable must be device-safe (a plain lambda
// capturing device pointers by value).  Needs a HIP compiler (hipcc) for the
// calling translation unit.
#include "stream.hpp"

#include <hip/hip_runtime.h>

namespace dyphur {
namespace detail {
template <typename F>
__global__ void parallel_for_kernel(size_t n, F f) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) f(i);
}
}  // namespace detail

template <typename F>
Event parallel_for(Stream &s, size_t n, F kernel) {
    if (n > 0) {
        const size_t blocks = (n + 255) / 256;
        detail::parallel_for_kernel<<<static_cast<unsigned>(blocks), 256, 0,
                                      reinterpret_cast<hipStream_t>(s.handle())>>>(n, kernel);
        check(static_cast<int>(hipGetLastError()), "parallel_for: launch");
    }
    return Event::record(s.handle());
}

}  // namespace dyphur
