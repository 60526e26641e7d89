#pragma once
// atomic_add / atomic_add_seq (compute/include/compute/atomics.hpp:9-24):
// device-scope atomic add returning the old value; device code only (hipcc).
#include <hip/hip_runtime.h>

namespace dyphur {

template <typename T>
__device__ T atomic_add(T *ptr, T val) {
    return __hip_atomic_fetch_add(ptr, val, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename T>
__device__ T atomic_add_seq(T *ptr, T val) {
    return __hip_atomic_fetch_add(ptr, val, __ATOMIC_SEQ_CST, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace dyphur
