// dy_internal.hpp — helpers shared by the HIP translation units of
// libdyphur_hip.so (not part of the public ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdlib>
#include <stdint.h>

#include "../../include/dyphur_hip.h"
#include "dy_math.hpp"

// The one-XCD handoffs below (xcd_barrier, the dataflow sweep's body records
// and versions) rely on gfx950's cache hierarchy: the vector L1 writes through
// to the XCD's L2, a store is in the L2 once s_waitcnt vmcnt(0) has retired
// it, and agent-scope (sc1) loads are served by the L2, so workgroups of one
// XCD see each other's completed stores without cache maintenance.  The HIP
// memory model does not promise this for plain stores, so the device code is
// built for gfx950 only.
#if defined(__HIP_DEVICE_COMPILE__) && !defined(__gfx950__)
#error "dyphur_amd's device code relies on gfx950's L1/L2 behaviour; build with --offload-arch=gfx950"
#endif

#define DY_TRY(expr)                                          \
    do {                                                      \
        hipError_t dy_e_ = (expr);                            \
        if (dy_e_ != hipSuccess) return static_cast<int>(dy_e_); \
    } while (0)

#define DY_TRY_STATUS(expr)              \
    do {                                 \
        int dy_s_ = (expr);              \
        if (dy_s_ != DY_OK) return dy_s_; \
    } while (0)

#define DY_LAUNCH_CHECK() DY_TRY(hipGetLastError())

// Launch of a kernel whose workgroups wait for each other (grid barriers,
// XCD check-ins, dataflow hand-offs).  Every caller sizes the grid to what the
// device holds at once (occupancy x CUs), so on a device that runs only this
// stream's work an ordinary launch makes the whole grid resident: the
// stream's previous kernel has finished and the dispatcher places every
// workgroup.  That is the default, because the runtime's cooperative path
// leaves the device idle for about 10 us before and after each such kernel.
// The guarantee itself needs the cooperative path: with DYPHUR_COOP_LAUNCH=1
// (callers that run other work on the device at the same time) the launch is
// cooperative, and may be refused: another process that has made cooperative
// launches on the device holds them (hipErrorIllegalState), or the grid does
// not fit.  Nothing is enqueued then; the runtime's last error is cleared so
// the refusal does not resurface at the next launch check, and the caller
// takes its non-cooperative path.  A grid kept from residency by other work
// for longer than the kernels' spin limits ends with the kernel's error flag
// set, not in a hang.
inline bool coop_launch_required() {
    const char *e = std::getenv("DYPHUR_COOP_LAUNCH");
    return e && std::atoi(e) != 0;
}
inline hipError_t try_coop_launch(const void *fn, dim3 grid, dim3 block, void **args, hipStream_t st) {
    const hipError_t e = coop_launch_required() ? hipLaunchCooperativeKernel(fn, grid, block, args, 0, st)
                                                : hipLaunchKernel(fn, grid, block, args, 0, st);
    if (e != hipSuccess) (void)hipGetLastError();
    return e;
}

namespace dy {

inline hipStream_t hs(dy_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

inline uint32_t ceil_div(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

// Device buffer that grows on demand (contents are not preserved).
template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t cap = 0;
    int ensure(size_t n) {
        if (n <= cap) return DY_OK;
        if (p) DY_TRY(hipFree(p));
        p = nullptr;
        cap = 0;
        size_t c = n < 64 ? 64 : n;
        DY_TRY(hipMalloc(reinterpret_cast<void **>(&p), c * sizeof(T)));
        cap = c;
        return DY_OK;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

// Primitives implemented in primitives.hip, used by the stage translation units.
// Exclusive scan of n u32 values; d_total receives the sum (may be null).
// Zeroes n words with one small kernel (a hipMemsetAsync of a few words can
// take two fill dispatches, and cannot be recorded into a graph).
int clear_words(hipStream_t s, uint32_t *p, uint32_t n);

// `tmp` must hold at least scan_tmp_words(n) u32.
size_t scan_tmp_words(size_t n);
int scan_exclusive(hipStream_t s, const uint32_t *in, uint32_t *out, size_t n, uint32_t *d_total,
                   uint32_t *tmp);

// Stable LSD radix sort in one workgroup of (key, value) pairs over key bits
// [begin_bit, end_bit).  Keys/values in (k0, v0); (k1, v1) are scratch of the
// same size.  The result ends in (k0, v0).  vals may be null for key-only sorts.
int radix_sort_u32(hipStream_t s, uint32_t *k0, uint32_t *v0, uint32_t *k1, uint32_t *v1, size_t n,
                   int begin_bit, int end_bit);
int radix_sort_u64(hipStream_t s, uint64_t *k0, uint32_t *v0, uint64_t *k1, uint32_t *v1, size_t n,
                   int begin_bit, int end_bit);

// Block-wide helpers ----------------------------------------------------------
__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Inclusive scan across the 64 lanes of a wavefront.
__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v) {
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t t = __shfl_up(v, d, 64);
        if (static_cast<int>(lane_id()) >= d) v += t;
    }
    return v;
}

// Exclusive scan across a workgroup of `nthreads` (multiple of 64, <= 1024).
// `wsum` is LDS scratch of at least 16 words.  Returns the exclusive prefix;
// `total` receives the block sum.  Contains __syncthreads().
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t *wsum, uint32_t &total) {
    const uint32_t lane = lane_id();
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t nw = blockDim.x >> 6;
    const uint32_t inc = wave_inclusive_scan(v);
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    if (wave == 0) {
        uint32_t w = lane < nw ? wsum[lane] : 0u;
        uint32_t wi = wave_inclusive_scan(w);
        if (lane < nw) wsum[lane] = wi - w;
        if (lane == nw - 1) wsum[16] = wi;
    }
    __syncthreads();
    total = wsum[16];
    const uint32_t r = wsum[wave] + inc - v;
    __syncthreads();
    return r;
}

// Grid-wide barrier for a cooperative launch (every workgroup of the grid
// co-resident): arrival counter bar[0] + generation bar[1], both zero before
// the launch.  Thread 0 of each workgroup publishes the workgroup's stores
// with an agent-scope release before arriving and acquires after the
// generation moves.  A wait that exceeds kGridSpinLimit polls sets err_bit in
// *err and gives up, so a broken launch ends instead of hanging.
constexpr uint32_t kGridSpinLimit = 1u << 22;
__device__ __forceinline__ void grid_barrier(uint32_t *bar, uint32_t &gen, uint32_t *err, uint32_t err_bit) {
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t g = gen;
        const uint32_t arrived = __hip_atomic_fetch_add(&bar[0], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (arrived == gridDim.x - 1) {
            __hip_atomic_store(&bar[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&bar[1], g + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            uint32_t spins = 0;
            while (__hip_atomic_load(&bar[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == g) {
                __builtin_amdgcn_s_sleep(1);
                if (++spins > kGridSpinLimit) {
                    atomicOr(err, err_bit);
                    break;
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
    }
    gen += 1u;
    __syncthreads();
}

// ---------------------------------------------------------------------------
// Work confined to one XCD.  The workgroups of one XCD share its L2: the
// vector L1 writes through to it and atomics execute in it, so they can hand
// data to each other through the L2 with no cache maintenance, provided every
// read of data written during the kernel skips the L1 (agent-scope loads) and
// every writer waits for its stores before arriving at a barrier.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t xcc_id() {
    return static_cast<uint32_t>(__builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20)) & 0xfu;  // HW_REG_XCC_ID
}

__device__ __forceinline__ void wait_vmem() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// Current value at the L2 (a returning atomic add of zero, which the compiler
// would otherwise turn into a load that may hit the L1).
__device__ __forceinline__ uint32_t l2_read(uint32_t *p) {
    uint32_t v;
    const uint32_t zero = 0u;
    asm volatile("global_atomic_add %0, %1, %2, off sc0\n\ts_waitcnt vmcnt(0)"
                 : "=v"(v) : "v"(p), "v"(zero) : "memory");
    return v;
}

template <typename T>
__device__ __forceinline__ T ld_agent(const T *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Barrier among the nblk workgroups of one XCD (bar[0] arrivals, bar[1]
// generation; zero before the launch).  With `fence`, thread 0 also acquires
// at agent scope after the wait, so that plain loads after the barrier see
// data written before it by the other workgroups.  False on a timeout (err
// gets err_bit).
__device__ __forceinline__ bool xcd_barrier(uint32_t *bar, uint32_t &gen, uint32_t nblk, uint32_t *err,
                                            uint32_t err_bit, bool fence = false) {
    __shared__ uint32_t s_ok;
    wait_vmem();
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t g = gen;
        uint32_t ok = 1;
        const uint32_t arrived = __hip_atomic_fetch_add(&bar[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (arrived == nblk - 1) {
            __hip_atomic_store(&bar[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            wait_vmem();
            __hip_atomic_store(&bar[1], g + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            wait_vmem();
        } else {
            uint32_t spins = 0;
            while (l2_read(&bar[1]) == g) {
                __builtin_amdgcn_s_sleep(1);
                if (++spins > kGridSpinLimit) {
                    atomicOr(err, err_bit);
                    ok = 0;
                    break;
                }
            }
        }
        if (fence) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        s_ok = ok;
    }
    gen += 1u;
    __syncthreads();
    return s_ok != 0;
}

// Check-in of a launch that works on XCD `xcc` only: every workgroup arrives
// on reg[1] (agent scope); the ones on that XCD take dense indices from
// reg[0] and wait until all have arrived.  Returns true for a participant,
// with its index and the participant count.  The check-in publishes no data
// (participants read what earlier kernels wrote, visible at the kernel
// boundary, and hand data to each other through the XCD's L2), so its atomics
// and polls are relaxed: no L2 write-back or invalidation.
__device__ __forceinline__ bool xcd_checkin(uint32_t xcc, uint32_t *reg, uint32_t *err, uint32_t err_bit,
                                            uint32_t &idx, uint32_t &n) {
    __shared__ uint32_t s_idx, s_n;
    const bool mine = xcc_id() == xcc;
    if (threadIdx.x == 0) {
        s_idx = mine ? __hip_atomic_fetch_add(&reg[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0u;
        __hip_atomic_fetch_add(&reg[1], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_n = 0;
        if (mine) {
            uint32_t spins = 0;
            while (__hip_atomic_load(&reg[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != gridDim.x) {
                __builtin_amdgcn_s_sleep(2);
                if (++spins > kGridSpinLimit) {
                    atomicOr(err, err_bit);
                    break;
                }
            }
            s_n = __hip_atomic_load(&reg[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    __syncthreads();
    idx = s_idx;
    n = s_n;
    return mine && n > 0;
}


}  // namespace dy
