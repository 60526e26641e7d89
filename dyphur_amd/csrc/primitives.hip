// primitives.hip — device-wide scan, single-workgroup LSD radix sort and the
// deterministic reduction behind the compute shim's sort_by_key / sort /
// reduce_sum (compute/include/compute/{sort,reduction}.hpp).
#include <vector>

#include "dy_internal.hpp"

namespace dy {

// ---------------------------------------------------------------------------
// Exclusive scan
// ---------------------------------------------------------------------------
constexpr uint32_t kScanThreads = 1024;
constexpr uint32_t kScanItems = 8;
constexpr uint32_t kScanTile = kScanThreads * kScanItems;  // one workgroup pass
constexpr size_t kScanSingleMax = 1u << 16;

// One workgroup scans up to kScanSingleMax values: every tile's values are
// loaded before the first tile is scanned (one memory latency for the whole
// input), then the tiles are scanned in order with a running carry.
// `total` (may be null) receives the sum.
__global__ __launch_bounds__(kScanThreads) void scan_one_block(const uint32_t *__restrict__ in,
                                                               uint32_t *__restrict__ out, uint32_t n,
                                                               uint32_t *__restrict__ total) {
    constexpr uint32_t kMaxTiles = static_cast<uint32_t>(kScanSingleMax / kScanTile);
    __shared__ uint32_t tile[kScanTile];
    __shared__ uint32_t wsum[17];
    const uint32_t t = threadIdx.x;
    const uint32_t ntiles = (n + kScanTile - 1u) / kScanTile;
    uint32_t in_v[kMaxTiles][kScanItems];
#pragma unroll
    for (uint32_t b = 0; b < kMaxTiles; ++b)
#pragma unroll
        for (uint32_t j = 0; j < kScanItems; ++j) {
            const uint32_t idx = b * kScanTile + j * kScanThreads + t;
            in_v[b][j] = idx < n ? in[idx] : 0u;
        }
    uint32_t carry = 0;
#pragma unroll
    for (uint32_t b = 0; b < kMaxTiles; ++b) {
        if (b >= ntiles) break;
        const uint32_t base = b * kScanTile;
#pragma unroll
        for (uint32_t j = 0; j < kScanItems; ++j) tile[j * kScanThreads + t] = in_v[b][j];
        __syncthreads();
        uint32_t v[kScanItems], sum = 0;
#pragma unroll
        for (uint32_t j = 0; j < kScanItems; ++j) { v[j] = tile[t * kScanItems + j]; sum += v[j]; }
        uint32_t btotal;
        uint32_t run = carry + block_exclusive_scan(sum, wsum, btotal);
#pragma unroll
        for (uint32_t j = 0; j < kScanItems; ++j) { tile[t * kScanItems + j] = run; run += v[j]; }
        __syncthreads();
#pragma unroll
        for (uint32_t j = 0; j < kScanItems; ++j) {
            const uint32_t idx = base + j * kScanThreads + t;
            if (idx < n) out[idx] = tile[j * kScanThreads + t];
        }
        carry += btotal;
        __syncthreads();
    }
    if (t == 0 && total) *total = carry;
}

// One workgroup, values staged in LDS (one pad word per 32, so that a
// thread's contiguous run falls on distinct banks): thread t scans the run
// [t*K, t*K + K) serially, one workgroup scan of the 1 024 run sums gives each
// run's base — four workgroup barriers in all, against several per 8 192-value
// tile in scan_one_block.  Up to kScanLdsMax values.
constexpr uint32_t kScanLdsMax = kScanThreads * 36u;  // 36 864 values, 148.5 KiB with padding

__global__ __launch_bounds__(kScanThreads) void scan_lds(const uint32_t *__restrict__ in, uint32_t *__restrict__ out,
                                                         uint32_t n, uint32_t *__restrict__ total) {
    __shared__ uint32_t buf[kScanLdsMax + kScanLdsMax / 32u];
    __shared__ uint32_t wsum[17];
    const uint32_t t = threadIdx.x;
    const uint32_t k_run = (n + kScanThreads - 1u) / kScanThreads;  // values per thread (<= 36)
    const uint32_t padded = k_run * kScanThreads;
    for (uint32_t i0 = 0; i0 < padded; i0 += 8u * kScanThreads) {  // coalesced, eight loads in flight
        uint32_t v[8];
#pragma unroll
        for (uint32_t k = 0; k < 8u; ++k) {
            const uint32_t i = i0 + k * kScanThreads + t;
            v[k] = i < n ? in[i] : 0u;
        }
#pragma unroll
        for (uint32_t k = 0; k < 8u; ++k) {
            const uint32_t i = i0 + k * kScanThreads + t;
            if (i < padded) buf[i + (i >> 5)] = v[k];
        }
    }
    __syncthreads();
    const uint32_t r0 = t * k_run;
    uint32_t sum = 0;
    for (uint32_t k = 0; k < k_run; ++k) {
        const uint32_t i = r0 + k;
        sum += buf[i + (i >> 5)];
    }
    uint32_t all = 0;
    uint32_t run = block_exclusive_scan(sum, wsum, all);
    for (uint32_t k = 0; k < k_run; ++k) {
        const uint32_t i = r0 + k, j = i + (i >> 5);
        const uint32_t v = buf[j];
        buf[j] = run;
        run += v;
    }
    __syncthreads();
    for (uint32_t i = t; i < n; i += kScanThreads) out[i] = buf[i + (i >> 5)];
    if (t == 0 && total) *total = all;
}

constexpr uint32_t kMbThreads = 256;
constexpr uint32_t kMbItems = 8;
constexpr uint32_t kMbTile = kMbThreads * kMbItems;

__global__ __launch_bounds__(kMbThreads) void scan_tile_sums(const uint32_t *__restrict__ in, uint32_t n,
                                                             uint32_t *__restrict__ sums) {
    __shared__ uint32_t red[kMbThreads / 64];
    const uint32_t base = blockIdx.x * kMbTile;
    uint32_t s = 0;
#pragma unroll
    for (uint32_t j = 0; j < kMbItems; ++j) {
        const uint32_t idx = base + j * kMbThreads + threadIdx.x;
        if (idx < n) s += in[idx];
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) s += __shfl_xor(s, d, 64);
    if (lane_id() == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (uint32_t w = 0; w < kMbThreads / 64; ++w) t += red[w];
        sums[blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(kMbThreads) void scan_tiles(const uint32_t *__restrict__ in,
                                                         uint32_t *__restrict__ out, uint32_t n,
                                                         const uint32_t *__restrict__ offs) {
    __shared__ uint32_t tile[kMbTile];
    __shared__ uint32_t wsum[17];
    const uint32_t base = blockIdx.x * kMbTile, t = threadIdx.x;
#pragma unroll
    for (uint32_t j = 0; j < kMbItems; ++j) {
        const uint32_t idx = base + j * kMbThreads + t;
        tile[j * kMbThreads + t] = idx < n ? in[idx] : 0u;
    }
    __syncthreads();
    uint32_t v[kMbItems], sum = 0;
#pragma unroll
    for (uint32_t j = 0; j < kMbItems; ++j) { v[j] = tile[t * kMbItems + j]; sum += v[j]; }
    uint32_t btotal;
    uint32_t run = offs[blockIdx.x] + block_exclusive_scan(sum, wsum, btotal);
#pragma unroll
    for (uint32_t j = 0; j < kMbItems; ++j) { tile[t * kMbItems + j] = run; run += v[j]; }
    __syncthreads();
#pragma unroll
    for (uint32_t j = 0; j < kMbItems; ++j) {
        const uint32_t idx = base + j * kMbThreads + t;
        if (idx < n) out[idx] = tile[j * kMbThreads + t];
    }
}

__global__ __launch_bounds__(64) void clear_words_kernel(uint32_t *__restrict__ p, uint32_t n) {
    for (uint32_t i = threadIdx.x; i < n; i += 64u) p[i] = 0u;
}

int clear_words(hipStream_t s, uint32_t *p, uint32_t n) {
    if (n == 0) return DY_OK;
    clear_words_kernel<<<1, 64, 0, s>>>(p, n);
    DY_LAUNCH_CHECK();
    return DY_OK;
}

size_t scan_tmp_words(size_t n) {
    const size_t nb = (n + kMbTile - 1) / kMbTile;
    return 2 * nb + 2;
}

int scan_exclusive(hipStream_t s, const uint32_t *in, uint32_t *out, size_t n, uint32_t *d_total,
                   uint32_t *tmp) {
    if (n == 0) {
        if (d_total) DY_TRY(hipMemsetAsync(d_total, 0, sizeof(uint32_t), s));
        return DY_OK;
    }
    if (n > 0xffffffffu) return DY_ERR_INVALID_ARG;
    if (n <= kScanLdsMax) {
        scan_lds<<<1, kScanThreads, 0, s>>>(in, out, static_cast<uint32_t>(n), d_total);
        DY_LAUNCH_CHECK();
        return DY_OK;
    }
    if (n <= kScanSingleMax) {
        scan_one_block<<<1, kScanThreads, 0, s>>>(in, out, static_cast<uint32_t>(n), d_total);
        DY_LAUNCH_CHECK();
        return DY_OK;
    }
    const uint32_t nb = ceil_div(static_cast<uint32_t>(n), kMbTile);
    uint32_t *sums = tmp, *offs = tmp + nb;
    scan_tile_sums<<<nb, kMbThreads, 0, s>>>(in, static_cast<uint32_t>(n), sums);
    DY_LAUNCH_CHECK();
    scan_one_block<<<1, kScanThreads, 0, s>>>(sums, offs, nb, d_total);
    DY_LAUNCH_CHECK();
    scan_tiles<<<nb, kMbThreads, 0, s>>>(in, out, static_cast<uint32_t>(n), offs);
    DY_LAUNCH_CHECK();
    return DY_OK;
}

// ---------------------------------------------------------------------------
// Single-workgroup LSD radix sort, 8-bit digits, stable.
// Tile element order is (wave, item, lane) so that a wave's ballot-based
// ranking of item j, then j+1, ... preserves input order.
// ---------------------------------------------------------------------------
constexpr uint32_t kSortThreads = 1024;
constexpr uint32_t kSortWaves = kSortThreads / 64;
constexpr uint32_t kSortItems = 4;
constexpr uint32_t kSortTile = kSortThreads * kSortItems;

template <typename K>
__device__ __forceinline__ uint32_t digit_of(K k, int bit, uint32_t mask) {
    return static_cast<uint32_t>(k >> bit) & mask;
}

template <typename K>
__global__ __launch_bounds__(kSortThreads) void radix_sort_one_block(K *k0, uint32_t *v0, K *k1, uint32_t *v1,
                                                                     uint32_t n, int begin_bit, int end_bit) {
    __shared__ uint32_t hist[256];
    __shared__ uint32_t base[256];
    __shared__ uint32_t wcnt[kSortWaves][256];
    const uint32_t t = threadIdx.x, lane = lane_id(), wave = t >> 6;
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    K *src_k = k0, *dst_k = k1;
    uint32_t *src_v = v0, *dst_v = v1;
    const bool has_v = v0 != nullptr;
    for (int bit = begin_bit; bit < end_bit; bit += 8) {
        const int nbits = (end_bit - bit) < 8 ? (end_bit - bit) : 8;
        const uint32_t mask = (1u << nbits) - 1u;
        if (t < 256) hist[t] = 0;
        __syncthreads();
        for (uint32_t i = t; i < n; i += kSortThreads) atomicAdd(&hist[digit_of(src_k[i], bit, mask)], 1u);
        __syncthreads();
        if (wave == 0) {
            uint32_t h[4], s = 0;
#pragma unroll
            for (int j = 0; j < 4; ++j) { h[j] = hist[lane * 4 + j]; s += h[j]; }
            uint32_t run = wave_inclusive_scan(s) - s;
#pragma unroll
            for (int j = 0; j < 4; ++j) { base[lane * 4 + j] = run; run += h[j]; }
        }
        __syncthreads();
        for (uint32_t tile = 0; tile < n; tile += kSortTile) {
            for (uint32_t i = t; i < kSortWaves * 256; i += kSortThreads) (&wcnt[0][0])[i] = 0;
            __syncthreads();
            K kk[kSortItems];
            uint32_t vv[kSortItems], dd[kSortItems], rr[kSortItems];
#pragma unroll
            for (uint32_t j = 0; j < kSortItems; ++j) {
                const uint32_t idx = tile + wave * (64 * kSortItems) + j * 64 + lane;
                const bool valid = idx < n;
                kk[j] = valid ? src_k[idx] : K(0);
                vv[j] = (valid && has_v) ? src_v[idx] : 0u;
                const uint32_t d = valid ? digit_of(kk[j], bit, mask) : 256u;
                uint64_t peers = ~0ull;
#pragma unroll
                for (int b = 0; b < 9; ++b) {
                    const bool set = (d >> b) & 1u;
                    const uint64_t bal = __ballot(set);
                    peers &= set ? bal : ~bal;
                }
                const uint32_t pre = static_cast<uint32_t>(__popcll(peers & lt_mask));
                const uint32_t c = valid ? wcnt[wave][valid ? d : 0u] : 0u;
                rr[j] = c + pre;
                dd[j] = d;
                if (valid && pre == 0) wcnt[wave][d] = c + static_cast<uint32_t>(__popcll(peers));
                __builtin_amdgcn_wave_barrier();
            }
            __syncthreads();
            if (t < 256) {
                uint32_t run = base[t];
                for (uint32_t w = 0; w < kSortWaves; ++w) {
                    const uint32_t c = wcnt[w][t];
                    wcnt[w][t] = run;
                    run += c;
                }
                base[t] = run;
            }
            __syncthreads();
#pragma unroll
            for (uint32_t j = 0; j < kSortItems; ++j) {
                if (dd[j] < 256u) {
                    const uint32_t pos = wcnt[wave][dd[j]] + rr[j];
                    dst_k[pos] = kk[j];
                    if (has_v) dst_v[pos] = vv[j];
                }
            }
            __syncthreads();
        }
        K *tk = src_k; src_k = dst_k; dst_k = tk;
        uint32_t *tv = src_v; src_v = dst_v; dst_v = tv;
        __syncthreads();
    }
    if (src_k != k0) {
        for (uint32_t i = t; i < n; i += kSortThreads) {
            k0[i] = src_k[i];
            if (has_v) v0[i] = src_v[i];
        }
    }
}

template <typename K>
static int radix_sort_t(hipStream_t s, K *k0, uint32_t *v0, K *k1, uint32_t *v1, size_t n, int b0, int b1) {
    if (n <= 1) return DY_OK;
    if (n > 0xffffffffu) return DY_ERR_INVALID_ARG;
    radix_sort_one_block<K><<<1, kSortThreads, 0, s>>>(k0, v0, k1, v1, static_cast<uint32_t>(n), b0, b1);
    DY_LAUNCH_CHECK();
    return DY_OK;
}

int radix_sort_u32(hipStream_t s, uint32_t *k0, uint32_t *v0, uint32_t *k1, uint32_t *v1, size_t n, int b0,
                   int b1) {
    return radix_sort_t<uint32_t>(s, k0, v0, k1, v1, n, b0, b1);
}
int radix_sort_u64(hipStream_t s, uint64_t *k0, uint32_t *v0, uint64_t *k1, uint32_t *v1, size_t n, int b0,
                   int b1) {
    return radix_sort_t<uint64_t>(s, k0, v0, k1, v1, n, b0, b1);
}

// ---------------------------------------------------------------------------
// Deterministic sum: 256-wide halving tree per workgroup, partials merged on
// the host in index order (compute/include/compute/reduction.hpp:22-62).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void reduce_sum_partials(const float *__restrict__ d, size_t n,
                                                           float *__restrict__ partial) {
    __shared__ float sm[256];
    const size_t gid = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    const uint32_t lid = threadIdx.x;
    sm[lid] = gid < n ? d[gid] : 0.f;
    __syncthreads();
    for (uint32_t st = 128; st > 0; st >>= 1) {
        if (lid < st) sm[lid] += sm[lid + st];
        __syncthreads();
    }
    if (lid == 0) partial[blockIdx.x] = sm[0];
}

}  // namespace dy

using namespace dy;

namespace {
template <typename K>
static int sort_api(dy_stream_t s, K *keys, uint32_t *vals, size_t n) {
    if (n <= 1) return DY_OK;
    if (!keys) return DY_ERR_INVALID_ARG;
    K *k1 = nullptr;
    uint32_t *v1 = nullptr;
    DY_TRY(hipMallocAsync(reinterpret_cast<void **>(&k1), n * sizeof(K), hs(s)));
    if (vals) DY_TRY(hipMallocAsync(reinterpret_cast<void **>(&v1), n * sizeof(uint32_t), hs(s)));
    const int st = radix_sort_t<K>(hs(s), keys, vals, k1, v1, n, 0, static_cast<int>(8 * sizeof(K)));
    DY_TRY(hipFreeAsync(k1, hs(s)));
    if (v1) DY_TRY(hipFreeAsync(v1, hs(s)));
    return st;
}

}  // namespace

extern "C" {

int dy_scan_exclusive_u32(dy_stream_t s, const uint32_t *d_in, uint32_t *d_out, size_t n, uint32_t *d_total) {
    if ((n && (!d_in || !d_out))) return DY_ERR_INVALID_ARG;
    uint32_t *tmp = nullptr;
    if (n > kScanSingleMax) DY_TRY(hipMallocAsync(reinterpret_cast<void **>(&tmp),
                                                  scan_tmp_words(n) * sizeof(uint32_t), hs(s)));
    const int st = scan_exclusive(hs(s), d_in, d_out, n, d_total, tmp);
    if (tmp) DY_TRY(hipFreeAsync(tmp, hs(s)));
    return st;
}

int dy_sort_by_key_u64_u32(dy_stream_t s, uint64_t *keys, uint32_t *vals, size_t n) {
    if (n > 1 && !vals) return DY_ERR_INVALID_ARG;
    return sort_api<uint64_t>(s, keys, vals, n);
}
int dy_sort_by_key_u32_u32(dy_stream_t s, uint32_t *keys, uint32_t *vals, size_t n) {
    if (n > 1 && !vals) return DY_ERR_INVALID_ARG;
    return sort_api<uint32_t>(s, keys, vals, n);
}
int dy_sort_u32(dy_stream_t s, uint32_t *keys, size_t n) { return sort_api<uint32_t>(s, keys, nullptr, n); }
int dy_sort_u64(dy_stream_t s, uint64_t *keys, size_t n) { return sort_api<uint64_t>(s, keys, nullptr, n); }

int dy_reduce_sum_f32(dy_stream_t s, const float *d_data, size_t n, float *out) {
    if (!out) return DY_ERR_INVALID_ARG;
    *out = 0.f;
    if (n == 0) return DY_OK;
    if (!d_data) return DY_ERR_INVALID_ARG;
    const size_t groups = (n + 255) / 256;
    float *d_part = nullptr;
    DY_TRY(hipMallocAsync(reinterpret_cast<void **>(&d_part), groups * sizeof(float), hs(s)));
    reduce_sum_partials<<<static_cast<uint32_t>(groups), 256, 0, hs(s)>>>(d_data, n, d_part);
    DY_LAUNCH_CHECK();
    std::vector<float> h(groups);
    DY_TRY(hipMemcpyAsync(h.data(), d_part, groups * sizeof(float), hipMemcpyDeviceToHost, hs(s)));
    DY_TRY(hipFreeAsync(d_part, hs(s)));
    DY_TRY(hipStreamSynchronize(hs(s)));
    float r = 0.f;
    for (size_t i = 0; i < groups; ++i) r += h[i];
    *out = r;
    return DY_OK;
}

}  // extern "C"
