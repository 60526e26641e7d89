// narrowphase.hip — primitive contact generation for gfx950
// (core/src/narrowphase.cpp:920-1127, sphere/box shapes).
//
// Each in-scope pair yields at most one contact (box-box manifolds reduce to
// their centroid), so the reference's append + (pair, sub) key sort + gather
// (narrowphase.cpp:1077-1127) reduces to an order-preserving compaction of the
// pair list.  It is done in one pass: workgroups claim 256-pair tiles in
// ascending order from a ticket, test one pair per lane, scan the hits of the
// tile, publish the tile's count and sum the earlier tiles' published counts
// for the tile's output offset (a single-pass scan after Merrill & Garland,
// "Single-pass Parallel Prefix Scan with Decoupled Look-back", 2016, reading
// every predecessor at once); each contact is then written straight to its
// slot.  Contacts come out in pair order by construction; no
// per-pair staging, no sort.  The launch is bounded by the device-held pair
// count: tiles past it are never claimed.
#include <cstdlib>

#include "narrowphase_pass.hpp"
#include "pair_contact.hpp"

namespace dy {

constexpr uint32_t kNpThreads = kNpTilePairs;
// Persistent workgroups at most: 2 per CU on 256 CUs (a sweep over 256, 512,
// 1024 and 2048 on the 32 768-body frame put 512 first).
constexpr uint32_t kNpMaxGrid = 512;
// Tile count word: the run's epoch in bits 32-63, the tile's hit count in
// bits 0-31; a word whose epoch is not the run's is not yet published (so the
// array needs no clearing between runs: each run has a new epoch, and memory
// from before the first run is zero, which no epoch takes).  The word carries
// everything it publishes, so it is written and read with relaxed agent-scope
// atomics: coherent across the XCDs' L2s by the memory model's
// single-location coherence, with no release/acquire fence (on gfx950 an
// agent-scope fence writes back or invalidates the whole XCD L2, which the
// other workgroups are reading bodies from).
constexpr uint32_t kNpSpinLimit = 1u << 24;

__device__ __forceinline__ void np_publish(uint64_t *status, uint32_t t, uint32_t epoch, uint32_t count) {
    __hip_atomic_store(&status[t], (static_cast<uint64_t>(epoch) << 32) | count, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}

// Exclusive prefix of tile t, by wave 0: the sum of the counts of tiles
// 0..t-1, all read at once (lane l reads tiles l, l + 64, ...), again until
// every one of them has been published.  Every tile publishes its count as
// soon as its pairs are tested, so this is about one round trip; the reads
// total t words per tile.
__device__ __forceinline__ uint32_t np_lookback(uint64_t *status, uint32_t t, uint32_t epoch, uint32_t *err) {
    const uint32_t lane = lane_id();
    for (uint32_t spins = 0;; ++spins) {
        uint32_t sum = 0;
        bool waiting = false;
#pragma unroll 8
        for (uint32_t j = lane; j < t; j += 64) {
            const uint64_t w = __hip_atomic_load(&status[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            waiting = waiting || static_cast<uint32_t>(w >> 32) != epoch;
            sum += static_cast<uint32_t>(w);
        }
        if (__ballot(waiting) == 0) {
#pragma unroll
            for (int d = 32; d >= 1; d >>= 1) sum += __shfl_xor(sum, d, 64);
            return sum;
        }
        if (spins > kNpSpinLimit) {
            if (lane == 0) atomicOr(err, 1u);
            return 0;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

// Persistent pass: claim tile t, test its pairs, publish the tile's count,
// look back for its offset, publish the inclusive count, write the contacts.
__global__ __launch_bounds__(kNpThreads) void np_fused(const dy_contact_pair *__restrict__ pairs, uint32_t n_bound,
                                                       const uint32_t *__restrict__ d_n, dy_body_view bv,
                                                       dy_shape_view sv, uint64_t *status, uint32_t epoch,
                                                       NpOut out) {
    __shared__ uint32_t wsum[17];
    __shared__ uint32_t s_tile, s_prefix;
    const uint32_t n = d_n ? min(*d_n, n_bound) : n_bound;
    const uint32_t n_tiles = (n + kNpThreads - 1) / kNpThreads;
    const uint32_t cap = out.cv.capacity;
    for (;;) {
        if (threadIdx.x == 0) s_tile = atomicAdd(&out.counts[kNpTicket], 1u);
        __syncthreads();
        const uint32_t t = s_tile;
        if (t >= n_tiles) break;  // uniform: every wave of the workgroup leaves here
        const uint32_t i = t * kNpThreads + threadIdx.x;
        uint32_t hit = 0, unsup = 0;
        Hit h{};
        dy_contact_pair p{0u, 0u};
        if (i < n) {
            p = pairs[i];
            hit = pair_contact(bv, sv, p.a, p.b, h, unsup) ? 1u : 0u;
        }
        if (unsup) atomicOr(&out.counts[2], 1u);
        uint32_t total;
        const uint32_t local = block_exclusive_scan(hit, wsum, total);
        if (threadIdx.x < 64) {
            if (threadIdx.x == 0) np_publish(status, t, epoch, total);
            const uint32_t prefix = t ? np_lookback(status, t, epoch, &out.counts[3]) : 0u;
            if (threadIdx.x == 0) {
                s_prefix = prefix;
                if (t == n_tiles - 1) {  // the last tile knows the true count
                    const uint32_t all = prefix + total;
                    out.counts[1] = all;
                    out.counts[0] = all < cap ? all : cap;
                }
            }
        }
        __syncthreads();
        const uint32_t pos = s_prefix + local;
        if (hit && pos < cap) {
            const dy_contact_view &cv = out.cv;
            cv.body_a[pos] = p.a;
            cv.body_b[pos] = p.b;
            cv.pos_x[pos] = h.px; cv.pos_y[pos] = h.py; cv.pos_z[pos] = h.pz;
            cv.norm_x[pos] = h.nx; cv.norm_y[pos] = h.ny; cv.norm_z[pos] = h.nz;
            cv.depth[pos] = h.depth;
            if (out.feat) out.feat[pos] = h.feat;
        }
        __syncthreads();  // s_tile and s_prefix are rewritten by the next claim
    }
}

// Persistent workgroups for a launch over at most n_bound pairs
// (DYPHUR_NP_GRID overrides the cap of kNpMaxGrid).
uint32_t np_grid(uint32_t n_bound) {
    static const uint32_t cap = [] {
        const char *env = std::getenv("DYPHUR_NP_GRID");
        const int v = env ? std::atoi(env) : 0;
        return v > 0 ? static_cast<uint32_t>(v) : kNpMaxGrid;
    }();
    const uint32_t tiles = ceil_div(n_bound, kNpThreads);
    return tiles < cap ? tiles : cap;
}

int np_enqueue(hipStream_t st, const dy_contact_pair *pairs, uint32_t n_bound, const uint32_t *d_n,
               const dy_body_view &bv, const dy_shape_view &sv, uint64_t *status, uint32_t epoch, NpOut out) {
    if (epoch == 0) return DY_ERR_INVALID_ARG;
    DY_TRY_STATUS(clear_words(st, out.counts, kNpCounts));
    if (n_bound == 0) return DY_OK;
    np_fused<<<np_grid(n_bound), kNpThreads, 0, st>>>(pairs, n_bound, d_n, bv, sv, status, epoch, out);
    DY_LAUNCH_CHECK();
    return DY_OK;
}

}  // namespace dy

using namespace dy;

struct dy_narrowphase_s {
    uint32_t cap = 0;
    // store
    uint32_t *body_a = nullptr, *body_b = nullptr;
    float *px = nullptr, *py = nullptr, *pz = nullptr, *nx = nullptr, *ny = nullptr, *nz = nullptr, *depth = nullptr;
    uint32_t *counts = nullptr;  // kNpCounts words
    uint32_t *feat = nullptr;    // feature words (dy_narrowphase_set_features), else null
    bool want_feat = false;
    // tile count words (grow with the launch bound; zeroed when allocated)
    DevBuf<uint64_t> status;
    uint32_t epoch = 0;
};

namespace {
int np_launch(dy_narrowphase_s *np, hipStream_t st, const dy_contact_pair *pairs, uint32_t n_bound,
              const uint32_t *d_n, const dy_body_view &bv, const dy_shape_view &sv) {
    const size_t tiles = ceil_div(n_bound > 0 ? n_bound : 1u, kNpThreads);
    if (np->status.cap < tiles) {
        DY_TRY_STATUS(np->status.ensure(tiles));
        DY_TRY(hipMemsetAsync(np->status.p, 0, np->status.cap * sizeof(uint64_t), st));
    }
    if (++np->epoch == 0) {  // after 2^32 - 1 runs: clear once and start again
        DY_TRY(hipMemsetAsync(np->status.p, 0, np->status.cap * sizeof(uint64_t), st));
        np->epoch = 1;
    }
    const dy_contact_view cv{np->body_a, np->body_b, np->px, np->py, np->pz, np->nx, np->ny, np->nz, np->depth,
                             np->counts, np->cap};
    return np_enqueue(st, pairs, n_bound, d_n, bv, sv, np->status.p, np->epoch,
                      NpOut{cv, np->counts, np->want_feat ? np->feat : nullptr});
}

void np_free(dy_narrowphase_s *np) {
    void *ptrs[] = {np->body_a, np->body_b, np->px, np->py, np->pz, np->nx, np->ny, np->nz, np->depth,
                    np->counts, np->feat};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    np->status.release();
}

}  // namespace

extern "C" {

int dy_narrowphase_create(dy_stream_t s, uint32_t max_contacts, dy_narrowphase_t *out) {
    if (!out) return DY_ERR_INVALID_ARG;
    *out = nullptr;
    dy_narrowphase_s *np = new dy_narrowphase_s;
    np->cap = max_contacts;
    const size_t c = max_contacts ? max_contacts : 1;
    hipError_t e = hipSuccess;
    uint32_t **u[] = {&np->body_a, &np->body_b};
    float **f[] = {&np->px, &np->py, &np->pz, &np->nx, &np->ny, &np->nz, &np->depth};
    for (auto p : u)
        if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(p), c * sizeof(uint32_t));
    for (auto p : f)
        if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(p), c * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&np->counts), kNpCounts * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemsetAsync(np->counts, 0, kNpCounts * sizeof(uint32_t), hs(s));
    if (e != hipSuccess) {
        np_free(np);
        delete np;
        return static_cast<int>(e);
    }
    *out = np;
    return DY_OK;
}

int dy_narrowphase_destroy(dy_narrowphase_t np) {
    if (!np) return DY_OK;
    (void)hipDeviceSynchronize();
    np_free(np);
    delete np;
    return DY_OK;
}

int dy_narrowphase_run(dy_narrowphase_t np, dy_stream_t s, const dy_contact_pair *d_pairs, uint32_t n_pairs,
                       const dy_body_view *bodies, const dy_shape_view *shapes) {
    if (!np || !bodies || !shapes || (n_pairs && !d_pairs)) return DY_ERR_INVALID_ARG;
    return np_launch(np, hs(s), d_pairs, n_pairs, nullptr, *bodies, *shapes);
}

int dy_narrowphase_run_device_count(dy_narrowphase_t np, dy_stream_t s, const dy_contact_pair *d_pairs,
                                    const uint32_t *d_n_pairs, uint32_t max_pairs, const dy_body_view *bodies,
                                    const dy_shape_view *shapes) {
    if (!np || !bodies || !shapes || !d_n_pairs || (max_pairs && !d_pairs)) return DY_ERR_INVALID_ARG;
    return np_launch(np, hs(s), d_pairs, max_pairs, d_n_pairs, *bodies, *shapes);
}

int dy_narrowphase_contacts(dy_narrowphase_t np, dy_contact_view *out) {
    if (!np || !out) return DY_ERR_INVALID_ARG;
    *out = dy_contact_view{np->body_a, np->body_b, np->px, np->py, np->pz, np->nx, np->ny, np->nz, np->depth,
                           np->counts, np->cap};
    return DY_OK;
}

int dy_narrowphase_set_features(dy_narrowphase_t np, int enable) {
    if (!np) return DY_ERR_INVALID_ARG;
    if (enable && !np->feat)
        DY_TRY(hipMalloc(reinterpret_cast<void **>(&np->feat), (np->cap ? np->cap : 1u) * sizeof(uint32_t)));
    np->want_feat = enable != 0;
    return DY_OK;
}

int dy_narrowphase_features(dy_narrowphase_t np, const uint32_t **d_features) {
    if (!np || !d_features) return DY_ERR_INVALID_ARG;
    *d_features = np->want_feat ? np->feat : nullptr;
    return np->want_feat ? DY_OK : DY_ERR_INVALID_ARG;
}

int dy_narrowphase_download_count(dy_narrowphase_t np, dy_stream_t s, uint32_t *out) {
    if (!np || !out) return DY_ERR_INVALID_ARG;
    uint32_t h[kNpCounts];
    DY_TRY(hipMemcpyAsync(h, np->counts, sizeof(h), hipMemcpyDeviceToHost, hs(s)));
    DY_TRY(hipStreamSynchronize(hs(s)));
    *out = h[0];
    if (h[3]) return DY_ERR_INTERNAL;
    return h[2] ? DY_ERR_UNSUPPORTED : DY_OK;
}

int dy_narrowphase_last_count(dy_narrowphase_t np, dy_stream_t s, uint32_t *out) {
    if (!np || !out) return DY_ERR_INVALID_ARG;
    uint32_t h[kNpCounts];
    DY_TRY(hipMemcpyAsync(h, np->counts, sizeof(h), hipMemcpyDeviceToHost, hs(s)));
    DY_TRY(hipStreamSynchronize(hs(s)));
    *out = h[1];
    if (h[3]) return DY_ERR_INTERNAL;
    return h[2] ? DY_ERR_UNSUPPORTED : DY_OK;
}

}  // extern "C"
