// solver.hip — XPBD contact solver for gfx950 (core/src/xpbd_solver.cpp:558-733).
//
// Both reference paths are executed as "class sweeps": every contact gets a
// class, contacts of one class share no dynamic body (so they run in parallel
// with disjoint writes), and classes run in increasing order, n_iters times.
//
//  * Coloured path (more than 256 contacts): class = the reference's
//    Jones-Plassmann colour.  With priorities (hash(k) << 32 | k), JP gives
//    each constraint the lowest colour unused by its higher-priority
//    neighbours, i.e. greedy colouring in decreasing priority order; the round
//    in which the reference colours k is the length of the longest
//    decreasing-priority chain ending at k.  assign_round computes exactly
//    that (colour and depth) with one pass per round that may also finish
//    deeper constraints early — the values do not depend on timing.  The
//    reference's serial fallbacks (a body needing > 64 colours, or more than
//    256 rounds) are detected from the same data.
//  * Reference-order path (<= 256 contacts or force_serial): run_serial
//    applies contacts 0..nc-1 in order.  Contacts only interact through shared
//    dynamic bodies, so class = dependency level (1 + the level of the latest
//    earlier contact on either dynamic body) reproduces the serial result bit
//    for bit while running each level in parallel.
//
// Quaternion renormalisation after every iteration is applied lazily: a
// dynamic body is brought to `it` renormalisations when a contact of
// iteration `it` first touches it, and every body to n_iters at the end.
// Static bodies are never written by contacts and their orientation only
// enters products with a zero inverse inertia, so they are renormalised at
// the end.  Every body ends with the reference's sequence of operations.
#include <vector>

#include "dy_internal.hpp"

namespace dy {

constexpr uint64_t kUnassigned = ~0ull;
constexpr uint32_t kSolveThreads = 256;
constexpr uint32_t kMetaWords = 16;   // [0] remaining [1] overflow [2] max class [3] max depth
constexpr uint32_t kHistDownload = 256;

__device__ __forceinline__ uint32_t cls_of(uint64_t s) { return static_cast<uint32_t>(s & 0xffffffffu); }
__device__ __forceinline__ uint32_t dep_of(uint64_t s) { return static_cast<uint32_t>(s >> 32); }

__global__ void solve_prep(uint32_t nb, uint32_t nc, uint32_t *__restrict__ deg, uint32_t *__restrict__ fill,
                           uint32_t *__restrict__ renorm, float *__restrict__ lam, uint64_t *__restrict__ state,
                           uint32_t *__restrict__ meta, uint32_t n_hist, uint32_t *__restrict__ hist) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nb) { deg[i] = 0; fill[i] = 0; renorm[i] = 0; }
    if (i < nc) { lam[i] = 0.f; state[i] = kUnassigned; }
    if (i < n_hist) hist[i] = 0;
    if (i < kMetaWords) meta[i] = (i == 0) ? nc : 0u;
}

// Reset the class assignment only (serial fallback after a failed colouring).
__global__ void solve_reassign_prep(uint32_t nc, uint64_t *__restrict__ state, uint32_t *__restrict__ meta,
                                    uint32_t n_hist, uint32_t *__restrict__ hist) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nc) state[i] = kUnassigned;
    if (i < n_hist) hist[i] = 0;
    if (i < kMetaWords) meta[i] = (i == 0) ? nc : 0u;
}

__global__ void solve_degree(dy_contact_view cv, uint32_t nc, const float *__restrict__ im, uint32_t *__restrict__ deg) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const uint32_t a = cv.body_a[c], b = cv.body_b[c];
    if (im[a] != 0.f) atomicAdd(&deg[a], 1u);
    if (im[b] != 0.f) atomicAdd(&deg[b], 1u);
}

__global__ void solve_fill(dy_contact_view cv, uint32_t nc, const float *__restrict__ im,
                           const uint32_t *__restrict__ off, uint32_t *__restrict__ fill, uint32_t *__restrict__ adj) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const uint32_t a = cv.body_a[c], b = cv.body_b[c];
    if (im[a] != 0.f) adj[off[a] + atomicAdd(&fill[a], 1u)] = c;
    if (im[b] != 0.f) adj[off[b] + atomicAdd(&fill[b], 1u)] = c;
}

__global__ __launch_bounds__(kSolveThreads) void assign_round(dy_contact_view cv, uint32_t nc, const float *__restrict__ im,
                                                              const uint32_t *__restrict__ off,
                                                              const uint32_t *__restrict__ deg,
                                                              const uint32_t *__restrict__ adj, uint64_t *state,
                                                              uint32_t *meta, uint32_t *hist, int serial) {
    if (__hip_atomic_load(&meta[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u) return;
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    if (__hip_atomic_load(&state[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != kUnassigned) return;
    const uint32_t ends[2] = {cv.body_a[c], cv.body_b[c]};
    const uint64_t pc = jp_priority(c);
    uint64_t used = 0;
    uint32_t depth = 0;
    for (int e = 0; e < 2; ++e) {
        const uint32_t x = ends[e];
        if (im[x] == 0.f) continue;
        const uint32_t beg = off[x], cnt = deg[x];
        for (uint32_t q = 0; q < cnt; ++q) {
            const uint32_t j = adj[beg + q];
            if (j == c) continue;
            const bool higher = serial ? (j < c) : (jp_priority(j) > pc);
            if (!higher) continue;
            const uint64_t sj = __hip_atomic_load(&state[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (sj == kUnassigned) return;  // a higher-priority neighbour is still open
            used |= 1ull << (cls_of(sj) & 63u);
            const uint32_t dj = dep_of(sj);
            depth = dj > depth ? dj : depth;
        }
    }
    uint32_t cls;
    if (serial) {
        cls = depth;  // level
    } else if (used == ~0ull) {
        atomicOr(&meta[1], 1u);
        cls = 64u;
    } else {
        cls = static_cast<uint32_t>(__builtin_ctzll(~used));
    }
    const uint32_t d = depth + 1u;
    __hip_atomic_store(&state[c], (static_cast<uint64_t>(d) << 32) | cls, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    atomicAdd(&hist[cls], 1u);
    atomicMax(&meta[2], cls);
    atomicMax(&meta[3], d);
    atomicSub(&meta[0], 1u);
}

__global__ void bucket_scatter(uint32_t nc, const uint64_t *__restrict__ state, const uint32_t *__restrict__ boff,
                               uint32_t *__restrict__ bfill, uint32_t *__restrict__ bucket) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const uint32_t k = cls_of(state[c]);
    bucket[boff[k] + atomicAdd(&bfill[k], 1u)] = c;
}

__device__ __forceinline__ BodyState load_body(const dy_body_view &bv, uint32_t i) {
    BodyState s;
    s.px = bv.pos_x[i]; s.py = bv.pos_y[i]; s.pz = bv.pos_z[i];
    s.qw = bv.rot_w[i]; s.qx = bv.rot_x[i]; s.qy = bv.rot_y[i]; s.qz = bv.rot_z[i];
    s.vx = bv.vel_x[i]; s.vy = bv.vel_y[i]; s.vz = bv.vel_z[i];
    s.wx = bv.ang_x[i]; s.wy = bv.ang_y[i]; s.wz = bv.ang_z[i];
    s.im = bv.inv_mass[i];
    s.ixx = bv.iI_xx[i]; s.iyy = bv.iI_yy[i]; s.izz = bv.iI_zz[i];
    s.ixy = bv.iI_xy[i]; s.ixz = bv.iI_xz[i]; s.iyz = bv.iI_yz[i];
    return s;
}

__device__ __forceinline__ void store_body(const dy_body_view &bv, uint32_t i, const BodyState &s) {
    bv.pos_x[i] = s.px; bv.pos_y[i] = s.py; bv.pos_z[i] = s.pz;
    bv.rot_w[i] = s.qw; bv.rot_x[i] = s.qx; bv.rot_y[i] = s.qy; bv.rot_z[i] = s.qz;
    bv.vel_x[i] = s.vx; bv.vel_y[i] = s.vy; bv.vel_z[i] = s.vz;
    bv.ang_x[i] = s.wx; bv.ang_y[i] = s.wy; bv.ang_z[i] = s.wz;
}

__device__ __forceinline__ void catch_up(BodyState &s, uint32_t done, uint32_t target) {
    for (uint32_t r = done; r < target; ++r) renormalise(s.qw, s.qx, s.qy, s.qz);
}

__global__ __launch_bounds__(kSolveThreads) void solve_bucket(const uint32_t *__restrict__ bucket, uint32_t count,
                                                              uint32_t it, dy_contact_view cv, dy_body_view bv,
                                                              float *__restrict__ lam, uint32_t *__restrict__ renorm,
                                                              float mu, float dt) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const uint32_t c = bucket[i];
    const uint32_t a = cv.body_a[c], b = cv.body_b[c];
    const ContactIn ci{cv.pos_x[c], cv.pos_y[c], cv.pos_z[c], cv.norm_x[c], cv.norm_y[c], cv.norm_z[c], cv.depth[c]};
    BodyState A = load_body(bv, a), B = load_body(bv, b);
    const bool da = A.im != 0.f, db = B.im != 0.f;
    if (da) catch_up(A, renorm[a], it);
    if (db) catch_up(B, renorm[b], it);
    float l = lam[c];
    project_contact(ci, A, B, mu, l, dt);
    lam[c] = l;
    if (da) { store_body(bv, a, A); renorm[a] = it; }
    if (db) { store_body(bv, b, B); renorm[b] = it; }
}

__global__ void final_renorm(dy_body_view bv, const uint32_t *__restrict__ renorm, uint32_t n_iters, int use_counts) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= bv.n) return;
    const uint32_t done = use_counts ? renorm[i] : 0u;
    if (done >= n_iters) return;
    float qw = bv.rot_w[i], qx = bv.rot_x[i], qy = bv.rot_y[i], qz = bv.rot_z[i];
    for (uint32_t r = done; r < n_iters; ++r) renormalise(qw, qx, qy, qz);
    bv.rot_w[i] = qw; bv.rot_x[i] = qx; bv.rot_y[i] = qy; bv.rot_z[i] = qz;
}

}  // namespace dy

using namespace dy;

struct dy_xpbd_s {
    int n_iters = 10;
    float mu = 0.5f;
    int force_serial = 0;
    uint32_t cap = 0;
    float *lam = nullptr;
    uint64_t *state = nullptr;
    uint32_t *adj = nullptr, *bucket = nullptr, *boff = nullptr, *bfill = nullptr;
    uint32_t *meta_hist = nullptr;  // kMetaWords meta + (cap + 65) histogram
    DevBuf<uint32_t> deg, off, fill, renorm, scan_tmp;
    dy_solve_stats stats{};
    std::vector<uint32_t> h_meta_hist;
    std::vector<uint32_t> h_boff;
};

namespace {
void xp_free(dy_xpbd_s *x) {
    void *ptrs[] = {x->lam, x->state, x->adj, x->bucket, x->boff, x->bfill, x->meta_hist};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    x->deg.release(); x->off.release(); x->fill.release(); x->renorm.release(); x->scan_tmp.release();
}

// Runs assignment rounds until every contact has a class.  Returns DY_OK with
// meta/hist downloaded into x->h_meta_hist, or 1 when the coloured path must
// fall back (too many rounds).
int assign_classes(dy_xpbd_s *x, hipStream_t st, const dy_contact_view &cv, uint32_t nc, const float *im, int serial,
                   uint32_t &k_classes) {
    uint32_t *meta = x->meta_hist, *hist = x->meta_hist + kMetaWords;
    const uint32_t grid = ceil_div(nc, kSolveThreads);
    const uint32_t max_rounds = serial ? nc + 1u : 256u;
    uint32_t rounds = 0, batch = serial ? 16u : 12u;
    x->h_meta_hist.resize(kMetaWords + kHistDownload);
    while (true) {
        for (uint32_t r = 0; r < batch && rounds < max_rounds; ++r, ++rounds) {
            assign_round<<<grid, kSolveThreads, 0, st>>>(cv, nc, im, x->off.p, x->deg.p, x->adj, x->state, meta, hist,
                                                         serial);
            DY_LAUNCH_CHECK();
        }
        DY_TRY(hipMemcpyAsync(x->h_meta_hist.data(), meta, (kMetaWords + kHistDownload) * sizeof(uint32_t),
                              hipMemcpyDeviceToHost, st));
        DY_TRY(hipStreamSynchronize(st));
        if (x->h_meta_hist[0] == 0u) break;
        if (rounds >= max_rounds) return serial ? DY_ERR_INTERNAL : 1;
        batch = 8;
    }
    k_classes = x->h_meta_hist[2] + 1u;
    if (k_classes > kHistDownload) {
        x->h_meta_hist.resize(kMetaWords + k_classes);
        DY_TRY(hipMemcpyAsync(x->h_meta_hist.data() + kMetaWords + kHistDownload, hist + kHistDownload,
                              (k_classes - kHistDownload) * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        DY_TRY(hipStreamSynchronize(st));
    }
    return DY_OK;
}
}  // namespace

extern "C" {

int dy_xpbd_create(dy_stream_t s, int n_iters, uint32_t max_contacts, float friction, dy_xpbd_t *out) {
    if (!out || n_iters < 0) return DY_ERR_INVALID_ARG;
    *out = nullptr;
    dy_xpbd_s *x = new dy_xpbd_s;
    x->n_iters = n_iters;
    x->mu = friction;
    x->cap = max_contacts;
    const size_t c = max_contacts ? max_contacts : 1;
    hipError_t e = hipMalloc(reinterpret_cast<void **>(&x->lam), c * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&x->state), c * sizeof(uint64_t));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&x->adj), 2 * c * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&x->bucket), c * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&x->boff), (c + 66) * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&x->bfill), (c + 66) * sizeof(uint32_t));
    const size_t hist_words = (c + 66) > kHistDownload ? (c + 66) : kHistDownload;
    if (e == hipSuccess)
        e = hipMalloc(reinterpret_cast<void **>(&x->meta_hist), (kMetaWords + hist_words) * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemsetAsync(x->lam, 0, c * sizeof(float), hs(s));
    if (e != hipSuccess) {
        xp_free(x);
        delete x;
        return static_cast<int>(e);
    }
    *out = x;
    return DY_OK;
}

int dy_xpbd_destroy(dy_xpbd_t x) {
    if (!x) return DY_OK;
    (void)hipDeviceSynchronize();
    xp_free(x);
    delete x;
    return DY_OK;
}

int dy_xpbd_set_force_serial(dy_xpbd_t x, int force_serial) {
    if (!x) return DY_ERR_INVALID_ARG;
    x->force_serial = force_serial ? 1 : 0;
    return DY_OK;
}

int dy_xpbd_solve(dy_xpbd_t x, dy_stream_t s, const dy_contact_view *cv, uint32_t n_joints, const dy_body_view *bv,
                  float dt) {
    if (!x || !cv || !bv) return DY_ERR_INVALID_ARG;
    if (n_joints != 0) return DY_ERR_UNSUPPORTED;
    hipStream_t st = hs(s);
    x->stats = dy_solve_stats{};
    uint32_t nc = 0;
    DY_TRY(hipMemcpyAsync(&nc, cv->n, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    DY_TRY(hipStreamSynchronize(st));
    x->stats.n_constraints = nc;
    if (nc == 0) return DY_OK;
    if (nc > x->cap) return DY_ERR_INVALID_ARG;
    const uint32_t nb = bv->n;
    DY_TRY_STATUS(x->deg.ensure(nb));
    DY_TRY_STATUS(x->off.ensure(nb));
    DY_TRY_STATUS(x->fill.ensure(nb));
    DY_TRY_STATUS(x->renorm.ensure(nb));
    DY_TRY_STATUS(x->scan_tmp.ensure(scan_tmp_words(nb > nc + 66u ? nb : nc + 66u)));
    uint32_t *meta = x->meta_hist, *hist = x->meta_hist + kMetaWords;
    const uint32_t n_hist = nc + 65u;
    const uint32_t prep_n = (nb > nc ? nb : nc) > n_hist ? (nb > nc ? nb : nc) : n_hist;
    solve_prep<<<ceil_div(prep_n, 256), 256, 0, st>>>(nb, nc, x->deg.p, x->fill.p, x->renorm.p, x->lam, x->state,
                                                      meta, n_hist, hist);
    DY_LAUNCH_CHECK();
    const uint32_t gc = ceil_div(nc, kSolveThreads);
    solve_degree<<<gc, kSolveThreads, 0, st>>>(*cv, nc, bv->inv_mass, x->deg.p);
    DY_LAUNCH_CHECK();
    DY_TRY_STATUS(scan_exclusive(st, x->deg.p, x->off.p, nb, nullptr, x->scan_tmp.p));
    solve_fill<<<gc, kSolveThreads, 0, st>>>(*cv, nc, bv->inv_mass, x->off.p, x->fill.p, x->adj);
    DY_LAUNCH_CHECK();

    int serial = (x->force_serial || nc <= 256u) ? 1 : 0;
    uint32_t k = 0;
    int r = assign_classes(x, st, *cv, nc, bv->inv_mass, serial, k);
    if (r < 0) return r;
    if (!serial) {
        const uint32_t overflow = x->h_meta_hist[1], depth = x->h_meta_hist[3];
        x->stats.rounds = r == 1 ? 257u : depth;
        x->stats.overflow = overflow;
        if (r == 1 || overflow || depth > 256u) {
            // The reference runs its serial path in these cases.
            serial = 1;
            solve_reassign_prep<<<ceil_div(n_hist, 256), 256, 0, st>>>(nc, x->state, meta, n_hist, hist);
            DY_LAUNCH_CHECK();
            r = assign_classes(x, st, *cv, nc, bv->inv_mass, 1, k);
            if (r != DY_OK) return r < 0 ? r : DY_ERR_INTERNAL;
        }
    }
    x->stats.used_serial = static_cast<uint32_t>(serial);
    x->stats.n_classes = k;
    if (serial) x->stats.rounds = x->h_meta_hist[3];

    // Bucket contacts by class: offsets scanned on the device from the same
    // histogram the host downloaded (the host copy only sizes the launches).
    const uint32_t *h_hist = x->h_meta_hist.data() + kMetaWords;
    x->h_boff.resize(k);
    uint32_t run = 0;
    for (uint32_t c = 0; c < k; ++c) { x->h_boff[c] = run; run += h_hist[c]; }
    if (run != nc) return DY_ERR_INTERNAL;
    DY_TRY_STATUS(scan_exclusive(st, hist, x->boff, k, nullptr, x->scan_tmp.p));
    DY_TRY(hipMemsetAsync(x->bfill, 0, k * sizeof(uint32_t), st));
    bucket_scatter<<<gc, kSolveThreads, 0, st>>>(nc, x->state, x->boff, x->bfill, x->bucket);
    DY_LAUNCH_CHECK();

    for (int it = 0; it < x->n_iters; ++it) {
        for (uint32_t c = 0; c < k; ++c) {
            const uint32_t cnt = h_hist[c];
            if (cnt == 0) continue;
            solve_bucket<<<ceil_div(cnt, kSolveThreads), kSolveThreads, 0, st>>>(
                x->bucket + x->h_boff[c], cnt, static_cast<uint32_t>(it), *cv, *bv, x->lam, x->renorm.p, x->mu, dt);
            DY_LAUNCH_CHECK();
        }
    }
    final_renorm<<<ceil_div(nb, 256), 256, 0, st>>>(*bv, x->renorm.p, static_cast<uint32_t>(x->n_iters), 1);
    DY_LAUNCH_CHECK();
    return DY_OK;
}

int dy_xpbd_last_stats(dy_xpbd_t x, dy_solve_stats *out) {
    if (!x || !out) return DY_ERR_INVALID_ARG;
    *out = x->stats;
    return DY_OK;
}

int dy_xpbd_download_classes(dy_xpbd_t x, dy_stream_t s, uint32_t *out, uint32_t n) {
    if (!x || (!out && n)) return DY_ERR_INVALID_ARG;
    if (n > x->cap) return DY_ERR_INVALID_ARG;
    std::vector<uint64_t> h(n);
    if (n) {
        DY_TRY(hipMemcpyAsync(h.data(), x->state, n * sizeof(uint64_t), hipMemcpyDeviceToHost, hs(s)));
        DY_TRY(hipStreamSynchronize(hs(s)));
    }
    for (uint32_t i = 0; i < n; ++i) out[i] = static_cast<uint32_t>(h[i] & 0xffffffffu);
    return DY_OK;
}

}  // extern "C"
