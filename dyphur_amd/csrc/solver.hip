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
//    decreasing-priority chain ending at k.  Both are computed from per-body
//    contact lists sorted by priority: a contact's higher-priority neighbours
//    are the contacts ahead of it on its bodies' lists, so round r colours
//    the contacts that head all their lists, one per body, reading and
//    extending one used-colour mask per body.  The values do not depend on
//    timing.  The reference's serial fallbacks (a contact whose neighbours
//    use all 64 colours, or more than 256 rounds) are detected on the way.
//  * Reference-order path (<= 256 contacts or force_serial): run_serial
//    applies contacts 0..nc-1 in order.  Contacts only interact through shared
//    dynamic bodies, so class = dependency level (1 + the level of the latest
//    earlier contact on either dynamic body) reproduces the serial result bit
//    for bit while running each level in parallel.
//
// Data layout during a solve: bodies are gathered into 64 B dynamic + 32 B
// static AoS records (one or two 128-bit loads per field group instead of 20
// scattered SoA loads), contacts into class-sorted 40 B records.  All I x K
// phases then run in one cooperative persistent kernel separated by a grid
// barrier; the minimum number of dependent phases is I x (max contacts on one
// body), so the per-phase cost is what matters.
//
// Quaternion renormalisation after every iteration is applied lazily: a
// dynamic body is brought to `it` renormalisations when a contact of
// iteration `it` first touches it, and every body to n_iters at the end.
// Static bodies are never written by contacts and their orientation only
// enters products with a zero inverse inertia, so they are renormalised at
// the end.  Every body ends with the reference's sequence of operations.
#include <cstdlib>
#include <vector>

#include "dy_internal.hpp"

namespace dy {

constexpr uint64_t kUnassigned = ~0ull;
constexpr uint32_t kSolveThreads = 256;
// meta: [0] remaining [1] colour overflow [2] max class [3] max depth [4] error flags [5] assigned
constexpr uint32_t kMetaWords = 16;
constexpr uint32_t kLdsClasses = 2048;   // block-aggregated class counting limit
constexpr uint32_t kErrBarrier = 2u;

__device__ __forceinline__ uint32_t cls_of(uint64_t s) { return static_cast<uint32_t>(s & 0xffffffffu); }
__device__ __forceinline__ uint32_t dep_of(uint64_t s) { return static_cast<uint32_t>(s >> 32); }
__device__ __forceinline__ uint64_t ld_state(const uint64_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void solve_prep(uint32_t nb, uint32_t nc, uint32_t *__restrict__ deg, uint32_t *__restrict__ fill,
                           uint64_t *__restrict__ used, float *__restrict__ lam, uint64_t *__restrict__ state,
                           uint32_t *__restrict__ meta) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nb) { deg[i] = 0; fill[i] = 0; used[i] = 0; }
    if (i < nc) { lam[i] = 0.f; state[i] = kUnassigned; }
    if (i < kMetaWords) meta[i] = (i == 0) ? nc : 0u;
}

__global__ void solve_reassign_prep(uint32_t nc, uint64_t *__restrict__ state, uint32_t *__restrict__ meta) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nc) state[i] = kUnassigned;
    if (i < 6 && i != 4) meta[i] = (i == 0) ? nc : 0u;  // keeps error flags and the endpoint total
}

__global__ void solve_degree(dy_contact_view cv, uint32_t nc, const float *__restrict__ im, uint32_t *__restrict__ deg) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const uint32_t a = cv.body_a[c], b = cv.body_b[c];
    if (im[a] != 0.f) atomicAdd(&deg[a], 1u);
    if (im[b] != 0.f) atomicAdd(&deg[b], 1u);
}

// Unordered per-body lists of dynamic endpoints: adj = (c << 1 | end), own = body.
__global__ void solve_fill(dy_contact_view cv, uint32_t nc, const float *__restrict__ im,
                           const uint32_t *__restrict__ off, uint32_t *__restrict__ fill, uint32_t *__restrict__ adj,
                           uint32_t *__restrict__ own) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const uint32_t a = cv.body_a[c], b = cv.body_b[c];
    if (im[a] != 0.f) {
        const uint32_t p = off[a] + atomicAdd(&fill[a], 1u);
        adj[p] = c << 1;
        own[p] = a;
    }
    if (im[b] != 0.f) {
        const uint32_t p = off[b] + atomicAdd(&fill[b], 1u);
        adj[p] = (c << 1) | 1u;
        own[p] = b;
    }
}

// Contact order used by the assignment: higher key goes first.  JP priority
// (hash(k) << 32 | k) on the coloured path; lower index first on the
// reference-order path.
__device__ __forceinline__ uint64_t order_key(uint32_t c, int serial) {
    return serial ? static_cast<uint64_t>(0xffffffffu - c) : jp_priority(c);
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        const uint32_t o = __shfl_xor(v, d, 64);
        v = o > v ? o : v;
    }
    return v;
}
__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

// Wave-aggregated append of up to N values per lane: one atomic per wave.
// Must be reached by every lane of the wave.
template <int N>
__device__ __forceinline__ void wave_append(const bool (&pred)[N], const uint32_t (&val)[N], uint32_t *out,
                                            uint32_t *count) {
    uint64_t m[N];
    uint32_t pre[N], total = 0;
#pragma unroll
    for (int t = 0; t < N; ++t) {
        m[t] = __ballot(pred[t]);
        pre[t] = total;
        total += static_cast<uint32_t>(__popcll(m[t]));
    }
    if (total == 0) return;
    const uint32_t lane = lane_id();
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(count, total);
    base = __shfl(base, 0, 64);
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
    for (int t = 0; t < N; ++t)
        if (pred[t]) out[base + pre[t] + static_cast<uint32_t>(__popcll(m[t] & lt))] = val[t];
}

// Per-body contact lists.  Every dynamic endpoint (contact c, end e) has an
// entry adj[p] = (c << 1 | e) in its body's segment [off[x], off[x] + deg[x]),
// and own[p] = x.  list_rank places c at its rank in the body's order
// (slot[off[x] + rank]) and records the rank per endpoint (crank[2c + e]).
// Cost per entry: one pass over the body's segment (L2-resident).
__global__ __launch_bounds__(kSolveThreads) void list_rank(const uint32_t *__restrict__ adj,
                                                           const uint32_t *__restrict__ own,
                                                           const uint32_t *__restrict__ off,
                                                           const uint32_t *__restrict__ deg,
                                                           const uint32_t *__restrict__ total, int serial,
                                                           uint32_t *__restrict__ slot, uint32_t *__restrict__ crank) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= *total) return;
    const uint32_t v = adj[p], x = own[p];
    const uint32_t c = v >> 1;
    const uint64_t kc = order_key(c, serial);
    const uint32_t beg = off[x], n = deg[x];
    const uint32_t self = p - beg;
    uint32_t rank = 0;
    for (uint32_t q = 0; q < n; ++q) {
        const uint64_t kq = order_key(adj[beg + q] >> 1, serial);
        rank += (kq > kc || (kq == kc && q < self)) ? 1u : 0u;  // ties (a == b) by position
    }
    slot[beg + rank] = c;
    crank[2u * c + (v & 1u)] = rank;
}

constexpr uint32_t kNone = 0xffffffffu;

// Links each contact to its successor in both endpoint lists and counts the
// lists it does not head yet (pending).  Contacts heading every list of a
// dynamic endpoint form the first frontier.  lrec[c] = (a or none, b or none,
// next on a, next on b); static endpoints are "none".
__global__ __launch_bounds__(kSolveThreads) void list_link(dy_contact_view cv, uint32_t nc,
                                                           const float *__restrict__ im,
                                                           const uint32_t *__restrict__ off,
                                                           const uint32_t *__restrict__ deg,
                                                           const uint32_t *__restrict__ slot,
                                                           const uint32_t *__restrict__ crank,
                                                           uint4 *__restrict__ lrec, uint32_t *__restrict__ pending,
                                                           uint32_t *__restrict__ frontier, uint32_t *fcount) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    bool ready = false;
    if (c < nc) {
        uint32_t ends[2] = {cv.body_a[c], cv.body_b[c]};
        uint32_t next[2] = {kNone, kNone};
        uint32_t wait = 0;
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            const uint32_t x = ends[e];
            if (im[x] == 0.f) {
                ends[e] = kNone;
                continue;
            }
            const uint32_t r = crank[2u * c + e];
            if (r + 1u < deg[x]) next[e] = slot[off[x] + r + 1u];
            wait += r != 0u ? 1u : 0u;
        }
        lrec[c] = make_uint4(ends[0], ends[1], next[0], next[1]);
        pending[c] = wait;
        ready = wait == 0u;
    }
    const bool pred[1] = {ready};
    const uint32_t val[1] = {c};
    wave_append<1>(pred, val, frontier, fcount);
}

// JP rounds over frontiers in one cooperative launch.  Round r assigns the
// contacts that head the ordered lists of all their dynamic endpoints: on a
// body's list every earlier contact is a higher-priority neighbour, so these
// are exactly the contacts whose higher-priority neighbours were all assigned
// before round r — the ones the reference colours in its round r.  A body
// heads at most one contact per round, so the contact owns its bodies' masks
// of used colours during the round: class = lowest colour free in
// used[a] | used[b] (round - 1 on the reference-order path), then both masks
// gain the class and both successors move one list closer to ready.  Values
// written in a round are read only after the barrier that ends it.
// fcnt[0..2] rotate as frontier sizes (round r appends to fcnt[r % 3]).
// meta[1] colour overflow, meta[2] max class, meta[3] rounds, meta[5] assigned.
__global__ __launch_bounds__(1024) void assign_lists(const uint4 *__restrict__ lrec, int serial, uint32_t max_rounds,
                                                     uint32_t *pending, uint64_t *used, uint64_t *state, uint32_t *fr0,
                                                     uint32_t *fr1, uint32_t *fcnt, uint32_t *meta, uint32_t *bar,
                                                     uint32_t *err) {
    uint32_t gen = 0;
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x, gstride = gridDim.x * blockDim.x;
    const uint32_t wave0 = gtid & ~63u, lane = lane_id();
    uint32_t *cur = fr0, *nxt = fr1;
    uint32_t ncur = fcnt[0];
    uint32_t my_max = 0, my_done = 0, round = 1;
    bool exceeded = false;
    for (; ncur > 0; ++round) {
        if (round > max_rounds) {  // deeper than the reference allows
            exceeded = true;
            break;
        }
        uint32_t *nxt_count = &fcnt[round % 3];
        if (gtid == 0) fcnt[(round + 1) % 3] = 0u;
        for (uint32_t i0 = wave0; i0 < ncur; i0 += gstride) {
            const uint32_t i = i0 + lane;
            bool ready[2] = {false, false};
            uint32_t nx[2] = {0u, 0u};
            if (i < ncur) {
                const uint32_t c = cur[i];
                const uint4 r = lrec[c];
                uint32_t cls;
                if (serial) {
                    cls = round - 1u;
                } else {
                    const uint64_t ua = r.x != kNone ? used[r.x] : 0ull;
                    const uint64_t ub = r.y != kNone ? used[r.y] : 0ull;
                    const uint64_t u = ua | ub;
                    if (u == ~0ull) {
                        atomicOr(&meta[1], 1u);
                        cls = 64u;
                    } else {
                        cls = static_cast<uint32_t>(__builtin_ctzll(~u));
                    }
                    const uint64_t bit = 1ull << (cls & 63u);
                    if (r.x != kNone) used[r.x] = ua | bit;
                    if (r.y != kNone) used[r.y] = ub | bit;
                }
                state[c] = (static_cast<uint64_t>(round) << 32) | cls;
                my_max = cls > my_max ? cls : my_max;
                ++my_done;
                nx[0] = r.z;
                nx[1] = r.w;
                const uint32_t oa = r.z != kNone ? atomicSub(&pending[r.z], 1u) : 0u;
                const uint32_t ob = r.w != kNone ? atomicSub(&pending[r.w], 1u) : 0u;
                ready[0] = r.z != kNone && oa == 1u;
                ready[1] = r.w != kNone && ob == 1u;
            }
            wave_append<2>(ready, nx, nxt, nxt_count);
        }
        grid_barrier(bar, gen, err, kErrBarrier);
        ncur = *nxt_count;
        uint32_t *t = cur; cur = nxt; nxt = t;
    }
    my_max = wave_max_u32(my_max);
    my_done = wave_sum_u32(my_done);
    if (lane == 0) {
        atomicMax(&meta[2], my_max);
        atomicAdd(&meta[5], my_done);
    }
    if (gtid == 0) meta[3] = exceeded ? max_rounds + 1u : round - 1u;
}

// One round per launch (used when a cooperative launch is unavailable or
// disabled): the frontier of round r is read from fr[r & 1] with its size in
// fcnt[(r - 1) % 3], the next one written to fr[(r + 1) & 1].
__global__ __launch_bounds__(kSolveThreads) void assign_lists_round(const uint4 *__restrict__ lrec, int serial,
                                                                    uint32_t round, uint32_t max_cur,
                                                                    uint32_t *pending, uint64_t *used,
                                                                    uint64_t *state, uint32_t *fr0, uint32_t *fr1,
                                                                    uint32_t *fcnt, uint32_t *meta) {
    const uint32_t ncur = fcnt[(round - 1u) % 3u];
    uint32_t *cur = (round & 1u) ? fr0 : fr1;
    uint32_t *nxt = (round & 1u) ? fr1 : fr0;
    uint32_t *nxt_count = &fcnt[round % 3u];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0 && ncur > 0) atomicMax(&meta[3], round);
    bool ready[2] = {false, false};
    uint32_t nx[2] = {0u, 0u};
    if (i < ncur && i < max_cur) {
        const uint32_t c = cur[i];
        const uint4 r = lrec[c];
        uint32_t cls;
        if (serial) {
            cls = round - 1u;
        } else {
            const uint64_t ua = r.x != kNone ? used[r.x] : 0ull;
            const uint64_t ub = r.y != kNone ? used[r.y] : 0ull;
            const uint64_t u = ua | ub;
            if (u == ~0ull) {
                atomicOr(&meta[1], 1u);
                cls = 64u;
            } else {
                cls = static_cast<uint32_t>(__builtin_ctzll(~u));
            }
            const uint64_t bit = 1ull << (cls & 63u);
            if (r.x != kNone) used[r.x] = ua | bit;
            if (r.y != kNone) used[r.y] = ub | bit;
        }
        state[c] = (static_cast<uint64_t>(round) << 32) | cls;
        atomicMax(&meta[2], cls);
        atomicAdd(&meta[5], 1u);
        nx[0] = r.z;
        nx[1] = r.w;
        const uint32_t oa = r.z != kNone ? atomicSub(&pending[r.z], 1u) : 0u;
        const uint32_t ob = r.w != kNone ? atomicSub(&pending[r.w], 1u) : 0u;
        ready[0] = r.z != kNone && oa == 1u;
        ready[1] = r.w != kNone && ob == 1u;
    }
    wave_append<2>(ready, nx, nxt, nxt_count);
}

// Clears the frontier count two rounds ahead (launched between rounds).
__global__ void clear_count(uint32_t *fcnt, uint32_t idx) { fcnt[idx] = 0u; }

// Class histogram, aggregated per workgroup in LDS when the class count fits.
__global__ __launch_bounds__(kSolveThreads) void class_count(uint32_t nc, uint32_t n_cls,
                                                             const uint64_t *__restrict__ state,
                                                             uint32_t *__restrict__ hist) {
    __shared__ uint32_t lh[kLdsClasses];
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    const bool local = n_cls <= kLdsClasses;
    if (local)
        for (uint32_t k = threadIdx.x; k < n_cls; k += blockDim.x) lh[k] = 0;
    __syncthreads();
    if (c < nc) {
        const uint32_t k = cls_of(state[c]);
        if (local) atomicAdd(&lh[k], 1u);
        else atomicAdd(&hist[k], 1u);
    }
    __syncthreads();
    if (local)
        for (uint32_t k = threadIdx.x; k < n_cls; k += blockDim.x)
            if (lh[k]) atomicAdd(&hist[k], lh[k]);
}

// Contact record for the sweeps, class-sorted: (a, b), (p, n.x), (n.y, n.z, depth, -).
struct CRec {
    uint2 *ab;
    float4 *p0, *p1;
};

__global__ __launch_bounds__(kSolveThreads) void class_gather(dy_contact_view cv, uint32_t nc, uint32_t n_cls,
                                                              const uint64_t *__restrict__ state,
                                                              const uint32_t *__restrict__ boff,
                                                              uint32_t *__restrict__ bfill, CRec rec,
                                                              uint32_t *__restrict__ rec_id) {
    __shared__ uint32_t lh[kLdsClasses];
    __shared__ uint32_t lbase[kLdsClasses];
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    const bool local = n_cls <= kLdsClasses;
    if (local)
        for (uint32_t k = threadIdx.x; k < n_cls; k += blockDim.x) lh[k] = 0;
    __syncthreads();
    uint32_t k = 0, rank = 0;
    if (c < nc) {
        k = cls_of(state[c]);
        rank = local ? atomicAdd(&lh[k], 1u) : atomicAdd(&bfill[k], 1u);
    }
    __syncthreads();
    if (local)
        for (uint32_t q = threadIdx.x; q < n_cls; q += blockDim.x)
            lbase[q] = lh[q] ? atomicAdd(&bfill[q], lh[q]) : 0u;
    __syncthreads();
    if (c >= nc) return;
    const uint32_t pos = boff[k] + (local ? lbase[k] + rank : rank);
    rec.ab[pos] = make_uint2(cv.body_a[c], cv.body_b[c]);
    rec.p0[pos] = make_float4(cv.pos_x[c], cv.pos_y[c], cv.pos_z[c], cv.norm_x[c]);
    rec.p1[pos] = make_float4(cv.norm_y[c], cv.norm_z[c], cv.depth[c], 0.f);
    rec_id[pos] = c;
}

// AoS body records: dyn = (p.xyz, q.w) (q.xyz, v.x) (v.yz, w.xy) (w.z, renorms, -, -);
// sta = (inv_mass, Ixx, Iyy, Izz) (Ixy, Ixz, Iyz, -).
__global__ void gather_bodies(dy_body_view bv, float4 *__restrict__ dyn, float4 *__restrict__ sta) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= bv.n) return;
    dyn[4 * i + 0] = make_float4(bv.pos_x[i], bv.pos_y[i], bv.pos_z[i], bv.rot_w[i]);
    dyn[4 * i + 1] = make_float4(bv.rot_x[i], bv.rot_y[i], bv.rot_z[i], bv.vel_x[i]);
    dyn[4 * i + 2] = make_float4(bv.vel_y[i], bv.vel_z[i], bv.ang_x[i], bv.ang_y[i]);
    dyn[4 * i + 3] = make_float4(bv.ang_z[i], __uint_as_float(0u), 0.f, 0.f);
    sta[2 * i + 0] = make_float4(bv.inv_mass[i], bv.iI_xx[i], bv.iI_yy[i], bv.iI_zz[i]);
    sta[2 * i + 1] = make_float4(bv.iI_xy[i], bv.iI_xz[i], bv.iI_yz[i], 0.f);
}

__device__ __forceinline__ BodyState load_aos(const float4 *__restrict__ dyn, const float4 *__restrict__ sta,
                                              uint32_t i, uint32_t &renorms) {
    const float4 d0 = dyn[4 * i], d1 = dyn[4 * i + 1], d2 = dyn[4 * i + 2], d3 = dyn[4 * i + 3];
    const float4 s0 = sta[2 * i], s1 = sta[2 * i + 1];
    BodyState s;
    s.px = d0.x; s.py = d0.y; s.pz = d0.z; s.qw = d0.w;
    s.qx = d1.x; s.qy = d1.y; s.qz = d1.z; s.vx = d1.w;
    s.vy = d2.x; s.vz = d2.y; s.wx = d2.z; s.wy = d2.w;
    s.wz = d3.x;
    renorms = __float_as_uint(d3.y);
    s.im = s0.x; s.ixx = s0.y; s.iyy = s0.z; s.izz = s0.w;
    s.ixy = s1.x; s.ixz = s1.y; s.iyz = s1.z;
    return s;
}

__device__ __forceinline__ void store_aos(float4 *__restrict__ dyn, uint32_t i, const BodyState &s, uint32_t renorms) {
    dyn[4 * i] = make_float4(s.px, s.py, s.pz, s.qw);
    dyn[4 * i + 1] = make_float4(s.qx, s.qy, s.qz, s.vx);
    dyn[4 * i + 2] = make_float4(s.vy, s.vz, s.wx, s.wy);
    dyn[4 * i + 3] = make_float4(s.wz, __uint_as_float(renorms), 0.f, 0.f);
}

__device__ __forceinline__ void catch_up(BodyState &s, uint32_t done, uint32_t target) {
    for (uint32_t r = done; r < target; ++r) renormalise(s.qw, s.qx, s.qy, s.qz);
}

__device__ __forceinline__ void solve_one(uint32_t i, uint32_t it, const CRec &rec, float *__restrict__ lam,
                                          float4 *__restrict__ dyn, const float4 *__restrict__ sta, float mu,
                                          float dt) {
    const uint2 ab = rec.ab[i];
    const float4 p0 = rec.p0[i], p1 = rec.p1[i];
    const ContactIn ci{p0.x, p0.y, p0.z, p0.w, p1.x, p1.y, p1.z};
    uint32_t ra, rb;
    BodyState A = load_aos(dyn, sta, ab.x, ra), B = load_aos(dyn, sta, ab.y, rb);
    const bool da = A.im != 0.f, db = B.im != 0.f;
    if (da) catch_up(A, ra, it);
    if (db) catch_up(B, rb, it);
    float l = lam[i];
    project_contact(ci, A, B, mu, l, dt);
    lam[i] = l;
    if (da) store_aos(dyn, ab.x, A, it);
    if (db) store_aos(dyn, ab.y, B, it);
}

// One class (fallback path: one launch per class).
__global__ __launch_bounds__(kSolveThreads) void sweep_class(const uint32_t *__restrict__ boff, uint32_t k, uint32_t it,
                                                             CRec rec, float *__restrict__ lam, float4 *__restrict__ dyn,
                                                             const float4 *__restrict__ sta, float mu, float dt) {
    const uint32_t beg = boff[k], end = boff[k + 1];
    for (uint32_t i = beg + blockIdx.x * blockDim.x + threadIdx.x; i < end; i += gridDim.x * blockDim.x)
        solve_one(i, it, rec, lam, dyn, sta, mu, dt);
}

__global__ __launch_bounds__(1024) void sweep_persistent(const uint32_t *__restrict__ boff, uint32_t n_cls,
                                                                  uint32_t n_iters, CRec rec, float *__restrict__ lam,
                                                                  float4 *__restrict__ dyn,
                                                                  const float4 *__restrict__ sta, float mu, float dt,
                                                                  uint32_t *bar, uint32_t *err) {
    uint32_t gen = 0;
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x, gstride = gridDim.x * blockDim.x;
    for (uint32_t it = 0; it < n_iters; ++it) {
        for (uint32_t k = 0; k < n_cls; ++k) {
            const uint32_t beg = boff[k], end = boff[k + 1];
            for (uint32_t i = beg + gtid; i < end; i += gstride) solve_one(i, it, rec, lam, dyn, sta, mu, dt);
            if (it + 1 < n_iters || k + 1 < n_cls) grid_barrier(bar, gen, err, kErrBarrier);
        }
    }
}

// Back to the SoA view, bringing every body to n_iters renormalisations.
__global__ void scatter_bodies(dy_body_view bv, const float4 *__restrict__ dyn, uint32_t n_iters) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= bv.n) return;
    const float4 d0 = dyn[4 * i], d1 = dyn[4 * i + 1], d2 = dyn[4 * i + 2], d3 = dyn[4 * i + 3];
    float qw = d0.w, qx = d1.x, qy = d1.y, qz = d1.z;
    for (uint32_t r = __float_as_uint(d3.y); r < n_iters; ++r) renormalise(qw, qx, qy, qz);
    bv.pos_x[i] = d0.x; bv.pos_y[i] = d0.y; bv.pos_z[i] = d0.z;
    bv.rot_w[i] = qw; bv.rot_x[i] = qx; bv.rot_y[i] = qy; bv.rot_z[i] = qz;
    bv.vel_x[i] = d1.w; bv.vel_y[i] = d2.x; bv.vel_z[i] = d2.y;
    bv.ang_x[i] = d2.z; bv.ang_y[i] = d2.w; bv.ang_z[i] = d3.x;
}

// Class of each contact id (for dy_xpbd_download_classes).
__global__ void unpack_classes(uint32_t n, const uint64_t *__restrict__ state, uint32_t *__restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = cls_of(state[i]);
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
    uint32_t *adj = nullptr, *own = nullptr, *slot = nullptr, *crank = nullptr;
    uint32_t *boff = nullptr, *bfill = nullptr, *hist = nullptr, *rec_id = nullptr;
    uint4 *lrec = nullptr;
    uint32_t *meta = nullptr;  // kMetaWords + barrier words
    uint2 *rec_ab = nullptr;
    float4 *rec_p0 = nullptr, *rec_p1 = nullptr;
    DevBuf<uint32_t> deg, off, fill, scan_tmp, pending, fr0, fr1;
    DevBuf<uint64_t> used;
    DevBuf<float4> dyn, sta;
    dy_solve_stats stats{};
    uint32_t h_meta[kMetaWords + 8] = {};
    int coop_assign_max = 0, coop_sweep_max = 0, device_cus = 0;
    int sweep_blocks_override = 0, assign_blocks_override = 0;
    int assign_threads = 256, sweep_threads = 256;
    int path = 0;  // 0 persistent sweep, 1 per-class launches
    // Kernel timing (dy_xpbd_set_timing): event pairs recorded around the
    // assignment and sweep kernels, read back once the next solve syncs.
    bool timing = false;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    bool ev_pending = false, ev_assign = false;
    double pending_bytes = 0.0;
    dy_solver_timing acc{};
};

namespace {
// Adds the event intervals of the previous solve to the totals (the events
// have completed once the stream reached a later sync point).
int collect_timing(dy_xpbd_s *x) {
    if (!x->ev_pending) return DY_OK;
    float a = 0.f, b = 0.f;
    DY_TRY(hipEventSynchronize(x->ev[3]));
    if (x->ev_assign) DY_TRY(hipEventElapsedTime(&a, x->ev[0], x->ev[1]));
    DY_TRY(hipEventElapsedTime(&b, x->ev[2], x->ev[3]));
    x->ev_assign = false;
    x->acc.solves += 1;
    x->acc.assign_ms += a;
    x->acc.sweep_ms += b;
    x->acc.sweep_bytes += x->pending_bytes;
    x->ev_pending = false;
    return DY_OK;
}

void xp_free(dy_xpbd_s *x) {
    for (hipEvent_t &e : x->ev)
        if (e) { (void)hipEventDestroy(e); e = nullptr; }
    void *ptrs[] = {x->lam, x->state, x->adj, x->own, x->slot, x->crank, x->lrec, x->boff, x->bfill, x->hist,
                    x->rec_id, x->meta, x->rec_ab, x->rec_p0, x->rec_p1};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    x->deg.release(); x->off.release(); x->fill.release(); x->scan_tmp.release();
    x->pending.release(); x->fr0.release(); x->fr1.release(); x->used.release();
    x->dyn.release(); x->sta.release();
}

int env_int(const char *name, int dflt) {
    const char *e = std::getenv(name);
    return e ? std::atoi(e) : dflt;
}

// Cooperative-launch limits; DYPHUR_{ASSIGN,SWEEP}_{BLOCKS,THREADS} and
// DYPHUR_SOLVER_PATH (1 = one launch per round / per class) override the
// launch shapes for measurement.
int query_coop(dy_xpbd_s *x) {
    int dev = 0, coop = 0;
    DY_TRY(hipGetDevice(&dev));
    DY_TRY(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
    DY_TRY(hipDeviceGetAttribute(&x->device_cus, hipDeviceAttributeMultiprocessorCount, dev));
    x->sweep_blocks_override = env_int("DYPHUR_SWEEP_BLOCKS", 0);
    x->assign_blocks_override = env_int("DYPHUR_ASSIGN_BLOCKS", 0);
    x->path = env_int("DYPHUR_SOLVER_PATH", 0);
    const int at = env_int("DYPHUR_ASSIGN_THREADS", 256), sw = env_int("DYPHUR_SWEEP_THREADS", 256);
    x->assign_threads = (at == 64 || at == 128 || at == 256 || at == 512 || at == 1024) ? at : 256;
    x->sweep_threads = (sw == 64 || sw == 128 || sw == 256 || sw == 512 || sw == 1024) ? sw : 256;
    if (!coop) return DY_OK;
    int per_cu = 0;
    DY_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(assign_lists),
                                                        x->assign_threads, 0));
    x->coop_assign_max = per_cu * x->device_cus;
    DY_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(sweep_persistent),
                                                        x->sweep_threads, 0));
    x->coop_sweep_max = per_cu * x->device_cus;
    return DY_OK;
}

// Blocks for a cooperative kernel: the override, else device_cus / 8 workgroups
// of 256 threads' worth, clamped to what can be co-resident.
uint32_t coop_blocks(const dy_xpbd_s *x, int override_blocks, int threads, int max_blocks) {
    int b = override_blocks > 0 ? override_blocks : (x->device_cus / 8) * 256 / threads;
    if (b < 1) b = 1;
    if (b > max_blocks) b = max_blocks;
    return static_cast<uint32_t>(b);
}

// Assigns a class to every contact; leaves meta downloaded in x->h_meta.
int assign_classes(dy_xpbd_s *x, hipStream_t st, const dy_contact_view &cv, uint32_t nc, const float *im, int serial) {
    const uint32_t gc = ceil_div(nc, kSolveThreads);
    uint32_t *fcnt = x->meta + kMetaWords + 4;
    DY_TRY(hipMemsetAsync(fcnt, 0, 3 * sizeof(uint32_t), st));
    list_rank<<<ceil_div(2u * nc, kSolveThreads), kSolveThreads, 0, st>>>(x->adj, x->own, x->off.p, x->deg.p,
                                                                          x->meta + 6, serial, x->slot, x->crank);
    DY_LAUNCH_CHECK();
    list_link<<<gc, kSolveThreads, 0, st>>>(cv, nc, im, x->off.p, x->deg.p, x->slot, x->crank, x->lrec,
                                             x->pending.p, x->fr0.p, fcnt);
    DY_LAUNCH_CHECK();
    const uint32_t max_rounds = serial ? nc + 1u : 256u;
    if (x->coop_assign_max > 0 && x->path == 0) {
        uint32_t *bar = x->meta + kMetaWords;
        DY_TRY(hipMemsetAsync(bar, 0, 2 * sizeof(uint32_t), st));
        const uint32_t blocks = coop_blocks(x, x->assign_blocks_override, x->assign_threads, x->coop_assign_max);
        const uint4 *lrecv = x->lrec;
        int serialv = serial;
        uint32_t maxr = max_rounds;
        uint32_t *pendv = x->pending.p, *f0 = x->fr0.p, *f1 = x->fr1.p, *metav = x->meta;
        uint64_t *usedv = x->used.p, *statev = x->state;
        uint32_t *errv = x->meta + kMetaWords + 2;
        void *args[] = {&lrecv, &serialv, &maxr, &pendv, &usedv, &statev, &f0, &f1, &fcnt, &metav, &bar, &errv};
        if (x->timing) DY_TRY(hipEventRecord(x->ev[0], st));
        DY_TRY(hipLaunchCooperativeKernel(reinterpret_cast<const void *>(assign_lists), dim3(blocks),
                                          dim3(static_cast<uint32_t>(x->assign_threads)), args, 0, st));
        if (x->timing) {
            DY_TRY(hipEventRecord(x->ev[1], st));
            x->ev_assign = true;
        }
        DY_TRY(hipMemcpyAsync(x->h_meta, x->meta, sizeof(x->h_meta), hipMemcpyDeviceToHost, st));
        DY_TRY(hipStreamSynchronize(st));
        if (x->h_meta[4] || x->h_meta[kMetaWords + 2]) return DY_ERR_INTERNAL;
        if (x->h_meta[3] > max_rounds) {
            if (serial) return DY_ERR_INTERNAL;
            x->h_meta[3] = 257u;  // the reference would stop after 256 rounds
            return DY_OK;
        }
        return x->h_meta[5] == nc ? DY_OK : DY_ERR_INTERNAL;
    }
    // One launch per round, checked every 16 rounds.
    for (uint32_t r = 1; r <= max_rounds; ++r) {
        clear_count<<<1, 1, 0, st>>>(fcnt, r % 3u);
        DY_LAUNCH_CHECK();
        assign_lists_round<<<gc, kSolveThreads, 0, st>>>(x->lrec, serial, r, nc, x->pending.p, x->used.p, x->state,
                                                          x->fr0.p, x->fr1.p, fcnt, x->meta);
        DY_LAUNCH_CHECK();
        if (r % 16u == 0u || r == max_rounds) {
            DY_TRY(hipMemcpyAsync(x->h_meta, x->meta, sizeof(x->h_meta), hipMemcpyDeviceToHost, st));
            DY_TRY(hipStreamSynchronize(st));
            if (x->h_meta[4]) return DY_ERR_INTERNAL;
            if (x->h_meta[5] == nc) return DY_OK;
        }
    }
    if (serial) return DY_ERR_INTERNAL;
    x->h_meta[3] = 257u;
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
    const size_t ncls = c + 66;
    hipError_t e = hipMalloc(reinterpret_cast<void **>(&x->lam), c * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&x->state), c * sizeof(uint64_t));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&x->adj), 2 * c * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&x->own), 2 * c * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&x->slot), 2 * c * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&x->crank), 2 * c * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&x->lrec), c * sizeof(uint4));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&x->boff), ncls * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&x->bfill), ncls * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&x->hist), ncls * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&x->rec_id), c * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&x->meta), (kMetaWords + 8) * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&x->rec_ab), c * sizeof(uint2));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&x->rec_p0), c * sizeof(float4));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&x->rec_p1), c * sizeof(float4));
    if (e == hipSuccess) e = hipMemsetAsync(x->meta, 0, (kMetaWords + 8) * sizeof(uint32_t), hs(s));
    if (e != hipSuccess) {
        xp_free(x);
        delete x;
        return static_cast<int>(e);
    }
    const int q = query_coop(x);
    if (q != DY_OK) {
        xp_free(x);
        delete x;
        return q;
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
    uint32_t head[2] = {0, 0};  // contact count; sweep error flags of the previous solve
    DY_TRY(hipMemcpyAsync(&head[0], cv->n, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    DY_TRY(hipMemcpyAsync(&head[1], x->meta + kMetaWords + 2, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    DY_TRY(hipStreamSynchronize(st));
    if (head[1]) return DY_ERR_INTERNAL;
    const uint32_t nc = head[0];
    x->stats.n_constraints = nc;
    if (nc == 0) return DY_OK;
    if (nc > x->cap) return DY_ERR_INVALID_ARG;
    const uint32_t nb = bv->n;
    DY_TRY_STATUS(x->deg.ensure(nb));
    DY_TRY_STATUS(x->off.ensure(nb));
    DY_TRY_STATUS(x->fill.ensure(nb));
    DY_TRY_STATUS(x->used.ensure(nb));
    DY_TRY_STATUS(x->pending.ensure(nc));
    DY_TRY_STATUS(x->fr0.ensure(nc));
    DY_TRY_STATUS(x->fr1.ensure(nc));
    DY_TRY_STATUS(x->dyn.ensure(4 * static_cast<size_t>(nb)));
    DY_TRY_STATUS(x->sta.ensure(2 * static_cast<size_t>(nb)));
    DY_TRY_STATUS(x->scan_tmp.ensure(scan_tmp_words(nb > nc + 66u ? nb : nc + 66u)));
    const uint32_t prep_n = nb > nc ? nb : nc;
    solve_prep<<<ceil_div(prep_n, 256), 256, 0, st>>>(nb, nc, x->deg.p, x->fill.p, x->used.p, x->lam, x->state,
                                                      x->meta);
    DY_LAUNCH_CHECK();
    const uint32_t gc = ceil_div(nc, kSolveThreads);
    solve_degree<<<gc, kSolveThreads, 0, st>>>(*cv, nc, bv->inv_mass, x->deg.p);
    DY_LAUNCH_CHECK();
    DY_TRY_STATUS(scan_exclusive(st, x->deg.p, x->off.p, nb, x->meta + 6, x->scan_tmp.p));
    solve_fill<<<gc, kSolveThreads, 0, st>>>(*cv, nc, bv->inv_mass, x->off.p, x->fill.p, x->adj, x->own);
    DY_LAUNCH_CHECK();
    gather_bodies<<<ceil_div(nb, 256), 256, 0, st>>>(*bv, x->dyn.p, x->sta.p);
    DY_LAUNCH_CHECK();

    int serial = (x->force_serial || nc <= 256u) ? 1 : 0;
    DY_TRY_STATUS(collect_timing(x));
    DY_TRY_STATUS(assign_classes(x, st, *cv, nc, bv->inv_mass, serial));
    if (!serial) {
        x->stats.overflow = x->h_meta[1];
        x->stats.rounds = x->h_meta[3];
        if (x->h_meta[1] || x->h_meta[3] > 256u) {
            // The reference runs its serial path in these cases.
            serial = 1;
            solve_reassign_prep<<<ceil_div(nc, 256), 256, 0, st>>>(nc, x->state, x->meta);
            DY_LAUNCH_CHECK();
            DY_TRY_STATUS(assign_classes(x, st, *cv, nc, bv->inv_mass, 1));
        }
    }
    const uint32_t k = x->h_meta[2] + 1u;
    x->stats.used_serial = static_cast<uint32_t>(serial);
    x->stats.n_classes = k;
    if (serial) x->stats.rounds = x->h_meta[3];

    // Class-sorted contact records.
    DY_TRY(hipMemsetAsync(x->hist, 0, (k + 1) * sizeof(uint32_t), st));
    DY_TRY(hipMemsetAsync(x->bfill, 0, k * sizeof(uint32_t), st));
    class_count<<<gc, kSolveThreads, 0, st>>>(nc, k, x->state, x->hist);
    DY_LAUNCH_CHECK();
    DY_TRY_STATUS(scan_exclusive(st, x->hist, x->boff, k + 1, nullptr, x->scan_tmp.p));
    CRec rec{x->rec_ab, x->rec_p0, x->rec_p1};
    class_gather<<<gc, kSolveThreads, 0, st>>>(*cv, nc, k, x->state, x->boff, x->bfill, rec, x->rec_id);
    DY_LAUNCH_CHECK();

    const uint32_t iters = static_cast<uint32_t>(x->n_iters);
    x->stats.dyn_endpoints = x->h_meta[6];
    if (x->timing) {
        // Algorithmic bytes of the sweep: per contact per iteration 36 B
        // record + 8 B lambda, and per dynamic endpoint 80 B read + 52 B
        // written; plus the per-iteration renormalisation (32 B per body).
        x->pending_bytes = static_cast<double>(iters) *
                           (44.0 * nc + 132.0 * x->h_meta[6] + 32.0 * static_cast<double>(nb));
        DY_TRY(hipEventRecord(x->ev[2], st));
    }
    if (iters > 0) {
        const bool persistent = x->path == 0 && x->coop_sweep_max > 0;
        if (persistent) {
            const uint32_t blocks = coop_blocks(x, x->sweep_blocks_override, x->sweep_threads, x->coop_sweep_max);
            uint32_t *bar = x->meta + kMetaWords;
            DY_TRY(hipMemsetAsync(bar, 0, 2 * sizeof(uint32_t), st));
            const uint32_t *boffv = x->boff;
            uint32_t kv = k, iv = iters;
            float *lamv = x->lam;
            float4 *dynv = x->dyn.p;
            const float4 *stav = x->sta.p;
            float muv = x->mu, dtv = dt;
            uint32_t *errv = x->meta + kMetaWords + 2;
            void *args[] = {&boffv, &kv, &iv, &rec, &lamv, &dynv, &stav, &muv, &dtv, &bar, &errv};
            DY_TRY(hipLaunchCooperativeKernel(reinterpret_cast<const void *>(sweep_persistent), dim3(blocks),
                                              dim3(static_cast<uint32_t>(x->sweep_threads)), args, 0, st));
        } else {
            // Per-class launches; each reads its range from the device offsets.
            const uint32_t blocks = gc < 1024u ? gc : 1024u;
            for (uint32_t it = 0; it < iters; ++it)
                for (uint32_t c = 0; c < k; ++c) {
                    sweep_class<<<blocks, kSolveThreads, 0, st>>>(x->boff, c, it, rec, x->lam, x->dyn.p, x->sta.p,
                                                                  x->mu, dt);
                    DY_LAUNCH_CHECK();
                }
        }
    }
    if (x->timing) {
        DY_TRY(hipEventRecord(x->ev[3], st));
        x->ev_pending = true;
    }
    scatter_bodies<<<ceil_div(nb, 256), 256, 0, st>>>(*bv, x->dyn.p, iters);
    DY_LAUNCH_CHECK();
    return DY_OK;
}

int dy_xpbd_set_timing(dy_xpbd_t x, int enable) {
    if (!x) return DY_ERR_INVALID_ARG;
    if (enable && !x->ev[0])
        for (hipEvent_t &e : x->ev) DY_TRY(hipEventCreate(&e));
    x->timing = enable != 0;
    x->ev_pending = false;
    x->acc = dy_solver_timing{};
    return DY_OK;
}

int dy_xpbd_get_timing(dy_xpbd_t x, dy_stream_t s, dy_solver_timing *out) {
    if (!x || !out) return DY_ERR_INVALID_ARG;
    DY_TRY(hipStreamSynchronize(hs(s)));
    DY_TRY_STATUS(collect_timing(x));
    *out = x->acc;
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
    if (n == 0) return DY_OK;
    unpack_classes<<<ceil_div(n, 256), 256, 0, hs(s)>>>(n, x->state, x->rec_id);
    DY_LAUNCH_CHECK();
    DY_TRY(hipMemcpyAsync(out, x->rec_id, n * sizeof(uint32_t), hipMemcpyDeviceToHost, hs(s)));
    DY_TRY(hipStreamSynchronize(hs(s)));
    return DY_OK;
}

}  // extern "C"
