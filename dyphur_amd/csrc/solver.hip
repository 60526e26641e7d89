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
// The whole solve is asynchronous on its stream on the cooperative path: the
// contact count is read on the device (clamped to capacity), the choice of
// path, the fallback to the reference order, the class count and the class
// offsets are made on the device, and the statistics come back with one
// asynchronous copy that the blocking queries wait for.
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

#include "dy_internal.hpp"
#include "joint_math.hpp"

namespace dy {

constexpr uint32_t kSolveThreads = 256;
constexpr uint32_t kLdsClasses = 2048;  // block-aggregated class counting limit
constexpr uint32_t kNone = 0xffffffffu;
constexpr uint32_t kMaxColours = 64;
constexpr uint32_t kMaxGridBlocks = 1024;  // grid-stride kernels over capacity-sized ranges
constexpr uint32_t kSweepXcc = 0;         // the XCD whose workgroups run the assignment and the sweep

// Solver metadata (device words).  solve_prep zeroes all but kMErr, which
// keeps error flags until the solver is destroyed.
enum : uint32_t {
    kMOverflow = 1,      // a contact's neighbours held all 64 colours
    kMMaxCls = 2,        // largest class
    kMRounds = 3,        // rounds of the assignment that was used
    kMErr = 4,           // error flags (sticky)
    kMAssigned = 5,      // contacts assigned
    kMDyn = 6,           // dynamic endpoints = entries of the per-body lists
    kMNcls = 7,          // classes swept
    kMSerial = 8,        // 1: reference contact order
    kMNc = 9,            // constraints solved: contacts (count clamped to capacity) + joints
    kMColourRounds = 10, // colouring rounds before a fallback (257: more than 256)
    kMNcc = 11,          // contacts solved; constraint k >= kMNcc is joint k - kMNcc
    kMNj = 12,           // joints solved
    kMJBad = 13,         // a joint of this solve failed validation
    kMLive = 14,         // a contact left live by its first iteration (sweep_flow)
    kMBarAssign = 16,    // 2 words: grid barrier of assign_lists
    kMBarSweep = 18,     // 2 words: grid barrier of the sweeps (sweep_flow: after iteration 0)
    kMFcnt = 20,         // 3 words: frontier sizes
    kMRegSweep = 24,     // 2 words: check-in of sweep_xcd (participants, arrivals)
    kMRegAssign = 26,    // 2 words: check-in of assign_lists
    kMRegMotor = 28,     // 2 words: check-in of joint_motors_xcd
    kMBarMotor = 30,     // 2 words: barrier of joint_motors_xcd
    kMDeep = 32,         // a contact's chain is longer than 256 rounds (colour_flow)
    kMQHead = 64,        // colour_flow's ready queue, one 128-byte line per counter:
    kMQTail = 96,        //   positions claimed, entries pushed (after the first
    kMQDone = 128,       //   frontier), contacts coloured
    kMWords = 160,
};
constexpr uint32_t kErrBarrier = 2u, kErrCapacity = 4u, kErrAssign = 8u, kErrFlow = 16u, kErrJoint = 32u;

__device__ __forceinline__ uint32_t cls_of(uint64_t s) { return static_cast<uint32_t>(s & 0xffffffffu); }

__device__ __forceinline__ uint32_t grid_stride() { return gridDim.x * blockDim.x; }
__device__ __forceinline__ uint32_t grid_tid() { return blockIdx.x * blockDim.x + threadIdx.x; }

// Zeroes the per-solve state and publishes the constraint counts and the
// path: the reference runs its constraints serially at <= 256 constraints
// (contacts + joints) or when forced (xpbd_solver.cpp:611-619).
// AoS body records: dyn = (p.xyz, q.w) (q.xyz, v.x) (v.yz, w.xy) (w.z, renorms, -, -);
// sta = (inv_mass, Ixx, Iyy, Izz) (Ixy, Ixz, Iyz, -).
__device__ __forceinline__ void gather_body(const dy_body_view &bv, uint32_t i, float4 *__restrict__ dyn,
                                            float4 *__restrict__ sta) {
    dyn[4 * i + 0] = make_float4(bv.pos_x[i], bv.pos_y[i], bv.pos_z[i], bv.rot_w[i]);
    dyn[4 * i + 1] = make_float4(bv.rot_x[i], bv.rot_y[i], bv.rot_z[i], bv.vel_x[i]);
    dyn[4 * i + 2] = make_float4(bv.vel_y[i], bv.vel_z[i], bv.ang_x[i], bv.ang_y[i]);
    dyn[4 * i + 3] = make_float4(bv.ang_z[i], __uint_as_float(0u), 0.f, 0.f);
    sta[2 * i + 0] = make_float4(bv.inv_mass[i], bv.iI_xx[i], bv.iI_yy[i], bv.iI_zz[i]);
    sta[2 * i + 1] = make_float4(bv.iI_xy[i], bv.iI_xz[i], bv.iI_yz[i], 0.f);
}

// Per-solve reset (per-body counters and masks, per-constraint lambdas, the
// metadata) and the bodies' AoS records.
__global__ void solve_prep(dy_body_view bv, const uint32_t *__restrict__ ncp, uint32_t cap, uint32_t nj,
                           int force_serial, uint32_t *__restrict__ deg, uint32_t *__restrict__ fill,
                           uint64_t *__restrict__ used, uint32_t *__restrict__ ver, float *__restrict__ lam,
                           uint32_t *__restrict__ hist, uint32_t *__restrict__ meta, float4 *__restrict__ dyn,
                           float4 *__restrict__ sta) {
    const uint32_t nb = bv.n;
    const uint32_t n_raw = *ncp;
    const uint32_t nc = n_raw < cap ? n_raw : cap;
    const uint32_t ncon = nc + nj;
    const uint32_t top = nb > ncon ? nb : ncon;
    for (uint32_t i = grid_tid(); i < top; i += grid_stride()) {
        if (i < nb) {
            deg[i] = 0;
            fill[i] = 0;
            used[i] = 0;
            ver[i] = 0;
            gather_body(bv, i, dyn, sta);
        }
        if (i < ncon) lam[i] = 0.f;
    }
    const uint32_t i = grid_tid();
    if (i <= kMaxColours) hist[i] = 0;
    if (i < kMWords && i != kMErr) {
        uint32_t v = 0;
        if (i == kMNc) v = ncon;
        if (i == kMNcc) v = nc;
        if (i == kMNj) v = nj;
        if (i == kMSerial) v = (force_serial || ncon <= 256u) ? 1u : 0u;
        meta[i] = v;
    }
    if (i == 0 && n_raw > cap) atomicOr(&meta[kMErr], kErrCapacity);
}

// Constraint k of a solve: contact k (k < kMNcc) or joint k - kMNcc, whose
// validated (parent, child) joint_pack wrote to jend.
struct Cons {
    dy_contact_view cv;
    const uint2 *jend;
};

__device__ __forceinline__ void cons_ends(const Cons &cs, uint32_t k, uint32_t ncc, uint32_t &a, uint32_t &b) {
    if (k < ncc) {
        a = cs.cv.body_a[k];
        b = cs.cv.body_b[k];
    } else {
        const uint2 e = cs.jend[k - ncc];
        a = e.x;
        b = e.y;
    }
}

// Joint parameters as 5 float4 per joint: (anchor_p xyz, anchor_c x)
// (anchor_c yz, axis xy) (axis z, lo, hi, target_pos) (target_vel,
// stiffness, damping, compliance_pos) (compliance_ang, type, -, -).  A joint
// with a body index out of range, parent == child or a type above 3 is
// flagged (kMJBad, kErrJoint) and the solve then runs without joints.
__global__ void joint_pack(dy_joint_view jv, uint32_t nb, float4 *__restrict__ jpk, uint2 *__restrict__ jend,
                           uint32_t *__restrict__ meta) {
    const uint32_t j = grid_tid();
    if (j >= jv.n) return;
    const uint32_t p = jv.body_parent[j], c = jv.body_child[j], t = jv.type[j];
    if (p >= nb || c >= nb || p == c || t > kJointBall) {
        atomicOr(&meta[kMJBad], 1u);
        atomicOr(&meta[kMErr], kErrJoint);
        jend[j] = make_uint2(0u, 0u);
        return;
    }
    jend[j] = make_uint2(p, c);
    jpk[5 * j + 0] = make_float4(jv.anchor_px[j], jv.anchor_py[j], jv.anchor_pz[j], jv.anchor_cx[j]);
    jpk[5 * j + 1] = make_float4(jv.anchor_cy[j], jv.anchor_cz[j], jv.axis_px[j], jv.axis_py[j]);
    jpk[5 * j + 2] = make_float4(jv.axis_pz[j], jv.limit_lo[j], jv.limit_hi[j], jv.target_pos[j]);
    jpk[5 * j + 3] = make_float4(jv.target_vel[j], jv.stiffness[j], jv.damping[j], jv.compliance_pos[j]);
    jpk[5 * j + 4] = make_float4(jv.compliance_ang[j], __uint_as_float(t), 0.f, 0.f);
}

// After a failed joint validation: the solve proceeds with the contacts only.
__global__ void joint_commit(int force_serial, uint32_t *__restrict__ meta) {
    if (threadIdx.x != 0 || meta[kMJBad] == 0u) return;
    const uint32_t nc = meta[kMNcc];
    meta[kMNj] = 0u;
    meta[kMNc] = nc;
    meta[kMSerial] = (force_serial || nc <= 256u) ? 1u : 0u;
}

__device__ __forceinline__ JointIn load_joint(const float4 *__restrict__ jpk, uint32_t j) {
    const float4 a = jpk[5 * j], b = jpk[5 * j + 1], c = jpk[5 * j + 2], d = jpk[5 * j + 3], e = jpk[5 * j + 4];
    JointIn r;
    r.apx = a.x; r.apy = a.y; r.apz = a.z; r.acx = a.w;
    r.acy = b.x; r.acz = b.y; r.lx = b.z; r.ly = b.w;
    r.lz = c.x; r.lo = c.y; r.hi = c.z; r.target_pos = c.w;
    r.target_vel = d.x; r.stiffness = d.y; r.damping = d.z; r.compliance_pos = d.w;
    r.compliance_ang = e.x;
    r.type = __float_as_uint(e.y);
    return r;
}

// Runs of equal keys over consecutive lanes (kNone: no key).  Contacts come
// grouped by their first body (the broadphase lists each body's partners in
// one row), so one atomic per run of body a replaces one per contact.
struct WaveRun {
    uint32_t start, len;  // first lane of this lane's run, its length
    bool head;            // this lane starts the run
};
__device__ __forceinline__ WaveRun wave_run(uint32_t key) {
    const uint32_t lane = lane_id();
    const uint32_t prev = __shfl_up(key, 1u, 64);
    const bool head = lane == 0 || prev != key;
    const uint64_t heads = __ballot(head);
    const uint64_t le = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);  // lanes 0..lane
    const uint32_t start = 63u - static_cast<uint32_t>(__builtin_clzll(heads & le));
    const uint64_t after = heads & ~le;
    const uint32_t end = after ? static_cast<uint32_t>(__builtin_ctzll(after)) : 64u;
    return WaveRun{start, end - start, head};
}

// One thread per constraint, every lane of a wave reaching the loop body
// (wave-uniform bounds), for the run aggregation.
__global__ void solve_degree(Cons cs, const uint32_t *__restrict__ meta, const float *__restrict__ im,
                             uint32_t *__restrict__ deg) {
    const uint32_t nc = meta[kMNc], ncc = meta[kMNcc];
    for (uint32_t c0 = grid_tid() & ~63u; c0 < nc; c0 += grid_stride()) {
        const uint32_t c = c0 + lane_id();
        uint32_t a = kNone, b = kNone;
        if (c < nc) {
            cons_ends(cs, c, ncc, a, b);
            if (im[a] == 0.f) a = kNone;
            if (im[b] == 0.f) b = kNone;
        }
        const WaveRun r = wave_run(a);
        if (r.head && a != kNone) atomicAdd(&deg[a], r.len);
        if (b != kNone) atomicAdd(&deg[b], 1u);
    }
}

// Unordered per-body lists of dynamic endpoints: adj = (c << 1 | end), own = body.
__global__ void solve_fill(Cons cs, const uint32_t *__restrict__ meta, const float *__restrict__ im,
                           const uint32_t *__restrict__ off, uint32_t *__restrict__ fill, uint32_t *__restrict__ adj,
                           uint32_t *__restrict__ own) {
    const uint32_t nc = meta[kMNc], ncc = meta[kMNcc];
    for (uint32_t c0 = grid_tid() & ~63u; c0 < nc; c0 += grid_stride()) {
        const uint32_t c = c0 + lane_id();
        uint32_t a = kNone, b = kNone;
        if (c < nc) {
            cons_ends(cs, c, ncc, a, b);
            if (im[a] == 0.f) a = kNone;
            if (im[b] == 0.f) b = kNone;
        }
        const WaveRun r = wave_run(a);
        uint32_t base = 0;
        if (r.head && a != kNone) base = atomicAdd(&fill[a], r.len);
        base = __shfl(base, static_cast<int>(r.start), 64);
        if (a != kNone) {
            const uint32_t p = off[a] + base + (lane_id() - r.start);
            adj[p] = c << 1;
            own[p] = a;
        }
        if (b != kNone) {
            const uint32_t p = off[b] + atomicAdd(&fill[b], 1u);
            adj[p] = (c << 1) | 1u;
            own[p] = b;
        }
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

// Workgroup-aggregated append of one value per thread: one atomic per
// workgroup (the counter is shared by the whole grid, so fewer arrivals on
// its address).  Must be reached by every thread of the workgroup.
__device__ __forceinline__ void block_append(bool pred, uint32_t val, uint32_t *out, uint32_t *count) {
    __shared__ uint32_t s_cnt[kSolveThreads / 64 + 1];
    const uint32_t lane = lane_id(), wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const uint64_t m = __ballot(pred);
    if (lane == 0) s_cnt[wave] = static_cast<uint32_t>(__popcll(m));
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t total = 0;
        for (uint32_t w = 0; w < nw; ++w) {
            const uint32_t c = s_cnt[w];
            s_cnt[w] = total;
            total += c;
        }
        s_cnt[nw] = total ? atomicAdd(count, total) : 0u;
    }
    __syncthreads();
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    if (pred) out[s_cnt[nw] + s_cnt[wave] + static_cast<uint32_t>(__popcll(m & lt))] = val;
    __syncthreads();
}

// Per-body contact lists.  Every dynamic endpoint (contact c, end e) has an
// entry adj[p] = (c << 1 | e) in its body's segment [off[x], off[x] + deg[x]),
// and own[p] = x.  rank_entry places c at its rank in the body's order
// (slot[off[x] + rank]) and records the rank per endpoint (crank[2c + e]).
// Cost per entry: one pass over the body's segment (L2-resident).
__device__ __forceinline__ void rank_entry(uint32_t p, const uint32_t *__restrict__ adj,
                                           const uint32_t *__restrict__ own, const uint32_t *__restrict__ off,
                                           const uint32_t *__restrict__ deg, int serial, uint32_t *__restrict__ slot,
                                           uint32_t *__restrict__ crank) {
    const uint32_t v = adj[p], x = own[p];
    const uint32_t c = v >> 1;
    const uint64_t kc = order_key(c, serial);
    const uint32_t beg = off[x], n = deg[x];
    const uint32_t self = p - beg;
    uint32_t rank = 0;
    for (uint32_t q0 = 0; q0 < n; q0 += 8u) {  // eight entries' loads in flight at a time
        uint32_t e[8];
#pragma unroll
        for (uint32_t k = 0; k < 8u; ++k) e[k] = q0 + k < n ? adj[beg + q0 + k] : 0u;
#pragma unroll
        for (uint32_t k = 0; k < 8u; ++k) {
            const uint32_t q = q0 + k;
            const uint64_t kq = order_key(e[k] >> 1, serial);
            rank += (q < n && (kq > kc || (kq == kc && q < self))) ? 1u : 0u;  // ties (a == b) by position
        }
    }
    slot[beg + rank] = c;
    crank[2u * c + (v & 1u)] = rank;
}

// Links contact c to its successor in both endpoint lists and counts the
// lists it does not head yet (pending); returns true when it heads all of
// them.  lrec[c] = (a or none, b or none, next on a, next on b); static
// endpoints are "none".
__device__ __forceinline__ bool link_contact(uint32_t c, const Cons &cs, uint32_t ncc, const float *__restrict__ im,
                                             const uint32_t *__restrict__ off, const uint32_t *__restrict__ deg,
                                             const uint32_t *__restrict__ slot, const uint32_t *__restrict__ crank,
                                             uint4 *__restrict__ lrec, uint32_t *__restrict__ pending) {
    uint32_t ends[2];
    cons_ends(cs, c, ncc, ends[0], ends[1]);
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
    return wait == 0u;
}

__global__ __launch_bounds__(kSolveThreads) void list_rank(const uint32_t *__restrict__ adj,
                                                           const uint32_t *__restrict__ own,
                                                           const uint32_t *__restrict__ off,
                                                           const uint32_t *__restrict__ deg,
                                                           const uint32_t *__restrict__ meta,
                                                           uint32_t *__restrict__ slot, uint32_t *__restrict__ crank) {
    const uint32_t total = meta[kMDyn];
    const int serial = static_cast<int>(meta[kMSerial]);
    for (uint32_t p = grid_tid(); p < total; p += grid_stride()) rank_entry(p, adj, own, off, deg, serial, slot, crank);
}

// Contacts heading every list of a dynamic endpoint form the first frontier
// (fr0, size in fcnt[0]); the dataflow colouring's queue (queue[0, nc)) is
// cleared to all ones.
__global__ __launch_bounds__(kSolveThreads) void list_link(Cons cs, const uint32_t *__restrict__ meta,
                                                           const float *__restrict__ im,
                                                           const uint32_t *__restrict__ off,
                                                           const uint32_t *__restrict__ deg,
                                                           const uint32_t *__restrict__ slot,
                                                           const uint32_t *__restrict__ crank,
                                                           uint4 *__restrict__ lrec, uint32_t *__restrict__ pending,
                                                           uint32_t *__restrict__ frontier, uint32_t *fcount,
                                                           uint64_t *__restrict__ queue) {
    const uint32_t nc = meta[kMNc], ncc = meta[kMNcc];
    for (uint32_t c0 = blockIdx.x * kSolveThreads; c0 < nc; c0 += gridDim.x * kSolveThreads) {
        const uint32_t c = c0 + threadIdx.x;
        const bool head = c < nc && link_contact(c, cs, ncc, im, off, deg, slot, crank, lrec, pending);
        if (c < nc) queue[c] = ~0ull;
        block_append(head, c, frontier, fcount);
    }
}

// Assigns round r's frontier entry c: class = lowest colour free in
// used[a] | used[b] (round - 1 on the reference-order path), then both
// successors lose one pending predecessor; those reaching zero are returned
// in (ready, nx) for the next frontier.
__device__ __forceinline__ uint32_t assign_one(uint32_t c, uint32_t round, int serial,
                                               const uint4 *__restrict__ lrec, uint32_t *pending, uint64_t *used,
                                               uint64_t *state, uint32_t *meta, bool (&ready)[2], uint32_t (&nx)[2]) {
    const uint4 r = lrec[c];
    uint32_t cls;
    if (serial) {
        cls = round - 1u;
    } else {
        const uint64_t ua = r.x != kNone ? ld_agent(&used[r.x]) : 0ull;
        const uint64_t ub = r.y != kNone ? ld_agent(&used[r.y]) : 0ull;
        const uint64_t u = ua | ub;
        if (u == ~0ull) {
            atomicOr(&meta[kMOverflow], 1u);
            cls = kMaxColours;
        } else {
            cls = static_cast<uint32_t>(__builtin_ctzll(~u));
        }
        const uint64_t bit = 1ull << (cls & 63u);
        if (r.x != kNone) used[r.x] = ua | bit;
        if (r.y != kNone) used[r.y] = ub | bit;
    }
    state[c] = (static_cast<uint64_t>(round) << 32) | cls;
    nx[0] = r.z;
    nx[1] = r.w;
    const uint32_t oa = r.z != kNone ? atomicSub(&pending[r.z], 1u) : 0u;
    const uint32_t ob = r.w != kNone ? atomicSub(&pending[r.w], 1u) : 0u;
    ready[0] = r.z != kNone && oa == 1u;
    ready[1] = r.w != kNone && ob == 1u;
    return cls;
}

// Buffers of the cooperative class assignment.
struct AssignBufs {
    Cons cons;
    const float *im;
    const uint32_t *adj, *own, *off, *deg;
    uint32_t *slot, *crank;
    uint4 *lrec;
    uint64_t *fq;  // colour_flow's queue: (round << 32) | contact
    uint32_t *pending;
    uint64_t *used, *state;
    uint32_t *fr0, *fr1;
    uint32_t *meta, *hist, *boff, *bfill;
    uint32_t nbody;
};

// Touches every 128-byte line of [p, p + bytes) once across the participants
// (gtid of gstride), so that the XCD's L2 holds the lines before the
// dependent accesses of colour_flow reach them.
__device__ __forceinline__ void touch_lines(const void *p, size_t bytes, uint32_t gtid, uint32_t gstride) {
    const uint8_t *b = static_cast<const uint8_t *>(p);
    for (size_t o = static_cast<size_t>(gtid) * 128u; o < bytes; o += static_cast<size_t>(gstride) * 128u) {
        const uint32_t v = *reinterpret_cast<const uint32_t *>(b + o);
        asm volatile("" ::"v"(v));
    }
}

// Dataflow colouring: the JP assignment without round barriers.  A contact
// can be coloured once the contacts ahead of it on both its lists are: then
// used[a] | used[b] holds exactly its higher-priority neighbours' colours, as
// in its JP round, and its round (the longest decreasing-priority chain
// ending at it) is one more than the larger of its list predecessors'
// rounds.  A predecessor decrements a successor's pending count and adds its
// own round above bit 8 in the same atomic, so the last of a contact's (at
// most two) predecessors learns the other's round from the returned value.
// That lane continues with the successor itself (its list record loaded
// while the predecessor was coloured); a second ready successor is pushed to
// a queue (positions fr0[0, n0) hold the first frontier, then queue[0, ..)
// entries (round << 32 | contact) in push order), from which lanes with
// nothing in hand claim positions, one atomic per wave.  A claimed position
// is read until its entry appears; it may never appear once every contact
// is coloured, which ends the pass.  Writers wait for their stores (masks,
// state) before the decrements; data written during the kernel is read with
// agent-scope loads (the one-XCD handoff of assign_lists' rounds).
// Progress: a claimed position below the queue's fill level is always
// filled, so while uncoloured contacts remain some lane holds a ready one.
// False on a timeout (kErrFlow).
__device__ bool colour_flow(const AssignBufs &B, uint32_t nc, uint32_t *lhist, uint32_t &my_max,
                            uint32_t &my_done, uint32_t &my_round) {
    uint32_t *meta = B.meta;
    uint64_t *queue = B.fq;
    const uint32_t lane = lane_id();
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const uint32_t n0 = ld_agent(&meta[kMFcnt]);
    uint32_t cur = kNone, pos = kNone, round = 0, spins = 0, unpublished = 0;
    uint4 rec = make_uint4(kNone, kNone, kNone, kNone);
    bool have_rec = false;
    for (;;) {
        // lanes with nothing in hand claim queue positions
        const bool need = cur == kNone && pos == kNone;
        const uint64_t nm = __ballot(need);
        if (nm) {
            const uint32_t lead = static_cast<uint32_t>(__builtin_ctzll(nm));
            uint32_t base = 0;
            if (lane == lead) base = atomicAdd(&meta[kMQHead], static_cast<uint32_t>(__popcll(nm)));
            base = __shfl(base, static_cast<int>(lead), 64);
            if (need) pos = base + static_cast<uint32_t>(__popcll(nm & lt));
        }
        if (pos < nc) {
            if (pos < n0) {
                cur = B.fr0[pos];
                round = 1u;
                pos = kNone;
                have_rec = false;
            } else {
                const uint64_t w = ld_agent(&queue[pos - n0]);
                if (w != ~0ull) {
                    cur = static_cast<uint32_t>(w);
                    round = static_cast<uint32_t>(w >> 32);
                    pos = kNone;
                    have_rec = false;
                }
            }
        }
        // colour the contacts in hand
        bool push = false;
        uint64_t pushed = 0;
        const bool work = cur != kNone;
        if (work) {
            if (!have_rec) rec = B.lrec[cur];
            const uint64_t ua = rec.x != kNone ? ld_agent(&B.used[rec.x]) : 0ull;
            const uint64_t ub = rec.y != kNone ? ld_agent(&B.used[rec.y]) : 0ull;
            const uint4 none4 = make_uint4(kNone, kNone, kNone, kNone);
            const uint4 rec_a = rec.z != kNone ? B.lrec[rec.z] : none4;  // the successors' records
            const uint4 rec_b = rec.w != kNone ? B.lrec[rec.w] : none4;
            const uint64_t u = ua | ub;
            uint32_t cls;
            if (u == ~0ull) {
                atomicOr(&meta[kMOverflow], 1u);
                cls = kMaxColours;
            } else {
                cls = static_cast<uint32_t>(__builtin_ctzll(~u));
            }
            const uint64_t bit = 1ull << (cls & 63u);
            if (rec.x != kNone) B.used[rec.x] = ua | bit;
            if (rec.y != kNone) B.used[rec.y] = ub | bit;
            B.state[cur] = (static_cast<uint64_t>(round) << 32) | cls;
            if (round > 256u) atomicOr(&meta[kMDeep], 1u);
            wait_vmem();  // masks and state in the L2 before a successor can start
            const uint32_t rr = round < 0xffffu ? round : 0xffffu;  // rounds past 256 only flag
            const uint32_t add = (rr << 8) - 1u;
            const uint32_t oa = rec.z != kNone ? atomicAdd(&B.pending[rec.z], add) : 0u;
            const uint32_t ob = rec.w != kNone ? atomicAdd(&B.pending[rec.w], add) : 0u;
            const bool ra = rec.z != kNone && (oa & 0xffu) == 1u, rb = rec.w != kNone && (ob & 0xffu) == 1u;
            const uint32_t na = 1u + (rr > (oa >> 8) ? rr : (oa >> 8));
            const uint32_t nb = 1u + (rr > (ob >> 8) ? rr : (ob >> 8));
            my_max = cls > my_max ? cls : my_max;
            my_round = round > my_round ? round : my_round;
            ++my_done;
            atomicAdd(&lhist[cls < kMaxColours ? cls : kMaxColours], 1u);
            const uint32_t succ_b = rec.w;
            if (ra) {
                cur = rec.z;
                rec = rec_a;
                round = na;
                have_rec = true;
                push = rb;
                pushed = (static_cast<uint64_t>(nb) << 32) | succ_b;
            } else if (rb) {
                cur = succ_b;
                rec = rec_b;
                round = nb;
                have_rec = true;
            } else {
                cur = kNone;
                have_rec = false;
            }
        }
        // a second ready successor goes to the queue
        const uint64_t pm = __ballot(push);
        if (pm) {
            const uint32_t lead = static_cast<uint32_t>(__builtin_ctzll(pm));
            uint32_t base = 0;
            if (lane == lead) base = atomicAdd(&meta[kMQTail], static_cast<uint32_t>(__popcll(pm)));
            base = __shfl(base, static_cast<int>(lead), 64);
            if (push) queue[base + static_cast<uint32_t>(__popcll(pm & lt))] = pushed;
        }
        unpublished += static_cast<uint32_t>(__popcll(__ballot(work)));
        if (!__ballot(cur != kNone)) {
            // the count is read only by idle waves: a wave adds its coloured
            // contacts when it runs out of work
            if (unpublished) {
                if (lane == 0) atomicAdd(&meta[kMQDone], unpublished);
                unpublished = 0;
            }
            if (ld_agent(&meta[kMQDone]) >= nc) return true;
            __builtin_amdgcn_s_sleep(1);
            if (++spins > kGridSpinLimit) {
                if (lane == 0) atomicOr(&meta[kMErr], kErrFlow);
                return false;
            }
            if ((spins & 255u) == 0u && (ld_agent(&meta[kMErr]) & kErrFlow)) return false;
        }
    }
}

// JP rounds over frontiers in one cooperative launch.  Round r assigns the
// contacts that head the ordered lists of all their dynamic endpoints: on a
// body's list every earlier contact is a higher-priority neighbour, so these
// are exactly the contacts whose higher-priority neighbours were all assigned
// before round r — the ones the reference colours in its round r.  A body
// heads at most one contact per round, so the contact owns its bodies' masks
// of used colours during the round.  Values written in a round are read only
// after the barrier that ends it.  Frontier sizes rotate over fcnt[0..2]
// (round r appends to fcnt[r % 3]).
//
// When the colouring overflows 64 colours or needs more than 256 rounds, the
// lists are re-ranked in contact order and the rounds re-run as dependency
// levels: the reference's run_serial order.  The kernel ends with the class
// offsets (boff[0..k]), zeroed per-class fill counters and the statistics.
// Like the sweep, it runs on the workgroups of one XCD (see xcd_barrier):
// frontier entries, used-colour masks and counters written during the kernel
// are read at agent scope (L2), pending counts are L2 atomics.
__global__ __launch_bounds__(1024) void assign_lists(AssignBufs B) {
    __shared__ uint32_t lhist[kMaxColours + 1];
    uint32_t *meta = B.meta;
    uint32_t *bar = meta + kMBarAssign, *fcnt = meta + kMFcnt, *err = meta + kMErr;
    uint32_t bidx = 0, nblk = 0;
    if (!xcd_checkin(kSweepXcc, meta + kMRegAssign, err, kErrBarrier, bidx, nblk)) return;
    uint32_t gen = 0;
    const uint32_t gtid = bidx * blockDim.x + threadIdx.x, gstride = nblk * blockDim.x;
    const uint32_t lane = lane_id(), wave0 = gtid & ~63u;
    const uint32_t nc = meta[kMNc];
    int serial = static_cast<int>(meta[kMSerial]);
    uint32_t my_max = 0, my_done = 0, rounds = 0, running = 0;
    bool exceeded = false;
    for (int pass = 0; pass < 2; ++pass) {
        for (uint32_t k = threadIdx.x; k <= kMaxColours; k += blockDim.x) lhist[k] = 0;
        __syncthreads();
        my_max = 0;
        my_done = 0;
        running = 0;
        exceeded = false;
        const uint32_t max_rounds = serial ? nc + 1u : 256u;
        uint32_t *cur = B.fr0, *nxt = B.fr1;
        uint32_t ncur = ld_agent(&fcnt[0]), round = 1;
        if (!serial) {  // the colouring as dataflow; rounds only for the reference order
            touch_lines(B.used, sizeof(uint64_t) * B.nbody, gtid, gstride);
            touch_lines(B.lrec, sizeof(uint4) * nc, gtid, gstride);
            touch_lines(B.pending, sizeof(uint32_t) * nc, gtid, gstride);
            touch_lines(B.fq, sizeof(uint64_t) * nc, gtid, gstride);
            touch_lines(B.fr0, sizeof(uint32_t) * ld_agent(&fcnt[0]), gtid, gstride);
            uint32_t my_round = 0;
            (void)colour_flow(B, nc, lhist, my_max, my_done, my_round);
            my_round = wave_max_u32(my_round);
            if (lane == 0 && my_round) atomicMax(&meta[kMRounds], my_round);
            ncur = 0;
        }
        for (; ncur > 0; ++round) {
            if (round > max_rounds) {  // deeper than the reference allows (uniform)
                exceeded = true;
                break;
            }
            if (serial && gtid == 0) B.boff[round - 1] = running;  // class round-1 = this frontier
            running += ncur;
            uint32_t *nxt_count = &fcnt[round % 3];
            if (gtid == 0) fcnt[(round + 1) % 3] = 0u;
            for (uint32_t i0 = wave0; i0 < ncur; i0 += gstride) {
                const uint32_t i = i0 + lane;
                bool ready[2] = {false, false};
                uint32_t nx[2] = {0u, 0u};
                if (i < ncur) {
                    const uint32_t cls = assign_one(ld_agent(&cur[i]), round, serial, B.lrec, B.pending, B.used,
                                                    B.state, meta, ready, nx);
                    my_max = cls > my_max ? cls : my_max;
                    ++my_done;
                    if (!serial) atomicAdd(&lhist[cls < kMaxColours ? cls : kMaxColours], 1u);
                }
                wave_append<2>(ready, nx, nxt, nxt_count);
            }
            xcd_barrier(bar, gen, nblk, err, kErrBarrier);
            ncur = ld_agent(&fcnt[round % 3]);
            uint32_t *t = cur; cur = nxt; nxt = t;
        }
        rounds = round - 1u;
        xcd_barrier(bar, gen, nblk, err, kErrBarrier);  // overflow flags of the last round
        if (!serial) {
            rounds = ld_agent(&meta[kMRounds]);
            exceeded = ld_agent(&meta[kMDeep]) != 0u;  // a timeout leaves contacts unassigned: kErrAssign
        }
        const bool fallback = !serial && (exceeded || ld_agent(&meta[kMOverflow]) != 0u);
        if (!fallback) break;
        // The reference's serial path: contact order, classes = dependency
        // levels.  The lists are rebuilt here, so the barriers around the
        // rebuild also make plain loads of them see the other workgroups'
        // stores.
        serial = 1;
        if (gtid == 0) {
            meta[kMSerial] = 1u;
            meta[kMColourRounds] = exceeded ? 257u : rounds;
            fcnt[0] = 0u;
            fcnt[1] = 0u;
            fcnt[2] = 0u;
        }
        const uint32_t total = meta[kMDyn];
        for (uint32_t p = gtid; p < total; p += gstride)
            rank_entry(p, B.adj, B.own, B.off, B.deg, 1, B.slot, B.crank);
        xcd_barrier(bar, gen, nblk, err, kErrBarrier, true);
        for (uint32_t c0 = wave0; c0 < nc; c0 += gstride) {
            const uint32_t c = c0 + lane;
            const bool pred[1] = {c < nc && link_contact(c, B.cons, meta[kMNcc], B.im, B.off, B.deg, B.slot,
                                                         B.crank, B.lrec, B.pending)};
            const uint32_t val[1] = {c};
            wave_append<1>(pred, val, B.fr0, &fcnt[0]);
        }
        xcd_barrier(bar, gen, nblk, err, kErrBarrier, true);
    }
    // Class sizes: per-block colour histograms, or the frontier sizes.
    if (!serial) {
        __syncthreads();
        for (uint32_t k = threadIdx.x; k <= kMaxColours; k += blockDim.x)
            if (lhist[k]) atomicAdd(&B.hist[k], lhist[k]);
    }
    my_max = wave_max_u32(my_max);
    my_done = wave_sum_u32(my_done);
    if (lane == 0 && my_done) {
        atomicMax(&meta[kMMaxCls], my_max);
        atomicAdd(&meta[kMAssigned], my_done);
    }
    xcd_barrier(bar, gen, nblk, err, kErrBarrier);
    const bool ok = !exceeded && ld_agent(&meta[kMAssigned]) == nc;
    uint32_t ncls = 0;
    if (ok) ncls = serial ? rounds : (nc ? ld_agent(&meta[kMMaxCls]) + 1u : 0u);
    if (gtid == 0) {
        if (!ok) atomicOr(err, kErrAssign);
        meta[kMRounds] = rounds;
        meta[kMNcls] = ncls;
        if (serial) B.boff[ncls] = ok ? running : 0u;
    }
    if (!serial && gtid < 64u) {
        // class offsets: exclusive scan of the <= 65 class sizes by one wave
        const uint32_t h0 = gtid < ncls ? ld_agent(&B.hist[gtid]) : 0u;
        const uint32_t h1 = gtid + 64u < ncls ? ld_agent(&B.hist[gtid + 64u]) : 0u;
        uint32_t s0 = h0, s1 = h1;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t t0 = __shfl_up(s0, d, 64), t1 = __shfl_up(s1, d, 64);
            if (lane >= static_cast<uint32_t>(d)) {
                s0 += t0;
                s1 += t1;
            }
        }
        const uint32_t tot0 = __shfl(s0, 63, 64), tot1 = __shfl(s1, 63, 64);
        if (gtid < ncls) B.boff[gtid] = s0 - h0;
        if (gtid + 64u < ncls) B.boff[gtid + 64u] = tot0 + s1 - h1;
        if (gtid == 0) B.boff[ncls] = tot0 + tot1;
    }
    for (uint32_t k = gtid; k < ncls; k += gstride) B.bfill[k] = 0u;
}

// One round per launch (used when a cooperative launch is unavailable or
// disabled): the frontier of round r is read from fr[r & 1] with its size in
// fcnt[(r - 1) % 3], the next one written to fr[(r + 1) & 1].
__global__ __launch_bounds__(kSolveThreads) void assign_lists_round(const uint4 *__restrict__ lrec, uint32_t round,
                                                                    uint32_t *pending, uint64_t *used,
                                                                    uint64_t *state, uint32_t *fr0, uint32_t *fr1,
                                                                    uint32_t *meta) {
    uint32_t *fcnt = meta + kMFcnt;
    const uint32_t ncur = fcnt[(round - 1u) % 3u];
    const int serial = static_cast<int>(meta[kMSerial]);
    uint32_t *cur = (round & 1u) ? fr0 : fr1;
    uint32_t *nxt = (round & 1u) ? fr1 : fr0;
    uint32_t *nxt_count = &fcnt[round % 3u];
    const uint32_t lane = lane_id();
    if (grid_tid() == 0 && ncur > 0) atomicMax(&meta[kMRounds], round);
    for (uint32_t i0 = grid_tid() & ~63u; i0 < ncur; i0 += grid_stride()) {
        const uint32_t i = i0 + lane;
        bool ready[2] = {false, false};
        uint32_t nx[2] = {0u, 0u};
        if (i < ncur) {
            const uint32_t cls = assign_one(cur[i], round, serial, lrec, pending, used, state, meta, ready, nx);
            atomicMax(&meta[kMMaxCls], cls);
            atomicAdd(&meta[kMAssigned], 1u);
        }
        wave_append<2>(ready, nx, nxt, nxt_count);
    }
}

// Sets fcnt[idx] = 0 and, when `restart` (1, or 2 after more than 256
// colouring rounds), switches to the reference order and clears the round
// statistics (launched between rounds by the host).
__global__ void round_control(uint32_t *meta, uint32_t idx, int restart) {
    meta[kMFcnt + idx] = 0u;
    if (restart) {
        meta[kMColourRounds] = restart == 2 ? 257u : meta[kMRounds];
        meta[kMSerial] = 1u;
        meta[kMRounds] = 0u;
        meta[kMMaxCls] = 0u;
        meta[kMAssigned] = 0u;
        meta[kMFcnt] = 0u;
        meta[kMFcnt + 1] = 0u;
        meta[kMFcnt + 2] = 0u;
    }
}

// Class histogram for the launch-per-round path.
__global__ __launch_bounds__(kSolveThreads) void class_count(const uint32_t *__restrict__ meta, uint32_t n_cls,
                                                             const uint64_t *__restrict__ state,
                                                             uint32_t *__restrict__ hist) {
    const uint32_t nc = meta[kMNc];
    for (uint32_t c = grid_tid(); c < nc; c += grid_stride()) {
        const uint32_t k = cls_of(state[c]);
        if (k < n_cls) atomicAdd(&hist[k], 1u);
    }
}

// Constraint record for the sweeps, class-sorted: (a, b), (p, n.x), (n.y,
// n.z, depth, 0) for a contact; (parent, child), (joint index, -, -, -),
// (-, -, -, kRecJoint) for a joint.
constexpr uint32_t kRecJoint = 1u;
struct CRec {
    uint2 *ab;
    float4 *p0, *p1;
};

__global__ __launch_bounds__(1024) void class_gather(Cons cs, uint32_t *meta,
                                                              const uint64_t *__restrict__ state,
                                                              const uint32_t *__restrict__ boff,
                                                              uint32_t *__restrict__ bfill, CRec rec,
                                                              uint32_t *__restrict__ rec_id,
                                                              uint32_t *__restrict__ cpos,
                                                              uint4 *__restrict__ rseq) {
    __shared__ uint32_t lh[kLdsClasses];
    __shared__ uint32_t lbase[kLdsClasses];
    const uint32_t nc = meta[kMNc], n_cls = meta[kMNcls], ncc = meta[kMNcc];
    const bool local = n_cls <= kLdsClasses;
    for (uint32_t c0 = blockIdx.x * blockDim.x; c0 < nc; c0 += grid_stride()) {
        const uint32_t c = c0 + threadIdx.x;
        if (local)
            for (uint32_t k = threadIdx.x; k < n_cls; k += blockDim.x) lh[k] = 0;
        __syncthreads();
        uint32_t k = 0, rank = 0;
        bool valid = c < nc;
        if (valid) {
            k = cls_of(state[c]);
            if (k >= n_cls) {
                atomicOr(&meta[kMErr], kErrAssign);
                valid = false;
            } else {
                rank = local ? atomicAdd(&lh[k], 1u) : atomicAdd(&bfill[k], 1u);
            }
        }
        __syncthreads();
        if (local)
            for (uint32_t q = threadIdx.x; q < n_cls; q += blockDim.x)
                lbase[q] = lh[q] ? atomicAdd(&bfill[q], lh[q]) : 0u;
        __syncthreads();
        if (valid) {
            const uint32_t pos = boff[k] + (local ? lbase[k] + rank : rank);
            if (pos < nc) {
                const dy_contact_view &cv = cs.cv;
                if (c < ncc) {
                    rec.ab[pos] = make_uint2(cv.body_a[c], cv.body_b[c]);
                    rec.p0[pos] = make_float4(cv.pos_x[c], cv.pos_y[c], cv.pos_z[c], cv.norm_x[c]);
                    rec.p1[pos] = make_float4(cv.norm_y[c], cv.norm_z[c], cv.depth[c], 0.f);
                } else {
                    rec.ab[pos] = cs.jend[c - ncc];
                    rec.p0[pos] = make_float4(__uint_as_float(c - ncc), 0.f, 0.f, 0.f);
                    rec.p1[pos] = make_float4(0.f, 0.f, 0.f, __uint_as_float(kRecJoint));
                }
                rec_id[pos] = c;
                cpos[c] = pos;
                rseq[pos] = make_uint4(0u, 0u, 0u, 0u);
            } else {
                atomicOr(&meta[kMErr], kErrAssign);
            }
        }
        __syncthreads();
    }
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

// Applies record (p0, p1) in iteration it to the loaded bodies, first
// brought to `it` renormalisations.  For a joint this includes a static
// body, in this private copy only: the joint reads its orientation, which
// the reference renormalises after every iteration like any other body's.
template <bool JOINTS>
__device__ __forceinline__ void apply_record(const float4 &p0, const float4 &p1, BodyState &A, BodyState &B,
                                             uint32_t ra, uint32_t rb, uint32_t it, float &l, float mu, float dt,
                                             const float4 *__restrict__ jpk) {
    if (JOINTS && __float_as_uint(p1.w) == kRecJoint) {
        catch_up(A, ra, it);
        catch_up(B, rb, it);
        solve_joint(load_joint(jpk, __float_as_uint(p0.x)), A, B, dt);
        return;
    }
    if (A.im != 0.f) catch_up(A, ra, it);
    if (B.im != 0.f) catch_up(B, rb, it);
    const ContactIn ci{p0.x, p0.y, p0.z, p0.w, p1.x, p1.y, p1.z};
    project_contact(ci, A, B, mu, l, dt);
}

template <bool JOINTS>
__device__ __forceinline__ void solve_one(uint32_t i, uint32_t it, const CRec &rec, float *__restrict__ lam,
                                          float4 *__restrict__ dyn, const float4 *__restrict__ sta, float mu,
                                          float dt, const float4 *__restrict__ jpk) {
    const uint2 ab = rec.ab[i];
    const float4 p0 = rec.p0[i], p1 = rec.p1[i];
    uint32_t ra, rb;
    BodyState A = load_aos(dyn, sta, ab.x, ra), B = load_aos(dyn, sta, ab.y, rb);
    const bool da = A.im != 0.f, db = B.im != 0.f;
    float l = lam[i];
    apply_record<JOINTS>(p0, p1, A, B, ra, rb, it, l, mu, dt, jpk);
    lam[i] = l;
    if (da) store_aos(dyn, ab.x, A, it);
    if (db) store_aos(dyn, ab.y, B, it);
}

// Joint motors, once per solve before the first iteration, class by class
// (xpbd_solver.cpp:566-577 in joint order, 703-719 in colour order; a
// body's joints run in index order on the reference-order path because the
// dependency levels follow it).  One workgroup: a class's joints share no
// dynamic body, and the workgroup barrier orders the classes.
__global__ __launch_bounds__(1024) void joint_motors(const uint32_t *__restrict__ boff, CRec rec,
                                                     float4 *__restrict__ dyn, const float4 *__restrict__ sta,
                                                     const float4 *__restrict__ jpk, float dt,
                                                     const uint32_t *__restrict__ meta) {
    if ((meta[kMErr] & kErrAssign) || meta[kMNj] == 0u) return;
    const uint32_t n_cls = meta[kMNcls];
    for (uint32_t k = 0; k < n_cls; ++k) {
        const uint32_t end = boff[k + 1];
        for (uint32_t i = boff[k] + threadIdx.x; i < end; i += blockDim.x) {
            if (__float_as_uint(rec.p1[i].w) != kRecJoint) continue;
            const JointIn j = load_joint(jpk, __float_as_uint(rec.p0[i].x));
            if (j.type != kJointRevolute && j.type != kJointPrismatic) continue;
            const uint2 ab = rec.ab[i];
            uint32_t ra, rb;
            BodyState A = load_aos(dyn, sta, ab.x, ra), B = load_aos(dyn, sta, ab.y, rb);
            joint_motor(A, B, j, dt);
            if (A.im != 0.f) store_aos(dyn, ab.x, A, ra);
            if (B.im != 0.f) store_aos(dyn, ab.y, B, rb);
        }
        __syncthreads();
    }
}

// One class (fallback path: one launch per class).
template <bool JOINTS>
__global__ __launch_bounds__(kSolveThreads) void sweep_class(const uint32_t *__restrict__ boff, uint32_t k, uint32_t it,
                                                             CRec rec, float *__restrict__ lam, float4 *__restrict__ dyn,
                                                             const float4 *__restrict__ sta, float mu, float dt,
                                                             const float4 *__restrict__ jpk) {
    const uint32_t beg = boff[k], end = boff[k + 1];
    for (uint32_t i = beg + blockIdx.x * blockDim.x + threadIdx.x; i < end; i += gridDim.x * blockDim.x)
        solve_one<JOINTS>(i, it, rec, lam, dyn, sta, mu, dt, jpk);
}

// All I x K phases in one cooperative launch; the class count comes from the
// assignment (meta), the phases are separated by grid barriers.  Nothing is
// swept after an assignment error.
template <bool JOINTS>
__global__ __launch_bounds__(1024) void sweep_persistent(const uint32_t *__restrict__ boff, uint32_t n_iters,
                                                         CRec rec, float *__restrict__ lam, float4 *__restrict__ dyn,
                                                         const float4 *__restrict__ sta, float mu, float dt,
                                                         uint32_t *meta, const float4 *__restrict__ jpk) {
    uint32_t gen = 0;
    uint32_t *bar = meta + kMBarSweep, *err = meta + kMErr;
    const uint32_t n_cls = (meta[kMErr] & kErrAssign) ? 0u : meta[kMNcls];
    const uint32_t gtid = grid_tid(), gstride = grid_stride();
    for (uint32_t it = 0; it < n_iters; ++it) {
        for (uint32_t k = 0; k < n_cls; ++k) {
            const uint32_t beg = boff[k], end = boff[k + 1];
            for (uint32_t i = beg + gtid; i < end; i += gstride)
                solve_one<JOINTS>(i, it, rec, lam, dyn, sta, mu, dt, jpk);
            if (it + 1 < n_iters || k + 1 < n_cls) grid_barrier(bar, gen, err, kErrBarrier);
        }
    }
}

// ---------------------------------------------------------------------------
// Sweep confined to one XCD.  The workgroups of one XCD share its L2: the
// vector L1 writes through to it and atomics execute in it, so workgroups of
// one XCD can hand data to each other through the L2 with no cache
// maintenance, provided every read of data written during the kernel skips
// the L1 (agent-scope loads) and every writer waits for its stores before
// arriving.  A grid barrier between the phases then costs one L2 atomic per
// workgroup plus an L2-atomic poll, instead of the agent-scope write-back and
// invalidate of the L2 at every phase.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float4 ld_agent4(const float4 *p) {
    const uint64_t *q = reinterpret_cast<const uint64_t *>(p);
    const uint64_t a = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint64_t b = __hip_atomic_load(q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return make_float4(__uint_as_float(static_cast<uint32_t>(a)), __uint_as_float(static_cast<uint32_t>(a >> 32)),
                       __uint_as_float(static_cast<uint32_t>(b)), __uint_as_float(static_cast<uint32_t>(b >> 32)));
}

__device__ __forceinline__ BodyState load_aos_agent(const float4 *__restrict__ dyn, const float4 *__restrict__ sta,
                                                    uint32_t i, uint32_t &renorms) {
    const float4 d0 = ld_agent4(&dyn[4 * i]), d1 = ld_agent4(&dyn[4 * i + 1]);
    const float4 d2 = ld_agent4(&dyn[4 * i + 2]), d3 = ld_agent4(&dyn[4 * i + 3]);
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

template <bool JOINTS>
__device__ __forceinline__ void solve_one_agent(uint32_t i, uint32_t it, const CRec &rec, float *__restrict__ lam,
                                                float4 *__restrict__ dyn, const float4 *__restrict__ sta, float mu,
                                                float dt, const float4 *__restrict__ jpk) {
    const uint2 ab = rec.ab[i];
    const float4 p0 = rec.p0[i], p1 = rec.p1[i];
    uint32_t ra, rb;
    BodyState A = load_aos_agent(dyn, sta, ab.x, ra), B = load_aos_agent(dyn, sta, ab.y, rb);
    const bool da = A.im != 0.f, db = B.im != 0.f;
    float l = lam[i];  // lam[i] is only ever touched by this thread
    apply_record<JOINTS>(p0, p1, A, B, ra, rb, it, l, mu, dt, jpk);
    lam[i] = l;
    if (da) store_aos(dyn, ab.x, A, it);
    if (db) store_aos(dyn, ab.y, B, it);
}

// A contact record and its lambda, loaded ahead of the phase that uses them.
struct CPre {
    uint2 ab;
    float4 p0, p1;
    float l;
    uint32_t i;
    bool valid;
};

__device__ __forceinline__ void pre_load(CPre &c, uint32_t i, uint32_t end, const CRec &rec,
                                         const float *__restrict__ lam) {
    c.i = i;
    c.valid = i < end;
    if (c.valid) {
        c.ab = rec.ab[i];
        c.p0 = rec.p0[i];
        c.p1 = rec.p1[i];
        c.l = lam[i];
    }
}

template <bool JOINTS>
__device__ __forceinline__ void solve_pre(const CPre &c, uint32_t it, float *__restrict__ lam,
                                          float4 *__restrict__ dyn, const float4 *__restrict__ sta, float mu,
                                          float dt, const float4 *__restrict__ jpk) {
    uint32_t ra, rb;
    BodyState A = load_aos_agent(dyn, sta, c.ab.x, ra), B = load_aos_agent(dyn, sta, c.ab.y, rb);
    const bool da = A.im != 0.f, db = B.im != 0.f;
    float l = c.l;
    apply_record<JOINTS>(c.p0, c.p1, A, B, ra, rb, it, l, mu, dt, jpk);
    lam[c.i] = l;
    if (da) store_aos(dyn, c.ab.x, A, it);
    if (db) store_aos(dyn, c.ab.y, B, it);
}

// Joint motors on the workgroups of XCD kSweepXcc (cooperative, like
// sweep_xcd): a class's motor joints spread over every participant, an L2
// barrier between classes, body records read at agent scope (other
// workgroups wrote them in earlier classes).  On the reference-order path
// the classes are dependency levels, possibly many and mostly small: the
// first participant alone runs them, with workgroup barriers.
__global__ __launch_bounds__(1024) void joint_motors_xcd(const uint32_t *__restrict__ boff, CRec rec,
                                                         float4 *__restrict__ dyn, const float4 *__restrict__ sta,
                                                         const float4 *__restrict__ jpk, float dt, uint32_t *meta) {
    uint32_t *bar = meta + kMBarMotor, *err = meta + kMErr;
    uint32_t b = 0, nblk = 0;
    if (!xcd_checkin(kSweepXcc, meta + kMRegMotor, err, kErrBarrier, b, nblk)) return;
    if ((meta[kMErr] & kErrAssign) || meta[kMNj] == 0u) return;
    const uint32_t n_cls = meta[kMNcls];
    const bool serial = meta[kMSerial] != 0u;
    if (serial && b != 0) return;
    const uint32_t first = serial ? threadIdx.x : b * blockDim.x + threadIdx.x;
    const uint32_t stride = serial ? blockDim.x : nblk * blockDim.x;
    uint32_t gen = 0;
    for (uint32_t k = 0; k < n_cls; ++k) {
        const uint32_t end = boff[k + 1];
        for (uint32_t i = boff[k] + first; i < end; i += stride) {
            if (__float_as_uint(rec.p1[i].w) != kRecJoint) continue;
            const JointIn j = load_joint(jpk, __float_as_uint(rec.p0[i].x));
            if (j.type != kJointRevolute && j.type != kJointPrismatic) continue;
            const uint2 ab = rec.ab[i];
            uint32_t ra, rb;
            BodyState A = load_aos_agent(dyn, sta, ab.x, ra), B = load_aos_agent(dyn, sta, ab.y, rb);
            joint_motor(A, B, j, dt);
            if (A.im != 0.f) store_aos(dyn, ab.x, A, ra);
            if (B.im != 0.f) store_aos(dyn, ab.y, B, rb);
        }
        if (k + 1 == n_cls) break;
        if (serial) {
            wait_vmem();
            __syncthreads();
        } else if (!xcd_barrier(bar, gen, nblk, err, kErrBarrier)) {
            return;
        }
    }
}

// All I x K phases on the workgroups of XCD kSweepXcc.  The launch has
// workgroups on every XCD (cooperative, all co-resident); each checks in once
// at agent scope, the others leave, and the participants number themselves
// densely.  Contact i of a phase always goes to the same thread.
template <bool JOINTS>
__global__ __launch_bounds__(1024) void sweep_xcd(const uint32_t *__restrict__ boff, uint32_t n_iters, CRec rec,
                                                  float *__restrict__ lam, float4 *__restrict__ dyn,
                                                  const float4 *__restrict__ sta, float mu, float dt,
                                                  uint32_t *meta, const float4 *__restrict__ jpk) {
    uint32_t *bar = meta + kMBarSweep, *err = meta + kMErr;
    uint32_t b = 0, nblk = 0;
    if (!xcd_checkin(kSweepXcc, meta + kMRegSweep, err, kErrBarrier, b, nblk)) return;
    const uint32_t n_cls = (meta[kMErr] & kErrAssign) ? 0u : meta[kMNcls];
    const uint32_t gtid = b * blockDim.x + threadIdx.x, gstride = nblk * blockDim.x;
    uint32_t gen = 0;
    // The record and lambda of this thread's first contact of the next phase
    // do not depend on the phase barrier: they are loaded before it.
    CPre pf;
    if (n_cls > 0 && n_iters > 0) pre_load(pf, boff[0] + gtid, boff[1], rec, lam);
    for (uint32_t it = 0; it < n_iters; ++it) {
        for (uint32_t k = 0; k < n_cls; ++k) {
            const uint32_t end = boff[k + 1];
            const CPre cur = pf;
            if (cur.valid) solve_pre<JOINTS>(cur, it, lam, dyn, sta, mu, dt, jpk);
            for (uint32_t i = cur.i + gstride; i < end; i += gstride)
                solve_one_agent<JOINTS>(i, it, rec, lam, dyn, sta, mu, dt, jpk);
            const bool last_k = k + 1 == n_cls;
            if (it + 1 == n_iters && last_k) break;
            const uint32_t kn = last_k ? 0u : k + 1u;
            pre_load(pf, boff[kn] + gtid, boff[kn + 1], rec, lam);
            if (!xcd_barrier(bar, gen, nblk, err, kErrBarrier)) return;
        }
    }
}

// ---------------------------------------------------------------------------
// Dataflow sweep.  Two contacts interact only through a shared dynamic body,
// and the result is the reference's as long as every body sees its contacts
// in class order, iteration after iteration.  Each dynamic body x keeps a
// version ver[x] = contacts applied to it so far; the contact with class rank
// seq on x (deg[x] contacts on x) runs in iteration it once ver[x] ==
// it * deg[x] + seq for each of its dynamic bodies, then stores the bodies,
// waits for the stores and advances the versions.  No phase barriers: a
// contact waits for its own two predecessors only.
//
// Progress: every dependency points to an earlier (iteration, record)
// position, so the earliest unfinished contact can always run; within a wave
// the lanes that are ready run while the others keep waiting, so a class
// boundary inside a wave cannot stall it.  All threads are co-resident
// (cooperative launch) and every wait is bounded (kErrFlow aborts the sweep).
// Without joints, one barrier after iteration 0 decides whether any contact
// is still live; usually none is, and the sweep ends there.
// Versions and body records move through the L2 of XCD kSweepXcc (agent-scope
// loads, stores acknowledged before the version store), as in sweep_xcd.
// ---------------------------------------------------------------------------

// rseq[pos] = (class rank on a, contacts on a, class rank on b, contacts on
// b) for the record at class-sorted position pos; zero counts for static
// ends.  One thread per per-body list entry; classes on one body are distinct.
__global__ __launch_bounds__(kSolveThreads) void flow_seq(const uint32_t *__restrict__ adj,
                                                          const uint32_t *__restrict__ own,
                                                          const uint32_t *__restrict__ off,
                                                          const uint32_t *__restrict__ deg,
                                                          const uint64_t *__restrict__ state,
                                                          const uint32_t *__restrict__ cpos,
                                                          const uint32_t *__restrict__ meta,
                                                          uint4 *__restrict__ rseq) {
    if (meta[kMErr] & kErrAssign) return;
    const uint32_t total = meta[kMDyn];
    for (uint32_t p = grid_tid(); p < total; p += grid_stride()) {
        const uint32_t v = adj[p], x = own[p];
        const uint32_t c = v >> 1;
        const uint32_t kc = cls_of(state[c]);
        const uint32_t beg = off[x], n = deg[x];
        uint32_t rank = 0;
        for (uint32_t q0 = 0; q0 < n; q0 += 8u) {  // eight entries' loads in flight at a time
            uint32_t e[8];
#pragma unroll
            for (uint32_t k = 0; k < 8u; ++k) e[k] = q0 + k < n ? adj[beg + q0 + k] : 0u;
            uint64_t sk[8];
#pragma unroll
            for (uint32_t k = 0; k < 8u; ++k) sk[k] = q0 + k < n ? state[e[k] >> 1] : 0ull;
#pragma unroll
            for (uint32_t k = 0; k < 8u; ++k) rank += (q0 + k < n && cls_of(sk[k]) < kc) ? 1u : 0u;
        }
        uint32_t *r = reinterpret_cast<uint32_t *>(&rseq[cpos[c]]);
        r[2u * (v & 1u)] = rank;
        r[2u * (v & 1u) + 1u] = n;
    }
}

template <bool JOINTS>
__global__ __launch_bounds__(1024) void sweep_flow(uint32_t n_iters, CRec rec, const uint4 *__restrict__ rseq,
                                                   float *__restrict__ lam, float4 *__restrict__ dyn,
                                                   const float4 *__restrict__ sta, uint32_t *ver, float mu,
                                                   float dt, uint32_t *meta, const float4 *__restrict__ jpk) {
    uint32_t *err = meta + kMErr;
    uint32_t b = 0, nblk = 0;
    if (!xcd_checkin(kSweepXcc, meta + kMRegSweep, err, kErrBarrier, b, nblk)) return;
    if (meta[kMErr] & kErrAssign) return;
    const uint32_t nc = meta[kMNc];
    const uint32_t gstride = nblk * blockDim.x;
    const uint32_t lane = lane_id();
    const uint32_t wave0 = b * blockDim.x + (threadIdx.x & ~63u);
    uint32_t gen = 0;
    for (uint32_t it = 0; it < n_iters; ++it) {
        if (!JOINTS && it == 1) {
            // A contact's lambda changes only in its own projection, which
            // sets it to the depth (contact_live then stays false for the
            // rest of the solve) unless project_contact declines to move the
            // bodies (lambda unchanged).  With no contact left live after
            // iteration 0 the later iterations change nothing: their
            // renormalisations are applied by the final scatter.
            if (!xcd_barrier(meta + kMBarSweep, gen, nblk, err, kErrBarrier)) return;
            if (ld_agent(&meta[kMLive]) == 0u) return;
        }
        for (uint32_t i0 = wave0; i0 < nc; i0 += gstride) {
            const uint32_t i = i0 + lane;
            bool todo = i < nc;
            uint2 ab = make_uint2(0u, 0u);
            uint4 sq = make_uint4(0u, 0u, 0u, 0u);
            float4 p0 = make_float4(0.f, 0.f, 0.f, 0.f), p1 = p0;
            float l = 0.f;
            if (todo) {
                ab = rec.ab[i];
                sq = rseq[i];
                p0 = rec.p0[i];
                p1 = rec.p1[i];
                l = lam[i];
            }
            const uint32_t ta = it * sq.y + sq.x, tb = it * sq.w + sq.z;
            uint32_t spins = 0;
            while (__ballot(todo)) {
                if (todo) {
                    const bool ready = (sq.y == 0u || ld_agent(&ver[ab.x]) == ta) &&
                                       (sq.w == 0u || ld_agent(&ver[ab.y]) == tb);
                    // a contact that changes nothing (contact_live: every contact after
                    // its first projection) moves its bodies' versions only; their
                    // deferred renormalisations are applied by the next live record
                    const bool live = (JOINTS && __float_as_uint(p1.w) == kRecJoint) || contact_live(p1.z, l);
                    if (ready && live) {
                        uint32_t ra, rb;
                        BodyState A = load_aos_agent(dyn, sta, ab.x, ra), B = load_aos_agent(dyn, sta, ab.y, rb);
                        const bool da = A.im != 0.f, db = B.im != 0.f;
                        apply_record<JOINTS>(p0, p1, A, B, ra, rb, it, l, mu, dt, jpk);
                        lam[i] = l;
                        if (da) store_aos(dyn, ab.x, A, it);
                        if (db) store_aos(dyn, ab.y, B, it);
                        if (!JOINTS && it == 0 && contact_live(p1.z, l)) atomicOr(&meta[kMLive], 1u);
                    }
                    if (ready) todo = false;
                }
                wait_vmem();  // body stores acknowledged by the L2 before the versions move
                if (!todo && i < nc && sq.y != ~0u) {
                    if (sq.y) __hip_atomic_store(&ver[ab.x], ta + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    if (sq.w) __hip_atomic_store(&ver[ab.y], tb + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    sq.y = ~0u;  // versions published
                }
                if (__ballot(todo)) {
                    __builtin_amdgcn_s_sleep(1);
                    if (++spins > kGridSpinLimit) {
                        if (lane == 0) atomicOr(err, kErrFlow);
                        return;
                    }
                    if ((spins & 255u) == 0u && (l2_read(err) & kErrFlow)) return;
                }
            }
        }
    }
}

// Contact-only dataflow sweep with a pair of lanes per record (lanes 2k and
// 2k + 1 of a wave take record i0 + k, one body each, project_contact_pair):
// each lane polls, loads, renormalises, stores and publishes its own body,
// so a pass of the projection is about half as long.  Otherwise sweep_flow.
__global__ __launch_bounds__(1024) void sweep_flow_pair(uint32_t n_iters, CRec rec, const uint4 *__restrict__ rseq,
                                                        float *__restrict__ lam, float4 *__restrict__ dyn,
                                                        const float4 *__restrict__ sta, uint32_t *ver, float mu,
                                                        float dt, uint32_t *meta,
                                                        const float4 *__restrict__ /* joints: none */) {
    uint32_t *err = meta + kMErr;
    uint32_t bidx = 0, nblk = 0;
    if (!xcd_checkin(kSweepXcc, meta + kMRegSweep, err, kErrBarrier, bidx, nblk)) return;
    if (meta[kMErr] & kErrAssign) return;
    const uint32_t nc = meta[kMNc];
    const uint32_t rstride = nblk * blockDim.x / 2u;  // records per round over all waves
    const uint32_t lane = lane_id(), side = lane & 1u;
    const uint32_t wave_rec0 = (bidx * blockDim.x + (threadIdx.x & ~63u)) / 2u;
    uint32_t gen = 0;
    for (uint32_t it = 0; it < n_iters; ++it) {
        if (it == 1) {
            if (!xcd_barrier(meta + kMBarSweep, gen, nblk, err, kErrBarrier)) return;
            if (ld_agent(&meta[kMLive]) == 0u) return;
        }
        for (uint32_t r0 = wave_rec0; r0 < nc; r0 += rstride) {
            const uint32_t i = r0 + (lane >> 1);
            bool todo = i < nc;  // the same on both lanes of a pair
            uint2 ab = make_uint2(0u, 0u);
            uint4 sq = make_uint4(0u, 0u, 0u, 0u);
            float4 p0 = make_float4(0.f, 0.f, 0.f, 0.f), p1 = p0;
            float l = 0.f;
            if (todo) {
                ab = rec.ab[i];
                sq = rseq[i];
                p0 = rec.p0[i];
                p1 = rec.p1[i];
                l = lam[i];
            }
            const uint32_t x = side ? ab.y : ab.x;             // this lane's body
            const bool dyn_end = (side ? sq.w : sq.y) != 0u;    // a dynamic end: it waits and publishes
            const uint32_t t = side ? it * sq.w + sq.z : it * sq.y + sq.x;
            const bool live = contact_live(p1.z, l);
            bool published = false;
            uint32_t spins = 0;
            while (__ballot(todo)) {
                const bool mine = todo && (!dyn_end || ld_agent(&ver[x]) == t);
                // the exchange runs on every lane: a lane inside a short-circuit would read an inactive partner
                const int partner = __shfl_xor(mine ? 1 : 0, 1, 64);
                const bool ready = mine && partner != 0;
                if (ready && live) {
                    uint32_t rs;
                    BodyState S = load_aos_agent(dyn, sta, x, rs);
                    if (S.im != 0.f) catch_up(S, rs, it);
                    const ContactIn ci{p0.x, p0.y, p0.z, p0.w, p1.x, p1.y, p1.z};
                    project_contact_pair(ci, S, side, mu, l, dt, [](float v) { return __shfl_xor(v, 1, 64); });
                    if (side == 0u) lam[i] = l;
                    if (S.im != 0.f) store_aos(dyn, x, S, it);
                    if (side == 0u && it == 0 && contact_live(p1.z, l)) atomicOr(&meta[kMLive], 1u);
                }
                if (ready) todo = false;
                wait_vmem();  // body stores acknowledged by the L2 before the versions move
                if (!todo && i < nc && !published) {
                    if (dyn_end) __hip_atomic_store(&ver[x], t + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    published = true;
                }
                if (__ballot(todo)) {
                    __builtin_amdgcn_s_sleep(1);
                    if (++spins > kGridSpinLimit) {
                        if (lane == 0) atomicOr(err, kErrFlow);
                        return;
                    }
                    if ((spins & 255u) == 0u && (l2_read(err) & kErrFlow)) return;
                }
            }
        }
    }
}

// Dataflow sweep with a record cursor per lane: within an iteration a lane
// moves on to its next record (stride = participating threads) as soon as
// its current one is done, so the lanes of a wave do not wait for each
// other.  The earliest unfinished record is always at its lane's cursor and
// ready, so the sweep always progresses.  Records that change nothing only
// advance their bodies' versions, and without joints one barrier after
// iteration 0 ends the sweep when no contact is left live, as in
// sweep_flow.  Waves sleep only when none of their lanes advanced; a wave
// with no progress for kGridSpinLimit polls aborts the sweep (kErrFlow).
template <bool JOINTS>
__global__ __launch_bounds__(1024) void sweep_flow_lanes(uint32_t n_iters, CRec rec, const uint4 *__restrict__ rseq,
                                                         float *__restrict__ lam, float4 *__restrict__ dyn,
                                                         const float4 *__restrict__ sta, uint32_t *ver, float mu,
                                                         float dt, uint32_t *meta, const float4 *__restrict__ jpk) {
    uint32_t *err = meta + kMErr;
    uint32_t b = 0, nblk = 0;
    if (!xcd_checkin(kSweepXcc, meta + kMRegSweep, err, kErrBarrier, b, nblk)) return;
    if (meta[kMErr] & kErrAssign) return;
    const uint32_t nc = meta[kMNc];
    const uint32_t gstride = nblk * blockDim.x, first = b * blockDim.x + threadIdx.x;
    const uint32_t lane = lane_id();
    uint32_t gen = 0;
    for (uint32_t it = 0; it < n_iters; ++it) {
        if (!JOINTS && it == 1) {
            if (!xcd_barrier(meta + kMBarSweep, gen, nblk, err, kErrBarrier)) return;
            if (ld_agent(&meta[kMLive]) == 0u) return;
        }
        uint32_t i = first, idle = 0, ta = 0, tb = 0;
        uint2 ab = make_uint2(0u, 0u);
        uint4 sq = make_uint4(0u, 0u, 0u, 0u);
        float4 p0 = make_float4(0.f, 0.f, 0.f, 0.f), p1 = p0;
        float l = 0.f;
        if (i < nc) {
            ab = rec.ab[i]; sq = rseq[i]; p0 = rec.p0[i]; p1 = rec.p1[i]; l = lam[i];
            ta = it * sq.y + sq.x; tb = it * sq.w + sq.z;
        }
        while (__ballot(i < nc)) {
            bool done_now = false;
            if (i < nc) {
                const bool ready = (sq.y == 0u || ld_agent(&ver[ab.x]) == ta) &&
                                   (sq.w == 0u || ld_agent(&ver[ab.y]) == tb);
                if (ready) {
                    const bool live = (JOINTS && __float_as_uint(p1.w) == kRecJoint) || contact_live(p1.z, l);
                    if (live) {
                        uint32_t ra, rb;
                        BodyState A = load_aos_agent(dyn, sta, ab.x, ra), B = load_aos_agent(dyn, sta, ab.y, rb);
                        const bool da = A.im != 0.f, db = B.im != 0.f;
                        apply_record<JOINTS>(p0, p1, A, B, ra, rb, it, l, mu, dt, jpk);
                        lam[i] = l;
                        if (da) store_aos(dyn, ab.x, A, it);
                        if (db) store_aos(dyn, ab.y, B, it);
                        if (!JOINTS && it == 0 && contact_live(p1.z, l)) atomicOr(&meta[kMLive], 1u);
                    }
                    done_now = true;
                }
            }
            wait_vmem();  // body stores acknowledged by the L2 before the versions move
            if (done_now) {
                if (sq.y) __hip_atomic_store(&ver[ab.x], ta + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                if (sq.w) __hip_atomic_store(&ver[ab.y], tb + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                i += gstride;
                if (i < nc) {  // the next record, loaded now
                    ab = rec.ab[i]; sq = rseq[i]; p0 = rec.p0[i]; p1 = rec.p1[i]; l = lam[i];
                    ta = it * sq.y + sq.x; tb = it * sq.w + sq.z;
                }
            }
            if (__ballot(done_now)) {
                idle = 0;
            } else {
                __builtin_amdgcn_s_sleep(1);
                if (++idle > kGridSpinLimit) {
                    if (lane == 0) atomicOr(err, kErrFlow);
                    return;
                }
                if ((idle & 255u) == 0u && (l2_read(err) & kErrFlow)) return;
            }
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
    uint32_t *adj = nullptr, *own = nullptr, *slot = nullptr, *crank = nullptr, *pending = nullptr;
    uint32_t *fr0 = nullptr, *fr1 = nullptr;
    uint32_t *boff = nullptr, *bfill = nullptr, *hist = nullptr, *rec_id = nullptr, *cpos = nullptr;
    uint32_t *meta = nullptr;  // kMWords device words, inside meta_block
    uint8_t *meta_block = nullptr;
    uint4 *lrec = nullptr, *rseq = nullptr;
    uint2 *rec_ab = nullptr;
    uint64_t *fq = nullptr;
    float4 *rec_p0 = nullptr, *rec_p1 = nullptr;
    DevBuf<uint32_t> deg, off, fill, scan_tmp, ver;
    DevBuf<uint64_t> used;
    DevBuf<float4> dyn, sta;
    // Metadata of the last solve, copied back asynchronously into pinned
    // memory; ev_meta completes when the copy has landed.
    uint32_t *h_meta = nullptr;
    hipEvent_t ev_meta = nullptr;
    bool meta_pending = false;
    uint32_t last_nb = 0;
    dy_solve_stats stats{};
    // Co-resident workgroup limits of the cooperative kernels, [0] contacts
    // only, [1] the variants that also solve joints.
    int coop_assign_max = 0, coop_sweep_max[2] = {0, 0}, coop_xcd_max[2] = {0, 0}, coop_flow_max[2] = {0, 0},
        coop_lanes_max[2] = {0, 0}, coop_pair_max = 0, coop_motor_max = 0, device_cus = 0;
    int sweep_pair = 1;     // contact-only dataflow sweeps use sweep_flow_pair (DYPHUR_SWEEP_PAIR=0: sweep_flow)
    int pair_threads = 512;  // its workgroup size (DYPHUR_SWEEP_PAIR_THREADS)
    // Constraint capacity (contacts + joints) of the constraint-indexed
    // buffers, and the joint capacity of jpk / jend.
    uint32_t ccap = 0, jcap = 0;
    float4 *jpk = nullptr;
    uint2 *jend = nullptr;
    int sweep_mode = 2;  // 2 sweep_flow (default), 0 sweep_xcd, 1 sweep_persistent (whole grid)
    int sweep_blocks_override = 0, assign_blocks_override = 0;
    int assign_threads = 256, sweep_threads = 256;
    int path = 0;  // 0 cooperative kernels, 1 one launch per round / per class
    bool coop_refused = false;  // a cooperative launch was refused: launches per round / per class from now on
    // Kernel timing (dy_xpbd_set_timing): event pairs recorded around the
    // assignment and sweep kernels, read back once the next solve syncs.
    bool timing = false;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    bool ev_pending = false;
    dy_solver_timing acc{};
};

namespace {

uint32_t grid_for(uint32_t n) {
    const uint32_t g = ceil_div(n ? n : 1u, kSolveThreads);
    return g < kMaxGridBlocks ? g : kMaxGridBlocks;
}

// Waits for the last solve's metadata copy and fills x->stats; reports the
// device error flags.
int sync_meta(dy_xpbd_s *x) {
    if (!x->meta_pending) return DY_OK;
    DY_TRY(hipEventSynchronize(x->ev_meta));
    x->meta_pending = false;
    const uint32_t *m = x->h_meta;
    x->stats.n_constraints = m[kMNc];
    x->stats.used_serial = m[kMSerial];
    x->stats.n_classes = m[kMNcls];
    x->stats.rounds = m[kMRounds];
    x->stats.overflow = m[kMOverflow];
    x->stats.dyn_endpoints = m[kMDyn];
    if (m[kMErr] == kErrJoint) return DY_ERR_INVALID_ARG;
    return m[kMErr] ? DY_ERR_INTERNAL : DY_OK;
}

// Adds the event intervals of the previous solve to the totals.
int collect_timing(dy_xpbd_s *x) {
    if (!x->ev_pending) return DY_OK;
    DY_TRY_STATUS(sync_meta(x));
    float a = 0.f, b = 0.f;
    DY_TRY(hipEventSynchronize(x->ev[3]));
    DY_TRY(hipEventElapsedTime(&a, x->ev[0], x->ev[1]));
    DY_TRY(hipEventElapsedTime(&b, x->ev[2], x->ev[3]));
    x->acc.solves += 1;
    x->acc.assign_ms += a;
    x->acc.sweep_ms += b;
    // Algorithmic bytes of the sweep: per contact per iteration 36 B record
    // + 8 B lambda, per dynamic endpoint 80 B read + 52 B written, plus the
    // per-iteration renormalisation (32 B per body).
    x->acc.sweep_bytes += static_cast<double>(x->n_iters) *
                          (44.0 * x->stats.n_constraints + 132.0 * x->stats.dyn_endpoints + 32.0 * x->last_nb);
    x->ev_pending = false;
    return DY_OK;
}

// The constraint-indexed buffers, for c constraints (contacts + joints).
void free_constraint_bufs(dy_xpbd_s *x) {
    void **ptrs[] = {(void **)&x->lam, (void **)&x->state, (void **)&x->adj, (void **)&x->own, (void **)&x->slot,
                     (void **)&x->crank, (void **)&x->pending, (void **)&x->fr0, (void **)&x->fr1,
                     (void **)&x->lrec, (void **)&x->fq, (void **)&x->boff, (void **)&x->bfill, (void **)&x->hist,
                     (void **)&x->rec_id, (void **)&x->cpos, (void **)&x->rseq, (void **)&x->rec_ab,
                     (void **)&x->rec_p0, (void **)&x->rec_p1};
    for (void **p : ptrs)
        if (*p) {
            (void)hipFree(*p);
            *p = nullptr;
        }
    x->ccap = 0;
}

hipError_t alloc_constraint_bufs(dy_xpbd_s *x, uint32_t n_con) {
    free_constraint_bufs(x);
    const size_t c = n_con ? n_con : 1;
    const size_t ncls = c + 66;  // dependency levels <= constraints; colours <= 65
    hipError_t e = hipSuccess;
    auto dalloc = [&e](auto *&p, size_t n) {
        if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&p), n * sizeof(*p));
    };
    dalloc(x->lam, c);
    dalloc(x->state, c);
    dalloc(x->adj, 2 * c);
    dalloc(x->own, 2 * c);
    dalloc(x->slot, 2 * c);
    dalloc(x->crank, 2 * c);
    dalloc(x->pending, c);
    dalloc(x->fr0, c);
    dalloc(x->fr1, c);
    dalloc(x->lrec, c);
    dalloc(x->fq, c);
    dalloc(x->boff, ncls);
    dalloc(x->bfill, ncls);
    dalloc(x->hist, ncls);
    dalloc(x->rec_id, c);
    dalloc(x->cpos, c);
    dalloc(x->rseq, c);
    dalloc(x->rec_ab, c);
    dalloc(x->rec_p0, c);
    dalloc(x->rec_p1, c);
    if (e == hipSuccess) x->ccap = static_cast<uint32_t>(c);
    return e;
}

// Room for nj joints: joint parameters and the constraint buffers for
// max_contacts + nj constraints (grown, never shrunk).
int ensure_joints(dy_xpbd_s *x, uint32_t nj) {
    if (nj <= x->jcap) return DY_OK;
    DY_TRY(hipDeviceSynchronize());  // nothing in flight still uses the buffers
    if (x->jpk) (void)hipFree(x->jpk);
    if (x->jend) (void)hipFree(x->jend);
    x->jpk = nullptr;
    x->jend = nullptr;
    x->jcap = 0;
    DY_TRY(hipMalloc(reinterpret_cast<void **>(&x->jpk), 5 * static_cast<size_t>(nj) * sizeof(float4)));
    DY_TRY(hipMalloc(reinterpret_cast<void **>(&x->jend), static_cast<size_t>(nj) * sizeof(uint2)));
    DY_TRY(alloc_constraint_bufs(x, x->cap + nj));
    x->jcap = nj;
    return DY_OK;
}

void xp_free(dy_xpbd_s *x) {
    for (hipEvent_t &e : x->ev)
        if (e) { (void)hipEventDestroy(e); e = nullptr; }
    if (x->ev_meta) (void)hipEventDestroy(x->ev_meta);
    if (x->h_meta) (void)hipHostFree(x->h_meta);
    free_constraint_bufs(x);
    void *ptrs[] = {x->meta_block, x->jpk, x->jend};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    x->deg.release(); x->off.release(); x->fill.release(); x->scan_tmp.release(); x->ver.release();
    x->used.release(); x->dyn.release(); x->sta.release();
}

// Placement of the metadata words.  Device atomics are served beyond the
// XCD's L2 and their latency depends on the address: of two lines 8 KiB
// apart, one is typically served about a third faster than the other from a
// given XCD, and which one depends on the physical page.  The assignment's
// queue counters (kMQHead, kMQTail, kMQDone) take an atomic per wave and
// claim, so the metadata goes into whichever half of a 16 KiB block answers
// faster from the assignment's XCD, measured once at creation by chains of
// dependent atomics on each half's queue-counter line.
constexpr uint32_t kMetaHalf = 8192u, kMetaBlock = 2u * kMetaHalf;
constexpr uint32_t kMetaProbeChain = 64u;
static_assert(kMWords * sizeof(uint32_t) <= kMetaHalf, "metadata fits half the block");

__global__ void meta_place_probe(uint32_t *block, uint32_t *elect, uint64_t *ticks) {
    if (threadIdx.x != 0 || xcc_id() != kSweepXcc) return;
    if (atomicCAS(elect, 0u, 1u) != 0u) return;  // one workgroup on that XCD measures
    for (uint32_t h = 0; h < 2u; ++h) {
        uint32_t *p = block + h * (kMetaHalf / 4u) + kMQHead;
        uint32_t v = 0;
        const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
        for (uint32_t i = 0; i < kMetaProbeChain; ++i) v = atomicAdd(p + (v >> 31), 1u);  // v < 2^31: dependent chain
        ticks[h] = __builtin_amdgcn_s_memrealtime() - t0 + (v >> 31);
    }
}

// Chooses x->meta inside x->meta_block (allocated, kMetaBlock bytes) and
// zeroes the block.  Synchronous (creation only).
hipError_t place_meta(dy_xpbd_s *x, hipStream_t st) {
    uint32_t *scratch = nullptr;  // [0] election, [2..5] two 64-bit tick counts
    hipError_t e = hipMalloc(reinterpret_cast<void **>(&scratch), 8 * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemsetAsync(x->meta_block, 0, kMetaBlock, st);
    if (e == hipSuccess) e = hipMemsetAsync(scratch, 0, 8 * sizeof(uint32_t), st);
    uint64_t ticks[2] = {0, 0};
    if (e == hipSuccess) {
        meta_place_probe<<<8, 64, 0, st>>>(reinterpret_cast<uint32_t *>(x->meta_block), scratch,
                                            reinterpret_cast<uint64_t *>(scratch + 2));
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(ticks, scratch + 2, sizeof(ticks), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (scratch) (void)hipFree(scratch);
    // the second half only when it measured clearly faster (both zero: no measurement)
    const bool second = ticks[0] > 0 && ticks[1] > 0 && ticks[1] * 10u < ticks[0] * 9u;
    x->meta = reinterpret_cast<uint32_t *>(x->meta_block + (second ? kMetaHalf : 0u));
    if (e == hipSuccess) e = hipMemsetAsync(x->meta_block, 0, kMetaBlock, st);
    return e;
}

int env_int(const char *name, int dflt) {
    const char *e = std::getenv(name);
    return e ? std::atoi(e) : dflt;
}

// Launch limits of the grid-synchronised kernels; DYPHUR_{ASSIGN,SWEEP}_
// {BLOCKS,THREADS}, DYPHUR_SWEEP_MODE (2 = dataflow sweep, the default; 3 =
// dataflow with a cursor per lane; 0 = class phases on one XCD; 1 = class
// phases over the whole grid), DYPHUR_SWEEP_PAIR (0: contact-only dataflow
// sweeps with one lane per record instead of a pair), DYPHUR_SWEEP_PAIR_THREADS
// and DYPHUR_SOLVER_PATH (1 = one launch per round / per class) override the
// launch shapes for measurement.
int query_coop(dy_xpbd_s *x) {
    int dev = 0, coop = 0;
    DY_TRY(hipGetDevice(&dev));
    DY_TRY(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
    DY_TRY(hipDeviceGetAttribute(&x->device_cus, hipDeviceAttributeMultiprocessorCount, dev));
    x->sweep_blocks_override = env_int("DYPHUR_SWEEP_BLOCKS", 0);
    x->assign_blocks_override = env_int("DYPHUR_ASSIGN_BLOCKS", 0);
    x->path = env_int("DYPHUR_SOLVER_PATH", 0);
    {
        const int m = env_int("DYPHUR_SWEEP_MODE", 2);
        x->sweep_mode = (m == 0 || m == 1 || m == 3) ? m : 2;
    }
    const int at = env_int("DYPHUR_ASSIGN_THREADS", 256), sw = env_int("DYPHUR_SWEEP_THREADS", 256);
    x->assign_threads = (at == 64 || at == 128 || at == 256 || at == 512 || at == 1024) ? at : 256;
    x->sweep_threads = (sw == 64 || sw == 128 || sw == 256 || sw == 512 || sw == 1024) ? sw : 256;
    x->sweep_pair = env_int("DYPHUR_SWEEP_PAIR", 1) != 0 ? 1 : 0;
    {
        const int pt = env_int("DYPHUR_SWEEP_PAIR_THREADS", 512);
        x->pair_threads = (pt == 128 || pt == 256 || pt == 512 || pt == 1024) ? pt : 512;
    }
    if (!coop) return DY_OK;
    int per_cu = 0;
    DY_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(assign_lists),
                                                        x->assign_threads, 0));
    x->coop_assign_max = per_cu * x->device_cus;
    const void *fns[4][2] = {
        {reinterpret_cast<const void *>(sweep_persistent<false>), reinterpret_cast<const void *>(sweep_persistent<true>)},
        {reinterpret_cast<const void *>(sweep_xcd<false>), reinterpret_cast<const void *>(sweep_xcd<true>)},
        {reinterpret_cast<const void *>(sweep_flow<false>), reinterpret_cast<const void *>(sweep_flow<true>)},
        {reinterpret_cast<const void *>(sweep_flow_lanes<false>),
         reinterpret_cast<const void *>(sweep_flow_lanes<true>)}};
    int *outs[4] = {x->coop_sweep_max, x->coop_xcd_max, x->coop_flow_max, x->coop_lanes_max};
    for (int k = 0; k < 4; ++k)
        for (int j = 0; j < 2; ++j) {
            DY_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fns[k][j], x->sweep_threads, 0));
            outs[k][j] = per_cu * x->device_cus;
        }
    DY_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(joint_motors_xcd),
                                                        x->sweep_threads, 0));
    x->coop_motor_max = per_cu * x->device_cus;
    DY_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(sweep_flow_pair),
                                                        x->pair_threads, 0));
    x->coop_pair_max = per_cu * x->device_cus;
    return DY_OK;
}

// Blocks for a cooperative kernel: the override, else device_cus / 8
// workgroups of 256 threads' worth, clamped to what can be co-resident.
uint32_t coop_blocks(const dy_xpbd_s *x, int override_blocks, int threads, int max_blocks) {
    int b = override_blocks > 0 ? override_blocks : (x->device_cus / 8) * 256 / threads;
    if (b < 1) b = 1;
    if (b > max_blocks) b = max_blocks;
    return static_cast<uint32_t>(b);
}

// Launch-per-round assignment: the host drives the rounds and decides the
// fallback to the reference order, reading the metadata every 16 rounds.
int assign_by_rounds(dy_xpbd_s *x, hipStream_t st, const Cons &cv, const float *im) {
    uint32_t h[kMWords];
    for (int pass = 0; pass < 2; ++pass) {
        DY_TRY(hipMemcpyAsync(h, x->meta, sizeof(h), hipMemcpyDeviceToHost, st));
        DY_TRY(hipStreamSynchronize(st));
        const uint32_t nc = h[kMNc];
        const int serial = static_cast<int>(h[kMSerial]);
        const uint32_t max_rounds = serial ? nc + 1u : 256u;
        const uint32_t grid = grid_for(nc);
        bool done = false, exceeded = false;
        for (uint32_t r = 1; !done; ++r) {
            if (r > max_rounds) {
                exceeded = true;
                break;
            }
            round_control<<<1, 1, 0, st>>>(x->meta, r % 3u, 0);
            DY_LAUNCH_CHECK();
            assign_lists_round<<<grid, kSolveThreads, 0, st>>>(x->lrec, r, x->pending, x->used.p, x->state, x->fr0,
                                                               x->fr1, x->meta);
            DY_LAUNCH_CHECK();
            if (r % 16u == 0u || r == max_rounds) {
                DY_TRY(hipMemcpyAsync(h, x->meta, sizeof(h), hipMemcpyDeviceToHost, st));
                DY_TRY(hipStreamSynchronize(st));
                if (h[kMErr] & ~kErrJoint) return DY_ERR_INTERNAL;  // a rejected joint is reported by the queries
                done = h[kMAssigned] == nc;
            }
        }
        DY_TRY(hipMemcpyAsync(h, x->meta, sizeof(h), hipMemcpyDeviceToHost, st));
        DY_TRY(hipStreamSynchronize(st));
        if (!serial && (exceeded || h[kMOverflow])) {
            // The reference's serial path: re-rank in contact order and rerun.
            round_control<<<1, 1, 0, st>>>(x->meta, 0, exceeded ? 2 : 1);
            DY_LAUNCH_CHECK();
            list_rank<<<grid_for(2u * nc), kSolveThreads, 0, st>>>(x->adj, x->own, x->off.p, x->deg.p, x->meta,
                                                                   x->slot, x->crank);
            DY_LAUNCH_CHECK();
            list_link<<<grid_for(nc), kSolveThreads, 0, st>>>(cv, x->meta, im, x->off.p, x->deg.p, x->slot, x->crank,
                                                              x->lrec, x->pending, x->fr0, x->meta + kMFcnt, x->fq);
            DY_LAUNCH_CHECK();
            continue;
        }
        if (h[kMAssigned] != nc) return DY_ERR_INTERNAL;
        // Class offsets from a histogram of the classes.
        const uint32_t k = serial ? h[kMRounds] : (nc ? h[kMMaxCls] + 1u : 0u);
        DY_TRY(hipMemsetAsync(x->hist, 0, (k + 1) * sizeof(uint32_t), st));
        DY_TRY(hipMemsetAsync(x->bfill, 0, (k + 1) * sizeof(uint32_t), st));
        class_count<<<grid_for(nc), kSolveThreads, 0, st>>>(x->meta, k, x->state, x->hist);
        DY_LAUNCH_CHECK();
        DY_TRY_STATUS(scan_exclusive(st, x->hist, x->boff, k + 1, nullptr, x->scan_tmp.p));
        DY_TRY(hipMemcpyAsync(x->meta + kMNcls, &k, sizeof(k), hipMemcpyHostToDevice, st));
        DY_TRY(hipStreamSynchronize(st));
        return DY_OK;
    }
    return DY_ERR_INTERNAL;
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
    hipError_t e = alloc_constraint_bufs(x, max_contacts);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&x->meta_block), kMetaBlock);
    if (e == hipSuccess) e = place_meta(x, hs(s));
    if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void **>(&x->h_meta), kMWords * sizeof(uint32_t));
    if (e == hipSuccess) e = hipEventCreateWithFlags(&x->ev_meta, hipEventDisableTiming);
    if (e == hipSuccess) e = hipStreamSynchronize(hs(s));
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

int dy_xpbd_solve(dy_xpbd_t x, dy_stream_t s, const dy_contact_view *cv, const dy_joint_view *jv,
                  const dy_body_view *bv, float dt) {
    if (!x || !cv || !bv || !cv->n) return DY_ERR_INVALID_ARG;
    const uint32_t nj = jv ? jv->n : 0u;
    if (nj && (!jv->body_parent || !jv->body_child || !jv->type || !jv->anchor_px || !jv->anchor_py ||
               !jv->anchor_pz || !jv->anchor_cx || !jv->anchor_cy || !jv->anchor_cz || !jv->axis_px ||
               !jv->axis_py || !jv->axis_pz || !jv->limit_lo || !jv->limit_hi || !jv->target_pos ||
               !jv->target_vel || !jv->stiffness || !jv->damping || !jv->compliance_pos || !jv->compliance_ang))
        return DY_ERR_INVALID_ARG;
    hipStream_t st = hs(s);
    // Errors of earlier solves surface here once their metadata has landed.
    if (x->timing) {
        DY_TRY_STATUS(collect_timing(x));
    } else if (x->meta_pending && hipEventQuery(x->ev_meta) == hipSuccess) {
        DY_TRY_STATUS(sync_meta(x));
    }
    DY_TRY_STATUS(ensure_joints(x, nj));
    const uint32_t nb = bv->n, cap = x->cap, ncap = cap + nj;
    DY_TRY_STATUS(x->deg.ensure(nb));
    DY_TRY_STATUS(x->off.ensure(nb));
    DY_TRY_STATUS(x->fill.ensure(nb));
    DY_TRY_STATUS(x->used.ensure(nb));
    DY_TRY_STATUS(x->ver.ensure(nb));
    DY_TRY_STATUS(x->dyn.ensure(4 * static_cast<size_t>(nb)));
    DY_TRY_STATUS(x->sta.ensure(2 * static_cast<size_t>(nb)));
    DY_TRY_STATUS(x->scan_tmp.ensure(scan_tmp_words(nb > ncap + 66u ? nb : ncap + 66u)));
    x->last_nb = nb;
    uint32_t top = nb > ncap ? nb : ncap;
    if (top < kMWords + kMaxColours + 1u) top = kMWords + kMaxColours + 1u;
    solve_prep<<<grid_for(top), kSolveThreads, 0, st>>>(*bv, cv->n, cap, nj, x->force_serial, x->deg.p, x->fill.p,
                                                         x->used.p, x->ver.p, x->lam, x->hist, x->meta, x->dyn.p,
                                                         x->sta.p);
    DY_LAUNCH_CHECK();
    if (nj) {
        joint_pack<<<ceil_div(nj, 256), 256, 0, st>>>(*jv, nb, x->jpk, x->jend, x->meta);
        DY_LAUNCH_CHECK();
        joint_commit<<<1, 64, 0, st>>>(x->force_serial, x->meta);
        DY_LAUNCH_CHECK();
    }
    const Cons cs{*cv, x->jend};
    solve_degree<<<grid_for(ncap), kSolveThreads, 0, st>>>(cs, x->meta, bv->inv_mass, x->deg.p);
    DY_LAUNCH_CHECK();
    DY_TRY_STATUS(scan_exclusive(st, x->deg.p, x->off.p, nb, x->meta + kMDyn, x->scan_tmp.p));
    solve_fill<<<grid_for(ncap), kSolveThreads, 0, st>>>(cs, x->meta, bv->inv_mass, x->off.p, x->fill.p, x->adj,
                                                          x->own);
    DY_LAUNCH_CHECK();
    list_rank<<<grid_for(2u * ncap), kSolveThreads, 0, st>>>(x->adj, x->own, x->off.p, x->deg.p, x->meta, x->slot,
                                                              x->crank);
    DY_LAUNCH_CHECK();
    list_link<<<grid_for(ncap), kSolveThreads, 0, st>>>(cs, x->meta, bv->inv_mass, x->off.p, x->deg.p, x->slot,
                                                         x->crank, x->lrec, x->pending, x->fr0, x->meta + kMFcnt,
                                                         x->fq);
    DY_LAUNCH_CHECK();

    const int J = nj ? 1 : 0;
    bool coop = x->path == 0 && !x->coop_refused && x->coop_assign_max > 0 && x->coop_sweep_max[J] > 0;
    if (x->timing) DY_TRY(hipEventRecord(x->ev[0], st));
    if (coop) {
        AssignBufs B{cs, bv->inv_mass, x->adj, x->own, x->off.p, x->deg.p, x->slot, x->crank, x->lrec, x->fq,
                     x->pending, x->used.p, x->state, x->fr0, x->fr1, x->meta, x->hist, x->boff, x->bfill, nb};
        void *args[] = {&B};
        // Participants on one XCD out of a launch with as many on every XCD:
        // by default 16 workgroups of 256 threads per XCD on a 256-CU device,
        // half the solver's usual share.  On the 32k-body frame 16 beat 8, 24
        // and 32 (fewer workgroups meet at each round barrier).
        const int assign_dflt = (x->device_cus / 16) * 256 / x->assign_threads;
        uint32_t blocks = 8u * coop_blocks(x, x->assign_blocks_override > 0 ? x->assign_blocks_override : assign_dflt,
                                           x->assign_threads, x->coop_assign_max);
        if (blocks > static_cast<uint32_t>(x->coop_assign_max)) blocks = static_cast<uint32_t>(x->coop_assign_max);
        if (try_coop_launch(reinterpret_cast<const void *>(assign_lists), dim3(blocks),
                            dim3(static_cast<uint32_t>(x->assign_threads)), args, st) != hipSuccess) {
            x->coop_refused = true;
            coop = false;
        }
    }
    if (!coop) DY_TRY_STATUS(assign_by_rounds(x, st, cs, bv->inv_mass));
    if (x->timing) DY_TRY(hipEventRecord(x->ev[1], st));

    // Class-sorted constraint records.
    CRec rec{x->rec_ab, x->rec_p0, x->rec_p1};
    // workgroups of 1 024: one class-count atomic per class present per 1 024 constraints
    class_gather<<<ceil_div(grid_for(ncap), 4u), 1024, 0, st>>>(cs, x->meta, x->state, x->boff, x->bfill, rec,
                                                                 x->rec_id, x->cpos, x->rseq);
    DY_LAUNCH_CHECK();
    const int *flow_max = x->sweep_mode == 3 ? x->coop_lanes_max : x->coop_flow_max;
    const bool flow = coop && (x->sweep_mode == 2 || x->sweep_mode == 3) && flow_max[J] > 0;
    if (flow) {
        flow_seq<<<grid_for(2u * ncap), kSolveThreads, 0, st>>>(x->adj, x->own, x->off.p, x->deg.p, x->state,
                                                                 x->cpos, x->meta, x->rseq);
        DY_LAUNCH_CHECK();
    }
    if (nj) {
        bool motors = false;
        if (coop && x->coop_motor_max > 0) {
            // P participants on one XCD out of a launch of P per XCD, as sweep_xcd.
            uint32_t blocks = 8u * coop_blocks(x, 0, x->sweep_threads, x->coop_motor_max);
            if (blocks > static_cast<uint32_t>(x->coop_motor_max)) blocks = static_cast<uint32_t>(x->coop_motor_max);
            const uint32_t *boffv = x->boff;
            float4 *dynv = x->dyn.p;
            const float4 *stav = x->sta.p, *jpkv0 = x->jpk;
            float dtv = dt;
            uint32_t *metav = x->meta;
            void *args[] = {&boffv, &rec, &dynv, &stav, &jpkv0, &dtv, &metav};
            motors = try_coop_launch(reinterpret_cast<const void *>(joint_motors_xcd), dim3(blocks),
                                     dim3(static_cast<uint32_t>(x->sweep_threads)), args, st) == hipSuccess;
            if (!motors) {
                x->coop_refused = true;
                coop = false;
            }
        }
        if (!motors) {
            joint_motors<<<1, 1024, 0, st>>>(x->boff, rec, x->dyn.p, x->sta.p, x->jpk, dt, x->meta);
            DY_LAUNCH_CHECK();
        }
    }

    const uint32_t iters = static_cast<uint32_t>(x->n_iters);
    const float4 *jpkv = x->jpk;
    if (x->timing) DY_TRY(hipEventRecord(x->ev[2], st));
    bool swept = iters == 0;
    if (!swept && coop && flow) {
        // contact-only: a pair of lanes per record, one workgroup of
        // pair_threads per CU of the XCD by default (the measured best on the
        // 32k-body frame: 32 x 512 against 16..48 x 256..1024)
        const bool pair = x->sweep_mode == 2 && !J && x->sweep_pair && x->coop_pair_max > 0;
        const int threads = pair ? x->pair_threads : x->sweep_threads;
        const int cap = pair ? x->coop_pair_max : flow_max[J];
        uint32_t blocks = 8u * coop_blocks(x, x->sweep_blocks_override > 0 ? x->sweep_blocks_override
                                              : pair                        ? x->device_cus / 8
                                                                            : 0,
                                           threads, cap);
        if (blocks > static_cast<uint32_t>(cap)) blocks = static_cast<uint32_t>(cap);
        uint32_t iv = iters;
        const uint4 *rseqv = x->rseq;
        float *lamv = x->lam;
        float4 *dynv = x->dyn.p;
        const float4 *stav = x->sta.p;
        uint32_t *verv = x->ver.p, *metav = x->meta;
        float muv = x->mu, dtv = dt;
        void *args[] = {&iv, &rec, &rseqv, &lamv, &dynv, &stav, &verv, &muv, &dtv, &metav, &jpkv};
        const void *fn;
        if (x->sweep_mode == 3)
            fn = J ? reinterpret_cast<const void *>(sweep_flow_lanes<true>)
                   : reinterpret_cast<const void *>(sweep_flow_lanes<false>);
        else if (pair)
            fn = reinterpret_cast<const void *>(sweep_flow_pair);
        else
            fn = J ? reinterpret_cast<const void *>(sweep_flow<true>) : reinterpret_cast<const void *>(sweep_flow<false>);
        swept = try_coop_launch(fn, dim3(blocks), dim3(static_cast<uint32_t>(threads)), args, st) == hipSuccess;
    } else if (!swept && coop) {
        // sweep_xcd: P participants on one XCD out of a launch of P per XCD
        // (8 XCDs on gfx950); sweep_persistent: P workgroups over the chip.
        const bool xcd = x->sweep_mode == 0 && x->coop_xcd_max[J] > 0;
        uint32_t blocks = coop_blocks(x, x->sweep_blocks_override, x->sweep_threads, x->coop_sweep_max[J]);
        if (xcd) {
            blocks *= 8u;
            if (blocks > static_cast<uint32_t>(x->coop_xcd_max[J])) blocks = static_cast<uint32_t>(x->coop_xcd_max[J]);
        }
        const uint32_t *boffv = x->boff;
        uint32_t iv = iters;
        float *lamv = x->lam;
        float4 *dynv = x->dyn.p;
        const float4 *stav = x->sta.p;
        float muv = x->mu, dtv = dt;
        uint32_t *metav = x->meta;
        void *args[] = {&boffv, &iv, &rec, &lamv, &dynv, &stav, &muv, &dtv, &metav, &jpkv};
        const void *fn;
        if (xcd)
            fn = J ? reinterpret_cast<const void *>(sweep_xcd<true>) : reinterpret_cast<const void *>(sweep_xcd<false>);
        else
            fn = J ? reinterpret_cast<const void *>(sweep_persistent<true>)
                   : reinterpret_cast<const void *>(sweep_persistent<false>);
        swept = try_coop_launch(fn, dim3(blocks), dim3(static_cast<uint32_t>(x->sweep_threads)), args, st) == hipSuccess;
    }
    if (!swept) {
        if (coop) x->coop_refused = true;
        // Per-class launches; the class count is read back first.
        uint32_t k = 0;
        DY_TRY(hipMemcpyAsync(&k, x->meta + kMNcls, sizeof(k), hipMemcpyDeviceToHost, st));
        DY_TRY(hipStreamSynchronize(st));
        const uint32_t blocks = grid_for(ncap);
        for (uint32_t it = 0; it < iters; ++it)
            for (uint32_t c = 0; c < k; ++c) {
                if (J)
                    sweep_class<true><<<blocks, kSolveThreads, 0, st>>>(x->boff, c, it, rec, x->lam, x->dyn.p,
                                                                        x->sta.p, x->mu, dt, jpkv);
                else
                    sweep_class<false><<<blocks, kSolveThreads, 0, st>>>(x->boff, c, it, rec, x->lam, x->dyn.p,
                                                                         x->sta.p, x->mu, dt, jpkv);
                DY_LAUNCH_CHECK();
            }
    }
    if (x->timing) {
        DY_TRY(hipEventRecord(x->ev[3], st));
        x->ev_pending = true;
    }
    scatter_bodies<<<ceil_div(nb, 256), 256, 0, st>>>(*bv, x->dyn.p, iters);
    DY_LAUNCH_CHECK();
    DY_TRY(hipMemcpyAsync(x->h_meta, x->meta, kMWords * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    DY_TRY(hipEventRecord(x->ev_meta, st));
    x->meta_pending = true;
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
    const int st = sync_meta(x);
    *out = x->stats;
    return st;
}

int dy_xpbd_download_classes(dy_xpbd_t x, dy_stream_t s, uint32_t *out, uint32_t n) {
    if (!x || (!out && n)) return DY_ERR_INVALID_ARG;
    if (n > x->ccap) return DY_ERR_INVALID_ARG;
    if (n == 0) return DY_OK;
    unpack_classes<<<ceil_div(n, 256), 256, 0, hs(s)>>>(n, x->state, x->rec_id);
    DY_LAUNCH_CHECK();
    DY_TRY(hipMemcpyAsync(out, x->rec_id, n * sizeof(uint32_t), hipMemcpyDeviceToHost, hs(s)));
    DY_TRY(hipStreamSynchronize(hs(s)));
    return DY_OK;
}

}  // extern "C"
