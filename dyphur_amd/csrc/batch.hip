// batch.hip — thousands of independent scenes stepped in one launch (gfx950).
//
// One workgroup per scene, persistent over the frames of a launch.  Every
// stage of the reference frame (examples/stress_test_blocks/main.cpp:264-273:
// integrate -> build_and_query -> sort_pairs -> narrowphase -> solve) runs
// inside the workgroup with __syncthreads() as its only barrier, on the
// scene's own segment of the concatenated SoA body arrays.  Scenes are
// independent, so no workgroup ever waits for another: a launch of F frames
// has no grid-wide synchronisation, and the hardware dispatcher hands the
// next scene (largest first) to whichever CU frees a slot.
//
// Per scene and substep h = dt / S:
//   1. integrate (core/src/integrator.cpp:17-70);
//   2. leaf boxes rounded to binary16 (core/src/broadphase.cpp:215-251), held
//      in LDS as the floats they represent;
//   3. candidate pairs.  The reference's pair set is every pair of leaves
//      whose binary16 boxes overlap, touching included (broadphase.cpp:290-345;
//      internal boxes are exact unions of their leaves, so the tree walk
//      prunes nothing that overlaps).  At the scene sizes batched here the
//      workgroup tests the pairs straight from LDS: a count pass over each
//      body's row of partners b > a, a scan of the row counts, an emit pass.
//      Rows in body order with partners ascending are exactly the canonical
//      (a, b) order of sort_pairs (broadphase.cpp:348-359); pairs beyond the
//      scene's capacity are dropped in that order and the true count kept;
//   4. narrowphase (core/src/narrowphase.cpp:920-1127): one pair per lane,
//      a block scan of the hit flags, contacts stored in pair order;
//   5. XPBD solve (core/src/xpbd_solver.cpp:409-733): each contact's
//      predecessor on each dynamic body is found from per-body contact lists
//      (ascending contact index for the reference contact order; descending
//      Jones-Plassmann priority for the coloured path), classes are assigned
//      in rounds (a contact is ready once its predecessors were assigned in
//      earlier rounds: the dependency level, or the colour from the
//      per-body masks), then I iterations sweep the classes in order, each
//      followed by the renormalisation of every quaternion.
#include <algorithm>
#include <numeric>
#include <vector>

#include "dy_internal.hpp"
#include "pair_contact.hpp"

namespace dy {

constexpr uint32_t kBatchMaxBodies = 2048;  // LDS boxes: 24 B per body
constexpr uint32_t kBatchSerialMax = 256;   // xpbd_solver.hpp:29 kSerialThreshold
constexpr uint32_t kBatchMaxRounds = 256;   // xpbd_solver.cpp:640
constexpr uint32_t kNotAssigned = 0xffffffffu;
constexpr uint32_t kBatchErrUnsupported = 1u, kBatchErrLists = 2u;

enum : uint32_t {
    kStPairs = 0, kStContacts, kStClasses, kStSerial, kStRounds, kStOverflow, kStError, kStSubsteps,
    kStWords
};

struct BatchArgs {
    dy_body_view bv;       // every scene's bodies, concatenated
    dy_shape_view sv;      // shape catalog; bv.shape holds catalog indices
    dy_contact_view cv;    // contact store, concatenated (cv.n unused)
    dy_contact_pair *pairs;
    const uint32_t *order; // launch slot -> scene
    const uint32_t *body_off, *pair_off, *contact_off;
    const uint64_t *scratch_off;  // 32-bit words, even
    uint32_t *scratch;
    uint32_t *stats;       // kStWords per scene
    uint32_t max_n;
    float gx, gy, gz, h;
    int n_sub, n_frames, n_iters, force_serial;
    float mu;
};

// Per-scene scratch, carved from the scene's slice: 6n + 10C + 2 words.
struct SceneScratch {
    uint64_t *used;                 // [n] colour masks
    uint32_t *row, *deg, *loff, *fill;  // [n]
    float *lam;                     // [C]
    int32_t *pred_a, *pred_b;       // [C]
    uint32_t *rnd, *cls, *order;    // [C]
    uint32_t *list;                 // [2C]
    uint32_t *ccnt, *coff;          // [C + 1]
};

__host__ __device__ inline uint64_t scene_scratch_words(uint32_t n, uint32_t c) {
    const uint64_t w = 6ull * n + 10ull * c + 2ull;
    return (w + 1ull) & ~1ull;
}

__device__ __forceinline__ SceneScratch carve(uint32_t *p, uint32_t n, uint32_t c) {
    SceneScratch s;
    s.used = reinterpret_cast<uint64_t *>(p); p += 2 * n;
    s.row = p; p += n;
    s.deg = p; p += n;
    s.loff = p; p += n;
    s.fill = p; p += n;
    s.lam = reinterpret_cast<float *>(p); p += c;
    s.pred_a = reinterpret_cast<int32_t *>(p); p += c;
    s.pred_b = reinterpret_cast<int32_t *>(p); p += c;
    s.rnd = p; p += c;
    s.cls = p; p += c;
    s.order = p; p += c;
    s.list = p; p += 2 * c;
    s.ccnt = p; p += c + 1;
    s.coff = p;
    return s;
}

__device__ __forceinline__ dy_body_view scene_bodies(const dy_body_view &g, uint32_t o, uint32_t n) {
    dy_body_view v = g;
    v.pos_x += o; v.pos_y += o; v.pos_z += o;
    v.rot_w += o; v.rot_x += o; v.rot_y += o; v.rot_z += o;
    v.vel_x += o; v.vel_y += o; v.vel_z += o;
    v.ang_x += o; v.ang_y += o; v.ang_z += o;
    v.inv_mass += o;
    v.iI_xx += o; v.iI_yy += o; v.iI_zz += o;
    v.iI_xy += o; v.iI_xz += o; v.iI_yz += o;
    v.shape += o; v.flags += o;
    v.n = n;
    return v;
}

__device__ __forceinline__ dy_contact_view scene_contacts(const dy_contact_view &g, uint32_t o, uint32_t c) {
    dy_contact_view v = g;
    v.body_a += o; v.body_b += o;
    v.pos_x += o; v.pos_y += o; v.pos_z += o;
    v.norm_x += o; v.norm_y += o; v.norm_z += o;
    v.depth += o;
    v.capacity = c;
    return v;
}

__device__ __forceinline__ BodyState load_body(const dy_body_view &v, uint32_t i) {
    BodyState b;
    b.px = v.pos_x[i]; b.py = v.pos_y[i]; b.pz = v.pos_z[i];
    b.qw = v.rot_w[i]; b.qx = v.rot_x[i]; b.qy = v.rot_y[i]; b.qz = v.rot_z[i];
    b.vx = v.vel_x[i]; b.vy = v.vel_y[i]; b.vz = v.vel_z[i];
    b.wx = v.ang_x[i]; b.wy = v.ang_y[i]; b.wz = v.ang_z[i];
    b.im = v.inv_mass[i];
    b.ixx = v.iI_xx[i]; b.iyy = v.iI_yy[i]; b.izz = v.iI_zz[i];
    b.ixy = v.iI_xy[i]; b.ixz = v.iI_xz[i]; b.iyz = v.iI_yz[i];
    return b;
}

__device__ __forceinline__ void store_body(const dy_body_view &v, uint32_t i, const BodyState &b) {
    v.pos_x[i] = b.px; v.pos_y[i] = b.py; v.pos_z[i] = b.pz;
    v.rot_w[i] = b.qw; v.rot_x[i] = b.qx; v.rot_y[i] = b.qy; v.rot_z[i] = b.qz;
    v.vel_x[i] = b.vx; v.vel_y[i] = b.vy; v.vel_z[i] = b.vz;
    v.ang_x[i] = b.wx; v.ang_y[i] = b.wy; v.ang_z[i] = b.wz;
}

// Exclusive scan of data[0, n) in place by the whole workgroup; returns the
// total.  Each entry is rewritten by the thread that read it: callers
// __syncthreads() before other threads read the result.
__device__ uint32_t scan_in_place(uint32_t *data, uint32_t n, uint32_t *wsum) {
    uint32_t carry = 0;
    for (uint32_t base = 0; base < n; base += blockDim.x) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t v = i < n ? data[i] : 0u;
        uint32_t tot = 0;
        const uint32_t ex = block_exclusive_scan(v, wsum, tot);
        if (i < n) data[i] = carry + ex;
        carry += tot;
    }
    return carry;
}

__device__ __forceinline__ uint32_t block_sum(uint32_t v, uint32_t *wsum) {
    uint32_t tot = 0;
    (void)block_exclusive_scan(v, wsum, tot);
    return tot;
}

// Workgroup sum of v and OR of f in one exchange (two barriers).  `wsum`
// needs 12 words; every thread gets both totals.
__device__ __forceinline__ uint32_t block_sum_or(uint32_t v, uint32_t f, uint32_t *wsum, uint32_t &f_all) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        v += __shfl_xor(v, d, 64);
        f |= __shfl_xor(f, d, 64);
    }
    const uint32_t wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if (lane_id() == 0) {
        wsum[wave] = v;
        wsum[8 + wave] = f;
    }
    __syncthreads();
    uint32_t sv = 0, so = 0;
    for (uint32_t w = 0; w < nw; ++w) {
        sv += wsum[w];
        so |= wsum[8 + w];
    }
    __syncthreads();
    f_all = so;
    return sv;
}

// Reference overlap test on two boxes (broadphase.cpp:322-325): separated on
// some axis, else touching or overlapping.
struct LdsBoxes { float *mnx, *mny, *mnz, *mxx, *mxy, *mxz; };

__device__ __forceinline__ bool touch(const float a[6], const LdsBoxes &L, uint32_t b) {
    return !(L.mnx[b] > a[3] || L.mxx[b] < a[0] || L.mny[b] > a[4] || L.mxy[b] < a[1] ||
             L.mnz[b] > a[5] || L.mxz[b] < a[2]);
}

// Sort one body's contact list and record each contact's predecessor on that
// body: ascending contact index (reference order) or descending colouring
// priority (xpbd_solver.cpp:31-34,625-695).
__device__ __forceinline__ void body_predecessors(uint32_t x, const SceneScratch &S, const uint32_t *ca,
                                                  bool jp) {
    uint32_t *L = S.list + S.loff[x];
    const uint32_t d = S.deg[x];
    for (uint32_t i = 1; i < d; ++i) {
        const uint32_t k = L[i];
        const uint64_t kk = jp ? ~jp_priority(k) : k;
        uint32_t j = i;
        while (j > 0) {
            const uint32_t o = L[j - 1];
            const uint64_t ko = jp ? ~jp_priority(o) : o;
            if (ko <= kk) break;
            L[j] = o;
            --j;
        }
        L[j] = k;
    }
    for (uint32_t p = 0; p < d; ++p) {
        const uint32_t k = L[p];
        const int32_t pred = p ? static_cast<int32_t>(L[p - 1]) : -1;
        if (ca[k] == x) S.pred_a[k] = pred;
        else S.pred_b[k] = pred;
    }
}

// Class assignment in rounds.  Returns the rounds run; `remaining` holds
// the contacts left unassigned (non-zero only for the coloured path at 256
// rounds or after a colour overflow, `over`).  Every exit condition is a
// workgroup-wide total, so all threads leave the loop together.
__device__ uint32_t assign_rounds(uint32_t nc, bool jp, const SceneScratch &S, const dy_contact_view &cv,
                                  const float *im, uint32_t *wsum, uint32_t &remaining, uint32_t &over,
                                  uint32_t &err) {
    uint32_t r = 0;
    remaining = nc;
    over = 0;
    while (remaining > 0) {
        if (jp && r >= kBatchMaxRounds) break;
        uint32_t mine = 0, my_over = 0;
        for (uint32_t k = threadIdx.x; k < nc; k += blockDim.x) {
            if (S.rnd[k] != kNotAssigned) continue;
            const int32_t pa = S.pred_a[k], pb = S.pred_b[k];
            // A predecessor assigned in this round reads as r or as unassigned.
            if (pa >= 0 && __hip_atomic_load(&S.rnd[pa], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) >= r)
                continue;
            if (pb >= 0 && __hip_atomic_load(&S.rnd[pb], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) >= r)
                continue;
            if (jp) {
                const uint32_t a = cv.body_a[k], b = cv.body_b[k];
                const bool da = im[a] != 0.f, db = im[b] != 0.f;
                const uint64_t used = (da ? S.used[a] : 0ull) | (db ? S.used[b] : 0ull);
                if (used == ~0ull) {  // a 65th colour: the reference falls back to run_serial
                    my_over = 1;
                    continue;
                }
                const uint32_t col = static_cast<uint32_t>(__builtin_ctzll(~used));
                const uint64_t bit = 1ull << col;
                if (da) S.used[a] |= bit;
                if (db) S.used[b] |= bit;
                S.cls[k] = col;
            } else {
                S.cls[k] = r;
            }
            __hip_atomic_store(&S.rnd[k], r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            ++mine;
        }
        uint32_t any_over = 0;
        const uint32_t got = block_sum_or(mine, my_over, wsum, any_over);
        remaining -= got;
        ++r;
        if (jp && any_over) {
            over = 1;
            break;
        }
        if (got == 0 && remaining > 0) {  // every round frees at least the first contact of the order
            err |= kBatchErrLists;
            break;
        }
    }
    return r;
}

template <int BS>
__global__ __launch_bounds__(BS) void batch_frames(BatchArgs A) {
    extern __shared__ float lds_box[];
    __shared__ uint32_t wsum[17];
    __shared__ uint32_t s_flags[4];  // [1] error bits, [2] class count
    const uint32_t t = threadIdx.x;
    const uint32_t s = A.order[blockIdx.x];
    const uint32_t b0 = A.body_off[s], n = A.body_off[s + 1] - b0;
    const uint32_t P = A.pair_off[s + 1] - A.pair_off[s];
    const uint32_t C = A.contact_off[s + 1] - A.contact_off[s];
    const dy_body_view bv = scene_bodies(A.bv, b0, n);
    const dy_contact_view cv = scene_contacts(A.cv, A.contact_off[s], C);
    dy_contact_pair *pairs = A.pairs + A.pair_off[s];
    const SceneScratch S = carve(A.scratch + A.scratch_off[s], n, C);
    uint32_t *st = A.stats + static_cast<size_t>(s) * kStWords;
    const LdsBoxes L{lds_box, lds_box + A.max_n, lds_box + 2 * A.max_n,
                     lds_box + 3 * A.max_n, lds_box + 4 * A.max_n, lds_box + 5 * A.max_n};
    const float h = A.h;
    const uint32_t n_steps = static_cast<uint32_t>(A.n_frames) * static_cast<uint32_t>(A.n_sub);

    for (uint32_t step = 0; step < n_steps; ++step) {
        // 1. integrate one substep (integrator.cpp:17-70, n_substeps = 1 at h).
        for (uint32_t i = t; i < n; i += BS) {
            if (bv.inv_mass[i] == 0.f || (bv.flags[i] & kFlagSleeping)) continue;
            Kin k{bv.pos_x[i], bv.pos_y[i], bv.pos_z[i], bv.rot_w[i], bv.rot_x[i], bv.rot_y[i],
                  bv.rot_z[i], bv.vel_x[i], bv.vel_y[i], bv.vel_z[i], bv.ang_x[i], bv.ang_y[i], bv.ang_z[i]};
            integrate_body(k, A.gx, A.gy, A.gz, h, 1);
            bv.pos_x[i] = k.px; bv.pos_y[i] = k.py; bv.pos_z[i] = k.pz;
            bv.rot_w[i] = k.qw; bv.rot_x[i] = k.qx; bv.rot_y[i] = k.qy; bv.rot_z[i] = k.qz;
            bv.vel_x[i] = k.vx; bv.vel_y[i] = k.vy; bv.vel_z[i] = k.vz;
        }
        if (t < 4) s_flags[t] = 0;
        __syncthreads();

        // 2. leaf boxes (broadphase.cpp:215-251), binary16-rounded.
        for (uint32_t i = t; i < n; i += BS) {
            const uint32_t sh = bv.shape[i];
            float ex, ey, ez;
            leaf_extent(A.sv.type[sh], A.sv.half_x[sh], A.sv.half_y[sh], A.sv.half_z[sh], bv.rot_w[i],
                        bv.rot_x[i], bv.rot_y[i], bv.rot_z[i], ex, ey, ez);
            const float px = bv.pos_x[i], py = bv.pos_y[i], pz = bv.pos_z[i];
            L.mnx[i] = f16_to_f32(f32_to_f16(px - ex));
            L.mny[i] = f16_to_f32(f32_to_f16(py - ey));
            L.mnz[i] = f16_to_f32(f32_to_f16(pz - ez));
            L.mxx[i] = f16_to_f32(f32_to_f16(px + ex));
            L.mxy[i] = f16_to_f32(f32_to_f16(py + ey));
            L.mxz[i] = f16_to_f32(f32_to_f16(pz + ez));
        }
        __syncthreads();

        // 3. candidate pairs: row counts, scan, ordered emission.
        for (uint32_t a = t; a < n; a += BS) {
            const float q[6] = {L.mnx[a], L.mny[a], L.mnz[a], L.mxx[a], L.mxy[a], L.mxz[a]};
            uint32_t c = 0;
            for (uint32_t b = a + 1; b < n; ++b) c += touch(q, L, b) ? 1u : 0u;
            S.row[a] = c;
        }
        __syncthreads();
        const uint32_t n_pairs = scan_in_place(S.row, n, wsum);
        __syncthreads();
        for (uint32_t a = t; a < n; a += BS) {
            uint32_t pos = S.row[a];
            if (pos >= P) continue;
            const float q[6] = {L.mnx[a], L.mny[a], L.mnz[a], L.mxx[a], L.mxy[a], L.mxz[a]};
            for (uint32_t b = a + 1; b < n && pos < P; ++b) {
                if (touch(q, L, b)) {
                    pairs[pos] = dy_contact_pair{a, b};
                    ++pos;
                }
            }
        }
        __syncthreads();

        // 4. narrowphase in pair order.
        const uint32_t np_n = n_pairs < P ? n_pairs : P;
        uint32_t run = 0;
        for (uint32_t base = 0; base < np_n; base += BS) {
            const uint32_t i = base + t;
            uint32_t hit = 0, unsup = 0;
            Hit hc;
            dy_contact_pair p{0u, 0u};
            if (i < np_n) {
                p = pairs[i];
                hit = pair_contact(bv, A.sv, p.a, p.b, hc, unsup) ? 1u : 0u;
            }
            if (unsup) atomicOr(&s_flags[1], kBatchErrUnsupported);
            uint32_t tot = 0;
            const uint32_t ex = block_exclusive_scan(hit, wsum, tot);
            const uint32_t slot = run + ex;
            if (hit && slot < C) {
                cv.body_a[slot] = p.a; cv.body_b[slot] = p.b;
                cv.pos_x[slot] = hc.px; cv.pos_y[slot] = hc.py; cv.pos_z[slot] = hc.pz;
                cv.norm_x[slot] = hc.nx; cv.norm_y[slot] = hc.ny; cv.norm_z[slot] = hc.nz;
                cv.depth[slot] = hc.depth;
            }
            run += tot;
        }
        const uint32_t nc = run < C ? run : C;

        // 5. solve (xpbd_solver.cpp:409-733), contacts only.
        uint32_t serial = 1, rounds = 0, n_cls = 0, overflow = 0, err = 0;
        if (nc > 0) {
            for (uint32_t x = t; x < n; x += BS) { S.deg[x] = 0; S.fill[x] = 0; S.used[x] = 0; }
            for (uint32_t k = t; k < nc; k += BS) {
                S.lam[k] = 0.f;
                S.rnd[k] = kNotAssigned;
                S.pred_a[k] = -1;
                S.pred_b[k] = -1;
                S.cls[k] = 0;
            }
            __syncthreads();
            for (uint32_t k = t; k < nc; k += BS) {
                const uint32_t a = cv.body_a[k], b = cv.body_b[k];
                if (bv.inv_mass[a] != 0.f) atomicAdd(&S.deg[a], 1u);
                if (bv.inv_mass[b] != 0.f) atomicAdd(&S.deg[b], 1u);
            }
            __syncthreads();
            for (uint32_t x = t; x < n; x += BS) S.loff[x] = S.deg[x];
            __syncthreads();
            (void)scan_in_place(S.loff, n, wsum);
            __syncthreads();
            for (uint32_t k = t; k < nc; k += BS) {
                const uint32_t a = cv.body_a[k], b = cv.body_b[k];
                if (bv.inv_mass[a] != 0.f) S.list[S.loff[a] + atomicAdd(&S.fill[a], 1u)] = k;
                if (bv.inv_mass[b] != 0.f) S.list[S.loff[b] + atomicAdd(&S.fill[b], 1u)] = k;
            }
            __syncthreads();
            const bool jp = !(A.force_serial || nc <= kBatchSerialMax);
            for (uint32_t x = t; x < n; x += BS) body_predecessors(x, S, cv.body_a, jp);
            __syncthreads();
            uint32_t remaining = 0, over = 0;
            rounds = assign_rounds(nc, jp, S, cv, bv.inv_mass, wsum, remaining, over, err);
            overflow = over;
            serial = jp ? 0u : 1u;
            if (jp && (remaining > 0 || over)) {
                // xpbd_solver.cpp:696-699: colouring failed -> reference order.
                serial = 1;
                for (uint32_t k = t; k < nc; k += BS) S.rnd[k] = kNotAssigned;
                __syncthreads();
                for (uint32_t x = t; x < n; x += BS) body_predecessors(x, S, cv.body_a, false);
                __syncthreads();
                (void)assign_rounds(nc, false, S, cv, bv.inv_mass, wsum, remaining, over, err);
            }
            if (remaining > 0) err |= kBatchErrLists;
            if (err == 0) {
                for (uint32_t k = t; k < nc; k += BS) atomicMax(&s_flags[2], S.cls[k] + 1u);
            }
            __syncthreads();
            n_cls = s_flags[2];
            if (err == 0) {
                // classes -> contiguous ranges of `order`
                for (uint32_t c = t; c < n_cls; c += BS) S.ccnt[c] = 0;
                __syncthreads();
                for (uint32_t k = t; k < nc; k += BS) atomicAdd(&S.ccnt[S.cls[k]], 1u);
                __syncthreads();
                for (uint32_t c = t; c < n_cls; c += BS) S.coff[c] = S.ccnt[c];
                __syncthreads();
                (void)scan_in_place(S.coff, n_cls, wsum);
                if (t == 0) S.coff[n_cls] = nc;
                for (uint32_t c = t; c < n_cls; c += BS) S.ccnt[c] = 0;
                __syncthreads();
                for (uint32_t k = t; k < nc; k += BS) {
                    const uint32_t c = S.cls[k];
                    S.order[S.coff[c] + atomicAdd(&S.ccnt[c], 1u)] = k;
                }
                __syncthreads();
                // I iterations: classes in order, then renormalise every body
                // (xpbd_solver.cpp:578-586,700-733).
                for (int it = 0; it < A.n_iters; ++it) {
                    for (uint32_t c = 0; c < n_cls; ++c) {
                        const uint32_t beg = S.coff[c], end = S.coff[c + 1];
                        for (uint32_t j = beg + t; j < end; j += BS) {
                            const uint32_t k = S.order[j];
                            const uint32_t a = cv.body_a[k], b = cv.body_b[k];
                            const ContactIn ci{cv.pos_x[k], cv.pos_y[k], cv.pos_z[k], cv.norm_x[k],
                                               cv.norm_y[k], cv.norm_z[k], cv.depth[k]};
                            BodyState Ba = load_body(bv, a), Bb = load_body(bv, b);
                            float l = S.lam[k];
                            if (project_contact(ci, Ba, Bb, A.mu, l, h)) {
                                S.lam[k] = l;
                                if (Ba.im != 0.f) store_body(bv, a, Ba);
                                if (Bb.im != 0.f) store_body(bv, b, Bb);
                            }
                        }
                        __syncthreads();
                    }
                    for (uint32_t i = t; i < n; i += BS) {
                        float qw = bv.rot_w[i], qx = bv.rot_x[i], qy = bv.rot_y[i], qz = bv.rot_z[i];
                        renormalise(qw, qx, qy, qz);
                        bv.rot_w[i] = qw; bv.rot_x[i] = qx; bv.rot_y[i] = qy; bv.rot_z[i] = qz;
                    }
                    __syncthreads();
                }
            }
        }
        if (err) atomicOr(&s_flags[1], err);
        __syncthreads();
        if (t == 0) {
            st[kStPairs] = n_pairs;
            st[kStContacts] = run;
            st[kStClasses] = n_cls;
            st[kStSerial] = nc > 0 ? serial : 0u;
            st[kStRounds] = rounds;
            st[kStOverflow] = overflow;
            st[kStError] |= s_flags[1];
            st[kStSubsteps] += 1u;
        }
        __syncthreads();
    }
}

template <int BS>
int launch_frames(const BatchArgs &a, uint32_t n_scenes, hipStream_t s) {
    const size_t lds = 6u * a.max_n * sizeof(float);
    hipLaunchKernelGGL(batch_frames<BS>, dim3(n_scenes), dim3(BS), lds, s, a);
    DY_LAUNCH_CHECK();
    return DY_OK;
}

}  // namespace dy

using namespace dy;

struct dy_batch_s {
    uint32_t n_scenes = 0, max_n = 0, block = 64;
    int n_iters = 10;
    float mu = 0.5f;
    int force_serial = 0;
    std::vector<uint32_t> body_off, pair_off, contact_off;
    uint32_t *d_order = nullptr, *d_body_off = nullptr, *d_pair_off = nullptr, *d_contact_off = nullptr;
    uint64_t *d_scratch_off = nullptr;
    uint32_t *d_scratch = nullptr, *d_stats = nullptr;
    dy_contact_pair *d_pairs = nullptr;
    void *d_contacts = nullptr;  // 9 arrays of contact_off.back() words
    dy_contact_view cv{};
};

namespace {

void batch_free(dy_batch_s *b) {
    void *ps[] = {b->d_order, b->d_body_off, b->d_pair_off, b->d_contact_off, b->d_scratch_off,
                  b->d_scratch, b->d_stats, b->d_pairs, b->d_contacts};
    for (void *p : ps)
        if (p) (void)hipFree(p);
}

template <typename T>
int dev_alloc(T *&p, size_t n) {
    DY_TRY(hipMalloc(reinterpret_cast<void **>(&p), (n ? n : 1) * sizeof(T)));
    return DY_OK;
}

int upload(hipStream_t s, void *dst, const void *src, size_t bytes) {
    if (bytes) DY_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    return DY_OK;
}

int batch_launch(dy_batch_s *b, hipStream_t s, const dy_body_view &bodies, const dy_shape_view &shapes, float gx,
                 float gy, float gz, float dt, int n_substeps, int n_frames) {
    BatchArgs a{};
    a.bv = bodies;
    a.sv = shapes;
    a.cv = b->cv;
    a.pairs = b->d_pairs;
    a.order = b->d_order;
    a.body_off = b->d_body_off;
    a.pair_off = b->d_pair_off;
    a.contact_off = b->d_contact_off;
    a.scratch_off = b->d_scratch_off;
    a.scratch = b->d_scratch;
    a.stats = b->d_stats;
    a.max_n = b->max_n;
    a.gx = gx; a.gy = gy; a.gz = gz;
    a.h = dt / static_cast<float>(n_substeps);
    a.n_sub = n_substeps;
    a.n_frames = n_frames;
    a.n_iters = b->n_iters;
    a.force_serial = b->force_serial;
    a.mu = b->mu;
    switch (b->block) {
        case 64: return launch_frames<64>(a, b->n_scenes, s);
        case 128: return launch_frames<128>(a, b->n_scenes, s);
        case 256: return launch_frames<256>(a, b->n_scenes, s);
        default: return DY_ERR_INVALID_ARG;
    }
}

int check_step_args(const dy_batch_s *b, const dy_body_view *bodies, const dy_shape_view *shapes, int n_substeps,
                    int n_frames) {
    if (!b || !bodies || !shapes || n_substeps <= 0 || n_frames < 0) return DY_ERR_INVALID_ARG;
    if (bodies->n != b->body_off.back()) return DY_ERR_INVALID_ARG;
    return DY_OK;
}

}  // namespace

extern "C" {

int dy_batch_create(dy_stream_t s, uint32_t n_scenes, const uint32_t *body_offset, const uint32_t *max_pairs,
                    const uint32_t *max_contacts, int n_iters, float friction, dy_batch_t *out) {
    if (!out || !body_offset || !max_pairs || !max_contacts || n_scenes == 0 || n_iters < 0)
        return DY_ERR_INVALID_ARG;
    *out = nullptr;
    auto *b = new dy_batch_s();
    b->n_scenes = n_scenes;
    b->n_iters = n_iters;
    b->mu = friction;
    b->body_off.assign(body_offset, body_offset + n_scenes + 1);
    b->pair_off.assign(n_scenes + 1, 0u);
    b->contact_off.assign(n_scenes + 1, 0u);
    std::vector<uint64_t> scratch_off(n_scenes + 1, 0ull);
    uint64_t pt = 0, ct = 0;
    int st = DY_OK;
    for (uint32_t i = 0; i < n_scenes && st == DY_OK; ++i) {
        const uint32_t lo = body_offset[i], hi = body_offset[i + 1];
        if (hi <= lo || hi - lo > kBatchMaxBodies) st = DY_ERR_INVALID_ARG;
        const uint32_t n = hi - lo;
        b->max_n = std::max(b->max_n, n);
        b->pair_off[i] = static_cast<uint32_t>(pt);
        b->contact_off[i] = static_cast<uint32_t>(ct);
        pt += max_pairs[i];
        ct += max_contacts[i];
        scratch_off[i + 1] = scratch_off[i] + scene_scratch_words(n, max_contacts[i]);
        if (pt >= 0xffffffffull || ct >= 0xffffffffull) st = DY_ERR_INVALID_ARG;
    }
    if (body_offset[0] != 0) st = DY_ERR_INVALID_ARG;
    if (st != DY_OK) {
        delete b;
        return st;
    }
    b->pair_off[n_scenes] = static_cast<uint32_t>(pt);
    b->contact_off[n_scenes] = static_cast<uint32_t>(ct);
    // Largest scenes first, so the longest workgroups start earliest.
    std::vector<uint32_t> order(n_scenes);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
        return body_offset[x + 1] - body_offset[x] > body_offset[y + 1] - body_offset[y];
    });
    const hipStream_t hsm = hs(s);
    const size_t C = ct;
    auto fail = [&](int code) {
        batch_free(b);
        delete b;
        return code;
    };
#define DY_BATCH_TRY(expr)            \
    do {                              \
        const int e_ = (expr);        \
        if (e_ != DY_OK) return fail(e_); \
    } while (0)
    DY_BATCH_TRY(dev_alloc(b->d_order, n_scenes));
    DY_BATCH_TRY(dev_alloc(b->d_body_off, n_scenes + 1));
    DY_BATCH_TRY(dev_alloc(b->d_pair_off, n_scenes + 1));
    DY_BATCH_TRY(dev_alloc(b->d_contact_off, n_scenes + 1));
    DY_BATCH_TRY(dev_alloc(b->d_scratch_off, n_scenes + 1));
    DY_BATCH_TRY(dev_alloc(b->d_scratch, scratch_off[n_scenes]));
    DY_BATCH_TRY(dev_alloc(b->d_stats, static_cast<size_t>(n_scenes) * kStWords));
    DY_BATCH_TRY(dev_alloc(b->d_pairs, pt));
    {
        char *p = nullptr;
        DY_BATCH_TRY(dev_alloc(p, 9 * (C ? C : 1) * 4));
        b->d_contacts = p;
        const size_t stride = (C ? C : 1) * 4;
        b->cv.body_a = reinterpret_cast<uint32_t *>(p);
        b->cv.body_b = reinterpret_cast<uint32_t *>(p + stride);
        b->cv.pos_x = reinterpret_cast<float *>(p + 2 * stride);
        b->cv.pos_y = reinterpret_cast<float *>(p + 3 * stride);
        b->cv.pos_z = reinterpret_cast<float *>(p + 4 * stride);
        b->cv.norm_x = reinterpret_cast<float *>(p + 5 * stride);
        b->cv.norm_y = reinterpret_cast<float *>(p + 6 * stride);
        b->cv.norm_z = reinterpret_cast<float *>(p + 7 * stride);
        b->cv.depth = reinterpret_cast<float *>(p + 8 * stride);
        b->cv.n = nullptr;
        b->cv.capacity = static_cast<uint32_t>(C);
    }
    DY_BATCH_TRY(upload(hsm, b->d_order, order.data(), n_scenes * 4));
    DY_BATCH_TRY(upload(hsm, b->d_body_off, b->body_off.data(), (n_scenes + 1) * 4));
    DY_BATCH_TRY(upload(hsm, b->d_pair_off, b->pair_off.data(), (n_scenes + 1) * 4));
    DY_BATCH_TRY(upload(hsm, b->d_contact_off, b->contact_off.data(), (n_scenes + 1) * 4));
    DY_BATCH_TRY(upload(hsm, b->d_scratch_off, scratch_off.data(), (n_scenes + 1) * 8));
    if (hipMemsetAsync(b->d_stats, 0, static_cast<size_t>(n_scenes) * kStWords * 4, hsm) != hipSuccess)
        return fail(DY_ERR_INTERNAL);
    if (hipStreamSynchronize(hsm) != hipSuccess) return fail(DY_ERR_INTERNAL);
#undef DY_BATCH_TRY
    *out = b;
    return DY_OK;
}

int dy_batch_destroy(dy_batch_t b) {
    if (!b) return DY_OK;
    batch_free(b);
    delete b;
    return DY_OK;
}

int dy_batch_set_force_serial(dy_batch_t b, int force_serial) {
    if (!b) return DY_ERR_INVALID_ARG;
    b->force_serial = force_serial ? 1 : 0;
    return DY_OK;
}

int dy_batch_set_block_threads(dy_batch_t b, uint32_t threads) {
    if (!b || (threads != 64 && threads != 128 && threads != 256)) return DY_ERR_INVALID_ARG;
    b->block = threads;
    return DY_OK;
}

int dy_batch_get_block_threads(dy_batch_t b, uint32_t *threads) {
    if (!b || !threads) return DY_ERR_INVALID_ARG;
    *threads = b->block;
    return DY_OK;
}

int dy_batch_step(dy_batch_t b, dy_stream_t s, const dy_body_view *bodies, const dy_shape_view *shapes, float gx,
                  float gy, float gz, float dt, int n_substeps, int n_frames) {
    DY_TRY_STATUS(check_step_args(b, bodies, shapes, n_substeps, n_frames));
    if (n_frames == 0) return DY_OK;
    return batch_launch(b, hs(s), *bodies, *shapes, gx, gy, gz, dt, n_substeps, n_frames);
}

int dy_batch_autotune(dy_batch_t b, dy_stream_t s, const dy_body_view *bodies, const dy_shape_view *shapes,
                      float gx, float gy, float gz, float dt, int n_substeps, int trial_frames,
                      uint32_t *chosen, float *ms_out) {
    DY_TRY_STATUS(check_step_args(b, bodies, shapes, n_substeps, trial_frames));
    if (trial_frames <= 0) return DY_ERR_INVALID_ARG;
    const hipStream_t hsm = hs(s);
    const size_t n = bodies->n;
    // The trials advance the state; keep the 13 mutable arrays and the stats.
    float *const arr[13] = {bodies->pos_x, bodies->pos_y, bodies->pos_z, bodies->rot_w, bodies->rot_x,
                            bodies->rot_y, bodies->rot_z, bodies->vel_x, bodies->vel_y, bodies->vel_z,
                            bodies->ang_x, bodies->ang_y, bodies->ang_z};
    float *save = nullptr;
    uint32_t *save_stats = nullptr;
    DY_TRY(hipMalloc(reinterpret_cast<void **>(&save), 13 * n * sizeof(float)));
    const size_t stat_bytes = static_cast<size_t>(b->n_scenes) * kStWords * 4;
    if (hipMalloc(reinterpret_cast<void **>(&save_stats), stat_bytes) != hipSuccess) {
        (void)hipFree(save);
        return DY_ERR_INTERNAL;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int st = DY_OK;
    auto copy_state = [&](bool to_save) -> int {
        for (int k = 0; k < 13; ++k) {
            float *bk = save + k * n;
            DY_TRY(hipMemcpyAsync(to_save ? bk : arr[k], to_save ? arr[k] : bk, n * sizeof(float),
                                  hipMemcpyDeviceToDevice, hsm));
        }
        DY_TRY(hipMemcpyAsync(to_save ? save_stats : b->d_stats, to_save ? b->d_stats : save_stats, stat_bytes,
                              hipMemcpyDeviceToDevice, hsm));
        return DY_OK;
    };
    const uint32_t cand[3] = {64, 128, 256};
    float best = 0.f;
    uint32_t best_bs = b->block;
    const uint32_t keep = b->block;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) st = DY_ERR_INTERNAL;
    if (st == DY_OK) st = copy_state(true);
    for (int c = 0; c < 3 && st == DY_OK; ++c) {
        b->block = cand[c];
        st = copy_state(false);
        if (st == DY_OK) st = batch_launch(b, hsm, *bodies, *shapes, gx, gy, gz, dt, n_substeps, 1);  // warm-up
        if (st == DY_OK) st = copy_state(false);
        if (st == DY_OK && hipEventRecord(e0, hsm) != hipSuccess) st = DY_ERR_INTERNAL;
        if (st == DY_OK) st = batch_launch(b, hsm, *bodies, *shapes, gx, gy, gz, dt, n_substeps, trial_frames);
        if (st == DY_OK && hipEventRecord(e1, hsm) != hipSuccess) st = DY_ERR_INTERNAL;
        if (st == DY_OK && hipEventSynchronize(e1) != hipSuccess) st = DY_ERR_INTERNAL;
        float ms = 0.f;
        if (st == DY_OK && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) st = DY_ERR_INTERNAL;
        if (ms_out) ms_out[c] = ms;
        if (st == DY_OK && (c == 0 || ms < best)) {
            best = ms;
            best_bs = cand[c];
        }
    }
    b->block = st == DY_OK ? best_bs : keep;
    if (st == DY_OK) st = copy_state(false);
    if (hipStreamSynchronize(hsm) != hipSuccess && st == DY_OK) st = DY_ERR_INTERNAL;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(save);
    (void)hipFree(save_stats);
    if (st == DY_OK && chosen) *chosen = b->block;
    return st;
}

int dy_batch_stats(dy_batch_t b, dy_stream_t s, dy_batch_scene_stats *out) {
    if (!b || !out) return DY_ERR_INVALID_ARG;
    static_assert(sizeof(dy_batch_scene_stats) == kStWords * 4, "stats layout");
    DY_TRY(hipMemcpyAsync(out, b->d_stats, static_cast<size_t>(b->n_scenes) * kStWords * 4,
                          hipMemcpyDeviceToHost, hs(s)));
    DY_TRY(hipStreamSynchronize(hs(s)));
    uint32_t any = 0;
    for (uint32_t i = 0; i < b->n_scenes; ++i) any |= out[i].error;
    if (any & kBatchErrUnsupported) return DY_ERR_UNSUPPORTED;
    if (any) return DY_ERR_INTERNAL;
    return DY_OK;
}

int dy_batch_layout_get(dy_batch_t b, dy_batch_layout *out) {
    if (!b || !out) return DY_ERR_INVALID_ARG;
    out->n_scenes = b->n_scenes;
    out->pairs = b->d_pairs;
    out->contacts = b->cv;
    out->body_offset = b->body_off.data();
    out->pair_offset = b->pair_off.data();
    out->contact_offset = b->contact_off.data();
    return DY_OK;
}

}  // extern "C"
