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
#include <cmath>
#include <map>
#include <numeric>
#include <type_traits>
#include <utility>
#include <vector>

#include "dy_internal.hpp"
#include "pair_contact.hpp"
#include "adjoint.hpp"

namespace dy {

constexpr uint32_t kBatchMaxBodies = 2048;  // LDS box records: 32 B per body
constexpr uint32_t kLdsMaxFloats = 16384;   // 64 KiB of dynamic LDS per workgroup
constexpr uint32_t kTuneShapes = 6;         // launch shapes the autotuner times
constexpr uint32_t kBatchSerialMax = 256;   // xpbd_solver.hpp:29 kSerialThreshold
constexpr uint32_t kBatchMaxRounds = 256;   // xpbd_solver.cpp:640
constexpr uint32_t kNotAssigned = 0xffffffffu;
// Scene error bits (stats word kStError, and a tape frame's header): a hull
// or mesh shape met; an internal check failed; more contacts in a step than
// the tape holds (the step ran, its record is incomplete).
constexpr uint32_t kBatchErrUnsupported = 1u, kBatchErrLists = 2u, kBatchErrTape = 4u;

// Timed stages: integrate + leaf boxes, pairs, narrowphase, class
// assignment, sweep.
constexpr uint32_t kStages = 5;

enum : uint32_t {
    kStPairs = 0, kStContacts, kStClasses, kStSerial, kStRounds, kStOverflow, kStError, kStSubsteps,
    kStSumPairs, kStSumContacts, kStSumDyn, kStSumSolves, kStGrid, kStWords
};

struct BatchArgs {
    dy_body_view bv;       // every scene's bodies, concatenated
    dy_shape_view sv;      // shape catalog; bv.shape holds catalog indices
    dy_contact_view cv;    // contact store, concatenated (cv.n unused)
    dy_contact_pair *pairs;
    const uint4 *units;    // per workgroup: up to four scenes (kNoScene = empty slot)
    const uint32_t *unit_kind;  // bit 0: whole-workgroup team; bit 1 + slot: slot's state in LDS
    const uint32_t *body_off, *pair_off, *contact_off;
    const uint64_t *scratch_off;  // 32-bit words, even
    uint32_t *scratch;
    uint32_t *stats;       // kStWords per scene
    uint64_t *stage_clk;   // kStages per scene when timed
    float *tape;           // differentiable form: per-frame records (TapeLayout)
    const uint64_t *tape_off;  // per scene, floats
    const uint32_t *tape_cap;  // per scene: contacts a tape step holds
    int keep_input;        // 1: the final state is not written back (a recompute from a checkpoint)
    uint32_t region_floats;  // LDS floats per team
    float gx, gy, gz, h;
    int n_sub, n_frames, n_iters, force_serial;
    float mu;
};

constexpr uint32_t kNoScene = 0xffffffffu;
constexpr uint32_t kBigThreads = 512;  // workgroup of the larger-LDS launch (8 wavefronts: <= 256 VGPRs each)
constexpr uint32_t kUnitBlock = 1u;
// LDS floats per body: a 32 B box record, plus the 20 float fields of the
// body when the scene's state is held in LDS for the whole launch.
constexpr uint32_t kBoxFloats = 8, kStateFloats = 20;

// Per-scene scratch, carved from the scene's slice.
struct SceneScratch {
    uint64_t *used;                     // [n] colour masks
    uint32_t *row, *deg, *loff, *fill;  // [n]
    float *lam;                         // [C]
    int32_t *pred_a, *pred_b;           // [C]
    uint32_t *rnd, *cls, *order;        // [C]
    uint32_t *list;                     // [2C]
    uint32_t *ccnt, *coff;              // [C + 1]
    uint32_t *feat;                     // [C] contact feature words (taped runs)
    uint32_t *bits;                     // [n * ceil(n / 32)] overlap rows
};

__host__ __device__ inline uint64_t scene_scratch_words(uint32_t n, uint32_t c) {
    const uint64_t w = 6ull * n + 11ull * c + 2ull + static_cast<uint64_t>(n) * ((n + 31u) / 32u);
    return (w + 1ull) & ~1ull;
}

// Per-frame record of one scene for the reverse sweep (floats, each part
// 16-byte aligned): frame-start state (13 fields x n, SoA), post-integrate
// pose (7 x n), quaternions before each iteration's renormalisation
// (iters x 4 x n), a header (contacts solved, classes, error bits, the
// iterations that projected a live contact: bit min(it, 31)), class offsets
// (C + 1), contacts in class order (8 words: p, n, depth, the feature word;
// and the body pair),
// and each contact's inputs at each iteration (28 floats: both bodies' 13
// mutable fields, lambda; written only for a live projection, see below),
// and each contact's lambda at each iteration.
struct TapeLayout {
    uint64_t st, po, qr, hdr, co, rec, ab, pre, lam, frame;
};

__host__ __device__ inline TapeLayout tape_layout(uint32_t n, uint32_t c, uint32_t iters) {
    TapeLayout L;
    uint64_t o = 0;
    L.st = o; o = (o + 13ull * n + 3ull) & ~3ull;
    L.po = o; o = (o + 7ull * n + 3ull) & ~3ull;
    L.qr = o; o = (o + 4ull * n * iters + 3ull) & ~3ull;
    L.hdr = o; o += 4;
    L.co = o; o = (o + c + 1ull + 3ull) & ~3ull;
    L.rec = o; o += 8ull * c;
    L.ab = o; o = (o + 2ull * c + 3ull) & ~3ull;
    L.pre = o; o += 28ull * c * iters;
    L.lam = o; o = (o + static_cast<uint64_t>(c) * iters + 3ull) & ~3ull;
    L.frame = o;
    return L;
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
    s.coff = p; p += c + 1;
    s.feat = p; p += c;
    s.bits = p;
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

// Workgroup sum of v and OR of f in one exchange (two barriers).  `wsum`
// needs 8 + waves words; every thread gets both totals.
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

// A team steps one scene.  WaveTeam: the 64 lanes of one wavefront, whose
// lanes see each other's memory operations in program order, so its barrier
// is a wavefront-scope fence pair that orders the compiler and emits no
// instruction.  BlockTeam: the whole workgroup, __syncthreads().  Each team
// owns 32 LDS words: [0, 17) scan partials, [20] error bits, [21] class count.
struct WaveTeam {
    static constexpr uint32_t kSize = 64;
    uint32_t rank;
    uint32_t *sh;
    __device__ __forceinline__ void sync() const {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    __device__ __forceinline__ uint32_t excl_scan(uint32_t v, uint32_t &tot) const {
        const uint32_t inc = wave_inclusive_scan(v);
        tot = __shfl(inc, 63, 64);
        return inc - v;
    }
    __device__ __forceinline__ uint32_t sum_or(uint32_t v, uint32_t f, uint32_t &f_all) const {
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) {
            v += __shfl_xor(v, d, 64);
            f |= __shfl_xor(f, d, 64);
        }
        f_all = f;
        return v;
    }
};

template <uint32_t NT>
struct BlockTeam {
    static constexpr uint32_t kSize = NT;
    uint32_t rank;
    uint32_t *sh;
    __device__ __forceinline__ void sync() const { __syncthreads(); }
    __device__ __forceinline__ uint32_t excl_scan(uint32_t v, uint32_t &tot) const {
        return block_exclusive_scan(v, sh, tot);
    }
    __device__ __forceinline__ uint32_t sum_or(uint32_t v, uint32_t f, uint32_t &f_all) const {
        return block_sum_or(v, f, sh, f_all);
    }
};

// Exclusive scan of data[0, n) in place by the team; returns the total.  Each
// entry is rewritten by the thread that read it: callers sync before other
// threads read the result.
template <class Team>
__device__ uint32_t team_scan(const Team &tm, uint32_t *data, uint32_t n, uint32_t stride = 1) {
    uint32_t carry = 0;
    for (uint32_t base = 0; base < n; base += Team::kSize) {
        const uint32_t i = base + tm.rank;
        const uint32_t v = i < n ? data[i * stride] : 0u;
        uint32_t tot = 0;
        const uint32_t ex = tm.excl_scan(v, tot);
        if (i < n) data[i * stride] = carry + ex;
        carry += tot;
    }
    return carry;
}

// Reference overlap test (broadphase.cpp:322-325): separated on some axis,
// else touching or overlapping.  Box records: mn xyz, mx xyz, 2 unused.
__device__ __forceinline__ bool touch(const float q[6], const float4 lo, const float2 hi) {
    return !(lo.x > q[3] || lo.w < q[0] || lo.y > q[4] || hi.x < q[1] || lo.z > q[5] || hi.y < q[2]);
}

// ---- candidate pairs of a larger scene on a uniform x-z grid --------------
// Order-preserving map of floats to u32 (the max of keys is the key of the max).
__device__ __forceinline__ uint32_t fkey(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float funkey(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = max(v, static_cast<uint32_t>(__shfl_xor(v, d, 64)));
    return v;
}
__device__ __forceinline__ float wave_sum_f32(float v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

constexpr uint32_t kGridMinBodies = 160;  // smaller scenes test every pair
constexpr uint32_t kRowList = 16;         // partners a gridded row may list instead of bit words

struct PairGrid {
    float x0, z0, wx, wz;
    uint32_t kx, kz;
    __device__ static uint32_t cells(float span, float w) {
        const float f = floorf(span / w);
        return f >= 65535.f ? 65536u : static_cast<uint32_t>(f) + 1u;
    }
    __device__ static uint32_t cell(float v, float v0, float w, uint32_t k) {
        const float u = floorf((v - v0) / w);
        return u > 0.f ? static_cast<uint32_t>(fminf(u, static_cast<float>(k - 1u))) : 0u;
    }
    __device__ uint32_t cx(float x) const { return cell(x, x0, wx, kx); }
    __device__ uint32_t cz(float z) const { return cell(z, z0, wz, kz); }
};

// The rows of S.bits (partners b > a whose boxes touch a's) and their counts
// S.row, found on a grid: every box more than 4 times the mean extent (a
// ground) is "large" and tested against every body; the others are binned by
// the x and z of their lower corner into cells 1.01 times the largest small
// extent wide, so that a small body's small partners lie in the 3 x 3 cells
// around its own (|lo_b - lo_a| <= that extent on each axis when the boxes
// touch; the 1 % covers the rounding of the cell coordinates).  The test and
// the words are the all-pairs pass's, so the pairs are the same.  Scratch:
// the box records' two unused words (8k + 7: cell counts, then offsets;
// 8p + 6: bodies by cell, then the large ones), S.deg
// (1: the row is a sorted list of its partners, else bit words) and the team
// words tf[0, 4) and tf[23, 31).  False, before writing any row, when the
// grid does not apply (no small body, non-finite boxes) or would not pay
// (too few cells): the caller tests every pair.  Every exit is team-uniform.
// Kept out of line: inlined into the step kernel, this compiler (ROCm 7.2,
// gfx950) fails instruction selection ("Operand has incorrect register
// class" on a compare with the LDS aperture).
typedef __attribute__((address_space(3))) float lds_f32;
typedef __attribute__((address_space(3))) uint32_t lds_u32;

// A box record from LDS (out of line, the pointer's address space is
// explicit so that the reads are LDS instructions, not flat ones).
__device__ __forceinline__ float4 lds_lo(const lds_f32 *box, uint32_t i) {
    const lds_f32 *p = box + 8u * i;
    return make_float4(p[0], p[1], p[2], p[3]);
}
__device__ __forceinline__ float2 lds_hi(const lds_f32 *box, uint32_t i) {
    const lds_f32 *p = box + 8u * i + 4u;
    return make_float2(p[0], p[1]);
}

template <class Team>
__device__ __attribute__((noinline)) bool grid_rows(const Team &tm, lds_f32 *box, uint32_t n, uint32_t W,
                                                    const SceneScratch &S, lds_u32 *tf) {
    constexpr uint32_t NT = Team::kSize;
    const uint32_t r = tm.rank;
    lds_u32 *bu = reinterpret_cast<lds_u32 *>(box);
    // 1. the large-box threshold: 4 x the mean of the x and z extents
    float se = 0.f;
    for (uint32_t i = r; i < n; i += NT) {
        const float4 lo = lds_lo(box, i);
        const float2 hi = lds_hi(box, i);
        se += (lo.w - lo.x) + (hi.y - lo.z);
    }
    se = wave_sum_f32(se);
    if (lane_id() == 0) tf[r >> 6] = __float_as_uint(se);  // per-wavefront partials (scan words, free here)
    tm.sync();
    float st = 0.f;
    for (uint32_t w = 0; w < NT / 64; ++w) st += __uint_as_float(tf[w]);
    const float T = 2.f * st / static_cast<float>(n);
    auto small = [&](float ex, float ez) { return ex <= T && ez <= T; };
    // 2. over the small boxes: largest extents, lower-corner ranges, count
    uint32_t k[6] = {0u, 0u, 0u, 0u, 0u, 0u};  // Ex, Ez, -min x, max x, -min z, max z (keys)
    uint32_t ns = 0;
    for (uint32_t i = r; i < n; i += NT) {
        const float4 lo = lds_lo(box, i);
        const float2 hi = lds_hi(box, i);
        const float ex = lo.w - lo.x, ez = hi.y - lo.z;
        if (!small(ex, ez)) continue;
        k[0] = max(k[0], fkey(ex));
        k[1] = max(k[1], fkey(ez));
        k[2] = max(k[2], ~fkey(lo.x));
        k[3] = max(k[3], fkey(lo.x));
        k[4] = max(k[4], ~fkey(lo.z));
        k[5] = max(k[5], fkey(lo.z));
        ++ns;
    }
#pragma unroll
    for (int j = 0; j < 6; ++j) k[j] = wave_max_u32(k[j]);
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) ns += __shfl_xor(ns, d, 64);
    if (lane_id() == 0) {
#pragma unroll
        for (int j = 0; j < 6; ++j) atomicMax((uint32_t *)&tf[23 + j], k[j]);
        atomicAdd((uint32_t *)&tf[29], ns);
    }
    tm.sync();
    ns = tf[29];
    if (ns == 0) return false;
    const float Ex = funkey(tf[23]), Ez = funkey(tf[24]);
    const float xmin = funkey(~tf[25]), xmax = funkey(tf[26]), zmin = funkey(~tf[27]), zmax = funkey(tf[28]);
    const float sx = xmax - xmin, sz = zmax - zmin;
    if (!(isfinite(sx) && isfinite(sz) && isfinite(Ex) && isfinite(Ez))) return false;
    PairGrid g;
    g.x0 = xmin;
    g.z0 = zmin;
    g.wx = fmaxf(1.01f * Ex, 1e-4f);
    g.wz = fmaxf(1.01f * Ez, 1e-4f);
    g.kx = PairGrid::cells(sx, g.wx);
    g.kz = PairGrid::cells(sz, g.wz);
    while (static_cast<uint64_t>(g.kx) * g.kz > n - 1u) {  // offsets k <= K live in words 8k + 7, k < n
        if (g.kx >= g.kz) {
            g.wx *= 2.f;
            g.kx = PairGrid::cells(sx, g.wx);
        } else {
            g.wz *= 2.f;
            g.kz = PairGrid::cells(sz, g.wz);
        }
    }
    const uint32_t K = g.kx * g.kz;
    // The grid pays when a row's 3 x 3 cells hold far fewer boxes than the
    // row would test otherwise (about n / 2): 9 ns / K <= n / 8.  A compact
    // pile spans few cells and tests every pair.
    if (72ull * ns > static_cast<uint64_t>(n) * K) return false;
    // 3. counting sort of the small bodies by cell (a body's place in its
    //    cell is the count its histogram increment returned, kept in a
    //    register: at most kBatchMaxBodies / NT bodies per thread); the large
    //    ones follow
    constexpr uint32_t kPer = kBatchMaxBodies / NT;
    for (uint32_t c = r; c <= K; c += NT) bu[8 * c + 7] = 0u;
    tm.sync();
    uint32_t cell_of[kPer], place[kPer];
#pragma unroll
    for (uint32_t j = 0; j < kPer; ++j) {
        const uint32_t i = r + j * NT;
        cell_of[j] = ~0u;
        place[j] = 0u;
        if (i >= n) continue;
        const float4 lo = lds_lo(box, i);
        const float2 hi = lds_hi(box, i);
        if (small(lo.w - lo.x, hi.y - lo.z)) {
            cell_of[j] = g.cz(lo.z) * g.kx + g.cx(lo.x);
            place[j] = atomicAdd((uint32_t *)&bu[8 * cell_of[j] + 7], 1u);
        }
    }
    tm.sync();
    (void)team_scan(tm, (uint32_t *)(bu + 7), K + 1u, 8u);  // exclusive; offset K = ns
    tm.sync();
#pragma unroll
    for (uint32_t j = 0; j < kPer; ++j) {
        const uint32_t i = r + j * NT;
        if (i >= n) continue;
        const uint32_t p = cell_of[j] != ~0u ? bu[8 * cell_of[j] + 7] + place[j]
                                             : ns + atomicAdd((uint32_t *)&tf[30], 1u);
        bu[8 * p + 6] = i;
    }
    tm.sync();
    // 4. the rows of the small bodies.  A row of at most kRowList partners
    //    that fit its words is stored as their sorted list (S.deg[a] = 1),
    //    else as bit words; a large body's row is cleared here and filled by
    //    the whole team below.
    for (uint32_t a = r; a < n; a += NT) {
        const float4 lo = lds_lo(box, a);
        const float2 hi = lds_hi(box, a);
        const float q[6] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y};
        const uint32_t w0 = (a + 1) >> 5;
        uint32_t *row = S.bits + static_cast<size_t>(a) * W;
        const bool sm = small(lo.w - lo.x, hi.y - lo.z);
        uint32_t x0 = 0, x1 = 0, z0 = 1, z1 = 0;  // no cells for a large row
        if (sm) {
            const uint32_t cx = g.cx(lo.x), cz = g.cz(lo.z);
            x0 = cx ? cx - 1u : 0u;
            x1 = cx + 1u < g.kx ? cx + 1u : cx;
            z0 = cz ? cz - 1u : 0u;
            z1 = cz + 1u < g.kz ? cz + 1u : cz;
        }
        // visit(b) for every candidate partner b > a whose box touches a's
        auto each_hit = [&](auto visit) {
            // list positions [p, p1): four candidates' reads issued together
            auto range = [&](uint32_t p, uint32_t p1) {
                for (; p + 4u <= p1; p += 4u) {
                    uint32_t bb[4];
                    float4 lo4[4];
                    float2 hi4[4];
#pragma unroll
                    for (uint32_t j = 0; j < 4; ++j) bb[j] = bu[8 * (p + j) + 6];
#pragma unroll
                    for (uint32_t j = 0; j < 4; ++j) {
                        lo4[j] = lds_lo(box, bb[j]);
                        hi4[j] = lds_hi(box, bb[j]);
                    }
#pragma unroll
                    for (uint32_t j = 0; j < 4; ++j)
                        if (bb[j] > a && touch(q, lo4[j], hi4[j])) visit(bb[j]);
                }
                for (; p < p1; ++p) {
                    const uint32_t b = bu[8 * p + 6];
                    if (b > a && touch(q, lds_lo(box, b), lds_hi(box, b))) visit(b);
                }
            };
            for (uint32_t z = z0; z <= z1; ++z)
                range(bu[8 * (z * g.kx + x0) + 7], bu[8 * (z * g.kx + x1 + 1u) + 7]);
            if (sm) range(ns, n);
        };
        uint32_t h[kRowList];  // the first hits, ascending
        uint32_t cnt = 0;
        each_hit([&](uint32_t b) {
            uint32_t c = b;
#pragma unroll
            for (uint32_t j = 0; j < kRowList; ++j) {
                if (j < cnt) {
                    const uint32_t lo_ = min(h[j], c), hi_ = max(h[j], c);
                    h[j] = lo_;
                    c = hi_;
                } else if (j == cnt) {
                    h[j] = c;
                }
            }
            ++cnt;
        });
        const bool listed = sm && cnt <= kRowList && cnt <= W - w0;
        if (listed) {
#pragma unroll
            for (uint32_t j = 0; j < kRowList; ++j)
                if (j < cnt) row[w0 + j] = h[j];
        } else {
            for (uint32_t w = w0; w < W; ++w) row[w] = 0u;
            each_hit([&](uint32_t b) { row[b >> 5] |= 1u << (b & 31u); });
        }
        S.deg[a] = listed ? 1u : 0u;
        S.row[a] = cnt;
    }
    tm.sync();
    // 5. the rows of the large bodies, each by the whole team
    for (uint32_t p = ns; p < n; ++p) {
        const uint32_t a = bu[8 * p + 6];
        const float4 lo = lds_lo(box, a);
        const float2 hi = lds_hi(box, a);
        const float q[6] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y};
        uint32_t cnt = 0;
        for (uint32_t b = a + 1u + r; b < n; b += NT) {
            if (touch(q, lds_lo(box, b), lds_hi(box, b))) {
                atomicOr(&S.bits[static_cast<size_t>(a) * W + (b >> 5)], 1u << (b & 31u));
                ++cnt;
            }
        }
        if (cnt) atomicAdd(&S.row[a], cnt);
    }
    return true;
}

// Sort one body's contact list and record each contact's predecessor on that
// body: ascending contact index (reference order) or descending colouring
// priority (xpbd_solver.cpp:31-34,625-695).
__device__ __forceinline__ void body_predecessors(uint32_t x, const SceneScratch &S, const uint32_t *ca,
                                                  bool jp) {
    uint32_t *L = S.list + S.loff[x];
    const uint32_t d = S.deg[x];
    constexpr uint32_t kRegSort = 8;
    if (d <= kRegSort) {
        // in registers: the list's loads in flight together, an odd-even
        // transposition network (the keys are distinct, so the order is the
        // insertion sort's), then the stores
        uint32_t v[kRegSort];
        uint64_t key[kRegSort];
#pragma unroll
        for (uint32_t i = 0; i < kRegSort; ++i) v[i] = i < d ? L[i] : 0u;
#pragma unroll
        for (uint32_t i = 0; i < kRegSort; ++i) key[i] = i < d ? (jp ? ~jp_priority(v[i]) : v[i]) : ~0ull;
#pragma unroll
        for (uint32_t round = 0; round < kRegSort; ++round) {
#pragma unroll
            for (uint32_t i = round & 1u; i + 1 < kRegSort; i += 2) {
                if (key[i] > key[i + 1]) {  // strict: padding (~0) never passes a list entry
                    const uint64_t tk = key[i]; key[i] = key[i + 1]; key[i + 1] = tk;
                    const uint32_t tv = v[i]; v[i] = v[i + 1]; v[i + 1] = tv;
                }
            }
        }
        uint32_t own[kRegSort];
#pragma unroll
        for (uint32_t p = 0; p < kRegSort; ++p) own[p] = p < d ? ca[v[p]] : 0u;
#pragma unroll
        for (uint32_t p = 0; p < kRegSort; ++p) {
            if (p >= d) break;
            L[p] = v[p];
            const int32_t pred = p ? static_cast<int32_t>(v[p - 1]) : -1;
            if (own[p] == x) S.pred_a[v[p]] = pred;
            else S.pred_b[v[p]] = pred;
        }
        return;
    }
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

// Every body's predecessors, by the team.  Out of line: the in-register
// sort's arrays would otherwise raise the step kernel's register peak (and
// halve its occupancy); one call per solve.
template <class Team>
__device__ __attribute__((noinline)) void team_predecessors(const Team &tm, uint32_t n, const SceneScratch &S,
                                                            const uint32_t *ca, bool jp) {
    for (uint32_t x = tm.rank; x < n; x += Team::kSize) body_predecessors(x, S, ca, jp);
}

// Class assignment in rounds.  Returns the rounds run; `remaining` holds the
// contacts left unassigned (non-zero only for the coloured path at 256 rounds
// or after a colour overflow, `over`).  Every exit condition is a team-wide
// total, so all threads leave the loop together.
template <class Team>
__device__ uint32_t assign_rounds(const Team &tm, uint32_t nc, bool jp, const SceneScratch &S,
                                  const dy_contact_view &cv, const float *im, uint32_t &remaining, uint32_t &over,
                                  uint32_t &err) {
    uint32_t r = 0;
    remaining = nc;
    over = 0;
    constexpr uint32_t kChunk = 8;  // contacts a thread decides on before committing any
    while (remaining > 0) {
        if (jp && r >= kBatchMaxRounds) break;
        uint32_t mine = 0, my_over = 0;
        for (uint32_t base = tm.rank; base < nc; base += kChunk * Team::kSize) {
            // decide: which of this thread's next kChunk contacts have every
            // predecessor assigned in an earlier round (independent loads,
            // in flight together; a predecessor assigned in this round reads
            // as r or as unassigned)
            uint32_t ready = 0;
#pragma unroll
            for (uint32_t j = 0; j < kChunk; ++j) {
                const uint32_t k = base + j * Team::kSize;
                if (k >= nc) break;
                if (__hip_atomic_load(&S.rnd[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) != kNotAssigned)
                    continue;
                const int32_t pa = S.pred_a[k], pb = S.pred_b[k];
                const bool ok_a =
                    pa < 0 || __hip_atomic_load(&S.rnd[pa], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) < r;
                const bool ok_b =
                    pb < 0 || __hip_atomic_load(&S.rnd[pb], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) < r;
                if (ok_a && ok_b) ready |= 1u << j;
            }
            // commit the ready ones (contacts of one round share no dynamic body)
            for (uint32_t rem_ = ready; rem_; rem_ &= rem_ - 1u) {
            const uint32_t k = base + static_cast<uint32_t>(__builtin_ctz(rem_)) * Team::kSize;
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
        }
        uint32_t any_over = 0;
        const uint32_t got = tm.sum_or(mine, my_over, any_over);
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

// A record the step writes once and only the reverse sweep reads, much later:
// a non-temporal store, so that it does not displace the step's working set.
typedef float v4f_t __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void nt_store4(float4 *p, float a, float b, float c, float d) {
    const v4f_t v = {a, b, c, d};
    __builtin_nontemporal_store(v, reinterpret_cast<v4f_t *>(p));
}

// One scene for the frames of the launch.  RES: the scene's 20 body fields
// live in the team's LDS region from the first substep to the last and are
// written back at the end; otherwise the stages read and write the global
// arrays.  Per substep: integrate, leaf boxes, pairs, narrowphase, solve.
template <class Team, bool RES, bool TIMED, bool TAPE>
__device__ void run_scene(const Team &tm, const BatchArgs &A, uint32_t s, float *region) {
    constexpr uint32_t NT = Team::kSize;
    const uint32_t r = tm.rank;
    uint32_t *tf = tm.sh;
    const uint32_t b0 = A.body_off[s], n = A.body_off[s + 1] - b0;
    const uint32_t P = A.pair_off[s + 1] - A.pair_off[s];
    const uint32_t C = A.contact_off[s + 1] - A.contact_off[s];
    const dy_body_view gv = scene_bodies(A.bv, b0, n);
    const dy_contact_view cv = scene_contacts(A.cv, A.contact_off[s], C);
    dy_contact_pair *pairs = A.pairs + A.pair_off[s];
    const SceneScratch S = carve(A.scratch + A.scratch_off[s], n, C);
    uint32_t *st = A.stats + static_cast<size_t>(s) * kStWords;
    float *box = region;
    dy_body_view bv = gv;
    if constexpr (RES) {
        float *p = region + kBoxFloats * n;
#define DY_TO_LDS(f, k)                                              \
    bv.f = p + (k) * n;                                              \
    for (uint32_t i = r; i < n; i += NT) bv.f[i] = gv.f[i];
        DY_TO_LDS(pos_x, 0) DY_TO_LDS(pos_y, 1) DY_TO_LDS(pos_z, 2)
        DY_TO_LDS(rot_w, 3) DY_TO_LDS(rot_x, 4) DY_TO_LDS(rot_y, 5) DY_TO_LDS(rot_z, 6)
        DY_TO_LDS(vel_x, 7) DY_TO_LDS(vel_y, 8) DY_TO_LDS(vel_z, 9)
        DY_TO_LDS(ang_x, 10) DY_TO_LDS(ang_y, 11) DY_TO_LDS(ang_z, 12)
        DY_TO_LDS(inv_mass, 13)
        DY_TO_LDS(iI_xx, 14) DY_TO_LDS(iI_yy, 15) DY_TO_LDS(iI_zz, 16)
        DY_TO_LDS(iI_xy, 17) DY_TO_LDS(iI_xz, 18) DY_TO_LDS(iI_yz, 19)
#undef DY_TO_LDS
        tm.sync();
    }
    const float h = A.h;
    const uint32_t W = (n + 31u) >> 5;
    const uint32_t n_steps = static_cast<uint32_t>(A.n_frames) * static_cast<uint32_t>(A.n_sub);
    uint64_t clk[kStages + 1] = {};
    auto mark = [&](int j) {
        if constexpr (TIMED) {
            if (r == 0) clk[j] = wall_clock64();
        }
    };

    [[maybe_unused]] TapeLayout TL{};
    [[maybe_unused]] float *tp = nullptr;
    [[maybe_unused]] uint32_t Ct = 0;  // contacts a tape step holds
    if constexpr (TAPE) {
        Ct = A.tape_cap[s];
        TL = tape_layout(n, Ct, static_cast<uint32_t>(A.n_iters));
        tp = A.tape + A.tape_off[s];
    }
    for (uint32_t step = 0; step < n_steps; ++step) {
        mark(0);
        [[maybe_unused]] float *fb = nullptr;
        if constexpr (TAPE) {
            fb = tp + step * TL.frame;
            float *st0 = fb + TL.st;
            for (uint32_t i = r; i < n; i += NT) {
                st0[0 * n + i] = bv.pos_x[i]; st0[1 * n + i] = bv.pos_y[i]; st0[2 * n + i] = bv.pos_z[i];
                st0[3 * n + i] = bv.rot_w[i]; st0[4 * n + i] = bv.rot_x[i]; st0[5 * n + i] = bv.rot_y[i];
                st0[6 * n + i] = bv.rot_z[i];
                st0[7 * n + i] = bv.vel_x[i]; st0[8 * n + i] = bv.vel_y[i]; st0[9 * n + i] = bv.vel_z[i];
                st0[10 * n + i] = bv.ang_x[i]; st0[11 * n + i] = bv.ang_y[i]; st0[12 * n + i] = bv.ang_z[i];
            }
        }
        // 1. integrate one substep (integrator.cpp:17-70, n_substeps = 1 at h).
        for (uint32_t i = r; i < n; i += NT) {
            if constexpr (TAPE) {
                float *po = fb + TL.po;
                po[0 * n + i] = bv.pos_x[i]; po[1 * n + i] = bv.pos_y[i]; po[2 * n + i] = bv.pos_z[i];
                po[3 * n + i] = bv.rot_w[i]; po[4 * n + i] = bv.rot_x[i]; po[5 * n + i] = bv.rot_y[i];
                po[6 * n + i] = bv.rot_z[i];
            }
            if (bv.inv_mass[i] == 0.f || (gv.flags[i] & kFlagSleeping)) continue;
            Kin k{bv.pos_x[i], bv.pos_y[i], bv.pos_z[i], bv.rot_w[i], bv.rot_x[i], bv.rot_y[i],
                  bv.rot_z[i], bv.vel_x[i], bv.vel_y[i], bv.vel_z[i], bv.ang_x[i], bv.ang_y[i], bv.ang_z[i]};
            integrate_body(k, A.gx, A.gy, A.gz, h, 1);
            bv.pos_x[i] = k.px; bv.pos_y[i] = k.py; bv.pos_z[i] = k.pz;
            bv.rot_w[i] = k.qw; bv.rot_x[i] = k.qx; bv.rot_y[i] = k.qy; bv.rot_z[i] = k.qz;
            bv.vel_x[i] = k.vx; bv.vel_y[i] = k.vy; bv.vel_z[i] = k.vz;
            if constexpr (TAPE) {
                float *po = fb + TL.po;
                po[0 * n + i] = k.px; po[1 * n + i] = k.py; po[2 * n + i] = k.pz;
                po[3 * n + i] = k.qw; po[4 * n + i] = k.qx; po[5 * n + i] = k.qy; po[6 * n + i] = k.qz;
            }
        }
        // 2. leaf boxes (broadphase.cpp:215-251), binary16-rounded.
        for (uint32_t i = r; i < n; i += NT) {
            const uint32_t sh = gv.shape[i];
            float ex, ey, ez;
            leaf_extent(A.sv.type[sh], A.sv.half_x[sh], A.sv.half_y[sh], A.sv.half_z[sh], bv.rot_w[i],
                        bv.rot_x[i], bv.rot_y[i], bv.rot_z[i], ex, ey, ez);
            const float px = bv.pos_x[i], py = bv.pos_y[i], pz = bv.pos_z[i];
            float4 lo, hi;
            lo.x = f16_to_f32(f32_to_f16(px - ex));
            lo.y = f16_to_f32(f32_to_f16(py - ey));
            lo.z = f16_to_f32(f32_to_f16(pz - ez));
            lo.w = f16_to_f32(f32_to_f16(px + ex));
            hi.x = f16_to_f32(f32_to_f16(py + ey));
            hi.y = f16_to_f32(f32_to_f16(pz + ez));
            hi.z = 0.f;
            hi.w = 0.f;
            reinterpret_cast<float4 *>(box)[2 * i] = lo;
            reinterpret_cast<float4 *>(box)[2 * i + 1] = hi;
        }
        if (r < 32) tf[r] = 0;
        tm.sync();
        mark(1);

        // 3. candidate pairs.  Lanes hold rows a; every lane tests the same
        //    partner b (an LDS broadcast), building 32-bit overlap words of
        //    the row; then a scan of the row counts and an emit pass over the
        //    words, in (a, b) order.  A scene of kGridMinBodies or more
        //    stepped by a workgroup fills the same words from a grid of its
        //    boxes instead (for a single wavefront's scenes, at most a few
        //    hundred bodies, testing every pair measured faster).
        bool gridded = false;
        if constexpr (NT > 64)
            gridded = n >= kGridMinBodies && grid_rows(tm, (lds_f32 *)box, n, W, S, (lds_u32 *)tf);
        for (uint32_t base = 0; base < n && !gridded; base += NT) {
            const uint32_t a = base + r;
            const bool have = a < n;
            float q[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
            if (have) {
                const float4 lo = reinterpret_cast<const float4 *>(box)[2 * a];
                const float4 hi = reinterpret_cast<const float4 *>(box)[2 * a + 1];
                q[0] = lo.x; q[1] = lo.y; q[2] = lo.z; q[3] = lo.w; q[4] = hi.x; q[5] = hi.y;
            }
            uint32_t cnt = 0;
            for (uint32_t w = (base + 1) >> 5; w < W; ++w) {
                const uint32_t b_lo = w << 5;
                const uint32_t m = n - b_lo < 32u ? n - b_lo : 32u;
                uint32_t word = 0;
                for (uint32_t j = 0; j < m; ++j) {
                    const float4 lo = reinterpret_cast<const float4 *>(box)[2 * (b_lo + j)];
                    const float2 hi = reinterpret_cast<const float2 *>(box)[4 * (b_lo + j) + 2];
                    word |= (touch(q, lo, hi) ? 1u : 0u) << j;
                }
                // keep partners b > a
                const int32_t lo_bit = static_cast<int32_t>(a) + 1 - static_cast<int32_t>(b_lo);
                const uint32_t keep = lo_bit <= 0 ? ~0u : (lo_bit >= 32 ? 0u : (~0u << lo_bit));
                word &= keep;
                if (have) S.bits[static_cast<size_t>(a) * W + w] = word;
                cnt += static_cast<uint32_t>(__popc(word));
            }
            if (have) S.row[a] = cnt;
        }
        tm.sync();
        const uint32_t n_pairs = team_scan(tm, S.row, n);
        tm.sync();
        for (uint32_t a = r; a < n; a += NT) {
            uint32_t pos = S.row[a];
            if (gridded && S.deg[a]) {  // a listed row: its partners, ascending
                const uint32_t *row = S.bits + static_cast<size_t>(a) * W + ((a + 1) >> 5);
                const uint32_t cnt = (a + 1 < n ? S.row[a + 1] : n_pairs) - pos;
                for (uint32_t j = 0; j < cnt && pos < P; ++j) pairs[pos++] = dy_contact_pair{a, row[j]};
                continue;
            }
            for (uint32_t w = (a + 1) >> 5; w < W && pos < P; ++w) {
                uint32_t word = S.bits[static_cast<size_t>(a) * W + w];
                while (word && pos < P) {
                    const uint32_t j = static_cast<uint32_t>(__builtin_ctz(word));
                    word &= word - 1u;
                    pairs[pos++] = dy_contact_pair{a, (w << 5) + j};
                }
            }
        }
        tm.sync();
        mark(2);

        // 4. narrowphase in pair order.
        const uint32_t np_n = n_pairs < P ? n_pairs : P;
        uint32_t run = 0;
        for (uint32_t base = 0; base < np_n; base += NT) {
            const uint32_t i = base + r;
            uint32_t hit = 0, unsup = 0;
            Hit hc;
            dy_contact_pair p{0u, 0u};
            if (i < np_n) {
                p = pairs[i];
                dy_body_view lv = bv;
                lv.shape = gv.shape;
                hit = pair_contact(lv, A.sv, p.a, p.b, hc, unsup) ? 1u : 0u;
            }
            if (unsup) atomicOr(&tf[20], kBatchErrUnsupported);
            uint32_t tot = 0;
            const uint32_t ex = tm.excl_scan(hit, tot);
            const uint32_t slot = run + ex;
            if (hit && slot < C) {
                cv.body_a[slot] = p.a; cv.body_b[slot] = p.b;
                cv.pos_x[slot] = hc.px; cv.pos_y[slot] = hc.py; cv.pos_z[slot] = hc.pz;
                cv.norm_x[slot] = hc.nx; cv.norm_y[slot] = hc.ny; cv.norm_z[slot] = hc.nz;
                cv.depth[slot] = hc.depth;
                if constexpr (TAPE) S.feat[slot] = hc.feat;
            }
            run += tot;
        }
        const uint32_t nc = run < C ? run : C;
        mark(3);
        mark(4);
        mark(5);

        // 5. solve (xpbd_solver.cpp:409-733), contacts only.
        uint32_t serial = 1, rounds = 0, n_cls = 0, overflow = 0, err = 0, n_dyn_end = 0;
        uint32_t live_its = 0;  // bit min(it, 31): this thread projected a live contact in iteration it
        if (nc > 0) {
            for (uint32_t x = r; x < n; x += NT) { S.deg[x] = 0; S.fill[x] = 0; S.used[x] = 0; }
            for (uint32_t k = r; k < nc; k += NT) {
                S.lam[k] = 0.f;
                S.rnd[k] = kNotAssigned;
                S.pred_a[k] = -1;
                S.pred_b[k] = -1;
                S.cls[k] = 0;
            }
            tm.sync();
            for (uint32_t k = r; k < nc; k += NT) {
                const uint32_t a = cv.body_a[k], b = cv.body_b[k];
                if (bv.inv_mass[a] != 0.f) atomicAdd(&S.deg[a], 1u);
                if (bv.inv_mass[b] != 0.f) atomicAdd(&S.deg[b], 1u);
            }
            tm.sync();
            for (uint32_t x = r; x < n; x += NT) S.loff[x] = S.deg[x];
            tm.sync();
            n_dyn_end = team_scan(tm, S.loff, n);
            tm.sync();
            for (uint32_t k = r; k < nc; k += NT) {
                const uint32_t a = cv.body_a[k], b = cv.body_b[k];
                if (bv.inv_mass[a] != 0.f) S.list[S.loff[a] + atomicAdd(&S.fill[a], 1u)] = k;
                if (bv.inv_mass[b] != 0.f) S.list[S.loff[b] + atomicAdd(&S.fill[b], 1u)] = k;
            }
            tm.sync();
            const bool jp = !(A.force_serial || nc <= kBatchSerialMax);
            team_predecessors(tm, n, S, cv.body_a, jp);
            tm.sync();
            uint32_t remaining = 0, over = 0;
            rounds = assign_rounds(tm, nc, jp, S, cv, bv.inv_mass, remaining, over, err);
            overflow = over;
            serial = jp ? 0u : 1u;
            if (jp && (remaining > 0 || over)) {
                // xpbd_solver.cpp:696-699: colouring failed -> reference order.
                serial = 1;
                for (uint32_t k = r; k < nc; k += NT) S.rnd[k] = kNotAssigned;
                tm.sync();
                team_predecessors(tm, n, S, cv.body_a, false);
                tm.sync();
                (void)assign_rounds(tm, nc, false, S, cv, bv.inv_mass, remaining, over, err);
            }
            if (remaining > 0) err |= kBatchErrLists;
            if (err == 0) {
                for (uint32_t k = r; k < nc; k += NT) atomicMax(&tf[21], S.cls[k] + 1u);
            }
            tm.sync();
            n_cls = tf[21];
            if (err == 0) {
                // classes -> contiguous ranges of `order`
                for (uint32_t c = r; c < n_cls; c += NT) S.ccnt[c] = 0;
                tm.sync();
                for (uint32_t k = r; k < nc; k += NT) atomicAdd(&S.ccnt[S.cls[k]], 1u);
                tm.sync();
                for (uint32_t c = r; c < n_cls; c += NT) S.coff[c] = S.ccnt[c];
                tm.sync();
                (void)team_scan(tm, S.coff, n_cls);
                if (r == 0) S.coff[n_cls] = nc;
                for (uint32_t c = r; c < n_cls; c += NT) S.ccnt[c] = 0;
                tm.sync();
                for (uint32_t k = r; k < nc; k += NT) {
                    const uint32_t c = S.cls[k];
                    S.order[S.coff[c] + atomicAdd(&S.ccnt[c], 1u)] = k;
                }
                [[maybe_unused]] bool taped = false;
                if constexpr (TAPE) {
                    taped = nc <= Ct;
                    if (!taped && r == 0) atomicOr(&tf[20], kBatchErrTape);
                    // the record gathers `order`, which other wavefronts of
                    // a workgroup team have just scattered
                    tm.sync();
                }
                if (taped) {
                    uint32_t *co = reinterpret_cast<uint32_t *>(fb + TL.co);
                    for (uint32_t c = r; c <= n_cls; c += NT) co[c] = S.coff[c];
                    float *rec = fb + TL.rec;
                    uint32_t *ab = reinterpret_cast<uint32_t *>(fb + TL.ab);
                    for (uint32_t j = r; j < nc; j += NT) {
                        const uint32_t k = S.order[j];
                        reinterpret_cast<float4 *>(rec)[2 * j] =
                            make_float4(cv.pos_x[k], cv.pos_y[k], cv.pos_z[k], cv.norm_x[k]);
                        reinterpret_cast<float4 *>(rec)[2 * j + 1] =
                            make_float4(cv.norm_y[k], cv.norm_z[k], cv.depth[k], __uint_as_float(S.feat[k]));
                        ab[2 * j] = cv.body_a[k];
                        ab[2 * j + 1] = cv.body_b[k];
                    }
                }
                tm.sync();
                mark(4);
                // I iterations: classes in order, then renormalise every body
                // (xpbd_solver.cpp:578-586,700-733).
                uint32_t still_live = 0;  // a contact left live by its iteration-0 projection
                for (int it = 0; it < A.n_iters; ++it) {
                    if (it > 0 && tf[19] == 0u) {
                        // A contact's lambda changes only in its own projection,
                        // which sets it to the depth unless project_contact
                        // declines to move the bodies; with no contact left live
                        // after iteration 0 the later iterations project nothing
                        // and only their renormalisations remain.
                        for (uint32_t i = r; i < n; i += NT) {
                            float qw = bv.rot_w[i], qx = bv.rot_x[i], qy = bv.rot_y[i], qz = bv.rot_z[i];
                            for (int k = it; k < A.n_iters; ++k) {
                                if constexpr (TAPE) {
                                    float *qr = fb + TL.qr + 4ull * n * static_cast<uint32_t>(k);
                                    qr[i] = qw; qr[n + i] = qx; qr[2 * n + i] = qy; qr[3 * n + i] = qz;
                                }
                                renormalise(qw, qx, qy, qz);
                            }
                            bv.rot_w[i] = qw; bv.rot_x[i] = qx; bv.rot_y[i] = qy; bv.rot_z[i] = qz;
                        }
                        tm.sync();
                        break;
                    }
                    for (uint32_t c = 0; c < n_cls; ++c) {
                        const uint32_t beg = S.coff[c], end = S.coff[c + 1];
                        for (uint32_t j = beg + r; j < end; j += NT) {
                            const uint32_t k = S.order[j];
                            const uint32_t a = cv.body_a[k], b = cv.body_b[k];
                            const ContactIn ci{cv.pos_x[k], cv.pos_y[k], cv.pos_z[k], cv.norm_x[k],
                                               cv.norm_y[k], cv.norm_z[k], cv.depth[k]};
                            float l = S.lam[k];
                            if (taped) fb[TL.lam + static_cast<uint64_t>(it) * Ct + j] = l;
                            // project_contact returns at once, changing nothing, unless the
                            // contact is live (its depth not yet covered by lambda: after its
                            // first projection lambda == depth for the rest of the substep)
                            if (!contact_live(ci.depth, l)) continue;
                            live_its |= 1u << (it < 31 ? it : 31);
                            BodyState Ba = load_body(bv, a), Bb = load_body(bv, b);
                            if (taped) {
                                float4 *pr = reinterpret_cast<float4 *>(
                                    fb + TL.pre + 28ull * (static_cast<uint64_t>(it) * Ct + j));
                                nt_store4(pr + 0, Ba.px, Ba.py, Ba.pz, Ba.qw);
                                nt_store4(pr + 1, Ba.qx, Ba.qy, Ba.qz, Ba.vx);
                                nt_store4(pr + 2, Ba.vy, Ba.vz, Ba.wx, Ba.wy);
                                nt_store4(pr + 3, Ba.wz, Bb.px, Bb.py, Bb.pz);
                                nt_store4(pr + 4, Bb.qw, Bb.qx, Bb.qy, Bb.qz);
                                nt_store4(pr + 5, Bb.vx, Bb.vy, Bb.vz, Bb.wx);
                                nt_store4(pr + 6, Bb.wy, Bb.wz, l, 0.f);
                            }
                            if (project_contact(ci, Ba, Bb, A.mu, l, h)) {
                                S.lam[k] = l;
                                if (Ba.im != 0.f) store_body(bv, a, Ba);
                                if (Bb.im != 0.f) store_body(bv, b, Bb);
                            }
                            if (it == 0 && contact_live(ci.depth, l)) still_live = 1u;
                        }
                        tm.sync();
                    }
                    if (it == 0 && still_live) atomicOr(&tf[19], 1u);
                    for (uint32_t i = r; i < n; i += NT) {
                        float qw = bv.rot_w[i], qx = bv.rot_x[i], qy = bv.rot_y[i], qz = bv.rot_z[i];
                        if constexpr (TAPE) {
                            float *qr = fb + TL.qr + 4ull * n * static_cast<uint32_t>(it);
                            qr[i] = qw; qr[n + i] = qx; qr[2 * n + i] = qy; qr[3 * n + i] = qz;
                        }
                        renormalise(qw, qx, qy, qz);
                        bv.rot_w[i] = qw; bv.rot_x[i] = qx; bv.rot_y[i] = qy; bv.rot_z[i] = qz;
                    }
                    tm.sync();
                }
            }
        }
        if (err) atomicOr(&tf[20], err);
        if (live_its) atomicOr(&tf[22], live_its);
        tm.sync();
        mark(5);
        if constexpr (TAPE) {
            if (r == 0) {
                const uint32_t terr = err | (nc > Ct ? kBatchErrTape : 0u);
                uint32_t *hdr = reinterpret_cast<uint32_t *>(fb + TL.hdr);
                hdr[0] = terr ? 0u : nc;
                hdr[1] = terr ? 0u : n_cls;
                hdr[2] = terr;
                hdr[3] = tf[22];
            }
        }
        if (r == 0) {
            if constexpr (TIMED) {
                for (uint32_t j = 0; j < kStages; ++j) {
                    const uint64_t x0 = clk[j], x1 = clk[j + 1];
                    A.stage_clk[static_cast<size_t>(s) * kStages + j] += x1 > x0 ? x1 - x0 : 0ull;
                }
            }
            st[kStPairs] = n_pairs;
            st[kStContacts] = run;
            st[kStClasses] = n_cls;
            st[kStSerial] = nc > 0 ? serial : 0u;
            st[kStRounds] = rounds;
            st[kStOverflow] = overflow;
            st[kStError] |= tf[20];
            st[kStSubsteps] += 1u;
            st[kStSumPairs] += n_pairs;
            st[kStSumContacts] += nc;
            st[kStSumDyn] += n_dyn_end;
            st[kStSumSolves] += nc > 0 ? 1u : 0u;
            st[kStGrid] += gridded ? 1u : 0u;
        }
        tm.sync();
    }
    if constexpr (RES) {
        // write back the mutable fields: pos, rot, vel, ang
        for (uint32_t i = r; i < n && !A.keep_input; i += NT) {
            gv.pos_x[i] = bv.pos_x[i]; gv.pos_y[i] = bv.pos_y[i]; gv.pos_z[i] = bv.pos_z[i];
            gv.rot_w[i] = bv.rot_w[i]; gv.rot_x[i] = bv.rot_x[i]; gv.rot_y[i] = bv.rot_y[i];
            gv.rot_z[i] = bv.rot_z[i];
            gv.vel_x[i] = bv.vel_x[i]; gv.vel_y[i] = bv.vel_y[i]; gv.vel_z[i] = bv.vel_z[i];
            gv.ang_x[i] = bv.ang_x[i]; gv.ang_y[i] = bv.ang_y[i]; gv.ang_z[i] = bv.ang_z[i];
        }
    }
}

// One workgroup per unit: either one scene stepped by the whole workgroup,
// or one scene per wavefront (each wavefront its own team and LDS region).
// RES: every scene of the launch keeps its state in LDS (the usual case);
// otherwise the launch holds the scenes too large for that, one per
// workgroup (or per 64-thread workgroup).
template <uint32_t NT, bool TIMED, bool RES, bool TAPE>
__global__ __launch_bounds__(NT) void batch_frames(BatchArgs A) {
    extern __shared__ float lds[];
    __shared__ uint32_t team_sh[NT / 64][32];
    const uint4 u = A.units[blockIdx.x];
    const uint32_t kind = A.unit_kind[blockIdx.x];
    if constexpr (NT > 64) {
        if (kind & kUnitBlock) {
            const BlockTeam<NT> tm{threadIdx.x, team_sh[0]};
            run_scene<BlockTeam<NT>, RES, TIMED, TAPE>(tm, A, u.x, lds);
            return;
        }
    }
    if constexpr (RES || NT == 64) {
        // Wave-uniform by construction; readfirstlane lets the compiler keep
        // the scene's pointers in scalar registers.
        const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
        const uint32_t sc = __builtin_amdgcn_readfirstlane(w == 0 ? u.x : w == 1 ? u.y : w == 2 ? u.z : u.w);
        if (sc == kNoScene) return;  // wavefront teams never wait on each other
        const WaveTeam tm{lane_id(), team_sh[w]};
        run_scene<WaveTeam, RES, TIMED, TAPE>(tm, A, sc, lds + static_cast<size_t>(w) * A.region_floats);
    }
}

// Dynamic LDS beyond 64 KiB (gfx950 gives a workgroup up to 160 KB) is
// declared to the runtime once per kernel instance.
template <class K>
int allow_lds(K kernel, size_t lds_bytes) {
    if (lds_bytes <= 65536) return DY_OK;
    DY_TRY(hipFuncSetAttribute(reinterpret_cast<const void *>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                               static_cast<int>(lds_bytes)));
    return DY_OK;
}

template <uint32_t NT, bool RES>
int launch_frames(const BatchArgs &a, uint32_t n_units, size_t lds_bytes, hipStream_t s) {
    if (n_units == 0) return DY_OK;
    if (a.tape) {
        DY_TRY_STATUS(allow_lds(batch_frames<NT, false, RES, true>, lds_bytes));
        hipLaunchKernelGGL((batch_frames<NT, false, RES, true>), dim3(n_units), dim3(NT), lds_bytes, s, a);
    } else if (a.stage_clk) {
        DY_TRY_STATUS(allow_lds(batch_frames<NT, true, RES, false>, lds_bytes));
        hipLaunchKernelGGL((batch_frames<NT, true, RES, false>), dim3(n_units), dim3(NT), lds_bytes, s, a);
    } else {
        DY_TRY_STATUS(allow_lds(batch_frames<NT, false, RES, false>, lds_bytes));
        hipLaunchKernelGGL((batch_frames<NT, false, RES, false>), dim3(n_units), dim3(NT), lds_bytes, s, a);
    }
    DY_LAUNCH_CHECK();
    return DY_OK;
}

// ---------------------------------------------------------------------------
// Reverse sweep of recorded frames (the differentiable form)
// ---------------------------------------------------------------------------
// Per-contact floats of the reverse sweep: lambda adjoint, the contact's
// (p, n, depth) adjoint, and its pose adjoints for both bodies (7 + 7).
constexpr uint32_t kCsStride = 22;

struct DiffArgs {
    dy_body_view adj;  // adjoint of the 13 mutable fields; inv_mass holds the inverse-mass adjoint
    float *cs;         // kCsStride per tape contact: 8 accumulators, then 14 pose gradients
    const uint64_t *cs_off;  // per scene, first tape contact (kCsStride floats each)
    uint32_t frames;   // recorded steps to reverse, newest first
    uint32_t lds_floats;  // dynamic LDS per workgroup
    uint32_t acc_lds;  // contacts whose 8 accumulators live in LDS after the body adjoints (the rest: cs)
    const uint32_t *scenes;  // the scenes this launch reverses, one per workgroup
    uint32_t *err;     // OR of the reversed steps' header errors
};

// The tape's inputs of one contact projection, loaded one phase ahead.
struct PreRec {
    uint32_t j, a, b;
    bool live;    // the projection changed something (else its adjoint is zero)
    float4 r0, r1;
    float4 m[7];  // both bodies' 13 fields and lambda (28 floats), live only
};

__device__ __forceinline__ void pre_fetch(PreRec &P, const float *fb, const TapeLayout &TL, const uint32_t *co,
                                          uint32_t C, uint32_t it, uint32_t c, uint32_t r) {
    P.j = co[c] + r;
    P.live = false;
    if (P.j >= co[c + 1]) return;
    const float4 *rec = reinterpret_cast<const float4 *>(fb + TL.rec);
    const uint64_t at = static_cast<uint64_t>(it) * C + P.j;
    P.r1 = rec[2 * P.j + 1];
    P.live = contact_live(P.r1.z, fb[TL.lam + at]);
    if (!P.live) return;
    const uint32_t *ab = reinterpret_cast<const uint32_t *>(fb + TL.ab);
    const float4 *pr = reinterpret_cast<const float4 *>(fb + TL.pre + 28ull * at);
    P.a = ab[2 * P.j];
    P.b = ab[2 * P.j + 1];
    P.r0 = rec[2 * P.j];
#pragma unroll
    for (int k = 0; k < 7; ++k) P.m[k] = pr[k];
}

__device__ __forceinline__ BodyBar<float> load_bar(const dy_body_view &v, uint32_t i) {
    BodyBar<float> b;
    b.p[0] = v.pos_x[i]; b.p[1] = v.pos_y[i]; b.p[2] = v.pos_z[i];
    b.q[0] = v.rot_w[i]; b.q[1] = v.rot_x[i]; b.q[2] = v.rot_y[i]; b.q[3] = v.rot_z[i];
    b.v[0] = v.vel_x[i]; b.v[1] = v.vel_y[i]; b.v[2] = v.vel_z[i];
    b.w[0] = v.ang_x[i]; b.w[1] = v.ang_y[i]; b.w[2] = v.ang_z[i];
    b.im = v.inv_mass[i];
    return b;
}

__device__ __forceinline__ void store_bar(const dy_body_view &v, uint32_t i, const BodyBar<float> &b) {
    v.pos_x[i] = b.p[0]; v.pos_y[i] = b.p[1]; v.pos_z[i] = b.p[2];
    v.rot_w[i] = b.q[0]; v.rot_x[i] = b.q[1]; v.rot_y[i] = b.q[2]; v.rot_z[i] = b.q[3];
    v.vel_x[i] = b.v[0]; v.vel_y[i] = b.v[1]; v.vel_z[i] = b.v[2];
    v.ang_x[i] = b.w[0]; v.ang_y[i] = b.w[1]; v.ang_z[i] = b.w[2];
    v.inv_mass[i] = b.im;
}

// Body state as a projection input: 13 recorded fields + the constants.
__device__ __forceinline__ BodyT<float> body_in(const float *m, const dy_body_view &g, uint32_t i) {
    BodyT<float> b;
    for (int k = 0; k < 3; ++k) { b.p[k] = m[k]; b.v[k] = m[7 + k]; b.w[k] = m[10 + k]; }
    for (int k = 0; k < 4; ++k) b.q[k] = m[3 + k];
    b.im = g.inv_mass[i];
    b.I[0] = g.iI_xx[i]; b.I[1] = g.iI_yy[i]; b.I[2] = g.iI_zz[i];
    b.I[3] = g.iI_xy[i]; b.I[4] = g.iI_xz[i]; b.I[5] = g.iI_yz[i];
    return b;
}

// One wavefront per scene reverses its recorded frames, newest first: per
// frame the solver iterations in reverse (each iteration's renormalisation,
// then its classes last to first, contacts of a class in parallel: they
// share no dynamic body), then the contact geometry (class by class), then
// the integrator.  Static bodies receive no adjoint.
// One contact's projection adjoint: inputs m (28 floats: both bodies' 13
// fields and lambda), accumulators acc (lambda adjoint, then the contact's
// p, n, depth adjoints), body adjoints in LDS.
__device__ __forceinline__ void contact_phase_adj(const BatchArgs &A, float h, const dy_body_view &gv,
                                                  dy_body_view &av, uint32_t a, uint32_t b, const float4 &r0,
                                                  const float4 &r1, const float *m, float *acc) {
    ContactT<float> ct;
    ct.p[0] = r0.x; ct.p[1] = r0.y; ct.p[2] = r0.z;
    ct.n[0] = r0.w; ct.n[1] = r1.x; ct.n[2] = r1.y;
    ct.depth = r1.z;
    const BodyT<float> A0 = body_in(m, gv, a), B0 = body_in(m + 13, gv, b);
    BodyBar<float> Ab = load_bar(av, a), Bb = load_bar(av, b);
    float lb = acc[0];
    ContactT<float> cb;
    for (int k = 0; k < 3; ++k) { cb.p[k] = 0.f; cb.n[k] = 0.f; }
    cb.depth = 0.f;
    project_adj(ct, A0, B0, A.mu, m[26], h, Ab, Bb, lb, cb);
    acc[0] = lb;
    for (int k = 0; k < 3; ++k) {
        acc[1 + k] += cb.p[k];
        acc[4 + k] += cb.n[k];
    }
    acc[7] += cb.depth;
    if (A0.im != 0.f) store_bar(av, a, Ab);
    if (B0.im != 0.f) store_bar(av, b, Bb);
}

// The reverse sweep of one scene by a team (a wavefront, or a 256-thread
// workgroup for a larger scene).  !GLB: the scene's adjoint (13 fields +
// inverse mass) and its constants (inverse mass, inverse inertia) live in
// LDS during the launch, beside its first contacts' accumulators; GLB (a
// scene too large for that): they stay in global memory, and the team
// barrier is the workgroup's, which orders global memory as well.
template <class Team, bool GLB>
__device__ void reverse_scene(const Team &tm, const BatchArgs &A, const DiffArgs &D, uint32_t s, float *lds) {
    constexpr uint32_t NT = Team::kSize;
    const uint32_t r = tm.rank;
    const uint32_t b0 = A.body_off[s], n = A.body_off[s + 1] - b0;
    const dy_body_view gl = scene_bodies(A.bv, b0, n);
    const dy_body_view ag = scene_bodies(D.adj, b0, n);
    dy_body_view av = ag, gv = gl;
    if (!GLB) {
        float *p = lds;
#define DY_ADJ_LDS(dst, src, k)                                         \
    dst = p + (k) * n;                                                  \
    for (uint32_t i = r; i < n; i += NT) dst[i] = src[i];
        DY_ADJ_LDS(av.pos_x, ag.pos_x, 0) DY_ADJ_LDS(av.pos_y, ag.pos_y, 1) DY_ADJ_LDS(av.pos_z, ag.pos_z, 2)
        DY_ADJ_LDS(av.rot_w, ag.rot_w, 3) DY_ADJ_LDS(av.rot_x, ag.rot_x, 4) DY_ADJ_LDS(av.rot_y, ag.rot_y, 5)
        DY_ADJ_LDS(av.rot_z, ag.rot_z, 6)
        DY_ADJ_LDS(av.vel_x, ag.vel_x, 7) DY_ADJ_LDS(av.vel_y, ag.vel_y, 8) DY_ADJ_LDS(av.vel_z, ag.vel_z, 9)
        DY_ADJ_LDS(av.ang_x, ag.ang_x, 10) DY_ADJ_LDS(av.ang_y, ag.ang_y, 11) DY_ADJ_LDS(av.ang_z, ag.ang_z, 12)
        DY_ADJ_LDS(av.inv_mass, ag.inv_mass, 13)
        DY_ADJ_LDS(gv.inv_mass, gl.inv_mass, 14)
        DY_ADJ_LDS(gv.iI_xx, gl.iI_xx, 15) DY_ADJ_LDS(gv.iI_yy, gl.iI_yy, 16) DY_ADJ_LDS(gv.iI_zz, gl.iI_zz, 17)
        DY_ADJ_LDS(gv.iI_xy, gl.iI_xy, 18) DY_ADJ_LDS(gv.iI_xz, gl.iI_xz, 19) DY_ADJ_LDS(gv.iI_yz, gl.iI_yz, 20)
#undef DY_ADJ_LDS
    }
    tm.sync();
    const uint32_t iters = static_cast<uint32_t>(A.n_iters);
    const uint32_t C = A.tape_cap[s];  // contacts a tape step holds
    const TapeLayout TL = tape_layout(n, C, iters);
    const float *tp = A.tape + A.tape_off[s];
    float *cs = D.cs + kCsStride * D.cs_off[s];
    // per contact: 8 accumulators (the first n_lds in LDS, the others in cs),
    // 14 pose gradients (cs)
    float *const acc_l = lds + 21u * n;
    const uint32_t n_lds = GLB ? 0u : D.acc_lds;
    auto acc_of = [&](uint32_t j) -> float * { return j < n_lds ? acc_l + 8u * j : cs + 8u * j; };
    float *pg = cs + 8ull * C;
    const float h = A.h;
    const float g[3] = {A.gx, A.gy, A.gz};
    for (int f = static_cast<int>(D.frames) - 1; f >= 0; --f) {
        const float *fb = tp + static_cast<uint64_t>(f) * TL.frame;
        const uint32_t *hdr = reinterpret_cast<const uint32_t *>(fb + TL.hdr);
        const uint32_t nc = hdr[0], ncls = hdr[1];
        if (r == 0 && hdr[2] != 0u) atomicOr(D.err, hdr[2]);
        if (nc > 0) {
            const uint32_t *co = reinterpret_cast<const uint32_t *>(fb + TL.co);
            const uint32_t *ab = reinterpret_cast<const uint32_t *>(fb + TL.ab);
            for (uint32_t j = r; j < 8u * nc; j += NT) acc_of(j >> 3)[j & 7u] = 0.f;
            tm.sync();
            // Iterations newest first.  A projection changes nothing once its
            // contact's lambda covers the depth (after the contact's first
            // projection in a substep), so most iterations project nothing:
            // hdr[3] bit min(it, 31) says iteration it projected a live
            // contact.  The iterations down to the next live one need only
            // their renormalisation adjoints, applied per body in one pass;
            // then the live iteration's classes last to first, each lane's
            // first contact of the next class loaded while this one computes.
            const uint32_t live_its = hdr[3];
            auto is_live = [&](int it) { return ((live_its >> (it < 31 ? it : 31)) & 1u) != 0u; };
            for (int it = static_cast<int>(iters) - 1; it >= 0;) {
                int lo = it;
                while (lo > 0 && !is_live(lo)) --lo;
                const bool live = is_live(lo);
                PreRec cur, nxt;
                cur.j = ~0u;
                if (live) pre_fetch(cur, fb, TL, co, C, static_cast<uint32_t>(lo), ncls - 1u, r);
                for (uint32_t i = r; i < n; i += NT) {
                    if (gv.inv_mass[i] == 0.f) continue;
                    float qb[4] = {av.rot_w[i], av.rot_x[i], av.rot_y[i], av.rot_z[i]};
                    for (int k = it; k >= lo; --k) {
                        const float *qr = fb + TL.qr + 4ull * n * static_cast<uint32_t>(k);
                        const float q[4] = {qr[i], qr[n + i], qr[2 * n + i], qr[3 * n + i]};
                        renorm_adj(q, qb);
                    }
                    av.rot_w[i] = qb[0]; av.rot_x[i] = qb[1]; av.rot_y[i] = qb[2]; av.rot_z[i] = qb[3];
                }
                tm.sync();
                if (live) {
                    for (int c = static_cast<int>(ncls) - 1; c >= 0; --c) {
                        if (c > 0) pre_fetch(nxt, fb, TL, co, C, static_cast<uint32_t>(lo), static_cast<uint32_t>(c - 1), r);
                        const uint32_t end = co[c + 1];
                        if (cur.j < end && cur.live)
                            contact_phase_adj(A, h, gv, av, cur.a, cur.b, cur.r0, cur.r1,
                                              reinterpret_cast<const float *>(cur.m), acc_of(cur.j));
                        for (uint32_t j = cur.j + NT; j < end; j += NT) {  // classes larger than the team
                            PreRec x;
                            pre_fetch(x, fb, TL, co, C, static_cast<uint32_t>(lo), static_cast<uint32_t>(c),
                                      r + (j - cur.j));
                            if (x.live)
                                contact_phase_adj(A, h, gv, av, x.a, x.b, x.r0, x.r1,
                                                  reinterpret_cast<const float *>(x.m), acc_of(j));
                        }
                        tm.sync();
                        if (c > 0) cur = nxt;
                    }
                }
                it = lo - 1;
            }
            // contact geometry: (p, n, depth) adjoints to the post-integrate
            // poses, every contact in parallel (written out per contact kind,
            // the branches fixed by the forward's feature word), then added
            // to the bodies class by class
            const float *po = fb + TL.po;
            for (uint32_t j = r; j < nc; j += NT) {
                const uint32_t a = ab[2 * j], b = ab[2 * j + 1];
                const uint32_t sa = gl.shape[a], sb = gl.shape[b];
                const uint32_t ta = A.sv.type[sa], tb = A.sv.type[sb];
                const float ha[3] = {A.sv.half_x[sa], A.sv.half_y[sa], A.sv.half_z[sa]};
                const float hb[3] = {A.sv.half_x[sb], A.sv.half_y[sb], A.sv.half_z[sb]};
                const float pa[3] = {po[a], po[n + a], po[2 * n + a]};
                const float qa[4] = {po[3 * n + a], po[4 * n + a], po[5 * n + a], po[6 * n + a]};
                const float pb[3] = {po[b], po[n + b], po[2 * n + b]};
                const float qb[4] = {po[3 * n + b], po[4 * n + b], po[5 * n + b], po[6 * n + b]};
                float hbar[7];
                const float *aj = acc_of(j);
                for (int k = 0; k < 7; ++k) hbar[k] = aj[1 + k];
                const uint32_t feat = __float_as_uint(reinterpret_cast<const float4 *>(fb + TL.rec)[2 * j + 1].w);
                float ga[7] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f}, gb2[7] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
                geom_adj(ta, tb, feat, ha, hb, pa, qa, pb, qb, hbar, ga, gb2);
                for (int k = 0; k < 7; ++k) {
                    pg[14u * j + k] = ga[k];
                    pg[14u * j + 7 + k] = gb2[k];
                }
            }
            tm.sync();
            for (uint32_t c = 0; c < ncls; ++c) {
                const uint32_t end = co[c + 1];
                for (uint32_t j = co[c] + r; j < end; j += NT) {
                    const uint32_t a = ab[2 * j], b = ab[2 * j + 1];
                    const float *ga = pg + 14u * j, *gb2 = pg + 14u * j + 7;
                    if (gv.inv_mass[a] != 0.f) {
                        av.pos_x[a] += ga[0]; av.pos_y[a] += ga[1]; av.pos_z[a] += ga[2];
                        av.rot_w[a] += ga[3]; av.rot_x[a] += ga[4]; av.rot_y[a] += ga[5]; av.rot_z[a] += ga[6];
                    }
                    if (gv.inv_mass[b] != 0.f) {
                        av.pos_x[b] += gb2[0]; av.pos_y[b] += gb2[1]; av.pos_z[b] += gb2[2];
                        av.rot_w[b] += gb2[3]; av.rot_x[b] += gb2[4]; av.rot_y[b] += gb2[5]; av.rot_z[b] += gb2[6];
                    }
                }
                tm.sync();
            }
        }
        // integrator substep
        const float *st0 = fb + TL.st;
        for (uint32_t i = r; i < n; i += NT) {
            if (gv.inv_mass[i] == 0.f || (gv.flags[i] & kFlagSleeping)) continue;
            BodyT<float> k0;
            for (int q = 0; q < 3; ++q) {
                k0.p[q] = st0[q * n + i]; k0.v[q] = st0[(7 + q) * n + i]; k0.w[q] = st0[(10 + q) * n + i];
            }
            for (int q = 0; q < 4; ++q) k0.q[q] = st0[(3 + q) * n + i];
            BodyBar<float> bb = load_bar(av, i);
            integrate_adj(k0, g, h, bb);
            store_bar(av, i, bb);
        }
        tm.sync();
    }
    if (!GLB)
        for (uint32_t i = r; i < n; i += NT) store_bar(ag, i, load_bar(av, i));
}

// Reverse kernels, one workgroup per scene of D.scenes.  diff_backward: a
// wavefront per small scene, its adjoint in LDS.  diff_backward_block: a
// 256-thread workgroup per larger scene whose adjoint (21 n floats) and
// first accumulators fit the launch's LDS.  diff_backward_global: the same
// for the scenes beyond that, their adjoint in global memory.  The launches
// cover disjoint scenes and run concurrently (separate kernels: one holding
// both 256-thread variants spills three times as much).
constexpr uint32_t kRevBlock = 256;
constexpr uint32_t kRevWaveMax = 128;  // bodies: larger scenes get a workgroup

__global__ __launch_bounds__(64) void diff_backward(BatchArgs A, DiffArgs D) {
    extern __shared__ float lds_adj[];
    const WaveTeam tm{threadIdx.x, nullptr};
    reverse_scene<WaveTeam, false>(tm, A, D, D.scenes[blockIdx.x], lds_adj);
}

__global__ __launch_bounds__(kRevBlock) void diff_backward_block(BatchArgs A, DiffArgs D) {
    extern __shared__ float lds_adj[];
    const BlockTeam<kRevBlock> tm{threadIdx.x, nullptr};
    reverse_scene<BlockTeam<kRevBlock>, false>(tm, A, D, D.scenes[blockIdx.x], lds_adj);
}

__global__ __launch_bounds__(kRevBlock) void diff_backward_global(BatchArgs A, DiffArgs D) {
    const BlockTeam<kRevBlock> tm{threadIdx.x, nullptr};
    reverse_scene<BlockTeam<kRevBlock>, true>(tm, A, D, D.scenes[blockIdx.x], nullptr);
}

}  // namespace dy

using namespace dy;

struct dy_batch_s {
    // Launch shape: workgroup threads, and the largest scene (bodies) stepped
    // by a single wavefront; larger scenes get a whole workgroup.
    struct Units {
        uint4 *d_units = nullptr;
        uint32_t *d_kind = nullptr;
        uint32_t n_units = 0, region_floats = 0;
        size_t lds_bytes = 0;
    };
    struct Shape {
        Units res;   // scenes whose state fits 64 KiB of LDS
        Units big;   // scenes whose state fits the larger LDS of one workgroup (one per workgroup)
        Units glob;  // the others: state in global memory
    };
    uint32_t n_scenes = 0, max_n = 0, block = 256, wave_max = 128;
    uint32_t lds_max_floats = kLdsMaxFloats;  // dynamic LDS one workgroup may take on this device
    int n_iters = 10;
    float mu = 0.5f;
    int force_serial = 0;
    int timing = 0;
    uint64_t *d_clk = nullptr;  // kStages per scene
    std::vector<uint32_t> body_off, pair_off, contact_off;
    std::map<std::pair<uint32_t, uint32_t>, Shape> shapes;
    uint32_t *d_body_off = nullptr, *d_pair_off = nullptr, *d_contact_off = nullptr;
    uint64_t *d_scratch_off = nullptr;
    uint32_t *d_scratch = nullptr, *d_stats = nullptr;
    dy_contact_pair *d_pairs = nullptr;
    void *d_contacts = nullptr;  // 9 arrays of contact_off.back() words
    dy_contact_view cv{};
    // The launch of the scenes kept in global memory runs beside the
    // LDS-resident one: a side stream forked from and joined back into the
    // caller's (a graph capture records the two as parallel nodes).
    hipStream_t side[2] = {nullptr, nullptr};  // default priority
    hipEvent_t ev_fork = nullptr, ev_join[2] = {nullptr, nullptr};
};

namespace {

void batch_free(dy_batch_s *b) {
    void *ps[] = {b->d_body_off, b->d_pair_off, b->d_contact_off, b->d_scratch_off,
                  b->d_scratch, b->d_stats, b->d_pairs, b->d_contacts, b->d_clk};
    for (void *p : ps)
        if (p) (void)hipFree(p);
    for (auto &kv : b->shapes)
        for (auto *u : {&kv.second.res, &kv.second.big, &kv.second.glob}) {
            if (u->d_units) (void)hipFree(u->d_units);
            if (u->d_kind) (void)hipFree(u->d_kind);
        }
    for (hipStream_t x : b->side)
        if (x) {
            (void)hipStreamSynchronize(x);
            (void)hipStreamDestroy(x);
        }
    for (hipEvent_t e : {b->ev_fork, b->ev_join[0], b->ev_join[1]})
        if (e) (void)hipEventDestroy(e);
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

int upload_units(dy_batch_s::Units &u, const std::vector<uint4> &units, const std::vector<uint32_t> &kind,
                 uint32_t region, uint32_t lds_floats) {
    u.n_units = static_cast<uint32_t>(units.size());
    u.region_floats = region;
    u.lds_bytes = static_cast<size_t>(lds_floats) * sizeof(float);
    if (units.empty()) return DY_OK;
    DY_TRY_STATUS(dev_alloc(u.d_units, units.size()));
    DY_TRY_STATUS(dev_alloc(u.d_kind, kind.size()));
    DY_TRY(hipMemcpy(u.d_units, units.data(), units.size() * sizeof(uint4), hipMemcpyHostToDevice));
    DY_TRY(hipMemcpy(u.d_kind, kind.data(), kind.size() * 4, hipMemcpyHostToDevice));
    return DY_OK;
}

// Work units of a launch shape.  Scenes of at most wave_max bodies are packed
// one per wavefront, threads/64 to a workgroup, in descending size (at 64
// threads every scene is a wavefront team); larger scenes get a workgroup
// each, largest first.  A wavefront team's LDS region holds its scene's box
// records and state; a workgroup team's scene keeps its state in LDS where
// 64 KiB allow.  Scenes beyond that go to a second launch, one per
// workgroup, whose LDS is sized for its largest scene up to what one
// workgroup may take (160 KB on gfx950), so that the first launch keeps its
// occupancy; scenes beyond that go to a third launch with their state in
// global memory.
int prepare_shape(dy_batch_s *b, uint32_t threads, uint32_t wave_max) {
    const auto key = std::make_pair(threads, wave_max);
    if (b->shapes.count(key)) return DY_OK;
    const uint32_t per = threads / 64;
    constexpr uint32_t kRes = kStateFloats + kBoxFloats;
    std::vector<uint32_t> order(b->n_scenes);
    std::iota(order.begin(), order.end(), 0u);
    auto nb = [&](uint32_t i) { return b->body_off[i + 1] - b->body_off[i]; };
    std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return nb(x) > nb(y); });
    const uint32_t wave_cap = kLdsMaxFloats / per / kRes;  // bodies a wavefront region can hold
    std::vector<uint32_t> big, small, large, huge;
    for (uint32_t i : order) {
        const bool wave = threads == 64 ? nb(i) <= wave_cap : nb(i) <= std::min(wave_max, wave_cap);
        if (wave) small.push_back(i);
        else if (kRes * nb(i) <= kLdsMaxFloats && threads > 64) big.push_back(i);
        else if (kRes * nb(i) <= b->lds_max_floats) large.push_back(i);
        else huge.push_back(i);
    }
    const uint32_t region = small.empty() ? 0u : kRes * nb(small.front());
    uint32_t lds = per * region;
    if (!big.empty()) lds = std::max(lds, kRes * nb(big.front()));
    std::vector<uint4> units;
    std::vector<uint32_t> kind;
    for (uint32_t i : big) {
        units.push_back(make_uint4(i, kNoScene, kNoScene, kNoScene));
        kind.push_back(kUnitBlock);
    }
    for (size_t j = 0; j < small.size(); j += per) {
        uint32_t sc[4] = {kNoScene, kNoScene, kNoScene, kNoScene};
        for (uint32_t w = 0; w < per && j + w < small.size(); ++w) sc[w] = small[j + w];
        units.push_back(make_uint4(sc[0], sc[1], sc[2], sc[3]));
        kind.push_back(0u);
    }
    auto singles = [&](const std::vector<uint32_t> &list, std::vector<uint4> &u, std::vector<uint32_t> &k) {
        for (uint32_t i : list) {
            u.push_back(make_uint4(i, kNoScene, kNoScene, kNoScene));
            k.push_back(threads > 64 ? kUnitBlock : 0u);
        }
    };
    std::vector<uint4> lunits, gunits;
    std::vector<uint32_t> lkind, gkind;
    singles(large, lunits, lkind);
    singles(huge, gunits, gkind);
    const uint32_t llds = large.empty() ? 0u : kRes * nb(large.front());
    const uint32_t glds = huge.empty() ? 0u : kBoxFloats * nb(huge.front());
    dy_batch_s::Shape sh;
    int st = upload_units(sh.res, units, kind, region, lds);
    if (st == DY_OK) st = upload_units(sh.big, lunits, lkind, llds, llds);
    if (st == DY_OK) st = upload_units(sh.glob, gunits, gkind, glds, glds);
    b->shapes[key] = sh;  // freed with the batch, also after a failure
    // one side stream per launch beyond the first, made here (set-up, never
    // inside a graph capture) and only as many as a shape needs: a process's
    // streams share HIP's few hardware queues
    const int n_launch = (units.empty() ? 0 : 1) + (lunits.empty() ? 0 : 1) + (gunits.empty() ? 0 : 1);
    for (int j = 0; j + 1 < n_launch && st == DY_OK; ++j)
        if (!b->side[j] && hipStreamCreateWithFlags(&b->side[j], hipStreamNonBlocking) != hipSuccess)
            st = DY_ERR_INTERNAL;
    return st;
}

BatchArgs batch_args(dy_batch_s *b, const dy_body_view &bodies, const dy_shape_view &shapes, float gx, float gy,
                     float gz, float dt, int n_substeps, int n_frames) {
    BatchArgs a{};
    a.bv = bodies;
    a.sv = shapes;
    a.cv = b->cv;
    a.pairs = b->d_pairs;
    a.body_off = b->d_body_off;
    a.pair_off = b->d_pair_off;
    a.contact_off = b->d_contact_off;
    a.scratch_off = b->d_scratch_off;
    a.scratch = b->d_scratch;
    a.stats = b->d_stats;
    a.stage_clk = b->timing ? b->d_clk : nullptr;
    a.gx = gx; a.gy = gy; a.gz = gz;
    a.h = dt / static_cast<float>(n_substeps);
    a.n_sub = n_substeps;
    a.n_frames = n_frames;
    a.n_iters = b->n_iters;
    a.force_serial = b->force_serial;
    a.mu = b->mu;
    return a;
}

// Every scene of the current launch shape runs LDS-resident (its state is
// read once and written back at the end): a recompute can then read a
// checkpoint in place.
bool batch_tape_ok(const dy_batch_s *b) {
    const auto it = b->shapes.find(std::make_pair(b->block, b->wave_max));
    return it != b->shapes.end() && it->second.glob.n_units == 0;
}

int batch_launch(dy_batch_s *b, hipStream_t s, const dy_body_view &bodies, const dy_shape_view &shapes, float gx,
                 float gy, float gz, float dt, int n_substeps, int n_frames, float *tape = nullptr,
                 const uint64_t *tape_off = nullptr, bool keep_input = false, const uint32_t *tape_cap = nullptr) {
    const auto it = b->shapes.find(std::make_pair(b->block, b->wave_max));
    if (it == b->shapes.end()) return DY_ERR_INTERNAL;
    const dy_batch_s::Shape &sh = it->second;
    BatchArgs a = batch_args(b, bodies, shapes, gx, gy, gz, dt, n_substeps, n_frames);
    a.tape = tape;
    a.tape_off = tape_off;
    a.tape_cap = tape_cap;
    a.keep_input = keep_input ? 1 : 0;
    if (tape && !tape_cap) return DY_ERR_INVALID_ARG;
    if (keep_input && sh.glob.n_units) return DY_ERR_INVALID_ARG;
    // Up to three launches (larger LDS-resident, global memory, LDS-resident;
    // the ones holding the largest scenes are enqueued first so that their
    // workgroups are placed first): the first non-empty one on the caller's
    // stream, the others on the side streams, forked from the caller's
    // stream before any of them and joined back into it.
    const dy_batch_s::Units *passes[3] = {&sh.big, &sh.glob, &sh.res};
    const bool residents[3] = {true, false, true};
    int n_nonempty = 0;
    for (const dy_batch_s::Units *u : passes) n_nonempty += u->n_units ? 1 : 0;
    if (n_nonempty > 1) DY_TRY(hipEventRecord(b->ev_fork, s));
    int n_launch = 0;
    for (int pass = 0; pass < 3; ++pass) {
        const dy_batch_s::Units &u = *passes[pass];
        if (u.n_units == 0) continue;
        hipStream_t ls = s;
        if (n_launch > 0) {
            ls = b->side[n_launch - 1];  // made by prepare_shape
            DY_TRY(hipStreamWaitEvent(ls, b->ev_fork, 0));
        }
        ++n_launch;
        a.units = u.d_units;
        a.unit_kind = u.d_kind;
        a.region_floats = u.region_floats;
        const bool res = residents[pass];
        int st = DY_ERR_INVALID_ARG;
        // the largest LDS-resident scenes (one per workgroup, the launch's
        // critical path) get 512 threads when the shape uses workgroup teams
        if (passes[pass] == &sh.big && b->block > 64) {
            DY_TRY_STATUS((launch_frames<kBigThreads, true>(a, u.n_units, u.lds_bytes, ls)));
            continue;
        }
        switch (b->block) {
            case 64: st = res ? launch_frames<64, true>(a, u.n_units, u.lds_bytes, ls)
                              : launch_frames<64, false>(a, u.n_units, u.lds_bytes, ls); break;
            case 128: st = res ? launch_frames<128, true>(a, u.n_units, u.lds_bytes, ls)
                               : launch_frames<128, false>(a, u.n_units, u.lds_bytes, ls); break;
            case 256: st = res ? launch_frames<256, true>(a, u.n_units, u.lds_bytes, ls)
                               : launch_frames<256, false>(a, u.n_units, u.lds_bytes, ls); break;
            default: break;
        }
        DY_TRY_STATUS(st);
    }
    for (int j = 0; j + 1 < n_launch; ++j) {
        DY_TRY(hipEventRecord(b->ev_join[j], b->side[j]));
        DY_TRY(hipStreamWaitEvent(s, b->ev_join[j], 0));
    }
    return DY_OK;
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
    DY_BATCH_TRY(dev_alloc(b->d_body_off, n_scenes + 1));
    DY_BATCH_TRY(dev_alloc(b->d_pair_off, n_scenes + 1));
    DY_BATCH_TRY(dev_alloc(b->d_contact_off, n_scenes + 1));
    DY_BATCH_TRY(dev_alloc(b->d_scratch_off, n_scenes + 1));
    DY_BATCH_TRY(dev_alloc(b->d_scratch, scratch_off[n_scenes]));
    DY_BATCH_TRY(dev_alloc(b->d_stats, static_cast<size_t>(n_scenes) * kStWords));
    DY_BATCH_TRY(dev_alloc(b->d_pairs, pt));
    DY_BATCH_TRY(dev_alloc(b->d_clk, static_cast<size_t>(n_scenes) * kStages));
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
    DY_BATCH_TRY(upload(hsm, b->d_body_off, b->body_off.data(), (n_scenes + 1) * 4));
    DY_BATCH_TRY(upload(hsm, b->d_pair_off, b->pair_off.data(), (n_scenes + 1) * 4));
    DY_BATCH_TRY(upload(hsm, b->d_contact_off, b->contact_off.data(), (n_scenes + 1) * 4));
    DY_BATCH_TRY(upload(hsm, b->d_scratch_off, scratch_off.data(), (n_scenes + 1) * 8));
    if (hipMemsetAsync(b->d_stats, 0, static_cast<size_t>(n_scenes) * kStWords * 4, hsm) != hipSuccess)
        return fail(DY_ERR_INTERNAL);
    if (hipStreamSynchronize(hsm) != hipSuccess) return fail(DY_ERR_INTERNAL);
    for (hipEvent_t *e : {&b->ev_fork, &b->ev_join[0], &b->ev_join[1]})
        if (hipEventCreateWithFlags(e, hipEventDisableTiming) != hipSuccess) return fail(DY_ERR_INTERNAL);
    {
        // dynamic LDS one workgroup may take, less 4 KiB for the kernels' static LDS
        int dev = 0, lds = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
            return fail(DY_ERR_INTERNAL);
        b->lds_max_floats = std::max<uint32_t>(kLdsMaxFloats, static_cast<uint32_t>(std::max(lds - 4096, 0)) / 4u);
    }
    DY_BATCH_TRY(prepare_shape(b, b->block, b->wave_max));
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

int dy_batch_set_shape(dy_batch_t b, uint32_t threads, uint32_t wave_scene_max) {
    if (!b || (threads != 64 && threads != 128 && threads != 256)) return DY_ERR_INVALID_ARG;
    if (threads == 64) wave_scene_max = 0;
    DY_TRY_STATUS(prepare_shape(b, threads, wave_scene_max));
    b->block = threads;
    b->wave_max = wave_scene_max;
    return DY_OK;
}

int dy_batch_set_lds_limit(dy_batch_t b, uint32_t bytes) {
    if (!b) return DY_ERR_INVALID_ARG;
    int dev = 0, lds = 0;
    DY_TRY(hipGetDevice(&dev));
    DY_TRY(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    const uint32_t dev_floats = std::max<uint32_t>(kLdsMaxFloats, static_cast<uint32_t>(std::max(lds - 4096, 0)) / 4u);
    const uint32_t want = bytes == 0 ? dev_floats : std::min(dev_floats, std::max(kLdsMaxFloats, bytes / 4u));
    if (want == b->lds_max_floats) return DY_OK;
    // the launch plans depend on the limit: drop them and plan the current shape again
    DY_TRY(hipDeviceSynchronize());
    for (auto &kv : b->shapes)
        for (auto *u : {&kv.second.res, &kv.second.big, &kv.second.glob}) {
            if (u->d_units) (void)hipFree(u->d_units);
            if (u->d_kind) (void)hipFree(u->d_kind);
        }
    b->shapes.clear();
    b->lds_max_floats = want;
    return prepare_shape(b, b->block, b->wave_max);
}

int dy_batch_get_shape(dy_batch_t b, uint32_t *threads, uint32_t *wave_scene_max) {
    if (!b || !threads || !wave_scene_max) return DY_ERR_INVALID_ARG;
    *threads = b->block;
    *wave_scene_max = b->wave_max;
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
                      uint32_t *chosen_threads, uint32_t *chosen_wave_max, float *ms_out) {
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
    const uint32_t cand[kTuneShapes][2] = {{64, 0}, {128, 64}, {128, 128}, {256, 64}, {256, 128}, {256, 256}};
    float best = 0.f;
    const uint32_t keep[2] = {b->block, b->wave_max};
    uint32_t best_shape[2] = {keep[0], keep[1]};
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) st = DY_ERR_INTERNAL;
    if (st == DY_OK) st = copy_state(true);
    for (uint32_t c = 0; c < kTuneShapes && st == DY_OK; ++c) {
        st = dy_batch_set_shape(b, cand[c][0], cand[c][1]);
        if (st == DY_OK) st = copy_state(false);
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
            best_shape[0] = cand[c][0];
            best_shape[1] = cand[c][1];
        }
    }
    if (st == DY_OK) st = dy_batch_set_shape(b, best_shape[0], best_shape[1]);
    else (void)dy_batch_set_shape(b, keep[0], keep[1]);
    if (st == DY_OK) st = copy_state(false);
    if (hipStreamSynchronize(hsm) != hipSuccess && st == DY_OK) st = DY_ERR_INTERNAL;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(save);
    (void)hipFree(save_stats);
    if (st == DY_OK && chosen_threads) *chosen_threads = b->block;
    if (st == DY_OK && chosen_wave_max) *chosen_wave_max = b->wave_max;
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

int dy_batch_set_timing(dy_batch_t b, dy_stream_t s, int enable) {
    if (!b) return DY_ERR_INVALID_ARG;
    b->timing = enable ? 1 : 0;
    if (enable)
        DY_TRY(hipMemsetAsync(b->d_clk, 0, static_cast<size_t>(b->n_scenes) * kStages * 8, hs(s)));
    return DY_OK;
}

int dy_batch_get_timing(dy_batch_t b, dy_stream_t s, uint64_t *ticks, double *tick_hz) {
    if (!b || !ticks || !tick_hz) return DY_ERR_INVALID_ARG;
    DY_TRY(hipMemcpyAsync(ticks, b->d_clk, static_cast<size_t>(b->n_scenes) * kStages * 8, hipMemcpyDeviceToHost,
                          hs(s)));
    DY_TRY(hipStreamSynchronize(hs(s)));
    int dev = 0, khz = 0;
    DY_TRY(hipGetDevice(&dev));
    DY_TRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
    *tick_hz = 1e3 * static_cast<double>(khz);
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

// ---- differentiable form -------------------------------------------------
struct dy_diff_s {
    dy_batch_s *b = nullptr;
    uint32_t max_frames = 0, every = 0, frames = 0, n_sub = 1;
    uint32_t tape_steps = 0;  // substeps one tape holds (whole frames)
    uint64_t step_bytes = 0;  // one recorded substep of every scene
    size_t n_total = 0;
    bool resident = true;     // every scene LDS-resident: recomputes read checkpoints in place
    bool cap_reduced = false; // some scene's tape holds fewer contacts than its store
    // Two tapes: the reverse sweep of one unit reads one while the next
    // (older) unit is recomputed into the other on the recompute stream.
    float *d_tape[2] = {nullptr, nullptr};
    // When every segment's record fits the memory of the two tapes, the
    // forward records them all (one tape per segment, tape_floats each) and
    // the backward only reverses: no recompute.
    bool record_all = false;
    float *d_all = nullptr;
    uint64_t tape_floats = 0;
    uint64_t *d_tape_off = nullptr, *d_cs_off = nullptr;
    uint32_t *d_tape_cap = nullptr;  // per scene: contacts a tape step holds
    uint32_t *d_err = nullptr;       // header errors of the reversed steps
    std::vector<uint32_t> tape_cap;
    float *d_ckpt = nullptr;  // checkpoints: 13 x n_total floats each
    float *d_sub = nullptr;   // sub-checkpoints inside a segment longer than a tape
    float *d_work = nullptr;  // recompute state when scenes run from global memory
    float *d_cs = nullptr;    // kCsStride per tape contact
    hipStream_t rs = nullptr;  // recompute stream (default priority)
    // Reverse launches (wavefront, workgroup in LDS, workgroup in global
    // memory): scenes of each (largest first, concatenated), LDS sizing, and
    // the streams beside the caller's that run the second and third.
    uint32_t *d_rev_scenes = nullptr;
    uint32_t n_rev[3] = {0, 0, 0}, acc_rev[3] = {0, 0, 0}, lds_rev[3] = {0, 0, 0};
    hipStream_t rev_st[2] = {nullptr, nullptr};  // default priority
    hipEvent_t ev_start = nullptr, ev_tape[2] = {nullptr, nullptr}, ev_free[2] = {nullptr, nullptr};
    hipEvent_t ev_fork = nullptr, ev_join[2] = {nullptr, nullptr};
    int newest = -1;  // tape holding the last forward's newest segment, or -1
};

namespace {

float *const *mutable_fields(const dy_body_view &v, float *out[13]) {
    float *f[13] = {v.pos_x, v.pos_y, v.pos_z, v.rot_w, v.rot_x, v.rot_y, v.rot_z,
                    v.vel_x, v.vel_y, v.vel_z, v.ang_x, v.ang_y, v.ang_z};
    for (int k = 0; k < 13; ++k) out[k] = f[k];
    return out;
}

void set_mutable_fields(dy_body_view &v, float *base, size_t n) {
    float **f[13] = {&v.pos_x, &v.pos_y, &v.pos_z, &v.rot_w, &v.rot_x, &v.rot_y, &v.rot_z,
                     &v.vel_x, &v.vel_y, &v.vel_z, &v.ang_x, &v.ang_y, &v.ang_z};
    for (int k = 0; k < 13; ++k) *f[k] = base + k * n;
}

int copy_state(hipStream_t s, float *const dst[13], float *const src[13], size_t n) {
    for (int k = 0; k < 13; ++k)
        DY_TRY(hipMemcpyAsync(dst[k], src[k], n * sizeof(float), hipMemcpyDeviceToDevice, s));
    return DY_OK;
}

// A 13-field state stored contiguously (checkpoint, sub-checkpoint, work).
dy_body_view state_view(const dy_body_view &bodies, float *base, size_t n) {
    dy_body_view v = bodies;
    set_mutable_fields(v, base, n);
    return v;
}

void diff_free(dy_diff_s *d) {
    if (d->rs) (void)hipStreamSynchronize(d->rs);
    for (hipStream_t x : d->rev_st)
        if (x) (void)hipStreamSynchronize(x);
    void *ps[] = {d->d_tape[0], d->d_tape[1], d->d_all, d->d_tape_off, d->d_cs_off, d->d_tape_cap, d->d_err,
                  d->d_ckpt, d->d_sub, d->d_work, d->d_cs, d->d_rev_scenes};
    for (void *p : ps)
        if (p) (void)hipFree(p);
    hipEvent_t evs[] = {d->ev_start, d->ev_tape[0], d->ev_tape[1], d->ev_free[0], d->ev_free[1],
                        d->ev_fork,  d->ev_join[0], d->ev_join[1]};
    for (hipEvent_t e : evs)
        if (e) (void)hipEventDestroy(e);
    if (d->rs) (void)hipStreamDestroy(d->rs);
    for (hipStream_t x : d->rev_st)
        if (x) (void)hipStreamDestroy(x);
}

}  // namespace

extern "C" {

int dy_diff_create(dy_batch_t b, uint32_t max_frames, uint32_t checkpoint_every, uint32_t n_substeps,
                   float tape_contacts_per_body, uint64_t tape_bytes, dy_diff_t *out) {
    if (!b || !out || max_frames == 0 || checkpoint_every == 0 || n_substeps == 0) return DY_ERR_INVALID_ARG;
    *out = nullptr;
    auto *d = new dy_diff_s();
    d->b = b;
    d->max_frames = max_frames;
    d->every = checkpoint_every;
    d->n_sub = n_substeps;
    d->n_total = b->body_off.back();
    d->resident = batch_tape_ok(b);
    // Per scene: the contacts one tape step holds (the contact store's
    // capacity, or fewer on request), and the floats of one recorded step.
    const uint32_t ns = b->n_scenes;
    d->tape_cap.resize(ns);
    std::vector<uint64_t> step_floats(ns), cs_off(ns + 1, 0ull);
    uint64_t step_total = 0;
    for (uint32_t i = 0; i < ns; ++i) {
        const uint32_t n = b->body_off[i + 1] - b->body_off[i];
        const uint32_t c = b->contact_off[i + 1] - b->contact_off[i];
        uint32_t ct = c;
        if (tape_contacts_per_body > 0.f) {
            const double want = std::ceil(static_cast<double>(tape_contacts_per_body) * n);
            ct = static_cast<uint32_t>(std::min<double>(c, std::max(1.0, want)));
        }
        d->cap_reduced = d->cap_reduced || ct < c;
        d->tape_cap[i] = ct;
        step_floats[i] = tape_layout(n, ct, b->n_iters).frame;
        step_total += step_floats[i];
        cs_off[i + 1] = cs_off[i] + ct;
    }
    // Substeps per tape: whole frames, at most a segment, within the budget
    // (24 GiB per tape unless given).
    const uint64_t budget = tape_bytes ? tape_bytes : (24ull << 30);
    const uint64_t seg_steps = static_cast<uint64_t>(checkpoint_every) * n_substeps;
    uint64_t fit = budget / (std::max<uint64_t>(step_total, 1ull) * sizeof(float));
    fit = fit / n_substeps * n_substeps;
    if (fit == 0) {
        delete d;
        return DY_ERR_INVALID_ARG;  // one frame's record exceeds the tape budget
    }
    d->tape_steps = static_cast<uint32_t>(std::min(fit, seg_steps));
    d->step_bytes = step_total * sizeof(float);
    std::vector<uint64_t> off(ns + 1, 0ull);
    for (uint32_t i = 0; i < ns; ++i) off[i + 1] = off[i] + d->tape_steps * step_floats[i];
    const uint32_t n_ckpt = (max_frames + checkpoint_every - 1) / checkpoint_every;
    const uint32_t tape_frames = d->tape_steps / n_substeps;
    const uint32_t n_subck = (checkpoint_every + tape_frames - 1) / tape_frames;
    d->tape_floats = off.back();
    d->record_all = d->tape_steps == seg_steps &&
                    static_cast<uint64_t>(n_ckpt) * d->tape_floats * sizeof(float) <= 2 * budget;
    int st = DY_OK;
    if (d->record_all) {
        if (st == DY_OK) st = dev_alloc(d->d_all, static_cast<size_t>(n_ckpt) * d->tape_floats);
    } else {
        if (st == DY_OK) st = dev_alloc(d->d_tape[0], d->tape_floats);
        if (st == DY_OK) st = dev_alloc(d->d_tape[1], d->tape_floats);
    }
    if (st == DY_OK) st = dev_alloc(d->d_tape_off, off.size());
    if (st == DY_OK) st = dev_alloc(d->d_cs_off, cs_off.size());
    if (st == DY_OK) st = dev_alloc(d->d_tape_cap, ns);
    if (st == DY_OK) st = dev_alloc(d->d_err, 1);
    if (st == DY_OK) st = dev_alloc(d->d_ckpt, static_cast<size_t>(n_ckpt) * 13 * d->n_total);
    if (st == DY_OK && n_subck > 1) st = dev_alloc(d->d_sub, static_cast<size_t>(n_subck) * 13 * d->n_total);
    if (st == DY_OK) st = dev_alloc(d->d_work, 13 * d->n_total);
    if (st == DY_OK) st = dev_alloc(d->d_cs, kCsStride * std::max<uint64_t>(cs_off.back(), 1ull));
    // Reverse launches: scenes of at most kRevWaveMax bodies one per
    // wavefront, the others one per workgroup, in LDS where their adjoint and
    // constants (21 floats per body) and up to 384 contacts' accumulators fit
    // what one workgroup may take (160 KB on gfx950: the workgroup fills its
    // CU at one wavefront per SIMD in any case), else in global memory; each
    // list largest first.
    constexpr uint32_t kAccLds = 384;
    const uint32_t kLdsFloats = b->lds_max_floats;
    std::vector<uint32_t> rev[3];
    uint32_t n_max[3] = {0, 0, 0}, c_max[3] = {0, 0, 0};
    auto nb = [&](uint32_t i) { return b->body_off[i + 1] - b->body_off[i]; };
    uint32_t acc_big = 0;  // the workgroup launch's LDS accumulators (one count for all its scenes)
    for (uint32_t i = 0; i < ns; ++i)
        if (nb(i) > kRevWaveMax) acc_big = std::max(acc_big, std::min(d->tape_cap[i], kAccLds));
    for (uint32_t i = 0; i < ns; ++i) {
        const int k = nb(i) <= kRevWaveMax ? 0 : (21ull * nb(i) + 8ull * acc_big <= kLdsFloats ? 1 : 2);
        rev[k].push_back(i);
        n_max[k] = std::max(n_max[k], nb(i));
        c_max[k] = std::max(c_max[k], d->tape_cap[i]);
    }
    std::vector<uint32_t> all;
    for (int k = 0; k < 3; ++k) {
        std::stable_sort(rev[k].begin(), rev[k].end(), [&](uint32_t x, uint32_t y) { return nb(x) > nb(y); });
        all.insert(all.end(), rev[k].begin(), rev[k].end());
        d->n_rev[k] = static_cast<uint32_t>(rev[k].size());
        // every scene of launch k fits 21 n_max[k] + 8 acc_rev[k] floats (<= kLdsFloats)
        d->acc_rev[k] = k == 0 ? std::min(c_max[0], kAccLds) : (k == 1 ? acc_big : 0u);
        d->lds_rev[k] = k == 2 ? 0u : 21u * n_max[k] + 8u * d->acc_rev[k];
    }
    if (st == DY_OK) st = dev_alloc(d->d_rev_scenes, all.size());
    if (st == DY_OK && hipMemcpy(d->d_rev_scenes, all.data(), all.size() * 4, hipMemcpyHostToDevice) != hipSuccess)
        st = DY_ERR_INTERNAL;
    if (st == DY_OK && hipStreamCreateWithFlags(&d->rs, hipStreamNonBlocking) != hipSuccess) st = DY_ERR_INTERNAL;
    // one side stream per reverse launch beyond the first (a process's
    // streams share HIP's few hardware queues, so idle ones are not made)
    int n_rev_launches = 0;
    for (uint32_t c : d->n_rev) n_rev_launches += c ? 1 : 0;
    for (int j = 0; j + 1 < n_rev_launches; ++j)
        if (st == DY_OK && hipStreamCreateWithFlags(&d->rev_st[j], hipStreamNonBlocking) != hipSuccess)
            st = DY_ERR_INTERNAL;
    hipEvent_t *evs[] = {&d->ev_start, &d->ev_tape[0], &d->ev_tape[1], &d->ev_free[0], &d->ev_free[1],
                         &d->ev_fork,  &d->ev_join[0], &d->ev_join[1]};
    for (hipEvent_t *e : evs)
        if (st == DY_OK && hipEventCreateWithFlags(e, hipEventDisableTiming) != hipSuccess) st = DY_ERR_INTERNAL;
    if (st == DY_OK && (hipMemcpy(d->d_tape_off, off.data(), off.size() * 8, hipMemcpyHostToDevice) != hipSuccess ||
                        hipMemcpy(d->d_cs_off, cs_off.data(), cs_off.size() * 8, hipMemcpyHostToDevice) != hipSuccess ||
                        hipMemcpy(d->d_tape_cap, d->tape_cap.data(), ns * 4, hipMemcpyHostToDevice) != hipSuccess))
        st = DY_ERR_INTERNAL;
    if (st != DY_OK) {
        diff_free(d);
        delete d;
        return st;
    }
    *out = d;
    return DY_OK;
}

int dy_diff_destroy(dy_diff_t d) {
    if (!d) return DY_OK;
    diff_free(d);
    delete d;
    return DY_OK;
}

int dy_diff_tape_steps(dy_diff_t d, uint32_t *steps, uint64_t *bytes_per_step) {
    if (!d || !steps) return DY_ERR_INVALID_ARG;
    *steps = d->tape_steps;
    if (bytes_per_step) *bytes_per_step = d->step_bytes;
    return DY_OK;
}

int dy_diff_forward(dy_diff_t d, dy_stream_t s, const dy_body_view *bodies, const dy_shape_view *shapes, float gx,
                    float gy, float gz, float dt, int n_frames) {
    if (!d) return DY_ERR_INVALID_ARG;
    DY_TRY_STATUS(check_step_args(d->b, bodies, shapes, static_cast<int>(d->n_sub), n_frames));
    if (n_frames <= 0 || static_cast<uint32_t>(n_frames) > d->max_frames) return DY_ERR_INVALID_ARG;
    const hipStream_t hsm = hs(s);
    float *src[13];
    mutable_fields(*bodies, src);
    d->newest = -1;
    const uint32_t n_seg = (static_cast<uint32_t>(n_frames) + d->every - 1) / d->every;
    const int sub = static_cast<int>(d->n_sub);
    for (uint32_t f0 = 0, k = 0; f0 < static_cast<uint32_t>(n_frames); f0 += d->every, ++k) {
        const int nf = static_cast<int>(std::min<uint32_t>(d->every, static_cast<uint32_t>(n_frames) - f0));
        if (d->record_all) {  // every segment recorded as it runs; no checkpoints needed
            DY_TRY_STATUS(batch_launch(d->b, hsm, *bodies, *shapes, gx, gy, gz, dt, sub, nf,
                                       d->d_all + static_cast<size_t>(k) * d->tape_floats, d->d_tape_off, false,
                                       d->d_tape_cap));
            continue;
        }
        dy_body_view ck{};
        set_mutable_fields(ck, d->d_ckpt + static_cast<size_t>(k) * 13 * d->n_total, d->n_total);
        float *dst[13];
        mutable_fields(ck, dst);
        DY_TRY_STATUS(copy_state(hsm, dst, src, d->n_total));
        // The newest segment is recorded as it runs when one tape holds it:
        // the backward then starts from it without recomputing it.
        const bool newest = k + 1 == n_seg && static_cast<uint32_t>(nf) * d->n_sub <= d->tape_steps;
        const uint32_t t = 0u;  // the tape the backward's first unit reads
        DY_TRY_STATUS(batch_launch(d->b, hsm, *bodies, *shapes, gx, gy, gz, dt, sub, nf,
                                   newest ? d->d_tape[t] : nullptr, newest ? d->d_tape_off : nullptr, false,
                                   newest ? d->d_tape_cap : nullptr));
        if (newest) d->newest = static_cast<int>(t);
    }
    d->frames = static_cast<uint32_t>(n_frames);
    return DY_OK;
}

int dy_diff_backward(dy_diff_t d, dy_stream_t s, const dy_body_view *bodies, const dy_shape_view *shapes, float gx,
                     float gy, float gz, float dt, const dy_body_view *adj_final, dy_body_view *adj_initial) {
    if (!d || !adj_final || !adj_initial || d->frames == 0) return DY_ERR_INVALID_ARG;
    DY_TRY_STATUS(check_step_args(d->b, bodies, shapes, static_cast<int>(d->n_sub), 1));
    if (adj_final->n != d->n_total || adj_initial->n != d->n_total) return DY_ERR_INVALID_ARG;
    const hipStream_t hsm = hs(s);
    const size_t n = d->n_total;
    const int sub = static_cast<int>(d->n_sub);
    float *fin[13], *ini[13];
    mutable_fields(*adj_final, fin);
    mutable_fields(*adj_initial, ini);
    if (fin[0] != ini[0]) DY_TRY_STATUS(copy_state(hsm, ini, fin, n));
    DY_TRY(hipMemsetAsync(adj_initial->inv_mass, 0, n * sizeof(float), hsm));
    DY_TRY(hipMemsetAsync(d->d_err, 0, sizeof(uint32_t), hsm));
    DiffArgs D{};
    D.adj = *adj_initial;
    D.cs = d->d_cs;
    D.cs_off = d->d_cs_off;
    D.err = d->d_err;
    DiffArgs Dk[3] = {D, D, D};
    for (int k = 0, o = 0; k < 3; o += static_cast<int>(d->n_rev[k]), ++k) {
        Dk[k].scenes = d->d_rev_scenes + o;
        Dk[k].acc_lds = d->acc_rev[k];
        Dk[k].lds_floats = d->lds_rev[k];
    }
    const uint32_t tape_frames = d->tape_steps / d->n_sub;

    // All segments recorded by the forward: reverse each from its own tape,
    // newest first.  Otherwise units of at most one tape, newest first:
    // segment k (checkpoint every `every` frames) in chunks of tape_frames
    // frames.  A unit's reverse
    // sweep (stream s) reads tape u % 2 while the next unit is recomputed
    // into the other tape on the recompute stream; events order the
    // hand-overs.  A segment longer than a tape is first replayed from its
    // checkpoint, saving a sub-checkpoint per chunk.  The reverse sweep reads
    // only the constant fields (inverse mass and inertia, shape, flags) of
    // the body view.
    struct Unit {
        uint32_t seg, f0, nf;  // frames [f0, f0 + nf) of segment seg (absolute frames)
        const float *src;      // state at f0: a checkpoint or a sub-checkpoint
    };
    DY_TRY(hipEventRecord(d->ev_start, hsm));
    DY_TRY(hipStreamWaitEvent(d->rs, d->ev_start, 0));
    const uint32_t n_seg = (d->frames + d->every - 1) / d->every;
    uint32_t u_index = 0;  // units enqueued so far (tape = u_index % 2)
    auto seg_units = [&](uint32_t k, std::vector<Unit> &out) -> int {
        const uint32_t f0 = k * d->every, nf = std::min(d->every, d->frames - f0);
        float *ck = d->d_ckpt + static_cast<size_t>(k) * 13 * n;
        out.clear();
        if (nf <= tape_frames) {
            out.push_back(Unit{k, f0, nf, ck});
            return DY_OK;
        }
        // replay the segment on the recompute stream, saving each chunk's start
        DY_TRY(hipMemcpyAsync(d->d_work, ck, 13 * n * sizeof(float), hipMemcpyDeviceToDevice, d->rs));
        const dy_body_view wv = state_view(*bodies, d->d_work, n);
        uint32_t c = 0;
        for (uint32_t g = 0; g < nf; g += tape_frames, ++c) {
            float *sc = d->d_sub + static_cast<size_t>(c) * 13 * n;
            DY_TRY(hipMemcpyAsync(sc, d->d_work, 13 * n * sizeof(float), hipMemcpyDeviceToDevice, d->rs));
            const uint32_t cf = std::min(tape_frames, nf - g);
            out.push_back(Unit{k, f0 + g, cf, sc});
            if (g + cf < nf)
                DY_TRY_STATUS(batch_launch(d->b, d->rs, wv, *shapes, gx, gy, gz, dt, sub, static_cast<int>(cf)));
        }
        std::reverse(out.begin(), out.end());
        return DY_OK;
    };
    auto recompute = [&](const Unit &u, int t) -> int {
        DY_TRY(hipStreamWaitEvent(d->rs, d->ev_free[t], 0));  // the reverse sweep two units back is done with it
        if (d->resident) {
            // straight from the saved state, which the recompute reads but does not overwrite
            const dy_body_view sv = state_view(*bodies, const_cast<float *>(u.src), n);
            DY_TRY_STATUS(batch_launch(d->b, d->rs, sv, *shapes, gx, gy, gz, dt, sub, static_cast<int>(u.nf),
                                       d->d_tape[t], d->d_tape_off, true, d->d_tape_cap));
        } else {
            DY_TRY(hipMemcpyAsync(d->d_work, u.src, 13 * n * sizeof(float), hipMemcpyDeviceToDevice, d->rs));
            const dy_body_view wv = state_view(*bodies, d->d_work, n);
            DY_TRY_STATUS(batch_launch(d->b, d->rs, wv, *shapes, gx, gy, gz, dt, sub, static_cast<int>(u.nf),
                                       d->d_tape[t], d->d_tape_off, false, d->d_tape_cap));
        }
        DY_TRY(hipEventRecord(d->ev_tape[t], d->rs));
        DY_TRY(hipStreamWaitEvent(hsm, d->ev_tape[t], 0));
        return DY_OK;
    };
    auto reverse = [&](const Unit &u, float *tape, int t) -> int {  // t: the tape to release, or -1
        BatchArgs a = batch_args(d->b, *bodies, *shapes, gx, gy, gz, dt, sub, static_cast<int>(u.nf));
        a.tape = tape;
        a.tape_off = d->d_tape_off;
        a.tape_cap = d->d_tape_cap;
        // the non-empty launches: the first on the caller's stream, the
        // others on the side streams, forked from and joined back into it
        // (the largest scenes' launches first, so that their workgroups are
        // placed first; the fork is recorded before any of them)
        const void *kern[3] = {reinterpret_cast<const void *>(&diff_backward),
                               reinterpret_cast<const void *>(&diff_backward_block),
                               reinterpret_cast<const void *>(&diff_backward_global)};
        const uint32_t threads[3] = {64, kRevBlock, kRevBlock};
        int n_nonempty = 0;
        for (uint32_t c : d->n_rev) n_nonempty += c ? 1 : 0;
        if (n_nonempty > 1) DY_TRY(hipEventRecord(d->ev_fork, hsm));
        int n_launch = 0;
        for (int k : {2, 1, 0}) {
            if (d->n_rev[k] == 0) continue;
            Dk[k].frames = u.nf * d->n_sub;
            hipStream_t st = hsm;
            if (n_launch > 0) {
                st = d->rev_st[n_launch - 1];
                DY_TRY(hipStreamWaitEvent(st, d->ev_fork, 0));
            }
            void *args[] = {&a, &Dk[k]};
            DY_TRY_STATUS(allow_lds(kern[k], static_cast<size_t>(d->lds_rev[k]) * sizeof(float)));
            DY_TRY(hipLaunchKernel(kern[k], dim3(d->n_rev[k]), dim3(threads[k]), args,
                                   static_cast<size_t>(d->lds_rev[k]) * sizeof(float), st));
            ++n_launch;
        }
        for (int j = 0; j + 1 < n_launch; ++j) {
            DY_TRY(hipEventRecord(d->ev_join[j], d->rev_st[j]));
            DY_TRY(hipStreamWaitEvent(hsm, d->ev_join[j], 0));
        }
        if (t >= 0) DY_TRY(hipEventRecord(d->ev_free[t], hsm));
        return DY_OK;
    };
    std::vector<Unit> units;
    bool first = true;
    for (int k = static_cast<int>(n_seg) - 1; k >= 0 && d->record_all; --k) {
        const uint32_t f0 = static_cast<uint32_t>(k) * d->every;
        const Unit u{static_cast<uint32_t>(k), f0, std::min(d->every, d->frames - f0), nullptr};
        DY_TRY_STATUS(reverse(u, d->d_all + static_cast<size_t>(k) * d->tape_floats, -1));
    }
    for (int k = static_cast<int>(n_seg) - 1; k >= 0 && !d->record_all; --k) {
        DY_TRY_STATUS(seg_units(static_cast<uint32_t>(k), units));
        for (const Unit &u : units) {
            const int t = static_cast<int>(u_index % 2u);
            // the forward taped the newest segment when one tape held it
            const bool taped = first && d->newest == t && units.size() == 1;
            if (!taped) DY_TRY_STATUS(recompute(u, t));
            DY_TRY_STATUS(reverse(u, d->d_tape[t], t));
            first = false;
            ++u_index;
        }
    }
    d->newest = -1;  // the tapes now hold older units
    if (d->cap_reduced) {
        // a step with more contacts than a reduced tape holds cannot be reversed
        uint32_t err = 0;
        DY_TRY(hipMemcpyAsync(&err, d->d_err, sizeof(err), hipMemcpyDeviceToHost, hsm));
        DY_TRY(hipStreamSynchronize(hsm));
        if (err & kBatchErrTape) return DY_ERR_INVALID_ARG;
        if (err) return DY_ERR_INTERNAL;
    }
    return DY_OK;
}

}  // extern "C"
