// ref_cpu — CPU restatement of the dyphur frame step.  TEST INFRASTRUCTURE:
// the parity oracle for the gfx950 kernels and the host-CPU baseline.  The
// product never links or loads it.  See ref_cpu.h for the reference files
// followed and the semantic choices at the SYCL boundary.
//
// Build: -O2 -ffp-contract=off -fopenmp (no -march, no fast-math), so every
// float expression below rounds after each operation, exactly as written.

#include "ref_cpu.h"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

// ----------------------------------------------------------------------------
// Numeric helpers
// ----------------------------------------------------------------------------

inline uint32_t fbits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline float    bitsf(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

// sycl::rsqrt on the reference's OpenMP backend: 1 / sqrt(x).
inline float rsqrt_ref(float x) { return 1.f / std::sqrt(x); }

// sycl::clz on a 64-bit operand; clz(0) = 64.
inline int clz64(uint64_t v) { return v ? __builtin_clzll(v) : 64; }

inline uint32_t next_pow2(uint32_t n) {
    if (n <= 1u) return 1u;
    uint32_t p = 1u;
    while (p < n) p <<= 1u;
    return p;
}

// Priority word used by the graph colouring (xpbd_solver.cpp:657).
inline uint64_t jp_priority(uint32_t k) {
    return (static_cast<uint64_t>(rc_hash_u32(k)) << 32) | k;
}

// Device atomics (relaxed) as the reference uses them.
inline uint32_t atomic_inc_u32(uint32_t *p) { return __atomic_fetch_add(p, 1u, __ATOMIC_SEQ_CST); }
inline void atomic_max_u64(uint64_t *p, uint64_t v) {
    uint64_t cur = __atomic_load_n(p, __ATOMIC_RELAXED);
    while (cur < v &&
           !__atomic_compare_exchange_n(p, &cur, v, true, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
    }
}
inline void atomic_max_u32(uint32_t *p, uint32_t v) {
    uint32_t cur = __atomic_load_n(p, __ATOMIC_RELAXED);
    while (cur < v &&
           !__atomic_compare_exchange_n(p, &cur, v, true, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
    }
}

constexpr uint32_t kFlagStatic = 1u << 0;    // body.hpp:9
constexpr uint32_t kFlagSleeping = 1u << 2;  // body.hpp:11
constexpr uint32_t kShapeBox = 0u, kShapeSphere = 1u;

// ----------------------------------------------------------------------------
// Multi-pass bitonic sort (compute/include/compute/sort.hpp:68-85): one
// parallel pass per (k, j) stage, as the reference's CPU path launches them.
// ----------------------------------------------------------------------------
template <typename K, typename V>
void bitonic_kv(K *keys, V *vals, size_t n) {
    if (n <= 1) return;
    for (size_t k = 2; k <= n; k <<= 1) {
        for (size_t j = k >> 1; j > 0; j >>= 1) {
#pragma omp parallel for schedule(static)
            for (long long ii = 0; ii < static_cast<long long>(n); ++ii) {
                size_t i = static_cast<size_t>(ii);
                size_t l = i ^ j;
                if (l <= i) continue;
                bool up = (i & k) == 0;
                bool swap = up ? (keys[i] > keys[l]) : (keys[i] < keys[l]);
                if (swap) {
                    std::swap(keys[i], keys[l]);
                    if (vals) std::swap(vals[i], vals[l]);
                }
            }
        }
    }
}

// ----------------------------------------------------------------------------
// Karras 2012 radix tree over distinct 64-bit keys (broadphase.cpp:35-76).
// ----------------------------------------------------------------------------
inline int lcp(const uint64_t *codes, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    return clz64(codes[i] ^ codes[j]);
}

inline void karras_node(const uint64_t *codes, int n, int i, int &lo, int &hi, int &gamma) {
    // Direction of the range: towards the neighbour sharing the longer prefix.
    int dir = (lcp(codes, n, i, i + 1) - lcp(codes, n, i, i - 1)) >= 0 ? 1 : -1;
    int dmin = lcp(codes, n, i, i - dir);
    // Exponential search for an upper bound of the range length, then binary search.
    int lmax = 2;
    while (lcp(codes, n, i, i + lmax * dir) > dmin) lmax *= 2;
    int len = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
        if (lcp(codes, n, i, i + (len + t) * dir) > dmin) len += t;
    int other = i + len * dir;
    lo = dir > 0 ? i : other;
    hi = dir > 0 ? other : i;
    // Split: last index in [lo, hi) whose prefix with lo exceeds the node prefix.
    int dnode = lcp(codes, n, lo, hi);
    int s = 0, step = hi - lo;
    do {
        step = (step + 1) / 2;
        if (lcp(codes, n, lo, lo + s + step) > dnode) s += step;
    } while (step > 1);
    gamma = lo + s;
}

// Orientation matrix rows as the broadphase leaf pass builds them
// (broadphase.cpp:229-237): row r holds R(q)[r][0..2].
struct Rot3 { float m[3][3]; };
inline Rot3 leaf_rotation(float qw, float qx, float qy, float qz) {
    Rot3 r;
    r.m[0][0] = 1.f - 2.f * (qy * qy + qz * qz);
    r.m[0][1] = 2.f * (qx * qy - qz * qw);
    r.m[0][2] = 2.f * (qx * qz + qy * qw);
    r.m[1][0] = 2.f * (qx * qy + qz * qw);
    r.m[1][1] = 1.f - 2.f * (qx * qx + qz * qz);
    r.m[1][2] = 2.f * (qy * qz - qx * qw);
    r.m[2][0] = 2.f * (qx * qz - qy * qw);
    r.m[2][1] = 2.f * (qy * qz + qx * qw);
    r.m[2][2] = 1.f - 2.f * (qx * qx + qy * qy);
    return r;
}

// Box axes for the narrowphase (narrowphase.cpp:913-917): axis row a is the
// world direction of local axis a, i.e. column a of R(q).
inline void box_axes(float qw, float qx, float qy, float qz, float E[3][3]) {
    E[0][0] = 1.f - 2.f * (qy * qy + qz * qz);
    E[0][1] = 2.f * (qx * qy + qz * qw);
    E[0][2] = 2.f * (qx * qz - qy * qw);
    E[1][0] = 2.f * (qx * qy - qz * qw);
    E[1][1] = 1.f - 2.f * (qx * qx + qz * qz);
    E[1][2] = 2.f * (qy * qz + qx * qw);
    E[2][0] = 2.f * (qx * qz + qy * qw);
    E[2][1] = 2.f * (qy * qz - qx * qw);
    E[2][2] = 1.f - 2.f * (qx * qx + qy * qy);
}

// ----------------------------------------------------------------------------
// Narrowphase primitives (narrowphase.cpp:46-306)
// ----------------------------------------------------------------------------
// feat: the contact's feature word (kind | axes | incident-vertex mask | SAT
// axis), defined in ref_cpu.h; the reference computes these quantities in
// box_box and keeps only the contact (narrowphase.cpp:153-305).
struct Hit { float px, py, pz, nx, ny, nz, depth; uint32_t feat; };

constexpr uint32_t kFeatFaceA = 0, kFeatFaceB = 1, kFeatEdge = 2, kFeatSphSph = 3, kFeatSphBoxOut = 4,
                   kFeatSphBoxIn = 5;
inline uint32_t feat_word(uint32_t kind, uint32_t i, uint32_t j, uint32_t mask, uint32_t sat) {
    return kind | (i << 3) | (j << 5) | (mask << 8) | (sat << 16);
}

bool sphere_vs_sphere(float ax, float ay, float az, float ra,
                      float bx, float by, float bz, float rb, Hit &h) {
    float dx = ax - bx, dy = ay - by, dz = az - bz;
    float d2 = dx * dx + dy * dy + dz * dz;
    float rsum = ra + rb;
    if (d2 >= rsum * rsum) return false;
    float d = std::sqrt(d2);
    if (d < 1e-8f) { h.nx = 0.f; h.ny = 1.f; h.nz = 0.f; }
    else { float inv = 1.f / d; h.nx = dx * inv; h.ny = dy * inv; h.nz = dz * inv; }
    h.depth = rsum - d;
    // Midpoint of the two surface points, per axis.
    h.px = 0.5f * ((ax - h.nx * ra) + (bx + h.nx * rb));
    h.py = 0.5f * ((ay - h.ny * ra) + (by + h.ny * rb));
    h.pz = 0.5f * ((az - h.nz * ra) + (bz + h.nz * rb));
    h.feat = feat_word(kFeatSphSph, 0, 0, 0, 0);
    return true;
}

// Sphere (centre s, radius r) against an oriented box (centre c, axes E rows,
// half extents he).  Normal points from the sphere into the box side (the
// caller flips it when the sphere is body A).
bool sphere_vs_box(float sx, float sy, float sz, float r,
                   float cx, float cy, float cz, const float E[3][3], const float he[3], Hit &h) {
    float dx = sx - cx, dy = sy - cy, dz = sz - cz;
    float q[3], cl[3];
    for (int a = 0; a < 3; ++a) q[a] = dx * E[a][0] + dy * E[a][1] + dz * E[a][2];
    for (int a = 0; a < 3; ++a) cl[a] = q[a] < -he[a] ? -he[a] : (q[a] > he[a] ? he[a] : q[a]);
    float lx = q[0] - cl[0], ly = q[1] - cl[1], lz = q[2] - cl[2];
    float d2 = lx * lx + ly * ly + lz * lz;
    if (d2 < 1e-12f) {
        // Centre inside: push out through the face of least penetration.
        float pen[3];
        for (int a = 0; a < 3; ++a) pen[a] = he[a] - std::fabs(q[a]);
        int f = (pen[0] <= pen[1] && pen[0] <= pen[2]) ? 0 : (pen[1] <= pen[2] ? 1 : 2);
        float sg = q[f] >= 0.f ? 1.f : -1.f;
        h.nx = -(sg * E[f][0]); h.ny = -(sg * E[f][1]); h.nz = -(sg * E[f][2]);
        h.depth = r + pen[f];
        h.px = sx - h.nx * r; h.py = sy - h.ny * r; h.pz = sz - h.nz * r;
        h.feat = feat_word(kFeatSphBoxIn, static_cast<uint32_t>(f), 0, 0, 0);
        return true;
    }
    if (d2 >= r * r) return false;
    float d = std::sqrt(d2);
    float inv = 1.f / d;
    float wx = lx * E[0][0] + ly * E[1][0] + lz * E[2][0];
    float wy = lx * E[0][1] + ly * E[1][1] + lz * E[2][1];
    float wz = lx * E[0][2] + ly * E[1][2] + lz * E[2][2];
    h.nx = -wx * inv; h.ny = -wy * inv; h.nz = -wz * inv;
    h.depth = r - d;
    h.px = cx + cl[0] * E[0][0] + cl[1] * E[1][0] + cl[2] * E[2][0];
    h.py = cy + cl[0] * E[0][1] + cl[1] * E[1][1] + cl[2] * E[2][1];
    h.pz = cz + cl[0] * E[0][2] + cl[1] * E[1][2] + cl[2] * E[2][2];
    h.feat = feat_word(kFeatSphBoxOut, 0, 0, 0, 0);
    return true;
}

// Projected radius of a box (axes E, half extents he) on direction L.
inline float box_radius(const float E[3][3], const float he[3], float lx, float ly, float lz) {
    return he[0] * std::fabs(lx * E[0][0] + ly * E[0][1] + lz * E[0][2])
         + he[1] * std::fabs(lx * E[1][0] + ly * E[1][1] + lz * E[1][2])
         + he[2] * std::fabs(lx * E[2][0] + ly * E[2][1] + lz * E[2][2]);
}

// Box-box SAT + manifold.  Returns the number of manifold points (0..4); the
// normal points from B towards A.  feat receives the feature word (SAT axis,
// reference face or edge pair, incident vertices that passed the clip).
int box_vs_box(float acx, float acy, float acz, const float A[3][3], const float ah[3],
               float bcx, float bcy, float bcz, const float B[3][3], const float bh[3],
               float opx[4], float opy[4], float opz[4], float &onx, float &ony, float &onz,
               float od[4], uint32_t &feat) {
    const float kPar = 1e-5f;
    float C[3][3];
    float dx = bcx - acx, dy = bcy - acy, dz = bcz - acz;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            C[i][j] = A[i][0] * B[j][0] + A[i][1] * B[j][1] + A[i][2] * B[j][2];

    float best = 1e30f;
    int btype = -1, bi = 0, bj = 0;
    float bnx = 0.f, bny = 1.f, bnz = 0.f;
    // One separating-axis test; returns false when the axis separates.
    auto axis = [&](float lx, float ly, float lz, int type, int i, int j) -> bool {
        float rA = box_radius(A, ah, lx, ly, lz);
        float rB = box_radius(B, bh, lx, ly, lz);
        float t = std::fabs(dx * lx + dy * ly + dz * lz);
        float ov = rA + rB - t;
        if (ov <= 0.f) return false;
        if (ov < best) {
            best = ov; btype = type; bi = i; bj = j;
            float sg = (dx * lx + dy * ly + dz * lz) >= 0.f ? -1.f : 1.f;
            bnx = sg * lx; bny = sg * ly; bnz = sg * lz;
        }
        return true;
    };
    for (int i = 0; i < 3; ++i) if (!axis(A[i][0], A[i][1], A[i][2], 0, i, 0)) return 0;
    for (int j = 0; j < 3; ++j) if (!axis(B[j][0], B[j][1], B[j][2], 1, j, 0)) return 0;
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) {
            float cx = A[i][1] * B[j][2] - A[i][2] * B[j][1];
            float cy = A[i][2] * B[j][0] - A[i][0] * B[j][2];
            float cz = A[i][0] * B[j][1] - A[i][1] * B[j][0];
            float l2 = cx * cx + cy * cy + cz * cz;
            if (l2 < kPar * kPar) continue;
            float il = rsqrt_ref(l2);
            if (!axis(cx * il, cy * il, cz * il, 2, i, j)) return 0;
        }
    }
    onx = bnx; ony = bny; onz = bnz;
    int n_out = 0;
    const uint32_t sat_axis = btype == 0 ? bi : (btype == 1 ? 3 + bi : 6 + 3 * bi + bj);
    uint32_t mask = 0;

    if (btype == 0 || btype == 1) {
        if (btype == 0) {
            // Prefer B's face as reference when its most-aligned face is larger.
            int ta0 = (bi + 1) % 3, ta1 = (bi + 2) % 3;
            int jb = 0;
            for (int j = 1; j < 3; ++j)
                if (std::fabs(C[bi][j]) > std::fabs(C[bi][jb])) jb = j;
            int tb0 = (jb + 1) % 3, tb1 = (jb + 2) % 3;
            if (bh[tb0] * bh[tb1] > ah[ta0] * ah[ta1]) { btype = 1; bi = jb; }
        }
        const bool a_ref = (btype == 0);
        const int rf = bi;
        const float (*R)[3] = a_ref ? A : B;
        const float *rh = a_ref ? ah : bh;
        const float (*I)[3] = a_ref ? B : A;
        const float *ih = a_ref ? bh : ah;
        const float rcx = a_ref ? acx : bcx, rcy = a_ref ? acy : bcy, rcz = a_ref ? acz : bcz;
        const float icx = a_ref ? bcx : acx, icy = a_ref ? bcy : acy, icz = a_ref ? bcz : acz;
        const int t0 = (rf + 1) % 3, t1 = (rf + 2) % 3;
        const float ref_n = rcx * bnx + rcy * bny + rcz * bnz;
        const float sd = a_ref ? 1.f : -1.f;
        for (int k = 0; k < 8; ++k) {
            float s0 = (k & 1) ? 1.f : -1.f, s1 = (k & 2) ? 1.f : -1.f, s2 = (k & 4) ? 1.f : -1.f;
            float vx = icx + s0 * ih[0] * I[0][0] + s1 * ih[1] * I[1][0] + s2 * ih[2] * I[2][0];
            float vy = icy + s0 * ih[0] * I[0][1] + s1 * ih[1] * I[1][1] + s2 * ih[2] * I[2][1];
            float vz = icz + s0 * ih[0] * I[0][2] + s1 * ih[1] * I[1][2] + s2 * ih[2] * I[2][2];
            float vn = vx * bnx + vy * bny + vz * bnz;
            float dv = rh[rf] + sd * (vn - ref_n);
            if (dv <= 0.f) continue;
            float ux = vx - rcx, uy = vy - rcy, uz = vz - rcz;
            float q0 = ux * R[t0][0] + uy * R[t0][1] + uz * R[t0][2];
            float q1 = ux * R[t1][0] + uy * R[t1][1] + uz * R[t1][2];
            if (std::fabs(q0) > rh[t0] + 1e-3f) continue;
            if (std::fabs(q1) > rh[t1] + 1e-3f) continue;
            mask |= 1u << k;
            if (n_out < 4) {
                opx[n_out] = vx - sd * dv * bnx;
                opy[n_out] = vy - sd * dv * bny;
                opz[n_out] = vz - sd * dv * bnz;
                od[n_out] = dv;
                ++n_out;
            }
        }
        feat = feat_word(a_ref ? kFeatFaceA : kFeatFaceB, static_cast<uint32_t>(rf), 0, mask, sat_axis);
    } else {
        // Edge-edge: midpoint of the closest points of the two supporting edges.
        const int i = bi, j = bj;
        const int ta0 = (i + 1) % 3, ta1 = (i + 2) % 3;
        float sa0 = (-bnx * A[ta0][0] - bny * A[ta0][1] - bnz * A[ta0][2]) >= 0.f ? 1.f : -1.f;
        float sa1 = (-bnx * A[ta1][0] - bny * A[ta1][1] - bnz * A[ta1][2]) >= 0.f ? 1.f : -1.f;
        float pax = acx + sa0 * ah[ta0] * A[ta0][0] + sa1 * ah[ta1] * A[ta1][0];
        float pay = acy + sa0 * ah[ta0] * A[ta0][1] + sa1 * ah[ta1] * A[ta1][1];
        float paz = acz + sa0 * ah[ta0] * A[ta0][2] + sa1 * ah[ta1] * A[ta1][2];
        const int tb0 = (j + 1) % 3, tb1 = (j + 2) % 3;
        float sb0 = (bnx * B[tb0][0] + bny * B[tb0][1] + bnz * B[tb0][2]) >= 0.f ? 1.f : -1.f;
        float sb1 = (bnx * B[tb1][0] + bny * B[tb1][1] + bnz * B[tb1][2]) >= 0.f ? 1.f : -1.f;
        float pbx = bcx + sb0 * bh[tb0] * B[tb0][0] + sb1 * bh[tb1] * B[tb1][0];
        float pby = bcy + sb0 * bh[tb0] * B[tb0][1] + sb1 * bh[tb1] * B[tb1][1];
        float pbz = bcz + sb0 * bh[tb0] * B[tb0][2] + sb1 * bh[tb1] * B[tb1][2];
        float c = C[i][j];
        float den = 1.f - c * c;
        float wx = pax - pbx, wy = pay - pby, wz = paz - pbz;
        float d1 = wx * A[i][0] + wy * A[i][1] + wz * A[i][2];
        float d2 = wx * B[j][0] + wy * B[j][1] + wz * B[j][2];
        float s = 0.f, t = 0.f;
        if (std::fabs(den) > 1e-6f) {
            s = (c * d2 - d1) / den;
            t = (d2 - c * d1) / den;
        }
        float qax = pax + s * A[i][0], qay = pay + s * A[i][1], qaz = paz + s * A[i][2];
        float qbx = pbx + t * B[j][0], qby = pby + t * B[j][1], qbz = pbz + t * B[j][2];
        opx[0] = 0.5f * (qax + qbx);
        opy[0] = 0.5f * (qay + qby);
        opz[0] = 0.5f * (qaz + qbz);
        od[0] = best;
        n_out = 1;
        feat = feat_word(kFeatEdge, static_cast<uint32_t>(i), static_cast<uint32_t>(j), 0, sat_axis);
    }
    return n_out;
}

// process_pair for the in-scope shapes (narrowphase.cpp:920-1000): at most one
// contact per pair.  Unsupported shape types produce nothing here (the GPU
// path rejects them before launch).
bool contact_for_pair(const rc_bodies &bv, const rc_shapes &sv, uint32_t ia, uint32_t ib, Hit &h) {
    uint32_t sa = bv.shape[ia], sb = bv.shape[ib];
    uint32_t ta = sv.type[sa], tb = sv.type[sb];
    if (ta == kShapeSphere && tb == kShapeSphere)
        return sphere_vs_sphere(bv.pos_x[ia], bv.pos_y[ia], bv.pos_z[ia], sv.half_x[sa],
                                bv.pos_x[ib], bv.pos_y[ib], bv.pos_z[ib], sv.half_x[sb], h);
    if (ta == kShapeSphere && tb == kShapeBox) {
        float E[3][3]; float he[3] = {sv.half_x[sb], sv.half_y[sb], sv.half_z[sb]};
        box_axes(bv.rot_w[ib], bv.rot_x[ib], bv.rot_y[ib], bv.rot_z[ib], E);
        if (!sphere_vs_box(bv.pos_x[ia], bv.pos_y[ia], bv.pos_z[ia], sv.half_x[sa],
                           bv.pos_x[ib], bv.pos_y[ib], bv.pos_z[ib], E, he, h)) return false;
        h.nx = -h.nx; h.ny = -h.ny; h.nz = -h.nz;
        return true;
    }
    if (ta == kShapeBox && tb == kShapeSphere) {
        float E[3][3]; float he[3] = {sv.half_x[sa], sv.half_y[sa], sv.half_z[sa]};
        box_axes(bv.rot_w[ia], bv.rot_x[ia], bv.rot_y[ia], bv.rot_z[ia], E);
        return sphere_vs_box(bv.pos_x[ib], bv.pos_y[ib], bv.pos_z[ib], sv.half_x[sb],
                             bv.pos_x[ia], bv.pos_y[ia], bv.pos_z[ia], E, he, h);
    }
    if (ta == kShapeBox && tb == kShapeBox) {
        float A[3][3], B[3][3];
        float ah[3] = {sv.half_x[sa], sv.half_y[sa], sv.half_z[sa]};
        float bh[3] = {sv.half_x[sb], sv.half_y[sb], sv.half_z[sb]};
        box_axes(bv.rot_w[ia], bv.rot_x[ia], bv.rot_y[ia], bv.rot_z[ia], A);
        box_axes(bv.rot_w[ib], bv.rot_x[ib], bv.rot_y[ib], bv.rot_z[ib], B);
        float px[4], py[4], pz[4], dd[4], nx, ny, nz;
        uint32_t feat = 0;
        int m = box_vs_box(bv.pos_x[ia], bv.pos_y[ia], bv.pos_z[ia], A, ah,
                           bv.pos_x[ib], bv.pos_y[ib], bv.pos_z[ib], B, bh,
                           px, py, pz, nx, ny, nz, dd, feat);
        if (m == 0) return false;
        // One contact per pair: centroid of the manifold, mean depth.
        float sx = 0.f, sy = 0.f, sz = 0.f, sdp = 0.f;
        for (int k = 0; k < m; ++k) { sx += px[k]; sy += py[k]; sz += pz[k]; sdp += dd[k]; }
        float inv = 1.f / static_cast<float>(m);
        h.px = sx * inv; h.py = sy * inv; h.pz = sz * inv;
        h.nx = nx; h.ny = ny; h.nz = nz;
        h.depth = sdp * inv;
        h.feat = feat;
        return true;
    }
    return false;
}

// ----------------------------------------------------------------------------
// XPBD contact projection (xpbd_solver.cpp:40-98, 413-521)
// ----------------------------------------------------------------------------

// World inverse inertia applied to v, built from the current (possibly
// un-normalised) quaternion (xpbd_solver.cpp:40-59).  With M the matrix whose
// row a is the world direction of body axis a (box_axes), the reference
// evaluates M * Ib^-1 * M^T * v (to the body frame with M's columns, back with
// its rows).  The textbook world inverse inertia is M^T * Ib^-1 * M; the two
// agree whenever Ib^-1 is isotropic (cubes, spheres).  Kept exactly as written.
inline void inv_inertia_world(const rc_bodies &b, uint32_t i, float vx, float vy, float vz,
                              float &ox, float &oy, float &oz) {
    float qw = b.rot_w[i], qx = b.rot_x[i], qy = b.rot_y[i], qz = b.rot_z[i];
    float M[3][3];
    box_axes(qw, qx, qy, qz, M);
    float lx = M[0][0] * vx + M[1][0] * vy + M[2][0] * vz;
    float ly = M[0][1] * vx + M[1][1] * vy + M[2][1] * vz;
    float lz = M[0][2] * vx + M[1][2] * vy + M[2][2] * vz;
    float jx = b.iI_xx[i] * lx + b.iI_xy[i] * ly + b.iI_xz[i] * lz;
    float jy = b.iI_xy[i] * lx + b.iI_yy[i] * ly + b.iI_yz[i] * lz;
    float jz = b.iI_xz[i] * lx + b.iI_yz[i] * ly + b.iI_zz[i] * lz;
    ox = M[0][0] * jx + M[0][1] * jy + M[0][2] * jz;
    oy = M[1][0] * jx + M[1][1] * jy + M[1][2] * jz;
    oz = M[2][0] * jx + M[2][1] * jy + M[2][2] * jz;
}

// q += 0.5 * (0, w) (x) q for a dynamic body (never writes inv_mass == 0).
inline void rotate_by(rc_bodies &b, uint32_t i, float wx, float wy, float wz) {
    if (b.inv_mass[i] == 0.f) return;
    float qw = b.rot_w[i], qx = b.rot_x[i], qy = b.rot_y[i], qz = b.rot_z[i];
    b.rot_w[i] += 0.5f * (-wx * qx - wy * qy - wz * qz);
    b.rot_x[i] += 0.5f * (wx * qw + wy * qz - wz * qy);
    b.rot_y[i] += 0.5f * (-wx * qz + wy * qw + wz * qx);
    b.rot_z[i] += 0.5f * (wx * qy - wy * qx + wz * qw);
}

inline void renormalise(rc_bodies &b, uint32_t i) {
    float qw = b.rot_w[i], qx = b.rot_x[i], qy = b.rot_y[i], qz = b.rot_z[i];
    float s = rsqrt_ref(qw * qw + qx * qx + qy * qy + qz * qz);
    b.rot_w[i] = qw * s; b.rot_x[i] = qx * s; b.rot_y[i] = qy * s; b.rot_z[i] = qz * s;
}

struct ContactSoA {
    std::vector<uint32_t> a, b, feat;
    std::vector<float> px, py, pz, nx, ny, nz, depth;
    void resize(size_t n) {
        a.resize(n); b.resize(n); px.resize(n); py.resize(n); pz.resize(n);
        nx.resize(n); ny.resize(n); nz.resize(n); depth.resize(n); feat.resize(n);
    }
};

void project_contact(uint32_t c, const ContactSoA &cs, rc_bodies &bv, float mu, float *lam, float dt) {
    const uint32_t ia = cs.a[c], ib = cs.b[c];
    const float depth = cs.depth[c];
    if (depth <= 0.f) return;
    const float rem = depth - lam[c];
    if (rem <= 1e-6f) return;
    const float nx = cs.nx[c], ny = cs.ny[c], nz = cs.nz[c];
    const float px = cs.px[c], py = cs.py[c], pz = cs.pz[c];
    const float wa = bv.inv_mass[ia], wb = bv.inv_mass[ib];
    // Lever arms from the current body positions to the fixed contact point.
    const float rax = px - bv.pos_x[ia], ray = py - bv.pos_y[ia], raz = pz - bv.pos_z[ia];
    const float rbx = px - bv.pos_x[ib], rby = py - bv.pos_y[ib], rbz = pz - bv.pos_z[ib];
    const float cax = ray * nz - raz * ny, cay = raz * nx - rax * nz, caz = rax * ny - ray * nx;
    const float cbx = rby * nz - rbz * ny, cby = rbz * nx - rbx * nz, cbz = rbx * ny - rby * nx;
    float ka_x, ka_y, ka_z, kb_x, kb_y, kb_z;
    inv_inertia_world(bv, ia, cax, cay, caz, ka_x, ka_y, ka_z);
    inv_inertia_world(bv, ib, cbx, cby, cbz, kb_x, kb_y, kb_z);
    const float wa_ang = cax * ka_x + cay * ka_y + caz * ka_z;
    const float wb_ang = cbx * kb_x + cby * kb_y + cbz * kb_z;
    const float w = wa + wb + wa_ang + wb_ang;
    if (w < 1e-10f) return;
    const float dl = rem / w;
    lam[c] += rem;

    // Positional projection along the normal.
    if (wa != 0.f) { bv.pos_x[ia] += wa * dl * nx; bv.pos_y[ia] += wa * dl * ny; bv.pos_z[ia] += wa * dl * nz; }
    if (wb != 0.f) { bv.pos_x[ib] -= wb * dl * nx; bv.pos_y[ib] -= wb * dl * ny; bv.pos_z[ib] -= wb * dl * nz; }
    rotate_by(bv, ia, dl * ka_x, dl * ka_y, dl * ka_z);
    rotate_by(bv, ib, -dl * kb_x, -dl * kb_y, -dl * kb_z);

    // Relative contact-point velocity (A minus B) after the projection.
    auto rel_vel = [&](float &rx, float &ry, float &rz) {
        float oax = bv.ang_x[ia], oay = bv.ang_y[ia], oaz = bv.ang_z[ia];
        float obx = bv.ang_x[ib], oby = bv.ang_y[ib], obz = bv.ang_z[ib];
        float vax = bv.vel_x[ia] + oay * raz - oaz * ray;
        float vay = bv.vel_y[ia] + oaz * rax - oax * raz;
        float vaz = bv.vel_z[ia] + oax * ray - oay * rax;
        float vbx = bv.vel_x[ib] + oby * rbz - obz * rby;
        float vby = bv.vel_y[ib] + obz * rbx - obx * rbz;
        float vbz = bv.vel_z[ib] + obx * rby - oby * rbx;
        rx = vax - vbx; ry = vay - vby; rz = vaz - vbz;
    };

    float vx, vy, vz;
    rel_vel(vx, vy, vz);
    const float vn = vx * nx + vy * ny + vz * nz;
    if (vn < 0.f) {
        // Inelastic normal impulse.
        const float jn = -vn / w;
        float ga_x, ga_y, ga_z, gb_x, gb_y, gb_z;
        inv_inertia_world(bv, ia, cax, cay, caz, ga_x, ga_y, ga_z);
        inv_inertia_world(bv, ib, cbx, cby, cbz, gb_x, gb_y, gb_z);
        if (wa != 0.f) {
            bv.vel_x[ia] += wa * jn * nx; bv.vel_y[ia] += wa * jn * ny; bv.vel_z[ia] += wa * jn * nz;
            bv.ang_x[ia] += jn * ga_x; bv.ang_y[ia] += jn * ga_y; bv.ang_z[ia] += jn * ga_z;
        }
        if (wb != 0.f) {
            bv.vel_x[ib] -= wb * jn * nx; bv.vel_y[ib] -= wb * jn * ny; bv.vel_z[ib] -= wb * jn * nz;
            bv.ang_x[ib] -= jn * gb_x; bv.ang_y[ib] -= jn * gb_y; bv.ang_z[ib] -= jn * gb_z;
        }
    }

    if (mu > 0.f) {
        // Coulomb friction, capped by mu times the accumulated normal impulse.
        float ux, uy, uz;
        rel_vel(ux, uy, uz);
        const float un = ux * nx + uy * ny + uz * nz;
        const float tx0 = ux - un * nx, ty0 = uy - un * ny, tz0 = uz - un * nz;
        const float t2 = tx0 * tx0 + ty0 * ty0 + tz0 * tz0;
        if (t2 > 1e-12f) {
            const float tl = std::sqrt(t2);
            const float tx = tx0 / tl, ty = ty0 / tl, tz = tz0 / tl;
            const float dax = ray * tz - raz * ty, day = raz * tx - rax * tz, daz = rax * ty - ray * tx;
            const float dbx = rby * tz - rbz * ty, dby = rbz * tx - rbx * tz, dbz = rbx * ty - rby * tx;
            float fa_x, fa_y, fa_z, fb_x, fb_y, fb_z;
            inv_inertia_world(bv, ia, dax, day, daz, fa_x, fa_y, fa_z);
            inv_inertia_world(bv, ib, dbx, dby, dbz, fb_x, fb_y, fb_z);
            const float wt = wa + wb + dax * fa_x + day * fa_y + daz * fa_z
                                     + dbx * fb_x + dby * fb_y + dbz * fb_z;
            if (wt > 1e-10f) {
                const float jcap = lam[c] / (dt * w);
                float jt = tl / wt;
                jt = std::fmin(jt, mu * jcap);
                if (wa != 0.f) {
                    bv.vel_x[ia] -= wa * jt * tx; bv.vel_y[ia] -= wa * jt * ty; bv.vel_z[ia] -= wa * jt * tz;
                    bv.ang_x[ia] -= jt * fa_x; bv.ang_y[ia] -= jt * fa_y; bv.ang_z[ia] -= jt * fa_z;
                }
                if (wb != 0.f) {
                    bv.vel_x[ib] += wb * jt * tx; bv.vel_y[ib] += wb * jt * ty; bv.vel_z[ib] += wb * jt * tz;
                    bv.ang_x[ib] += jt * fb_x; bv.ang_y[ib] += jt * fb_y; bv.ang_z[ib] += jt * fb_z;
                }
            }
        }
    }
}

// ----------------------------------------------------------------------------
// XPBD joints (xpbd_solver.cpp:60-99 helpers, 100-407 constraints, 523-556
// solve_joint).  Bodies are read and written through the SoA arrays in the
// reference's order; bodies with inv_mass == 0 are never written.
// ----------------------------------------------------------------------------
enum : uint8_t { kJFixed = 0, kJRevolute = 1, kJPrismatic = 2, kJBall = 3 };

// atan2 of joint_current_value (xpbd_solver.cpp:317).  The reference calls
// sycl::atan2, on its OpenMP backend the C library's float atan2, which is
// not correctly rounded everywhere.  The oracle's atan2 is the correctly
// rounded one (the float64 atan2 rounded once to float), the function the
// gfx950 kernels reproduce; rc_set_libc_atan2(1) switches to the C library's
// float atan2 to reproduce the reference's own golden articulation hash.
int g_libc_atan2 = 0;
inline float joint_atan2(float y, float x) {
    if (g_libc_atan2) return std::atan2(y, x);
    return static_cast<float>(std::atan2(static_cast<double>(y), static_cast<double>(x)));
}

// Body-local v in world axes with body i's current quaternion
// (qrot_to_world, xpbd_solver.cpp:62-73): the columns of box_axes.
inline void to_world(const rc_bodies &b, uint32_t i, float vx, float vy, float vz, float &ox, float &oy,
                     float &oz) {
    float E[3][3];
    box_axes(b.rot_w[i], b.rot_x[i], b.rot_y[i], b.rot_z[i], E);
    ox = E[0][0] * vx + E[1][0] * vy + E[2][0] * vz;
    oy = E[0][1] * vx + E[1][1] * vy + E[2][1] * vz;
    oz = E[0][2] * vx + E[1][2] * vy + E[2][2] * vz;
}

// inv_mass + (r x n) . I_w^-1 (r x n)  (gen_inv_mass, xpbd_solver.cpp:90-98).
inline float gen_inv_mass(const rc_bodies &b, uint32_t i, float rx, float ry, float rz, float nx, float ny,
                          float nz) {
    const float cx = ry * nz - rz * ny, cy = rz * nx - rx * nz, cz = rx * ny - ry * nx;
    float kx, ky, kz;
    inv_inertia_world(b, i, cx, cy, cz, kx, ky, kz);
    return b.inv_mass[i] + (cx * kx + cy * ky + cz * kz);
}

// Positional impulse sign * dl * n through lever arm r: position, orientation
// and the matching velocity change dx / h (apply_pos_impulse, :103-133).
inline void pos_impulse(rc_bodies &b, uint32_t i, float sign, float dl, float nx, float ny, float nz, float rx,
                        float ry, float rz, float inv_dt) {
    const float im = b.inv_mass[i];
    if (im == 0.f) return;
    const float dx = sign * im * dl * nx, dy = sign * im * dl * ny, dz = sign * im * dl * nz;
    b.pos_x[i] += dx; b.pos_y[i] += dy; b.pos_z[i] += dz;
    b.vel_x[i] += dx * inv_dt; b.vel_y[i] += dy * inv_dt; b.vel_z[i] += dz * inv_dt;
    const float cx = ry * nz - rz * ny, cy = rz * nx - rx * nz, cz = rx * ny - ry * nx;
    float kx, ky, kz;
    inv_inertia_world(b, i, cx, cy, cz, kx, ky, kz);
    const float ax = sign * dl * kx, ay = sign * dl * ky, az = sign * dl * kz;
    rotate_by(b, i, ax, ay, az);
    b.ang_x[i] += ax * inv_dt; b.ang_y[i] += ay * inv_dt; b.ang_z[i] += az * inv_dt;
}

// Anchor coincidence (solve_pos_constraint, :137-164).
inline void pos_constraint(rc_bodies &b, uint32_t ia, uint32_t ib, float rax, float ray, float raz, float rbx,
                           float rby, float rbz, float alpha, float inv_dt) {
    const float eps = 1e-7f;
    const float cx = (b.pos_x[ia] + rax) - (b.pos_x[ib] + rbx);
    const float cy = (b.pos_y[ia] + ray) - (b.pos_y[ib] + rby);
    const float cz = (b.pos_z[ia] + raz) - (b.pos_z[ib] + rbz);
    const float c2 = cx * cx + cy * cy + cz * cz;
    if (c2 < eps * eps) return;
    const float cl = std::sqrt(c2);
    const float nx = cx / cl, ny = cy / cl, nz = cz / cl;
    const float wa = gen_inv_mass(b, ia, rax, ray, raz, nx, ny, nz);
    const float wb = gen_inv_mass(b, ib, rbx, rby, rbz, nx, ny, nz);
    const float w = wa + wb + alpha;
    if (w < 1e-10f) return;
    const float dl = -cl / w;
    pos_impulse(b, ia, +1.f, dl, nx, ny, nz, rax, ray, raz, inv_dt);
    pos_impulse(b, ib, -1.f, dl, nx, ny, nz, rbx, rby, rbz, inv_dt);
}

// Anchor coincidence across the axis only: the prismatic slide stays free
// (solve_transverse_pos, :169-196).
inline void transverse_constraint(rc_bodies &b, uint32_t ia, uint32_t ib, float rax, float ray, float raz,
                                  float rbx, float rby, float rbz, float ux, float uy, float uz, float alpha,
                                  float inv_dt) {
    const float cx = (b.pos_x[ia] + rax) - (b.pos_x[ib] + rbx);
    const float cy = (b.pos_y[ia] + ray) - (b.pos_y[ib] + rby);
    const float cz = (b.pos_z[ia] + raz) - (b.pos_z[ib] + rbz);
    const float along = cx * ux + cy * uy + cz * uz;
    const float tx = cx - along * ux, ty = cy - along * uy, tz = cz - along * uz;
    const float t2 = tx * tx + ty * ty + tz * tz;
    if (t2 < 1e-14f) return;
    const float tl = std::sqrt(t2);
    const float nx = tx / tl, ny = ty / tl, nz = tz / tl;
    const float wa = gen_inv_mass(b, ia, rax, ray, raz, nx, ny, nz);
    const float wb = gen_inv_mass(b, ib, rbx, rby, rbz, nx, ny, nz);
    const float w = wa + wb + alpha;
    if (w < 1e-10f) return;
    const float dl = -tl / w;
    pos_impulse(b, ia, +1.f, dl, nx, ny, nz, rax, ray, raz, inv_dt);
    pos_impulse(b, ib, -1.f, dl, nx, ny, nz, rbx, rby, rbz, inv_dt);
}

// Rotation correction along the world vector c, |c| = error angle
// (solve_ang_constraint, :200-230).  The angular velocities change by the
// rotation times inv_dt (0 for the limit correction).
inline void ang_constraint(rc_bodies &b, uint32_t ia, uint32_t ib, float cx, float cy, float cz, float alpha,
                           float inv_dt) {
    const float c2 = cx * cx + cy * cy + cz * cz;
    if (c2 < 1e-14f) return;
    const float cl = std::sqrt(c2);
    const float nx = cx / cl, ny = cy / cl, nz = cz / cl;
    float ax, ay, az, bx, by, bz;
    inv_inertia_world(b, ia, nx, ny, nz, ax, ay, az);
    inv_inertia_world(b, ib, nx, ny, nz, bx, by, bz);
    const float wa = nx * ax + ny * ay + nz * az;
    const float wb = nx * bx + ny * by + nz * bz;
    const float w = wa + wb + alpha;
    if (w < 1e-10f) return;
    const float dl = -cl / w;
    rotate_by(b, ia, -dl * ax, -dl * ay, -dl * az);
    rotate_by(b, ib, dl * bx, dl * by, dl * bz);
    if (b.inv_mass[ia] != 0.f) {
        b.ang_x[ia] += (-dl * ax) * inv_dt; b.ang_y[ia] += (-dl * ay) * inv_dt; b.ang_z[ia] += (-dl * az) * inv_dt;
    }
    if (b.inv_mass[ib] != 0.f) {
        b.ang_x[ib] += (dl * bx) * inv_dt; b.ang_y[ib] += (dl * by) * inv_dt; b.ang_z[ib] += (dl * bz) * inv_dt;
    }
}

// Relative orientation locked: error 2 * vec(conj(q_a) q_b) in a's world
// axes, shortest arc (solve_fixed_ang, :234-252).
inline void fixed_ang(rc_bodies &b, uint32_t ia, uint32_t ib, float alpha, float inv_dt) {
    const float aw = b.rot_w[ia], ax = b.rot_x[ia], ay = b.rot_y[ia], az = b.rot_z[ia];
    const float bw = b.rot_w[ib], bx = b.rot_x[ib], by = b.rot_y[ib], bz = b.rot_z[ib];
    const float rw = aw * bw + ax * bx + ay * by + az * bz;
    float rx = -ax * bw + aw * bx - az * by + ay * bz;
    float ry = -ay * bw + az * bx + aw * by - ax * bz;
    float rz = -az * bw - ay * bx + ax * by + aw * bz;
    if (rw < 0.f) { rx = -rx; ry = -ry; rz = -rz; }
    float cx, cy, cz;
    to_world(b, ia, rx, ry, rz, cx, cy, cz);
    cx *= 2.f; cy *= 2.f; cz *= 2.f;
    ang_constraint(b, ia, ib, cx, cy, cz, alpha, inv_dt);
}

// Hinge axes aligned: error = axis_a x axis_b, same hemisphere
// (solve_revolute_ang, :257-273).
inline void revolute_ang(rc_bodies &b, uint32_t ia, uint32_t ib, float lx, float ly, float lz, float alpha,
                         float inv_dt) {
    float ax, ay, az, bx, by, bz;
    to_world(b, ia, lx, ly, lz, ax, ay, az);
    to_world(b, ib, lx, ly, lz, bx, by, bz);
    if (ax * bx + ay * by + az * bz < 0.f) { bx = -bx; by = -by; bz = -bz; }
    const float cx = ay * bz - az * by, cy = az * bx - ax * bz, cz = ax * by - ay * bx;
    ang_constraint(b, ia, ib, cx, cy, cz, alpha, inv_dt);
}

// Joint coordinate: offset along the axis (prismatic) or the hinge angle
// between a reference direction seen from both bodies (revolute)
// (joint_current_value, :278-318).
inline float joint_value(const rc_bodies &b, uint32_t ia, uint32_t ib, float lx, float ly, float lz,
                         uint8_t type) {
    float ux, uy, uz;
    to_world(b, ia, lx, ly, lz, ux, uy, uz);
    if (type == kJPrismatic)
        return (b.pos_x[ib] - b.pos_x[ia]) * ux + (b.pos_y[ib] - b.pos_y[ia]) * uy + (b.pos_z[ib] - b.pos_z[ia]) * uz;
    float px, py, pz;
    if (std::fabs(lx) < std::fabs(ly) && std::fabs(lx) < std::fabs(lz)) {
        const float len = std::sqrt(ly * ly + lz * lz);
        px = 0.f; py = -lz / len; pz = ly / len;
    } else if (std::fabs(ly) < std::fabs(lz)) {
        const float len = std::sqrt(lx * lx + lz * lz);
        px = lz / len; py = 0.f; pz = -lx / len;
    } else {
        const float len = std::sqrt(lx * lx + ly * ly);
        px = -ly / len; py = lx / len; pz = 0.f;
    }
    float rx, ry, rz, cx, cy, cz;
    to_world(b, ia, px, py, pz, rx, ry, rz);
    to_world(b, ib, px, py, pz, cx, cy, cz);
    const float sx = ry * cz - rz * cy, sy = rz * cx - rx * cz, sz = rx * cy - ry * cx;
    const float sn = sx * ux + sy * uy + sz * uz;
    const float cs = rx * cx + ry * cy + rz * cz;
    return joint_atan2(sn, cs);
}

// One-sided limit on the joint coordinate (apply_limit, :322-352).
inline void joint_limit(rc_bodies &b, uint32_t ia, uint32_t ib, float lx, float ly, float lz, float lo, float hi,
                        uint8_t type, float inv_dt) {
    if (lo > hi) return;
    const float v = joint_value(b, ia, ib, lx, ly, lz, type);
    float viol = 0.f;
    if (v < lo) viol = v - lo;
    else if (v > hi) viol = v - hi;
    if (std::fabs(viol) < 1e-7f) return;
    float ux, uy, uz;
    to_world(b, ia, lx, ly, lz, ux, uy, uz);
    if (type == kJPrismatic) {
        const float wa = gen_inv_mass(b, ia, 0.f, 0.f, 0.f, ux, uy, uz);
        const float wb = gen_inv_mass(b, ib, 0.f, 0.f, 0.f, ux, uy, uz);
        const float w = wa + wb;
        if (w < 1e-10f) return;
        const float dl = -viol / w;
        pos_impulse(b, ia, +1.f, dl, ux, uy, uz, 0.f, 0.f, 0.f, inv_dt);
        pos_impulse(b, ib, -1.f, dl, ux, uy, uz, 0.f, 0.f, 0.f, inv_dt);
    } else {
        ang_constraint(b, ia, ib, viol * ux, viol * uy, viol * uz, 0.f, 0.f);
    }
}

// PD motor: one velocity impulse along the axis per solve (apply_motor,
// :356-407).
inline void joint_motor(rc_bodies &b, uint32_t ia, uint32_t ib, float lx, float ly, float lz, float target_pos,
                        float target_vel, float stiffness, float damping, float h, uint8_t type) {
    if (stiffness < 1e-6f && damping < 1e-6f) return;
    float ux, uy, uz;
    to_world(b, ia, lx, ly, lz, ux, uy, uz);
    float vrel;
    if (type == kJPrismatic) {
        const float vx = b.vel_x[ib] - b.vel_x[ia], vy = b.vel_y[ib] - b.vel_y[ia], vz = b.vel_z[ib] - b.vel_z[ia];
        vrel = vx * ux + vy * uy + vz * uz;
    } else {
        const float ox = b.ang_x[ib] - b.ang_x[ia], oy = b.ang_y[ib] - b.ang_y[ia], oz = b.ang_z[ib] - b.ang_z[ia];
        vrel = ox * ux + oy * uy + oz * uz;
    }
    float pos_err = 0.f;
    if (stiffness > 1e-6f) pos_err = target_pos - joint_value(b, ia, ib, lx, ly, lz, type);
    const float vel_err = target_vel - vrel;
    const float ima = b.inv_mass[ia], imb = b.inv_mass[ib];
    if (type == kJPrismatic) {
        const float w = ima + imb;
        if (w < 1e-10f) return;
        const float j = (stiffness * pos_err + damping * vel_err) * h / w;
        if (ima != 0.f) { b.vel_x[ia] -= ima * j * ux; b.vel_y[ia] -= ima * j * uy; b.vel_z[ia] -= ima * j * uz; }
        if (imb != 0.f) { b.vel_x[ib] += imb * j * ux; b.vel_y[ib] += imb * j * uy; b.vel_z[ib] += imb * j * uz; }
    } else {
        float ax, ay, az, bx, by, bz;
        inv_inertia_world(b, ia, ux, uy, uz, ax, ay, az);
        inv_inertia_world(b, ib, ux, uy, uz, bx, by, bz);
        const float wa = ux * ax + uy * ay + uz * az;
        const float wb = ux * bx + uy * by + uz * bz;
        const float w = wa + wb;
        if (w < 1e-10f) return;
        const float j = (stiffness * pos_err + damping * vel_err) * h / w;
        if (ima != 0.f) { b.ang_x[ia] -= j * ax; b.ang_y[ia] -= j * ay; b.ang_z[ia] -= j * az; }
        if (imb != 0.f) { b.ang_x[ib] += j * bx; b.ang_y[ib] += j * by; b.ang_z[ib] += j * bz; }
    }
}

// The motor of joint j where it has one (revolute, prismatic).
inline void motor_of(const rc_joints &J, uint32_t j, rc_bodies &b, float dt) {
    const uint8_t t = J.type[j];
    if (t != kJRevolute && t != kJPrismatic) return;
    joint_motor(b, J.body_parent[j], J.body_child[j], J.axis_px[j], J.axis_py[j], J.axis_pz[j], J.target_pos[j],
                J.target_vel[j], J.stiffness[j], J.damping[j], dt, t);
}

// One joint constraint (solve_joint, :524-556).  The lever arms are taken
// once, before any correction of this joint.
inline void solve_joint(const rc_joints &J, uint32_t j, rc_bodies &b, float dt) {
    const uint32_t ia = J.body_parent[j], ib = J.body_child[j];
    const uint8_t t = J.type[j];
    float rax, ray, raz, rbx, rby, rbz;
    to_world(b, ia, J.anchor_px[j], J.anchor_py[j], J.anchor_pz[j], rax, ray, raz);
    to_world(b, ib, J.anchor_cx[j], J.anchor_cy[j], J.anchor_cz[j], rbx, rby, rbz);
    const float alpha_pos = J.compliance_pos[j] / (dt * dt);
    const float alpha_ang = J.compliance_ang[j] / (dt * dt);
    const float lx = J.axis_px[j], ly = J.axis_py[j], lz = J.axis_pz[j];
    const float inv_dt = 1.f / dt;
    if (t == kJPrismatic) {
        float ux, uy, uz;
        to_world(b, ia, lx, ly, lz, ux, uy, uz);
        fixed_ang(b, ia, ib, alpha_ang, inv_dt);
        transverse_constraint(b, ia, ib, rax, ray, raz, rbx, rby, rbz, ux, uy, uz, alpha_pos, inv_dt);
        joint_limit(b, ia, ib, lx, ly, lz, J.limit_lo[j], J.limit_hi[j], t, inv_dt);
    } else {
        pos_constraint(b, ia, ib, rax, ray, raz, rbx, rby, rbz, alpha_pos, inv_dt);
        if (t == kJFixed) {
            fixed_ang(b, ia, ib, alpha_ang, inv_dt);
        } else if (t == kJRevolute) {
            revolute_ang(b, ia, ib, lx, ly, lz, alpha_ang, inv_dt);
            joint_limit(b, ia, ib, lx, ly, lz, J.limit_lo[j], J.limit_hi[j], t, inv_dt);
        }
    }
}

}  // namespace

// ============================================================================
// World: buffers of Broadphase / Narrowphase / XpbdSolver
// ============================================================================
struct rc_world {
    uint32_t max_bodies = 0, max_padded = 0, max_pairs = 0, max_contacts = 0;
    int n_iters = 10;
    float mu = 0.5f;
    // broadphase
    std::vector<uint64_t> morton;
    std::vector<uint32_t> sorted_idx;
    std::vector<int32_t> left, right, parent;
    int32_t root = 0;
    std::vector<uint16_t> aabb;  // 8 per node
    std::vector<uint32_t> ready;
    uint32_t refit_rounds = 0;
    std::vector<rc_pair> pairs;
    std::vector<uint64_t> pair_keys;
    uint32_t pair_count = 0;
    // narrowphase
    ContactSoA store, scratch;
    std::vector<uint64_t> ckeys;
    std::vector<uint32_t> cperm;
    uint32_t n_contacts = 0, last_count = 0;
    // solver
    std::vector<float> lambda;
    std::vector<uint32_t> color;
    std::vector<uint64_t> bpri, bmask;
};

extern "C" {

uint16_t rc_half_from_float(float f) {
    const uint32_t x = fbits(f);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t ax = x & 0x7fffffffu;
    if (ax > 0x7f800000u) return static_cast<uint16_t>(sign | 0x7e00u | ((ax >> 13) & 0x3ffu));
    if (ax >= 0x477ff000u) return static_cast<uint16_t>(sign | 0x7c00u);  // rounds to +-inf
    if (ax >= 0x38800000u) {
        // Normal binary16: re-bias the exponent, round the 13 dropped bits to nearest even.
        uint32_t r = ax + 0x0fffu + ((ax >> 13) & 1u);
        return static_cast<uint16_t>(sign | ((r - 0x38000000u) >> 13));
    }
    const uint32_t e = ax >> 23;
    if (e < 102u) return static_cast<uint16_t>(sign);  // below half the smallest subnormal
    const uint32_t m = (ax & 0x7fffffu) | 0x800000u;
    const uint32_t sh = 126u - e;  // 14..24
    uint32_t q = m >> sh;
    const uint32_t rem = m & ((1u << sh) - 1u);
    const uint32_t half = 1u << (sh - 1u);
    if (rem > half || (rem == half && (q & 1u))) ++q;
    return static_cast<uint16_t>(sign | q);
}

float rc_half_to_float(uint16_t h) {
    const uint32_t sign = static_cast<uint32_t>(h & 0x8000u) << 16;
    const uint32_t e = (h >> 10) & 0x1fu;
    uint32_t m = h & 0x3ffu;
    if (e == 0x1fu) return bitsf(sign | 0x7f800000u | (m << 13));
    if (e != 0) return bitsf(sign | ((e + 112u) << 23) | (m << 13));
    if (m == 0) return bitsf(sign);
    // Subnormal: normalise.
    int ee = 113;
    while (!(m & 0x400u)) { m <<= 1; --ee; }
    m &= 0x3ffu;
    return bitsf(sign | (static_cast<uint32_t>(ee) << 23) | (m << 13));
}

void rc_half_from_float_array(const float *in, uint16_t *out, size_t n) {
    for (size_t i = 0; i < n; ++i) out[i] = rc_half_from_float(in[i]);
}

uint32_t rc_hash_u32(uint32_t x) {
    // Murmur3 fmix32 finaliser (xpbd_solver.cpp:31-34).
    x ^= x >> 16; x *= 0x85ebca6bu;
    x ^= x >> 13; x *= 0xc2b2ae35u;
    x ^= x >> 16;
    return x;
}

uint32_t rc_morton3(uint32_t x, uint32_t y, uint32_t z) {
    // Bit b of x goes to bit 3b, of y to 3b+1, of z to 3b+2 (10 bits each).
    uint32_t m = 0;
    for (uint32_t b = 0; b < 10; ++b) {
        m |= ((x >> b) & 1u) << (3 * b);
        m |= ((y >> b) & 1u) << (3 * b + 1);
        m |= ((z >> b) & 1u) << (3 * b + 2);
    }
    return m;
}

uint64_t rc_fnv1a64(uint64_t h, const void *data, size_t n) {
    const unsigned char *p = static_cast<const unsigned char *>(data);
    for (size_t i = 0; i < n; ++i) { h ^= p[i]; h *= 0x00000100000001B3ULL; }
    return h;
}

uint64_t rc_fnv1a64_state(const rc_bodies *b) {
    // examples/stress_test_blocks/main.cpp:301-311: pos xyz, rot wxyz, all bodies.
    uint64_t h = 0xcbf29ce484222325ULL;
    const size_t bytes = static_cast<size_t>(b->n) * sizeof(float);
    const float *f[7] = {b->pos_x, b->pos_y, b->pos_z, b->rot_w, b->rot_x, b->rot_y, b->rot_z};
    for (int k = 0; k < 7; ++k) h = rc_fnv1a64(h, f[k], bytes);
    return h;
}

void rc_body_setup(uint32_t n, const float *mass, const float *I, const uint32_t *flags_in, rc_bodies *o) {
    for (uint32_t k = 0; k < n; ++k) {
        const bool st = (flags_in[k] & kFlagStatic) || mass[k] <= 0.f;
        o->inv_mass[k] = st ? 0.f : 1.f / mass[k];
        if (st) {
            o->iI_xx[k] = o->iI_yy[k] = o->iI_zz[k] = 0.f;
            o->iI_xy[k] = o->iI_xz[k] = o->iI_yz[k] = 0.f;
        } else {
            // Adjugate / determinant (mat3.hpp:61-79), entries (0,0),(1,1),(2,2),(0,1),(0,2),(1,2).
            const float *m = I + 9 * static_cast<size_t>(k);
            const float det = m[0] * (m[4] * m[8] - m[5] * m[7])
                            - m[1] * (m[3] * m[8] - m[5] * m[6])
                            + m[2] * (m[3] * m[7] - m[4] * m[6]);
            o->iI_xx[k] = (m[4] * m[8] - m[5] * m[7]) / det;
            o->iI_xy[k] = -(m[1] * m[8] - m[2] * m[7]) / det;
            o->iI_xz[k] = (m[1] * m[5] - m[2] * m[4]) / det;
            o->iI_yy[k] = (m[0] * m[8] - m[2] * m[6]) / det;
            o->iI_yz[k] = -(m[0] * m[5] - m[2] * m[3]) / det;
            o->iI_zz[k] = (m[0] * m[4] - m[1] * m[3]) / det;
        }
        o->flags[k] = st ? (flags_in[k] | kFlagStatic) : flags_in[k];
    }
}

void rc_integrate(rc_bodies *b, float gx, float gy, float gz, float dt, int n_substeps) {
    if (b->n == 0) return;
    const float h = dt / static_cast<float>(n_substeps);
    const float hh = 0.5f * h;
    for (int sub = 0; sub < n_substeps; ++sub) {
#pragma omp parallel for schedule(static)
        for (long long ii = 0; ii < static_cast<long long>(b->n); ++ii) {
            const size_t i = static_cast<size_t>(ii);
            if (b->inv_mass[i] == 0.f) continue;
            if (b->flags[i] & kFlagSleeping) continue;
            const float vx = b->vel_x[i] + gx * h;
            const float vy = b->vel_y[i] + gy * h;
            const float vz = b->vel_z[i] + gz * h;
            b->vel_x[i] = vx; b->vel_y[i] = vy; b->vel_z[i] = vz;
            b->pos_x[i] += vx * h; b->pos_y[i] += vy * h; b->pos_z[i] += vz * h;
            const float wx = b->ang_x[i], wy = b->ang_y[i], wz = b->ang_z[i];
            float qw = b->rot_w[i], qx = b->rot_x[i], qy = b->rot_y[i], qz = b->rot_z[i];
            // (0, w) (x) q
            const float ew = -(wx * qx + wy * qy + wz * qz);
            const float ex = (wx * qw + wy * qz - wz * qy);
            const float ey = (wy * qw + wz * qx - wx * qz);
            const float ez = (wz * qw + wx * qy - wy * qx);
            qw += hh * ew; qx += hh * ex; qy += hh * ey; qz += hh * ez;
            const float s = rsqrt_ref(qw * qw + qx * qx + qy * qy + qz * qz);
            b->rot_w[i] = qw * s; b->rot_x[i] = qx * s; b->rot_y[i] = qy * s; b->rot_z[i] = qz * s;
        }
    }
}

void rc_bitonic_sort_u64_u32(uint64_t *keys, uint32_t *vals, size_t n) { bitonic_kv(keys, vals, n); }
void rc_bitonic_sort_u32_u32(uint32_t *keys, uint32_t *vals, size_t n) { bitonic_kv(keys, vals, n); }
void rc_bitonic_sort_u32(uint32_t *keys, size_t n) { bitonic_kv<uint32_t, uint32_t>(keys, nullptr, n); }

rc_world *rc_world_create(uint32_t max_bodies, uint32_t max_pairs, uint32_t max_contacts,
                          int n_iters, float mu) {
    rc_world *w = new rc_world;
    w->max_bodies = max_bodies;
    w->max_padded = next_pow2(max_bodies);
    w->max_pairs = max_pairs;
    w->max_contacts = max_contacts;
    w->n_iters = n_iters;
    w->mu = mu;
    const uint32_t nb = max_bodies > 1 ? max_bodies : 1;
    w->morton.resize(w->max_padded);
    w->sorted_idx.resize(w->max_padded);
    w->left.resize(nb > 1 ? nb - 1 : 1);
    w->right.resize(nb > 1 ? nb - 1 : 1);
    w->parent.resize(nb > 1 ? 2 * nb - 1 : 1);
    w->aabb.resize(8 * static_cast<size_t>(nb > 1 ? 2 * nb - 1 : 1));
    w->ready.resize(nb > 1 ? nb - 1 : 1);
    w->pairs.resize(max_pairs);
    w->pair_keys.resize(next_pow2(max_pairs > 0 ? max_pairs : 1));
    w->store.resize(max_contacts);
    w->scratch.resize(max_contacts);
    w->ckeys.resize(next_pow2(max_contacts));
    w->cperm.resize(next_pow2(max_contacts));
    w->lambda.resize(max_contacts);
    return w;
}

void rc_world_destroy(rc_world *w) { delete w; }

uint32_t rc_bp_build_and_query(rc_world *w, const rc_bodies *bp, const rc_shapes *sp, const rc_aabb *bd) {
    const rc_bodies &b = *bp;
    const rc_shapes &s = *sp;
    const uint32_t n = b.n;
    w->pair_count = 0;
    w->refit_rounds = 0;
    if (n < 2) return 0;
    const int ni = static_cast<int>(n);
    const uint32_t npad = w->max_padded;

    // 1. Morton keys (morton30 << 32 | body), pads = all ones (broadphase.cpp:128-146).
    const float x0 = bd->min_x, x1 = bd->max_x, y0 = bd->min_y, y1 = bd->max_y, z0 = bd->min_z, z1 = bd->max_z;
#pragma omp parallel for schedule(static)
    for (long long ii = 0; ii < static_cast<long long>(npad); ++ii) {
        const uint32_t i = static_cast<uint32_t>(ii);
        if (i < n) {
            float tx = (b.pos_x[i] - x0) / (x1 - x0);
            float ty = (b.pos_y[i] - y0) / (y1 - y0);
            float tz = (b.pos_z[i] - z0) / (z1 - z0);
            tx = tx < 0.f ? 0.f : (tx > 1.f ? 1.f : tx);
            ty = ty < 0.f ? 0.f : (ty > 1.f ? 1.f : ty);
            tz = tz < 0.f ? 0.f : (tz > 1.f ? 1.f : tz);
            const uint32_t qx = static_cast<uint32_t>(tx * 1023.f);
            const uint32_t qy = static_cast<uint32_t>(ty * 1023.f);
            const uint32_t qz = static_cast<uint32_t>(tz * 1023.f);
            w->morton[i] = (static_cast<uint64_t>(rc_morton3(qx, qy, qz)) << 32) | i;
            w->sorted_idx[i] = i;
        } else {
            w->morton[i] = ~0ULL;
            w->sorted_idx[i] = ~0u;
        }
    }
    // 2. Key sort.
    bitonic_kv(w->morton.data(), w->sorted_idx.data(), npad);

    // 3-5. Karras tree, parents, root.
    std::fill(w->parent.begin(), w->parent.begin() + (2 * n - 1), -1);
    std::fill(w->ready.begin(), w->ready.begin() + (n - 1), 0u);
    const uint64_t *codes = w->morton.data();
#pragma omp parallel for schedule(static)
    for (int i = 0; i < ni - 1; ++i) {
        int lo, hi, g;
        karras_node(codes, ni, i, lo, hi, g);
        const int l = (lo == g) ? (ni - 1 + lo) : g;
        const int r = (g + 1 == hi) ? (ni - 1 + g + 1) : (g + 1);
        w->left[i] = l;
        w->right[i] = r;
        w->parent[l] = i;
        w->parent[r] = i;
    }
    w->root = 0;
    for (int i = 0; i < ni - 1; ++i)
        if (w->parent[i] == -1) w->root = i;

    // 6a. Leaf boxes in fp16 (broadphase.cpp:216-251).
    uint16_t *box = w->aabb.data();
#pragma omp parallel for schedule(static)
    for (int k = 0; k < ni; ++k) {
        const uint32_t body = w->sorted_idx[k];
        const float px = b.pos_x[body], py = b.pos_y[body], pz = b.pos_z[body];
        const uint32_t sh = b.shape[body];
        const float hx = s.half_x[sh];
        float ex, ey, ez;
        if (s.type[sh] == kShapeSphere) {
            ex = ey = ez = hx;
        } else {
            const float hy = s.half_y[sh], hz = s.half_z[sh];
            const Rot3 R = leaf_rotation(b.rot_w[body], b.rot_x[body], b.rot_y[body], b.rot_z[body]);
            ex = std::fabs(R.m[0][0]) * hx + std::fabs(R.m[0][1]) * hy + std::fabs(R.m[0][2]) * hz;
            ey = std::fabs(R.m[1][0]) * hx + std::fabs(R.m[1][1]) * hy + std::fabs(R.m[1][2]) * hz;
            ez = std::fabs(R.m[2][0]) * hx + std::fabs(R.m[2][1]) * hy + std::fabs(R.m[2][2]) * hz;
        }
        uint16_t *o = box + 8 * static_cast<size_t>(ni - 1 + k);
        o[0] = rc_half_from_float(px - ex); o[1] = rc_half_from_float(py - ey); o[2] = rc_half_from_float(pz - ez);
        o[3] = rc_half_from_float(px + ex); o[4] = rc_half_from_float(py + ey); o[5] = rc_half_from_float(pz + ez);
        o[6] = 0; o[7] = 0;
    }
    // 6b. Level-synchronous merge rounds (broadphase.cpp:257-287).  Each round
    // reads the readiness recorded by previous rounds only.
    {
        std::vector<uint32_t> prev(n - 1);
        const int root = w->root;
        uint32_t rounds = 0;
        for (int r = 0; r < 64 && !w->ready[root]; ++r) {
            std::copy(w->ready.begin(), w->ready.begin() + (n - 1), prev.begin());
#pragma omp parallel for schedule(static)
            for (int p = 0; p < ni - 1; ++p) {
                if (prev[p]) continue;
                const int lc = w->left[p], rc = w->right[p];
                const bool lr = lc >= ni - 1 || prev[lc];
                const bool rr = rc >= ni - 1 || prev[rc];
                if (!lr || !rr) continue;
                const uint16_t *L = box + 8 * static_cast<size_t>(lc);
                const uint16_t *R = box + 8 * static_cast<size_t>(rc);
                uint16_t *o = box + 8 * static_cast<size_t>(p);
                for (int a = 0; a < 3; ++a)
                    o[a] = rc_half_from_float(std::fmin(rc_half_to_float(L[a]), rc_half_to_float(R[a])));
                for (int a = 3; a < 6; ++a)
                    o[a] = rc_half_from_float(std::fmax(rc_half_to_float(L[a]), rc_half_to_float(R[a])));
                o[6] = 0; o[7] = 0;
                w->ready[p] = 1u;
            }
            ++rounds;
        }
        w->refit_rounds = rounds;
    }

    // 7. Per-leaf traversal with an explicit stack (broadphase.cpp:304-344).
    uint32_t count = 0;
    const uint32_t cap = w->max_pairs;
#pragma omp parallel for schedule(dynamic, 64)
    for (int j = 0; j < ni; ++j) {
        const uint16_t *J = box + 8 * static_cast<size_t>(ni - 1 + j);
        const float jmn[3] = {rc_half_to_float(J[0]), rc_half_to_float(J[1]), rc_half_to_float(J[2])};
        const float jmx[3] = {rc_half_to_float(J[3]), rc_half_to_float(J[4]), rc_half_to_float(J[5])};
        int stack[64];
        int top = 0;
        stack[top++] = w->root;
        while (top > 0) {
            const int node = stack[--top];
            const uint16_t *B = box + 8 * static_cast<size_t>(node);
            bool miss = false;
            for (int a = 0; a < 3 && !miss; ++a)
                miss = rc_half_to_float(B[a]) > jmx[a] || rc_half_to_float(B[a + 3]) < jmn[a];
            if (miss) continue;
            if (node >= ni - 1) {
                const int i = node - (ni - 1);
                if (i < j) {
                    const uint32_t slot = atomic_inc_u32(&count);
                    if (slot < cap) {
                        uint32_t a = w->sorted_idx[i], c = w->sorted_idx[j];
                        if (a > c) std::swap(a, c);
                        w->pairs[slot] = rc_pair{a, c};
                    }
                }
                continue;
            }
            stack[top++] = w->left[node];
            stack[top++] = w->right[node];
        }
    }
    w->pair_count = count;
    return count;
}

void rc_bp_copy_out(const rc_world *w, uint32_t n, uint64_t *morton, uint32_t *sorted_idx,
                    int32_t *left, int32_t *right, int32_t *parent, int32_t *root,
                    uint16_t *aabb, uint32_t *refit_rounds) {
    if (morton) std::copy(w->morton.begin(), w->morton.begin() + n, morton);
    if (sorted_idx) std::copy(w->sorted_idx.begin(), w->sorted_idx.begin() + n, sorted_idx);
    if (n >= 2) {
        if (left) std::copy(w->left.begin(), w->left.begin() + (n - 1), left);
        if (right) std::copy(w->right.begin(), w->right.begin() + (n - 1), right);
        if (parent) std::copy(w->parent.begin(), w->parent.begin() + (2 * n - 1), parent);
        if (aabb) std::copy(w->aabb.begin(), w->aabb.begin() + 8 * static_cast<size_t>(2 * n - 1), aabb);
    }
    if (root) *root = w->root;
    if (refit_rounds) *refit_rounds = w->refit_rounds;
}

void rc_bp_sort_pairs(rc_world *w, uint32_t n_pairs) {
    // broadphase.cpp:348-359 (count clamped to the buffer, see DESIGN notes).
    if (n_pairs > w->max_pairs) n_pairs = w->max_pairs;
    if (n_pairs <= 1) return;
    const uint32_t ns = next_pow2(n_pairs);
    std::vector<rc_pair> tmp(ns, rc_pair{~0u, ~0u});
    std::copy(w->pairs.begin(), w->pairs.begin() + n_pairs, tmp.begin());
    std::fill(w->pair_keys.begin(), w->pair_keys.begin() + ns, ~0ULL);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < static_cast<long long>(n_pairs); ++i)
        w->pair_keys[i] = (static_cast<uint64_t>(tmp[i].a) << 32) | tmp[i].b;
    bitonic_kv(w->pair_keys.data(), tmp.data(), ns);
    std::copy(tmp.begin(), tmp.begin() + n_pairs, w->pairs.begin());
}

void rc_bp_copy_pairs(const rc_world *w, rc_pair *out, uint32_t n) {
    if (n > w->max_pairs) n = w->max_pairs;
    std::copy(w->pairs.begin(), w->pairs.begin() + n, out);
}

const rc_pair *rc_bp_pairs(const rc_world *w) { return w->pairs.data(); }

uint32_t rc_np_run(rc_world *w, const rc_pair *pairs, uint32_t n_pairs, const rc_bodies *bp, const rc_shapes *sp) {
    w->n_contacts = 0;
    w->last_count = 0;
    if (n_pairs == 0) return 0;
    const uint32_t cap = w->max_contacts;
    uint32_t cnt = 0;
    ContactSoA &sc = w->scratch;
    // 1. Per-pair tests appending to scratch with key (pair << 32 | sub).
#pragma omp parallel for schedule(dynamic, 256)
    for (long long ii = 0; ii < static_cast<long long>(n_pairs); ++ii) {
        const uint32_t i = static_cast<uint32_t>(ii);
        Hit h;
        const uint32_t ia = pairs[i].a, ib = pairs[i].b;
        if (!contact_for_pair(*bp, *sp, ia, ib, h)) continue;
        const uint32_t slot = atomic_inc_u32(&cnt);
        if (slot >= cap) continue;
        sc.a[slot] = ia; sc.b[slot] = ib;
        sc.px[slot] = h.px; sc.py[slot] = h.py; sc.pz[slot] = h.pz;
        sc.nx[slot] = h.nx; sc.ny[slot] = h.ny; sc.nz[slot] = h.nz;
        sc.depth[slot] = h.depth;
        sc.feat[slot] = h.feat;
        w->ckeys[slot] = static_cast<uint64_t>(i) << 32;  // sub index 0
    }
    w->last_count = cnt;
    uint32_t nc = cnt > cap ? cap : cnt;
    if (nc == 0) return 0;
    // 2. Key sort (power-of-two padded), then gather into the store.
    const uint32_t ns = next_pow2(nc);
    for (uint32_t i = nc; i < ns; ++i) w->ckeys[i] = ~0ULL;
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < static_cast<long long>(ns); ++i) w->cperm[i] = static_cast<uint32_t>(i);
    bitonic_kv(w->ckeys.data(), w->cperm.data(), ns);
    ContactSoA &st = w->store;
#pragma omp parallel for schedule(static)
    for (long long kk = 0; kk < static_cast<long long>(nc); ++kk) {
        const uint32_t k = static_cast<uint32_t>(kk), src = w->cperm[k];
        st.a[k] = sc.a[src]; st.b[k] = sc.b[src];
        st.px[k] = sc.px[src]; st.py[k] = sc.py[src]; st.pz[k] = sc.pz[src];
        st.nx[k] = sc.nx[src]; st.ny[k] = sc.ny[src]; st.nz[k] = sc.nz[src];
        st.depth[k] = sc.depth[src];
        st.feat[k] = sc.feat[src];
    }
    w->n_contacts = nc;
    return nc;
}

void rc_np_features(const rc_world *w, uint32_t *out, uint32_t n) {
    if (n > w->n_contacts) n = w->n_contacts;
    std::copy(w->store.feat.begin(), w->store.feat.begin() + n, out);
}

void rc_np_contacts(const rc_world *w, rc_contacts *o) {
    const ContactSoA &s = w->store;
    o->body_a = s.a.data(); o->body_b = s.b.data();
    o->pos_x = s.px.data(); o->pos_y = s.py.data(); o->pos_z = s.pz.data();
    o->norm_x = s.nx.data(); o->norm_y = s.ny.data(); o->norm_z = s.nz.data();
    o->depth = s.depth.data();
    o->n = w->n_contacts;
    o->last_count = w->last_count;
}

void rc_np_set_contacts(rc_world *w, uint32_t n, const uint32_t *a, const uint32_t *b,
                        const float *px, const float *py, const float *pz,
                        const float *nx, const float *ny, const float *nz, const float *depth) {
    if (n > w->max_contacts) n = w->max_contacts;
    ContactSoA &s = w->store;
    std::copy(a, a + n, s.a.begin()); std::copy(b, b + n, s.b.begin());
    std::copy(px, px + n, s.px.begin()); std::copy(py, py + n, s.py.begin()); std::copy(pz, pz + n, s.pz.begin());
    std::copy(nx, nx + n, s.nx.begin()); std::copy(ny, ny + n, s.ny.begin()); std::copy(nz, nz + n, s.nz.begin());
    std::copy(depth, depth + n, s.depth.begin());
    w->n_contacts = n;
    w->last_count = n;
}

static void run_serial(rc_world *w, rc_bodies *bv, const rc_joints &J, float dt, uint32_t nc) {
    // xpbd_solver.cpp:559-587 on one work-item: motors once, then per
    // iteration contacts, joints, renormalisation of every body.
    for (uint32_t j = 0; j < J.n; ++j) motor_of(J, j, *bv, dt);
    for (int it = 0; it < w->n_iters; ++it) {
        for (uint32_t c = 0; c < nc; ++c) project_contact(c, w->store, *bv, w->mu, w->lambda.data(), dt);
        for (uint32_t j = 0; j < J.n; ++j) solve_joint(J, j, *bv, dt);
        for (uint32_t i = 0; i < bv->n; ++i) renormalise(*bv, i);
    }
}

void rc_solve(rc_world *w, rc_bodies *bv, float dt, int force_serial, rc_solve_stats *st) {
    rc_solve_joints(w, bv, nullptr, dt, force_serial, st);
}

void rc_set_libc_atan2(int on) { g_libc_atan2 = on ? 1 : 0; }
float rc_joint_atan2(float y, float x) { return joint_atan2(y, x); }

void rc_solve_joints(rc_world *w, rc_bodies *bv, const rc_joints *joints, float dt, int force_serial,
                     rc_solve_stats *st) {
    rc_joints none{};
    const rc_joints &J = joints ? *joints : none;
    std::fill(w->lambda.begin(), w->lambda.end(), 0.f);
    const uint32_t nc = w->n_contacts, nj = J.n;
    // Constraint k: contact k for k < nc, joint k - nc after (xpbd_solver.cpp:611-614).
    const uint32_t n_con = nc + nj;
    rc_solve_stats local{};
    local.n_constraints = n_con;
    if (n_con == 0) { if (st) *st = local; return; }
    if (force_serial || n_con <= 256u) {
        run_serial(w, bv, J, dt, nc);
        local.used_serial = 1;
        if (st) *st = local;
        return;
    }
    // Jones-Plassmann colouring in rounds of two passes (xpbd_solver.cpp:625-695).
    const uint32_t kUncol = 0xFFFFFFFFu;
    const uint32_t nb = bv->n;
    if (w->color.size() < n_con) w->color.resize(n_con);
    if (w->bpri.size() < nb) { w->bpri.resize(nb); w->bmask.resize(nb); }
    uint32_t *color = w->color.data();
    uint64_t *bpri = w->bpri.data();
    uint64_t *bmask = w->bmask.data();
    std::fill(color, color + n_con, kUncol);
    std::fill(bmask, bmask + nb, 0ull);
    uint32_t meta_max = 0, meta_over = 0, meta_rem = n_con;
    const ContactSoA &cs = w->store;
    const float *im = bv->inv_mass;
    auto ends = [&](uint32_t k, uint32_t &a, uint32_t &b) {
        if (k < nc) { a = cs.a[k]; b = cs.b[k]; }
        else { a = J.body_parent[k - nc]; b = J.body_child[k - nc]; }
    };
    uint32_t r = 0, h_rem = n_con;
    while (h_rem > 0 && r < 256u) {
        for (uint32_t bi = 0; bi < 8u && r < 256u; ++bi, ++r) {
            std::fill(bpri, bpri + nb, 0ull);
#pragma omp parallel for schedule(static)
            for (long long kk = 0; kk < static_cast<long long>(n_con); ++kk) {
                const uint32_t k = static_cast<uint32_t>(kk);
                if (color[k] != kUncol) continue;
                uint32_t a, b;
                ends(k, a, b);
                const uint64_t pk = jp_priority(k);
                if (im[a] != 0.f) atomic_max_u64(&bpri[a], pk);
                if (im[b] != 0.f) atomic_max_u64(&bpri[b], pk);
            }
#pragma omp parallel for schedule(static)
            for (long long kk = 0; kk < static_cast<long long>(n_con); ++kk) {
                const uint32_t k = static_cast<uint32_t>(kk);
                if (color[k] != kUncol) continue;
                uint32_t a, b;
                ends(k, a, b);
                const bool da = im[a] != 0.f, db = im[b] != 0.f;
                const uint64_t pk = jp_priority(k);
                if ((da && bpri[a] != pk) || (db && bpri[b] != pk)) continue;
                const uint64_t used = (da ? __atomic_load_n(&bmask[a], __ATOMIC_RELAXED) : 0ull) |
                                      (db ? __atomic_load_n(&bmask[b], __ATOMIC_RELAXED) : 0ull);
                const uint64_t fr = ~used;
                if (fr == 0ull) { __atomic_store_n(&meta_over, 1u, __ATOMIC_RELAXED); continue; }
                const uint32_t col = static_cast<uint32_t>(__builtin_ctzll(fr));
                color[k] = col;
                const uint64_t bit = 1ull << col;
                if (da) __atomic_fetch_or(&bmask[a], bit, __ATOMIC_RELAXED);
                if (db) __atomic_fetch_or(&bmask[b], bit, __ATOMIC_RELAXED);
                atomic_max_u32(&meta_max, col);
                __atomic_fetch_sub(&meta_rem, 1u, __ATOMIC_RELAXED);
            }
        }
        h_rem = meta_rem;
    }
    local.rounds = r;
    local.overflow = meta_over;
    if (h_rem > 0 || meta_over) {
        run_serial(w, bv, J, dt, nc);
        local.used_serial = 1;
        if (st) *st = local;
        return;
    }
    const uint32_t ncol = meta_max + 1;
    local.n_colors = ncol;
    float *lam = w->lambda.data();
    // Motor pre-pass, colour by colour, joints only (xpbd_solver.cpp:703-719).
    if (nj > 0) {
        for (uint32_t ci = 0; ci < ncol; ++ci) {
#pragma omp parallel for schedule(static)
            for (long long kk = nc; kk < static_cast<long long>(n_con); ++kk) {
                const uint32_t k = static_cast<uint32_t>(kk);
                if (color[k] == ci) motor_of(J, k - nc, *bv, dt);
            }
        }
    }
    for (int it = 0; it < w->n_iters; ++it) {
        for (uint32_t ci = 0; ci < ncol; ++ci) {
#pragma omp parallel for schedule(static)
            for (long long kk = 0; kk < static_cast<long long>(n_con); ++kk) {
                const uint32_t k = static_cast<uint32_t>(kk);
                if (color[k] != ci) continue;
                if (k < nc) project_contact(k, cs, *bv, w->mu, lam, dt);
                else solve_joint(J, k - nc, *bv, dt);
            }
        }
#pragma omp parallel for schedule(static)
        for (long long i = 0; i < static_cast<long long>(nb); ++i) renormalise(*bv, static_cast<uint32_t>(i));
    }
    if (st) *st = local;
}

void rc_solve_colors(const rc_world *w, uint32_t *out, uint32_t n) {
    if (n > w->color.size()) n = static_cast<uint32_t>(w->color.size());
    std::copy(w->color.begin(), w->color.begin() + n, out);
}

void rc_frame(rc_world *w, rc_bodies *b, const rc_shapes *s, const rc_aabb *bounds,
              float gx, float gy, float gz, float dt, int n_substeps, int force_serial,
              rc_frame_stats *st) {
    rc_frame_joints(w, b, s, bounds, nullptr, gx, gy, gz, dt, n_substeps, force_serial, st);
}

void rc_frame_joints(rc_world *w, rc_bodies *b, const rc_shapes *s, const rc_aabb *bounds, const rc_joints *joints,
                     float gx, float gy, float gz, float dt, int n_substeps, int force_serial,
                     rc_frame_stats *st) {
    rc_integrate(b, gx, gy, gz, dt, n_substeps);
    const uint32_t np = rc_bp_build_and_query(w, b, s, bounds);
    rc_bp_sort_pairs(w, np);
    const uint32_t npc = np > w->max_pairs ? w->max_pairs : np;
    const uint32_t nc = rc_np_run(w, w->pairs.data(), npc, b, s);
    rc_solve_stats ss{};
    rc_solve_joints(w, b, joints, dt, force_serial, &ss);
    if (st) { st->n_pairs = np; st->n_contacts = nc; st->n_colors = ss.n_colors; st->used_serial = ss.used_serial; }
}

// ---- ray queries over the broadphase's LBVH (core/src/ray_query.cpp:12-195) ----
namespace {

// Slab test of a ray against a box (ray_query.cpp:14-36): entered before
// best_t and not entirely behind the origin.
bool rq_aabb_hit(float ox, float oy, float oz, float ix, float iy, float iz, float mnx, float mny, float mnz,
                 float mxx, float mxy, float mxz, float best_t) {
    const float t1x = (mnx - ox) * ix, t2x = (mxx - ox) * ix;
    const float t1y = (mny - oy) * iy, t2y = (mxy - oy) * iy;
    const float t1z = (mnz - oz) * iz, t2z = (mxz - oz) * iz;
    float tmin = t1x < t2x ? t1x : t2x;
    float tmax = t1x < t2x ? t2x : t1x;
    const float ylo = t1y < t2y ? t1y : t2y, yhi = t1y < t2y ? t2y : t1y;
    tmin = tmin > ylo ? tmin : ylo;
    tmax = tmax < yhi ? tmax : yhi;
    const float zlo = t1z < t2z ? t1z : t2z, zhi = t1z < t2z ? t2z : t1z;
    tmin = tmin > zlo ? tmin : zlo;
    tmax = tmax < zhi ? tmax : zhi;
    return tmax >= 0.f && tmax >= tmin && tmin < best_t;
}

// v rotated by the conjugate of q (quat.hpp:31,39-42: t = 2 u x v,
// v + w t + u x t with u = -xyz).
void rq_rotate_inv(float w, float x, float y, float z, const float v[3], float o[3]) {
    const float ux = -x, uy = -y, uz = -z;
    const float tx = (uy * v[2] - uz * v[1]) * 2.f;
    const float ty = (uz * v[0] - ux * v[2]) * 2.f;
    const float tz = (ux * v[1] - uy * v[0]) * 2.f;
    o[0] = (v[0] + tx * w) + (uy * tz - uz * ty);
    o[1] = (v[1] + ty * w) + (uz * tx - ux * tz);
    o[2] = (v[2] + tz * w) + (ux * ty - uy * tx);
}

// Oriented box (ray_query.cpp:39-77); -1 on a miss.
float rq_box_hit(const float o[3], const float d[3], const float p[3], const float q[4], const float h[3],
                 float best_t) {
    const float oc[3] = {o[0] - p[0], o[1] - p[1], o[2] - p[2]};
    float lo[3], ld[3];
    rq_rotate_inv(q[0], q[1], q[2], q[3], oc, lo);
    rq_rotate_inv(q[0], q[1], q[2], q[3], d, ld);
    const float ix = 1.f / ld[0], iy = 1.f / ld[1], iz = 1.f / ld[2];
    const float t1x = (-h[0] - lo[0]) * ix, t2x = (h[0] - lo[0]) * ix;
    const float t1y = (-h[1] - lo[1]) * iy, t2y = (h[1] - lo[1]) * iy;
    const float t1z = (-h[2] - lo[2]) * iz, t2z = (h[2] - lo[2]) * iz;
    float tmin = t1x < t2x ? t1x : t2x;
    float tmax = t1x < t2x ? t2x : t1x;
    const float ylo = t1y < t2y ? t1y : t2y, yhi = t1y < t2y ? t2y : t1y;
    tmin = tmin > ylo ? tmin : ylo;
    tmax = tmax < yhi ? tmax : yhi;
    const float zlo = t1z < t2z ? t1z : t2z, zhi = t1z < t2z ? t2z : t1z;
    tmin = tmin > zlo ? tmin : zlo;
    tmax = tmax < zhi ? tmax : zhi;
    if (tmax < 0.f || tmin > tmax) return -1.f;
    const float t = tmin >= 0.f ? tmin : tmax;
    return (t >= 0.f && t < best_t) ? t : -1.f;
}

// Sphere (ray_query.cpp:80-95); -1 on a miss.
float rq_sphere_hit(const float o[3], const float d[3], const float c[3], float r, float best_t) {
    const float ocx = o[0] - c[0], ocy = o[1] - c[1], ocz = o[2] - c[2];
    const float a = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    const float b = d[0] * ocx + d[1] * ocy + d[2] * ocz;
    const float cc = ocx * ocx + ocy * ocy + ocz * ocz - r * r;
    const float disc = b * b - a * cc;
    if (disc < 0.f) return -1.f;
    const float sq = sqrtf(disc);
    float t = (-b - sq) / a;
    if (t < 0.f) t = (-b + sq) / a;
    return (t >= 0.f && t < best_t) ? t : -1.f;
}

void rq_leaf(uint32_t bi, const float o[3], const float d[3], const rc_bodies &b, const rc_shapes &s, float &best_t,
             uint32_t &best) {
    const uint32_t sh = b.shape[bi];
    const float p[3] = {b.pos_x[bi], b.pos_y[bi], b.pos_z[bi]};
    float t;
    if (s.type[sh] == 1u) {
        t = rq_sphere_hit(o, d, p, s.half_x[sh], best_t);
    } else {
        const float q[4] = {b.rot_w[bi], b.rot_x[bi], b.rot_y[bi], b.rot_z[bi]};
        const float h[3] = {s.half_x[sh], s.half_y[sh], s.half_z[sh]};
        t = rq_box_hit(o, d, p, q, h, best_t);
    }
    if (t >= 0.f) {
        best_t = t;
        best = bi;
    }
}

}  // namespace

void rc_ray_query(const rc_world *w, uint32_t n, const rc_bodies *b, const rc_shapes *s, const rc_ray *rays,
                  rc_ray_hit *hits, uint32_t n_rays) {
    // ray_query.cpp:126-195: depth-first from the root with a 64-entry stack,
    // right child pushed first so the left subtree is visited first.
#pragma omp parallel for schedule(static)
    for (long long ii = 0; ii < static_cast<long long>(n_rays); ++ii) {
        const rc_ray r = rays[ii];
        const float o[3] = {r.ox, r.oy, r.oz}, d[3] = {r.dx, r.dy, r.dz};
        float best_t = r.t_max;
        uint32_t best = 0xFFFFFFFFu;
        const float ix = 1.f / r.dx, iy = 1.f / r.dy, iz = 1.f / r.dz;
        if (n == 1) {
            rq_leaf(0u, o, d, *b, *s, best_t, best);
        } else if (n > 1) {
            int32_t stack[64];
            int top = 0;
            stack[top++] = w->root;
            const uint32_t n_int = n - 1u;
            while (top > 0) {
                const int32_t node = stack[--top];
                const uint16_t *bx = &w->aabb[8 * static_cast<size_t>(node)];
                const float mn[3] = {rc_half_to_float(bx[0]), rc_half_to_float(bx[1]), rc_half_to_float(bx[2])};
                const float mx[3] = {rc_half_to_float(bx[3]), rc_half_to_float(bx[4]), rc_half_to_float(bx[5])};
                if (!rq_aabb_hit(o[0], o[1], o[2], ix, iy, iz, mn[0], mn[1], mn[2], mx[0], mx[1], mx[2], best_t))
                    continue;
                if (static_cast<uint32_t>(node) >= n_int) {
                    rq_leaf(w->sorted_idx[node - static_cast<int32_t>(n_int)], o, d, *b, *s, best_t, best);
                } else if (top + 1 < 64) {
                    stack[top++] = w->right[node];
                    stack[top++] = w->left[node];
                }
            }
        }
        hits[ii].t = best_t;
        hits[ii].body = best;
    }
}

int rc_pair_contact_feat(uint32_t ta, uint32_t tb, const float *ha, const float *hb, const float *pa,
                         const float *pb, const float *qa, const float *qb, float *out, uint32_t *feat) {
    float px[2] = {pa[0], pb[0]}, py[2] = {pa[1], pb[1]}, pz[2] = {pa[2], pb[2]};
    float rw[2] = {qa[0], qb[0]}, rx[2] = {qa[1], qb[1]}, ry[2] = {qa[2], qb[2]}, rz[2] = {qa[3], qb[3]};
    uint32_t shape[2] = {0, 1}, flags[2] = {0, 0};
    rc_bodies b{};
    b.pos_x = px; b.pos_y = py; b.pos_z = pz;
    b.rot_w = rw; b.rot_x = rx; b.rot_y = ry; b.rot_z = rz;
    b.shape = shape; b.flags = flags; b.n = 2;
    uint32_t type[2] = {ta, tb}, ext[2] = {0, 0};
    float hx[2] = {ha[0], hb[0]}, hy[2] = {ha[1], hb[1]}, hz[2] = {ha[2], hb[2]};
    rc_shapes sh{type, hx, hy, hz, ext, 2};
    Hit h{};
    if (!contact_for_pair(b, sh, 0, 1, h)) return 0;
    out[0] = h.px; out[1] = h.py; out[2] = h.pz;
    out[3] = h.nx; out[4] = h.ny; out[5] = h.nz;
    out[6] = h.depth;
    if (feat) *feat = h.feat;
    return 1;
}

int rc_pair_contact(uint32_t ta, uint32_t tb, const float *ha, const float *hb, const float *pa,
                    const float *pb, const float *qa, const float *qb, float *out) {
    return rc_pair_contact_feat(ta, tb, ha, hb, pa, pb, qa, qb, out, nullptr);
}

void rc_project_contact(const float *contact, float *a, float *b, float mu, float *lam, float dt) {
    float f[20][2];
    for (int k = 0; k < 20; ++k) { f[k][0] = a[k]; f[k][1] = b[k]; }
    uint32_t shape[2] = {0, 0}, flags[2] = {0, 0};
    rc_bodies bv{};
    float **dst[20] = {&bv.pos_x, &bv.pos_y, &bv.pos_z, &bv.rot_w, &bv.rot_x, &bv.rot_y, &bv.rot_z,
                       &bv.vel_x, &bv.vel_y, &bv.vel_z, &bv.ang_x, &bv.ang_y, &bv.ang_z, &bv.inv_mass,
                       &bv.iI_xx, &bv.iI_yy, &bv.iI_zz, &bv.iI_xy, &bv.iI_xz, &bv.iI_yz};
    for (int k = 0; k < 20; ++k) *dst[k] = f[k];
    bv.shape = shape; bv.flags = flags; bv.n = 2;
    ContactSoA cs;
    cs.resize(1);
    cs.a[0] = 0; cs.b[0] = 1;
    cs.px[0] = contact[0]; cs.py[0] = contact[1]; cs.pz[0] = contact[2];
    cs.nx[0] = contact[3]; cs.ny[0] = contact[4]; cs.nz[0] = contact[5];
    cs.depth[0] = contact[6];
    project_contact(0, cs, bv, mu, lam, dt);
    for (int k = 0; k < 20; ++k) { a[k] = f[k][0]; b[k] = f[k][1]; }
}

void rc_leaf_extent(uint32_t stype, const float *he, const float *q, float *out) {
    if (stype == kShapeSphere) { out[0] = out[1] = out[2] = he[0]; return; }
    const Rot3 R = leaf_rotation(q[0], q[1], q[2], q[3]);
    out[0] = std::fabs(R.m[0][0]) * he[0] + std::fabs(R.m[0][1]) * he[1] + std::fabs(R.m[0][2]) * he[2];
    out[1] = std::fabs(R.m[1][0]) * he[0] + std::fabs(R.m[1][1]) * he[1] + std::fabs(R.m[1][2]) * he[2];
    out[2] = std::fabs(R.m[2][0]) * he[0] + std::fabs(R.m[2][1]) * he[1] + std::fabs(R.m[2][2]) * he[2];
}

void rc_karras_node(const uint64_t *codes, int n, int i, int32_t *out) {
    int lo, hi, g;
    karras_node(codes, n, i, lo, hi, g);
    out[0] = lo; out[1] = hi; out[2] = g;
    out[3] = (lo == g) ? (n - 1 + lo) : g;
    out[4] = (g + 1 == hi) ? (n - 1 + g + 1) : (g + 1);
}

int rc_max_threads(void) { return omp_get_max_threads(); }
void rc_set_threads(int n) { omp_set_num_threads(n); }

}  // extern "C"
