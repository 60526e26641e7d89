// dy_math.hpp — per-element arithmetic of the frame step, shared by every
// gfx950 kernel.  Everything here is __host__ __device__ so the identical code
// can also be compiled for the CPU and checked against the oracle without a
// GPU.  Compile with -ffp-contract=off: each float operation below must round
// on its own, in the order written, to stay bit-exact with the reference.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define DY_HD __host__ __device__ __forceinline__

namespace dy {

constexpr uint32_t kBox = 0u, kSphere = 1u;
constexpr uint32_t kFlagStatic = 1u, kFlagSleeping = 4u;
constexpr int32_t kEnd = -1;  // traversal terminator

// sycl::rsqrt of the reference's OpenMP backend: a correctly rounded sqrt then
// a correctly rounded divide (hipcc's default fp32 divide/sqrt are IEEE).
DY_HD float rsqrt_ieee(float x) { return 1.f / sqrtf(x); }

DY_HD int clz64(uint64_t v) { return v ? __builtin_clzll(v) : 64; }

// IEEE binary16 conversions, round to nearest even.
DY_HD uint16_t f32_to_f16(float f) {
    _Float16 h = static_cast<_Float16>(f);
    return __builtin_bit_cast(uint16_t, h);
}
DY_HD float f16_to_f32(uint16_t u) { return static_cast<float>(__builtin_bit_cast(_Float16, u)); }

// ---------------------------------------------------------------------------
// Morton codes (core/src/broadphase.cpp:21-33,128-146)
// ---------------------------------------------------------------------------
DY_HD uint32_t spread3(uint32_t v) {
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}
DY_HD uint32_t morton3(uint32_t x, uint32_t y, uint32_t z) {
    return spread3(x) | (spread3(y) << 1) | (spread3(z) << 2);
}
DY_HD uint32_t quantize10(float p, float lo, float hi) {
    float t = (p - lo) / (hi - lo);
    t = t < 0.f ? 0.f : (t > 1.f ? 1.f : t);
    return static_cast<uint32_t>(t * 1023.f);
}

// ---------------------------------------------------------------------------
// Karras radix tree node (core/src/broadphase.cpp:35-76,164-180)
// ---------------------------------------------------------------------------
DY_HD int lcp(const uint64_t *codes, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    return clz64(codes[i] ^ codes[j]);
}
// Direction of internal node i's range: +1 if it starts at i.
DY_HD int karras_dir(const uint64_t *codes, int n, int i) {
    return (lcp(codes, n, i, i + 1) - lcp(codes, n, i, i - 1)) >= 0 ? 1 : -1;
}
DY_HD void karras_node(const uint64_t *codes, int n, int i, int &lo, int &hi, int &split) {
    const int dir = karras_dir(codes, n, i);
    const int dmin = lcp(codes, n, i, i - dir);
    int lmax = 2;
    while (lcp(codes, n, i, i + lmax * dir) > dmin) lmax *= 2;
    int len = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
        if (lcp(codes, n, i, i + (len + t) * dir) > dmin) len += t;
    const int other = i + len * dir;
    lo = dir > 0 ? i : other;
    hi = dir > 0 ? other : i;
    const int dnode = lcp(codes, n, lo, hi);
    int s = 0, step = hi - lo;
    do {
        step = (step + 1) / 2;
        if (lcp(codes, n, lo, lo + s + step) > dnode) s += step;
    } while (step > 1);
    split = lo + s;
}
// Node visited after the subtree covering leaves [.., hi] in depth-first order:
// the right child of the lowest common ancestor of leaves hi and hi+1, which is
// internal node hi+1 when that node's range starts at hi+1, else leaf hi+1.
DY_HD int32_t karras_escape(const uint64_t *codes, int n, int hi) {
    if (hi >= n - 1) return kEnd;
    const int m = hi + 1;
    if (m <= n - 2 && karras_dir(codes, n, m) > 0) return m;
    return n - 1 + m;
}

// ---------------------------------------------------------------------------
// Rotations
// ---------------------------------------------------------------------------
// Rows of R(q) as the broadphase leaf pass builds them (broadphase.cpp:229-237).
struct M3 { float m[3][3]; };
DY_HD M3 leaf_rotation(float qw, float qx, float qy, float qz) {
    M3 r;
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
// Row a = world direction of body axis a (narrowphase.cpp:913-917; the same
// matrix xpbd_solver.cpp:40-59 builds).
DY_HD M3 body_axes(float qw, float qx, float qy, float qz) {
    M3 e;
    e.m[0][0] = 1.f - 2.f * (qy * qy + qz * qz);
    e.m[0][1] = 2.f * (qx * qy + qz * qw);
    e.m[0][2] = 2.f * (qx * qz - qy * qw);
    e.m[1][0] = 2.f * (qx * qy - qz * qw);
    e.m[1][1] = 1.f - 2.f * (qx * qx + qz * qz);
    e.m[1][2] = 2.f * (qy * qz + qx * qw);
    e.m[2][0] = 2.f * (qx * qz + qy * qw);
    e.m[2][1] = 2.f * (qy * qz - qx * qw);
    e.m[2][2] = 1.f - 2.f * (qx * qx + qy * qy);
    return e;
}

// Leaf half-extents of a body's world AABB (broadphase.cpp:222-244).
DY_HD void leaf_extent(uint32_t stype, float hx, float hy, float hz, float qw, float qx, float qy,
                       float qz, float &ex, float &ey, float &ez) {
    if (stype == kSphere) {
        ex = ey = ez = hx;
        return;
    }
    const M3 R = leaf_rotation(qw, qx, qy, qz);
    ex = fabsf(R.m[0][0]) * hx + fabsf(R.m[0][1]) * hy + fabsf(R.m[0][2]) * hz;
    ey = fabsf(R.m[1][0]) * hx + fabsf(R.m[1][1]) * hy + fabsf(R.m[1][2]) * hz;
    ez = fabsf(R.m[2][0]) * hx + fabsf(R.m[2][1]) * hy + fabsf(R.m[2][2]) * hz;
}

// ---------------------------------------------------------------------------
// Narrowphase primitives (core/src/narrowphase.cpp:46-306)
// ---------------------------------------------------------------------------
// feat: the contact's feature word (dy_contact_feature in dyphur_hip.h: kind,
// axes, incident-vertex mask, SAT axis); the reference's box_box computes
// these and keeps only the contact (narrowphase.cpp:153-305).
struct Hit { float px, py, pz, nx, ny, nz, depth; uint32_t feat; };

constexpr uint32_t kFeatFaceA = 0, kFeatFaceB = 1, kFeatEdge = 2, kFeatSphSph = 3, kFeatSphBoxOut = 4,
                   kFeatSphBoxIn = 5;
DY_HD uint32_t feat_word(uint32_t kind, uint32_t i, uint32_t j, uint32_t mask, uint32_t sat) {
    return kind | (i << 3) | (j << 5) | (mask << 8) | (sat << 16);
}

DY_HD bool sphere_sphere(float ax, float ay, float az, float ra, float bx, float by, float bz,
                         float rb, Hit &h) {
    const float dx = ax - bx, dy = ay - by, dz = az - bz;
    const float d2 = dx * dx + dy * dy + dz * dz;
    const float rs = ra + rb;
    if (d2 >= rs * rs) return false;
    const float d = sqrtf(d2);
    if (d < 1e-8f) { h.nx = 0.f; h.ny = 1.f; h.nz = 0.f; }
    else { const float inv = 1.f / d; h.nx = dx * inv; h.ny = dy * inv; h.nz = dz * inv; }
    h.depth = rs - d;
    h.px = 0.5f * ((ax - h.nx * ra) + (bx + h.nx * rb));
    h.py = 0.5f * ((ay - h.ny * ra) + (by + h.ny * rb));
    h.pz = 0.5f * ((az - h.nz * ra) + (bz + h.nz * rb));
    h.feat = feat_word(kFeatSphSph, 0, 0, 0, 0);
    return true;
}

DY_HD bool sphere_box(float sx, float sy, float sz, float r, float cx, float cy, float cz,
                      const M3 &E, const float he[3], Hit &h) {
    const float dx = sx - cx, dy = sy - cy, dz = sz - cz;
    float q[3], cl[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) q[a] = dx * E.m[a][0] + dy * E.m[a][1] + dz * E.m[a][2];
#pragma unroll
    for (int a = 0; a < 3; ++a) cl[a] = q[a] < -he[a] ? -he[a] : (q[a] > he[a] ? he[a] : q[a]);
    const float lx = q[0] - cl[0], ly = q[1] - cl[1], lz = q[2] - cl[2];
    const float d2 = lx * lx + ly * ly + lz * lz;
    if (d2 < 1e-12f) {
        float pen[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) pen[a] = he[a] - fabsf(q[a]);
        const int f = (pen[0] <= pen[1] && pen[0] <= pen[2]) ? 0 : (pen[1] <= pen[2] ? 1 : 2);
        const float qf = f == 0 ? q[0] : (f == 1 ? q[1] : q[2]);
        const float pf = f == 0 ? pen[0] : (f == 1 ? pen[1] : pen[2]);
        const float sg = qf >= 0.f ? 1.f : -1.f;
        h.nx = -(sg * E.m[f][0]); h.ny = -(sg * E.m[f][1]); h.nz = -(sg * E.m[f][2]);
        h.depth = r + pf;
        h.px = sx - h.nx * r; h.py = sy - h.ny * r; h.pz = sz - h.nz * r;
        h.feat = feat_word(kFeatSphBoxIn, static_cast<uint32_t>(f), 0, 0, 0);
        return true;
    }
    if (d2 >= r * r) return false;
    const float d = sqrtf(d2);
    const float inv = 1.f / d;
    const float wx = lx * E.m[0][0] + ly * E.m[1][0] + lz * E.m[2][0];
    const float wy = lx * E.m[0][1] + ly * E.m[1][1] + lz * E.m[2][1];
    const float wz = lx * E.m[0][2] + ly * E.m[1][2] + lz * E.m[2][2];
    h.nx = -wx * inv; h.ny = -wy * inv; h.nz = -wz * inv;
    h.depth = r - d;
    h.px = cx + cl[0] * E.m[0][0] + cl[1] * E.m[1][0] + cl[2] * E.m[2][0];
    h.py = cy + cl[0] * E.m[0][1] + cl[1] * E.m[1][1] + cl[2] * E.m[2][1];
    h.pz = cz + cl[0] * E.m[0][2] + cl[1] * E.m[1][2] + cl[2] * E.m[2][2];
    h.feat = feat_word(kFeatSphBoxOut, 0, 0, 0, 0);
    return true;
}

DY_HD float box_radius(const M3 &E, const float he[3], float lx, float ly, float lz) {
    return he[0] * fabsf(lx * E.m[0][0] + ly * E.m[0][1] + lz * E.m[0][2])
         + he[1] * fabsf(lx * E.m[1][0] + ly * E.m[1][1] + lz * E.m[1][2])
         + he[2] * fabsf(lx * E.m[2][0] + ly * E.m[2][1] + lz * E.m[2][2]);
}

// SAT state of the box-box test.
struct Sat {
    float best;
    int type, i, j;
    float nx, ny, nz;
};

// One axis test; false when the axis separates the boxes.
DY_HD bool sat_axis(Sat &s, const M3 &A, const float ah[3], const M3 &B, const float bh[3],
                    float dx, float dy, float dz, float lx, float ly, float lz, int type, int i,
                    int j) {
    const float rA = box_radius(A, ah, lx, ly, lz);
    const float rB = box_radius(B, bh, lx, ly, lz);
    const float t = fabsf(dx * lx + dy * ly + dz * lz);
    const float ov = rA + rB - t;
    if (ov <= 0.f) return false;
    if (ov < s.best) {
        s.best = ov; s.type = type; s.i = i; s.j = j;
        const float sg = (dx * lx + dy * ly + dz * lz) >= 0.f ? -1.f : 1.f;
        s.nx = sg * lx; s.ny = sg * ly; s.nz = sg * lz;
    }
    return true;
}

DY_HD float row_comp(const M3 &E, int r, int c) { return E.m[r][c]; }

// Box-box SAT + vertex-face / edge-edge manifold reduced to one contact:
// centroid of the manifold points and their mean depth (narrowphase.cpp:130-306,
// 980-998).  Normal points from B towards A.
DY_HD bool box_box(float acx, float acy, float acz, const M3 &A, const float ah[3], float bcx,
                   float bcy, float bcz, const M3 &B, const float bh[3], Hit &h) {
    const float kPar = 1e-5f;
    const float dx = bcx - acx, dy = bcy - acy, dz = bcz - acz;
    float C[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            C[i][j] = A.m[i][0] * B.m[j][0] + A.m[i][1] * B.m[j][1] + A.m[i][2] * B.m[j][2];
    Sat s{1e30f, -1, 0, 0, 0.f, 1.f, 0.f};
#pragma unroll
    for (int i = 0; i < 3; ++i)
        if (!sat_axis(s, A, ah, B, bh, dx, dy, dz, A.m[i][0], A.m[i][1], A.m[i][2], 0, i, 0)) return false;
#pragma unroll
    for (int j = 0; j < 3; ++j)
        if (!sat_axis(s, A, ah, B, bh, dx, dy, dz, B.m[j][0], B.m[j][1], B.m[j][2], 1, j, 0)) return false;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const float cx = A.m[i][1] * B.m[j][2] - A.m[i][2] * B.m[j][1];
            const float cy = A.m[i][2] * B.m[j][0] - A.m[i][0] * B.m[j][2];
            const float cz = A.m[i][0] * B.m[j][1] - A.m[i][1] * B.m[j][0];
            const float l2 = cx * cx + cy * cy + cz * cz;
            if (l2 < kPar * kPar) continue;
            const float il = rsqrt_ieee(l2);
            if (!sat_axis(s, A, ah, B, bh, dx, dy, dz, cx * il, cy * il, cz * il, 2, i, j)) return false;
        }
    }
    const float nx = s.nx, ny = s.ny, nz = s.nz;
    float sx = 0.f, sy = 0.f, sz = 0.f, sd = 0.f;
    int m = 0;
    const uint32_t sat = static_cast<uint32_t>(s.type == 0 ? s.i : (s.type == 1 ? 3 + s.i : 6 + 3 * s.i + s.j));
    uint32_t feat;
    if (s.type == 0 || s.type == 1) {
        int type = s.type, ref = s.i;
        if (type == 0) {
            const int ta0 = (ref + 1) % 3, ta1 = (ref + 2) % 3;
            int jb = 0;
            for (int j = 1; j < 3; ++j)
                if (fabsf(C[ref][j]) > fabsf(C[ref][jb])) jb = j;
            const int tb0 = (jb + 1) % 3, tb1 = (jb + 2) % 3;
            if (bh[tb0] * bh[tb1] > ah[ta0] * ah[ta1]) { type = 1; ref = jb; }
        }
        const bool a_ref = type == 0;
        const M3 &R = a_ref ? A : B;
        const M3 &I = a_ref ? B : A;
        const float *rh = a_ref ? ah : bh;
        const float *ih = a_ref ? bh : ah;
        const float rcx = a_ref ? acx : bcx, rcy = a_ref ? acy : bcy, rcz = a_ref ? acz : bcz;
        const float icx = a_ref ? bcx : acx, icy = a_ref ? bcy : acy, icz = a_ref ? bcz : acz;
        const int t0 = (ref + 1) % 3, t1 = (ref + 2) % 3;
        const float ref_n = rcx * nx + rcy * ny + rcz * nz;
        const float sdir = a_ref ? 1.f : -1.f;
        const float rh_ref = rh[ref], rh_t0 = rh[t0], rh_t1 = rh[t1];
        float r0[3], r1[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) { r0[c] = R.m[t0][c]; r1[c] = R.m[t1][c]; }
        uint32_t mask = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const float s0 = (k & 1) ? 1.f : -1.f, s1 = (k & 2) ? 1.f : -1.f, s2 = (k & 4) ? 1.f : -1.f;
            const float vx = icx + s0 * ih[0] * I.m[0][0] + s1 * ih[1] * I.m[1][0] + s2 * ih[2] * I.m[2][0];
            const float vy = icy + s0 * ih[0] * I.m[0][1] + s1 * ih[1] * I.m[1][1] + s2 * ih[2] * I.m[2][1];
            const float vz = icz + s0 * ih[0] * I.m[0][2] + s1 * ih[1] * I.m[1][2] + s2 * ih[2] * I.m[2][2];
            const float vn = vx * nx + vy * ny + vz * nz;
            const float dv = rh_ref + sdir * (vn - ref_n);
            if (dv <= 0.f) continue;
            const float ux = vx - rcx, uy = vy - rcy, uz = vz - rcz;
            const float q0 = ux * r0[0] + uy * r0[1] + uz * r0[2];
            const float q1 = ux * r1[0] + uy * r1[1] + uz * r1[2];
            if (fabsf(q0) > rh_t0 + 1e-3f) continue;
            if (fabsf(q1) > rh_t1 + 1e-3f) continue;
            mask |= 1u << k;
            if (m < 4) {
                // Running centroid sums in emission order (narrowphase.cpp:993).
                sx += vx - sdir * dv * nx;
                sy += vy - sdir * dv * ny;
                sz += vz - sdir * dv * nz;
                sd += dv;
                ++m;
            }
        }
        feat = feat_word(a_ref ? kFeatFaceA : kFeatFaceB, static_cast<uint32_t>(ref), 0, mask, sat);
    } else {
        const int i = s.i, j = s.j;
        const int ta0 = (i + 1) % 3, ta1 = (i + 2) % 3;
        const float sa0 = (-nx * A.m[ta0][0] - ny * A.m[ta0][1] - nz * A.m[ta0][2]) >= 0.f ? 1.f : -1.f;
        const float sa1 = (-nx * A.m[ta1][0] - ny * A.m[ta1][1] - nz * A.m[ta1][2]) >= 0.f ? 1.f : -1.f;
        const float pax = acx + sa0 * ah[ta0] * A.m[ta0][0] + sa1 * ah[ta1] * A.m[ta1][0];
        const float pay = acy + sa0 * ah[ta0] * A.m[ta0][1] + sa1 * ah[ta1] * A.m[ta1][1];
        const float paz = acz + sa0 * ah[ta0] * A.m[ta0][2] + sa1 * ah[ta1] * A.m[ta1][2];
        const int tb0 = (j + 1) % 3, tb1 = (j + 2) % 3;
        const float sb0 = (nx * B.m[tb0][0] + ny * B.m[tb0][1] + nz * B.m[tb0][2]) >= 0.f ? 1.f : -1.f;
        const float sb1 = (nx * B.m[tb1][0] + ny * B.m[tb1][1] + nz * B.m[tb1][2]) >= 0.f ? 1.f : -1.f;
        const float pbx = bcx + sb0 * bh[tb0] * B.m[tb0][0] + sb1 * bh[tb1] * B.m[tb1][0];
        const float pby = bcy + sb0 * bh[tb0] * B.m[tb0][1] + sb1 * bh[tb1] * B.m[tb1][1];
        const float pbz = bcz + sb0 * bh[tb0] * B.m[tb0][2] + sb1 * bh[tb1] * B.m[tb1][2];
        const float c = C[i][j];
        const float den = 1.f - c * c;
        const float wx = pax - pbx, wy = pay - pby, wz = paz - pbz;
        const float d1 = wx * A.m[i][0] + wy * A.m[i][1] + wz * A.m[i][2];
        const float d2 = wx * B.m[j][0] + wy * B.m[j][1] + wz * B.m[j][2];
        float u = 0.f, v = 0.f;
        if (fabsf(den) > 1e-6f) {
            u = (c * d2 - d1) / den;
            v = (d2 - c * d1) / den;
        }
        const float qax = pax + u * A.m[i][0], qay = pay + u * A.m[i][1], qaz = paz + u * A.m[i][2];
        const float qbx = pbx + v * B.m[j][0], qby = pby + v * B.m[j][1], qbz = pbz + v * B.m[j][2];
        sx += 0.5f * (qax + qbx);
        sy += 0.5f * (qay + qby);
        sz += 0.5f * (qaz + qbz);
        sd += s.best;
        m = 1;
        feat = feat_word(kFeatEdge, static_cast<uint32_t>(i), static_cast<uint32_t>(j), 0, sat);
    }
    if (m == 0) return false;
    const float inv = 1.f / static_cast<float>(m);
    h.px = sx * inv; h.py = sy * inv; h.pz = sz * inv;
    h.nx = nx; h.ny = ny; h.nz = nz;
    h.depth = sd * inv;
    h.feat = feat;
    return true;
}

// ---------------------------------------------------------------------------
// XPBD contact projection (core/src/xpbd_solver.cpp:40-98, 413-521)
// ---------------------------------------------------------------------------
// The state of one body as the projection reads and writes it.
struct BodyState {
    float px, py, pz;
    float qw, qx, qy, qz;
    float vx, vy, vz;
    float wx, wy, wz;
    float im;
    float ixx, iyy, izz, ixy, ixz, iyz;
};

// M * Ib^-1 * M^T * v with M = body_axes(q), the reference's world inverse
// inertia built from the current, possibly un-normalised quaternion.
DY_HD void inv_inertia_world(const BodyState &b, float vx, float vy, float vz, float &ox,
                             float &oy, float &oz) {
    const M3 M = body_axes(b.qw, b.qx, b.qy, b.qz);
    const float lx = M.m[0][0] * vx + M.m[1][0] * vy + M.m[2][0] * vz;
    const float ly = M.m[0][1] * vx + M.m[1][1] * vy + M.m[2][1] * vz;
    const float lz = M.m[0][2] * vx + M.m[1][2] * vy + M.m[2][2] * vz;
    const float jx = b.ixx * lx + b.ixy * ly + b.ixz * lz;
    const float jy = b.ixy * lx + b.iyy * ly + b.iyz * lz;
    const float jz = b.ixz * lx + b.iyz * ly + b.izz * lz;
    ox = M.m[0][0] * jx + M.m[0][1] * jy + M.m[0][2] * jz;
    oy = M.m[1][0] * jx + M.m[1][1] * jy + M.m[1][2] * jz;
    oz = M.m[2][0] * jx + M.m[2][1] * jy + M.m[2][2] * jz;
}

// q += 0.5 * (0, w) (x) q; bodies with inv_mass == 0 are never written.
DY_HD void rotate_by(BodyState &b, float wx, float wy, float wz) {
    if (b.im == 0.f) return;
    const float qw = b.qw, qx = b.qx, qy = b.qy, qz = b.qz;
    b.qw += 0.5f * (-wx * qx - wy * qy - wz * qz);
    b.qx += 0.5f * (wx * qw + wy * qz - wz * qy);
    b.qy += 0.5f * (-wx * qz + wy * qw + wz * qx);
    b.qz += 0.5f * (wx * qy - wy * qx + wz * qw);
}

DY_HD void renormalise(float &qw, float &qx, float &qy, float &qz) {
    const float s = rsqrt_ieee(qw * qw + qx * qx + qy * qy + qz * qz);
    qw = qw * s; qx = qx * s; qy = qy * s; qz = qz * s;
}

struct ContactIn { float px, py, pz, nx, ny, nz, depth; };

// One contact projection: position, inelastic normal velocity, Coulomb
// friction.  lam is the contact's accumulated depth; returns true when the
// bodies were modified.
// project_contact's own first test: a contact with depth <= 0, or whose depth
// lambda already covers (after its first projection in a substep, lambda ==
// depth), changes nothing.  Callers may skip such a contact's loads and
// stores; the operations are project_contact's, so the decision is its.
DY_HD inline bool contact_live(float depth, float lam) { return !(depth <= 0.f) && !(depth - lam <= 1e-6f); }

DY_HD bool project_contact(const ContactIn &c, BodyState &A, BodyState &B, float mu, float &lam,
                           float dt) {
    if (c.depth <= 0.f) return false;
    const float rem = c.depth - lam;
    if (rem <= 1e-6f) return false;
    const float nx = c.nx, ny = c.ny, nz = c.nz;
    const float wa = A.im, wb = B.im;
    const float rax = c.px - A.px, ray = c.py - A.py, raz = c.pz - A.pz;
    const float rbx = c.px - B.px, rby = c.py - B.py, rbz = c.pz - B.pz;
    const float cax = ray * nz - raz * ny, cay = raz * nx - rax * nz, caz = rax * ny - ray * nx;
    const float cbx = rby * nz - rbz * ny, cby = rbz * nx - rbx * nz, cbz = rbx * ny - rby * nx;
    float kax, kay, kaz, kbx, kby, kbz;
    inv_inertia_world(A, cax, cay, caz, kax, kay, kaz);
    inv_inertia_world(B, cbx, cby, cbz, kbx, kby, kbz);
    const float wa_ang = cax * kax + cay * kay + caz * kaz;
    const float wb_ang = cbx * kbx + cby * kby + cbz * kbz;
    const float w = wa + wb + wa_ang + wb_ang;
    if (w < 1e-10f) return false;
    const float dl = rem / w;
    lam += rem;

    if (wa != 0.f) { A.px += wa * dl * nx; A.py += wa * dl * ny; A.pz += wa * dl * nz; }
    if (wb != 0.f) { B.px -= wb * dl * nx; B.py -= wb * dl * ny; B.pz -= wb * dl * nz; }
    rotate_by(A, dl * kax, dl * kay, dl * kaz);
    rotate_by(B, -dl * kbx, -dl * kby, -dl * kbz);

    {
        const float vax = A.vx + A.wy * raz - A.wz * ray;
        const float vay = A.vy + A.wz * rax - A.wx * raz;
        const float vaz = A.vz + A.wx * ray - A.wy * rax;
        const float vbx = B.vx + B.wy * rbz - B.wz * rby;
        const float vby = B.vy + B.wz * rbx - B.wx * rbz;
        const float vbz = B.vz + B.wx * rby - B.wy * rbx;
        const float rx = vax - vbx, ry = vay - vby, rz = vaz - vbz;
        const float vn = rx * nx + ry * ny + rz * nz;
        if (vn < 0.f) {
            const float jn = -vn / w;
            float gax, gay, gaz, gbx, gby, gbz;
            inv_inertia_world(A, cax, cay, caz, gax, gay, gaz);
            inv_inertia_world(B, cbx, cby, cbz, gbx, gby, gbz);
            if (wa != 0.f) {
                A.vx += wa * jn * nx; A.vy += wa * jn * ny; A.vz += wa * jn * nz;
                A.wx += jn * gax; A.wy += jn * gay; A.wz += jn * gaz;
            }
            if (wb != 0.f) {
                B.vx -= wb * jn * nx; B.vy -= wb * jn * ny; B.vz -= wb * jn * nz;
                B.wx -= jn * gbx; B.wy -= jn * gby; B.wz -= jn * gbz;
            }
        }
    }
    if (mu > 0.f) {
        const float vax = A.vx + A.wy * raz - A.wz * ray;
        const float vay = A.vy + A.wz * rax - A.wx * raz;
        const float vaz = A.vz + A.wx * ray - A.wy * rax;
        const float vbx = B.vx + B.wy * rbz - B.wz * rby;
        const float vby = B.vy + B.wz * rbx - B.wx * rbz;
        const float vbz = B.vz + B.wx * rby - B.wy * rbx;
        const float ux = vax - vbx, uy = vay - vby, uz = vaz - vbz;
        const float un = ux * nx + uy * ny + uz * nz;
        const float tx0 = ux - un * nx, ty0 = uy - un * ny, tz0 = uz - un * nz;
        const float t2 = tx0 * tx0 + ty0 * ty0 + tz0 * tz0;
        if (t2 > 1e-12f) {
            const float tl = sqrtf(t2);
            const float tx = tx0 / tl, ty = ty0 / tl, tz = tz0 / tl;
            const float dax = ray * tz - raz * ty, day = raz * tx - rax * tz, daz = rax * ty - ray * tx;
            const float dbx = rby * tz - rbz * ty, dby = rbz * tx - rbx * tz, dbz = rbx * ty - rby * tx;
            float fax, fay, faz, fbx, fby, fbz;
            inv_inertia_world(A, dax, day, daz, fax, fay, faz);
            inv_inertia_world(B, dbx, dby, dbz, fbx, fby, fbz);
            const float wt = wa + wb + dax * fax + day * fay + daz * faz + dbx * fbx + dby * fby + dbz * fbz;
            if (wt > 1e-10f) {
                const float jcap = lam / (dt * w);
                float jt = tl / wt;
                jt = fminf(jt, mu * jcap);
                if (wa != 0.f) {
                    A.vx -= wa * jt * tx; A.vy -= wa * jt * ty; A.vz -= wa * jt * tz;
                    A.wx -= jt * fax; A.wy -= jt * fay; A.wz -= jt * faz;
                }
                if (wb != 0.f) {
                    B.vx += wb * jt * tx; B.vy += wb * jt * ty; B.vz += wb * jt * tz;
                    B.wx += jt * fbx; B.wy += jt * fby; B.wz += jt * fbz;
                }
            }
        }
    }
    return true;
}

// project_contact by a pair of lanes: the lane of side 0 holds body A, the
// lane of side 1 body B; each computes its own body's terms and updates its
// own body.  pair_xchg(v) returns the partner lane's v (a lane shuffle on the
// device, a rendezvous of two threads in the host check).  The sums
// project_contact forms left to right (w, the relative velocities, wt) are
// formed on both lanes in the same order from exchanged terms, so every value
// is bit-equal to project_contact's.  Every branch around an exchange depends
// on shared values only, so the two lanes always take it together.
template <class Xchg>
DY_HD void project_contact_pair(const ContactIn &c, BodyState &S, uint32_t side, float mu, float &lam, float dt,
                                Xchg &&pair_xchg) {
    if (c.depth <= 0.f) return;
    const float rem = c.depth - lam;
    if (rem <= 1e-6f) return;
    const float nx = c.nx, ny = c.ny, nz = c.nz;
    const float ws = S.im, wo = pair_xchg(ws);
    const float wa = side ? wo : ws, wb = side ? ws : wo;
    const float rx = c.px - S.px, ry = c.py - S.py, rz = c.pz - S.pz;
    const float cx = ry * nz - rz * ny, cy = rz * nx - rx * nz, cz = rx * ny - ry * nx;
    float kx, ky, kz;
    inv_inertia_world(S, cx, cy, cz, kx, ky, kz);
    const float ang_s = cx * kx + cy * ky + cz * kz, ang_o = pair_xchg(ang_s);
    const float wa_ang = side ? ang_o : ang_s, wb_ang = side ? ang_s : ang_o;
    const float w = wa + wb + wa_ang + wb_ang;
    if (w < 1e-10f) return;
    const float dl = rem / w;
    lam += rem;
    if (side == 0u) {
        if (wa != 0.f) { S.px += wa * dl * nx; S.py += wa * dl * ny; S.pz += wa * dl * nz; }
        rotate_by(S, dl * kx, dl * ky, dl * kz);
    } else {
        if (wb != 0.f) { S.px -= wb * dl * nx; S.py -= wb * dl * ny; S.pz -= wb * dl * nz; }
        rotate_by(S, -dl * kx, -dl * ky, -dl * kz);
    }
    {
        const float vsx = S.vx + S.wy * rz - S.wz * ry;
        const float vsy = S.vy + S.wz * rx - S.wx * rz;
        const float vsz = S.vz + S.wx * ry - S.wy * rx;
        const float vox = pair_xchg(vsx), voy = pair_xchg(vsy), voz = pair_xchg(vsz);
        const float vax = side ? vox : vsx, vay = side ? voy : vsy, vaz = side ? voz : vsz;
        const float vbx = side ? vsx : vox, vby = side ? vsy : voy, vbz = side ? vsz : voz;
        const float qx = vax - vbx, qy = vay - vby, qz = vaz - vbz;
        const float vn = qx * nx + qy * ny + qz * nz;
        if (vn < 0.f) {
            const float jn = -vn / w;
            float gx, gy, gz;
            inv_inertia_world(S, cx, cy, cz, gx, gy, gz);
            if (ws != 0.f) {
                if (side == 0u) {
                    S.vx += wa * jn * nx; S.vy += wa * jn * ny; S.vz += wa * jn * nz;
                    S.wx += jn * gx; S.wy += jn * gy; S.wz += jn * gz;
                } else {
                    S.vx -= wb * jn * nx; S.vy -= wb * jn * ny; S.vz -= wb * jn * nz;
                    S.wx -= jn * gx; S.wy -= jn * gy; S.wz -= jn * gz;
                }
            }
        }
    }
    if (mu > 0.f) {
        const float vsx = S.vx + S.wy * rz - S.wz * ry;
        const float vsy = S.vy + S.wz * rx - S.wx * rz;
        const float vsz = S.vz + S.wx * ry - S.wy * rx;
        const float vox = pair_xchg(vsx), voy = pair_xchg(vsy), voz = pair_xchg(vsz);
        const float vax = side ? vox : vsx, vay = side ? voy : vsy, vaz = side ? voz : vsz;
        const float vbx = side ? vsx : vox, vby = side ? vsy : voy, vbz = side ? vsz : voz;
        const float ux = vax - vbx, uy = vay - vby, uz = vaz - vbz;
        const float un = ux * nx + uy * ny + uz * nz;
        const float tx0 = ux - un * nx, ty0 = uy - un * ny, tz0 = uz - un * nz;
        const float t2 = tx0 * tx0 + ty0 * ty0 + tz0 * tz0;
        if (t2 > 1e-12f) {
            const float tl = sqrtf(t2);
            const float tx = tx0 / tl, ty = ty0 / tl, tz = tz0 / tl;
            const float dx = ry * tz - rz * ty, dy = rz * tx - rx * tz, dz = rx * ty - ry * tx;
            float fx, fy, fz;
            inv_inertia_world(S, dx, dy, dz, fx, fy, fz);
            const float p0 = dx * fx, p1 = dy * fy, p2 = dz * fz;
            const float o0 = pair_xchg(p0), o1 = pair_xchg(p1), o2 = pair_xchg(p2);
            const float a0 = side ? o0 : p0, a1 = side ? o1 : p1, a2 = side ? o2 : p2;
            const float b0 = side ? p0 : o0, b1 = side ? p1 : o1, b2 = side ? p2 : o2;
            const float wt = wa + wb + a0 + a1 + a2 + b0 + b1 + b2;
            if (wt > 1e-10f) {
                const float jcap = lam / (dt * w);
                float jt = tl / wt;
                jt = fminf(jt, mu * jcap);
                if (ws != 0.f) {
                    if (side == 0u) {
                        S.vx -= wa * jt * tx; S.vy -= wa * jt * ty; S.vz -= wa * jt * tz;
                        S.wx -= jt * fx; S.wy -= jt * fy; S.wz -= jt * fz;
                    } else {
                        S.vx += wb * jt * tx; S.vy += wb * jt * ty; S.vz += wb * jt * tz;
                        S.wx += jt * fx; S.wy += jt * fy; S.wz += jt * fz;
                    }
                }
            }
        }
    }
}

// Symplectic Euler substeps of one dynamic body with world-frame angular
// velocity and quaternion renormalisation (core/src/integrator.cpp:17-70).
struct Kin { float px, py, pz, qw, qx, qy, qz, vx, vy, vz, wx, wy, wz; };
DY_HD void integrate_body(Kin &k, float gx, float gy, float gz, float h, int n_substeps) {
    const float hh = 0.5f * h;
    for (int s = 0; s < n_substeps; ++s) {
        k.vx = k.vx + gx * h; k.vy = k.vy + gy * h; k.vz = k.vz + gz * h;
        k.px += k.vx * h; k.py += k.vy * h; k.pz += k.vz * h;
        const float ew = -(k.wx * k.qx + k.wy * k.qy + k.wz * k.qz);
        const float ex = (k.wx * k.qw + k.wy * k.qz - k.wz * k.qy);
        const float ey = (k.wy * k.qw + k.wz * k.qx - k.wx * k.qz);
        const float ez = (k.wz * k.qw + k.wx * k.qy - k.wy * k.qx);
        k.qw += hh * ew; k.qx += hh * ex; k.qy += hh * ey; k.qz += hh * ez;
        renormalise(k.qw, k.qx, k.qy, k.qz);
    }
}

// Murmur3 fmix32 (xpbd_solver.cpp:31-34) and the colouring priority word.
DY_HD uint32_t hash_u32(uint32_t x) {
    x ^= x >> 16; x *= 0x85ebca6bu;
    x ^= x >> 13; x *= 0xc2b2ae35u;
    x ^= x >> 16;
    return x;
}
DY_HD uint64_t jp_priority(uint32_t k) { return (static_cast<uint64_t>(hash_u32(k)) << 32) | k; }

}  // namespace dy
