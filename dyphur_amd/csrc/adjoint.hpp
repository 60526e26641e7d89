// adjoint.hpp — derivatives of the frame step's per-element arithmetic, for
// the differentiable form.
//
// Reverse mode, written by hand, for the three maps a frame applies to body
// state: the integrator substep (core/src/integrator.cpp:17-70), the contact
// projection (core/src/xpbd_solver.cpp:413-521) and the quaternion
// renormalisation (xpbd_solver.cpp:578-586,726-731).  Each *_adj function
// recomputes its forward from the saved inputs (same operation order, so the
// same branches as the forward that ran), then propagates output adjoints to
// input adjoints.  Inverse inertia is a constant of the problem and gets no
// adjoint; inverse mass does.
//
// The contact geometry (point, normal, depth of a pair as a function of the
// two poses, core/src/narrowphase.cpp:46-306) is differentiated in forward
// mode: Dual<S, 14> carries the 14 pose tangents (centre xyz, quaternion wxyz
// of both bodies) through a replay of the contact arithmetic along the branch
// the forward took (for box-box: the winning separating axis found by the
// float search, then the manifold of that axis).  The replay functions are
// templates of the forward arithmetic; instantiated at float they give the
// forward's bits (checked by the CPU tests).
//
// Everything is templated on the scalar S / T so the CPU tests can check it
// in double against central differences.
#pragma once

#include "dy_math.hpp"

namespace dy {

// ---------------------------------------------------------------------------
// Scalar helpers that also accept dual numbers
// ---------------------------------------------------------------------------
DY_HD float sq_root(float x) { return sqrtf(x); }
DY_HD double sq_root(double x) { return sqrt(x); }
DY_HD float absval(float x) { return fabsf(x); }
DY_HD double absval(double x) { return fabs(x); }
DY_HD float value(float x) { return x; }
DY_HD double value(double x) { return x; }

template <class S, int N>
struct Dual {
    S v;
    S d[N];
    DY_HD Dual() : v(S(0)) {
        #pragma unroll
        for (int i = 0; i < N; ++i) d[i] = S(0);
    }
    DY_HD Dual(S x) : v(x) {
        #pragma unroll
        for (int i = 0; i < N; ++i) d[i] = S(0);
    }
};

template <class S, int N> DY_HD S value(const Dual<S, N> &a) { return a.v; }
template <class S, int N> DY_HD Dual<S, N> operator+(const Dual<S, N> &a, const Dual<S, N> &b) {
    Dual<S, N> r; r.v = a.v + b.v;
    #pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i];
    return r;
}
template <class S, int N> DY_HD Dual<S, N> operator-(const Dual<S, N> &a, const Dual<S, N> &b) {
    Dual<S, N> r; r.v = a.v - b.v;
    #pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i];
    return r;
}
template <class S, int N> DY_HD Dual<S, N> operator-(const Dual<S, N> &a) {
    Dual<S, N> r; r.v = -a.v;
    #pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = -a.d[i];
    return r;
}
template <class S, int N> DY_HD Dual<S, N> operator*(const Dual<S, N> &a, const Dual<S, N> &b) {
    Dual<S, N> r; r.v = a.v * b.v;
    #pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
    return r;
}
template <class S, int N> DY_HD Dual<S, N> operator/(const Dual<S, N> &a, const Dual<S, N> &b) {
    Dual<S, N> r; r.v = a.v / b.v;
    #pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) / b.v;
    return r;
}
#define DY_DUAL_MIXED(op)                                                                              \
    template <class S, int N> DY_HD Dual<S, N> operator op(const Dual<S, N> &a, S b) { return a op Dual<S, N>(b); } \
    template <class S, int N> DY_HD Dual<S, N> operator op(S a, const Dual<S, N> &b) { return Dual<S, N>(a) op b; }
DY_DUAL_MIXED(+)
DY_DUAL_MIXED(-)
DY_DUAL_MIXED(*)
DY_DUAL_MIXED(/)
#undef DY_DUAL_MIXED
template <class S, int N> DY_HD Dual<S, N> &operator+=(Dual<S, N> &a, const Dual<S, N> &b) { return a = a + b; }
template <class S, int N> DY_HD Dual<S, N> &operator-=(Dual<S, N> &a, const Dual<S, N> &b) { return a = a - b; }
template <class S, int N> DY_HD bool operator<(const Dual<S, N> &a, const Dual<S, N> &b) { return a.v < b.v; }
template <class S, int N> DY_HD bool operator>(const Dual<S, N> &a, const Dual<S, N> &b) { return a.v > b.v; }
template <class S, int N> DY_HD bool operator<=(const Dual<S, N> &a, const Dual<S, N> &b) { return a.v <= b.v; }
template <class S, int N> DY_HD bool operator>=(const Dual<S, N> &a, const Dual<S, N> &b) { return a.v >= b.v; }
template <class S, int N> DY_HD bool operator<(const Dual<S, N> &a, S b) { return a.v < b; }
template <class S, int N> DY_HD bool operator>(const Dual<S, N> &a, S b) { return a.v > b; }
template <class S, int N> DY_HD bool operator<=(const Dual<S, N> &a, S b) { return a.v <= b; }
template <class S, int N> DY_HD bool operator>=(const Dual<S, N> &a, S b) { return a.v >= b; }
template <class S, int N> DY_HD bool operator!=(const Dual<S, N> &a, S b) { return a.v != b; }
template <class S, int N> DY_HD bool operator==(const Dual<S, N> &a, S b) { return a.v == b; }
template <class S, int N> DY_HD bool operator!=(const Dual<S, N> &a, const Dual<S, N> &b) { return a.v != b.v; }
template <class S, int N> DY_HD Dual<S, N> sq_root(const Dual<S, N> &a) {
    Dual<S, N> r; r.v = sq_root(a.v);
    #pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] / (S(2) * r.v);
    return r;
}
template <class S, int N> DY_HD Dual<S, N> absval(const Dual<S, N> &a) { return a.v >= S(0) ? a : -a; }

// ---------------------------------------------------------------------------
// Templated replay of the contact arithmetic (same operation order as
// dy_math.hpp's float functions)
// ---------------------------------------------------------------------------
template <class T> struct M3T { T m[3][3]; };
template <class T> struct HitT { T px, py, pz, nx, ny, nz, depth; uint32_t feat; };

// Element i of a 3-array and row i of a matrix for a run-time i, by selects:
// no run-time indexing, so everything stays in registers.
template <class T> DY_HD T pick3(const T a[3], int i) { return i == 0 ? a[0] : (i == 1 ? a[1] : a[2]); }
template <class T> DY_HD T pickm(const M3T<T> &M, int r, int c) {
    return r == 0 ? M.m[0][c] : (r == 1 ? M.m[1][c] : M.m[2][c]);
}
template <class T> DY_HD M3T<T> pick_mat(bool first, const M3T<T> &A, const M3T<T> &B) {
    M3T<T> R;
    #pragma unroll
    for (int i = 0; i < 3; ++i)
        #pragma unroll
        for (int j = 0; j < 3; ++j) R.m[i][j] = first ? A.m[i][j] : B.m[i][j];
    return R;
}

template <class T, class S>
DY_HD M3T<T> axes_t(const T &qw, const T &qx, const T &qy, const T &qz, S one) {
    const S two = one + one;
    M3T<T> e;
    e.m[0][0] = one - two * (qy * qy + qz * qz);
    e.m[0][1] = two * (qx * qy + qz * qw);
    e.m[0][2] = two * (qx * qz - qy * qw);
    e.m[1][0] = two * (qx * qy - qz * qw);
    e.m[1][1] = one - two * (qx * qx + qz * qz);
    e.m[1][2] = two * (qy * qz + qx * qw);
    e.m[2][0] = two * (qx * qz + qy * qw);
    e.m[2][1] = two * (qy * qz - qx * qw);
    e.m[2][2] = one - two * (qx * qx + qy * qy);
    return e;
}

template <class T, class S>
DY_HD bool sphere_sphere_t(const T &ax, const T &ay, const T &az, S ra, const T &bx, const T &by, const T &bz,
                           S rb, HitT<T> &h) {
    const T dx = ax - bx, dy = ay - by, dz = az - bz;
    const T d2 = dx * dx + dy * dy + dz * dz;
    const S rs = ra + rb;
    if (d2 >= rs * rs) return false;
    const T d = sq_root(d2);
    if (d < S(1e-8f)) { h.nx = T(S(0)); h.ny = T(S(1)); h.nz = T(S(0)); }
    else { const T inv = S(1) / d; h.nx = dx * inv; h.ny = dy * inv; h.nz = dz * inv; }
    h.depth = rs - d;
    h.px = S(0.5f) * ((ax - h.nx * ra) + (bx + h.nx * rb));
    h.py = S(0.5f) * ((ay - h.ny * ra) + (by + h.ny * rb));
    h.pz = S(0.5f) * ((az - h.nz * ra) + (bz + h.nz * rb));
    h.feat = feat_word(kFeatSphSph, 0, 0, 0, 0);
    return true;
}

template <class T, class S>
DY_HD bool sphere_box_t(const T &sx, const T &sy, const T &sz, S r, const T &cx, const T &cy, const T &cz,
                        const M3T<T> &E, const S he[3], HitT<T> &h) {
    const T dx = sx - cx, dy = sy - cy, dz = sz - cz;
    T q[3], cl[3];
    #pragma unroll
    for (int a = 0; a < 3; ++a) q[a] = dx * E.m[a][0] + dy * E.m[a][1] + dz * E.m[a][2];
    #pragma unroll
    for (int a = 0; a < 3; ++a) cl[a] = q[a] < -he[a] ? T(-he[a]) : (q[a] > he[a] ? T(he[a]) : q[a]);
    const T lx = q[0] - cl[0], ly = q[1] - cl[1], lz = q[2] - cl[2];
    const T d2 = lx * lx + ly * ly + lz * lz;
    if (d2 < S(1e-12f)) {
        T pen[3];
        #pragma unroll
        for (int a = 0; a < 3; ++a) pen[a] = he[a] - absval(q[a]);
        const int f = (pen[0] <= pen[1] && pen[0] <= pen[2]) ? 0 : (pen[1] <= pen[2] ? 1 : 2);
        const T qf = pick3(q, f);
        const T pf = pick3(pen, f);
        const S sg = qf >= S(0) ? S(1) : S(-1);
        h.nx = -(sg * pickm(E, f, 0)); h.ny = -(sg * pickm(E, f, 1)); h.nz = -(sg * pickm(E, f, 2));
        h.depth = r + pf;
        h.px = sx - h.nx * r; h.py = sy - h.ny * r; h.pz = sz - h.nz * r;
        h.feat = feat_word(kFeatSphBoxIn, static_cast<uint32_t>(f), 0, 0, 0);
        return true;
    }
    if (d2 >= r * r) return false;
    const T d = sq_root(d2);
    const T inv = S(1) / d;
    const T wx = lx * E.m[0][0] + ly * E.m[1][0] + lz * E.m[2][0];
    const T wy = lx * E.m[0][1] + ly * E.m[1][1] + lz * E.m[2][1];
    const T wz = lx * E.m[0][2] + ly * E.m[1][2] + lz * E.m[2][2];
    h.nx = -wx * inv; h.ny = -wy * inv; h.nz = -wz * inv;
    h.depth = r - d;
    h.px = cx + cl[0] * E.m[0][0] + cl[1] * E.m[1][0] + cl[2] * E.m[2][0];
    h.py = cy + cl[0] * E.m[0][1] + cl[1] * E.m[1][1] + cl[2] * E.m[2][1];
    h.pz = cz + cl[0] * E.m[0][2] + cl[1] * E.m[1][2] + cl[2] * E.m[2][2];
    h.feat = feat_word(kFeatSphBoxOut, 0, 0, 0, 0);
    return true;
}

template <class T, class S>
DY_HD T box_radius_t(const M3T<T> &E, const S he[3], const T &lx, const T &ly, const T &lz) {
    return he[0] * absval(lx * E.m[0][0] + ly * E.m[0][1] + lz * E.m[0][2])
         + he[1] * absval(lx * E.m[1][0] + ly * E.m[1][1] + lz * E.m[1][2])
         + he[2] * absval(lx * E.m[2][0] + ly * E.m[2][1] + lz * E.m[2][2]);
}

template <class T> struct SatT {
    T best;
    int type, i, j;
    T nx, ny, nz;
};

template <class T, class S>
DY_HD bool sat_axis_t(SatT<T> &s, const M3T<T> &A, const S ah[3], const M3T<T> &B, const S bh[3], const T &dx,
                      const T &dy, const T &dz, const T &lx, const T &ly, const T &lz, int type, int i, int j) {
    const T rA = box_radius_t(A, ah, lx, ly, lz);
    const T rB = box_radius_t(B, bh, lx, ly, lz);
    const T t = absval(dx * lx + dy * ly + dz * lz);
    const T ov = rA + rB - t;
    if (ov <= S(0)) return false;
    if (ov < s.best) {
        s.best = ov; s.type = type; s.i = i; s.j = j;
        const S sg = (dx * lx + dy * ly + dz * lz) >= S(0) ? S(-1) : S(1);
        s.nx = sg * lx; s.ny = sg * ly; s.nz = sg * lz;
    }
    return true;
}

// Edge-edge axis A_i x B_j of the separating-axis test (skipped, as a pass,
// when the edges are parallel).
template <class T, class S>
DY_HD bool edge_axis_t(SatT<T> &s, const M3T<T> &A, const S ah[3], const M3T<T> &B, const S bh[3], const T &dx,
                       const T &dy, const T &dz, int i, int j, S kPar) {
    const T a0 = pickm(A, i, 0), a1 = pickm(A, i, 1), a2 = pickm(A, i, 2);
    const T b0 = pickm(B, j, 0), b1 = pickm(B, j, 1), b2 = pickm(B, j, 2);
    const T cx = a1 * b2 - a2 * b1;
    const T cy = a2 * b0 - a0 * b2;
    const T cz = a0 * b1 - a1 * b0;
    const T l2 = cx * cx + cy * cy + cz * cz;
    if (l2 < kPar * kPar) return true;
    const T il = S(1) / sq_root(l2);
    return sat_axis_t(s, A, ah, B, bh, dx, dy, dz, cx * il, cy * il, cz * il, 2, i, j);
}

// Box-box contact (dy_math.hpp box_box).  axis < 0: full separating-axis
// search; otherwise only the axis type*16 + i*4 + j is evaluated (the one the
// search picked).  *winner (may be null) receives the axis code.
template <class T, class S>
DY_HD bool box_box_t(const T &acx, const T &acy, const T &acz, const M3T<T> &A, const S ah[3], const T &bcx,
                     const T &bcy, const T &bcz, const M3T<T> &B, const S bh[3], HitT<T> &h, int axis,
                     int *winner) {
    const S kPar = S(1e-5f);
    const T dx = bcx - acx, dy = bcy - acy, dz = bcz - acz;
    T C[3][3];
    #pragma unroll
    for (int i = 0; i < 3; ++i)
        #pragma unroll
        for (int j = 0; j < 3; ++j)
            C[i][j] = A.m[i][0] * B.m[j][0] + A.m[i][1] * B.m[j][1] + A.m[i][2] * B.m[j][2];
    SatT<T> s{T(S(1e30f)), -1, 0, 0, T(S(0)), T(S(1)), T(S(0))};
    if (axis < 0) {
        #pragma unroll
        for (int i = 0; i < 3; ++i)
            if (!sat_axis_t(s, A, ah, B, bh, dx, dy, dz, A.m[i][0], A.m[i][1], A.m[i][2], 0, i, 0)) return false;
        #pragma unroll
        for (int j = 0; j < 3; ++j)
            if (!sat_axis_t(s, A, ah, B, bh, dx, dy, dz, B.m[j][0], B.m[j][1], B.m[j][2], 1, j, 0)) return false;
        #pragma unroll
        for (int i = 0; i < 3; ++i)
            #pragma unroll
            for (int j = 0; j < 3; ++j)
                if (!edge_axis_t(s, A, ah, B, bh, dx, dy, dz, i, j, kPar)) return false;
    } else {
        const int ty = axis >> 4, ai = (axis >> 2) & 3, aj = axis & 3;
        bool ok = true;
        if (ty == 0)
            ok = sat_axis_t(s, A, ah, B, bh, dx, dy, dz, pickm(A, ai, 0), pickm(A, ai, 1), pickm(A, ai, 2), 0, ai, 0);
        else if (ty == 1)
            ok = sat_axis_t(s, A, ah, B, bh, dx, dy, dz, pickm(B, ai, 0), pickm(B, ai, 1), pickm(B, ai, 2), 1, ai, 0);
        else ok = edge_axis_t(s, A, ah, B, bh, dx, dy, dz, ai, aj, kPar);
        if (!ok || s.type < 0) return false;
    }
    if (winner) *winner = s.type * 16 + s.i * 4 + s.j;
    const uint32_t sat = static_cast<uint32_t>(s.type == 0 ? s.i : (s.type == 1 ? 3 + s.i : 6 + 3 * s.i + s.j));
    uint32_t feat;
    const T nx = s.nx, ny = s.ny, nz = s.nz;
    T sx = T(S(0)), sy = T(S(0)), sz = T(S(0)), sd = T(S(0));
    int m = 0;
    if (s.type == 0 || s.type == 1) {
        int type = s.type, ref = s.i;
        if (type == 0) {
            const int ta0 = (ref + 1) % 3, ta1 = (ref + 2) % 3;
            T cr[3];
            #pragma unroll
            for (int j = 0; j < 3; ++j) cr[j] = ref == 0 ? C[0][j] : (ref == 1 ? C[1][j] : C[2][j]);
            int jb = 0;
            if (absval(cr[1]) > absval(cr[0])) jb = 1;
            if (absval(cr[2]) > absval(pick3(cr, jb))) jb = 2;
            const int tb0 = (jb + 1) % 3, tb1 = (jb + 2) % 3;
            if (pick3(bh, tb0) * pick3(bh, tb1) > pick3(ah, ta0) * pick3(ah, ta1)) { type = 1; ref = jb; }
        }
        const bool a_ref = type == 0;
        const M3T<T> R = pick_mat(a_ref, A, B);
        const M3T<T> I = pick_mat(a_ref, B, A);
        S rh[3], ih[3];
        #pragma unroll
        for (int c = 0; c < 3; ++c) { rh[c] = a_ref ? ah[c] : bh[c]; ih[c] = a_ref ? bh[c] : ah[c]; }
        const T rcx = a_ref ? acx : bcx, rcy = a_ref ? acy : bcy, rcz = a_ref ? acz : bcz;
        const T icx = a_ref ? bcx : acx, icy = a_ref ? bcy : acy, icz = a_ref ? bcz : acz;
        const int t0 = (ref + 1) % 3, t1 = (ref + 2) % 3;
        const T ref_n = rcx * nx + rcy * ny + rcz * nz;
        const S sdir = a_ref ? S(1) : S(-1);
        const S rh_ref = pick3(rh, ref), rh_t0 = pick3(rh, t0), rh_t1 = pick3(rh, t1);
        T r0[3], r1[3];
        #pragma unroll
        for (int c = 0; c < 3; ++c) { r0[c] = pickm(R, t0, c); r1[c] = pickm(R, t1, c); }
        uint32_t mask = 0;
        #pragma unroll
        for (int k = 0; k < 8; ++k) {
            const S s0 = (k & 1) ? S(1) : S(-1), s1 = (k & 2) ? S(1) : S(-1), s2 = (k & 4) ? S(1) : S(-1);
            const T vx = icx + s0 * ih[0] * I.m[0][0] + s1 * ih[1] * I.m[1][0] + s2 * ih[2] * I.m[2][0];
            const T vy = icy + s0 * ih[0] * I.m[0][1] + s1 * ih[1] * I.m[1][1] + s2 * ih[2] * I.m[2][1];
            const T vz = icz + s0 * ih[0] * I.m[0][2] + s1 * ih[1] * I.m[1][2] + s2 * ih[2] * I.m[2][2];
            const T vn = vx * nx + vy * ny + vz * nz;
            const T dv = rh_ref + sdir * (vn - ref_n);
            if (dv <= S(0)) continue;
            const T ux = vx - rcx, uy = vy - rcy, uz = vz - rcz;
            const T q0 = ux * r0[0] + uy * r0[1] + uz * r0[2];
            const T q1 = ux * r1[0] + uy * r1[1] + uz * r1[2];
            if (absval(q0) > rh_t0 + S(1e-3f)) continue;
            if (absval(q1) > rh_t1 + S(1e-3f)) continue;
            mask |= 1u << k;
            if (m < 4) {
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
        T A0[3], A1[3], Ai[3], B0[3], B1[3], Bj[3];
        #pragma unroll
        for (int k = 0; k < 3; ++k) {
            A0[k] = pickm(A, ta0, k); A1[k] = pickm(A, ta1, k); Ai[k] = pickm(A, i, k);
        }
        const S sa0 = (-nx * A0[0] - ny * A0[1] - nz * A0[2]) >= S(0) ? S(1) : S(-1);
        const S sa1 = (-nx * A1[0] - ny * A1[1] - nz * A1[2]) >= S(0) ? S(1) : S(-1);
        const S ah0 = pick3(ah, ta0), ah1 = pick3(ah, ta1);
        const T pax = acx + sa0 * ah0 * A0[0] + sa1 * ah1 * A1[0];
        const T pay = acy + sa0 * ah0 * A0[1] + sa1 * ah1 * A1[1];
        const T paz = acz + sa0 * ah0 * A0[2] + sa1 * ah1 * A1[2];
        const int tb0 = (j + 1) % 3, tb1 = (j + 2) % 3;
        #pragma unroll
        for (int k = 0; k < 3; ++k) {
            B0[k] = pickm(B, tb0, k); B1[k] = pickm(B, tb1, k); Bj[k] = pickm(B, j, k);
        }
        const S sb0 = (nx * B0[0] + ny * B0[1] + nz * B0[2]) >= S(0) ? S(1) : S(-1);
        const S sb1 = (nx * B1[0] + ny * B1[1] + nz * B1[2]) >= S(0) ? S(1) : S(-1);
        const S bh0 = pick3(bh, tb0), bh1 = pick3(bh, tb1);
        const T pbx = bcx + sb0 * bh0 * B0[0] + sb1 * bh1 * B1[0];
        const T pby = bcy + sb0 * bh0 * B0[1] + sb1 * bh1 * B1[1];
        const T pbz = bcz + sb0 * bh0 * B0[2] + sb1 * bh1 * B1[2];
        const T c = Ai[0] * Bj[0] + Ai[1] * Bj[1] + Ai[2] * Bj[2];
        const T den = S(1) - c * c;
        const T wx = pax - pbx, wy = pay - pby, wz = paz - pbz;
        const T d1 = wx * Ai[0] + wy * Ai[1] + wz * Ai[2];
        const T d2 = wx * Bj[0] + wy * Bj[1] + wz * Bj[2];
        T u = T(S(0)), v = T(S(0));
        if (absval(den) > S(1e-6f)) {
            u = (c * d2 - d1) / den;
            v = (d2 - c * d1) / den;
        }
        const T qax = pax + u * Ai[0], qay = pay + u * Ai[1], qaz = paz + u * Ai[2];
        const T qbx = pbx + v * Bj[0], qby = pby + v * Bj[1], qbz = pbz + v * Bj[2];
        sx += S(0.5f) * (qax + qbx);
        sy += S(0.5f) * (qay + qby);
        sz += S(0.5f) * (qaz + qbz);
        sd += s.best;
        m = 1;
        feat = feat_word(kFeatEdge, static_cast<uint32_t>(i), static_cast<uint32_t>(j), 0, sat);
    }
    if (m == 0) return false;
    const S inv = S(1) / static_cast<S>(m);
    h.px = sx * inv; h.py = sy * inv; h.pz = sz * inv;
    h.nx = nx; h.ny = ny; h.nz = nz;
    h.depth = sd * inv;
    h.feat = feat;
    return true;
}

// One pair's contact from the two poses (pair_contact dispatch: sphere-box
// with the sphere as A negates the normal, narrowphase.cpp:963).
template <class T, class S>
DY_HD bool pair_contact_t(uint32_t ta, uint32_t tb, const S ha[3], const S hb[3], const T pa[3], const T qa[4],
                          const T pb[3], const T qb[4], HitT<T> &h, int axis, int *winner) {
    const S one = S(1);
    if (ta == kSphere && tb == kSphere) return sphere_sphere_t(pa[0], pa[1], pa[2], ha[0], pb[0], pb[1], pb[2], hb[0], h);
    if (ta == kSphere) {
        const M3T<T> E = axes_t(qb[0], qb[1], qb[2], qb[3], one);
        if (!sphere_box_t(pa[0], pa[1], pa[2], ha[0], pb[0], pb[1], pb[2], E, hb, h)) return false;
        h.nx = -h.nx; h.ny = -h.ny; h.nz = -h.nz;
        return true;
    }
    if (tb == kSphere) {
        const M3T<T> E = axes_t(qa[0], qa[1], qa[2], qa[3], one);
        return sphere_box_t(pb[0], pb[1], pb[2], hb[0], pa[0], pa[1], pa[2], E, ha, h);
    }
    const M3T<T> A = axes_t(qa[0], qa[1], qa[2], qa[3], one);
    const M3T<T> B = axes_t(qb[0], qb[1], qb[2], qb[3], one);
    return box_box_t(pa[0], pa[1], pa[2], A, ha, pb[0], pb[1], pb[2], B, hb, h, axis, winner);
}

// Adjoint of one contact's geometry: given the adjoints of (p xyz, n xyz,
// depth), accumulate the adjoints of (pa xyz, qa wxyz) into ga[7] and of
// (pb, qb) into gb[7].  The plain search picks the branch; replays of that
// branch with NT pose tangents at a time (14 in all) give the Jacobian.
// Returns false when the pair does not touch.
template <class S, int NT>
DY_HD bool contact_adj(uint32_t ta, uint32_t tb, const S ha[3], const S hb[3], const S pa[3], const S qa[4],
                       const S pb[3], const S qb[4], const S hbar[7], S ga[7], S gb[7]) {
    typedef Dual<S, NT> D;
    int axis = -1;
    {
        HitT<S> h;
        if (!pair_contact_t(ta, tb, ha, hb, pa, qa, pb, qb, h, -1, &axis)) return false;
    }
    for (int k0 = 0; k0 < 14; k0 += NT) {
        D dpa[3], dqa[4], dpb[3], dqb[4];
        #pragma unroll
        for (int k = 0; k < 3; ++k) dpa[k] = D(pa[k]);
        #pragma unroll
        for (int k = 0; k < 4; ++k) dqa[k] = D(qa[k]);
        #pragma unroll
        for (int k = 0; k < 3; ++k) dpb[k] = D(pb[k]);
        #pragma unroll
        for (int k = 0; k < 4; ++k) dqb[k] = D(qb[k]);
        #pragma unroll
        for (int t = 0; t < NT; ++t) {
            #pragma unroll
            for (int k = 0; k < 3; ++k) {
                dpa[k].d[t] = k == k0 + t ? S(1) : S(0);
                dpb[k].d[t] = k + 7 == k0 + t ? S(1) : S(0);
            }
            #pragma unroll
            for (int k = 0; k < 4; ++k) {
                dqa[k].d[t] = k + 3 == k0 + t ? S(1) : S(0);
                dqb[k].d[t] = k + 10 == k0 + t ? S(1) : S(0);
            }
        }
        HitT<D> h;
        if (!pair_contact_t(ta, tb, ha, hb, dpa, dqa, dpb, dqb, h, axis, nullptr)) return false;
        #pragma unroll
        for (int t = 0; t < NT; ++t) {
            const S acc = hbar[0] * h.px.d[t] + hbar[1] * h.py.d[t] + hbar[2] * h.pz.d[t] + hbar[3] * h.nx.d[t] +
                          hbar[4] * h.ny.d[t] + hbar[5] * h.nz.d[t] + hbar[6] * h.depth.d[t];
            #pragma unroll
            for (int k = 0; k < 7; ++k) {
                ga[k] += k == k0 + t ? acc : S(0);
                gb[k] += k + 7 == k0 + t ? acc : S(0);
            }
        }
    }
    return true;
}

// ---------------------------------------------------------------------------
// Contact geometry in reverse mode, written out per contact kind
// ---------------------------------------------------------------------------
// The forward's feature word (dy_math.hpp feat_word: kind, face or edge
// indices, passing incident vertices, winning separating axis) fixes every
// branch the forward took, so the adjoints below search nothing: they
// recompute only the quantities their reverse needs, with the forward's
// arithmetic, then propagate the adjoints of (p xyz, n xyz, depth) to the two
// poses.  Sign choices (clamps, |x|, which side of a face) are piecewise
// constant and take the forward's branch; |x| at x == 0 takes +1, as the
// dual numbers' absval does.

// Sphere-sphere (dy_math.hpp sphere_sphere): n = d / |d| with d = a - b
// (fixed (0, 1, 0) when |d| < 1e-8), depth = ra + rb - |d|,
// p = (a + b) / 2 + (rb - ra) n / 2.
template <class S>
DY_HD void sphere_sphere_adj(S ra, S rb, const S pa[3], const S pb[3], const S hbar[7], S ga[7], S gb[7]) {
    const S d[3] = {pa[0] - pb[0], pa[1] - pb[1], pa[2] - pb[2]};
    const S len = sq_root(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    S db[3] = {S(0), S(0), S(0)};
    if (len > S(0)) {
        const S inv = S(1) / len;
        const S n[3] = {d[0] * inv, d[1] * inv, d[2] * inv};
        const S half_dr = S(0.5f) * (rb - ra);
        const S nb[3] = {hbar[3] + half_dr * hbar[0], hbar[4] + half_dr * hbar[1], hbar[5] + half_dr * hbar[2]};
        const S nn = n[0] * nb[0] + n[1] * nb[1] + n[2] * nb[2];
        const bool unit = !(len < S(1e-8f));
#pragma unroll
        for (int c = 0; c < 3; ++c) db[c] = -hbar[6] * n[c] + (unit ? (nb[c] - n[c] * nn) * inv : S(0));
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        ga[c] += S(0.5f) * hbar[c] + db[c];
        gb[c] += S(0.5f) * hbar[c] - db[c];
    }
}

// Sphere (centre sc, radius r) against a box (centre bc, rotation qb, half
// extents he), the normal pointing from the box to the sphere as
// dy_math.hpp sphere_box returns it.  Outside: q = E (sc - bc), l = q -
// clamp(q), n = -E^T l / |l|, depth = r - |l|, p = bc + E^T clamp(q).
// Inside (face f of least penetration): n = -sg E_f, depth = r + he_f - |q_f|,
// p = sc - r n.  gs[3] receives the sphere centre's adjoint, gb[7] the box's
// centre and quaternion adjoints.
template <class S>
DY_HD void sphere_box_adj(uint32_t feat, S r, const S sc[3], const S bc[3], const S qb[4], const S he[3],
                          const S hbar[7], S gs[3], S gb[7]) {
    const M3T<S> E = axes_t(qb[0], qb[1], qb[2], qb[3], S(1));
    const S d[3] = {sc[0] - bc[0], sc[1] - bc[1], sc[2] - bc[2]};
    S q[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) q[a] = d[0] * E.m[a][0] + d[1] * E.m[a][1] + d[2] * E.m[a][2];
    M3T<S> Eb{};
    S qbar[3] = {S(0), S(0), S(0)};
    S cb[3] = {hbar[0], hbar[1], hbar[2]};  // outside: p = bc + E^T clamp(q)
    if ((feat & 7u) == kFeatSphBoxIn) {
        const int f = static_cast<int>((feat >> 3) & 3u);
        const S qf = pick3(q, f);
        const S sg = qf >= S(0) ? S(1) : S(-1);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            gs[c] += hbar[c];
            cb[c] = S(0);
            const S nb = hbar[3 + c] - r * hbar[c];  // p = sc - r n
#pragma unroll
            for (int a = 0; a < 3; ++a) Eb.m[a][c] += a == f ? -sg * nb : S(0);
        }
        const S qfb = -hbar[6] * sg;
        qbar[0] = f == 0 ? qfb : S(0);
        qbar[1] = f == 1 ? qfb : S(0);
        qbar[2] = f == 2 ? qfb : S(0);
    } else {
        S cl[3], l[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            cl[a] = q[a] < -he[a] ? -he[a] : (q[a] > he[a] ? he[a] : q[a]);
            l[a] = q[a] - cl[a];
        }
        const S len = sq_root(l[0] * l[0] + l[1] * l[1] + l[2] * l[2]);
        const S inv = S(1) / len;
        S w[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) w[c] = l[0] * E.m[0][c] + l[1] * E.m[1][c] + l[2] * E.m[2][c];
        // n = -w / |l|, depth = r - |l|
        const S nw = hbar[3] * w[0] + hbar[4] * w[1] + hbar[5] * w[2];
        const S lenb = nw * inv * inv - hbar[6];
        S wb[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) wb[c] = -hbar[3 + c] * inv;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const S clb = hbar[0] * E.m[a][0] + hbar[1] * E.m[a][1] + hbar[2] * E.m[a][2];
            const S lb = l[a] * lenb * inv + wb[0] * E.m[a][0] + wb[1] * E.m[a][1] + wb[2] * E.m[a][2];
            const bool free_axis = !(q[a] < -he[a]) && !(q[a] > he[a]);
            qbar[a] = lb + (free_axis ? clb - lb : S(0));
#pragma unroll
            for (int c = 0; c < 3; ++c) Eb.m[a][c] += cl[a] * hbar[c] + l[a] * wb[c];
        }
    }
    // q_a = (sc - bc) . E_a
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            gs[c] += qbar[a] * E.m[a][c];
            cb[c] -= qbar[a] * E.m[a][c];
            Eb.m[a][c] += qbar[a] * d[c];
        }
#pragma unroll
    for (int c = 0; c < 3; ++c) gb[c] += cb[c];
    S qbb[4] = {S(0), S(0), S(0), S(0)};
    axes_adj(qb, Eb, qbb);
#pragma unroll
    for (int k = 0; k < 4; ++k) gb[3 + k] += qbb[k];
}

// Box-box edge contact on axis l = A_i x B_j / |A_i x B_j| (dy_math.hpp
// box_box, edge branch): n = sg l, depth = sum_m ah_m |A_m.l| + sum_m bh_m
// |B_m.l| - |(cb - ca).l| (the m = i and m = j terms are identically zero
// and have no derivative), p = the midpoint of the closest points of the two
// support edges.
template <class S>
DY_HD void box_edge_adj(uint32_t feat, const S ha[3], const S hb[3], const S pa[3], const S qa[4], const S pb[3],
                        const S qb[4], const S hbar[7], S ga[7], S gb[7]) {
    const int i = static_cast<int>((feat >> 3) & 3u), j = static_cast<int>((feat >> 5) & 3u);
    const M3T<S> A = axes_t(qa[0], qa[1], qa[2], qa[3], S(1));
    const M3T<S> B = axes_t(qb[0], qb[1], qb[2], qb[3], S(1));
    const int ta0 = (i + 1) % 3, ta1 = (i + 2) % 3, tb0 = (j + 1) % 3, tb1 = (j + 2) % 3;
    S Ai[3], Bj[3], A0[3], A1[3], B0[3], B1[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        Ai[c] = pickm(A, i, c); A0[c] = pickm(A, ta0, c); A1[c] = pickm(A, ta1, c);
        Bj[c] = pickm(B, j, c); B0[c] = pickm(B, tb0, c); B1[c] = pickm(B, tb1, c);
    }
    // forward: the axis, its sign and the support edges
    const S cr[3] = {Ai[1] * Bj[2] - Ai[2] * Bj[1], Ai[2] * Bj[0] - Ai[0] * Bj[2], Ai[0] * Bj[1] - Ai[1] * Bj[0]};
    const S il = S(1) / sq_root(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
    const S l[3] = {cr[0] * il, cr[1] * il, cr[2] * il};
    const S dv[3] = {pb[0] - pa[0], pb[1] - pa[1], pb[2] - pa[2]};
    const S e = dv[0] * l[0] + dv[1] * l[1] + dv[2] * l[2];
    const S sg = e >= S(0) ? S(-1) : S(1);
    const S n[3] = {sg * l[0], sg * l[1], sg * l[2]};
    const S sa0 = (-n[0] * A0[0] - n[1] * A0[1] - n[2] * A0[2]) >= S(0) ? S(1) : S(-1);
    const S sa1 = (-n[0] * A1[0] - n[1] * A1[1] - n[2] * A1[2]) >= S(0) ? S(1) : S(-1);
    const S sb0 = (n[0] * B0[0] + n[1] * B0[1] + n[2] * B0[2]) >= S(0) ? S(1) : S(-1);
    const S sb1 = (n[0] * B1[0] + n[1] * B1[1] + n[2] * B1[2]) >= S(0) ? S(1) : S(-1);
    const S ka0 = sa0 * pick3(ha, ta0), ka1 = sa1 * pick3(ha, ta1);
    const S kb0 = sb0 * pick3(hb, tb0), kb1 = sb1 * pick3(hb, tb1);
    S ea[3], eb[3], w[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        ea[c] = pa[c] + ka0 * A0[c] + ka1 * A1[c];
        eb[c] = pb[c] + kb0 * B0[c] + kb1 * B1[c];
        w[c] = ea[c] - eb[c];
    }
    const S cc = Ai[0] * Bj[0] + Ai[1] * Bj[1] + Ai[2] * Bj[2];
    const S den = S(1) - cc * cc;
    const S d1 = w[0] * Ai[0] + w[1] * Ai[1] + w[2] * Ai[2];
    const S d2 = w[0] * Bj[0] + w[1] * Bj[1] + w[2] * Bj[2];
    const bool solve = absval(den) > S(1e-6f);
    const S u = solve ? (cc * d2 - d1) / den : S(0);
    const S v = solve ? (d2 - cc * d1) / den : S(0);
    // reverse: p = (ea + u Ai + eb + v Bj) / 2
    S Aib[3], Bjb[3], eab[3], ebb[3];
    const S pbh[3] = {S(0.5f) * hbar[0], S(0.5f) * hbar[1], S(0.5f) * hbar[2]};
    const S ub = pbh[0] * Ai[0] + pbh[1] * Ai[1] + pbh[2] * Ai[2];
    const S vb = pbh[0] * Bj[0] + pbh[1] * Bj[1] + pbh[2] * Bj[2];
    S ccb = S(0), d1b = S(0), d2b = S(0);
    if (solve) {
        const S iden = S(1) / den;
        ccb = (ub * (d2 + S(2) * cc * u) + vb * (S(2) * cc * v - d1)) * iden;
        d1b = (-ub - cc * vb) * iden;
        d2b = (cc * ub + vb) * iden;
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const S wb = d1b * Ai[c] + d2b * Bj[c];
        Aib[c] = u * pbh[c] + d1b * w[c] + ccb * Bj[c];
        Bjb[c] = v * pbh[c] + d2b * w[c] + ccb * Ai[c];
        eab[c] = pbh[c] + wb;
        ebb[c] = pbh[c] - wb;
    }
    M3T<S> Ab{}, Bb{};
    // ea = ca + ka0 A_ta0 + ka1 A_ta1, eb likewise
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            Ab.m[r][c] += (r == ta0 ? ka0 : (r == ta1 ? ka1 : S(0))) * eab[c];
            Bb.m[r][c] += (r == tb0 ? kb0 : (r == tb1 ? kb1 : S(0))) * ebb[c];
        }
    // n = sg l; depth = rA + rB - |e|
    const S deb = hbar[6];
    const S se = e >= S(0) ? S(1) : S(-1);
    S lb[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) lb[c] = sg * hbar[3 + c] - deb * se * dv[c];
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        const S tA = A.m[m][0] * l[0] + A.m[m][1] * l[1] + A.m[m][2] * l[2];
        const S tB = B.m[m][0] * l[0] + B.m[m][1] * l[1] + B.m[m][2] * l[2];
        const S kA = m == i ? S(0) : deb * ha[m] * (tA >= S(0) ? S(1) : S(-1));
        const S kB = m == j ? S(0) : deb * hb[m] * (tB >= S(0) ? S(1) : S(-1));
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            Ab.m[m][c] += kA * l[c];
            Bb.m[m][c] += kB * l[c];
            lb[c] += kA * A.m[m][c] + kB * B.m[m][c];
        }
    }
    // l = cr / |cr|, cr = Ai x Bj
    const S ll = l[0] * lb[0] + l[1] * lb[1] + l[2] * lb[2];
    S crb[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) crb[c] = (lb[c] - l[c] * ll) * il;
    Aib[0] += Bj[1] * crb[2] - Bj[2] * crb[1];
    Aib[1] += Bj[2] * crb[0] - Bj[0] * crb[2];
    Aib[2] += Bj[0] * crb[1] - Bj[1] * crb[0];
    Bjb[0] += crb[1] * Ai[2] - crb[2] * Ai[1];
    Bjb[1] += crb[2] * Ai[0] - crb[0] * Ai[2];
    Bjb[2] += crb[0] * Ai[1] - crb[1] * Ai[0];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            Ab.m[r][c] += r == i ? Aib[c] : S(0);
            Bb.m[r][c] += r == j ? Bjb[c] : S(0);
        }
    // the depth's -|e| with e = (cb - ca).l
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const S dvb = -deb * se * l[c];
        ga[c] += eab[c] - dvb;
        gb[c] += ebb[c] + dvb;
    }
    S qab[4] = {S(0), S(0), S(0), S(0)}, qbb[4] = {S(0), S(0), S(0), S(0)};
    axes_adj(qa, Ab, qab);
    axes_adj(qb, Bb, qbb);
#pragma unroll
    for (int k = 0; k < 4; ++k) { ga[3 + k] += qab[k]; gb[3 + k] += qbb[k]; }
}

// Box-box face contact (the usual case of resting boxes): the manifold's
// point and depth are means over the kept incident vertices v_k = c_I +
// sum_m s_km h_m I_m, and the normal is a signed row of one box's axes, so
// the adjoint is linear in the kept vertices.  The winning axis, the
// reference face and the passing vertices come from the feature word; the
// first four passing vertices in vertex order are the kept ones.
template <class S>
DY_HD void box_face_adj(uint32_t feat, const S ha[3], const S hb[3], const S pa[3], const S qa[4], const S pb[3],
                        const S qb[4], const S hbar[7], S ga[7], S gb[7]) {
    const S one = S(1);
    const M3T<S> A = axes_t(qa[0], qa[1], qa[2], qa[3], one);
    const M3T<S> B = axes_t(qb[0], qb[1], qb[2], qb[3], one);
    const uint32_t sat = (feat >> 16) & 0xffu;
    const int ty = sat < 3u ? 0 : 1, ai = static_cast<int>(sat < 3u ? sat : sat - 3u);
    const bool a_ref = (feat & 7u) == kFeatFaceA;
    const int ref = static_cast<int>((feat >> 3) & 3u);
    // normal: sg * row ai of A (ty 0) or of B (ty 1), sign from the centre offset
    const S dx = pb[0] - pa[0], dy = pb[1] - pa[1], dz = pb[2] - pa[2];
    S l[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) l[c] = ty == 0 ? pickm(A, ai, c) : pickm(B, ai, c);
    const S sg = (dx * l[0] + dy * l[1] + dz * l[2]) >= S(0) ? S(-1) : S(1);
    const S n[3] = {sg * l[0], sg * l[1], sg * l[2]};
    const M3T<S> I = pick_mat(a_ref, B, A);
    S ih[3], rc[3], ic[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        ih[c] = a_ref ? hb[c] : ha[c];
        rc[c] = a_ref ? pa[c] : pb[c];
        ic[c] = a_ref ? pb[c] : pa[c];
    }
    const S ref_n = rc[0] * n[0] + rc[1] * n[1] + rc[2] * n[2];
    const S sdir = a_ref ? S(1) : S(-1);
    const S rh_ref = a_ref ? pick3(ha, ref) : pick3(hb, ref);
    // kept vertices: the first four passing ones, and their dv
    uint32_t kept = 0;
    S dvk[8];
    int m = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        dvk[k] = S(0);
        if (!((feat >> (8 + k)) & 1u) || m >= 4) continue;
        const S s0 = (k & 1) ? S(1) : S(-1), s1 = (k & 2) ? S(1) : S(-1), s2 = (k & 4) ? S(1) : S(-1);
        const S vx = ic[0] + s0 * ih[0] * I.m[0][0] + s1 * ih[1] * I.m[1][0] + s2 * ih[2] * I.m[2][0];
        const S vy = ic[1] + s0 * ih[0] * I.m[0][1] + s1 * ih[1] * I.m[1][1] + s2 * ih[2] * I.m[2][1];
        const S vz = ic[2] + s0 * ih[0] * I.m[0][2] + s1 * ih[1] * I.m[1][2] + s2 * ih[2] * I.m[2][2];
        const S vn = vx * n[0] + vy * n[1] + vz * n[2];
        kept |= 1u << k;
        dvk[k] = rh_ref + sdir * (vn - ref_n);
        ++m;
    }
    if (m == 0) return;
    // reverse: p = mean(v - sdir dv n), depth = mean(dv)
    const S inv = S(1) / static_cast<S>(m);
    const S sb[3] = {hbar[0] * inv, hbar[1] * inv, hbar[2] * inv};
    const S sdb = hbar[6] * inv;
    const S sbn = sb[0] * n[0] + sb[1] * n[1] + sb[2] * n[2];
    S nb[3] = {hbar[3], hbar[4], hbar[5]}, icb[3] = {S(0), S(0), S(0)}, rcb[3] = {S(0), S(0), S(0)};
    M3T<S> Ib{};
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        if (!((kept >> k) & 1u)) continue;
        const S sk[3] = {(k & 1) ? S(1) : S(-1), (k & 2) ? S(1) : S(-1), (k & 4) ? S(1) : S(-1)};
        S v[3];
#pragma unroll
        for (int c = 0; c < 3; ++c)
            v[c] = ic[c] + sk[0] * ih[0] * I.m[0][c] + sk[1] * ih[1] * I.m[1][c] + sk[2] * ih[2] * I.m[2][c];
        const S dv = dvk[k];
        const S dvb = sdb - sdir * sbn;
        S vb[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            nb[c] += -sdir * dv * sb[c] + sdir * dvb * (v[c] - rc[c]);
            vb[c] = sb[c] + sdir * dvb * n[c];
            rcb[c] += -sdir * dvb * n[c];
            icb[c] += vb[c];
        }
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) Ib.m[r][c] += sk[r] * ih[r] * vb[c];
    }
    // normal n = sg * l: the adjoint of that axis row
    M3T<S> Ab{}, Bb{};
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const S lb = (r == ai) ? sg * nb[c] : S(0);
            if (ty == 0) Ab.m[r][c] += lb;
            else Bb.m[r][c] += lb;
            if (a_ref) Bb.m[r][c] += Ib.m[r][c];
            else Ab.m[r][c] += Ib.m[r][c];
        }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        ga[c] += a_ref ? rcb[c] : icb[c];
        gb[c] += a_ref ? icb[c] : rcb[c];
    }
    S qab[4] = {S(0), S(0), S(0), S(0)}, qbb[4] = {S(0), S(0), S(0), S(0)};
    axes_adj(qa, Ab, qab);
    axes_adj(qb, Bb, qbb);
#pragma unroll
    for (int k = 0; k < 4; ++k) { ga[3 + k] += qab[k]; gb[3 + k] += qbb[k]; }
}

// Reverse mode of one contact's geometry (pair_contact_t's dispatch, the
// sphere-box normal negated when the sphere is body A): accumulates the
// adjoints of (pa xyz, qa wxyz) into ga[7] and of (pb, qb) into gb[7], given
// the adjoints hbar[7] of (p xyz, n xyz, depth) and the forward's feature
// word.
template <class S>
DY_HD void geom_adj(uint32_t ta, uint32_t tb, uint32_t feat, const S ha[3], const S hb[3], const S pa[3],
                    const S qa[4], const S pb[3], const S qb[4], const S hbar[7], S ga[7], S gb[7]) {
    if (ta == kSphere && tb == kSphere) {
        sphere_sphere_adj(ha[0], hb[0], pa, pb, hbar, ga, gb);
    } else if (ta == kSphere) {
        const S hneg[7] = {hbar[0], hbar[1], hbar[2], -hbar[3], -hbar[4], -hbar[5], hbar[6]};
        sphere_box_adj(feat, ha[0], pa, pb, qb, hb, hneg, ga, gb);
    } else if (tb == kSphere) {
        sphere_box_adj(feat, hb[0], pb, pa, qa, ha, hbar, gb, ga);
    } else if ((feat & 7u) == kFeatEdge) {
        box_edge_adj(feat, ha, hb, pa, qa, pb, qb, hbar, ga, gb);
    } else {
        box_face_adj(feat, ha, hb, pa, qa, pb, qb, hbar, ga, gb);
    }
}

// ---------------------------------------------------------------------------
// Body state, the forward maps, and their adjoints
// ---------------------------------------------------------------------------
// Mutable state p xyz, q wxyz, v xyz, w xyz (13) plus inverse mass and the
// body-frame inverse inertia (xx yy zz xy xz yz).
template <class T> struct BodyT {
    T p[3], q[4], v[3], w[3];
    T im;
    T I[6];
};
// Adjoint of the mutable state and of the inverse mass.
template <class T> struct BodyBar {
    T p[3], q[4], v[3], w[3];
    T im;
};

template <class T> DY_HD void zero(BodyBar<T> &b) {
    #pragma unroll
    for (int k = 0; k < 3; ++k) { b.p[k] = T(0); b.v[k] = T(0); b.w[k] = T(0); }
    #pragma unroll
    for (int k = 0; k < 4; ++k) b.q[k] = T(0);
    b.im = T(0);
}

template <class T> DY_HD void cross3(const T a[3], const T b[3], T o[3]) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
template <class T> DY_HD T dot3(const T a[3], const T b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// Velocity of a body point at lever r: v + w x r, in project_contact's
// operation order.
template <class T> DY_HD void point_vel(const T v[3], const T w[3], const T r[3], T o[3]) {
    o[0] = v[0] + w[1] * r[2] - w[2] * r[1];
    o[1] = v[1] + w[2] * r[0] - w[0] * r[2];
    o[2] = v[2] + w[0] * r[1] - w[1] * r[0];
}

// o = M D M^T v (inv_inertia_world: M = axes(q), D the symmetric inverse inertia).
template <class T> DY_HD void iw_apply(const M3T<T> &M, const T I[6], const T v[3], T o[3]) {
    const T lx = M.m[0][0] * v[0] + M.m[1][0] * v[1] + M.m[2][0] * v[2];
    const T ly = M.m[0][1] * v[0] + M.m[1][1] * v[1] + M.m[2][1] * v[2];
    const T lz = M.m[0][2] * v[0] + M.m[1][2] * v[1] + M.m[2][2] * v[2];
    const T jx = I[0] * lx + I[3] * ly + I[4] * lz;
    const T jy = I[3] * lx + I[1] * ly + I[5] * lz;
    const T jz = I[4] * lx + I[5] * ly + I[2] * lz;
    o[0] = M.m[0][0] * jx + M.m[0][1] * jy + M.m[0][2] * jz;
    o[1] = M.m[1][0] * jx + M.m[1][1] * jy + M.m[1][2] * jz;
    o[2] = M.m[2][0] * jx + M.m[2][1] * jy + M.m[2][2] * jz;
}
// Adjoint of iw_apply for output adjoint ob: vbar += M D M^T ob, Mbar += ...
template <class T> DY_HD void iw_adj(const M3T<T> &M, const T I[6], const T v[3], const T ob[3], T vbar[3],
                                     M3T<T> &Mbar) {
#pragma clang fp contract(fast)  // adjoint-only arithmetic (see project_adj)
    T l[3], j[3], jb[3], lb[3];
    #pragma unroll
    for (int k = 0; k < 3; ++k) l[k] = M.m[0][k] * v[0] + M.m[1][k] * v[1] + M.m[2][k] * v[2];
    j[0] = I[0] * l[0] + I[3] * l[1] + I[4] * l[2];
    j[1] = I[3] * l[0] + I[1] * l[1] + I[5] * l[2];
    j[2] = I[4] * l[0] + I[5] * l[1] + I[2] * l[2];
    #pragma unroll
    for (int k = 0; k < 3; ++k) jb[k] = M.m[0][k] * ob[0] + M.m[1][k] * ob[1] + M.m[2][k] * ob[2];
    #pragma unroll
    for (int i = 0; i < 3; ++i)
        #pragma unroll
        for (int k = 0; k < 3; ++k) Mbar.m[i][k] += ob[i] * j[k];
    lb[0] = I[0] * jb[0] + I[3] * jb[1] + I[4] * jb[2];
    lb[1] = I[3] * jb[0] + I[1] * jb[1] + I[5] * jb[2];
    lb[2] = I[4] * jb[0] + I[5] * jb[1] + I[2] * jb[2];
    #pragma unroll
    for (int i = 0; i < 3; ++i) {
        vbar[i] += M.m[i][0] * lb[0] + M.m[i][1] * lb[1] + M.m[i][2] * lb[2];
        #pragma unroll
        for (int k = 0; k < 3; ++k) Mbar.m[i][k] += v[i] * lb[k];
    }
}
// Adjoint of axes_t: qbar += dM/dq^T Mbar (q = w x y z).
template <class T> DY_HD void axes_adj(const T q[4], const M3T<T> &B, T qb[4]) {
#pragma clang fp contract(fast)  // adjoint-only arithmetic (see project_adj)
    const T w = q[0], x = q[1], y = q[2], z = q[3];
    const T two = T(2), four = T(4);
    qb[0] += two * (z * B.m[0][1] - y * B.m[0][2] - z * B.m[1][0] + x * B.m[1][2] + y * B.m[2][0] - x * B.m[2][1]);
    qb[1] += two * (y * B.m[0][1] + z * B.m[0][2] + y * B.m[1][0] + w * B.m[1][2] + z * B.m[2][0] - w * B.m[2][1])
           - four * x * (B.m[1][1] + B.m[2][2]);
    qb[2] += two * (x * B.m[0][1] - w * B.m[0][2] + x * B.m[1][0] + z * B.m[1][2] + w * B.m[2][0] + z * B.m[2][1])
           - four * y * (B.m[0][0] + B.m[2][2]);
    qb[3] += two * (w * B.m[0][1] + x * B.m[0][2] - w * B.m[1][0] + y * B.m[1][2] + x * B.m[2][0] + y * B.m[2][1])
           - four * z * (B.m[0][0] + B.m[1][1]);
}

// q += 0.5 * (0, w) (x) q (rotate_by) and its adjoint.
template <class T> DY_HD void rot_apply(T q[4], const T w[3]) {
    const T qw = q[0], qx = q[1], qy = q[2], qz = q[3];
    const T h = T(0.5f);
    q[0] = qw + h * (-w[0] * qx - w[1] * qy - w[2] * qz);
    q[1] = qx + h * (w[0] * qw + w[1] * qz - w[2] * qy);
    q[2] = qy + h * (-w[0] * qz + w[1] * qw + w[2] * qx);
    q[3] = qz + h * (w[0] * qy - w[1] * qx + w[2] * qw);
}
// Given q (before), w and the adjoint q1b of the result: adds to qb and wb.
template <class T> DY_HD void rot_adj(const T q[4], const T w[3], const T q1b[4], T qb[4], T wb[3]) {
#pragma clang fp contract(fast)  // adjoint-only arithmetic (see project_adj)
    const T h = T(0.5f);
    const T bw = q1b[0], bx = q1b[1], by = q1b[2], bz = q1b[3];
    qb[0] += bw + h * (w[0] * bx + w[1] * by + w[2] * bz);
    qb[1] += bx + h * (-w[0] * bw + w[2] * by - w[1] * bz);
    qb[2] += by + h * (-w[1] * bw - w[2] * bx + w[0] * bz);
    qb[3] += bz + h * (-w[2] * bw + w[1] * bx - w[0] * by);
    wb[0] += h * (-q[1] * bw + q[0] * bx - q[3] * by + q[2] * bz);
    wb[1] += h * (-q[2] * bw + q[3] * bx + q[0] * by - q[1] * bz);
    wb[2] += h * (-q[3] * bw - q[2] * bx + q[1] * by + q[0] * bz);
}

// Quaternion renormalisation q' = q / |q| and its adjoint (in place on qb).
template <class T> DY_HD void renorm_t(T q[4]) {
    const T s = T(1) / sq_root(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    #pragma unroll
    for (int k = 0; k < 4; ++k) q[k] = q[k] * s;
}
template <class T> DY_HD void renorm_adj(const T q[4], T qb[4]) {
    const T s = T(1) / sq_root(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    T u[4];
    #pragma unroll
    for (int k = 0; k < 4; ++k) u[k] = q[k] * s;
    const T d = u[0] * qb[0] + u[1] * qb[1] + u[2] * qb[2] + u[3] * qb[3];
    #pragma unroll
    for (int k = 0; k < 4; ++k) qb[k] = s * (qb[k] - u[k] * d);
}

// One integrator substep of a dynamic body (integrate_body) and its adjoint.
template <class T> DY_HD void integrate_t(BodyT<T> &b, const T g[3], T h) {
    #pragma unroll
    for (int k = 0; k < 3; ++k) b.v[k] = b.v[k] + g[k] * h;
    #pragma unroll
    for (int k = 0; k < 3; ++k) b.p[k] += b.v[k] * h;
    const T hh = T(0.5f) * h;
    const T qw = b.q[0], qx = b.q[1], qy = b.q[2], qz = b.q[3];
    const T ew = -(b.w[0] * qx + b.w[1] * qy + b.w[2] * qz);
    const T ex = (b.w[0] * qw + b.w[1] * qz - b.w[2] * qy);
    const T ey = (b.w[1] * qw + b.w[2] * qx - b.w[0] * qz);
    const T ez = (b.w[2] * qw + b.w[0] * qy - b.w[1] * qx);
    b.q[0] += hh * ew; b.q[1] += hh * ex; b.q[2] += hh * ey; b.q[3] += hh * ez;
    renorm_t(b.q);
}
// b0: the state before the substep; bb: adjoint of the state after, replaced
// by the adjoint of the state before.
template <class T> DY_HD void integrate_adj(const BodyT<T> &b0, const T g[3], T h, BodyBar<T> &bb) {
    const T hh = T(0.5f) * h;
    const T *q = b0.q, *w = b0.w;
    const T e[4] = {-(w[0] * q[1] + w[1] * q[2] + w[2] * q[3]), (w[0] * q[0] + w[1] * q[3] - w[2] * q[2]),
                    (w[1] * q[0] + w[2] * q[1] - w[0] * q[3]), (w[2] * q[0] + w[0] * q[2] - w[1] * q[1])};
    T qt[4];
    #pragma unroll
    for (int k = 0; k < 4; ++k) qt[k] = q[k] + hh * e[k];
    renorm_adj(qt, bb.q);  // bb.q is now the adjoint of qt
    T eb[4];
    #pragma unroll
    for (int k = 0; k < 4; ++k) eb[k] = hh * bb.q[k];
    // q-bar: identity part stays in bb.q; add Omega(w)^T eb
    T qb[4] = {bb.q[0], bb.q[1], bb.q[2], bb.q[3]};
    qb[0] += w[0] * eb[1] + w[1] * eb[2] + w[2] * eb[3];
    qb[1] += -w[0] * eb[0] + w[2] * eb[2] - w[1] * eb[3];
    qb[2] += -w[1] * eb[0] - w[2] * eb[1] + w[0] * eb[3];
    qb[3] += -w[2] * eb[0] + w[1] * eb[1] - w[0] * eb[2];
    bb.w[0] += -q[1] * eb[0] + q[0] * eb[1] - q[3] * eb[2] + q[2] * eb[3];
    bb.w[1] += -q[2] * eb[0] + q[3] * eb[1] + q[0] * eb[2] - q[1] * eb[3];
    bb.w[2] += -q[3] * eb[0] - q[2] * eb[1] + q[1] * eb[2] + q[0] * eb[3];
    #pragma unroll
    for (int k = 0; k < 4; ++k) bb.q[k] = qb[k];
    #pragma unroll
    for (int k = 0; k < 3; ++k) bb.v[k] += h * bb.p[k];  // p1 = p + v1 h; v1 = v + g h
}

// wa + wb + da.fa + db.fb summed left to right as project_contact does.
template <class T> DY_HD T fric_w(T wa, T wb, const T da[3], const T fa[3], const T db[3], const T fb[3]) {
    return wa + wb + da[0] * fa[0] + da[1] * fa[1] + da[2] * fa[2] + db[0] * fb[0] + db[1] * fb[1] + db[2] * fb[2];
}

// Contact input of a projection: point, normal, depth.
template <class T> struct ContactT { T p[3], n[3], depth; };

// project_contact (dy_math.hpp) on BodyT; returns true when the bodies changed.
template <class T> DY_HD bool project_t(const ContactT<T> &c, BodyT<T> &A, BodyT<T> &B, T mu, T &lam, T dt) {
    if (c.depth <= T(0)) return false;
    const T rem = c.depth - lam;
    if (rem <= T(1e-6f)) return false;
    const T *n = c.n;
    const T wa = A.im, wb = B.im;
    T ra[3], rb[3], ca[3], cb[3], ka[3], kb[3];
    #pragma unroll
    for (int k = 0; k < 3; ++k) { ra[k] = c.p[k] - A.p[k]; rb[k] = c.p[k] - B.p[k]; }
    cross3(ra, n, ca);
    cross3(rb, n, cb);
    iw_apply(axes_t(A.q[0], A.q[1], A.q[2], A.q[3], T(1)), A.I, ca, ka);
    iw_apply(axes_t(B.q[0], B.q[1], B.q[2], B.q[3], T(1)), B.I, cb, kb);
    const T w = wa + wb + dot3(ca, ka) + dot3(cb, kb);
    if (w < T(1e-10f)) return false;
    const T dl = rem / w;
    lam += rem;
    if (wa != T(0)) for (int k = 0; k < 3; ++k) A.p[k] += wa * dl * n[k];
    if (wb != T(0)) for (int k = 0; k < 3; ++k) B.p[k] -= wb * dl * n[k];
    if (wa != T(0)) { const T om[3] = {dl * ka[0], dl * ka[1], dl * ka[2]}; rot_apply(A.q, om); }
    if (wb != T(0)) { const T om[3] = {-dl * kb[0], -dl * kb[1], -dl * kb[2]}; rot_apply(B.q, om); }
    {
        T va[3], vb[3], rv[3];
        point_vel(A.v, A.w, ra, va);
        point_vel(B.v, B.w, rb, vb);
        #pragma unroll
        for (int k = 0; k < 3; ++k) rv[k] = va[k] - vb[k];
        const T vn = dot3(rv, n);
        if (vn < T(0)) {
            const T jn = -vn / w;
            T ga[3], gb[3];
            iw_apply(axes_t(A.q[0], A.q[1], A.q[2], A.q[3], T(1)), A.I, ca, ga);
            iw_apply(axes_t(B.q[0], B.q[1], B.q[2], B.q[3], T(1)), B.I, cb, gb);
            if (wa != T(0)) for (int k = 0; k < 3; ++k) { A.v[k] += wa * jn * n[k]; A.w[k] += jn * ga[k]; }
            if (wb != T(0)) for (int k = 0; k < 3; ++k) { B.v[k] -= wb * jn * n[k]; B.w[k] -= jn * gb[k]; }
        }
    }
    if (mu > T(0)) {
        T va[3], vb[3], u[3], t0[3];
        point_vel(A.v, A.w, ra, va);
        point_vel(B.v, B.w, rb, vb);
        #pragma unroll
        for (int k = 0; k < 3; ++k) u[k] = va[k] - vb[k];
        const T un = dot3(u, n);
        #pragma unroll
        for (int k = 0; k < 3; ++k) t0[k] = u[k] - un * n[k];
        const T t2 = dot3(t0, t0);
        if (t2 > T(1e-12f)) {
            const T tl = sq_root(t2);
            T t[3], da[3], db[3], fa[3], fb[3];
            #pragma unroll
            for (int k = 0; k < 3; ++k) t[k] = t0[k] / tl;
            cross3(ra, t, da);
            cross3(rb, t, db);
            iw_apply(axes_t(A.q[0], A.q[1], A.q[2], A.q[3], T(1)), A.I, da, fa);
            iw_apply(axes_t(B.q[0], B.q[1], B.q[2], B.q[3], T(1)), B.I, db, fb);
            const T wt = fric_w(wa, wb, da, fa, db, fb);
            if (wt > T(1e-10f)) {
                const T jcap = lam / (dt * w);
                const T jt0 = tl / wt;
                const T jt = jt0 < mu * jcap ? jt0 : mu * jcap;
                if (wa != T(0)) for (int k = 0; k < 3; ++k) { A.v[k] -= wa * jt * t[k]; A.w[k] -= jt * fa[k]; }
                if (wb != T(0)) for (int k = 0; k < 3; ++k) { B.v[k] += wb * jt * t[k]; B.w[k] += jt * fb[k]; }
            }
        }
    }
    return true;
}

// Adjoint of project_t.  A0, B0, lam0: the inputs of the projection.  Ab, Bb
// hold the adjoints of the outputs on entry and of the inputs on return
// (im-bar accumulates); lamb likewise; cb accumulates the contact's adjoint.
template <class T>
DY_HD void project_adj(const ContactT<T> &c, const BodyT<T> &A0, const BodyT<T> &B0, T mu, T lam0, T dt,
                       BodyBar<T> &Ab, BodyBar<T> &Bb, T &lamb, ContactT<T> &cb) {
    if (c.depth <= T(0)) return;
    const T rem = c.depth - lam0;
    if (rem <= T(1e-6f)) return;
    const T *n = c.n;
    const T wa = A0.im, wb = B0.im;
    const bool dA = wa != T(0), dB = wb != T(0);
    T ra[3], rb[3], ca[3], cbv[3], ka[3], kb[3];
    #pragma unroll
    for (int k = 0; k < 3; ++k) { ra[k] = c.p[k] - A0.p[k]; rb[k] = c.p[k] - B0.p[k]; }
    cross3(ra, n, ca);
    cross3(rb, n, cbv);
    const M3T<T> MA = axes_t(A0.q[0], A0.q[1], A0.q[2], A0.q[3], T(1));
    const M3T<T> MB = axes_t(B0.q[0], B0.q[1], B0.q[2], B0.q[3], T(1));
    iw_apply(MA, A0.I, ca, ka);
    iw_apply(MB, B0.I, cbv, kb);
    const T w = wa + wb + dot3(ca, ka) + dot3(cbv, kb);
    if (w < T(1e-10f)) return;
    const T dl = rem / w;
    const T lam1 = lam0 + rem;
    // forward states after the position stage
    T qA1[4], qB1[4], omA[3], omB[3];
    #pragma unroll
    for (int k = 0; k < 4; ++k) { qA1[k] = A0.q[k]; qB1[k] = B0.q[k]; }
    #pragma unroll
    for (int k = 0; k < 3; ++k) { omA[k] = dl * ka[k]; omB[k] = -dl * kb[k]; }
    if (dA) rot_apply(qA1, omA);
    if (dB) rot_apply(qB1, omB);
    const M3T<T> MA1 = axes_t(qA1[0], qA1[1], qA1[2], qA1[3], T(1));
    const M3T<T> MB1 = axes_t(qB1[0], qB1[1], qB1[2], qB1[3], T(1));
    // velocity stage
    T vA1[3], wA1[3], vB1[3], wB1[3], rv[3], ga[3], gb[3];
    {
        T va[3], vb[3];
        point_vel(A0.v, A0.w, ra, va);
        point_vel(B0.v, B0.w, rb, vb);
        #pragma unroll
        for (int k = 0; k < 3; ++k) rv[k] = va[k] - vb[k];
    }
    const T vn = dot3(rv, n);
    const bool vel = vn < T(0);
    const T jn = vel ? -vn / w : T(0);
    iw_apply(MA1, A0.I, ca, ga);
    iw_apply(MB1, B0.I, cbv, gb);
    #pragma unroll
    for (int k = 0; k < 3; ++k) {
        vA1[k] = A0.v[k]; wA1[k] = A0.w[k]; vB1[k] = B0.v[k]; wB1[k] = B0.w[k];
        if (vel && dA) { vA1[k] += wa * jn * n[k]; wA1[k] += jn * ga[k]; }
        if (vel && dB) { vB1[k] -= wb * jn * n[k]; wB1[k] -= jn * gb[k]; }
    }
    // friction stage
    bool fric = false, cap = false;
    T u[3], t0[3], un = T(0), tl = T(1), t[3], da[3], db[3], fa[3], fb[3], wt = T(1), jt = T(0), jcap = T(0),
        jt0 = T(0);
    if (mu > T(0)) {
        T va[3], vb[3];
        point_vel(vA1, wA1, ra, va);
        point_vel(vB1, wB1, rb, vb);
        #pragma unroll
        for (int k = 0; k < 3; ++k) u[k] = va[k] - vb[k];
        un = dot3(u, n);
        #pragma unroll
        for (int k = 0; k < 3; ++k) t0[k] = u[k] - un * n[k];
        const T t2 = dot3(t0, t0);
        if (t2 > T(1e-12f)) {
            tl = sq_root(t2);
            #pragma unroll
            for (int k = 0; k < 3; ++k) t[k] = t0[k] / tl;
            cross3(ra, t, da);
            cross3(rb, t, db);
            iw_apply(MA1, A0.I, da, fa);
            iw_apply(MB1, B0.I, db, fb);
            wt = fric_w(wa, wb, da, fa, db, fb);
            if (wt > T(1e-10f)) {
                fric = true;
                jcap = lam1 / (dt * w);
                jt0 = tl / wt;
                cap = !(jt0 < mu * jcap);
                jt = cap ? mu * jcap : jt0;
            }
        }
    }

    // ---- reverse ----
    // Contracted (fused multiply-adds): the forward replay above must repeat
    // the forward's bits, so that every branch is the forward's; the
    // adjoint's own rounding is pinned by no reference and only shortens the
    // dependent chain.
    {
#pragma clang fp contract(fast)
    // one reciprocal per denominator (the forward replay divides exactly)
    const T rw = T(1) / w;
    T rab[3] = {T(0), T(0), T(0)}, rbb[3] = {T(0), T(0), T(0)};
    T cab[3] = {T(0), T(0), T(0)}, cbb[3] = {T(0), T(0), T(0)};
    T wbar = T(0), dlb = T(0), remb = T(0);
    M3T<T> MAb{}, MBb{}, MA1b{}, MB1b{};
    // running adjoints of the velocity state (start: outputs)
    T avb[3], awb[3], bvb[3], bwb[3];
    #pragma unroll
    for (int k = 0; k < 3; ++k) { avb[k] = Ab.v[k]; awb[k] = Ab.w[k]; bvb[k] = Bb.v[k]; bwb[k] = Bb.w[k]; }
    T lam1b = lamb;
    if (fric) {
        T jtb = T(0), tb[3] = {T(0), T(0), T(0)}, fab[3] = {T(0), T(0), T(0)}, fbb[3] = {T(0), T(0), T(0)};
        if (dA) {
            const T dvt = dot3(avb, t);
            jtb += -wa * dvt - dot3(awb, fa);
            Ab.im += -jt * dvt;
            #pragma unroll
            for (int k = 0; k < 3; ++k) { tb[k] += -(wa * jt) * avb[k]; fab[k] += -jt * awb[k]; }
        }
        if (dB) {
            const T dvt = dot3(bvb, t);
            jtb += wb * dvt + dot3(bwb, fb);
            Bb.im += jt * dvt;
            #pragma unroll
            for (int k = 0; k < 3; ++k) { tb[k] += (wb * jt) * bvb[k]; fbb[k] += jt * bwb[k]; }
        }
        T jt0b = T(0), jcapb = T(0);
        if (cap) jcapb = mu * jtb;
        else jt0b = jtb;
        // jcap = lam1 / (dt w)
        lam1b += jcapb * rw / dt;
        wbar += -jcapb * jcap * rw;
        // jt0 = tl / wt
        const T rwt = T(1) / wt;
        T tlb = jt0b * rwt;
        const T wtb = -jt0b * jt0 * rwt;
        // wt = wa + wb + da.fa + db.fb
        Ab.im += wtb;
        Bb.im += wtb;
        T dab[3], dbb[3];
        #pragma unroll
        for (int k = 0; k < 3; ++k) {
            dab[k] = wtb * fa[k];
            fab[k] += wtb * da[k];
            dbb[k] = wtb * fb[k];
            fbb[k] += wtb * db[k];
        }
        iw_adj(MA1, A0.I, da, fab, dab, MA1b);
        iw_adj(MB1, B0.I, db, fbb, dbb, MB1b);
        // da = ra x t: ra-bar += t x da-bar; t-bar += da-bar x ra
        T tmp[3];
        cross3(t, dab, tmp); for (int k = 0; k < 3; ++k) rab[k] += tmp[k];
        cross3(dab, ra, tmp); for (int k = 0; k < 3; ++k) tb[k] += tmp[k];
        cross3(t, dbb, tmp); for (int k = 0; k < 3; ++k) rbb[k] += tmp[k];
        cross3(dbb, rb, tmp); for (int k = 0; k < 3; ++k) tb[k] += tmp[k];
        // t = t0 / tl
        T t0b[3];
        T s = T(0);
        const T rtl = T(1) / tl;
        #pragma unroll
        for (int k = 0; k < 3; ++k) { t0b[k] = tb[k] * rtl; s += tb[k] * t0[k]; }
        tlb += -s * rtl * rtl;
        // tl = sqrt(t2); t2 = t0.t0
        const T t2b = tlb * (T(0.5f) * rtl);
        #pragma unroll
        for (int k = 0; k < 3; ++k) t0b[k] += T(2) * t2b * t0[k];
        // t0 = u - un n
        T ub[3];
        const T unb = -dot3(t0b, n);
        #pragma unroll
        for (int k = 0; k < 3; ++k) { ub[k] = t0b[k] + unb * n[k]; cb.n[k] += -un * t0b[k] + unb * u[k]; }
        // u = (vA1 + wA1 x ra) - (vB1 + wB1 x rb)
        #pragma unroll
        for (int k = 0; k < 3; ++k) { avb[k] += ub[k]; bvb[k] -= ub[k]; }
        cross3(ra, ub, tmp); for (int k = 0; k < 3; ++k) awb[k] += tmp[k];
        cross3(ub, wA1, tmp); for (int k = 0; k < 3; ++k) rab[k] += tmp[k];
        T nub[3] = {-ub[0], -ub[1], -ub[2]};
        cross3(rb, nub, tmp); for (int k = 0; k < 3; ++k) bwb[k] += tmp[k];
        cross3(nub, wB1, tmp); for (int k = 0; k < 3; ++k) rbb[k] += tmp[k];
    }
    // velocity stage reverse: v1 = v + wa jn n, w1 = w + jn ga (A); minus for B
    T gab[3] = {T(0), T(0), T(0)}, gbb[3] = {T(0), T(0), T(0)};
    if (vel) {
        T jnb = T(0);
        if (dA) {
            const T dvn = dot3(avb, n);
            jnb += wa * dvn + dot3(awb, ga);
            Ab.im += jn * dvn;
            #pragma unroll
            for (int k = 0; k < 3; ++k) { cb.n[k] += (wa * jn) * avb[k]; gab[k] += jn * awb[k]; }
        }
        if (dB) {
            const T dvn = dot3(bvb, n);
            jnb += -wb * dvn - dot3(bwb, gb);
            Bb.im += -jn * dvn;
            #pragma unroll
            for (int k = 0; k < 3; ++k) { cb.n[k] += -(wb * jn) * bvb[k]; gbb[k] += -jn * bwb[k]; }
        }
        iw_adj(MA1, A0.I, ca, gab, cab, MA1b);
        iw_adj(MB1, B0.I, cbv, gbb, cbb, MB1b);
        // jn = -vn / w
        const T vnb = -jnb * rw;
        wbar += jnb * vn * rw * rw;
        // vn = rv . n; rv = (vA + wA x ra) - (vB + wB x rb)
        T rvb[3], tmp[3];
        #pragma unroll
        for (int k = 0; k < 3; ++k) { rvb[k] = vnb * n[k]; cb.n[k] += vnb * rv[k]; }
        #pragma unroll
        for (int k = 0; k < 3; ++k) { avb[k] += rvb[k]; bvb[k] -= rvb[k]; }
        cross3(ra, rvb, tmp); for (int k = 0; k < 3; ++k) awb[k] += tmp[k];
        cross3(rvb, A0.w, tmp); for (int k = 0; k < 3; ++k) rab[k] += tmp[k];
        T nrv[3] = {-rvb[0], -rvb[1], -rvb[2]};
        cross3(rb, nrv, tmp); for (int k = 0; k < 3; ++k) bwb[k] += tmp[k];
        cross3(nrv, B0.w, tmp); for (int k = 0; k < 3; ++k) rbb[k] += tmp[k];
    }
    #pragma unroll
    for (int k = 0; k < 3; ++k) { Ab.v[k] = avb[k]; Ab.w[k] = awb[k]; Bb.v[k] = bvb[k]; Bb.w[k] = bwb[k]; }
    // quaternions: q1-bar = output adjoint + axes adjoint at q1, then rotate_by reverse
    T qA1b[4], qB1b[4];
    #pragma unroll
    for (int k = 0; k < 4; ++k) { qA1b[k] = Ab.q[k]; qB1b[k] = Bb.q[k]; }
    axes_adj(qA1, MA1b, qA1b);
    axes_adj(qB1, MB1b, qB1b);
    T qAb[4] = {T(0), T(0), T(0), T(0)}, qBb[4] = {T(0), T(0), T(0), T(0)};
    T kab[3] = {T(0), T(0), T(0)}, kbb[3] = {T(0), T(0), T(0)};
    if (dA) {
        T omb[3] = {T(0), T(0), T(0)};
        rot_adj(A0.q, omA, qA1b, qAb, omb);
        dlb += dot3(omb, ka);
        #pragma unroll
        for (int k = 0; k < 3; ++k) kab[k] += dl * omb[k];
    } else {
        #pragma unroll
        for (int k = 0; k < 4; ++k) qAb[k] = qA1b[k];
    }
    if (dB) {
        T omb[3] = {T(0), T(0), T(0)};
        rot_adj(B0.q, omB, qB1b, qBb, omb);
        dlb += -dot3(omb, kb);
        #pragma unroll
        for (int k = 0; k < 3; ++k) kbb[k] += -dl * omb[k];
    } else {
        #pragma unroll
        for (int k = 0; k < 4; ++k) qBb[k] = qB1b[k];
    }
    // positions: pA1 = pA + wa dl n; pB1 = pB - wb dl n
    if (dA) {
        const T dpn = dot3(Ab.p, n);
        Ab.im += dl * dpn;
        dlb += wa * dpn;
        #pragma unroll
        for (int k = 0; k < 3; ++k) cb.n[k] += (wa * dl) * Ab.p[k];
    }
    if (dB) {
        const T dpn = dot3(Bb.p, n);
        Bb.im += -dl * dpn;
        dlb += -wb * dpn;
        #pragma unroll
        for (int k = 0; k < 3; ++k) cb.n[k] += -(wb * dl) * Bb.p[k];
    }
    // lam1 = lam0 + rem; dl = rem / w
    lamb = lam1b;
    remb += lam1b;
    remb += dlb * rw;
    wbar += -dlb * dl * rw;
    // w = wa + wb + ca.ka + cb.kb
    Ab.im += wbar;
    Bb.im += wbar;
    #pragma unroll
    for (int k = 0; k < 3; ++k) {
        cab[k] += wbar * ka[k];
        kab[k] += wbar * ca[k];
        cbb[k] += wbar * kb[k];
        kbb[k] += wbar * cbv[k];
    }
    iw_adj(MA, A0.I, ca, kab, cab, MAb);
    iw_adj(MB, B0.I, cbv, kbb, cbb, MBb);
    axes_adj(A0.q, MAb, qAb);
    axes_adj(B0.q, MBb, qBb);
    #pragma unroll
    for (int k = 0; k < 4; ++k) { Ab.q[k] = qAb[k]; Bb.q[k] = qBb[k]; }
    // ca = ra x n: ra-bar += n x ca-bar; n-bar += ca-bar x ra
    T tmp[3];
    cross3(n, cab, tmp); for (int k = 0; k < 3; ++k) rab[k] += tmp[k];
    cross3(cab, ra, tmp); for (int k = 0; k < 3; ++k) cb.n[k] += tmp[k];
    cross3(n, cbb, tmp); for (int k = 0; k < 3; ++k) rbb[k] += tmp[k];
    cross3(cbb, rb, tmp); for (int k = 0; k < 3; ++k) cb.n[k] += tmp[k];
    // ra = p - pA; rb = p - pB
    #pragma unroll
    for (int k = 0; k < 3; ++k) {
        cb.p[k] += rab[k] + rbb[k];
        Ab.p[k] += -rab[k];
        Bb.p[k] += -rbb[k];
    }
    // rem = depth - lam0
    cb.depth += remb;
    lamb += -remb;
    }
}

}  // namespace dy
