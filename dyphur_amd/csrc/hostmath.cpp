// hostmath.cpp — the per-element arithmetic of dy_math.hpp compiled for the
// host (hipcc --cuda-host-only), exported with a C ABI so the CPU test suite
// can compare the exact code the kernels run against the oracle without a GPU.
#include "adjoint.hpp"
#include "joint_math.hpp"

#include <condition_variable>
#include <cstring>
#include <mutex>
#include <random>
#include <thread>
#include <vector>

using namespace dy;

extern "C" {

uint16_t hm_f32_to_f16(float f) { return f32_to_f16(f); }
float hm_f16_to_f32(uint16_t h) { return f16_to_f32(h); }
void hm_f32_to_f16_array(const float *in, uint16_t *out, size_t n) {
    for (size_t i = 0; i < n; ++i) out[i] = f32_to_f16(in[i]);
}
uint32_t hm_morton3(uint32_t x, uint32_t y, uint32_t z) { return morton3(x, y, z); }
uint32_t hm_quantize10(float p, float lo, float hi) { return quantize10(p, lo, hi); }
uint32_t hm_hash_u32(uint32_t x) { return hash_u32(x); }

// Karras node i over n sorted keys: out[0..4] = lo, hi, split, escape(left), escape(right).
void hm_karras_node(const uint64_t *codes, int n, int i, int32_t *out) {
    int lo, hi, g;
    karras_node(codes, n, i, lo, hi, g);
    out[0] = lo; out[1] = hi; out[2] = g;
    out[3] = karras_escape(codes, n, g);
    out[4] = karras_escape(codes, n, hi);
}

void hm_leaf_extent(uint32_t stype, const float *half3, const float *q4, float *out3) {
    leaf_extent(stype, half3[0], half3[1], half3[2], q4[0], q4[1], q4[2], q4[3], out3[0], out3[1], out3[2]);
}

// Contact of one pair: types, half extents (3 each), centres (3 each),
// quaternions wxyz (4 each).  out[7] = p xyz, n xyz, depth; *feat (may be
// null) the feature word.  Returns 1 on hit.
int hm_pair_contact_feat(uint32_t ta, uint32_t tb, const float *ha, const float *hb, const float *pa,
                         const float *pb, const float *qa, const float *qb, float *out, uint32_t *feat) {
    Hit h{};
    bool hit = false;
    if (ta == kSphere && tb == kSphere) {
        hit = sphere_sphere(pa[0], pa[1], pa[2], ha[0], pb[0], pb[1], pb[2], hb[0], h);
    } else if (ta == kSphere && tb == kBox) {
        const M3 E = body_axes(qb[0], qb[1], qb[2], qb[3]);
        hit = sphere_box(pa[0], pa[1], pa[2], ha[0], pb[0], pb[1], pb[2], E, hb, h);
        if (hit) { h.nx = -h.nx; h.ny = -h.ny; h.nz = -h.nz; }
    } else if (ta == kBox && tb == kSphere) {
        const M3 E = body_axes(qa[0], qa[1], qa[2], qa[3]);
        hit = sphere_box(pb[0], pb[1], pb[2], hb[0], pa[0], pa[1], pa[2], E, ha, h);
    } else if (ta == kBox && tb == kBox) {
        const M3 A = body_axes(qa[0], qa[1], qa[2], qa[3]);
        const M3 B = body_axes(qb[0], qb[1], qb[2], qb[3]);
        hit = box_box(pa[0], pa[1], pa[2], A, ha, pb[0], pb[1], pb[2], B, hb, h);
    }
    if (!hit) return 0;
    out[0] = h.px; out[1] = h.py; out[2] = h.pz;
    out[3] = h.nx; out[4] = h.ny; out[5] = h.nz;
    out[6] = h.depth;
    if (feat) *feat = h.feat;
    return 1;
}

int hm_pair_contact(uint32_t ta, uint32_t tb, const float *ha, const float *hb, const float *pa, const float *pb,
                    const float *qa, const float *qb, float *out) {
    return hm_pair_contact_feat(ta, tb, ha, hb, pa, pb, qa, qb, out, nullptr);
}

// One contact projection.  contact[7] = p xyz, n xyz, depth; body[20] =
// pos xyz, rot wxyz, vel xyz, ang xyz, inv_mass, iI xx yy zz xy xz yz.
int hm_project_contact(const float *contact, float *a, float *b, float mu, float *lam, float dt) {
    const ContactIn c{contact[0], contact[1], contact[2], contact[3], contact[4], contact[5], contact[6]};
    BodyState A, B;
    static_assert(sizeof(BodyState) == 20 * sizeof(float), "BodyState layout");
    __builtin_memcpy(&A, a, sizeof(A));
    __builtin_memcpy(&B, b, sizeof(B));
    const bool r = project_contact(c, A, B, mu, *lam, dt);
    __builtin_memcpy(a, &A, sizeof(A));
    __builtin_memcpy(b, &B, sizeof(B));
    return r ? 1 : 0;
}

// kin[13] = pos xyz, rot wxyz, vel xyz, ang xyz.
void hm_integrate_body(float *kin, float gx, float gy, float gz, float h, int n_substeps) {
    Kin k;
    static_assert(sizeof(Kin) == 13 * sizeof(float), "Kin layout");
    __builtin_memcpy(&k, kin, sizeof(k));
    integrate_body(k, gx, gy, gz, h, n_substeps);
    __builtin_memcpy(kin, &k, sizeof(k));
}

void hm_renormalise(float *q4) { renormalise(q4[0], q4[1], q4[2], q4[3]); }

float hm_atan2_cr(float y, float x) { return atan2_cr(y, x); }
void hm_atan2_cr_array(const float *y, const float *x, float *out, size_t n) {
    for (size_t i = 0; i < n; ++i) out[i] = atan2_cr(y[i], x[i]);
}

// A two-body solve with one joint (parent a, child b; body[20] as above):
// the motor once, then iters x (the joint, renormalisation of both bodies),
// the reference's run_serial sequence.  jp[17] = anchor_p xyz, anchor_c xyz,
// axis xyz, lo, hi, target_pos, target_vel, stiffness, damping,
// compliance_pos, compliance_ang.
void hm_joint_pair(float *a, float *b, const float *jp, uint32_t type, float dt, int iters) {
    JointIn j;
    static_assert(sizeof(JointIn) == 18 * sizeof(float), "JointIn layout");
    __builtin_memcpy(&j, jp, 17 * sizeof(float));
    j.type = type;
    BodyState A, B;
    __builtin_memcpy(&A, a, sizeof(A));
    __builtin_memcpy(&B, b, sizeof(B));
    joint_motor(A, B, j, dt);
    for (int it = 0; it < iters; ++it) {
        solve_joint(j, A, B, dt);
        renormalise(A.qw, A.qx, A.qy, A.qz);
        renormalise(B.qw, B.qx, B.qy, B.qz);
    }
    __builtin_memcpy(a, &A, sizeof(A));
    __builtin_memcpy(b, &B, sizeof(B));
}

}  // extern "C"

// ---- derivatives (adjoint.hpp), float and double --------------------------
namespace {

template <class S>
int contact_fwd(uint32_t ta, uint32_t tb, const S *ha, const S *hb, const S *pa, const S *pb, const S *qa,
                const S *qb, S *out, int *winner) {
    HitT<S> h{};
    if (!pair_contact_t(ta, tb, ha, hb, pa, qa, pb, qb, h, -1, winner)) return 0;
    out[0] = h.px; out[1] = h.py; out[2] = h.pz;
    out[3] = h.nx; out[4] = h.ny; out[5] = h.nz;
    out[6] = h.depth;
    return 1;
}

template <class S> BodyT<S> body_from(const S *a) {
    BodyT<S> b;
    static_assert(sizeof(BodyT<S>) == 20 * sizeof(S), "BodyT layout");
    __builtin_memcpy(&b, a, sizeof(b));
    return b;
}

template <class S> ContactT<S> contact_from(const S *c) {
    ContactT<S> k;
    for (int i = 0; i < 3; ++i) { k.p[i] = c[i]; k.n[i] = c[3 + i]; }
    k.depth = c[6];
    return k;
}

template <class S> int project_fwd(const S *c, S *a, S *b, S mu, S *lam, S dt) {
    BodyT<S> A = body_from(a), B = body_from(b);
    const bool r = project_t(contact_from(c), A, B, mu, *lam, dt);
    __builtin_memcpy(a, &A, sizeof(A));
    __builtin_memcpy(b, &B, sizeof(B));
    return r ? 1 : 0;
}

// bar[14] = adjoints of p xyz, q wxyz, v xyz, w xyz, inverse mass.
template <class S> BodyBar<S> bar_from(const S *x) {
    BodyBar<S> b;
    static_assert(sizeof(BodyBar<S>) == 14 * sizeof(S), "BodyBar layout");
    __builtin_memcpy(&b, x, sizeof(b));
    return b;
}

}  // namespace

extern "C" {

// Contact of one pair through the templated replay (full search); winner
// receives the box-box axis code.  Same argument order as hm_pair_contact.
int ha_pair_contact_f(uint32_t ta, uint32_t tb, const float *ha, const float *hb, const float *pa, const float *pb,
                      const float *qa, const float *qb, float *out, int *winner) {
    return contact_fwd(ta, tb, ha, hb, pa, pb, qa, qb, out, winner);
}
int ha_pair_contact_d(uint32_t ta, uint32_t tb, const double *ha, const double *hb, const double *pa,
                      const double *pb, const double *qa, const double *qb, double *out, int *winner) {
    return contact_fwd(ta, tb, ha, hb, pa, pb, qa, qb, out, winner);
}
// Contact geometry adjoint: ga[7], gb[7] (+=) = J^T hbar, nt pose tangents per pass (1, 2, 7 or 14).
int ha_contact_adj_d(uint32_t ta, uint32_t tb, const double *ha, const double *hb, const double *pa,
                     const double *pb, const double *qa, const double *qb, const double *hbar, double *ga,
                     double *gb, int nt) {
    switch (nt) {
        case 1: return contact_adj<double, 1>(ta, tb, ha, hb, pa, qa, pb, qb, hbar, ga, gb) ? 1 : 0;
        case 2: return contact_adj<double, 2>(ta, tb, ha, hb, pa, qa, pb, qb, hbar, ga, gb) ? 1 : 0;
        case 7: return contact_adj<double, 7>(ta, tb, ha, hb, pa, qa, pb, qb, hbar, ga, gb) ? 1 : 0;
        default: return contact_adj<double, 14>(ta, tb, ha, hb, pa, qa, pb, qb, hbar, ga, gb) ? 1 : 0;
    }
}
// Feature word of one pair's contact through the templated replay (float or
// double); returns 1 on hit.
int ha_pair_feat_f(uint32_t ta, uint32_t tb, const float *ha, const float *hb, const float *pa, const float *pb,
                   const float *qa, const float *qb, uint32_t *feat) {
    HitT<float> h{};
    if (!pair_contact_t(ta, tb, ha, hb, pa, qa, pb, qb, h, -1, nullptr)) return 0;
    *feat = h.feat;
    return 1;
}
int ha_pair_feat_d(uint32_t ta, uint32_t tb, const double *ha, const double *hb, const double *pa,
                   const double *pb, const double *qa, const double *qb, uint32_t *feat) {
    HitT<double> h{};
    if (!pair_contact_t(ta, tb, ha, hb, pa, qa, pb, qb, h, -1, nullptr)) return 0;
    *feat = h.feat;
    return 1;
}
// Written-out geometry adjoint for the contact kind the feature word names:
// ga[7], gb[7] (+=) = J^T hbar.
void ha_geom_adj_d(uint32_t ta, uint32_t tb, uint32_t feat, const double *ha, const double *hb, const double *pa,
                   const double *pb, const double *qa, const double *qb, const double *hbar, double *ga,
                   double *gb) {
    geom_adj(ta, tb, feat, ha, hb, pa, qa, pb, qb, hbar, ga, gb);
}
void ha_geom_adj_f(uint32_t ta, uint32_t tb, uint32_t feat, const float *ha, const float *hb, const float *pa,
                   const float *pb, const float *qa, const float *qb, const float *hbar, float *ga, float *gb) {
    geom_adj(ta, tb, feat, ha, hb, pa, qa, pb, qb, hbar, ga, gb);
}
int ha_contact_adj_f(uint32_t ta, uint32_t tb, const float *ha, const float *hb, const float *pa, const float *pb,
                     const float *qa, const float *qb, const float *hbar, float *ga, float *gb) {
    return contact_adj<float, 2>(ta, tb, ha, hb, pa, qa, pb, qb, hbar, ga, gb) ? 1 : 0;
}
// Contact projection on 20-value bodies (hm_project_contact layout), in place.
int ha_project_f(const float *c, float *a, float *b, float mu, float *lam, float dt) {
    return project_fwd(c, a, b, mu, lam, dt);
}
int ha_project_d(const double *c, double *a, double *b, double mu, double *lam, double dt) {
    return project_fwd(c, a, b, mu, lam, dt);
}
// Adjoint of the projection from inputs (c, a, b, lam): abar/bbar[14] hold
// the output adjoints on entry and the input adjoints on return; *lamb
// likewise; cbar[7] receives the contact's adjoint.
void ha_project_adj_d(const double *c, const double *a, const double *b, double mu, double lam, double dt,
                      double *abar, double *bbar, double *lamb, double *cbar) {
    BodyBar<double> Ab = bar_from(abar), Bb = bar_from(bbar);
    ContactT<double> cb{};
    project_adj(contact_from(c), body_from(a), body_from(b), mu, lam, dt, Ab, Bb, *lamb, cb);
    __builtin_memcpy(abar, &Ab, sizeof(Ab));
    __builtin_memcpy(bbar, &Bb, sizeof(Bb));
    for (int i = 0; i < 3; ++i) { cbar[i] = cb.p[i]; cbar[3 + i] = cb.n[i]; }
    cbar[6] = cb.depth;
}
// One integrator substep on 13 values (hm_integrate_body layout) and its adjoint.
void ha_integrate_f(float *kin, const float *g, float h) {
    BodyT<float> b{};
    __builtin_memcpy(&b, kin, 13 * sizeof(float));
    integrate_t(b, g, h);
    __builtin_memcpy(kin, &b, 13 * sizeof(float));
}
void ha_integrate_d(double *kin, const double *g, double h) {
    BodyT<double> b{};
    __builtin_memcpy(&b, kin, 13 * sizeof(double));
    integrate_t(b, g, h);
    __builtin_memcpy(kin, &b, 13 * sizeof(double));
}
void ha_integrate_adj_d(const double *kin0, const double *g, double h, double *bar) {
    BodyT<double> b{};
    __builtin_memcpy(&b, kin0, 13 * sizeof(double));
    BodyBar<double> bb{};
    __builtin_memcpy(&bb, bar, 13 * sizeof(double));
    integrate_adj(b, g, h, bb);
    __builtin_memcpy(bar, &bb, 13 * sizeof(double));
}
void ha_renorm_d(double *q) { renorm_t(q); }
void ha_renorm_adj_d(const double *q, double *qb) { renorm_adj(q, qb); }

}  // extern "C"

// ---- a host reference of whole frames, float64 or forward-mode -------------
namespace {

template <class S> S val(S x) { return x; }
template <class S, int N> S val(const Dual<S, N> &x) { return x.v; }

// frames of one scene: st[13][n] SoA state (pos, rot, vel, ang), im[n],
// I6[6][n], per-body shape type and half extents hs[3][n].  The frame of the
// stress driver, the pair set from binary16 boxes of the float-rounded state,
// contacts in pair order, reference contact order for <= 256 contacts (or
// forced) and Jones-Plassmann colours otherwise (with the reference's
// fallbacks).
template <class T, class S>
void host_frames(int n, T *st, const T *im, const double *I6, const uint32_t *stype, const double *hs_d,
                 const uint32_t *flags, const double *g, double dt, int frames, int iters, double mu, int force_serial,
                 uint32_t max_pairs, uint32_t max_contacts) {
    auto F = [&](int f, int i) -> T & { return st[static_cast<size_t>(f) * n + i]; };
    std::vector<S> hs(3 * static_cast<size_t>(n));
    for (size_t i = 0; i < hs.size(); ++i) hs[i] = static_cast<S>(hs_d[i]);
    auto body = [&](int i) {
        BodyT<T> b;
        for (int k = 0; k < 3; ++k) { b.p[k] = F(k, i); b.v[k] = F(7 + k, i); b.w[k] = F(10 + k, i); }
        for (int k = 0; k < 4; ++k) b.q[k] = F(3 + k, i);
        b.im = im[i];
        for (int k = 0; k < 6; ++k) b.I[k] = T(static_cast<S>(I6[static_cast<size_t>(k) * n + i]));
        return b;
    };
    auto put = [&](int i, const BodyT<T> &b) {
        for (int k = 0; k < 3; ++k) { F(k, i) = b.p[k]; F(7 + k, i) = b.v[k]; F(10 + k, i) = b.w[k]; }
        for (int k = 0; k < 4; ++k) F(3 + k, i) = b.q[k];
    };
    const T gT[3] = {T(static_cast<S>(g[0])), T(static_cast<S>(g[1])), T(static_cast<S>(g[2]))};
    const T dtT = T(static_cast<S>(dt)), muT = T(static_cast<S>(mu));
    for (int f = 0; f < frames; ++f) {
        for (int i = 0; i < n; ++i) {
            if (val(im[i]) == 0.0 || (flags[i] & kFlagSleeping)) continue;
            BodyT<T> b = body(i);
            integrate_t(b, gT, dtT);
            put(i, b);
        }
        // pairs from the binary16 boxes of the float-rounded state
        std::vector<float> bx(6 * n);
        for (int i = 0; i < n; ++i) {
            float ex, ey, ez;
            leaf_extent(stype[i], static_cast<float>(hs[i]), static_cast<float>(hs[n + i]),
                        static_cast<float>(hs[2 * n + i]), static_cast<float>(val(F(3, i))),
                        static_cast<float>(val(F(4, i))), static_cast<float>(val(F(5, i))),
                        static_cast<float>(val(F(6, i))), ex, ey, ez);
            const float p[3] = {static_cast<float>(val(F(0, i))), static_cast<float>(val(F(1, i))),
                                static_cast<float>(val(F(2, i)))};
            const float e[3] = {ex, ey, ez};
            for (int k = 0; k < 3; ++k) {
                bx[6 * i + k] = f16_to_f32(f32_to_f16(p[k] - e[k]));
                bx[6 * i + 3 + k] = f16_to_f32(f32_to_f16(p[k] + e[k]));
            }
        }
        std::vector<std::pair<int, int>> pairs;
        for (int a = 0; a < n; ++a)
            for (int b = a + 1; b < n; ++b) {
                const float *A = &bx[6 * a], *B = &bx[6 * b];
                const bool sep = B[0] > A[3] || B[3] < A[0] || B[1] > A[4] || B[4] < A[1] || B[2] > A[5] ||
                                 B[5] < A[2];
                if (!sep && pairs.size() < max_pairs) pairs.emplace_back(a, b);
            }
        std::vector<ContactT<T>> con;
        std::vector<std::pair<int, int>> cab;
        for (auto &pr : pairs) {
            const int a = pr.first, b = pr.second;
            const S ha[3] = {hs[a], hs[n + a], hs[2 * n + a]}, hb[3] = {hs[b], hs[n + b], hs[2 * n + b]};
            const T pa[3] = {F(0, a), F(1, a), F(2, a)}, pb[3] = {F(0, b), F(1, b), F(2, b)};
            const T qa[4] = {F(3, a), F(4, a), F(5, a), F(6, a)}, qb[4] = {F(3, b), F(4, b), F(5, b), F(6, b)};
            HitT<T> h;
            if (!pair_contact_t(stype[a], stype[b], ha, hb, pa, qa, pb, qb, h, -1, nullptr)) continue;
            if (con.size() >= max_contacts) continue;
            ContactT<T> c;
            c.p[0] = h.px; c.p[1] = h.py; c.p[2] = h.pz;
            c.n[0] = h.nx; c.n[1] = h.ny; c.n[2] = h.nz;
            c.depth = h.depth;
            con.push_back(c);
            cab.emplace_back(a, b);
        }
        const int nc = static_cast<int>(con.size());
        if (nc == 0) continue;
        // solve order: reference order, or colours
        std::vector<int> order(nc);
        for (int k = 0; k < nc; ++k) order[k] = k;
        if (!(force_serial || nc <= 256)) {
            std::vector<int> col(nc, -1);
            std::vector<uint64_t> mask(n, 0ull);
            int left = nc, rounds = 0;
            bool over = false;
            while (left > 0 && rounds < 256 && !over) {
                ++rounds;
                std::vector<uint64_t> best(n, 0ull);
                for (int k = 0; k < nc; ++k) {
                    if (col[k] >= 0) continue;
                    const uint64_t pk = jp_priority(static_cast<uint32_t>(k));
                    for (int e : {cab[k].first, cab[k].second})
                        if (val(im[e]) != 0.0 && pk > best[e]) best[e] = pk;
                }
                std::vector<int> win;
                for (int k = 0; k < nc; ++k) {
                    if (col[k] >= 0) continue;
                    const uint64_t pk = jp_priority(static_cast<uint32_t>(k));
                    bool ok = true;
                    for (int e : {cab[k].first, cab[k].second})
                        if (val(im[e]) != 0.0 && best[e] != pk) ok = false;
                    if (ok) win.push_back(k);
                }
                for (int k : win) {
                    uint64_t used = 0;
                    for (int e : {cab[k].first, cab[k].second})
                        if (val(im[e]) != 0.0) used |= mask[e];
                    if (used == ~0ull) { over = true; continue; }
                    const int c = __builtin_ctzll(~used);
                    col[k] = c;
                    for (int e : {cab[k].first, cab[k].second})
                        if (val(im[e]) != 0.0) mask[e] |= 1ull << c;
                    --left;
                }
            }
            if (left == 0 && !over) {
                std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return col[x] < col[y]; });
            }
        }
        std::vector<T> lam(nc, T(S(0)));
        for (int it = 0; it < iters; ++it) {
            for (int k : order) {
                BodyT<T> A = body(cab[k].first), B = body(cab[k].second);
                if (project_t(con[k], A, B, muT, lam[k], dtT)) {
                    if (val(A.im) != 0.0) put(cab[k].first, A);
                    if (val(B.im) != 0.0) put(cab[k].second, B);
                }
            }
            for (int i = 0; i < n; ++i) {
                T q[4] = {F(3, i), F(4, i), F(5, i), F(6, i)};
                renorm_t(q);
                for (int k = 0; k < 4; ++k) F(3 + k, i) = q[k];
            }
        }
    }
}

}  // namespace

extern "C" {

// float64 frames of one scene, state in place (13 x n SoA).
void ha_sim_d(int n, double *st, const double *im, const double *I6, const uint32_t *stype, const double *hs,
              const uint32_t *flags, const double *g, double dt, int frames, int iters, double mu, int force_serial,
              uint32_t max_pairs, uint32_t max_contacts) {
    host_frames<double, double>(n, st, im, I6, stype, hs, flags, g, dt, frames, iters, mu, force_serial, max_pairs,
                                max_contacts);
}

// Forward mode: the same frames carrying one tangent; st_dot[13 x n] and
// im_dot[n] seed it, st_dot receives the final tangent.
void ha_sim_dir_d(int n, double *st, double *st_dot, const double *im, const double *im_dot, const double *I6,
                  const uint32_t *stype, const double *hs, const uint32_t *flags, const double *g, double dt,
                  int frames, int iters, double mu, int force_serial, uint32_t max_pairs, uint32_t max_contacts) {
    typedef Dual<double, 1> D;
    std::vector<D> s(13 * static_cast<size_t>(n)), m(n);
    for (size_t i = 0; i < s.size(); ++i) { s[i] = D(st[i]); s[i].d[0] = st_dot[i]; }
    for (int i = 0; i < n; ++i) { m[i] = D(im[i]); m[i].d[0] = im_dot[i]; }
    host_frames<D, double>(n, s.data(), m.data(), I6, stype, hs, flags, g, dt, frames, iters, mu, force_serial,
                           max_pairs, max_contacts);
    for (size_t i = 0; i < s.size(); ++i) { st[i] = s[i].v; st_dot[i] = s[i].d[0]; }
}

// The same in float: values follow the gfx950 kernels' float arithmetic
// (identical operations), so this is the float forward-mode derivative of
// the exact trajectory the kernels compute.
void ha_sim_dir_f(int n, double *st, double *st_dot, const double *im, const double *im_dot, const double *I6,
                  const uint32_t *stype, const double *hs, const uint32_t *flags, const double *g, double dt,
                  int frames, int iters, double mu, int force_serial, uint32_t max_pairs, uint32_t max_contacts) {
    typedef Dual<float, 1> D;
    std::vector<D> s(13 * static_cast<size_t>(n)), m(n);
    for (size_t i = 0; i < s.size(); ++i) { s[i] = D(static_cast<float>(st[i])); s[i].d[0] = static_cast<float>(st_dot[i]); }
    for (int i = 0; i < n; ++i) { m[i] = D(static_cast<float>(im[i])); m[i].d[0] = static_cast<float>(im_dot[i]); }
    host_frames<D, float>(n, s.data(), m.data(), I6, stype, hs, flags, g, dt, frames, iters, mu, force_serial,
                          max_pairs, max_contacts);
    for (size_t i = 0; i < s.size(); ++i) { st[i] = s[i].v; st_dot[i] = s[i].d[0]; }
}

}  // extern "C"

// project_contact_pair against project_contact: the two lanes run as two
// threads that meet at every exchange (each writes its value, both read,
// then both leave before either writes again), over n random contacts
// (seeded; every third with a static second body, every fifth with a lambda
// already accumulated).  Returns the number of contacts whose lambda or any
// body word differs in its bits from project_contact's.
namespace {
struct LaneRendezvous {
    std::mutex m;
    std::condition_variable cv;
    float val[2] = {0.f, 0.f};
    int in1 = 0, in2 = 0;
    long gen1 = 0, gen2 = 0;
    float swap(uint32_t side, float v) {
        std::unique_lock<std::mutex> lk(m);
        val[side] = v;
        long g = gen1;
        if (++in1 == 2) { in1 = 0; ++gen1; cv.notify_all(); } else { cv.wait(lk, [&] { return gen1 != g; }); }
        const float r = val[side ^ 1u];
        g = gen2;
        if (++in2 == 2) { in2 = 0; ++gen2; cv.notify_all(); } else { cv.wait(lk, [&] { return gen2 != g; }); }
        return r;
    }
};
uint32_t f_bits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
}  // namespace

extern "C" int hm_pair_projection_mismatches(int n, uint32_t seed) {
    std::mt19937 rng(seed);
    std::uniform_real_distribution<float> U(-1.f, 1.f);
    auto body = [&](BodyState &s, bool dyn) {
        s.px = U(rng); s.py = U(rng); s.pz = U(rng);
        const float q[4] = {U(rng), U(rng), U(rng), U(rng)};
        const float qn = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
        s.qw = q[0] / qn; s.qx = q[1] / qn; s.qy = q[2] / qn; s.qz = q[3] / qn;
        s.vx = U(rng); s.vy = U(rng); s.vz = U(rng); s.wx = U(rng); s.wy = U(rng); s.wz = U(rng);
        s.im = dyn ? 0.5f + 0.5f * U(rng) : 0.f;
        s.ixx = dyn ? 1.f + 0.5f * U(rng) : 0.f; s.iyy = dyn ? 1.f + 0.5f * U(rng) : 0.f;
        s.izz = dyn ? 1.f + 0.5f * U(rng) : 0.f; s.ixy = dyn ? 0.1f * U(rng) : 0.f;
        s.ixz = dyn ? 0.1f * U(rng) : 0.f; s.iyz = dyn ? 0.1f * U(rng) : 0.f;
    };
    int bad = 0;
    for (int t = 0; t < n; ++t) {
        BodyState A{}, B{};
        body(A, true);
        body(B, t % 3 != 0);
        const float nv[3] = {U(rng), U(rng), U(rng)};
        const float nl = std::sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
        const ContactIn c{U(rng), U(rng), U(rng), nv[0] / nl, nv[1] / nl, nv[2] / nl, 0.05f + 0.05f * U(rng)};
        const float mu = 0.5f, dt = 1.f / 60.f, lam0 = (t % 5 == 0) ? 0.01f : 0.f;
        BodyState A1 = A, B1 = B;
        float l1 = lam0;
        project_contact(c, A1, B1, mu, l1, dt);
        BodyState A2 = A, B2 = B;
        float la = lam0, lb = lam0;
        LaneRendezvous X;
        std::thread other([&] { project_contact_pair(c, B2, 1u, mu, lb, dt, [&](float v) { return X.swap(1u, v); }); });
        project_contact_pair(c, A2, 0u, mu, la, dt, [&](float v) { return X.swap(0u, v); });
        other.join();
        bool ok = f_bits(la) == f_bits(l1) && f_bits(lb) == f_bits(l1);
        const float *a1 = &A1.px, *a2 = &A2.px, *b1 = &B1.px, *b2 = &B2.px;
        for (int k = 0; k < 13; ++k) ok = ok && f_bits(a1[k]) == f_bits(a2[k]) && f_bits(b1[k]) == f_bits(b2[k]);
        bad += ok ? 0 : 1;
    }
    return bad;
}
