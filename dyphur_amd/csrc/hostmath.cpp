// hostmath.cpp — the per-element arithmetic of dy_math.hpp compiled for the
// host (hipcc --cuda-host-only), exported with a C ABI so the CPU test suite
// can compare the exact code the kernels run against the oracle without a GPU.
#include "dy_math.hpp"

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
// quaternions wxyz (4 each).  out[7] = p xyz, n xyz, depth.  Returns 1 on hit.
int hm_pair_contact(uint32_t ta, uint32_t tb, const float *ha, const float *hb, const float *pa, const float *pb,
                    const float *qa, const float *qb, float *out) {
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
    return 1;
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

}  // extern "C"
