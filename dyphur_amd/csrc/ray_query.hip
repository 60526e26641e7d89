// ray_query.hip — closest-hit ray queries over the broadphase's LBVH for
// gfx950 (core/src/ray_query.cpp:12-195, core/include/core/ray_query.hpp).
//
// One ray per lane: depth-first walk from the root with a 64-entry stack,
// right child pushed first (the left subtree is visited first, which fixes
// which of two equally distant hits is reported), node boxes read as the
// 16-byte binary16 NodeBox records, leaves tested exactly (oriented box slab
// test in the body frame, analytic sphere).  Arithmetic in the reference's
// operation order; compiled without contraction, so results are bit-equal to
// a CPU evaluation of the same steps.
#include "dy_internal.hpp"

namespace dy {

struct RqBox { float mn[3], mx[3]; };

__device__ __forceinline__ bool rq_slab(const float o[3], const float inv[3], const float mn[3], const float mx[3],
                                        float &tmin, float &tmax) {
    const float t1x = (mn[0] - o[0]) * inv[0], t2x = (mx[0] - o[0]) * inv[0];
    const float t1y = (mn[1] - o[1]) * inv[1], t2y = (mx[1] - o[1]) * inv[1];
    const float t1z = (mn[2] - o[2]) * inv[2], t2z = (mx[2] - o[2]) * inv[2];
    tmin = t1x < t2x ? t1x : t2x;
    tmax = t1x < t2x ? t2x : t1x;
    const float ylo = t1y < t2y ? t1y : t2y, yhi = t1y < t2y ? t2y : t1y;
    tmin = tmin > ylo ? tmin : ylo;
    tmax = tmax < yhi ? tmax : yhi;
    const float zlo = t1z < t2z ? t1z : t2z, zhi = t1z < t2z ? t2z : t1z;
    tmin = tmin > zlo ? tmin : zlo;
    tmax = tmax < zhi ? tmax : zhi;
    return true;
}

// v rotated by the conjugate of q: t = 2 u x v, v + w t + u x t, u = -xyz
// (core/include/core/math/quat.hpp:31,39-42).
__device__ __forceinline__ void rq_rotate_inv(float w, float x, float y, float z, const float v[3], float o[3]) {
    const float ux = -x, uy = -y, uz = -z;
    const float tx = (uy * v[2] - uz * v[1]) * 2.f;
    const float ty = (uz * v[0] - ux * v[2]) * 2.f;
    const float tz = (ux * v[1] - uy * v[0]) * 2.f;
    o[0] = (v[0] + tx * w) + (uy * tz - uz * ty);
    o[1] = (v[1] + ty * w) + (uz * tx - ux * tz);
    o[2] = (v[2] + tz * w) + (ux * ty - uy * tx);
}

__device__ __forceinline__ void rq_leaf(uint32_t bi, const float o[3], const float d[3], const dy_body_view &bv,
                                        const dy_shape_view &sv, float &best_t, uint32_t &best) {
    const uint32_t sh = bv.shape[bi];
    const float p[3] = {bv.pos_x[bi], bv.pos_y[bi], bv.pos_z[bi]};
    float t = -1.f;
    if (sv.type[sh] == kSphere) {
        const float r = sv.half_x[sh];
        const float ocx = o[0] - p[0], ocy = o[1] - p[1], ocz = o[2] - p[2];
        const float a = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        const float b = d[0] * ocx + d[1] * ocy + d[2] * ocz;
        const float c = ocx * ocx + ocy * ocy + ocz * ocz - r * r;
        const float disc = b * b - a * c;
        if (disc >= 0.f) {
            const float sq = sqrtf(disc);
            t = (-b - sq) / a;
            if (t < 0.f) t = (-b + sq) / a;
            if (!(t >= 0.f && t < best_t)) t = -1.f;
        }
    } else {
        const float oc[3] = {o[0] - p[0], o[1] - p[1], o[2] - p[2]};
        float lo[3], ld[3];
        const float qw = bv.rot_w[bi], qx = bv.rot_x[bi], qy = bv.rot_y[bi], qz = bv.rot_z[bi];
        rq_rotate_inv(qw, qx, qy, qz, oc, lo);
        rq_rotate_inv(qw, qx, qy, qz, d, ld);
        const float inv[3] = {1.f / ld[0], 1.f / ld[1], 1.f / ld[2]};
        const float h[3] = {sv.half_x[sh], sv.half_y[sh], sv.half_z[sh]};
        const float mn[3] = {-h[0], -h[1], -h[2]};
        float tmin, tmax;
        rq_slab(lo, inv, mn, h, tmin, tmax);
        if (!(tmax < 0.f || tmin > tmax)) {
            t = tmin >= 0.f ? tmin : tmax;
            if (!(t >= 0.f && t < best_t)) t = -1.f;
        }
    }
    if (t >= 0.f) {
        best_t = t;
        best = bi;
    }
}

__global__ __launch_bounds__(256) void ray_query_kernel(dy_bvh_view bvh, dy_body_view bv, dy_shape_view sv,
                                                        const dy_ray *__restrict__ rays, dy_ray_hit *__restrict__ hits,
                                                        uint32_t n_rays) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_rays) return;
    const dy_ray r = rays[k];
    const float o[3] = {r.ox, r.oy, r.oz}, d[3] = {r.dx, r.dy, r.dz};
    float best_t = r.t_max;
    uint32_t best = 0xffffffffu;
    const float inv[3] = {1.f / r.dx, 1.f / r.dy, 1.f / r.dz};
    const uint32_t n = bvh.n;
    if (n == 1) {
        rq_leaf(0u, o, d, bv, sv, best_t, best);
    } else if (n > 1) {
        int32_t stack[64];
        int top = 0;
        stack[top++] = bvh.root[0];
        const uint32_t n_int = n - 1u;
        while (top > 0) {
            const int32_t node = stack[--top];
            const uint4 nb = *reinterpret_cast<const uint4 *>(&bvh.aabb[node]);
            const float mn[3] = {f16_to_f32(static_cast<uint16_t>(nb.x & 0xffffu)),
                                 f16_to_f32(static_cast<uint16_t>(nb.x >> 16)),
                                 f16_to_f32(static_cast<uint16_t>(nb.y & 0xffffu))};
            const float mx[3] = {f16_to_f32(static_cast<uint16_t>(nb.y >> 16)),
                                 f16_to_f32(static_cast<uint16_t>(nb.z & 0xffffu)),
                                 f16_to_f32(static_cast<uint16_t>(nb.z >> 16))};
            float tmin, tmax;
            rq_slab(o, inv, mn, mx, tmin, tmax);
            if (!(tmax >= 0.f && tmax >= tmin && tmin < best_t)) continue;
            if (static_cast<uint32_t>(node) >= n_int) {
                rq_leaf(bvh.sorted_idx[node - static_cast<int32_t>(n_int)], o, d, bv, sv, best_t, best);
            } else if (top + 1 < 64) {
                stack[top++] = bvh.right[node];
                stack[top++] = bvh.left[node];
            }
        }
    }
    hits[k] = dy_ray_hit{best_t, best};
}

}  // namespace dy

using namespace dy;

extern "C" {

int dy_ray_query(dy_stream_t s, const dy_bvh_view *bvh, const dy_body_view *bodies, const dy_shape_view *shapes,
                 const dy_ray *d_rays, dy_ray_hit *d_hits, uint32_t n_rays) {
    if (!bvh || !bodies || !shapes || (n_rays && (!d_rays || !d_hits))) return DY_ERR_INVALID_ARG;
    if (n_rays == 0) return DY_OK;
    if (bvh->n > bodies->n) return DY_ERR_INVALID_ARG;
    ray_query_kernel<<<ceil_div(n_rays, 256), 256, 0, hs(s)>>>(*bvh, *bodies, *shapes, d_rays, d_hits, n_rays);
    DY_LAUNCH_CHECK();
    return DY_OK;
}

}  // extern "C"
