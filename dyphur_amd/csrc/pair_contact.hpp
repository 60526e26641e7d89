// pair_contact.hpp — contact of one body pair, shared by the per-stage
// narrowphase and the batched frame kernel (core/src/narrowphase.cpp:920-1000:
// dispatch on the shape types; sphere-box with the sphere as A negates the
// normal, :963).  Hull and mesh types set `unsupported`.
#pragma once
#include "dy_internal.hpp"

namespace dy {

__device__ __forceinline__ bool pair_contact(const dy_body_view &bv, const dy_shape_view &sv, uint32_t ia,
                                             uint32_t ib, Hit &h, uint32_t &unsupported) {
    const uint32_t sa = bv.shape[ia], sb = bv.shape[ib];
    const uint32_t ta = sv.type[sa], tb = sv.type[sb];
    if (ta > kSphere || tb > kSphere) {
        unsupported = 1u;
        return false;
    }
    const float ax = bv.pos_x[ia], ay = bv.pos_y[ia], az = bv.pos_z[ia];
    const float bx = bv.pos_x[ib], by = bv.pos_y[ib], bz = bv.pos_z[ib];
    if (ta == kSphere && tb == kSphere) return sphere_sphere(ax, ay, az, sv.half_x[sa], bx, by, bz, sv.half_x[sb], h);
    if (ta == kSphere) {
        const M3 E = body_axes(bv.rot_w[ib], bv.rot_x[ib], bv.rot_y[ib], bv.rot_z[ib]);
        const float he[3] = {sv.half_x[sb], sv.half_y[sb], sv.half_z[sb]};
        if (!sphere_box(ax, ay, az, sv.half_x[sa], bx, by, bz, E, he, h)) return false;
        h.nx = -h.nx; h.ny = -h.ny; h.nz = -h.nz;
        return true;
    }
    if (tb == kSphere) {
        const M3 E = body_axes(bv.rot_w[ia], bv.rot_x[ia], bv.rot_y[ia], bv.rot_z[ia]);
        const float he[3] = {sv.half_x[sa], sv.half_y[sa], sv.half_z[sa]};
        return sphere_box(bx, by, bz, sv.half_x[sb], ax, ay, az, E, he, h);
    }
    const M3 A = body_axes(bv.rot_w[ia], bv.rot_x[ia], bv.rot_y[ia], bv.rot_z[ia]);
    const M3 B = body_axes(bv.rot_w[ib], bv.rot_x[ib], bv.rot_y[ib], bv.rot_z[ib]);
    const float ah[3] = {sv.half_x[sa], sv.half_y[sa], sv.half_z[sa]};
    const float bh[3] = {sv.half_x[sb], sv.half_y[sb], sv.half_z[sb]};
    return box_box(ax, ay, az, A, ah, bx, by, bz, B, bh, h);
}

}  // namespace dy
