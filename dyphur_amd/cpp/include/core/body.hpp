#pragma once
// BodyFlag, BodyParams, BodyView (core/include/core/body.hpp:8-46).
#include "math/math.hpp"

#include <dyphur_hip.h>

#include <cstddef>
#include <cstdint>

namespace dyphur {

namespace BodyFlag {
constexpr uint32_t Static = DY_BODY_STATIC;
constexpr uint32_t Kinematic = DY_BODY_KINEMATIC;
constexpr uint32_t Sleeping = DY_BODY_SLEEPING;
}  // namespace BodyFlag

struct BodyParams {
    Vec3f position = Vec3f::zero();
    Quatf rotation = Quatf::identity();
    Vec3f linear_velocity = Vec3f::zero();
    Vec3f angular_velocity = Vec3f::zero();
    float mass = 1.f;
    Mat3f inertia = Mat3f::identity();  // body frame
    uint32_t shape_handle = 0;
    uint32_t flags = 0;
};

struct BodyView {
    float *pos_x; float *pos_y; float *pos_z;
    float *rot_w; float *rot_x; float *rot_y; float *rot_z;
    float *vel_x; float *vel_y; float *vel_z;
    float *ang_x; float *ang_y; float *ang_z;
    float *inv_mass;
    float *iI_xx; float *iI_yy; float *iI_zz;
    float *iI_xy; float *iI_xz; float *iI_yz;
    uint32_t *shape;
    uint32_t *flags;
    uint32_t n;
};
static_assert(sizeof(BodyView) == sizeof(dy_body_view) && offsetof(BodyView, n) == offsetof(dy_body_view, n),
              "BodyView layout");

inline const dy_body_view *c_view(const BodyView &v) { return reinterpret_cast<const dy_body_view *>(&v); }

}  // namespace dyphur
