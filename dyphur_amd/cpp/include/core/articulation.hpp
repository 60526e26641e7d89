#pragma once
// JointType, JointParams, JointView (core/include/core/articulation.hpp:7-70).
// JointView is the device SoA the solver reads; its layout is dy_joint_view's.
#include "math/math.hpp"

#include <dyphur_hip.h>

#include <cstdint>
#include <type_traits>

namespace dyphur {

enum class JointType : uint8_t {
    Fixed = 0,      // no relative motion
    Revolute = 1,   // rotation about the axis
    Prismatic = 2,  // translation along the axis
    Ball = 3,       // free rotation about the anchor
};

// Host-facing joint description, passed to JointStore::add().
struct JointParams {
    uint32_t body_parent = 0;
    uint32_t body_child = 0;
    Vec3f anchor_parent = Vec3f::zero();  // parent body frame
    Vec3f anchor_child = Vec3f::zero();   // child body frame
    Vec3f axis = {0.f, 1.f, 0.f};         // parent body frame (revolute, prismatic)
    float limit_lo = 1.f;                 // lo > hi disables the limits
    float limit_hi = -1.f;
    float target_pos = 0.f;  // PD motor target (rad or m) and gains
    float target_vel = 0.f;
    float stiffness = 0.f;
    float damping = 0.f;
    float compliance_pos = 0.f;  // XPBD compliance (inverse stiffness)
    float compliance_ang = 0.f;
    JointType type = JointType::Fixed;
};

struct JointView {
    uint32_t *body_parent = nullptr;
    uint32_t *body_child = nullptr;
    float *anchor_px = nullptr, *anchor_py = nullptr, *anchor_pz = nullptr;
    float *anchor_cx = nullptr, *anchor_cy = nullptr, *anchor_cz = nullptr;
    float *axis_px = nullptr, *axis_py = nullptr, *axis_pz = nullptr;
    float *limit_lo = nullptr;
    float *limit_hi = nullptr;
    float *target_pos = nullptr;
    float *target_vel = nullptr;
    float *stiffness = nullptr;
    float *damping = nullptr;
    float *compliance_pos = nullptr;
    float *compliance_ang = nullptr;
    uint8_t *type = nullptr;
    uint32_t n = 0;
};
static_assert(sizeof(JointView) == sizeof(dy_joint_view), "JointView layout");
static_assert(std::is_trivially_copyable_v<JointView>, "JointView is captured by value");

inline const dy_joint_view *c_view(const JointView &v) { return reinterpret_cast<const dy_joint_view *>(&v); }

}  // namespace dyphur
