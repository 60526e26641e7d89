#pragma once
// JointView (core/include/core/articulation.hpp:47-...).  Joints are outside
// the gfx950 solver: XpbdSolver::solve throws for joints.n > 0.
#include <cstdint>

namespace dyphur {

struct JointView {
    uint32_t *body_parent = nullptr;
    uint32_t *body_child = nullptr;
    uint32_t n = 0;
};

}  // namespace dyphur
