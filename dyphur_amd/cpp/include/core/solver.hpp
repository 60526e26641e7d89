#pragma once
// IConstraintSolver (core/include/core/solver.hpp:9-21).
#include "articulation.hpp"
#include "body.hpp"
#include "contact.hpp"

#include <compute/stream.hpp>

namespace dyphur {

class IConstraintSolver {
public:
    virtual ~IConstraintSolver() = default;
    virtual void solve(Stream &s, const ContactView &contacts, const JointView &joints, BodyView bodies,
                       float dt) = 0;
};

}  // namespace dyphur
