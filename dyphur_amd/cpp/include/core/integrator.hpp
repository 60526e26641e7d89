#pragma once
// IntegratorParams, integrate (core/include/core/integrator.hpp:7-25): all
// substeps of every body in one asynchronous launch.
#include "body.hpp"

#include <compute/stream.hpp>

namespace dyphur {

struct IntegratorParams {
    Vec3f gravity = {0.f, -9.81f, 0.f};
    float dt = 1.f / 60.f;
    int n_substeps = 4;
};

inline void integrate(Stream &s, BodyView bodies, const IntegratorParams &params) {
    check(dy_integrate(s.handle(), c_view(bodies), params.gravity.x, params.gravity.y, params.gravity.z, params.dt,
                       params.n_substeps),
          "integrate");
}

}  // namespace dyphur
