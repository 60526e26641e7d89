#pragma once
// One frame of the drivers' per-substep sequence (examples/manipulator_pick/main.cpp:683-760,
// examples/stress_test_blocks/main.cpp:264-273): for each of params.n_substeps
// substeps of dt / n_substeps, integrate one substep, build the LBVH and query
// pairs, run the narrowphase on the device-resident pair count, and solve.
// Asynchronous: nothing waits on the host (errors surface at the next
// blocking call).  With n_substeps = 1 this is the stress driver's frame.
#include "broadphase.hpp"
#include "integrator.hpp"
#include "narrowphase.hpp"
#include "xpbd_solver.hpp"

namespace dyphur {

inline void step(Stream &s, BodyView bodies, const ShapeView &shapes, Broadphase &bp, Narrowphase &np,
                 XpbdSolver &solver, const AABB &bounds, const IntegratorParams &params) {
    const int n_sub = params.n_substeps > 0 ? params.n_substeps : 1;
    const float sub_dt = params.dt / static_cast<float>(n_sub);
    const IntegratorParams one{params.gravity, sub_dt, 1};
    for (int k = 0; k < n_sub; ++k) {
        integrate(s, bodies, one);
        bp.build_and_query(s, bodies, shapes, bounds);
        check(dy_narrowphase_run_device_count(np.handle(), s.handle(),
                                              reinterpret_cast<const dy_contact_pair *>(bp.pairs_ptr()),
                                              bp.count_ptr(), bp.max_pairs(), c_view(bodies), c_view(shapes)),
              "step: narrowphase");
        solver.solve(s, np.contacts(), JointView{}, bodies, sub_dt);
    }
}

// FNV-1a-64 over pos_x, pos_y, pos_z, rot_w, rot_x, rot_y, rot_z of all
// bodies (examples/stress_test_blocks/main.cpp:34-41,301-311); blocking.
inline uint64_t state_hash(Stream &s, const BodyView &bodies) {
    uint64_t h = 0;
    check(dy_state_hash_fnv1a64(s.handle(), c_view(bodies), &h), "state_hash");
    return h;
}

}  // namespace dyphur
