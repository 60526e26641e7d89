#pragma once
// XpbdSolver (core/include/core/xpbd_solver.hpp:29-48) over the gfx950 class
// sweeps: <= kSerialThreshold constraints (contacts + joints) or
// set_force_serial gives the reference's serial-order results, larger scenes
// its Jones-Plassmann colour order.  solve() is asynchronous on the stream;
// errors (invalid joints among them) surface at the next solve or
// last_stats().
#include "solver.hpp"

#include <memory>

namespace dyphur {

class XpbdSolver final : public IConstraintSolver {
public:
    explicit XpbdSolver(Stream &s, int n_iters = 10, uint32_t max_contacts = 16384, float friction = 0.5f) {
        dy_xpbd_t h = nullptr;
        check(dy_xpbd_create(s.handle(), n_iters, max_contacts, friction, &h), "XpbdSolver");
        h_ = std::shared_ptr<dy_xpbd_s>(h, [](dy_xpbd_t p) { (void)dy_xpbd_destroy(p); });
    }

    void solve(Stream &s, const ContactView &contacts, const JointView &joints, BodyView bodies,
               float dt) override {
        check(dy_xpbd_solve(h_.get(), s.handle(), c_view(contacts), joints.n ? c_view(joints) : nullptr,
                            c_view(bodies), dt),
              "XpbdSolver::solve");
    }

    static constexpr uint32_t kSerialThreshold = 256;
    void set_force_serial(bool v) { check(dy_xpbd_set_force_serial(h_.get(), v ? 1 : 0), "set_force_serial"); }

    // Extensions: statistics of the last solve and per-contact classes.
    dy_solve_stats last_stats() const {
        dy_solve_stats st{};
        check(dy_xpbd_last_stats(h_.get(), &st), "XpbdSolver::last_stats");
        return st;
    }

private:
    std::shared_ptr<dy_xpbd_s> h_;
};

}  // namespace dyphur
