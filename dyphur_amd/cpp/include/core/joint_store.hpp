#pragma once
// JointStore (core/include/core/joint_store.hpp:17-59, core/src/articulation.cpp):
// host arrays filled by add(), one asynchronous upload of the device SoA,
// set_targets() to update the PD targets.  add() past the capacity throws.
#include "articulation.hpp"

#include <compute/buffer.hpp>
#include <compute/stream.hpp>

#include <algorithm>
#include <vector>

namespace dyphur {

class JointStore {
public:
    JointStore() = default;
    JointStore(Stream &s, uint32_t capacity)
        : cap_(capacity), d_u_(s, static_cast<size_t>(capacity) * 2), d_f_(s, static_cast<size_t>(capacity) * kF),
          d_t_(s, capacity), h_u_(static_cast<size_t>(capacity) * 2, 0u), h_f_(static_cast<size_t>(capacity) * kF, 0.f),
          h_t_(capacity, 0u) {}
    JointStore(const JointStore &) = delete;
    JointStore &operator=(const JointStore &) = delete;
    JointStore(JointStore &&) = default;
    JointStore &operator=(JointStore &&) = default;

    uint32_t add(const JointParams &p) {
        if (count_ >= cap_) throw Error(DY_ERR_INVALID_ARG, "JointStore::add: capacity exceeded");
        const uint32_t i = count_++;
        h_u_[i] = p.body_parent;
        h_u_[cap_ + i] = p.body_child;
        const float v[kF] = {p.anchor_parent.x, p.anchor_parent.y, p.anchor_parent.z,
                             p.anchor_child.x,  p.anchor_child.y,  p.anchor_child.z,
                             p.axis.x,          p.axis.y,          p.axis.z,
                             p.limit_lo,        p.limit_hi,        p.target_pos,
                             p.target_vel,      p.stiffness,       p.damping,
                             p.compliance_pos,  p.compliance_ang};
        for (uint32_t f = 0; f < kF; ++f) h_f_[static_cast<size_t>(f) * cap_ + i] = v[f];
        h_t_[i] = static_cast<uint8_t>(p.type);
        return i;
    }

    void upload() {
        if (count_ == 0) return;
        d_u_.upload(h_u_.data());
        d_f_.upload(h_f_.data());
        d_t_.upload(h_t_.data());
    }

    // New PD targets for all count() joints (host arrays), uploaded.
    void set_targets(const float *pos, const float *vel) {
        std::copy(pos, pos + count_, h_f_.begin() + static_cast<size_t>(kTargetPos) * cap_);
        std::copy(vel, vel + count_, h_f_.begin() + static_cast<size_t>(kTargetPos + 1) * cap_);
        d_f_.upload(h_f_.data());
    }

    JointView view() noexcept {
        JointView v;
        uint32_t *u = d_u_.data();
        float *f = d_f_.data();
        const size_t c = cap_;
        v.body_parent = u;
        v.body_child = u ? u + c : nullptr;
        float **fields[kF] = {&v.anchor_px, &v.anchor_py, &v.anchor_pz, &v.anchor_cx, &v.anchor_cy, &v.anchor_cz,
                              &v.axis_px,   &v.axis_py,   &v.axis_pz,   &v.limit_lo,  &v.limit_hi,  &v.target_pos,
                              &v.target_vel, &v.stiffness, &v.damping, &v.compliance_pos, &v.compliance_ang};
        for (uint32_t k = 0; k < kF; ++k) *fields[k] = f ? f + k * c : nullptr;
        v.type = d_t_.data();
        v.n = count_;
        return v;
    }

    uint32_t count() const noexcept { return count_; }
    uint32_t capacity() const noexcept { return cap_; }

private:
    static constexpr uint32_t kF = 17, kTargetPos = 11;
    uint32_t cap_ = 0, count_ = 0;
    Buffer<uint32_t> d_u_;  // body_parent, body_child
    Buffer<float> d_f_;     // 17 float fields in JointView order, capacity each
    Buffer<uint8_t> d_t_;   // type
    std::vector<uint32_t> h_u_;
    std::vector<float> h_f_;
    std::vector<uint8_t> h_t_;
};

}  // namespace dyphur
