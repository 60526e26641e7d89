#pragma once
// BodyStore (core/include/core/body_store.hpp:10-73, core/src/body_store.cpp:28-109):
// host shadow arrays filled by add(), one asynchronous bulk upload, a
// device SoA view.  add() derives inv_mass and the inverse body-frame inertia
// exactly as the reference does (dy_body_dynamics_host).
#include "body.hpp"

#include <compute/buffer.hpp>
#include <compute/stream.hpp>

#include <vector>

namespace dyphur {

class BodyStore {
public:
    BodyStore() = default;
    BodyStore(Stream &s, uint32_t capacity)
        : cap_(capacity), d_f_(s, static_cast<size_t>(capacity) * 20), d_u_(s, static_cast<size_t>(capacity) * 2),
          h_f_(static_cast<size_t>(capacity) * 20, 0.f), h_u_(static_cast<size_t>(capacity) * 2, 0u) {}
    BodyStore(const BodyStore &) = delete;
    BodyStore &operator=(const BodyStore &) = delete;
    BodyStore(BodyStore &&) = default;
    BodyStore &operator=(BodyStore &&) = default;

    uint32_t add(const BodyParams &p) {
        if (count_ >= cap_) throw Error(DY_ERR_INVALID_ARG, "BodyStore::add: capacity exceeded");
        const uint32_t i = count_++;
        float inv_mass = 0.f, iI[6] = {};
        uint32_t flags = 0;
        check(dy_body_dynamics_host(1, &p.mass, p.inertia.data_.data(), &p.flags, &inv_mass, iI, &flags),
              "BodyStore::add");
        const float vals[20] = {p.position.x, p.position.y, p.position.z,
                                p.rotation.w, p.rotation.xyz.x, p.rotation.xyz.y, p.rotation.xyz.z,
                                p.linear_velocity.x, p.linear_velocity.y, p.linear_velocity.z,
                                p.angular_velocity.x, p.angular_velocity.y, p.angular_velocity.z,
                                inv_mass, iI[0], iI[1], iI[2], iI[3], iI[4], iI[5]};
        for (int f = 0; f < 20; ++f) h_f_[static_cast<size_t>(f) * cap_ + i] = vals[f];
        h_u_[i] = p.shape_handle;
        h_u_[cap_ + i] = flags;
        h_shape_.push_back(p.shape_handle);
        return i;
    }

    void upload() {
        if (count_ == 0) return;
        d_f_.upload(h_f_.data());
        d_u_.upload(h_u_.data());
    }

    BodyView view() noexcept {
        float *f = d_f_.data();
        uint32_t *u = d_u_.data();
        const size_t c = cap_;
        BodyView v;
        float **fields[20] = {&v.pos_x, &v.pos_y, &v.pos_z, &v.rot_w, &v.rot_x, &v.rot_y, &v.rot_z,
                              &v.vel_x, &v.vel_y, &v.vel_z, &v.ang_x, &v.ang_y, &v.ang_z, &v.inv_mass,
                              &v.iI_xx, &v.iI_yy, &v.iI_zz, &v.iI_xy, &v.iI_xz, &v.iI_yz};
        for (size_t k = 0; k < 20; ++k) *fields[k] = f ? f + k * c : nullptr;
        v.shape = u;
        v.flags = u ? u + c : nullptr;
        v.n = count_;
        return v;
    }

    uint32_t count() const noexcept { return count_; }
    uint32_t capacity() const noexcept { return cap_; }
    const std::vector<uint32_t> &body_shapes() const noexcept { return h_shape_; }

private:
    uint32_t cap_ = 0, count_ = 0;
    Buffer<float> d_f_;     // 20 float fields, capacity each, in BodyView order
    Buffer<uint32_t> d_u_;  // shape, flags
    std::vector<float> h_f_;
    std::vector<uint32_t> h_u_;
    std::vector<uint32_t> h_shape_;
};

}  // namespace dyphur
