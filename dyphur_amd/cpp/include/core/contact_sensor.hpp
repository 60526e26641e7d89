#pragma once
// ContactSensor (core/include/core/contact_sensor.hpp, core/src/contact_sensor.cpp:6-59):
// downloads the narrowphase contacts and keeps, in narrowphase order, those
// with a body whose shape index is shape_idx; host arrays for zero-copy views.
#include "narrowphase.hpp"

#include <compute/stream.hpp>

#include <cstdint>
#include <vector>

namespace dyphur {

class ContactSensor {
public:
    explicit ContactSensor(uint32_t shape_idx) : shape_idx_(shape_idx) {}

    uint32_t query(Stream &s, const Narrowphase &np, const std::vector<uint32_t> &body_shapes) {
        const uint32_t total = np.download_count(s);
        for (auto *v : {&h_body_a_, &h_body_b_}) v->clear();
        for (auto *v : {&h_pos_x_, &h_pos_y_, &h_pos_z_, &h_norm_x_, &h_norm_y_, &h_norm_z_, &h_depth_}) v->clear();
        count_ = 0;
        if (total == 0) return 0;
        ContactView cv = const_cast<Narrowphase &>(np).contacts();
        std::vector<uint32_t> ba(total), bb(total);
        std::vector<float> f[7];
        float *src[7] = {cv.pos_x, cv.pos_y, cv.pos_z, cv.norm_x, cv.norm_y, cv.norm_z, cv.depth};
        s.queue().memcpy(ba.data(), cv.body_a, total * sizeof(uint32_t));
        s.queue().memcpy(bb.data(), cv.body_b, total * sizeof(uint32_t));
        for (int k = 0; k < 7; ++k) {
            f[k].resize(total);
            s.queue().memcpy(f[k].data(), src[k], total * sizeof(float));
        }
        s.wait();
        std::vector<float> *dst[7] = {&h_pos_x_, &h_pos_y_, &h_pos_z_, &h_norm_x_, &h_norm_y_, &h_norm_z_, &h_depth_};
        for (uint32_t i = 0; i < total; ++i) {
            const uint32_t a = ba[i], b = bb[i];
            const bool hit = (a < body_shapes.size() && body_shapes[a] == shape_idx_) ||
                             (b < body_shapes.size() && body_shapes[b] == shape_idx_);
            if (!hit) continue;
            h_body_a_.push_back(a);
            h_body_b_.push_back(b);
            for (int k = 0; k < 7; ++k) dst[k]->push_back(f[k][i]);
        }
        count_ = static_cast<uint32_t>(h_depth_.size());
        return count_;
    }

    uint32_t count() const noexcept { return count_; }
    uint32_t shape_idx() const noexcept { return shape_idx_; }
    const uint32_t *body_a() const noexcept { return h_body_a_.data(); }
    const uint32_t *body_b() const noexcept { return h_body_b_.data(); }
    const float *pos_x() const noexcept { return h_pos_x_.data(); }
    const float *pos_y() const noexcept { return h_pos_y_.data(); }
    const float *pos_z() const noexcept { return h_pos_z_.data(); }
    const float *norm_x() const noexcept { return h_norm_x_.data(); }
    const float *norm_y() const noexcept { return h_norm_y_.data(); }
    const float *norm_z() const noexcept { return h_norm_z_.data(); }
    const float *depth() const noexcept { return h_depth_.data(); }

private:
    uint32_t shape_idx_;
    uint32_t count_ = 0;
    std::vector<uint32_t> h_body_a_, h_body_b_;
    std::vector<float> h_pos_x_, h_pos_y_, h_pos_z_, h_norm_x_, h_norm_y_, h_norm_z_, h_depth_;
};

}  // namespace dyphur
