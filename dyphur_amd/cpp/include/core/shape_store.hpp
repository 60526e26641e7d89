#pragma once
// ShapeStore (core/include/core/shape_store.hpp, core/src/shapes.cpp:19-41).
#include "shapes.hpp"

#include <compute/buffer.hpp>
#include <compute/stream.hpp>

#include <vector>

namespace dyphur {

class ShapeStore {
public:
    ShapeStore() = default;
    ShapeStore(Stream &s, uint32_t capacity)
        : cap_(capacity), d_u_(s, static_cast<size_t>(capacity) * 2), d_f_(s, static_cast<size_t>(capacity) * 3),
          h_u_(static_cast<size_t>(capacity) * 2, 0u), h_f_(static_cast<size_t>(capacity) * 3, 0.f) {}

    uint32_t add(const ShapeParams &p) {
        if (count_ >= cap_) throw Error(DY_ERR_INVALID_ARG, "ShapeStore::add: capacity exceeded");
        const uint32_t i = count_++;
        h_u_[i] = static_cast<uint32_t>(p.type);
        h_u_[cap_ + i] = p.ext_id;
        h_f_[i] = p.half_x;
        h_f_[cap_ + i] = p.half_y;
        h_f_[2 * static_cast<size_t>(cap_) + i] = p.half_z;
        return i;
    }
    void upload() {
        if (count_ == 0) return;
        d_u_.upload(h_u_.data());
        d_f_.upload(h_f_.data());
    }
    ShapeView view() noexcept {
        const size_t c = cap_;
        return ShapeView{d_u_.data(), d_f_.data(), d_f_.data() ? d_f_.data() + c : nullptr,
                         d_f_.data() ? d_f_.data() + 2 * c : nullptr, d_u_.data() ? d_u_.data() + c : nullptr,
                         count_};
    }
    uint32_t count() const noexcept { return count_; }
    uint32_t capacity() const noexcept { return cap_; }

private:
    uint32_t cap_ = 0, count_ = 0;
    Buffer<uint32_t> d_u_;  // type, ext_id
    Buffer<float> d_f_;     // half_x, half_y, half_z
    std::vector<uint32_t> h_u_;
    std::vector<float> h_f_;
};

}  // namespace dyphur
