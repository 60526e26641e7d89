#pragma once
// ContactPair, NodeBox, BvhView, Broadphase (core/include/core/broadphase.hpp:12-82)
// over the gfx950 LBVH.  build_and_query is asynchronous; pairs come out in
// ascending (a, b) order, so sort_pairs only checks its argument.
#include "aabb.hpp"
#include "body.hpp"
#include "half.hpp"
#include "shapes.hpp"

#include <compute/stream.hpp>

#include <cstring>
#include <memory>

namespace dyphur {

struct ContactPair {
    uint32_t a, b;  // a < b
};
static_assert(sizeof(ContactPair) == sizeof(dy_contact_pair), "ContactPair layout");

struct alignas(16) NodeBox {
    half mn_x, mn_y, mn_z;
    half mx_x, mx_y, mx_z;
    half pad0, pad1;
};
static_assert(sizeof(NodeBox) == 16 && sizeof(NodeBox) == sizeof(dy_node_box), "NodeBox layout");

struct BvhView {
    const int32_t *left;
    const int32_t *right;
    const int32_t *parent;
    const int32_t *root;
    const uint32_t *sorted_idx;
    const NodeBox *aabb;
    uint32_t n;
};
static_assert(sizeof(BvhView) == sizeof(dy_bvh_view), "BvhView layout");

class Broadphase {
public:
    Broadphase() = default;
    Broadphase(Stream &s, uint32_t max_bodies, uint32_t max_pairs) : max_pairs_(max_pairs) {
        dy_broadphase_t h = nullptr;
        check(dy_broadphase_create(s.handle(), max_bodies, max_pairs, &h), "Broadphase");
        h_ = std::shared_ptr<dy_broadphase_s>(h, [](dy_broadphase_t p) { (void)dy_broadphase_destroy(p); });
        const dy_contact_pair *pairs = nullptr;
        const uint32_t *count = nullptr;
        uint32_t mp = 0;
        check(dy_broadphase_pairs(h, &pairs, &count, &mp), "Broadphase: pairs");
        pairs_ = reinterpret_cast<const ContactPair *>(pairs);
        count_ = count;
    }

    void build_and_query(Stream &s, const BodyView &bodies, const ShapeView &shapes, const AABB &scene_bounds) {
        check(dy_broadphase_build_and_query(h_.get(), s.handle(), c_view(bodies), c_view(shapes),
                                            reinterpret_cast<const dy_aabb *>(&scene_bounds)),
              "Broadphase::build_and_query");
    }
    const ContactPair *pairs_ptr() const noexcept { return pairs_; }
    const uint32_t *count_ptr() const noexcept { return count_; }
    uint32_t max_pairs() const noexcept { return max_pairs_; }
    void sort_pairs(Stream &s, uint32_t n_pairs) {
        check(dy_broadphase_sort_pairs(h_.get(), s.handle(), n_pairs), "Broadphase::sort_pairs");
    }
    uint32_t download_count(Stream &s) const {
        uint32_t n = 0;
        check(dy_broadphase_download_count(h_.get(), s.handle(), &n), "Broadphase::download_count");
        return n;
    }
    BvhView bvh_view(uint32_t n_active) const {
        dy_bvh_view v{};
        check(dy_broadphase_bvh_view(h_.get(), n_active, &v), "Broadphase::bvh_view");
        BvhView out;
        static_assert(sizeof out == sizeof v, "BvhView layout");
        std::memcpy(&out, &v, sizeof v);
        return out;
    }
    dy_broadphase_t handle() const noexcept { return h_.get(); }

private:
    std::shared_ptr<dy_broadphase_s> h_;
    const ContactPair *pairs_ = nullptr;
    const uint32_t *count_ = nullptr;
    uint32_t max_pairs_ = 0;
};

}  // namespace dyphur
