#pragma once
// Narrowphase (core/include/core/narrowphase.hpp:26-58): one contact per
// touching sphere/box pair, in pair order.  Hull and mesh catalogs must be
// empty and hull/mesh shapes make the run fail (dyphur::Error).
#include "body.hpp"
#include "broadphase.hpp"
#include "contact_store.hpp"
#include "convex_hull_store.hpp"
#include "mesh_bvh.hpp"
#include "shapes.hpp"

#include <compute/stream.hpp>

#include <cstring>
#include <memory>

namespace dyphur {

class Narrowphase {
public:
    Narrowphase() = default;
    explicit Narrowphase(Stream &s, uint32_t max_contacts) : s_(s.handle()), cap_(max_contacts) {
        dy_narrowphase_t h = nullptr;
        check(dy_narrowphase_create(s.handle(), max_contacts, &h), "Narrowphase");
        h_ = std::shared_ptr<dy_narrowphase_s>(h, [](dy_narrowphase_t p) { (void)dy_narrowphase_destroy(p); });
    }

    void run(Stream &s, const ContactPair *d_pairs, uint32_t n_pairs, const BodyView &bodies,
             const ShapeView &shapes, ConvexHullView hulls = {}, MeshBvhCatalogView meshes = {}) {
        if (hulls.n_hulls || meshes.n_meshes)
            throw Error(DY_ERR_UNSUPPORTED, "Narrowphase::run: convex hulls and meshes are not supported");
        s_ = s.handle();
        check(dy_narrowphase_run(h_.get(), s.handle(), reinterpret_cast<const dy_contact_pair *>(d_pairs), n_pairs,
                                 c_view(bodies), c_view(shapes)),
              "Narrowphase::run");
    }
    // The reference reads the count back and runs the host-count path
    // (core/src/narrowphase.cpp:1135-1144); so does this.
    void run(Stream &s, const ContactPair *d_pairs, const uint32_t *d_n_pairs, const BodyView &bodies,
             const ShapeView &shapes, ConvexHullView hulls = {}, MeshBvhCatalogView meshes = {}) {
        uint32_t n = 0;
        check(dy_memcpy_d2h(&n, d_n_pairs, sizeof n, s.handle()), "Narrowphase::run: pair count");
        run(s, d_pairs, n, bodies, shapes, hulls, meshes);
    }

    ContactView contacts() noexcept {
        dy_contact_view v{};
        (void)dy_narrowphase_contacts(h_.get(), &v);
        ContactView out;
        std::memcpy(&out, &v, sizeof v);
        return out;
    }
    uint32_t download_count(Stream &s) const {
        uint32_t n = 0;
        check(dy_narrowphase_download_count(h_.get(), s.handle(), &n), "Narrowphase::download_count");
        return n;
    }
    // True count of the last run before clamping to capacity (blocking here).
    uint32_t last_contact_count() const {
        uint32_t n = 0;
        check(dy_narrowphase_last_count(h_.get(), s_, &n), "Narrowphase::last_contact_count");
        return n;
    }
    bool overflowed() const { return last_contact_count() > cap_; }

    dy_narrowphase_t handle() const noexcept { return h_.get(); }

private:
    std::shared_ptr<dy_narrowphase_s> h_;
    dy_stream_t s_ = nullptr;
    uint32_t cap_ = 0;
};

}  // namespace dyphur
