// RayQuery through the C++ layer (core/tests/ray_query_test.cpp:44-181):
// the reference's four known answers, each on a two-body scene whose LBVH is
// built by Broadphase::build_and_query.
#include "mini_test.hpp"

#include <compute/buffer.hpp>
#include <compute/device.hpp>
#include <core/body_store.hpp>
#include <core/broadphase.hpp>
#include <core/ray_query.hpp>
#include <core/shape_store.hpp>

#include <cstdint>
#include <vector>

using namespace dyphur;

namespace {

struct RayScene {
    Stream s;
    ShapeStore shapes;
    BodyStore bodies;
    Broadphase bp;
    RayScene(ShapeType type, Vec3f a, Vec3f b, AABB bounds)
        : s(Device::default_gpu().make_stream()), shapes(s, 1), bodies(s, 2), bp(s, 2, 8) {
        ShapeParams sp;
        sp.type = type;
        sp.half_x = sp.half_y = sp.half_z = 1.f;
        const uint32_t h = shapes.add(sp);
        for (const Vec3f &p : {a, b}) {
            BodyParams bd;
            bd.position = p;
            bd.shape_handle = h;
            bodies.add(bd);
        }
        shapes.upload();
        bodies.upload();
        bp.build_and_query(s, bodies.view(), shapes.view(), bounds);
        s.wait();
    }
    std::vector<RayHit> cast(const std::vector<Ray> &rays) {
        Buffer<Ray> d_rays(s, rays.size());
        Buffer<RayHit> d_hits(s, rays.size());
        d_rays.upload(rays.data());
        ray_query(s, bp.bvh_view(2), bodies.view(), shapes.view(), d_rays.data(), d_hits.data(),
                  uint32_t(rays.size()));
        s.wait();
        std::vector<RayHit> h(rays.size());
        d_hits.download(h.data());
        return h;
    }
};

constexpr float kTol = 1e-3f;

}  // namespace

TEST_CASE("RayQuery: sphere hit along +x, a parallel ray misses", "[ray_query]") {
    RayScene sc(ShapeType::Sphere, {0.f, 0.f, 0.f}, {0.f, 50.f, 0.f}, AABB{-6.f, -6.f, -6.f, 56.f, 56.f, 56.f});
    auto h = sc.cast({Ray{-5.f, 0.f, 0.f, 1.f, 0.f, 0.f}, Ray{-5.f, 3.f, 0.f, 1.f, 0.f, 0.f}});
    REQUIRE(h[0].body_idx == 0u);
    REQUIRE_THAT(h[0].t, WithinAbs(4.f, kTol));
    REQUIRE(h[1].body_idx == UINT32_MAX);
}

TEST_CASE("RayQuery: box hit along +x, a parallel ray misses", "[ray_query]") {
    RayScene sc(ShapeType::Box, {0.f, 0.f, 0.f}, {0.f, 50.f, 0.f}, AABB{-6.f, -6.f, -6.f, 56.f, 56.f, 56.f});
    auto h = sc.cast({Ray{-5.f, 0.f, 0.f, 1.f, 0.f, 0.f}, Ray{-5.f, 2.f, 0.f, 1.f, 0.f, 0.f}});
    REQUIRE(h[0].body_idx == 0u);
    REQUIRE_THAT(h[0].t, WithinAbs(4.f, kTol));
    REQUIRE(h[1].body_idx == UINT32_MAX);
}

TEST_CASE("RayQuery: the closer of two spheres on the ray is returned", "[ray_query]") {
    RayScene sc(ShapeType::Sphere, {0.f, 0.f, 0.f}, {10.f, 0.f, 0.f}, AABB{-6.f, -6.f, -6.f, 16.f, 6.f, 6.f});
    auto h = sc.cast({Ray{-5.f, 0.f, 0.f, 1.f, 0.f, 0.f}});
    REQUIRE(h[0].body_idx == 0u);
    REQUIRE_THAT(h[0].t, WithinAbs(4.f, kTol));
}

TEST_CASE("RayQuery: t_max caps the hit distance", "[ray_query]") {
    RayScene sc(ShapeType::Sphere, {0.f, 0.f, 0.f}, {50.f, 0.f, 0.f}, AABB{-6.f, -6.f, -6.f, 56.f, 6.f, 6.f});
    auto h = sc.cast({Ray{-5.f, 0.f, 0.f, 1.f, 0.f, 0.f, 2.f}});
    REQUIRE(h[0].body_idx == UINT32_MAX);
}
