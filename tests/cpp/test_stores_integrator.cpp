// BodyStore / ShapeStore / integrate through the C++ layer
// (core/tests/body_store_test.cpp:11-110, core/tests/integrator_test.cpp:26-144).
#include "mini_test.hpp"

#include <compute/device.hpp>
#include <core/body_store.hpp>
#include <core/integrator.hpp>
#include <core/shape_store.hpp>

#include <vector>

using namespace dyphur;

static constexpr float kEps = 1e-5f;

template <typename T>
static std::vector<T> fetch(Stream &s, const T *d, size_t n) {
    std::vector<T> h(n);
    s.queue().memcpy(h.data(), d, n * sizeof(T)).wait();
    return h;
}

TEST_CASE("Device: no CPU device, a GPU is preferred", "[device]") {
    REQUIRE_THROWS_AS(Device::default_cpu(), Error);
    auto dev = Device::preferred();
    REQUIRE(dev.is_gpu());
    REQUIRE_FALSE(dev.is_cpu());
    REQUIRE(dev.max_compute_units() > 0);
    REQUIRE(dev.global_mem_bytes() > (size_t(1) << 30));
}

TEST_CASE("BodyStore: construction and capacity", "[body]") {
    auto s = Device::default_gpu().make_stream();
    BodyStore store(s, 16);
    REQUIRE(store.count() == 0);
    REQUIRE(store.capacity() == 16);
}

TEST_CASE("BodyStore: add and upload round-trip", "[body]") {
    auto s = Device::default_gpu().make_stream();
    BodyStore store(s, 8);
    BodyParams p;
    p.position = {1.f, 2.f, 3.f};
    p.linear_velocity = {0.5f, 0.f, 0.f};
    p.mass = 2.f;
    p.inertia = Mat3f::identity() * 4.f;
    REQUIRE(store.add(p) == 0);
    BodyParams q;
    q.position = {0.f, 10.f, 0.f};
    q.flags = BodyFlag::Static;
    REQUIRE(store.add(q) == 1);
    REQUIRE(store.count() == 2);
    store.upload();
    s.wait();
    BodyView v = store.view();
    REQUIRE(v.n == 2);
    auto px = fetch(s, v.pos_x, 2), py = fetch(s, v.pos_y, 2), vx = fetch(s, v.vel_x, 2);
    auto im = fetch(s, v.inv_mass, 2), ixx = fetch(s, v.iI_xx, 2);
    auto fl = fetch(s, v.flags, 2);
    REQUIRE_THAT(px[0], WithinAbs(1.f, kEps));
    REQUIRE_THAT(py[1], WithinAbs(10.f, kEps));
    REQUIRE_THAT(vx[0], WithinAbs(0.5f, kEps));
    REQUIRE_THAT(im[0], WithinAbs(0.5f, kEps));
    REQUIRE_THAT(im[1], WithinAbs(0.f, kEps));
    REQUIRE_THAT(ixx[0], WithinAbs(0.25f, kEps));
    REQUIRE_THAT(ixx[1], WithinAbs(0.f, kEps));
    REQUIRE((fl[1] & BodyFlag::Static) != 0);
    REQUIRE((fl[0] & BodyFlag::Static) == 0);
}

TEST_CASE("BodyStore: diagonal inverse inertia", "[body]") {
    auto s = Device::default_gpu().make_stream();
    BodyStore store(s, 1);
    BodyParams p;
    p.inertia = Mat3f::diagonal(2.f, 4.f, 8.f);
    store.add(p);
    store.upload();
    BodyView v = store.view();
    REQUIRE_THAT(fetch(s, v.iI_xx, 1)[0], WithinAbs(0.5f, kEps));
    REQUIRE_THAT(fetch(s, v.iI_yy, 1)[0], WithinAbs(0.25f, kEps));
    REQUIRE_THAT(fetch(s, v.iI_zz, 1)[0], WithinAbs(0.125f, kEps));
    REQUIRE_THAT(fetch(s, v.iI_xy, 1)[0], WithinAbs(0.f, kEps));
}

TEST_CASE("BodyStore: capacity overflow throws", "[body]") {
    auto s = Device::default_gpu().make_stream();
    BodyStore store(s, 1);
    store.add(BodyParams{});
    REQUIRE_THROWS_AS(store.add(BodyParams{}), Error);
}

TEST_CASE("integrate: static body does not move", "[integrator]") {
    auto s = Device::default_gpu().make_stream();
    BodyStore store(s, 1);
    BodyParams p;
    p.position = {1.f, 2.f, 3.f};
    p.mass = 0.f;
    p.flags = BodyFlag::Static;
    store.add(p);
    store.upload();
    integrate(s, store.view(), IntegratorParams{{0.f, -9.81f, 0.f}, 1.f / 60.f, 1});
    s.wait();
    BodyView v = store.view();
    REQUIRE_THAT(fetch(s, v.pos_x, 1)[0], WithinAbs(1.f, 1e-6));
    REQUIRE_THAT(fetch(s, v.pos_y, 1)[0], WithinAbs(2.f, 1e-6));
}

// Semi-implicit Euler over N substeps of dt/N: y_N = y0 + g dt^2 (N + 1) / (2N).
TEST_CASE("integrate: free fall matches the closed form", "[integrator]") {
    auto s = Device::default_gpu().make_stream();
    BodyStore store(s, 1);
    BodyParams p;
    p.position = {0.f, 10.f, 0.f};
    store.add(p);
    store.upload();
    integrate(s, store.view(), IntegratorParams{{0.f, -9.81f, 0.f}, 1.f, 60});
    s.wait();
    const float expected = 10.f + (-9.81f) * 61.f / 120.f;
    REQUIRE_THAT(fetch(s, store.view().pos_y, 1)[0], WithinAbs(expected, 1e-3));
}

TEST_CASE("integrate: constant velocity without gravity", "[integrator]") {
    auto s = Device::default_gpu().make_stream();
    BodyStore store(s, 1);
    BodyParams p;
    p.linear_velocity = {3.f, 0.f, 0.f};
    store.add(p);
    store.upload();
    integrate(s, store.view(), IntegratorParams{{0.f, 0.f, 0.f}, 1.f, 1});
    s.wait();
    REQUIRE_THAT(fetch(s, store.view().pos_x, 1)[0], WithinAbs(3.f, kEps));
}

TEST_CASE("integrate: spinning quaternion stays unit length", "[integrator]") {
    auto s = Device::default_gpu().make_stream();
    BodyStore store(s, 1);
    BodyParams p;
    p.angular_velocity = {0.f, 1.f, 0.f};
    store.add(p);
    store.upload();
    for (int i = 0; i < 60; ++i) integrate(s, store.view(), IntegratorParams{{0.f, 0.f, 0.f}, 1.f / 60.f, 4});
    s.wait();
    BodyView v = store.view();
    const float w = fetch(s, v.rot_w, 1)[0], x = fetch(s, v.rot_x, 1)[0];
    const float y = fetch(s, v.rot_y, 1)[0], z = fetch(s, v.rot_z, 1)[0];
    REQUIRE_THAT(w * w + x * x + y * y + z * z, WithinAbs(1.f, 1e-4));
    REQUIRE(std::fabs(y) > 0.1f);
}

TEST_CASE("ShapeStore: add, upload, view", "[shapes]") {
    auto s = Device::default_gpu().make_stream();
    ShapeStore ss(s, 4);
    ShapeParams a;
    a.type = ShapeType::Sphere;
    a.half_x = 0.25f;
    REQUIRE(ss.add(a) == 0);
    ShapeParams b;
    b.half_y = 2.f;
    REQUIRE(ss.add(b) == 1);
    ss.upload();
    ShapeView v = ss.view();
    REQUIRE(v.n == 2);
    auto t = fetch(s, v.type, 2);
    REQUIRE(t[0] == 1u);
    REQUIRE(t[1] == 0u);
    REQUIRE_THAT(fetch(s, v.half_x, 1)[0], WithinAbs(0.25f, 0.0));
    REQUIRE_THAT(fetch(s, v.half_y, 2)[1], WithinAbs(2.f, 0.0));
}
