// JointStore and joints in XpbdSolver through the C++ layer
// (core/tests/articulation_test.cpp:23-320, core/tests/smoke_test.cpp:151-173,
// core/tests/sim_determinism_test.cpp:130-270).
#include "mini_test.hpp"

#include <compute/device.hpp>
#include <core/body_store.hpp>
#include <core/contact_store.hpp>
#include <core/integrator.hpp>
#include <core/joint_store.hpp>
#include <core/step.hpp>
#include <core/xpbd_solver.hpp>

#include <cstdint>
#include <vector>

using namespace dyphur;

namespace {

constexpr float kEps = 1e-5f;

template <typename T>
T fetch1(Stream &s, const T *d, uint32_t i = 0) {
    T v{};
    s.queue().memcpy(&v, d + i, sizeof(T)).wait();
    return v;
}

// Two unit-mass bodies with identity inertia, body 0 optionally static
// (articulation_test.cpp:124-137).
BodyStore two_bodies(Stream &s, Vec3f a, Vec3f b, bool static_a) {
    BodyStore bs(s, 4);
    BodyParams p;
    p.mass = static_a ? 0.f : 1.f;
    p.inertia = Mat3f::identity();
    p.flags = static_a ? BodyFlag::Static : 0u;
    p.position = a;
    bs.add(p);
    p.mass = 1.f;
    p.flags = 0u;
    p.position = b;
    bs.add(p);
    bs.upload();
    s.wait();
    return bs;
}

// One solve of a single joint with anchors (0, 0.5, 0) / (0, -0.5, 0).
void solve_one(Stream &s, BodyStore &bs, JointType type, int iters) {
    JointStore js(s, 2);
    JointParams jp;
    jp.body_parent = 0;
    jp.body_child = 1;
    jp.anchor_parent = {0.f, 0.5f, 0.f};
    jp.anchor_child = {0.f, -0.5f, 0.f};
    jp.type = type;
    js.add(jp);
    js.upload();
    s.wait();
    XpbdSolver solver(s, iters);
    ContactStore cs(s, 1);
    solver.solve(s, cs.view(), js.view(), bs.view(), 1.f / 60.f);
    s.wait();
    REQUIRE(solver.last_stats().n_constraints == 1u);
}

}  // namespace

TEST_CASE("JointType: enum values are distinct", "[joints][smoke]") {
    REQUIRE(static_cast<uint8_t>(JointType::Fixed) == 0);
    REQUIRE(static_cast<uint8_t>(JointType::Revolute) == 1);
    REQUIRE(static_cast<uint8_t>(JointType::Prismatic) == 2);
    REQUIRE(static_cast<uint8_t>(JointType::Ball) == 3);
    JointParams p;
    REQUIRE(p.type == JointType::Fixed);
    REQUIRE(p.limit_lo > p.limit_hi);
}

TEST_CASE("JointStore: construction and capacity", "[joints][smoke]") {
    auto s = Device::default_gpu().make_stream();
    JointStore store(s, 8);
    REQUIRE(store.count() == 0u);
    REQUIRE(store.capacity() == 8u);
    JointParams p;
    for (int i = 0; i < 8; ++i) store.add(p);
    REQUIRE_THROWS_AS(store.add(p), Error);
}

TEST_CASE("JointStore: add and upload round-trip", "[joints][smoke]") {
    auto s = Device::default_gpu().make_stream();
    JointStore store(s, 4);
    JointParams p;
    p.body_parent = 0;
    p.body_child = 1;
    p.anchor_parent = {0.f, 0.5f, 0.f};
    p.anchor_child = {0.f, -0.5f, 0.f};
    p.axis = {1.f, 0.f, 0.f};
    p.limit_lo = -1.57f;
    p.limit_hi = 1.57f;
    p.stiffness = 100.f;
    p.damping = 10.f;
    p.compliance_pos = 1e-4f;
    p.type = JointType::Revolute;
    REQUIRE(store.add(p) == 0u);
    JointParams q;
    q.body_parent = 1;
    q.body_child = 2;
    q.type = JointType::Ball;
    REQUIRE(store.add(q) == 1u);
    store.upload();
    s.wait();
    JointView v = store.view();
    REQUIRE(v.n == 2u);
    REQUIRE(fetch1(s, v.body_parent, 0) == 0u);
    REQUIRE(fetch1(s, v.body_child, 0) == 1u);
    REQUIRE_THAT(fetch1(s, v.anchor_py, 0), WithinAbs(0.5f, kEps));
    REQUIRE_THAT(fetch1(s, v.anchor_cy, 0), WithinAbs(-0.5f, kEps));
    REQUIRE_THAT(fetch1(s, v.axis_px, 0), WithinAbs(1.f, kEps));
    REQUIRE_THAT(fetch1(s, v.limit_lo, 0), WithinAbs(-1.57f, 1e-4f));
    REQUIRE_THAT(fetch1(s, v.limit_hi, 0), WithinAbs(1.57f, 1e-4f));
    REQUIRE_THAT(fetch1(s, v.stiffness, 0), WithinAbs(100.f, kEps));
    REQUIRE_THAT(fetch1(s, v.compliance_pos, 0), WithinAbs(1e-4f, 1e-8f));
    REQUIRE(fetch1(s, v.type, 0) == static_cast<uint8_t>(JointType::Revolute));
    REQUIRE(fetch1(s, v.body_parent, 1) == 1u);
    REQUIRE(fetch1(s, v.body_child, 1) == 2u);
    REQUIRE(fetch1(s, v.type, 1) == static_cast<uint8_t>(JointType::Ball));
    const float tp[2] = {0.25f, -0.75f}, tv[2] = {1.f, 2.f};
    store.set_targets(tp, tv);
    s.wait();
    REQUIRE(fetch1(s, v.target_pos, 1) == -0.75f);
    REQUIRE(fetch1(s, v.target_vel, 0) == 1.f);
}

TEST_CASE("Fixed joint: holds two bodies together after violation", "[joints]") {
    auto s = Device::default_gpu().make_stream();
    BodyStore bs = two_bodies(s, {0.f, 0.f, 0.f}, {0.f, 1.2f, 0.f}, false);
    solve_one(s, bs, JointType::Fixed, 20);
    BodyView bv = bs.view();
    const float ay0 = fetch1(s, bv.pos_y, 0) + 0.5f, ay1 = fetch1(s, bv.pos_y, 1) - 0.5f;
    REQUIRE_THAT(ay0, WithinAbs(ay1, 1e-3f));
}

TEST_CASE("Fixed joint: static parent, dynamic child pulled to constraint", "[joints]") {
    auto s = Device::default_gpu().make_stream();
    BodyStore bs = two_bodies(s, {0.f, 0.f, 0.f}, {0.5f, 1.f, 0.f}, true);
    solve_one(s, bs, JointType::Fixed, 100);
    BodyView bv = bs.view();
    REQUIRE_THAT(fetch1(s, bv.pos_x, 0), WithinAbs(0.f, kEps));
    REQUIRE_THAT(fetch1(s, bv.pos_y, 0), WithinAbs(0.f, kEps));
    REQUIRE_THAT(fetch1(s, bv.pos_x, 1), WithinAbs(0.f, 1e-3f));
    REQUIRE_THAT(fetch1(s, bv.pos_y, 1), WithinAbs(1.f, 1e-3f));
}

TEST_CASE("Revolute joint: anchor constraint satisfied, axis alignment preserved", "[joints]") {
    auto s = Device::default_gpu().make_stream();
    BodyStore bs = two_bodies(s, {0.f, 0.f, 0.f}, {0.f, 1.f, 0.f}, true);
    solve_one(s, bs, JointType::Revolute, 10);
    BodyView bv = bs.view();
    REQUIRE_THAT(fetch1(s, bv.pos_x, 1), WithinAbs(0.f, 1e-3f));
    REQUIRE_THAT(fetch1(s, bv.pos_y, 1) - 0.5f, WithinAbs(0.5f, 1e-3f));
    REQUIRE_THAT(fetch1(s, bv.pos_z, 1), WithinAbs(0.f, 1e-3f));
}

TEST_CASE("Ball joint: violation corrected to constraint satisfaction", "[joints]") {
    auto s = Device::default_gpu().make_stream();
    BodyStore bs = two_bodies(s, {0.f, 0.f, 0.f}, {0.f, 1.3f, 0.f}, true);
    solve_one(s, bs, JointType::Ball, 20);
    BodyView bv = bs.view();
    REQUIRE_THAT(fetch1(s, bv.pos_x, 1), WithinAbs(0.f, 1e-3f));
    REQUIRE_THAT(fetch1(s, bv.pos_y, 1), WithinAbs(1.f, 1e-3f));
    REQUIRE_THAT(fetch1(s, bv.pos_z, 1), WithinAbs(0.f, 1e-3f));
}

namespace {

// run_artic_sim (sim_determinism_test.cpp:130-218): static base, two
// revolute links about x with PD motors, integrate + solve for 20 frames.
uint64_t run_arm() {
    auto s = Device::default_gpu().make_stream();
    constexpr float kHy = 0.15f, kSep = 2.f * kHy, kMass = 1.f;
    BodyStore bs(s, 3);
    {
        BodyParams p;
        p.mass = 0.f;
        p.flags = BodyFlag::Static;
        p.inertia = Mat3f::identity();
        bs.add(p);
    }
    const float idiag = kMass / 3.f * 2.f * kHy * kHy;
    for (int i = 1; i <= 2; ++i) {
        BodyParams p;
        p.position = {0.f, static_cast<float>(i) * kSep, 0.f};
        p.mass = kMass;
        p.inertia = Mat3f{idiag, 0, 0, 0, idiag, 0, 0, 0, idiag};
        bs.add(p);
    }
    bs.upload();
    s.wait();
    JointStore js(s, 2);
    const float targets[2] = {0.5f, -0.5f};
    for (uint32_t i = 0; i < 2; ++i) {
        JointParams p;
        p.body_parent = i;
        p.body_child = i + 1;
        p.anchor_parent = {0.f, kHy, 0.f};
        p.anchor_child = {0.f, -kHy, 0.f};
        p.axis = {1.f, 0.f, 0.f};
        p.limit_lo = -2.9f;
        p.limit_hi = 2.9f;
        p.target_pos = targets[i];
        p.stiffness = 200.f;
        p.damping = 30.f;
        p.type = JointType::Revolute;
        js.add(p);
    }
    js.upload();
    s.wait();
    ContactStore cs(s, 1);
    XpbdSolver solver(s, 10);
    IntegratorParams ip;
    ip.gravity = {0.f, -9.81f, 0.f};
    ip.dt = 1.f / 60.f;
    ip.n_substeps = 1;
    for (int f = 0; f < 20; ++f) {
        integrate(s, bs.view(), ip);
        solver.solve(s, cs.view(), js.view(), bs.view(), ip.dt);
    }
    s.wait();
    return state_hash(s, bs.view());
}

}  // namespace

TEST_CASE("sim determinism: 2-link articulated arm, 20 frames", "[joints][determinism]") {
    const uint64_t h1 = run_arm(), h2 = run_arm();
    REQUIRE(h1 == h2);
    // The oracle's hash of the same run with the correctly rounded atan2;
    // with the C library's float atan2 it gives the reference's golden
    // 0x44d086af21f338cf (core/tests/sim_determinism_test.cpp:246).
    REQUIRE(h1 == 0x7dcb8c29e0193450ull);
}

TEST_CASE("XpbdSolver: a joint naming a missing body is reported, not run", "[joints]") {
    auto s = Device::default_gpu().make_stream();
    BodyStore bs = two_bodies(s, {0.f, 0.f, 0.f}, {0.f, 1.3f, 0.f}, true);
    JointStore js(s, 1);
    JointParams jp;
    jp.body_parent = 0;
    jp.body_child = 5;  // two bodies only
    jp.type = JointType::Ball;
    js.add(jp);
    js.upload();
    XpbdSolver solver(s, 4);
    ContactStore cs(s, 1);
    solver.solve(s, cs.view(), js.view(), bs.view(), 1.f / 60.f);
    REQUIRE_THROWS_AS(solver.last_stats(), Error);
    REQUIRE_THAT(fetch1(s, bs.view().pos_y, 1), WithinAbs(1.3f, 1e-6f));  // untouched: no joint was solved
}
