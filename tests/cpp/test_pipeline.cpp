// Broadphase / Narrowphase / XpbdSolver / frame determinism through the C++
// layer (core/tests/broadphase_test.cpp:21-155, narrowphase_test.cpp:58-210,
// xpbd_solver_test.cpp:37-230, sim_determinism_test.cpp:218-241).
#include "mini_test.hpp"

#include <compute/buffer.hpp>
#include <compute/device.hpp>
#include <core/aabb.hpp>
#include <core/body_store.hpp>
#include <core/broadphase.hpp>
#include <core/integrator.hpp>
#include <core/narrowphase.hpp>
#include <core/shape_store.hpp>
#include <core/xpbd_solver.hpp>

#include <algorithm>
#include <cstdint>
#include <set>
#include <utility>
#include <vector>

using namespace dyphur;

static constexpr float kEps = 1e-4f;

template <typename T>
static std::vector<T> fetch(Stream &s, const T *d, size_t n) {
    std::vector<T> h(n);
    if (n) s.queue().memcpy(h.data(), d, n * sizeof(T)).wait();
    return h;
}

static Buffer<ContactPair> make_pairs(Stream &s, const std::vector<ContactPair> &v) {
    Buffer<ContactPair> b(s, v.size());
    b.upload(v.data()).wait();
    return b;
}

struct Scene {
    Stream s;
    ShapeStore shapes;
    BodyStore bodies;
    Scene(uint32_t n_shapes, uint32_t n_bodies)
        : s(Device::default_gpu().make_stream()), shapes(s, n_shapes), bodies(s, n_bodies) {}
};

static void add_spheres(Scene &sc, const std::vector<Vec3f> &at, float r) {
    ShapeParams sp;
    sp.type = ShapeType::Sphere;
    sp.half_x = r;
    const uint32_t h = sc.shapes.add(sp);
    for (const auto &p : at) {
        BodyParams b;
        b.position = p;
        b.shape_handle = h;
        sc.bodies.add(b);
    }
    sc.shapes.upload();
    sc.bodies.upload();
    sc.s.wait();
}

TEST_CASE("AABB: overlap, touching and merge", "[broadphase][aabb]") {
    AABB a = {0.f, 0.f, 0.f, 1.f, 1.f, 1.f};
    AABB b = {0.5f, 0.f, 0.f, 1.5f, 1.f, 1.f};
    AABB c = {2.f, 0.f, 0.f, 3.f, 1.f, 1.f};
    AABB d = {1.f, 0.f, 0.f, 2.f, 1.f, 1.f};
    REQUIRE(a.overlaps(b));
    REQUIRE_FALSE(a.overlaps(c));
    REQUIRE(a.overlaps(d));
    AABB m = AABB::merge(a, AABB{0.5f, -1.f, 0.f, 2.f, 0.5f, 3.f});
    REQUIRE(m.min_y == -1.f);
    REQUIRE(m.max_x == 2.f);
    REQUIRE(m.max_z == 3.f);
}

TEST_CASE("Broadphase: two close spheres and one far away give one pair", "[broadphase]") {
    Scene sc(1, 3);
    add_spheres(sc, {{0.f, 0.f, 0.f}, {1.f, 0.f, 0.f}, {100.f, 0.f, 0.f}}, 1.f);
    Broadphase bp(sc.s, 3, 32);
    bp.build_and_query(sc.s, sc.bodies.view(), sc.shapes.view(), AABB{-2.f, -2.f, -2.f, 102.f, 2.f, 2.f});
    const uint32_t n = bp.download_count(sc.s);
    REQUIRE(n == 1u);
    bp.sort_pairs(sc.s, n);
    auto p = fetch(sc.s, bp.pairs_ptr(), n);
    REQUIRE(p[0].a == 0u);
    REQUIRE(p[0].b == 1u);
}

TEST_CASE("Broadphase: four mutually overlapping spheres give all six pairs", "[broadphase]") {
    Scene sc(1, 4);
    add_spheres(sc, {{0.f, 0.f, 0.f}, {0.5f, 0.f, 0.f}, {0.f, 0.5f, 0.f}, {0.f, 0.f, 0.5f}}, 1.f);
    Broadphase bp(sc.s, 4, 32);
    bp.build_and_query(sc.s, sc.bodies.view(), sc.shapes.view(), AABB{-2.f, -2.f, -2.f, 2.f, 2.f, 2.f});
    const uint32_t n = bp.download_count(sc.s);
    REQUIRE(n == 6u);
    bp.sort_pairs(sc.s, n);
    auto p = fetch(sc.s, bp.pairs_ptr(), n);
    std::set<std::pair<uint32_t, uint32_t>> uniq;
    for (auto &q : p) {
        REQUIRE(q.a < q.b);
        uniq.insert({q.a, q.b});
    }
    REQUIRE(uniq.size() == 6u);
    REQUIRE(std::is_sorted(p.begin(), p.end(), [](const ContactPair &x, const ContactPair &y) {
        return x.a != y.a ? x.a < y.a : x.b < y.b;
    }));
}

TEST_CASE("Broadphase: disjoint spheres give no pair; BVH view is consistent", "[broadphase]") {
    Scene sc(1, 2);
    add_spheres(sc, {{0.f, 0.f, 0.f}, {5.f, 0.f, 0.f}}, 0.4f);
    Broadphase bp(sc.s, 2, 32);
    bp.build_and_query(sc.s, sc.bodies.view(), sc.shapes.view(), AABB{-1.f, -1.f, -1.f, 6.f, 1.f, 1.f});
    REQUIRE(bp.download_count(sc.s) == 0u);
    BvhView v = bp.bvh_view(2);
    REQUIRE(v.n == 2u);
    auto root = fetch(sc.s, v.root, 1);
    REQUIRE(root[0] == 0);
    auto box = fetch(sc.s, v.aabb, 3);
    // leaf of body 0 (sorted first: smaller x): [-0.4, 0.4] in fp16
    auto idx = fetch(sc.s, v.sorted_idx, 2);
    const NodeBox &leaf0 = box[1 + (idx[0] == 0 ? 0 : 1)];
    REQUIRE_THAT(float(leaf0.mn_x), WithinAbs(-0.4f, 1e-3));
    REQUIRE_THAT(float(leaf0.mx_x), WithinAbs(0.4f, 1e-3));
    REQUIRE_THAT(float(box[0].mx_x), WithinAbs(5.4f, 1e-2));
}

struct NpCase {
    Scene sc;
    Narrowphase np;
    NpCase(ShapeType ta, float ra, ShapeType tb, float rb, Vec3f pb)
        : sc(2, 2), np(sc.s, 8) {
        ShapeParams a, b;
        a.type = ta;
        a.half_x = a.half_y = a.half_z = ra;
        b.type = tb;
        b.half_x = b.half_y = b.half_z = rb;
        const uint32_t ha = sc.shapes.add(a), hb = sc.shapes.add(b);
        BodyParams p;
        p.shape_handle = ha;
        sc.bodies.add(p);
        p.position = pb;
        p.shape_handle = hb;
        sc.bodies.add(p);
        sc.shapes.upload();
        sc.bodies.upload();
        auto pairs = make_pairs(sc.s, {{0u, 1u}});
        np.run(sc.s, pairs.data(), 1u, sc.bodies.view(), sc.shapes.view());
        sc.s.wait();
    }
    float get(float *ContactView::*f) {
        ContactView c = np.contacts();
        return fetch(sc.s, c.*f, 1)[0];
    }
};

TEST_CASE("Narrowphase: sphere-sphere contact", "[narrowphase]") {
    NpCase t(ShapeType::Sphere, 1.f, ShapeType::Sphere, 1.f, {1.5f, 0.f, 0.f});
    REQUIRE(t.np.download_count(t.sc.s) == 1u);
    REQUIRE_THAT(t.get(&ContactView::depth), WithinAbs(0.5f, kEps));
    REQUIRE_THAT(t.get(&ContactView::norm_x), WithinAbs(-1.f, kEps));
    REQUIRE_THAT(t.get(&ContactView::norm_y), WithinAbs(0.f, kEps));
    REQUIRE_THAT(t.get(&ContactView::pos_x), WithinAbs(0.75f, kEps));
}

TEST_CASE("Narrowphase: sphere-sphere no contact", "[narrowphase]") {
    NpCase t(ShapeType::Sphere, 1.f, ShapeType::Sphere, 1.f, {3.f, 0.f, 0.f});
    REQUIRE(t.np.download_count(t.sc.s) == 0u);
}

TEST_CASE("Narrowphase: sphere-box contact", "[narrowphase]") {
    NpCase t(ShapeType::Box, 1.f, ShapeType::Sphere, 0.5f, {1.3f, 0.f, 0.f});
    REQUIRE(t.np.download_count(t.sc.s) == 1u);
    REQUIRE_THAT(t.get(&ContactView::depth), WithinAbs(0.2f, kEps));
    REQUIRE_THAT(t.get(&ContactView::norm_x), WithinAbs(-1.f, kEps));
    REQUIRE_THAT(t.get(&ContactView::pos_x), WithinAbs(1.f, kEps));
}

TEST_CASE("Narrowphase: box-box face contact", "[narrowphase]") {
    NpCase t(ShapeType::Box, 1.f, ShapeType::Box, 1.f, {1.8f, 0.f, 0.f});
    REQUIRE(t.np.download_count(t.sc.s) == 1u);
    REQUIRE_THAT(t.get(&ContactView::depth), WithinAbs(0.2f, kEps));
    REQUIRE_THAT(t.get(&ContactView::norm_x), WithinAbs(-1.f, kEps));
    REQUIRE_THAT(t.get(&ContactView::pos_x), WithinAbs(1.f, kEps));
    REQUIRE(t.np.last_contact_count() == 1u);
    REQUIRE_FALSE(t.np.overflowed());
}

TEST_CASE("Narrowphase: box-box separated", "[narrowphase]") {
    NpCase t(ShapeType::Box, 1.f, ShapeType::Box, 1.f, {5.f, 0.f, 0.f});
    REQUIRE(t.np.download_count(t.sc.s) == 0u);
}

TEST_CASE("Narrowphase: convex hull catalogs are rejected", "[narrowphase]") {
    NpCase t(ShapeType::Sphere, 1.f, ShapeType::Sphere, 1.f, {1.5f, 0.f, 0.f});
    ConvexHullView hulls;
    hulls.n_hulls = 1;
    auto pairs = make_pairs(t.sc.s, {{0u, 1u}});
    REQUIRE_THROWS_AS(t.np.run(t.sc.s, pairs.data(), 1u, t.sc.bodies.view(), t.sc.shapes.view(), hulls), Error);
}

// Two unit spheres overlapping by 0.5 along y, one solve pass.
static std::vector<float> solve_pair(Vec3f vb, float mass_a, uint32_t flags_a, int iters, const char *field) {
    Scene sc(1, 2);
    ShapeParams sp;
    sp.type = ShapeType::Sphere;
    sp.half_x = 1.f;
    sc.shapes.add(sp);
    BodyParams a;
    a.mass = mass_a;
    a.flags = flags_a;
    sc.bodies.add(a);
    BodyParams b;
    b.position = {0.f, 1.5f, 0.f};
    b.linear_velocity = vb;
    sc.bodies.add(b);
    sc.shapes.upload();
    sc.bodies.upload();
    Narrowphase np(sc.s, 16);
    auto pairs = make_pairs(sc.s, {{0u, 1u}});
    np.run(sc.s, pairs.data(), 1u, sc.bodies.view(), sc.shapes.view());
    XpbdSolver solver(sc.s, iters, 16, 0.5f);
    solver.solve(sc.s, np.contacts(), JointView{}, sc.bodies.view(), 1.f / 60.f);
    sc.s.wait();
    BodyView v = sc.bodies.view();
    return fetch(sc.s, std::string(field) == "vel_y" ? v.vel_y : v.pos_y, 2);
}

TEST_CASE("XpbdSolver: overlapping spheres are separated", "[solver]") {
    auto y = solve_pair({0.f, 0.f, 0.f}, 1.f, 0u, 1, "pos_y");
    REQUIRE(y[1] - y[0] >= 2.f - 1e-3f);
}

TEST_CASE("XpbdSolver: approaching velocity is removed", "[solver]") {
    auto v = solve_pair({0.f, -2.f, 0.f}, 1.f, 0u, 1, "vel_y");
    REQUIRE(v[1] - v[0] >= -1e-3f);
}

TEST_CASE("XpbdSolver: static body is not moved", "[solver]") {
    auto y = solve_pair({0.f, 0.f, 0.f}, 0.f, BodyFlag::Static, 1, "pos_y");
    REQUIRE_THAT(y[0], WithinAbs(0.f, 1e-3));
    REQUIRE(y[1] >= 1.f - 1e-3f);
}

TEST_CASE("XpbdSolver: a joint view without arrays is rejected", "[solver]") {
    Scene sc(1, 1);
    add_spheres(sc, {{0.f, 0.f, 0.f}}, 1.f);
    Narrowphase np(sc.s, 4);
    XpbdSolver solver(sc.s);
    JointView j;
    j.n = 1;
    REQUIRE_THROWS_AS(solver.solve(sc.s, np.contacts(), j, sc.bodies.view(), 1.f / 60.f), Error);
}

// 300 disjoint sphere pairs: the coloured path (> kSerialThreshold contacts)
// and the forced serial path give the same bits.
TEST_CASE("XpbdSolver: coloured equals serial on independent pairs", "[solver]") {
    std::vector<std::vector<float>> runs;
    for (bool serial : {false, true, false}) {
        Scene sc(1, 600);
        std::vector<Vec3f> at;
        for (int k = 0; k < 300; ++k) {
            at.push_back({0.f, 10.f * k, 0.f});
            at.push_back({0.f, 10.f * k + 0.8f, 0.f});
        }
        add_spheres(sc, at, 0.5f);
        std::vector<ContactPair> pv;
        for (uint32_t k = 0; k < 300; ++k) pv.push_back({2 * k, 2 * k + 1});
        auto pairs = make_pairs(sc.s, pv);
        Narrowphase np(sc.s, 1024);
        np.run(sc.s, pairs.data(), 300u, sc.bodies.view(), sc.shapes.view());
        REQUIRE(np.download_count(sc.s) == 300u);
        XpbdSolver solver(sc.s, 10, 1024, 0.5f);
        solver.set_force_serial(serial);
        solver.solve(sc.s, np.contacts(), JointView{}, sc.bodies.view(), 1.f / 60.f);
        REQUIRE(solver.last_stats().used_serial == (serial ? 1u : 0u));
        runs.push_back(fetch(sc.s, sc.bodies.view().pos_y, 600));
    }
    REQUIRE(runs[0] == runs[1]);
    REQUIRE(runs[0] == runs[2]);
}

// The reference's determinism scene: 2x2x2 boxes (h = 0.4, 1 m pitch) over a
// static ground plate, 20 frames of the stress driver's call sequence; the
// FNV-1a-64 of pos/rot is the reference's golden value.
static uint64_t golden_scene_hash() {
    Scene sc(2, 9);
    ShapeParams box;
    box.half_x = box.half_y = box.half_z = 0.4f;
    ShapeParams ground;
    ground.half_x = 5.f;
    ground.half_y = 0.5f;
    ground.half_z = 5.f;
    const uint32_t hb = sc.shapes.add(box), hg = sc.shapes.add(ground);
    const float idiag = 1.f / 3.f * 2.f * 0.4f * 0.4f;
    for (int iy = 0; iy < 2; ++iy)
        for (int ix = 0; ix < 2; ++ix)
            for (int iz = 0; iz < 2; ++iz) {
                BodyParams p;
                p.position = {(float(ix - 1) + 0.5f) * 1.f, 0.5f + float(iy + 1) * 1.f, (float(iz - 1) + 0.5f) * 1.f};
                p.inertia = Mat3f::diagonal(idiag, idiag, idiag);
                p.shape_handle = hb;
                sc.bodies.add(p);
            }
    BodyParams g;
    g.mass = 0.f;
    g.shape_handle = hg;
    g.flags = BodyFlag::Static;
    sc.bodies.add(g);
    sc.shapes.upload();
    sc.bodies.upload();
    const AABB bounds{-20.f, -2.f, -20.f, 20.f, 30.f, 20.f};
    Broadphase bp(sc.s, 9, 9 * 32 + 256);
    Narrowphase np(sc.s, 16384);
    XpbdSolver solver(sc.s, 10, 16384, 0.5f);
    const float dt = 1.f / 60.f;
    for (int f = 0; f < 20; ++f) {
        integrate(sc.s, sc.bodies.view(), IntegratorParams{{0.f, -9.81f, 0.f}, dt, 1});
        bp.build_and_query(sc.s, sc.bodies.view(), sc.shapes.view(), bounds);
        const uint32_t n = bp.download_count(sc.s);
        bp.sort_pairs(sc.s, n);
        np.run(sc.s, bp.pairs_ptr(), std::min(n, bp.max_pairs()), sc.bodies.view(), sc.shapes.view());
        solver.solve(sc.s, np.contacts(), JointView{}, sc.bodies.view(), dt);
    }
    sc.s.wait();
    BodyView v = sc.bodies.view();
    uint64_t h = 0xcbf29ce484222325ull;
    for (float *field : {v.pos_x, v.pos_y, v.pos_z, v.rot_w, v.rot_x, v.rot_y, v.rot_z}) {
        auto vals = fetch(sc.s, field, 9);
        const auto *bytes = reinterpret_cast<const unsigned char *>(vals.data());
        for (size_t i = 0; i < vals.size() * sizeof(float); ++i) {
            h ^= bytes[i];
            h *= 0x100000001b3ull;
        }
    }
    return h;
}

TEST_CASE("Determinism: golden free-fall hash, run to run", "[determinism]") {
    const uint64_t a = golden_scene_hash(), b = golden_scene_hash();
    REQUIRE(a == b);
    REQUIRE(a == 0xcbcd209c3665c818ull);
}
