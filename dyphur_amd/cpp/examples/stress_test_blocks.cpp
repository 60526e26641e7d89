// stress_test_blocks — the reference's stress driver (examples/stress_test_blocks/main.cpp)
// on the gfx950 C++ layer: a stack of boxes on a static ground plate, the
// per-frame stage sequence of the reference, a trajectory snapshot every 2nd
// frame, the final-state FNV-1a-64 in <prefix>.golden and run metrics in
// <prefix>.metrics.json.
//
//   stress_test_blocks [prefix] [n_frames] [--grid nx ny nz] [--pitch p] [--profile]
//
// --profile runs 30 warm-up + n_frames frames with a stream wait after every
// stage and prints the mean per-stage times instead of writing files.  --cpu
// is refused: this build has no CPU device.
#include <compute/device.hpp>
#include <core/body_store.hpp>
#include <core/broadphase.hpp>
#include <core/integrator.hpp>
#include <core/narrowphase.hpp>
#include <core/shape_store.hpp>
#include <core/step.hpp>
#include <core/xpbd_solver.hpp>
#include <io/scene_io.hpp>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

using namespace dyphur;
using Clock = std::chrono::steady_clock;

int main(int argc, char **argv) try {
    std::string prefix = "stress_test_blocks";
    bool prefix_set = false, profile = false;
    int n_frames = 300, nx = 8, ny = 16, nz = 8;
    float pitch = 1.0f;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--cpu") {
            std::fprintf(stderr, "stress_test_blocks: --cpu is not available (gfx950 GPUs only)\n");
            return 2;
        }
        if (a == "--profile") { profile = true; continue; }
        if (a == "--grid" && i + 3 < argc) {
            nx = std::atoi(argv[++i]); ny = std::atoi(argv[++i]); nz = std::atoi(argv[++i]);
            continue;
        }
        if (a == "--pitch" && i + 1 < argc) { pitch = static_cast<float>(std::atof(argv[++i])); continue; }
        if (!prefix_set) { prefix = a; prefix_set = true; }
        else n_frames = std::atoi(a.c_str());
    }
    auto s = Device::preferred().make_stream();

    const uint32_t n_dyn = static_cast<uint32_t>(nx * ny * nz), n = n_dyn + 1;
    const float box_h = 0.4f, gnd_hy = 0.5f;
    const float gnd_hx = std::max(20.0f, 0.5f * static_cast<float>(std::max(nx, nz)) + 2.0f);
    const uint32_t max_pairs = std::max<uint32_t>(65536u, n_dyn * 48u);
    const uint32_t max_contacts = std::max<uint32_t>(16384u, n_dyn * 16u);
    const float dt = 1.0f / 60.0f;

    ShapeStore ss(s, 2);
    ShapeParams box, ground;
    box.half_x = box.half_y = box.half_z = box_h;
    ground.half_x = gnd_hx;
    ground.half_y = gnd_hy;
    ground.half_z = gnd_hx;
    const uint32_t h_box = ss.add(box), h_gnd = ss.add(ground);
    ss.upload();

    BodyStore bs(s, n);
    const float mass = 1.0f, idiag = mass / 3.0f * (box_h * box_h + box_h * box_h);
    for (int iy = 0; iy < ny; ++iy)
        for (int ix = 0; ix < nx; ++ix)
            for (int iz = 0; iz < nz; ++iz) {
                BodyParams p;
                p.position = {(static_cast<float>(ix - nx / 2) + 0.5f) * pitch,
                              gnd_hy + static_cast<float>(iy + 1) * pitch,
                              (static_cast<float>(iz - nz / 2) + 0.5f) * pitch};
                p.mass = mass;
                p.inertia = Mat3f::diagonal(idiag, idiag, idiag);
                p.shape_handle = h_box;
                bs.add(p);
            }
    BodyParams g;
    g.mass = 0.0f;
    g.shape_handle = h_gnd;
    g.flags = BodyFlag::Static;
    bs.add(g);
    bs.upload();
    s.wait();

    const float bound_xz = gnd_hx + 5.0f;
    const float bound_y = std::max(22.0f, gnd_hy + static_cast<float>(ny) + 5.0f);
    const AABB bounds{-bound_xz, -2.0f, -bound_xz, bound_xz, bound_y, bound_xz};
    Broadphase bp(s, n, max_pairs);
    Narrowphase np(s, max_contacts);
    XpbdSolver solver(s, 10, max_contacts, 0.5f);
    const IntegratorParams ip{{0.0f, -9.81f, 0.0f}, dt, 1};
    BodyView bv = bs.view();
    ShapeView sv = ss.view();

    if (profile) {
        const int n_warm = 30;
        double t[6] = {0, 0, 0, 0, 0, 0};
        int measured = 0;
        auto lap = [&](int k, auto &&fn, bool meas) {
            const auto a = Clock::now();
            fn();
            s.wait();
            if (meas) t[k] += std::chrono::duration<double>(Clock::now() - a).count();
        };
        for (int f = 0; f < n_warm + n_frames; ++f) {
            const bool meas = f >= n_warm;
            uint32_t n_pairs = 0;
            lap(0, [&] { integrate(s, bv, ip); }, meas);
            lap(1, [&] { bp.build_and_query(s, bv, sv, bounds); }, meas);
            lap(2, [&] { n_pairs = bp.download_count(s); }, meas);
            lap(3, [&] { bp.sort_pairs(s, n_pairs); }, meas);
            lap(4, [&] { np.run(s, bp.pairs_ptr(), std::min(n_pairs, max_pairs), bv, sv); }, meas);
            lap(5, [&] { solver.solve(s, np.contacts(), JointView{}, bv, dt); }, meas);
            measured += meas ? 1 : 0;
        }
        const double k = 1000.0 / std::max(measured, 1);
        const char *names[6] = {"integrate", "build_and_query", "download_count", "sort_pairs", "narrowphase",
                                "xpbd_solve"};
        double tot = 0;
        for (int i = 0; i < 6; ++i) {
            std::printf("%-16s %8.3f ms/frame\n", names[i], t[i] * k);
            tot += t[i] * k;
        }
        std::printf("%-16s %8.3f ms/frame (%u bodies, %u contacts at the end)\n", "total", tot, n,
                    np.download_count(s));
        return 0;
    }

    std::vector<uint32_t> body_shape(bs.body_shapes());
    const ShapeParams shapes[2] = {box, ground};
    write_scene(prefix, n, body_shape.data(), shapes, 2);
    TrajectoryWriter traj(prefix, n, static_cast<uint32_t>(n_frames));
    std::vector<float> h[7];
    float *fields[7] = {bv.pos_x, bv.pos_y, bv.pos_z, bv.rot_w, bv.rot_x, bv.rot_y, bv.rot_z};
    for (auto &v : h) v.resize(n);
    auto download_state = [&] {
        for (int k = 0; k < 7; ++k) s.queue().memcpy(h[k].data(), fields[k], n * sizeof(float));
        s.wait();
    };

    uint32_t peak_pairs = 0, peak_contacts = 0;
    const auto t0 = Clock::now();
    for (int f = 0; f < n_frames; ++f) {
        integrate(s, bv, ip);
        bp.build_and_query(s, bv, sv, bounds);
        const uint32_t n_pairs = bp.download_count(s);
        bp.sort_pairs(s, n_pairs);
        np.run(s, bp.pairs_ptr(), std::min(n_pairs, max_pairs), bv, sv);
        peak_pairs = std::max(peak_pairs, n_pairs);
        solver.solve(s, np.contacts(), JointView{}, bv, dt);
        if (f % 2 == 0) {
            download_state();
            traj.write_frame(h[0].data(), h[1].data(), h[2].data(), h[3].data(), h[4].data(), h[5].data(),
                             h[6].data());
        }
    }
    s.wait();
    peak_contacts = np.last_contact_count();
    const double avg_cnt = static_cast<double>(np.download_count(s));
    const double wall = std::chrono::duration<double>(Clock::now() - t0).count();
    if (peak_pairs >= max_pairs || peak_contacts >= max_contacts)
        std::fprintf(stderr, "stress_test_blocks: capacity reached (pairs %u/%u, contacts %u/%u)\n", peak_pairs,
                     max_pairs, peak_contacts, max_contacts);
    const uint64_t hash = state_hash(s, bv);
    write_golden(prefix, hash);
    RunMetrics m;
    m.n_bodies = n;
    m.n_frames = static_cast<uint32_t>(n_frames);
    m.simulation_seconds = n_frames * static_cast<double>(dt);
    m.wall_clock_seconds = wall;
    m.frames_per_second = n_frames / wall;
    m.avg_contacts_per_frame = avg_cnt;
    m.realtime_factor = m.frames_per_second / 60.0;
    write_metrics(prefix, m);
    std::printf("stress_test_blocks: %d frames, %u bodies, %.1f fps, state %016llx\n", n_frames, n,
                m.frames_per_second, static_cast<unsigned long long>(hash));
    return 0;
} catch (const std::exception &e) {
    std::fprintf(stderr, "stress_test_blocks: %s\n", e.what());
    return 1;
}
