#pragma once
// On-disk formats of the reference's drivers and viz tools
// (examples/common/scene_io.hpp:23-134, examples/stress_test_blocks/main.cpp:172-197,301-336):
//   <prefix>.scene       u32 n_bodies, n_shapes; u32 body_shape_idx[n_bodies];
//                        per shape 20 B: u32 type, f32 half_x, half_y, half_z, u32 0
//   <prefix>.meshmap     text lines "<body_idx> <path>" ('#' comments)
//   <prefix>.trajectory  u32 n_bodies, n_frames (simulation frames); then one
//                        record per written frame: n_bodies x 7 f32
//                        (pos xyz, rot wxyz), the drivers write every 2nd frame
//   <prefix>.golden      16 lower-case hex digits + '\n' (FNV-1a-64 of the state)
//   <prefix>.metrics.json  n_bodies, n_frames, simulation_seconds,
//                        wall_clock_seconds, frames_per_second,
//                        avg_contacts_per_frame, realtime_factor
// All binary fields are little-endian host order, as the reference writes them.
#include <core/shapes.hpp>

#include <cinttypes>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

namespace dyphur {

struct SceneFileDesc {
    uint32_t n_bodies = 0;
    uint32_t n_shapes = 0;
    std::vector<uint32_t> body_shape_idx;
    std::vector<ShapeParams> shapes;
};

inline void write_scene(const std::string &prefix, uint32_t n_bodies, const uint32_t *body_shape_idx,
                        const ShapeParams *shapes, uint32_t n_shapes) {
    std::ofstream f(prefix + ".scene", std::ios::binary);
    if (!f) throw std::runtime_error("scene file not writable: " + prefix + ".scene");
    const uint32_t hdr[2] = {n_bodies, n_shapes};
    f.write(reinterpret_cast<const char *>(hdr), sizeof hdr);
    f.write(reinterpret_cast<const char *>(body_shape_idx), static_cast<std::streamsize>(n_bodies) * 4);
    for (uint32_t i = 0; i < n_shapes; ++i) {
        const uint32_t type = static_cast<uint32_t>(shapes[i].type), pad = 0;
        const float h[3] = {shapes[i].half_x, shapes[i].half_y, shapes[i].half_z};
        f.write(reinterpret_cast<const char *>(&type), 4);
        f.write(reinterpret_cast<const char *>(h), 12);
        f.write(reinterpret_cast<const char *>(&pad), 4);
    }
    if (!f) throw std::runtime_error("scene file incomplete after writing: " + prefix + ".scene");
}

inline SceneFileDesc read_scene(const std::string &prefix) {
    std::ifstream f(prefix + ".scene", std::ios::binary);
    if (!f) throw std::runtime_error("scene file not readable: " + prefix + ".scene");
    SceneFileDesc d;
    uint32_t hdr[2] = {0, 0};
    f.read(reinterpret_cast<char *>(hdr), sizeof hdr);
    d.n_bodies = hdr[0];
    d.n_shapes = hdr[1];
    d.body_shape_idx.resize(d.n_bodies);
    f.read(reinterpret_cast<char *>(d.body_shape_idx.data()), static_cast<std::streamsize>(d.n_bodies) * 4);
    d.shapes.resize(d.n_shapes);
    for (uint32_t i = 0; i < d.n_shapes; ++i) {
        uint32_t type = 0, pad = 0;
        float h[3] = {0.f, 0.f, 0.f};
        f.read(reinterpret_cast<char *>(&type), 4);
        f.read(reinterpret_cast<char *>(h), 12);
        f.read(reinterpret_cast<char *>(&pad), 4);
        d.shapes[i].type = static_cast<ShapeType>(type);
        d.shapes[i].half_x = h[0];
        d.shapes[i].half_y = h[1];
        d.shapes[i].half_z = h[2];
    }
    if (!f) throw std::runtime_error("scene file shorter than its header says: " + prefix + ".scene");
    return d;
}

inline void write_meshmap(const std::string &prefix, const std::vector<std::pair<uint32_t, std::string>> &entries) {
    std::ofstream f(prefix + ".meshmap");
    if (!f) throw std::runtime_error("mesh map not writable: " + prefix + ".meshmap");
    for (const auto &e : entries) f << e.first << " " << e.second << "\n";
}

inline std::unordered_map<uint32_t, std::string> read_meshmap(const std::string &prefix) {
    std::unordered_map<uint32_t, std::string> out;
    std::ifstream f(prefix + ".meshmap");
    if (!f) return out;  // absent: shape-based rendering
    std::string line;
    while (std::getline(f, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream ss(line);
        uint32_t idx = 0;
        std::string path;
        if (ss >> idx >> path) out[idx] = path;
    }
    return out;
}

struct TrajHeader {
    uint32_t n_bodies;
    uint32_t n_frames;  // simulation frames; the file holds the frames written
};

inline TrajHeader read_traj_header(const std::string &prefix) {
    std::ifstream f(prefix + ".trajectory", std::ios::binary);
    if (!f) throw std::runtime_error("trajectory not readable: " + prefix + ".trajectory");
    TrajHeader h{};
    f.read(reinterpret_cast<char *>(&h), sizeof h);
    return h;
}

// Streams trajectory records: header at open, one n_bodies x 7 f32 record per
// write_frame (SoA host arrays in, interleaved per body out).
class TrajectoryWriter {
public:
    TrajectoryWriter(const std::string &prefix, uint32_t n_bodies, uint32_t n_frames)
        : f_(prefix + ".trajectory", std::ios::binary), n_(n_bodies), row_(7 * static_cast<size_t>(n_bodies)) {
        if (!f_) throw std::runtime_error("trajectory not writable: " + prefix + ".trajectory");
        const uint32_t hdr[2] = {n_bodies, n_frames};
        f_.write(reinterpret_cast<const char *>(hdr), sizeof hdr);
    }
    void write_frame(const float *px, const float *py, const float *pz, const float *qw, const float *qx,
                     const float *qy, const float *qz) {
        for (uint32_t i = 0; i < n_; ++i) {
            float *r = &row_[7 * static_cast<size_t>(i)];
            r[0] = px[i]; r[1] = py[i]; r[2] = pz[i];
            r[3] = qw[i]; r[4] = qx[i]; r[5] = qy[i]; r[6] = qz[i];
        }
        f_.write(reinterpret_cast<const char *>(row_.data()), static_cast<std::streamsize>(row_.size() * 4));
        if (!f_) throw std::runtime_error("trajectory frame not written in full");
    }

private:
    std::ofstream f_;
    uint32_t n_;
    std::vector<float> row_;
};

inline void write_golden(const std::string &prefix, uint64_t hash) {
    std::ofstream f(prefix + ".golden");
    if (!f) throw std::runtime_error("golden file not writable: " + prefix + ".golden");
    char buf[32];
    std::snprintf(buf, sizeof buf, "%016" PRIx64 "\n", hash);
    f << buf;
}

struct RunMetrics {
    uint32_t n_bodies = 0, n_frames = 0;
    double simulation_seconds = 0, wall_clock_seconds = 0, frames_per_second = 0, avg_contacts_per_frame = 0,
           realtime_factor = 0;
};

inline void write_metrics(const std::string &prefix, const RunMetrics &m) {
    std::ofstream f(prefix + ".metrics.json");
    if (!f) throw std::runtime_error("metrics file not writable: " + prefix + ".metrics.json");
    char buf[512];
    std::snprintf(buf, sizeof buf,
                  "{\n  \"n_bodies\": %u,\n  \"n_frames\": %u,\n  \"simulation_seconds\": %.4f,\n"
                  "  \"wall_clock_seconds\": %.4f,\n  \"frames_per_second\": %.2f,\n"
                  "  \"avg_contacts_per_frame\": %.1f,\n  \"realtime_factor\": %.3f\n}\n",
                  m.n_bodies, m.n_frames, m.simulation_seconds, m.wall_clock_seconds, m.frames_per_second,
                  m.avg_contacts_per_frame, m.realtime_factor);
    f << buf;
}

}  // namespace dyphur
