#pragma once
// ConvexHullView (core/include/core/convex_hull_store.hpp:12-20): accepted in
// signatures only; the gfx950 narrowphase rejects hull shapes.
#include <cstdint>

namespace dyphur {

struct ConvexHullView {
    const float *vtx_x = nullptr;
    const float *vtx_y = nullptr;
    const float *vtx_z = nullptr;
    const uint32_t *hull_start = nullptr;
    const uint32_t *hull_count = nullptr;
    uint32_t n_hulls = 0;
    uint32_t n_verts = 0;
};

}  // namespace dyphur
