#pragma once
// ShapeType, ShapeParams, ShapeView (core/include/core/shapes.hpp:6-28).
// ConvexHull and TriangleMesh are rejected by the gfx950 narrowphase.
#include <dyphur_hip.h>

#include <cstddef>
#include <cstdint>

namespace dyphur {

enum class ShapeType : uint32_t { Box = 0, Sphere = 1, ConvexHull = 2, TriangleMesh = 3 };

struct ShapeParams {
    ShapeType type = ShapeType::Box;
    float half_x = 0.5f;  // box half extent x; sphere radius
    float half_y = 0.5f;
    float half_z = 0.5f;
    uint32_t ext_id = 0;
};

struct ShapeView {
    const uint32_t *type;
    const float *half_x;
    const float *half_y;
    const float *half_z;
    const uint32_t *ext_id;
    uint32_t n;
};
static_assert(sizeof(ShapeView) == sizeof(dy_shape_view) && offsetof(ShapeView, n) == offsetof(dy_shape_view, n),
              "ShapeView layout");

inline const dy_shape_view *c_view(const ShapeView &v) { return reinterpret_cast<const dy_shape_view *>(&v); }

}  // namespace dyphur
