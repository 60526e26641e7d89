#pragma once
// Ray, RayHit, ray_query (core/include/core/ray_query.hpp:9-41) over the
// gfx950 LBVH: closest exact hit (oriented box, sphere) per ray.
// Asynchronous on s; call s.wait() before reading d_hits.
#include "body.hpp"
#include "broadphase.hpp"
#include "shapes.hpp"

#include <compute/stream.hpp>

#include <cstdint>

namespace dyphur {

struct Ray {
    float ox, oy, oz;    // origin
    float dx, dy, dz;    // direction (need not be normalised; t is in direction units)
    float t_max = 1e9f;  // a hit is recorded only if t < t_max
};
static_assert(sizeof(Ray) == sizeof(dy_ray), "Ray layout");

struct RayHit {
    float t;            // distance along the ray (direction units)
    uint32_t body_idx;  // UINT32_MAX when nothing is hit
};
static_assert(sizeof(RayHit) == sizeof(dy_ray_hit), "RayHit layout");

inline void ray_query(Stream &s, const BvhView &bvh, const BodyView &bodies, const ShapeView &shapes,
                      const Ray *d_rays, RayHit *d_hits, uint32_t n_rays) {
    check(dy_ray_query(s.handle(), reinterpret_cast<const dy_bvh_view *>(&bvh), c_view(bodies), c_view(shapes),
                       reinterpret_cast<const dy_ray *>(d_rays), reinterpret_cast<dy_ray_hit *>(d_hits), n_rays),
          "ray_query");
}

}  // namespace dyphur
