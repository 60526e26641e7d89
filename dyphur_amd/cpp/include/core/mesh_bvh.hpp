#pragma once
// MeshBvhCatalogView (core/include/core/mesh_bvh.hpp:38-...): accepted in
// signatures only; the gfx950 narrowphase rejects triangle meshes.
#include <cstdint>

namespace dyphur {

struct MeshBvhCatalogView {
    const uint32_t *vtx_offset = nullptr;
    const uint32_t *tri_offset = nullptr;
    const uint32_t *node_offset = nullptr;
    const uint32_t *root_node = nullptr;
    uint32_t n_meshes = 0;
};

}  // namespace dyphur
