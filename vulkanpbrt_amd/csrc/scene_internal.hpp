// Host-side packed scene shared by the scene builder and the kernels' launchers.
#pragma once

#include <vector>

#include "../../include/pbrt_c.h"
#include "scene_types.hpp"

namespace pbrt {

// Per-thread traversal stack entries (LDS, 4 B each).  The BVH builder bounds
// the tree depth by this value (median splits once SAH splits could overrun
// it), and a traversal holds at most depth - 1 deferred children.  32 entries
// keep the stack at 8 KB per 64-lane wave.
constexpr int kTraversalStackSize = 32;

struct PackedScene {
    std::vector<DeviceTexture> textures;
    std::vector<uint32_t> texels;
    std::vector<DeviceMesh> meshes;
    std::vector<DeviceMaterial> materials;
    std::vector<float> positions, normals, uvs;
    std::vector<uint32_t> index_words;
    std::vector<bool> mesh_opaque;
    std::vector<DeviceInstance> instances;
    std::vector<DeviceLight> lights;
    float light_strength_sum = 0.0f;
    std::vector<float> world_tris;  // 9 floats per triangle, instance order
    std::vector<uint32_t> tri_inst, tri_prim;
};

void pack_scene(const pbrt_scene_desc& d, PackedScene& ps);
const DeviceScene& scene_device_view(const pbrt_scene* s);

}  // namespace pbrt
