// Host-side packed scene shared by the scene builder and the kernels' launchers.
#pragma once

#include <vector>

#include "../../include/pbrt_c.h"
#include "scene_types.hpp"

namespace pbrt {

// Per-thread traversal stack depth (LDS); the BVH builder keeps the tree
// shallower than this.
constexpr int kTraversalStackSize = 64;

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
