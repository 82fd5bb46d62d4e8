// Device-resident scene layout (all arrays live in HBM, 16-byte aligned).
//
//   tris      3 x float4 per BVH-ordered world-space triangle:
//             {v0.xyz, instance index bits}, {e1 = v1 - v0, primitive index bits},
//             {e2 = v2 - v0, flags (bit 0: instance is alpha tested)}
//   nodes     BVH2, 2 x uint4 per node (quantized, see bvh_quant.hip): per-axis
//             origin + scale exponent, child 0 reference, 8-bit child bounds,
//             child 1 reference.  The builders emit full-precision nodes of
//             4 x float4 first: {c0.lo.x, c0.hi.x, c0.lo.y, c0.hi.y},
//             {c1.lo.x, c1.hi.x, c1.lo.y, c1.hi.y}, {c0.lo.z, c0.hi.z, c1.lo.z, c1.hi.z},
//             {child0, child1, 0, 0} as int bits.
//             child >= 0: inner node index; child < 0: leaf ~(first_tri << 4 | (count - 1)).
//   instances per instance: object matrix, normal matrix, mesh id, index stride, flags
//   meshes    per mesh: offsets into the concatenated vertex/uv/index pools + texture ids
//   materials WaveFrontMaterialPacked per mesh
//   textures  RGBA8 texel pools with per-texture {offset, width, height}
//   lights    PackedLight list with the inclusive strength prefix in strengths.w
#pragma once

#include "pbrt_math.hpp"

namespace pbrt {

struct DeviceLight {
    vec4 v0_type;
    vec4 v1_strength;
    vec4 v2_angle;
    vec4 dir_angle2;
    vec4 col_ambient;
    vec4 col_diffuse;
    vec4 col_specular;
    vec4 strengths;
};
static_assert(sizeof(DeviceLight) == 128, "PackedLight is 128 bytes");

struct DeviceMaterial {
    vec4 ambient_roughness;
    vec4 diffuse_ior;
    vec4 specular_dissolve;
    vec4 transmittance_illum;
    vec4 emission_texture_id;
    uint32_t category_id;
    float pad[3];
};
static_assert(sizeof(DeviceMaterial) == 96, "WaveFrontMaterialPacked is 96 bytes");

struct DeviceInstance {
    float object_mat[16];  // column-major object-to-world
    float normal_mat[9];   // column-major transpose(inverse(upper 3x3)), computed in double
    int32_t mesh_id;
    uint32_t index_stride;
    uint32_t flags;        // bit 0: alpha tested (not opaque)
    uint32_t pad[4];
};
static_assert(sizeof(DeviceInstance) % 16 == 0, "instance records stay 16-byte aligned");

struct DeviceMesh {
    uint32_t vertex_offset;  // into positions / normals (x3) and uvs (x2)
    uint32_t index_word_offset;  // into the 32-bit index word pool
    uint32_t index_count;
    uint32_t index_stride;
    int32_t diffuse_tex;   // texture table index (0 == 1x1 white default)
    int32_t normal_tex;
    int32_t specular_tex;
    int32_t emissive_tex;
};

struct DeviceTexture {
    uint32_t offset;  // texel offset into the texel pool (uint32 RGBA8 words)
    uint32_t width;
    uint32_t height;
    uint32_t pad;
};

struct DeviceScene {
    const float4* tris;
    const uint4* nodes;  // quantized BVH2
    const DeviceInstance* instances;
    const DeviceMesh* meshes;
    const DeviceMaterial* materials;
    const float* positions;
    const float* normals;
    const float* uvs;
    const uint32_t* index_words;
    const DeviceTexture* textures;
    const uint32_t* texels;
    const DeviceLight* lights;
    uint32_t light_count;
    float light_strength_sum;
    uint32_t triangle_count;
    uint32_t node_count;
    uint32_t has_transmissive;  // any material with illum 7 (bounces that do not count toward depth)
    uint32_t pad[3];
};

constexpr int kLeafCountBits = 4;
constexpr int kMaxLeafSize = 1 << kLeafCountBits;
constexpr uint32_t kTriFlagAlphaTested = 1u;
constexpr uint32_t kNoHit = 0xFFFFFFFFu;

}  // namespace pbrt
