/*
 * pbrt_scenegen.h - seed-driven synthetic scenes for benchmarks and tests.
 *
 * The reference loads assets through assimp (source/VulkanPBRT.cpp:199-210);
 * no assets ship with it, so this library generates the workloads itself from
 * a named PRNG (splitmix64-seeded PCG32).  This header has no counterpart in
 * the reference: it produces the input of pbrt_scene_create().
 *
 *   config 1: Cornell-style box (12 quads) + 16 subdivision-2 icospheres,
 *             one 0.2 x 0.2 ceiling area light                 (~5.1k triangles)
 *   config 2: atrium of instanced columns / arches / boxes with 8 procedural
 *             1024^2 RGBA8 textures (two alpha masks), 3 area quads + 1 sun (~260k)
 *   config 5: city of instanced buildings with 64k emissive quads  (~5M)
 */
#ifndef PBRT_SCENEGEN_H
#define PBRT_SCENEGEN_H

#include "pbrt_c.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pbrt_generated_scene pbrt_generated_scene;

typedef struct pbrt_camera_pose {
    float eye[3];
    float center[3];
    float up[3];
    float fovy_degrees;
    float near_distance;
    float far_distance;
} pbrt_camera_pose;

int32_t pbrt_generate_scene(uint32_t config, uint64_t seed, pbrt_generated_scene** out_scene);
/* Fills a scene description whose arrays stay valid until the generated scene is destroyed. */
int32_t pbrt_generated_scene_desc(const pbrt_generated_scene* scene, pbrt_scene_desc* out_desc);
/* Camera pose for frame `frame` of the scene's camera path (static for config 1). */
int32_t pbrt_generated_scene_camera(const pbrt_generated_scene* scene, uint32_t frame, pbrt_camera_pose* out_pose);
int32_t pbrt_generated_scene_destroy(pbrt_generated_scene* scene);

#ifdef __cplusplus
}
#endif

#endif /* PBRT_SCENEGEN_H */
