// Per-frame constants and ray generation shared by the megakernel and the
// wavefront tracer.
#pragma once

#include "../../include/pbrt_c.h"
#include "traverse.hpp"

namespace pbrt {

struct TraceConsts {
    float inv_view[16];
    float inv_proj[16];
    uint32_t width, height;
    uint32_t frame_number, sample_number;
    uint32_t max_depth, min_depth;
    uint32_t light_sampling;
    uint32_t output;
    uint32_t write_gbuffer;
    uint32_t pad[3];
};

constexpr int kTileX = 8;
constexpr int kTileY = 8;
constexpr int kTileThreads = kTileX * kTileY;

__device__ __forceinline__ uint8_t unorm8(float x) {
    x = gclamp(x, 0.0f, 1.0f);
    return (uint8_t)rintf(x * 255.0f);
}

// createRay (camera.glsl:6-16)
__device__ __forceinline__ void create_ray(const TraceConsts& c, uint32_t px, uint32_t py, bool anti_alias, Rng& re,
                                           vec3& pos, vec3& dir) {
    vec2 center = make_vec2((float)px + 0.5f, (float)py + 0.5f);
    if (anti_alias) {
        vec2 j = rng_vec2(re);
        center = center + (j - make_vec2(0.5f, 0.5f));
    }
    vec2 ndc = make_vec2(center.x / (float)c.width, center.y / (float)c.height);
    vec2 clip = make_vec2(ndc.x * 2.0f - 1.0f, ndc.y * 2.0f - 1.0f);
    pos = make_vec3(c.inv_view[12], c.inv_view[13], c.inv_view[14]);
    vec4 d = mat4_mul(c.inv_proj, make_vec4(clip.x, clip.y, 1.0f, 1.0f));
    vec4 w = mat4_mul(c.inv_view, make_vec4(normalize(xyz(d)), 0.0f));
    dir = xyz(w);
}

__device__ __forceinline__ LightSample sample_light(const DeviceScene& sc, uint32_t mode, vec3 pos, vec3 n, Rng& re) {
    if (mode == PBRT_LIGHT_SAMPLE_SURFACE_STRENGTH)
        return sample_light_surface_strength(sc.lights, sc.light_count, pos, n, re);
    if (mode == PBRT_LIGHT_SAMPLE_LIGHT_STRENGTH)
        return sample_light_light_strength(sc.lights, sc.light_count, sc.light_strength_sum, pos, n, re);
    return sample_light_uniform(sc.lights, sc.light_count, pos, n, re);
}

class TraceTiming;
size_t wavefront_workspace_size(uint32_t w, uint32_t h);
void wavefront_trace(const DeviceScene& sc, const TraceConsts& c, const pbrt_trace_images& img,
                     unsigned long long* counters, TraceTiming* tm, uint32_t block_threads, uint32_t queue_paths,
                     hipStream_t st);

}  // namespace pbrt
