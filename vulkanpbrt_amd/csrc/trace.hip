// Path tracing kernels and their C ABI entry points (pbrt_trace,
// pbrt_intersect, pbrt_occluded, pbrt_rng_sequence).
//
// The megakernel follows ptRaygen.rgen:24-104 per pixel: camera ray, primary
// hit, next-event estimation (lighting.glsl:246-261), G-buffer stores, the
// bounce loop of indirectLighting (lighting.glsl:263-287) and the final
// clamp / demodulation.  Project semantics where the reference is undefined:
//   * a path ends when indirectLighting returns without tracing (rejected
//     BRDF sample or Russian roulette) or when the traced ray misses; the
//     reference instead continues with normalize(vec3(0)) = NaN state
//     (ptRaygen.rgen:70), whose result depends on the driver's NaN handling.
//   * the motion/alpha/format rules of the other passes live in their files.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <vector>

#include "capi_common.hpp"
#include "scene_internal.hpp"
#include "trace_common.hpp"

namespace pbrt {

// kPlanes: the node-visit variant (visit_inner); both kernels keep the four per-ray
// plane selectors in registers
template <bool kCount, int kPlanes>
__device__ __forceinline__ void trace_closest(const DeviceScene& sc, vec3 o, vec3 d, float tmin, float tmax,
                                              int* stack, SurfaceInfo& s, vec3& position, uint32_t& category,
                                              TraversalCounters& cnt) {
    Hit h;
    if (traverse<false, kCount, kPlanes>(sc, o, d, tmin, tmax, h, stack, kTileThreads, cnt))
        shade_hit(sc, h, d, s, position, category);
    else
        miss_surface(s, position, category);
}

// nextEventEsitmation with the shadow ray traced inline.
template <bool kCount, int kPlanes>
__device__ __forceinline__ vec3 nee_inline(const DeviceScene& sc, uint32_t mode, vec3 pos, vec3 o,
                                           const SurfaceInfo& s, vec3 throughput, Rng& re, int* stack,
                                           TraversalCounters& cnt, uint32_t& shadow_rays) {
    if (is_miss(s)) return sky_color(-o) * throughput;
    LightSample ls = sample_light(sc, mode, pos, s.normal, re);
    if (!ls.valid) return make_vec3(0.0f);
    ++shadow_rays;
    Hit h;
    bool occluded = traverse<true, kCount, kPlanes>(sc, pos, ls.l, 0.001f, ls.tmax, h, stack, kTileThreads, cnt);
    if (occluded) return make_vec3(0.0f);
    return nee_unshadowed(ls, o, s, throughput);
}

// kStack: traversal stack entries per lane (kShortStackSize: five waves per SIMD)
template <bool kCount, int kStack>
__global__ void __launch_bounds__(kTileThreads) __attribute__((amdgpu_waves_per_eu(kStack <= kShortStackSize ? 5 : 4)))
pt_megakernel(DeviceScene sc, TraceConsts c, float* depth_img,
                                                              float* normal_img, uint8_t* material_img,
                                                              uint8_t* albedo_img, float4* illum_img,
                                                              unsigned long long* counters) {
    __shared__ int stack_mem[kStack * kTileThreads];
    const uint32_t lid = threadIdx.x;
    int* stack = stack_mem + lid;
    const uint32_t px = blockIdx.x * kTileX + (lid % kTileX);
    const uint32_t py = blockIdx.y * kTileY + (lid / kTileX);
    if (px >= c.width || py >= c.height) return;
    const size_t pix = (size_t)py * c.width + px;
    TraversalCounters cnt{0, 0}, cnt_sh{0, 0};
    uint32_t rays = 0, shadow_rays = 0;

    Rng re = rng_init(px, py, c.frame_number);
    vec3 throughput = make_vec3(1.0f);
    bool anti_alias = c.output == PBRT_OUTPUT_FINAL_IMAGE && c.sample_number > 0;
    vec3 cam_pos, dir;
    create_ray(c, px, py, anti_alias, re, cam_pos, dir);
    SurfaceInfo s;
    vec3 hit_pos;
    uint32_t category;
    ++rays;
    trace_closest<kCount, 1>(sc, cam_pos, dir, 0.001f, 10000.0f, stack, s, hit_pos, category, cnt);
    vec3 color = make_vec3(0.0f);
    color += nee_inline<kCount, 1>(sc, c.light_sampling, hit_pos, -normalize(dir), s, throughput, re, stack, cnt_sh,
                                shadow_rays);
    color += s.emissive_color;
    vec3 albedo = s.diffuse_color + s.specular_color;
    if (c.write_gbuffer) {
        depth_img[pix] = distance(hit_pos, cam_pos);
        normal_img[2 * pix + 0] = acosf(s.normal.z);
        normal_img[2 * pix + 1] = atan2f(s.normal.y, s.normal.x);
        uint8_t* mat = material_img + 4 * pix;
        mat[0] = unorm8((float)category / 255.0f);
        mat[1] = 0;
        mat[2] = 0;
        mat[3] = 0;
        uint8_t* alb = albedo_img + 4 * pix;
        alb[0] = unorm8(albedo.x);
        alb[1] = unorm8(albedo.y);
        alb[2] = unorm8(albedo.z);
        alb[3] = 255;
    }
    if (!is_miss(s)) {
        int trans_depth = 0;
        vec3 wpos = cam_pos;
        for (int i = 0; i < (int)c.max_depth && trans_depth < 10; ++i) {
            if (s.illumination_type == 7) {
                --i;
                ++trans_depth;
            }
            vec3 v = normalize(wpos - hit_pos);
            wpos = hit_pos;
            // indirectLighting (lighting.glsl:263-287)
            vec3 l;
            float pdf;
            vec3 brdf = sample_brdf(s, re, v, l, pdf);
            if (brdf == make_vec3(0.0f) || pdf < kEpsilon) break;
            float tc = dot(l, s.normal);
            if (s.illumination_type == 7) tc = 1.0f;
            vec3 path_tp = throughput * (brdf * tc) / pdf;
            if ((uint32_t)i > c.min_depth) {
                if (rng_float(re) < kMinTermination) break;
                path_tp /= 1.0f - kMinTermination;
            }
            throughput = path_tp;
            ++rays;
            trace_closest<kCount, 1>(sc, wpos, l, 0.001f, 10000.0f, stack, s, hit_pos, category, cnt);
            color += nee_inline<kCount, 1>(sc, c.light_sampling, hit_pos, -l, s, throughput, re, stack, cnt_sh,
                                        shadow_rays) +
                     s.emissive_color * throughput;
            if (is_miss(s)) break;
        }
    }
    color = vclamp(color, 0.0f, kMaxRadiance);
    if (c.output == PBRT_OUTPUT_DEMOD_ILLUMINATION_FLOAT) {
        vec3 demod = color;
        if (!isinf(hit_pos.x)) demod = vmin(color / (albedo + make_vec3(kEpsilon)), 1e3f);
        illum_img[pix] = make_float4(demod.x, demod.y, demod.z, 1.0f);
    } else {
        vec3 fc = color;
        if (c.sample_number > 0) {
            float4 prev = illum_img[pix];
            float alpha = 1.0f / (float)(c.sample_number + 1);
            vec3 pl = srgb_to_linear(make_vec3(prev.x, prev.y, prev.z));
            fc = vmix(pl, fc, alpha);
        }
        fc = linear_to_srgb(fc);
        fc = vclamp(fc, 0.0f, 1.0f);
        illum_img[pix] = make_float4(fc.x, fc.y, fc.z, 1.0f);
    }
    if (kCount) {
        atomicAdd(&counters[0], (unsigned long long)rays);
        atomicAdd(&counters[1], (unsigned long long)shadow_rays);
        atomicAdd(&counters[2], (unsigned long long)cnt.nodes);
        atomicAdd(&counters[3], (unsigned long long)cnt.tris);
        atomicAdd(&counters[4], (unsigned long long)cnt_sh.nodes);
        atomicAdd(&counters[5], (unsigned long long)cnt_sh.tris);
    }
}

// Pixel of the id-th path of a frame: the frame is cut into bands of 8 rows and
// each band into tiles of 8 columns (narrower at the right edge, shorter in the
// last band); ids run tile by tile, row-major inside a tile, so consecutive ids
// are neighbouring pixels and ids 0 .. width * height - 1 cover the frame once.
__device__ __forceinline__ void persistent_pixel(uint32_t id, uint32_t w, uint32_t h, uint32_t& px, uint32_t& py) {
    const uint32_t band = id / (8u * w);
    const uint32_t rows = min(8u, h - 8u * band);
    const uint32_t l = id - band * 8u * w;
    const uint32_t tc = l / (8u * rows);
    const uint32_t tw = min(8u, w - 8u * tc);
    const uint32_t k = l - tc * 8u * rows;
    px = 8u * tc + k % tw;
    py = 8u * band + k / tw;
}

// The megakernel with path regeneration.  Each 64-lane workgroup stays resident
// and runs the path of ptRaygen.rgen:24-104 one ray segment per iteration: trace
// (primary or bounce ray), shading, next-event estimation, then the next
// bounce's BRDF sample and Russian roulette, exactly as pt_megakernel orders
// them (the random draws of a pixel come in the same sequence).  A lane whose
// path has ended writes its pixel and takes the next pixel id from work[0]
// (one atomic per wave for all lanes that need one), so lanes stay busy until
// the frame runs out instead of idling while the wave's longest path finishes.
// The wave that finishes last sets work[0..1] back to zero for the next launch.
template <bool kCount>
__global__ void __launch_bounds__(kTileThreads) __attribute__((amdgpu_waves_per_eu(4))) pt_persistent(DeviceScene sc, TraceConsts c, float* depth_img,
                                                              float* normal_img, uint8_t* material_img,
                                                              uint8_t* albedo_img, float4* illum_img,
                                                              unsigned long long* counters, uint32_t* work) {
    __shared__ int stack_mem[kTraversalStackSize * kTileThreads];
    const uint32_t lane = threadIdx.x;
    int* stack = stack_mem + lane;
    const uint32_t total = c.width * c.height;
    const bool anti_alias = c.output == PBRT_OUTPUT_FINAL_IMAGE && c.sample_number > 0;
    TraversalCounters cnt{0, 0}, cnt_sh{0, 0};
    uint32_t rays = 0, shadow_rays = 0;

    bool alive = false, exhausted = false, primary = false;
    uint32_t pix = 0;
    Rng re{0, 0, 0, 0};
    vec3 ro = make_vec3(0.0f), rd = make_vec3(0.0f), throughput = make_vec3(1.0f), color = make_vec3(0.0f),
         albedo = make_vec3(0.0f);
    int depth_i = 0, trans_depth = 0;
    while (true) {
        const unsigned long long need = __ballot(!alive && !exhausted);
        if (need != 0ull) {
            const int leader = __ffsll(need) - 1;
            uint32_t base = 0;
            if ((int)lane == leader) base = atomicAdd(work, (uint32_t)__popcll(need));
            base = __shfl(base, leader);
            if (!alive && !exhausted) {
                const uint32_t id = base + (uint32_t)__popcll(need & ((1ull << lane) - 1ull));
                if (id < total) {
                    uint32_t px, py;
                    persistent_pixel(id, c.width, c.height, px, py);
                    pix = py * c.width + px;
                    re = rng_init(px, py, c.frame_number);
                    throughput = make_vec3(1.0f);
                    create_ray(c, px, py, anti_alias, re, ro, rd);
                    alive = true;
                    primary = true;
                } else {
                    exhausted = true;
                }
            }
        }
        if (!__any(alive)) break;
        if (!alive) continue;

        SurfaceInfo s;
        vec3 hit_pos;
        uint32_t category;
        ++rays;
        trace_closest<kCount, 1>(sc, ro, rd, 0.001f, 10000.0f, stack, s, hit_pos, category, cnt);
        // throughput is (1, 1, 1) on the primary segment
        const vec3 ne = nee_inline<kCount, 1>(sc, c.light_sampling, hit_pos, primary ? -normalize(rd) : -rd, s,
                                           throughput, re, stack, cnt_sh, shadow_rays);
        bool finish;
        if (primary) {
            color = make_vec3(0.0f);
            color += ne;
            color += s.emissive_color;
            albedo = s.diffuse_color + s.specular_color;
            if (c.write_gbuffer) {
                depth_img[pix] = distance(hit_pos, ro);
                normal_img[2 * pix + 0] = acosf(s.normal.z);
                normal_img[2 * pix + 1] = atan2f(s.normal.y, s.normal.x);
                uint8_t* mat = material_img + 4 * (size_t)pix;
                mat[0] = unorm8((float)category / 255.0f);
                mat[1] = 0;
                mat[2] = 0;
                mat[3] = 0;
                uint8_t* alb = albedo_img + 4 * (size_t)pix;
                alb[0] = unorm8(albedo.x);
                alb[1] = unorm8(albedo.y);
                alb[2] = unorm8(albedo.z);
                alb[3] = 255;
            }
            primary = false;
            depth_i = 0;
            trans_depth = 0;
            finish = is_miss(s);
        } else {
            color += ne + s.emissive_color * throughput;
            finish = is_miss(s);
            ++depth_i;
        }
        // head of the bounce loop (lighting.glsl:263-287 as in pt_megakernel)
        if (!finish) finish = !(depth_i < (int)c.max_depth && trans_depth < 10);
        if (!finish) {
            if (s.illumination_type == 7) {
                --depth_i;
                ++trans_depth;
            }
            const vec3 v = normalize(ro - hit_pos);
            vec3 l;
            float pdf;
            const vec3 brdf = sample_brdf(s, re, v, l, pdf);
            if (brdf == make_vec3(0.0f) || pdf < kEpsilon) {
                finish = true;
            } else {
                float tc = dot(l, s.normal);
                if (s.illumination_type == 7) tc = 1.0f;
                vec3 path_tp = throughput * (brdf * tc) / pdf;
                if ((uint32_t)depth_i > c.min_depth) {
                    if (rng_float(re) < kMinTermination)
                        finish = true;
                    else
                        path_tp /= 1.0f - kMinTermination;
                }
                if (!finish) {
                    throughput = path_tp;
                    ro = hit_pos;
                    rd = l;
                }
            }
        }
        if (finish) {
            color = vclamp(color, 0.0f, kMaxRadiance);
            if (c.output == PBRT_OUTPUT_DEMOD_ILLUMINATION_FLOAT) {
                vec3 demod = color;
                if (!isinf(hit_pos.x)) demod = vmin(color / (albedo + make_vec3(kEpsilon)), 1e3f);
                illum_img[pix] = make_float4(demod.x, demod.y, demod.z, 1.0f);
            } else {
                vec3 fc = color;
                if (c.sample_number > 0) {
                    float4 prev = illum_img[pix];
                    float alpha = 1.0f / (float)(c.sample_number + 1);
                    vec3 pl = srgb_to_linear(make_vec3(prev.x, prev.y, prev.z));
                    fc = vmix(pl, fc, alpha);
                }
                fc = linear_to_srgb(fc);
                fc = vclamp(fc, 0.0f, 1.0f);
                illum_img[pix] = make_float4(fc.x, fc.y, fc.z, 1.0f);
            }
            alive = false;
        }
    }
    if (kCount) {
        atomicAdd(&counters[0], (unsigned long long)rays);
        atomicAdd(&counters[1], (unsigned long long)shadow_rays);
        atomicAdd(&counters[2], (unsigned long long)cnt.nodes);
        atomicAdd(&counters[3], (unsigned long long)cnt.tris);
        atomicAdd(&counters[4], (unsigned long long)cnt_sh.nodes);
        atomicAdd(&counters[5], (unsigned long long)cnt_sh.tris);
    }
    if (lane == 0) {
        // every other wave has drawn its last id before it counts itself finished
        if (atomicAdd(work + 1, 1u) == gridDim.x - 1) {
            atomicExch(work, 0u);
            atomicExch(work + 1, 0u);
        }
    }
}

__global__ void __launch_bounds__(kTileThreads) k_intersect(DeviceScene sc, const float* rays, uint32_t n,
                                                            uint32_t* hits, float* bary) {
    __shared__ int stack_mem[kTraversalStackSize * kTileThreads];
    uint32_t i = blockIdx.x * kTileThreads + threadIdx.x;
    if (i >= n) return;
    const float* r = rays + 8 * (size_t)i;
    TraversalCounters cnt{0, 0};
    Hit h;
    bool any = traverse<false, false>(sc, make_vec3(r[0], r[1], r[2]), make_vec3(r[4], r[5], r[6]), r[3], r[7], h,
                                      stack_mem + threadIdx.x, kTileThreads, cnt);
    uint32_t* o = hits + 4 * (size_t)i;
    if (any) {
        const float4 t0 = sc.tris[3 * h.tri + 0];
        const float4 t1 = sc.tris[3 * h.tri + 1];
        o[0] = __float_as_uint(h.t);
        o[1] = __float_as_uint(t0.w);
        o[2] = __float_as_uint(t1.w);
        o[3] = h.tri;
        bary[2 * (size_t)i] = h.u;
        bary[2 * (size_t)i + 1] = h.v;
    } else {
        o[0] = __float_as_uint(r[7]);
        o[1] = kNoHit;
        o[2] = kNoHit;
        o[3] = kNoHit;
        bary[2 * (size_t)i] = 0.0f;
        bary[2 * (size_t)i + 1] = 0.0f;
    }
}

__global__ void __launch_bounds__(kTileThreads) k_occluded(DeviceScene sc, const float* rays, uint32_t n,
                                                           uint8_t* occ) {
    __shared__ int stack_mem[kTraversalStackSize * kTileThreads];
    uint32_t i = blockIdx.x * kTileThreads + threadIdx.x;
    if (i >= n) return;
    const float* r = rays + 8 * (size_t)i;
    TraversalCounters cnt{0, 0};
    Hit h;
    bool any = traverse<true, false>(sc, make_vec3(r[0], r[1], r[2]), make_vec3(r[4], r[5], r[6]), r[3], r[7], h,
                                     stack_mem + threadIdx.x, kTileThreads, cnt);
    occ[i] = any ? 1 : 0;
}

__global__ void k_rng_sequence(uint32_t width, uint32_t height, uint32_t frame, uint32_t count, uint32_t* out) {
    uint32_t px = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t py = blockIdx.y;
    if (px >= width || py >= height) return;
    Rng re = rng_init(px, py, frame);
    uint32_t* o = out + ((size_t)py * width + px) * count;
    for (uint32_t k = 0; k < count; ++k) o[k] = float_as_uint(rng_float(re));
}

}  // namespace pbrt

using namespace pbrt;

namespace {

// Workgroups of the persistent tracer: as many as stay resident on the current
// device at once (occupancy of the kernel variant times the compute units).
uint32_t persistent_waves(bool counting) {
    int dev = 0;
    PBRT_HIP(hipGetDevice(&dev));
    static std::mutex m;
    static std::map<std::pair<int, bool>, uint32_t> cache;
    std::lock_guard<std::mutex> lock(m);
    auto it = cache.find({dev, counting});
    if (it != cache.end()) return it->second;
    int per_cu = 0, cus = 0;
    if (counting)
        PBRT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, pt_persistent<true>, kTileThreads, 0));
    else
        PBRT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, pt_persistent<false>, kTileThreads, 0));
    PBRT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const uint32_t n = (uint32_t)std::max(1, per_cu) * (uint32_t)std::max(1, cus);
    cache[{dev, counting}] = n;
    return n;
}

}  // namespace

extern "C" int32_t pbrt_trace(const pbrt_scene* scene, const pbrt_trace_params* p, const pbrt_push_constants* pc,
                              const pbrt_trace_images* img, unsigned long long* counters, void* stream) {
    return guard([&] {
        PBRT_REQUIRE(scene && p && pc && img, "pbrt_trace: NULL argument");
        PBRT_REQUIRE(p->width > 0 && p->height > 0, "pbrt_trace: zero-size image");
        PBRT_REQUIRE(p->output == PBRT_OUTPUT_DEMOD_ILLUMINATION_FLOAT || p->output == PBRT_OUTPUT_FINAL_IMAGE,
                     "pbrt_trace: illumination buffer not supported");
        PBRT_REQUIRE(img->illumination, "pbrt_trace: illumination image is NULL");
        PBRT_REQUIRE(!p->write_gbuffer || (img->depth && img->normal && img->material && img->albedo),
                     "pbrt_trace: G-buffer requested but an image is NULL");
        PBRT_REQUIRE(p->tracer == PBRT_TRACER_MEGAKERNEL || p->tracer == PBRT_TRACER_WAVEFRONT ||
                         p->tracer == PBRT_TRACER_PERSISTENT,
                     "pbrt_trace: unknown tracer");
        PBRT_REQUIRE((uint64_t)p->width * p->height < (1ull << 31), "pbrt_trace: more than 2^31 pixels");
        const DeviceScene& sc = scene_device_view(scene);
        TraceConsts c{};
        std::memcpy(c.inv_view, pc->view_inverse, sizeof(c.inv_view));
        std::memcpy(c.inv_proj, pc->proj_inverse, sizeof(c.inv_proj));
        c.width = p->width;
        c.height = p->height;
        c.frame_number = pc->frame_number;
        c.sample_number = pc->sample_number;
        c.max_depth = p->max_recursion_depth;
        c.min_depth = p->min_recursion_depth;
        uint32_t mode = p->light_sampling;
        if (mode == PBRT_LIGHT_SAMPLE_AUTO)
            mode = sc.light_count > 800 ? PBRT_LIGHT_SAMPLE_UNIFORM : PBRT_LIGHT_SAMPLE_SURFACE_STRENGTH;
        PBRT_REQUIRE(mode <= PBRT_LIGHT_SAMPLE_UNIFORM, "pbrt_trace: unknown light sampling mode");
        c.light_sampling = mode;
        c.output = p->output;
        c.write_gbuffer = p->write_gbuffer;
        dim3 grid((p->width + kTileX - 1) / kTileX, (p->height + kTileY - 1) / kTileY);
        hipStream_t st = (hipStream_t)stream;
        float4* illum = reinterpret_cast<float4*>(img->illumination);
        TraceTiming* tm = scene_timing(scene);
        if (p->tracer == PBRT_TRACER_WAVEFRONT) {
            wavefront_trace(sc, c, *img, counters, tm, p->block_threads, p->queue_paths, st);
        } else if (p->tracer == PBRT_TRACER_PERSISTENT) {
            const uint32_t pixels = p->width * p->height;
            const uint32_t waves =
                std::min(persistent_waves(counters != nullptr), (pixels + kTileThreads - 1) / kTileThreads);
            uint32_t* work = scene_work_counter(scene);
            if (tm) tm->frame_begin(st);
            if (counters)
                hipLaunchKernelGGL(pt_persistent<true>, dim3(waves), dim3(kTileThreads), 0, st, sc, c, img->depth,
                                   img->normal, img->material, img->albedo, illum, counters, work);
            else
                hipLaunchKernelGGL(pt_persistent<false>, dim3(waves), dim3(kTileThreads), 0, st, sc, c, img->depth,
                                   img->normal, img->material, img->albedo, illum, counters, work);
            if (tm) tm->mark(kTimeMegakernel, st);
        } else {
            if (tm) tm->frame_begin(st);
            // a tree of depth d needs d - 1 stack entries at most
            const bool short_stack = scene_bvh_depth(scene) <= (uint32_t)kShortStackSize;
            auto kernel = counters ? (short_stack ? pt_megakernel<true, kShortStackSize>
                                                  : pt_megakernel<true, kTraversalStackSize>)
                                   : (short_stack ? pt_megakernel<false, kShortStackSize>
                                                  : pt_megakernel<false, kTraversalStackSize>);
            hipLaunchKernelGGL(kernel, grid, dim3(kTileThreads), 0, st, sc, c, img->depth, img->normal,
                               img->material, img->albedo, illum, counters);
            if (tm) tm->mark(kTimeMegakernel, st);
        }
        PBRT_HIP(hipGetLastError());
    });
}

extern "C" int32_t pbrt_intersect(const pbrt_scene* scene, const float* rays, uint32_t n, uint32_t* hits,
                                  float* barycentrics, void* stream) {
    return guard([&] {
        PBRT_REQUIRE(scene && (n == 0 || (rays && hits && barycentrics)), "pbrt_intersect: NULL argument");
        if (n == 0) return;
        const DeviceScene& sc = scene_device_view(scene);
        hipLaunchKernelGGL(k_intersect, dim3((n + kTileThreads - 1) / kTileThreads), dim3(kTileThreads), 0,
                           (hipStream_t)stream, sc, rays, n, hits, barycentrics);
        PBRT_HIP(hipGetLastError());
    });
}

extern "C" int32_t pbrt_occluded(const pbrt_scene* scene, const float* rays, uint32_t n, uint8_t* occluded,
                                 void* stream) {
    return guard([&] {
        PBRT_REQUIRE(scene && (n == 0 || (rays && occluded)), "pbrt_occluded: NULL argument");
        if (n == 0) return;
        const DeviceScene& sc = scene_device_view(scene);
        hipLaunchKernelGGL(k_occluded, dim3((n + kTileThreads - 1) / kTileThreads), dim3(kTileThreads), 0,
                           (hipStream_t)stream, sc, rays, n, occluded);
        PBRT_HIP(hipGetLastError());
    });
}

extern "C" int32_t pbrt_rng_sequence(uint32_t width, uint32_t height, uint32_t frame_index, uint32_t count,
                                     uint32_t* out, void* stream) {
    return guard([&] {
        PBRT_REQUIRE(out != nullptr, "pbrt_rng_sequence: NULL output");
        if (width == 0 || height == 0 || count == 0) return;
        dim3 grid((width + 255) / 256, height);
        hipLaunchKernelGGL(k_rng_sequence, grid, dim3(256), 0, (hipStream_t)stream, width, height, frame_index, count,
                           out);
        PBRT_HIP(hipGetLastError());
    });
}

extern "C" int32_t pbrt_autotune_trace(const pbrt_scene* scene, const pbrt_trace_params* p,
                                       const pbrt_push_constants* pc, const pbrt_trace_images* img, uint32_t repeats,
                                       pbrt_trace_tuning* out, void* stream) {
    return guard([&] {
        PBRT_REQUIRE(scene && p && pc && img && out, "pbrt_autotune_trace: NULL argument");
        PBRT_REQUIRE(repeats >= 1 && repeats <= 1000, "pbrt_autotune_trace: repeats must be in [1, 1000]");
        hipStream_t st = (hipStream_t)stream;
        *out = pbrt_trace_tuning{};
        struct Cand {
            uint32_t tracer, block, queue;
        };
        std::vector<Cand> cands{{PBRT_TRACER_MEGAKERNEL, 0, 0}, {PBRT_TRACER_PERSISTENT, 0, 0}};
        const uint32_t pixels = p->width * p->height;
        if (img->workspace) {
            for (uint32_t block : {64u, 128u, 256u})
                for (uint32_t div : {1u, 2u, 4u})
                    cands.push_back({PBRT_TRACER_WAVEFRONT, block, div == 1 ? 0u : (pixels + div - 1) / div});
        }
        PBRT_REQUIRE(cands.size() <= PBRT_TUNING_MAX_CANDIDATES, "pbrt_autotune_trace: too many candidates");
        // the events are released on every exit, also when a candidate is rejected
        struct Events {
            hipEvent_t e[2] = {nullptr, nullptr};
            ~Events() {
                for (hipEvent_t ev : e)
                    if (ev) (void)hipEventDestroy(ev);
            }
        } ev;
        PBRT_HIP(hipEventCreate(&ev.e[0]));
        PBRT_HIP(hipEventCreate(&ev.e[1]));
        hipEvent_t e0 = ev.e[0], e1 = ev.e[1];
        float best = 3.4e38f;
        for (size_t i = 0; i < cands.size(); ++i) {
            pbrt_trace_params q = *p;
            q.tracer = cands[i].tracer;
            q.block_threads = cands[i].block;
            q.queue_paths = cands[i].queue;
            int32_t rc = pbrt_trace(scene, &q, pc, img, nullptr, stream);  // warm-up
            if (rc != PBRT_OK) fail(rc, std::string("pbrt_autotune_trace: ") + pbrt_last_error());
            PBRT_HIP(hipEventRecord(e0, st));
            for (uint32_t r = 0; r < repeats; ++r) {
                rc = pbrt_trace(scene, &q, pc, img, nullptr, stream);
                if (rc != PBRT_OK) fail(rc, std::string("pbrt_autotune_trace: ") + pbrt_last_error());
            }
            PBRT_HIP(hipEventRecord(e1, st));
            PBRT_HIP(hipEventSynchronize(e1));
            float ms = 0.0f;
            PBRT_HIP(hipEventElapsedTime(&ms, e0, e1));
            ms /= (float)repeats;
            out->tracer[i] = q.tracer;
            out->block_threads[i] = q.block_threads;
            out->queue_paths[i] = q.queue_paths;
            out->ms[i] = ms;
            if (ms < best) {
                best = ms;
                out->best = q;
            }
        }
        out->candidates = (uint32_t)cands.size();
    });
}
