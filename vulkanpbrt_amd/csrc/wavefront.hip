// Wavefront path tracer: the per-pixel program of ptRaygen.rgen split into
// queue-driven launches per bounce so that traversal runs in lean kernels and
// shading in its own:
//
//   k_wf_generate   camera rays (createRay) for every pixel, 8x8-tile ordered queue
//   k_wf_extend     closest-hit traversal of the current ray queue -> hit records
//   k_wf_shade      hit shading (ptClosesthit / ptMiss), G-buffer at the primary
//                   vertex, next-event estimation (light sample -> shadow queue with
//                   the unshadowed contribution), emission, BRDF sample + Russian
//                   roulette -> compacted ray queue of the next bounce
//   k_wf_shadow     any-hit traversal of the shadow queue; unoccluded samples add
//                   their contribution
//   k_wf_finalize   clamp, demodulation / progressive final image
//
// Queue appends are wave-aggregated (one atomic per wave).  Queue sizes stay on
// the device; every launch is sized for the worst case and threads past the
// live count exit, so a frame needs no host synchronisation.  The random stream
// of each path is consumed in the same order as the per-pixel program:
// light sample, then BRDF sample (on a copy) and the roulette draw.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "capi_common.hpp"
#include "scene_internal.hpp"
#include "trace_common.hpp"

namespace pbrt {

constexpr int kWfThreads = 64;
constexpr int kMaxWfBounces = 32;
constexpr uint32_t kDeadPath = 0xFFFFFFFFu;

struct WfBuffers {
    float4* col;         // accumulated radiance, per path
    float4* alb;         // primary albedo (diffuse + specular), per path
    // path state carried with its ray through the queues (read and written at the
    // queue slot, so coherently): RNG state, and throughput with w holding the
    // next bounce's loop index (i + 1 >= 0: a transmissive hit at i = 0 leaves
    // i = -1; low 16 bits) and the transmissive depth (high 16 bits)
    uint4* ray_rng[2];
    float4* ray_tp[2];
    float4* ray_o[2];
    float4* ray_d[2];
    uint32_t* ray_path[2];
    float4* hit;
    float4* sh_o;
    float4* sh_d;
    float4* sh_pend;
    uint32_t* sh_path;
    uint32_t* counters;  // [0, kMaxWfBounces): ray queue sizes; [kMaxWfBounces, 2k): shadow queue sizes
    uint32_t capacity;
};

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

static size_t frame_tiles(uint32_t w, uint32_t h) {
    return (size_t)((w + kTileX - 1) / kTileX) * ((h + kTileY - 1) / kTileY);
}

// queue_tiles: tiles traced at once (0 = the whole frame)
static size_t wf_layout(uint32_t w, uint32_t h, size_t queue_tiles, char* base, WfBuffers* b) {
    const size_t tiles = queue_tiles ? std::min(queue_tiles, frame_tiles(w, h)) : frame_tiles(w, h);
    const size_t n = tiles * kTileThreads;  // queue capacity: the tile-padded pixel count
    const size_t np = (size_t)w * h;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        char* p = base ? base + off : nullptr;
        off = align256(off + bytes);
        return p;
    };
    WfBuffers t{};
    t.col = (float4*)take(np * 16);
    t.alb = (float4*)take(np * 16);
    for (int q = 0; q < 2; ++q) {
        t.ray_rng[q] = (uint4*)take(n * 16);
        t.ray_tp[q] = (float4*)take(n * 16);
        t.ray_o[q] = (float4*)take(n * 16);
        t.ray_d[q] = (float4*)take(n * 16);
        t.ray_path[q] = (uint32_t*)take(n * 4);
    }
    t.hit = (float4*)take(n * 16);
    t.sh_o = (float4*)take(n * 16);
    t.sh_d = (float4*)take(n * 16);
    t.sh_pend = (float4*)take(n * 16);
    t.sh_path = (uint32_t*)take(n * 4);
    t.counters = (uint32_t*)take(2 * kMaxWfBounces * 4);
    t.capacity = (uint32_t)n;
    if (b) *b = t;
    return off;
}

// Append `want` from each lane to a queue with one atomic per wave.
__device__ __forceinline__ uint32_t wave_append(uint32_t* counter, bool want) {
    const unsigned long long mask = __ballot(want);
    if (mask == 0) return 0;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((long long)mask) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = __shfl(base, leader);
    const unsigned long long below = lane ? (mask & ((~0ull) >> (64 - lane))) : 0ull;
    return base + (uint32_t)__popcll(below);
}

// slice_rays: queue entries of this slice (its tiles x 64 threads)
__global__ void k_wf_init(WfBuffers b, uint32_t slice_rays) {
    const int i = threadIdx.x;
    if (i < 2 * kMaxWfBounces) b.counters[i] = i == 0 ? slice_rays : 0u;
}

__global__ void __launch_bounds__(kWfThreads) k_wf_generate(TraceConsts c, WfBuffers b, uint32_t tile_begin,
                                                            uint32_t slice_rays) {
    const uint32_t q = blockIdx.x * kWfThreads + threadIdx.x;
    if (q >= slice_rays) return;
    const uint32_t tiles_x = (c.width + kTileX - 1) / kTileX;
    const uint32_t tile = tile_begin + q / kTileThreads, lid = q % kTileThreads;
    const uint32_t px = (tile % tiles_x) * kTileX + lid % kTileX;
    const uint32_t py = (tile / tiles_x) * kTileY + lid / kTileX;
    if (px >= c.width || py >= c.height) {
        b.ray_path[0][q] = kDeadPath;
        return;
    }
    const uint32_t path = py * c.width + px;
    Rng re = rng_init(px, py, c.frame_number);
    bool anti_alias = c.output == PBRT_OUTPUT_FINAL_IMAGE && c.sample_number > 0;
    vec3 pos, dir;
    create_ray(c, px, py, anti_alias, re, pos, dir);
    b.ray_o[0][q] = make_float4(pos.x, pos.y, pos.z, 0.001f);
    b.ray_d[0][q] = make_float4(dir.x, dir.y, dir.z, 10000.0f);
    b.ray_path[0][q] = path;
    b.ray_rng[0][q] = make_uint4(re.x, re.y, re.z, re.w);
    b.ray_tp[0][q] = make_float4(1.0f, 1.0f, 1.0f, __uint_as_float(0u));
    b.col[path] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

template <bool kCount, int kThreads>
__global__ void __launch_bounds__(kThreads) k_wf_extend(DeviceScene sc, WfBuffers b, int bounce,
                                                        unsigned long long* counters) {
    __shared__ int stack_mem[kTraversalStackSize * kThreads];
    const uint32_t q = blockIdx.x * kThreads + threadIdx.x;
    const int qi = bounce & 1;
    TraversalCounters cnt{0, 0};
    if (q < b.counters[bounce] && b.ray_path[qi][q] != kDeadPath) {
        const float4 o = b.ray_o[qi][q];
        const float4 d = b.ray_d[qi][q];
        Hit h;
        traverse<false, kCount>(sc, make_vec3(o.x, o.y, o.z), make_vec3(d.x, d.y, d.z), o.w, d.w, h,
                                stack_mem + threadIdx.x, kThreads, cnt);
        b.hit[q] = make_float4(h.t, __uint_as_float(h.tri), h.u, h.v);
    }
    if (kCount) {
        atomicAdd(&counters[2], (unsigned long long)cnt.nodes);
        atomicAdd(&counters[3], (unsigned long long)cnt.tris);
    }
}

__global__ void __launch_bounds__(kWfThreads) k_wf_shade(DeviceScene sc, TraceConsts c, WfBuffers b, int bounce,
                                                         float* depth_img, float* normal_img, uint8_t* material_img,
                                                         uint8_t* albedo_img) {
    const uint32_t q = blockIdx.x * kWfThreads + threadIdx.x;
    const int qi = bounce & 1, qn = qi ^ 1;
    bool live = q < b.counters[bounce];
    uint32_t path = live ? b.ray_path[qi][q] : kDeadPath;
    live = live && path != kDeadPath;
    bool want_shadow = false, want_ray = false;
    float4 sh_o{}, sh_d{}, sh_p{}, nr_o{}, nr_d{}, nr_tp{};
    uint4 nr_rng{};
    if (live) {
        const float4 ro = b.ray_o[qi][q];
        const float4 rd = b.ray_d[qi][q];
        const float4 hr = b.hit[q];
        const vec3 origin = make_vec3(ro.x, ro.y, ro.z);
        const vec3 dir = make_vec3(rd.x, rd.y, rd.z);
        Hit h;
        h.t = hr.x;
        h.tri = __float_as_uint(hr.y);
        h.u = hr.z;
        h.v = hr.w;
        SurfaceInfo s;
        vec3 pos;
        uint32_t category;
        if (h.tri != kNoHit)
            shade_hit(sc, h, dir, s, pos, category);
        else
            miss_surface(s, pos, category);
        const uint4 r4 = b.ray_rng[qi][q];
        Rng re{r4.x, r4.y, r4.z, r4.w};
        const float4 tp4 = b.ray_tp[qi][q];
        vec3 tp = make_vec3(tp4.x, tp4.y, tp4.z);
        const uint32_t st = __float_as_uint(tp4.w);
        const vec3 o = bounce == 0 ? -normalize(dir) : -dir;
        if (bounce == 0) {
            vec3 albedo = s.diffuse_color + s.specular_color;
            b.alb[path] = make_float4(albedo.x, albedo.y, albedo.z, 0.0f);
            if (c.write_gbuffer) {
                depth_img[path] = distance(pos, origin);
                normal_img[2 * (size_t)path + 0] = acosf(s.normal.z);
                normal_img[2 * (size_t)path + 1] = atan2f(s.normal.y, s.normal.x);
                uchar4 m;
                m.x = unorm8((float)category / 255.0f);
                m.y = m.z = m.w = 0;
                reinterpret_cast<uchar4*>(material_img)[path] = m;
                uchar4 a;
                a.x = unorm8(albedo.x);
                a.y = unorm8(albedo.y);
                a.z = unorm8(albedo.z);
                a.w = 255;
                reinterpret_cast<uchar4*>(albedo_img)[path] = a;
            }
        }
        // next-event estimation; the path's radiance is read and written only when
        // something is added (sky, emission), in the reference's order of additions
        vec3 add_sky = make_vec3(0.0f);
        if (is_miss(s)) {
            add_sky = sky_color(-o) * tp;
        } else {
            LightSample ls = sample_light(sc, c.light_sampling, pos, s.normal, re);
            if (ls.valid) {
                vec3 pend = nee_unshadowed(ls, o, s, tp);
                if (pend != make_vec3(0.0f)) {
                    want_shadow = true;
                    sh_o = make_float4(pos.x, pos.y, pos.z, 0.001f);
                    sh_d = make_float4(ls.l.x, ls.l.y, ls.l.z, ls.tmax);
                    sh_p = make_float4(pend.x, pend.y, pend.z, 0.0f);
                }
            }
        }
        const vec3 add_emit = bounce == 0 ? s.emissive_color : s.emissive_color * tp;
        if (add_sky != make_vec3(0.0f) || add_emit != make_vec3(0.0f)) {
            const float4 c4 = b.col[path];
            vec3 col = make_vec3(c4.x, c4.y, c4.z);
            col += add_sky;
            col += add_emit;
            b.col[path] = make_float4(col.x, col.y, col.z, 0.0f);
        }
        // the bounce loop of ptRaygen.rgen:66-75
        if (!is_miss(s)) {
            int i = bounce == 0 ? 0 : (int)(st & 0xFFFFu);
            int trans = (int)(st >> 16);
            if (i < (int)c.max_depth && trans < 10) {
                if (s.illumination_type == 7) {
                    --i;
                    ++trans;
                }
                vec3 v = normalize(origin - pos);
                vec3 l;
                float pdf;
                vec3 brdf = sample_brdf(s, re, v, l, pdf);
                if (!(brdf == make_vec3(0.0f) || pdf < kEpsilon)) {
                    float tc = dot(l, s.normal);
                    if (s.illumination_type == 7) tc = 1.0f;
                    vec3 ptp = tp * (brdf * tc) / pdf;
                    bool alive = true;
                    if ((uint32_t)i > c.min_depth) {
                        if (rng_float(re) < kMinTermination)
                            alive = false;
                        else
                            ptp /= 1.0f - kMinTermination;
                    }
                    if (alive) {
                        want_ray = true;
                        nr_o = make_float4(pos.x, pos.y, pos.z, 0.001f);
                        nr_d = make_float4(l.x, l.y, l.z, 10000.0f);
                        nr_tp = make_float4(ptp.x, ptp.y, ptp.z,
                                            __uint_as_float((uint32_t)(i + 1) | (uint32_t)trans << 16));
                        nr_rng = make_uint4(re.x, re.y, re.z, re.w);
                    }
                }
            }
        }
    }
    const uint32_t rs = wave_append(&b.counters[bounce + 1], want_ray);
    if (want_ray) {
        b.ray_o[qn][rs] = nr_o;
        b.ray_d[qn][rs] = nr_d;
        b.ray_path[qn][rs] = path;
        b.ray_rng[qn][rs] = nr_rng;
        b.ray_tp[qn][rs] = nr_tp;
    }
    const uint32_t ss = wave_append(&b.counters[kMaxWfBounces + bounce], want_shadow);
    if (want_shadow) {
        b.sh_o[ss] = sh_o;
        b.sh_d[ss] = sh_d;
        b.sh_pend[ss] = sh_p;
        b.sh_path[ss] = path;
    }
}

template <bool kCount, int kThreads>
__global__ void __launch_bounds__(kThreads) k_wf_shadow(DeviceScene sc, WfBuffers b, int bounce,
                                                        unsigned long long* counters) {
    __shared__ int stack_mem[kTraversalStackSize * kThreads];
    const uint32_t q = blockIdx.x * kThreads + threadIdx.x;
    TraversalCounters cnt{0, 0};
    if (q < b.counters[kMaxWfBounces + bounce]) {
        const float4 o = b.sh_o[q];
        const float4 d = b.sh_d[q];
        Hit h;
        bool occ = traverse<true, kCount>(sc, make_vec3(o.x, o.y, o.z), make_vec3(d.x, d.y, d.z), o.w, d.w, h,
                                          stack_mem + threadIdx.x, kThreads, cnt);
        if (!occ) {
            const uint32_t path = b.sh_path[q];
            const float4 p = b.sh_pend[q];
            float4 c = b.col[path];
            b.col[path] = make_float4(c.x + p.x, c.y + p.y, c.z + p.z, 0.0f);
        }
    }
    if (kCount) {
        atomicAdd(&counters[4], (unsigned long long)cnt.nodes);
        atomicAdd(&counters[5], (unsigned long long)cnt.tris);
    }
}

__global__ void k_wf_finalize(TraceConsts c, WfBuffers b, float4* illum_img) {
    const uint32_t path = blockIdx.x * blockDim.x + threadIdx.x;
    if (path >= c.width * c.height) return;
    float4 c4 = b.col[path];
    vec3 color = vclamp(make_vec3(c4.x, c4.y, c4.z), 0.0f, kMaxRadiance);
    if (c.output == PBRT_OUTPUT_DEMOD_ILLUMINATION_FLOAT) {
        // the miss position (1e10) is finite, so demodulation always applies (ptRaygen.rgen:85)
        float4 a = b.alb[path];
        vec3 demod = vmin(color / (make_vec3(a.x, a.y, a.z) + make_vec3(kEpsilon)), 1e3f);
        illum_img[path] = make_float4(demod.x, demod.y, demod.z, 1.0f);
    } else {
        vec3 fc = color;
        if (c.sample_number > 0) {
            float4 prev = illum_img[path];
            float alpha = 1.0f / (float)(c.sample_number + 1);
            fc = vmix(srgb_to_linear(make_vec3(prev.x, prev.y, prev.z)), fc, alpha);
        }
        fc = vclamp(linear_to_srgb(fc), 0.0f, 1.0f);
        illum_img[path] = make_float4(fc.x, fc.y, fc.z, 1.0f);
    }
}

__global__ void k_wf_count(WfBuffers b, int bounces, uint32_t pixels, unsigned long long* counters) {
    if (threadIdx.x != 0) return;
    unsigned long long rays = pixels, shadow = 0;
    for (int i = 1; i < bounces; ++i) rays += b.counters[i];
    for (int i = 0; i < bounces; ++i) shadow += b.counters[kMaxWfBounces + i];
    counters[0] += rays;
    counters[1] += shadow;
}

size_t wavefront_workspace_size(uint32_t w, uint32_t h) { return wf_layout(w, h, 0, nullptr, nullptr); }

template <int kThreads>
static void launch_traversal(bool shadow, bool count, const DeviceScene& sc, const WfBuffers& b, int bounce,
                             unsigned long long* counters, hipStream_t st) {
    const dim3 grid((b.capacity + kThreads - 1) / kThreads);
    if (shadow) {
        if (count)
            hipLaunchKernelGGL((k_wf_shadow<true, kThreads>), grid, dim3(kThreads), 0, st, sc, b, bounce, counters);
        else
            hipLaunchKernelGGL((k_wf_shadow<false, kThreads>), grid, dim3(kThreads), 0, st, sc, b, bounce, counters);
    } else {
        if (count)
            hipLaunchKernelGGL((k_wf_extend<true, kThreads>), grid, dim3(kThreads), 0, st, sc, b, bounce, counters);
        else
            hipLaunchKernelGGL((k_wf_extend<false, kThreads>), grid, dim3(kThreads), 0, st, sc, b, bounce, counters);
    }
}

static void traversal(uint32_t block_threads, bool shadow, bool count, const DeviceScene& sc, const WfBuffers& b,
                      int bounce, unsigned long long* counters, hipStream_t st) {
    if (block_threads == 256)
        launch_traversal<256>(shadow, count, sc, b, bounce, counters, st);
    else if (block_threads == 128)
        launch_traversal<128>(shadow, count, sc, b, bounce, counters, st);
    else
        launch_traversal<64>(shadow, count, sc, b, bounce, counters, st);
}

void wavefront_trace(const DeviceScene& sc, const TraceConsts& c, const pbrt_trace_images& img,
                     unsigned long long* counters, TraceTiming* tm, uint32_t block_threads, uint32_t queue_paths,
                     hipStream_t st) {
    PBRT_REQUIRE(block_threads == 0 || block_threads == 64 || block_threads == 128 || block_threads == 256,
                 "pbrt_trace: block_threads must be 0, 64, 128 or 256");
    const size_t all_tiles = frame_tiles(c.width, c.height);
    const size_t slice_tiles = queue_paths ? std::min(all_tiles, ((size_t)queue_paths + kTileThreads - 1) / kTileThreads)
                                           : all_tiles;
    WfBuffers b;
    size_t need = wf_layout(c.width, c.height, slice_tiles, static_cast<char*>(img.workspace), &b);
    PBRT_REQUIRE(img.workspace != nullptr && img.workspace_bytes >= need,
                 "pbrt_trace: the wavefront tracer needs a workspace of pbrt_trace_workspace_size() bytes");
    const int bounces = (int)c.max_depth + 1 + (sc.has_transmissive ? 10 : 0);
    PBRT_REQUIRE(bounces < kMaxWfBounces, "pbrt_trace: max_recursion_depth too large for the wavefront tracer");
    const dim3 grid((b.capacity + kWfThreads - 1) / kWfThreads);
    float4* illum = reinterpret_cast<float4*>(img.illumination);
    const uint32_t tiles_x = (c.width + kTileX - 1) / kTileX;
    if (tm) tm->frame_begin(st);
    for (size_t t0 = 0; t0 < all_tiles; t0 += slice_tiles) {
        const size_t nt = std::min(slice_tiles, all_tiles - t0);
        const uint32_t slice_rays = (uint32_t)(nt * kTileThreads);
        hipLaunchKernelGGL(k_wf_init, dim3(1), dim3(2 * kMaxWfBounces), 0, st, b, slice_rays);
        hipLaunchKernelGGL(k_wf_generate, grid, dim3(kWfThreads), 0, st, c, b, (uint32_t)t0, slice_rays);
        if (tm) tm->mark(kTimeOther, st);
        for (int bounce = 0; bounce < bounces; ++bounce) {
            traversal(block_threads, false, counters != nullptr, sc, b, bounce, counters, st);
            if (tm) tm->mark(kTimeExtend, st);
            hipLaunchKernelGGL(k_wf_shade, grid, dim3(kWfThreads), 0, st, sc, c, b, bounce, img.depth, img.normal,
                               img.material, img.albedo);
            if (tm) tm->mark(kTimeShade, st);
            traversal(block_threads, true, counters != nullptr, sc, b, bounce, counters, st);
            if (tm) tm->mark(kTimeShadow, st);
        }
        if (counters) {
            // primary rays: the slice's pixels inside the image
            uint32_t pixels = 0;
            for (size_t t = t0; t < t0 + nt; ++t) {
                const uint32_t x = (uint32_t)(t % tiles_x) * kTileX, y = (uint32_t)(t / tiles_x) * kTileY;
                pixels += std::min<uint32_t>(kTileX, c.width - x) * std::min<uint32_t>(kTileY, c.height - y);
            }
            hipLaunchKernelGGL(k_wf_count, dim3(1), dim3(64), 0, st, b, bounces, pixels, counters);
        }
    }
    const uint32_t np = c.width * c.height;
    hipLaunchKernelGGL(k_wf_finalize, dim3((np + 255) / 256), dim3(256), 0, st, c, b, illum);
    if (tm) tm->mark(kTimeOther, st);
}

}  // namespace pbrt

using namespace pbrt;

extern "C" int32_t pbrt_trace_workspace_size(uint32_t width, uint32_t height, size_t* out) {
    return guard([&] {
        PBRT_REQUIRE(out, "pbrt_trace_workspace_size: NULL output");
        PBRT_REQUIRE(width > 0 && height > 0, "pbrt_trace_workspace_size: zero-size image");
        *out = wavefront_workspace_size(width, height);
    });
}
