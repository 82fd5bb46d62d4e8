// Scene packing and BVH build: the host half of PBRTPipeline::setup_pipeline.
//
// Packing rules follow RayTracingSceneDescriptorCreationVisitor
// (source/scene/RayTracingVisitor.cpp):
//   * mesh buffers are stored once per mesh; instances reference them (:36-48)
//   * a zero first normal requests area-weighted vertex normals (:52-118)
//   * missing uvs get the per-triangle (0,0),(0,1),(1,0) assignment (:122-159)
//   * every triangle of an emissive mesh instance becomes an area light in world
//     space with attenuation (0, 0, 1) (:170-221)
//   * a diffuse texture with any alpha < 0.01 marks the instance alpha tested (:319-343)
//   * no lights => default sun, then the inclusive strength prefix sum (:414-441)
// and the light-sampling switch of PBRTPipeline.cpp:92-96 (> 800 lights => uniform).
//
// Instances are flattened into one world-space triangle set with a single BVH
// (instead of the driver's BLAS-per-mesh + TLAS): each triangle keeps its
// (instance, primitive) pair so hit shading still reads the instance's mesh
// buffers and object matrix exactly as ptClosesthit.rchit does.
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "bvh_host.hpp"
#include "bvh_quant.hpp"
#include "lbvh.hpp"
#include "ploc.hpp"
#include "capi_common.hpp"
#include "scene_internal.hpp"

namespace pbrt {

static thread_local std::string g_last_error;
void set_last_error(const std::string& msg) { g_last_error = msg; }

namespace {

void mat4_mul_point_d(const float* m, const float* p, double* out) {
    for (int r = 0; r < 3; ++r)
        out[r] = (double)m[r] * p[0] + (double)m[4 + r] * p[1] + (double)m[8 + r] * p[2] + (double)m[12 + r];
}

// transpose(inverse(M3)) of the upper-left 3x3, in double, column-major out.
void normal_matrix(const float* m, float* out9) {
    double a[3][3];  // a[row][col]
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) a[r][c] = m[4 * c + r];
    double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) - a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                 a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
    double inv[3][3];
    inv[0][0] = (a[1][1] * a[2][2] - a[1][2] * a[2][1]) / det;
    inv[0][1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) / det;
    inv[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) / det;
    inv[1][0] = (a[1][2] * a[2][0] - a[1][0] * a[2][2]) / det;
    inv[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) / det;
    inv[1][2] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) / det;
    inv[2][0] = (a[1][0] * a[2][1] - a[1][1] * a[2][0]) / det;
    inv[2][1] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) / det;
    inv[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) / det;
    // N = transpose(inv): N[row][col] = inv[col][row]; column-major storage out[col*3+row]
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) out9[3 * c + r] = (float)inv[c][r];
}

uint32_t read_index(const pbrt_mesh_desc& m, uint32_t i) {
    if (m.index_stride == 2) return static_cast<const uint16_t*>(m.indices)[i];
    return static_cast<const uint32_t*>(m.indices)[i];
}

}  // namespace

void pack_scene(const pbrt_scene_desc& d, PackedScene& ps) {
    PBRT_REQUIRE(d.mesh_count == 0 || d.meshes, "scene desc: meshes is NULL");
    PBRT_REQUIRE(d.instance_count == 0 || d.instances, "scene desc: instances is NULL");
    PBRT_REQUIRE(d.texture_count == 0 || d.textures, "scene desc: textures is NULL");
    PBRT_REQUIRE(d.light_count == 0 || d.lights, "scene desc: lights is NULL");

    // texture table: slot 0 is the 1x1 white default (RayTracingVisitor.cpp:5-15)
    ps.textures.clear();
    ps.texels.clear();
    ps.textures.push_back(DeviceTexture{0, 1, 1, 0});
    ps.texels.push_back(0xFFFFFFFFu);
    std::vector<bool> tex_has_alpha_hole(d.texture_count + 1, false);
    for (uint32_t t = 0; t < d.texture_count; ++t) {
        const pbrt_texture_desc& td = d.textures[t];
        PBRT_REQUIRE(td.width > 0 && td.height > 0 && td.rgba8, "texture " + std::to_string(t) + " is empty");
        DeviceTexture dt{(uint32_t)ps.texels.size(), td.width, td.height, 0};
        size_t n = (size_t)td.width * td.height;
        size_t off = ps.texels.size();
        ps.texels.resize(off + n);
        std::memcpy(&ps.texels[off], td.rgba8, n * 4);
        bool hole = false;
        for (size_t i = 0; i < n && !hole; ++i)
            if ((float)td.rgba8[4 * i + 3] < 0.01f * 255.0f) hole = true;
        tex_has_alpha_hole[t + 1] = hole;
        ps.textures.push_back(dt);
    }
    auto tex_slot = [&](int32_t t) -> int32_t {
        PBRT_REQUIRE(t >= -1 && t < (int32_t)d.texture_count, "mesh texture index out of range");
        return t + 1;
    };

    // meshes
    ps.meshes.clear();
    ps.materials.clear();
    ps.positions.clear();
    ps.normals.clear();
    ps.uvs.clear();
    ps.index_words.clear();
    ps.mesh_opaque.clear();
    for (uint32_t mi = 0; mi < d.mesh_count; ++mi) {
        const pbrt_mesh_desc& m = d.meshes[mi];
        PBRT_REQUIRE(m.vertex_count > 0 && m.positions, "mesh " + std::to_string(mi) + " has no positions");
        PBRT_REQUIRE(m.index_count % 3 == 0 && m.index_count > 0 && m.indices,
                     "mesh " + std::to_string(mi) + " needs a non-empty triangle index list");
        PBRT_REQUIRE(m.index_stride == 2 || m.index_stride == 4, "index_stride must be 2 or 4");
        for (uint32_t i = 0; i < m.index_count; ++i)
            PBRT_REQUIRE(read_index(m, i) < m.vertex_count, "mesh " + std::to_string(mi) + " index out of range");
        DeviceMesh dm;
        dm.vertex_offset = (uint32_t)(ps.positions.size() / 3);
        dm.index_word_offset = (uint32_t)ps.index_words.size();
        dm.index_count = m.index_count;
        dm.index_stride = m.index_stride;
        dm.diffuse_tex = tex_slot(m.diffuse_texture);
        dm.normal_tex = tex_slot(m.normal_texture);
        dm.specular_tex = tex_slot(m.specular_texture);
        dm.emissive_tex = tex_slot(m.emissive_texture);
        ps.positions.insert(ps.positions.end(), m.positions, m.positions + 3 * (size_t)m.vertex_count);

        // normals: area-weighted accumulation when the first normal is zero
        std::vector<float> nors(3 * (size_t)m.vertex_count, 0.0f);
        if (m.normals) std::memcpy(nors.data(), m.normals, nors.size() * sizeof(float));
        if (nors[0] * nors[0] + nors[1] * nors[1] + nors[2] * nors[2] == 0.0f) {
            std::vector<float> wsum(m.vertex_count, 0.0f);
            for (uint32_t tri = 0; tri < m.index_count / 3; ++tri) {
                uint32_t ia = read_index(m, 3 * tri), ib = read_index(m, 3 * tri + 1), ic = read_index(m, 3 * tri + 2);
                const float* a = m.positions + 3 * (size_t)ia;
                const float* b = m.positions + 3 * (size_t)ib;
                const float* c = m.positions + 3 * (size_t)ic;
                vec3 e1 = make_vec3(b[0] - a[0], b[1] - a[1], b[2] - a[2]);
                vec3 e2 = make_vec3(c[0] - a[0], c[1] - a[1], c[2] - a[2]);
                vec3 fn = cross(e1, e2);
                float w = length(fn);
                fn = fn / w;
                for (uint32_t idx : {ia, ib, ic}) {
                    float nw = w + wsum[idx];
                    float ka = wsum[idx] / nw, kb = w / nw;
                    for (int k = 0; k < 3; ++k) {
                        float fk = k == 0 ? fn.x : (k == 1 ? fn.y : fn.z);
                        nors[3 * idx + k] = nors[3 * idx + k] * ka + fk * kb;
                    }
                    wsum[idx] = nw;
                }
            }
        }
        ps.normals.insert(ps.normals.end(), nors.begin(), nors.end());

        // uvs: default per-triangle assignment when absent
        std::vector<float> uv(2 * (size_t)m.vertex_count, 0.0f);
        if (m.uvs) {
            std::memcpy(uv.data(), m.uvs, uv.size() * sizeof(float));
        } else {
            const float corner[3][2] = {{0, 0}, {0, 1}, {1, 0}};
            for (uint32_t tri = 0; tri < m.index_count / 3; ++tri)
                for (int k = 0; k < 3; ++k) {
                    uint32_t idx = read_index(m, 3 * tri + k);
                    uv[2 * idx] = corner[k][0];
                    uv[2 * idx + 1] = corner[k][1];
                }
        }
        ps.uvs.insert(ps.uvs.end(), uv.begin(), uv.end());

        // index words: u32 indices verbatim, u16 indices packed two per word (little endian)
        if (m.index_stride == 4) {
            const uint32_t* src = static_cast<const uint32_t*>(m.indices);
            ps.index_words.insert(ps.index_words.end(), src, src + m.index_count);
        } else {
            const uint16_t* src = static_cast<const uint16_t*>(m.indices);
            for (uint32_t i = 0; i < m.index_count; i += 2) {
                uint32_t lo = src[i];
                uint32_t hi = (i + 1 < m.index_count) ? src[i + 1] : 0u;
                ps.index_words.push_back(lo | (hi << 16));
            }
            ps.index_words.push_back(0u);  // unpackIndex may read word p + 1 of the last triangle
        }
        ps.meshes.push_back(dm);
        DeviceMaterial mat;
        static_assert(sizeof(DeviceMaterial) == sizeof(pbrt_material), "material layout");
        std::memcpy(&mat, &m.material, sizeof(mat));
        ps.materials.push_back(mat);
        ps.mesh_opaque.push_back(!tex_has_alpha_hole[dm.diffuse_tex]);
    }

    // instances, world triangles and emissive lights
    ps.instances.clear();
    ps.lights.clear();
    ps.world_tris.clear();
    ps.tri_inst.clear();
    ps.tri_prim.clear();
    for (uint32_t ii = 0; ii < d.instance_count; ++ii) {
        const pbrt_instance_desc& in = d.instances[ii];
        PBRT_REQUIRE(in.mesh < d.mesh_count, "instance " + std::to_string(ii) + " references a missing mesh");
        const pbrt_mesh_desc& m = d.meshes[in.mesh];
        DeviceInstance di{};
        std::memcpy(di.object_mat, in.transform, sizeof(di.object_mat));
        normal_matrix(in.transform, di.normal_mat);
        di.mesh_id = (int32_t)in.mesh;
        di.index_stride = m.index_stride;
        di.flags = ps.mesh_opaque[in.mesh] ? 0u : kTriFlagAlphaTested;
        ps.instances.push_back(di);
        uint32_t ntri = m.index_count / 3;
        for (uint32_t t = 0; t < ntri; ++t) {
            for (int k = 0; k < 3; ++k) {
                double w[3];
                mat4_mul_point_d(in.transform, m.positions + 3 * (size_t)read_index(m, 3 * t + k), w);
                for (int a = 0; a < 3; ++a) ps.world_tris.push_back((float)w[a]);
            }
            ps.tri_inst.push_back(ii);
            ps.tri_prim.push_back(t);
        }
        const float* em = m.material.emission_texture_id;
        if (em[0] + em[1] + em[2] != 0.0f) {
            for (uint32_t t = 0; t < ntri; ++t) {
                DeviceLight l{};
                vec3 col = make_vec3(em[0], em[1], em[2]);
                l.col_ambient = make_vec4(col, 0.0f);
                l.col_diffuse = make_vec4(col, 0.0f);
                l.col_specular = make_vec4(col, 0.0f);
                l.strengths = make_vec4(0.0f, 0.0f, 1.0f, 0.0f);
                vec4* corners[3] = {&l.v0_type, &l.v1_strength, &l.v2_angle};
                for (int k = 0; k < 3; ++k) {
                    double w[3];
                    mat4_mul_point_d(in.transform, m.positions + 3 * (size_t)read_index(m, 3 * t + k), w);
                    *corners[k] = make_vec4((float)w[0], (float)w[1], (float)w[2], 0.0f);
                }
                l.v0_type.w = 5.0f;  // LightSourceType::Area
                ps.lights.push_back(l);
            }
        }
    }
    for (uint32_t i = 0; i < d.light_count; ++i) {
        DeviceLight l;
        std::memcpy(&l, &d.lights[i], sizeof(l));
        ps.lights.push_back(l);
    }
    if (ps.lights.empty()) {
        DeviceLight l{};
        vec3 dir = normalize(make_vec3(0.1f, 1.0f, -5.1f));
        l.v0_type = make_vec4(0.0f, 0.0f, 0.0f, 1.0f);  // Directional
        l.dir_angle2 = make_vec4(dir, 0.0f);
        l.col_ambient = make_vec4(2.0f, 2.0f, 2.0f, 0.0f);
        l.col_diffuse = make_vec4(2.0f, 2.0f, 2.0f, 0.0f);
        l.col_specular = make_vec4(0.0f, 0.0f, 0.0f, 0.0f);
        l.strengths = make_vec4(0.5f, 0.0f, 0.0f, 0.0f);
        ps.lights.push_back(l);
    }
    float strength_sum = 0.0f;
    for (DeviceLight& l : ps.lights) {
        float s = l.col_ambient.x + l.col_ambient.y + l.col_ambient.z + l.col_diffuse.x + l.col_diffuse.y +
                  l.col_diffuse.z + l.col_specular.x + l.col_specular.y + l.col_specular.z;
        strength_sum += s;
        l.strengths.w = strength_sum;
    }
    ps.light_strength_sum = strength_sum;
}

}  // namespace pbrt

using namespace pbrt;

struct pbrt_scene {
    PackedScene host;
    BvhHost bvh;
    DeviceBuffer tris, nodes, instances, meshes, materials, positions, normals, uvs, index_words, textures, texels,
        lights;
    DeviceScene dev{};
    pbrt_scene_info info{};
    mutable TraceTiming timing;
    // Pixel counters of the persistent tracer: a ring of kWorkSlots slots of
    // (next pixel, finished waves), 64 B apart; each launch takes the next slot
    // and its last wave clears it.
    DeviceBuffer work;
    mutable std::atomic<uint32_t> work_next{0};
};

namespace pbrt {
const DeviceScene& scene_device_view(const pbrt_scene* s) { return s->dev; }

uint32_t scene_bvh_depth(const pbrt_scene* s) { return s->bvh.depth; }

TraceTiming* scene_timing(const pbrt_scene* s) { return s->timing.active() ? &s->timing : nullptr; }

uint32_t* scene_work_counter(const pbrt_scene* s) {
    const uint32_t slot = s->work_next.fetch_add(1, std::memory_order_relaxed) % kWorkSlots;
    return s->work.as<uint32_t>() + (size_t)slot * kWorkSlotWords;
}

TraceTiming::~TraceTiming() {
    for (hipEvent_t e : events_) (void)hipEventDestroy(e);
}

void TraceTiming::start() {
    used_ = 0;
    frames_ = 0;
    segments_.clear();
    active_ = true;
}

uint32_t TraceTiming::record(hipStream_t st) {
    if (used_ == events_.size()) {
        hipEvent_t e;
        PBRT_HIP(hipEventCreate(&e));
        events_.push_back(e);
    }
    PBRT_HIP(hipEventRecord(events_[used_], st));
    return used_++;
}

void TraceTiming::frame_begin(hipStream_t st) {
    last_ = record(st);
    ++frames_;
}

void TraceTiming::mark(int kind, hipStream_t st) {
    const uint32_t e = record(st);
    segments_.push_back({kind, last_, e});
    last_ = e;
}

void TraceTiming::finish(pbrt_trace_timing* out) {
    active_ = false;
    double ms[kTimeKinds] = {0, 0, 0, 0, 0};
    uint64_t n[kTimeKinds] = {0, 0, 0, 0, 0};
    if (used_ > 0) PBRT_HIP(hipEventSynchronize(events_[used_ - 1]));
    for (const Segment& g : segments_) {
        float t = 0.0f;
        PBRT_HIP(hipEventElapsedTime(&t, events_[g.a], events_[g.b]));
        ms[g.kind] += t;
        n[g.kind] += 1;
    }
    *out = pbrt_trace_timing{};
    out->extend_ms = ms[kTimeExtend];
    out->shade_ms = ms[kTimeShade];
    out->shadow_ms = ms[kTimeShadow];
    out->megakernel_ms = ms[kTimeMegakernel];
    out->other_ms = ms[kTimeOther];
    out->extend_launches = n[kTimeExtend];
    out->shade_launches = n[kTimeShade];
    out->shadow_launches = n[kTimeShadow];
    out->megakernel_launches = n[kTimeMegakernel];
    out->frames = frames_;
    segments_.clear();
    used_ = 0;
}
}  // namespace pbrt

extern "C" int32_t pbrt_trace_timing_begin(const pbrt_scene* scene) {
    return guard([&] {
        PBRT_REQUIRE(scene, "pbrt_trace_timing_begin: NULL scene");
        scene->timing.start();
    });
}

extern "C" int32_t pbrt_trace_timing_end(const pbrt_scene* scene, pbrt_trace_timing* out) {
    return guard([&] {
        PBRT_REQUIRE(scene && out, "pbrt_trace_timing_end: NULL argument");
        PBRT_REQUIRE(scene->timing.active(), "pbrt_trace_timing_end: timing was not started");
        scene->timing.finish(out);
    });
}

extern "C" int32_t pbrt_abi_version(void) { return PBRT_ABI_VERSION; }

extern "C" const char* pbrt_last_error(void) { return g_last_error.c_str(); }

extern "C" int32_t pbrt_scene_create(const pbrt_scene_desc* desc, pbrt_scene** out_scene) {
    return guard([&] {
        PBRT_REQUIRE(desc != nullptr && out_scene != nullptr, "pbrt_scene_create: NULL argument");
        *out_scene = nullptr;
        PBRT_REQUIRE(desc->bvh_builder == PBRT_BVH_SAH_HOST || desc->bvh_builder == PBRT_BVH_LBVH_GPU ||
                         desc->bvh_builder == PBRT_BVH_SBVH_HOST || desc->bvh_builder == PBRT_BVH_PLOC_GPU,
                     "pbrt_scene_create: unknown bvh_builder");
        PBRT_REQUIRE(desc->instance_count > 0, "pbrt_scene_create: scene has no instances");
        auto s = std::make_unique<pbrt_scene>();
        pack_scene(*desc, s->host);
        PackedScene& ps = s->host;
        uint32_t ntri = (uint32_t)ps.tri_inst.size();
        PBRT_REQUIRE(ntri > 0, "pbrt_scene_create: scene has no triangles");
        PBRT_REQUIRE(ntri < (1u << 27), "pbrt_scene_create: more than 2^27 triangles");

        std::vector<float> lo(3 * (size_t)ntri), hi(3 * (size_t)ntri);
        for (uint32_t t = 0; t < ntri; ++t)
            for (int a = 0; a < 3; ++a) {
                float v0 = ps.world_tris[9 * (size_t)t + a], v1 = ps.world_tris[9 * (size_t)t + 3 + a],
                      v2 = ps.world_tris[9 * (size_t)t + 6 + a];
                lo[3 * (size_t)t + a] = std::fmin(v0, std::fmin(v1, v2));
                hi[3 * (size_t)t + a] = std::fmax(v0, std::fmax(v1, v2));
            }
        // A traversal holds at most depth - 1 deferred children, so a tree of depth
        // <= kTraversalStackSize fits the per-thread stack.
        const auto t0 = std::chrono::steady_clock::now();
        uint32_t node_count = 0;
        if (desc->bvh_builder == PBRT_BVH_LBVH_GPU && ntri >= 2) {
            std::vector<uint32_t> flags(ntri);
            for (uint32_t t = 0; t < ntri; ++t) flags[t] = ps.instances[ps.tri_inst[t]].flags;
            LbvhResult r = build_lbvh(ps.world_tris, ps.tri_inst, ps.tri_prim, flags, kTraversalStackSize, s->tris,
                                      s->nodes);
            s->bvh.depth = r.depth;
            node_count = r.node_count;
            s->info.bvh_build_ms = r.build_ms;
            s->info.bvh_builder = PBRT_BVH_LBVH_GPU;
        } else if (desc->bvh_builder == PBRT_BVH_PLOC_GPU && ntri >= 2) {
            std::vector<uint32_t> flags(ntri);
            for (uint32_t t = 0; t < ntri; ++t) flags[t] = ps.instances[ps.tri_inst[t]].flags;
            PlocResult r = build_ploc(ps.world_tris, ps.tri_inst, ps.tri_prim, flags, kPlocMaxLeaf,
                                      kTraversalStackSize, s->tris, s->nodes);
            s->bvh.depth = r.depth;
            node_count = r.node_count;
            s->info.bvh_build_ms = r.build_ms;
            s->info.bvh_builder = PBRT_BVH_PLOC_GPU;
        } else {
            const bool spatial = desc->bvh_builder == PBRT_BVH_SBVH_HOST;
            // spatial-split leaves of at most 2 triangles (atrium trace: 4.23 ms against 4.29 with
            // 4, 4.25 with 3, 4.93 with 1)
            s->bvh = spatial ? build_bvh_sbvh(ps.world_tris.data(), ntri, 2, kTraversalStackSize)
                             : build_bvh_sah(lo.data(), hi.data(), ntri, 4, kTraversalStackSize);
            PBRT_REQUIRE(s->bvh.depth <= (uint32_t)kTraversalStackSize,
                         "pbrt_scene_create: BVH deeper than the traversal stack");
            // BVH-ordered triangle records (one per reference: spatial splits repeat a
            // triangle in several leaves)
            PBRT_REQUIRE(s->bvh.tri_order.size() < (1u << 27), "pbrt_scene_create: more than 2^27 triangle references");
            const uint32_t nref = (uint32_t)s->bvh.tri_order.size();
            std::vector<float4> tri(3 * (size_t)nref);
            for (uint32_t k = 0; k < nref; ++k) {
                uint32_t t = s->bvh.tri_order[k];
                const float* p = &ps.world_tris[9 * (size_t)t];
                uint32_t inst = ps.tri_inst[t];
                uint32_t flags = ps.instances[inst].flags;
                tri[3 * (size_t)k + 0] = make_float4(p[0], p[1], p[2], uint_as_float(inst));
                tri[3 * (size_t)k + 1] =
                    make_float4(p[3] - p[0], p[4] - p[1], p[5] - p[2], uint_as_float(ps.tri_prim[t]));
                tri[3 * (size_t)k + 2] = make_float4(p[6] - p[0], p[7] - p[1], p[8] - p[2], uint_as_float(flags));
            }
            s->tris.upload(tri.data(), tri.size() * sizeof(float4));
            s->nodes.upload(s->bvh.nodes.data(), s->bvh.nodes.size() * sizeof(float));
            node_count = (uint32_t)(s->bvh.nodes.size() / 16);
            s->info.bvh_build_ms =
                (float)std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            s->info.bvh_builder = spatial ? PBRT_BVH_SBVH_HOST : PBRT_BVH_SAH_HOST;
        }
        quantize_bvh_nodes(s->nodes, node_count);
        s->instances.upload(ps.instances.data(), ps.instances.size() * sizeof(DeviceInstance));
        s->meshes.upload(ps.meshes.data(), ps.meshes.size() * sizeof(DeviceMesh));
        s->materials.upload(ps.materials.data(), ps.materials.size() * sizeof(DeviceMaterial));
        s->positions.upload(ps.positions.data(), ps.positions.size() * sizeof(float));
        s->normals.upload(ps.normals.data(), ps.normals.size() * sizeof(float));
        s->uvs.upload(ps.uvs.data(), ps.uvs.size() * sizeof(float));
        s->index_words.upload(ps.index_words.data(), ps.index_words.size() * sizeof(uint32_t));
        s->textures.upload(ps.textures.data(), ps.textures.size() * sizeof(DeviceTexture));
        s->texels.upload(ps.texels.data(), ps.texels.size() * sizeof(uint32_t));
        s->lights.upload(ps.lights.data(), ps.lights.size() * sizeof(DeviceLight));
        {
            const std::vector<uint32_t> zeros((size_t)kWorkSlots * kWorkSlotWords, 0u);
            s->work.upload(zeros.data(), zeros.size() * sizeof(uint32_t));
        }

        DeviceScene& dv = s->dev;
        dv.tris = s->tris.as<float4>();
        dv.nodes = s->nodes.as<uint4>();
        dv.instances = s->instances.as<DeviceInstance>();
        dv.meshes = s->meshes.as<DeviceMesh>();
        dv.materials = s->materials.as<DeviceMaterial>();
        dv.positions = s->positions.as<float>();
        dv.normals = s->normals.as<float>();
        dv.uvs = s->uvs.as<float>();
        dv.index_words = s->index_words.as<uint32_t>();
        dv.textures = s->textures.as<DeviceTexture>();
        dv.texels = s->texels.as<uint32_t>();
        dv.lights = s->lights.as<DeviceLight>();
        dv.light_count = (uint32_t)ps.lights.size();
        dv.light_strength_sum = ps.light_strength_sum;
        dv.triangle_count = ntri;
        dv.node_count = node_count;
        dv.has_transmissive = 0;
        for (const DeviceMaterial& m : ps.materials)
            if ((int)m.transmittance_illum.w == 7) dv.has_transmissive = 1;

        pbrt_scene_info& info = s->info;
        info.triangle_count = ntri;
        info.instance_count = desc->instance_count;
        info.mesh_count = desc->mesh_count;
        info.light_count = dv.light_count;
        info.light_strength_sum = dv.light_strength_sum;
        info.bvh_node_count = dv.node_count;
        info.bvh_depth = s->bvh.depth;
        info.light_sampling = dv.light_count > 800 ? PBRT_LIGHT_SAMPLE_UNIFORM : PBRT_LIGHT_SAMPLE_SURFACE_STRENGTH;
        info.device_bytes = s->tris.bytes + s->nodes.bytes + s->instances.bytes + s->meshes.bytes +
                            s->materials.bytes + s->positions.bytes + s->normals.bytes + s->uvs.bytes +
                            s->index_words.bytes + s->textures.bytes + s->texels.bytes + s->lights.bytes;
        // the host-side world triangle copy is only needed for the BVH build
        std::vector<float>().swap(ps.world_tris);
        *out_scene = s.release();
    });
}

extern "C" int32_t pbrt_scene_destroy(pbrt_scene* scene) {
    return guard([&] { delete scene; });
}

extern "C" int32_t pbrt_scene_get_info(const pbrt_scene* scene, pbrt_scene_info* out_info) {
    return guard([&] {
        PBRT_REQUIRE(scene && out_info, "pbrt_scene_get_info: NULL argument");
        *out_info = scene->info;
    });
}

extern "C" int32_t pbrt_scene_get_lights(const pbrt_scene* scene, pbrt_light* out_lights, uint32_t capacity) {
    return guard([&] {
        PBRT_REQUIRE(scene && out_lights, "pbrt_scene_get_lights: NULL argument");
        PBRT_REQUIRE(capacity >= scene->host.lights.size(), "pbrt_scene_get_lights: capacity too small");
        std::memcpy(out_lights, scene->host.lights.data(), scene->host.lights.size() * sizeof(pbrt_light));
    });
}
