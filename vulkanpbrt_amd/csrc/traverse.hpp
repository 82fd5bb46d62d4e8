// BVH traversal and hit shading on gfx950.
//
// traverse<>() is the software replacement of the driver traversal behind
// traceRayEXT: stack-based BVH2 walk (near child first, far child pushed to a
// per-thread LDS stack), Moller-Trumbore triangle tests, and the alpha test of
// ptAlphaHit.rahit:14-52 applied to every candidate on alpha-tested instances
// (the project rule for the undefined Opaque|NoOpaque ray flags,
// ptRaygen.rgen:16).  Closest-hit mode keeps the nearest t in [tmin, tmax];
// any-hit mode (shadow rays: TerminateOnFirstHit | SkipClosestHit) returns on
// the first accepted candidate.
#pragma once

#include <climits>

#include "pt_device.hpp"
#include "scene_types.hpp"

namespace pbrt {

struct Hit {
    float t;
    uint32_t tri;  // BVH-ordered triangle index, kNoHit on miss
    float u, v;    // hitAttributeEXT barycentrics
};

struct TraversalCounters {
    uint32_t nodes;
    uint32_t tris;
};

// ---------------------------------------------------------------------------
// Texture fetch: RGBA8 UNORM, LINEAR filter, REPEAT addressing, mip 0.
// 1x1 textures return their texel exactly (all four taps coincide).
// ---------------------------------------------------------------------------
// b / 255 correctly rounded for a byte b: q = b * fl(1/255) corrected once by the
// exact residual b - 255 q (one fma) equals the IEEE quotient for all 256 bytes
// (checked exhaustively), for three operations instead of a full division.
__device__ __forceinline__ float unorm8_value(uint32_t b) {
    const float c = (float)b, r = 1.0f / 255.0f;
    const float q = c * r;
    return fmaf(fmaf(-q, 255.0f, c), r, q);
}

__device__ __forceinline__ vec4 unpack_rgba8(uint32_t c) {
    return make_vec4(unorm8_value(c & 0xFFu), unorm8_value((c >> 8) & 0xFFu), unorm8_value((c >> 16) & 0xFFu),
                     unorm8_value(c >> 24));
}

__device__ __forceinline__ int wrap_repeat(int i, int n) {
    int r = i % n;
    return r < 0 ? r + n : r;
}

__device__ __forceinline__ vec4 lerp4(vec4 a, vec4 b, float t) {
    float s = 1.0f - t;
    return make_vec4(a.x * s + b.x * t, a.y * s + b.y * t, a.z * s + b.z * t, a.w * s + b.w * t);
}

__device__ __forceinline__ vec4 texture_sample(const DeviceScene& sc, int tex, vec2 uv) {
    const DeviceTexture t = sc.textures[tex];
    const uint32_t* texels = sc.texels + t.offset;
    if (t.width == 1 && t.height == 1) return unpack_rgba8(texels[0]);
    float x = uv.x * (float)t.width - 0.5f;
    float y = uv.y * (float)t.height - 0.5f;
    float fx0 = floorf(x), fy0 = floorf(y);
    float fx = x - fx0, fy = y - fy0;
    // the second tap wraps from the first: (i + 1) mod n = i mod n + 1, or 0 at n
    const int x0 = wrap_repeat((int)fx0, (int)t.width), y0 = wrap_repeat((int)fy0, (int)t.height);
    const int x1 = x0 + 1 == (int)t.width ? 0 : x0 + 1, y1 = y0 + 1 == (int)t.height ? 0 : y0 + 1;
    vec4 c00 = unpack_rgba8(texels[(size_t)y0 * t.width + x0]);
    vec4 c10 = unpack_rgba8(texels[(size_t)y0 * t.width + x1]);
    vec4 c01 = unpack_rgba8(texels[(size_t)y1 * t.width + x0]);
    vec4 c11 = unpack_rgba8(texels[(size_t)y1 * t.width + x1]);
    return lerp4(lerp4(c00, c10, fx), lerp4(c01, c11, fx), fy);
}

// ---------------------------------------------------------------------------
// Index / vertex unpacking (geometry.glsl:4-40)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void unpack_index(const DeviceScene& sc, const DeviceMesh& m, uint32_t prim,
                                             uint32_t& ix, uint32_t& iy, uint32_t& iz) {
    const uint32_t* w = sc.index_words + m.index_word_offset;
    if (m.index_stride == 4) {
        ix = w[3 * prim];
        iy = w[3 * prim + 1];
        iz = w[3 * prim + 2];
    } else {
        uint32_t full = 3 * prim;
        uint32_t p = full >> 1;
        if (full & 1u) {
            ix = w[p] >> 16;
            uint32_t t = w[p + 1];
            iz = t >> 16;
            iy = t & 0xFFFFu;
        } else {
            uint32_t t = w[p];
            iy = t >> 16;
            ix = t & 0xFFFFu;
            iz = w[p + 1] & 0xFFFFu;
        }
    }
}

__device__ __forceinline__ vec3 load3(const float* p, uint32_t i) { return make_vec3(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
__device__ __forceinline__ vec2 load2(const float* p, uint32_t i) { return make_vec2(p[2 * i], p[2 * i + 1]); }

// Alpha test of ptAlphaHit.rahit:14-52 (diffuse alpha, no sRGB conversion).
__device__ __forceinline__ bool alpha_rejects(const DeviceScene& sc, uint32_t inst, uint32_t prim, float u, float v) {
    const DeviceInstance& in = sc.instances[inst];
    const DeviceMesh m = sc.meshes[in.mesh_id];
    uint32_t ix, iy, iz;
    unpack_index(sc, m, prim, ix, iy, iz);
    const float* uvs = sc.uvs + 2 * (size_t)m.vertex_offset;
    vec2 uv0 = load2(uvs, ix), uv1 = load2(uvs, iy), uv2 = load2(uvs, iz);
    float b0 = 1.0f - u - v;
    vec2 tc = uv0 * b0 + uv1 * u + uv2 * v;
    return texture_sample(sc, m.diffuse_tex, tc).w < 0.01f;
}

// ---------------------------------------------------------------------------
// Traversal
// ---------------------------------------------------------------------------
// two fp16 values in one register (a quantized node's planes, see make_traversal_ray)
typedef _Float16 PlanePair __attribute__((ext_vector_type(2)));

__device__ __forceinline__ float safe_rcp(float d) {
    return 1.0f / (fabsf(d) > 1e-20f ? d : copysignf(1e-20f, d));
}

// A ray prepared for slab tests: origin, direction, reciprocal direction,
// origin * reciprocal, and the byte selectors (v_perm_b32) that pick each child's
// near and far planes out of a quantized node for this ray's direction octant.
struct TraversalRay {
    vec3 o, d;
    float ix, iy, iz, ox, oy, oz, tmin;
    uint32_t sel_near_xy, sel_far_xy, sel_z0, sel_z1;
};

// A quantized node's plane bytes (second half, words x, y, z) are grouped by axis:
// x word = child 0 (lo x, hi x, lo y, hi y), z word = the same for child 1, y word
// = child 0 (lo z, hi z), child 1 (lo z, hi z).  A selector builds two fp16 values
// from two of those bytes, each byte as the significand of an fp16 subnormal with
// selector byte 0x0c (zero) above it, i.e. q 2^-24 exactly.  sel_near_xy picks the
// near x and y planes (x into the low half), sel_far_xy the far ones; sel_z0 child
// 0's near and far z planes, sel_z1 child 1's.
__device__ __forceinline__ TraversalRay make_traversal_ray(vec3 o, vec3 d, float tmin) {
    TraversalRay r;
    r.o = o;
    r.d = d;
    r.ix = safe_rcp(d.x);
    r.iy = safe_rcp(d.y);
    r.iz = safe_rcp(d.z);
    r.ox = o.x * r.ix;
    r.oy = o.y * r.iy;
    r.oz = o.z * r.iz;
    r.tmin = tmin;
    // safe_rcp never returns 0 or NaN, so the sign decides which plane is near:
    // the lower bound with a positive reciprocal, the upper with a negative one
    const uint32_t nx = r.ix < 0.0f, ny = r.iy < 0.0f, nz = r.iz < 0.0f;
    r.sel_near_xy = nx | (2u + ny) << 16 | 0x0c000c00u;
    r.sel_far_xy = r.sel_near_xy ^ 0x00010001u;
    r.sel_z0 = nz | (1u - nz) << 16 | 0x0c000c00u;
    r.sel_z1 = r.sel_z0 + 0x00020002u;
    return r;
}

// Lane holds no node to visit: its stack is empty (or it has no ray).  Leaf
// codes are ~(first << kLeafCountBits | count - 1) with first + count < 2^27,
// so INT_MIN is never a leaf.
constexpr int kNodeDone = INT_MIN;

__device__ __forceinline__ int pop_node(const int* stack, int stride, int& sp) {
    if (sp == 0) return kNodeDone;
    --sp;
    return stack[sp * stride];
}

// One inner node (quantized layout, 2 x uint4): slab-test both child boxes
// against [tmin, tcur]; returns the nearer hit child (the farther one pushed) or
// the next node off the stack.  A plane at origin + q 2^e along an axis is met at
// t = (origin + q 2^e - o) / d, evaluated as one v_fma_mix_f32 of the fp16
// subnormal q 2^-24, S = 2^(e + 24) / d (exact: a power-of-two scaling) and
// C = fma(origin, 1/d, -o/d) (once per axis and node); the product is exact, so t
// is (origin - o) / d + q 2^e / d rounded once.  The origin word is used whole (its
// low byte, e + 24, is part of the origin the quantization pass checked the boxes
// against).
//   kPlanes 1 and 2: only the near and far plane of each axis is formed, picked by
//   the ray's octant (fma rounding being monotonic, this equals min / max over both
//   planes).  1 keeps the four selectors in registers; 2 keeps sel_near_xy and
//   sel_z0 and derives the other two per node (2 operations), for the persistent
//   kernel, whose occupancy two more registers would lower.
//   kPlanes 0: all twelve planes formed and ordered by min / max.
//   kOrdered: closest-hit walks take the nearer child first; any-hit (shadow) walks
//   the farther one, since an occluder tends to lie toward the light rather than
//   next to the surface the ray leaves (atrium: 25.2 node visits per shadow ray
//   against 26.4 nearer-first and 25.9 in stored order).
template <int kPlanes, bool kOrdered = true>
__device__ __forceinline__ int visit_inner(const uint4* __restrict__ nodes, int node, const TraversalRay& r,
                                           float tcur, int* stack, int stride, int& sp) {
    // a 32-bit byte offset from the array's base (node indices stay below 2^27), so
    // the loads take the scalar-base addressing form
    const char* base = reinterpret_cast<const char*>(nodes);
    const uint32_t off = (uint32_t)node << 5;
    const uint4 na = *reinterpret_cast<const uint4*>(base + off);
    const uint4 nb = *reinterpret_cast<const uint4*>(base + off + 16u);
    // the origin word's low byte, signed, is e + 24
    const float sx = __builtin_amdgcn_ldexpf(r.ix, (int)(int8_t)na.x),
                sy = __builtin_amdgcn_ldexpf(r.iy, (int)(int8_t)na.y),
                sz = __builtin_amdgcn_ldexpf(r.iz, (int)(int8_t)na.z);
    const float cx = fmaf(__uint_as_float(na.x), r.ix, -r.ox), cy = fmaf(__uint_as_float(na.y), r.iy, -r.oy),
                cz = fmaf(__uint_as_float(na.z), r.iz, -r.oz);
    // two planes as fp16 subnormals; the widening folds into v_fma_mix_f32
    auto pair = [](uint32_t w, uint32_t sel) { return __builtin_bit_cast(PlanePair, __builtin_amdgcn_perm(w, w, sel)); };
    auto tx = [&](_Float16 h) { return fmaf((float)h, sx, cx); };
    auto ty = [&](_Float16 h) { return fmaf((float)h, sy, cy); };
    auto tz = [&](_Float16 h) { return fmaf((float)h, sz, cz); };
    float near0, far0, near1, far1;
    if constexpr (kPlanes != 0) {
        uint32_t sn = r.sel_near_xy, sf = r.sel_far_xy, sz0 = r.sel_z0, sz1 = r.sel_z1;
        if constexpr (kPlanes == 2) {
            // an opaque copy, so the derived selectors are not hoisted out of the walk
            asm volatile("" : "+v"(sn), "+v"(sz0));
            sf = sn ^ 0x00010001u;
            sz1 = sz0 + 0x00020002u;
        }
        const PlanePair n0 = pair(nb.x, sn), f0 = pair(nb.x, sf), z0 = pair(nb.y, sz0);
        const PlanePair n1 = pair(nb.z, sn), f1 = pair(nb.z, sf), z1 = pair(nb.y, sz1);
        near0 = fmaxf(fmaxf(tx(n0.x), ty(n0.y)), fmaxf(tz(z0.x), r.tmin));
        far0 = fminf(fminf(tx(f0.x), ty(f0.y)), fminf(tz(z0.y), tcur)) * 1.00000024f;
        near1 = fmaxf(fmaxf(tx(n1.x), ty(n1.y)), fmaxf(tz(z1.x), r.tmin));
        far1 = fminf(fminf(tx(f1.x), ty(f1.y)), fminf(tz(z1.y), tcur)) * 1.00000024f;
    } else {
        // (lo, hi) pairs: x and y of each child, z of both
        const PlanePair x0 = pair(nb.x, 0x0c010c00u), y0 = pair(nb.x, 0x0c030c02u);
        const PlanePair x1 = pair(nb.z, 0x0c010c00u), y1 = pair(nb.z, 0x0c030c02u);
        const PlanePair z0 = pair(nb.y, 0x0c010c00u), z1 = pair(nb.y, 0x0c030c02u);
        const float a0 = tx(x0.x), a1 = tx(x0.y), b0 = ty(y0.x), b1 = ty(y0.y), c0 = tz(z0.x), c1 = tz(z0.y);
        near0 = fmaxf(fmaxf(fminf(a0, a1), fminf(b0, b1)), fmaxf(fminf(c0, c1), r.tmin));
        far0 = fminf(fminf(fmaxf(a0, a1), fmaxf(b0, b1)), fminf(fmaxf(c0, c1), tcur)) * 1.00000024f;
        const float d0 = tx(x1.x), d1 = tx(x1.y), e0 = ty(y1.x), e1 = ty(y1.y), f0 = tz(z1.x), f1 = tz(z1.y);
        near1 = fmaxf(fmaxf(fminf(d0, d1), fminf(e0, e1)), fmaxf(fminf(f0, f1), r.tmin));
        far1 = fminf(fminf(fmaxf(d0, d1), fmaxf(e0, e1)), fminf(fmaxf(f0, f1), tcur)) * 1.00000024f;
    }
    bool h0 = near0 <= far0;
    bool h1 = near1 <= far1;
    int ch0 = (int)na.w;
    int ch1 = (int)nb.w;
    if (h0 && h1) {
        bool swap = kOrdered ? near1 < near0 : near1 > near0;
        stack[sp * stride] = swap ? ch0 : ch1;
        ++sp;
        return swap ? ch1 : ch0;
    }
    if (h0) return ch0;
    if (h1) return ch1;
    return pop_node(stack, stride, sp);
}

// A candidate at the same t as the current closest hit replaces it when its
// (instance, primitive) is lower; nothing replaces "no hit" at exactly tmax.
__device__ __forceinline__ bool wins_tie(const float4* __restrict__ tris, uint32_t best, uint32_t inst,
                                         uint32_t prim) {
    if (best == kNoHit) return false;
    const uint32_t bi = __float_as_uint(tris[3 * (size_t)best].w), bp = __float_as_uint(tris[3 * (size_t)best + 1].w);
    return inst < bi || (inst == bi && prim < bp);
}

// The triangles of one leaf (Moller-Trumbore, alpha test on alpha-tested
// instances); closest hit keeps the nearest accepted t, any-hit returns true on
// the first accepted candidate.
template <bool kAnyHit, bool kCount>
__device__ __forceinline__ bool test_leaf(const DeviceScene& sc, int node, const TraversalRay& r, Hit& hit,
                                          TraversalCounters& cnt) {
    const float4* __restrict__ tris = sc.tris;
    const vec3 o = r.o, d = r.d;
    uint32_t code = ~(uint32_t)node;
    uint32_t first = code >> kLeafCountBits;
    uint32_t count = (code & (kMaxLeafSize - 1)) + 1;
    for (uint32_t k = first; k < first + count; ++k) {
        if (kCount) cnt.tris++;
        const float4 t0 = tris[3 * k + 0];
        const float4 t1 = tris[3 * k + 1];
        const float4 t2 = tris[3 * k + 2];
        vec3 v0 = make_vec3(t0.x, t0.y, t0.z);
        vec3 e1 = make_vec3(t1.x, t1.y, t1.z);
        vec3 e2 = make_vec3(t2.x, t2.y, t2.z);
        vec3 pv = cross(d, e2);
        float det = dot(e1, pv);
        if (det == 0.0f) continue;
        float inv = 1.0f / det;
        vec3 tv = o - v0;
        float u = dot(tv, pv) * inv;
        if (!(u >= 0.0f && u <= 1.0f)) continue;
        vec3 qv = cross(tv, e1);
        float v = dot(d, qv) * inv;
        if (!(v >= 0.0f && u + v <= 1.0f)) continue;
        float t = dot(e2, qv) * inv;
        if constexpr (kAnyHit) {
            if (!(t >= r.tmin && t < hit.t)) continue;
        } else {
            // at exactly the current t the lower (instance, primitive) wins, so the
            // closest hit does not depend on the order leaves are visited in (BVH
            // builder, traversal order, triangles repeated by spatial splits)
            if (!(t >= r.tmin && t <= hit.t)) continue;
            if (t == hit.t && !wins_tie(tris, hit.tri, __float_as_uint(t0.w), __float_as_uint(t1.w))) continue;
        }
        uint32_t flags = __float_as_uint(t2.w);
        if (flags & kTriFlagAlphaTested) {
            if (alpha_rejects(sc, __float_as_uint(t0.w), __float_as_uint(t1.w), u, v)) continue;
        }
        hit.t = t;
        hit.tri = k;
        hit.u = u;
        hit.v = v;
        if (kAnyHit) return true;
    }
    return false;
}

// Per-ray walk, organised for wave64 (speculative while-while): a lane walks
// inner nodes and sets the first leaf it meets aside, continuing down the tree
// until every lane of the wave holds a leaf (or has no inner node left); then
// the lanes test one leaf each.  Leaf tests thus run for many lanes at once
// instead of interleaving with node visits.  The triangles tested are those of
// the plain nearest-child-first walk; only the order of leaves can differ, which
// changes nothing but the choice among hits at exactly equal t.
template <bool kAnyHit, bool kCount, int kPlanes = 1>
__device__ bool traverse(const DeviceScene& sc, vec3 o, vec3 d, float tmin, float tmax, Hit& hit, int* stack,
                         int stride, TraversalCounters& cnt) {
    const TraversalRay r = make_traversal_ray(o, d, tmin);
    hit.tri = kNoHit;
    hit.t = tmax;
    int sp = 0;
    int node = 0;
    int leaf = 0;  // leaf set aside (0: none; leaf codes are negative)
    while (node != kNodeDone || leaf != 0) {
        while (node >= 0) {
            if (kCount) cnt.nodes++;
            node = visit_inner<kPlanes, !kAnyHit>(sc.nodes, node, r, hit.t, stack, stride, sp);
            if (node < 0 && node != kNodeDone && leaf == 0) {
                leaf = node;
                node = pop_node(stack, stride, sp);
            }
            // leave only when no lane still walking lacks a set-aside leaf: every lane
            // then holds a leaf, a leaf code in `node` or kNodeDone, never an inner
            // node, which the leaf test below would read as a triangle range
            if (!__any(leaf == 0)) break;
        }
        const int lf = leaf != 0 ? leaf : (node != kNodeDone ? node : 0);
        if (lf != 0) {
            if (test_leaf<kAnyHit, kCount>(sc, lf, r, hit, cnt)) return true;
            if (leaf != 0)
                leaf = 0;
            else
                node = pop_node(stack, stride, sp);
        }
    }
    return hit.tri != kNoHit;
}

// ---------------------------------------------------------------------------
// Closest-hit shading: ptClosesthit.rchit:18-92 with geometry.glsl helpers.
// ---------------------------------------------------------------------------
__device__ __forceinline__ vec3 get_tangent(vec3 a, vec3 b, vec3 c, vec2 auv, vec2 buv, vec2 cuv) {
    float bv_cv = buv.y - cuv.y;
    if (bv_cv == 0.0f) return (b - c) / (buv.x - cuv.x);
    float quotient = (auv.y - cuv.y) / bv_cv;
    vec3 dd = c + (b - c) * quotient;
    vec2 duv = cuv + (buv - cuv) * quotient;
    return (dd - a) / (duv.x - auv.x);
}

__device__ __forceinline__ vec3 project_onto(vec3 a, vec3 b) { return dot(a, b) / dot(b, b) * b; }

__device__ __forceinline__ vec3 mat3_mul(const float* m, vec3 v) {
    return make_vec3(m[0] * v.x + m[3] * v.y + m[6] * v.z, m[1] * v.x + m[4] * v.y + m[7] * v.z,
                     m[2] * v.x + m[5] * v.y + m[8] * v.z);
}

__device__ void shade_hit(const DeviceScene& sc, const Hit& hit, vec3 ray_dir, SurfaceInfo& s, vec3& position,
                          uint32_t& category) {
    const float4 t0 = sc.tris[3 * hit.tri + 0];
    const float4 t1 = sc.tris[3 * hit.tri + 1];
    uint32_t inst_id = __float_as_uint(t0.w);
    uint32_t prim = __float_as_uint(t1.w);
    const DeviceInstance& inst = sc.instances[inst_id];
    const uint32_t obj = (uint32_t)inst.mesh_id;
    const DeviceMesh m = sc.meshes[obj];
    uint32_t ix, iy, iz;
    unpack_index(sc, m, prim, ix, iy, iz);
    const float* P = sc.positions + 3 * (size_t)m.vertex_offset;
    const float* N = sc.normals + 3 * (size_t)m.vertex_offset;
    const float* UV = sc.uvs + 2 * (size_t)m.vertex_offset;
    vec3 p0 = load3(P, ix), p1 = load3(P, iy), p2 = load3(P, iz);
    vec3 n0 = load3(N, ix), n1 = load3(N, iy), n2 = load3(N, iz);
    vec2 uv0 = load2(UV, ix), uv1 = load2(UV, iy), uv2 = load2(UV, iz);
    const float bx = 1.0f - hit.u - hit.v, by = hit.u, bz = hit.v;
    vec2 tc = uv0 * bx + uv1 * by + uv2 * bz;
    vec4 diffuse = texture_sample(sc, m.diffuse_tex, tc);
    vec3 drgb = srgb_to_linear(make_vec3(diffuse.x, diffuse.y, diffuse.z));
    drgb *= diffuse.w;
    vec3 pos = p0 * bx + p1 * by + p2 * bz;
    position = xyz(mat4_mul(inst.object_mat, make_vec4(pos, 1.0f)));
    vec3 normal = normalize(n0 * bx + n1 * by + n2 * bz);
    if (isinf(normal.x) || isnan(normal.x)) normal = make_vec3(0.0f, 1.0f, 0.0f);
    normal = normalize(mat3_mul(inst.normal_mat, normal));
    const float eps = 1e-6f;
    if (uv0 == uv1) uv1 = uv1 + make_vec2(eps, 0.0f);
    if (uv0 == uv2) uv2 = uv2 + make_vec2(0.0f, eps);
    if (uv1 == uv2) uv2 = uv2 + make_vec2(eps, eps);
    vec3 T = mat3_mul(inst.normal_mat, get_tangent(p0, p1, p2, uv0, uv1, uv2));
    vec3 B = mat3_mul(inst.normal_mat, get_tangent(p0, p2, p1, make_vec2(uv0.y, uv0.x), make_vec2(uv2.y, uv2.x),
                                                   make_vec2(uv1.y, uv1.x)));
    // gramSchmidt (geometry.glsl:78-82)
    vec3 tt = T - project_onto(T, normal);
    vec3 bb = B - project_onto(B, tt) - project_onto(B, normal);
    mat3 tbn;
    tbn.c[0] = normalize(tt);
    tbn.c[1] = normalize(bb);
    tbn.c[2] = normal;
    // getNormal (geometry.glsl:84-92)
    const DeviceTexture ntex = sc.textures[m.normal_tex];
    if (!(ntex.width == 1 && ntex.height == 1)) {
        vec4 tn = texture_sample(sc, m.normal_tex, tc);
        vec3 tangent_normal = make_vec3(tn.x * 2.0f - 1.0f, tn.y * 2.0f - 1.0f, tn.z * 2.0f - 1.0f);
        normal = normalize(tbn * tangent_normal);
    } else {
        normal = tbn.c[2];
    }
    const DeviceMaterial& mat = sc.materials[obj];
    drgb = drgb * xyz(mat.diffuse_ior);
    vec4 specular;
    const DeviceTexture stex = sc.textures[m.specular_tex];
    if (stex.width == 1 && stex.height == 1) {
        specular = make_vec4(xyz(mat.specular_dissolve), mat.ambient_roughness.w);
    } else {
        vec4 st = texture_sample(sc, m.specular_tex, tc);
        vec3 sl = srgb_to_linear(make_vec3(st.x, st.y, st.z));
        specular = make_vec4(sl, st.w);
    }
    float perceptual_roughness = specular.w;
    vec3 srgb = make_vec3(specular.x, specular.y, specular.z);
    float max_specular = max3(srgb);
    float metallic = convert_metallic(drgb, srgb, max_specular);
    vec3 diff_part = drgb * ((1.0f - max_specular) / (1.0f - kMinRoughness) / fmaxf(1.0f - metallic, eps));
    vec3 spec_part = srgb - (make_vec3(kMinRoughness) * (1.0f - metallic) * (1.0f / fmaxf(metallic, eps)));
    vec3 base = vmix(diff_part, spec_part, metallic * metallic);
    const vec3 f0 = make_vec3(0.04f);
    vec3 diffuse_color = base * (make_vec3(1.0f) - f0);
    diffuse_color *= 1.0f - metallic;
    float alpha_roughness = perceptual_roughness * perceptual_roughness;
    vec3 specular_color = vmix(f0, base, metallic);
    float reflectance = max3(specular_color);
    float reflectance90 = gclamp(reflectance * 25.0f, 0.0f, 1.0f);
    vec3 v = normalize(-ray_dir);
    vec4 em = texture_sample(sc, m.emissive_tex, tc);
    vec3 emissive = xyz(mat.emission_texture_id) * srgb_to_linear(make_vec3(em.x, em.y, em.z));
    if (dot(v, normal) < 0.0f) emissive = make_vec3(0.0f);
    s.perceptual_roughness = perceptual_roughness;
    s.metalness = metallic;
    s.alpha_roughness = alpha_roughness;
    s.illumination_type = (int)mat.transmittance_illum.w;
    s.reflectance0 = specular_color;
    s.reflectance90 = make_vec3(reflectance90);
    s.diffuse_color = diffuse_color;
    s.specular_color = specular_color;
    s.emissive_color = emissive;
    s.transmissive_color = xyz(mat.transmittance_illum);
    s.normal = normal;
    s.basis = tbn;
    s.index_of_refraction = mat.diffuse_ior.w;
    category = mat.category_id;
}

}  // namespace pbrt
