// Image texel formats and sampler emulation shared by the screen-space passes.
//
// Storage conventions (each image is stored in the byte order of the VkFormat
// the reference gives it):
//   R32F / RG32F / RGBA32F  float, float2, float4
//   RG16F / RGBA16F          IEEE binary16 bits (uint16), round-to-nearest-even stores
//   R8 / RGBA8 UNORM         bytes, stores round(clamp(x, 0, 1) * 255) to nearest even
//   B8G8R8A8 UNORM           bytes in B, G, R, A order (logical r, g, b, a)
// texture() reads emulate the default vsg::Sampler (LINEAR, REPEAT, mip 0,
// external/vsg/include/vsg/state/Sampler.h:29-43) with FP32 weights.
#pragma once

#include <hip/hip_fp16.h>

#include "pbrt_math.hpp"

namespace pbrt {

__device__ __forceinline__ float h2f(uint16_t h) { return __half2float(__ushort_as_half(h)); }
__device__ __forceinline__ uint16_t f2h(float f) { return __half_as_ushort(__float2half_rn(f)); }

__device__ __forceinline__ uint8_t unorm8_rne(float x) {
    x = fminf(fmaxf(x, 0.0f), 1.0f);
    return (uint8_t)rintf(x * 255.0f);
}
// c / kD rounded to nearest for a small integer c (0 <= c <= 255, 1 <= kD <= 255):
// q = c * r with r = 1/kD rounded to nearest, then one correction by the residual
// c - q*kD, which the inner FMA computes exactly.  With r within half an ulp of 1/kD
// and q within one ulp of c/kD the corrected quotient is the correctly rounded one
// (Markstein's theorem); every pair (c, kD) this code divides is also checked against
// IEEE division on the host.  Three instructions against a full division's ten.
template <int kD>
__device__ __forceinline__ float div_small(float c) {
    constexpr float r = 1.0f / (float)kD;
    const float q = c * r;
    return fmaf(fmaf(-q, (float)kD, c), r, q);
}

__device__ __forceinline__ float unorm8_to_float(uint32_t c) { return div_small<255>((float)c); }

// mirror() of bmfrGeneral.comp and bfr.comp: reflect an index that lies at most
// one image size outside [0, s)
__device__ __forceinline__ int mirror_coord(int x, int s) {
    if (x < 0) return -x - 1;
    if (x >= s) return 2 * s - x - 1;
    return x;
}

__device__ __forceinline__ int wrap_coord(int i, int n) {
    int r = i % n;
    return r < 0 ? r + n : r;
}

// Bilinear fetch with REPEAT addressing at normalized (u, v); fetch(x, y) returns vec4.
template <class F>
__device__ __forceinline__ vec4 sample_bilinear(F fetch, int w, int h, float u, float v) {
    float x = u * (float)w - 0.5f;
    float y = v * (float)h - 0.5f;
    float x0f = floorf(x), y0f = floorf(y);
    float fx = x - x0f, fy = y - y0f;
    int x0 = wrap_coord((int)x0f, w), x1 = wrap_coord((int)x0f + 1, w);
    int y0 = wrap_coord((int)y0f, h), y1 = wrap_coord((int)y0f + 1, h);
    vec4 a = fetch(x0, y0), b = fetch(x1, y0), c = fetch(x0, y1), d = fetch(x1, y1);
    float gx = 1.0f - fx, gy = 1.0f - fy;
    vec4 top = make_vec4(a.x * gx + b.x * fx, a.y * gx + b.y * fx, a.z * gx + b.z * fx, a.w * gx + b.w * fx);
    vec4 bot = make_vec4(c.x * gx + d.x * fx, c.y * gx + d.y * fx, c.z * gx + d.z * fx, c.w * gx + d.w * fx);
    return make_vec4(top.x * gy + bot.x * fy, top.y * gy + bot.y * fy, top.z * gy + bot.z * fy,
                     top.w * gy + bot.w * fy);
}

__device__ __forceinline__ vec4 load_half4(const uint16_t* p, size_t i) {
    const ushort4 v = reinterpret_cast<const ushort4*>(p)[i];
    return make_vec4(h2f(v.x), h2f(v.y), h2f(v.z), h2f(v.w));
}
__device__ __forceinline__ void store_half4(uint16_t* p, size_t i, vec4 v) {
    ushort4 o;
    o.x = f2h(v.x);
    o.y = f2h(v.y);
    o.z = f2h(v.z);
    o.w = f2h(v.w);
    reinterpret_cast<ushort4*>(p)[i] = o;
}
__device__ __forceinline__ vec4 load_rgba8(const uint8_t* p, size_t i) {
    const uchar4 c = reinterpret_cast<const uchar4*>(p)[i];
    return make_vec4(unorm8_to_float(c.x), unorm8_to_float(c.y), unorm8_to_float(c.z), unorm8_to_float(c.w));
}
__device__ __forceinline__ vec4 load_bgra8(const uint8_t* p, size_t i) {
    const uchar4 c = reinterpret_cast<const uchar4*>(p)[i];
    return make_vec4(unorm8_to_float(c.z), unorm8_to_float(c.y), unorm8_to_float(c.x), unorm8_to_float(c.w));
}
__device__ __forceinline__ void store_bgra8(uint8_t* p, size_t i, vec4 v) {
    uchar4 c;
    c.x = unorm8_rne(v.z);
    c.y = unorm8_rne(v.y);
    c.z = unorm8_rne(v.x);
    c.w = unorm8_rne(v.w);
    reinterpret_cast<uchar4*>(p)[i] = c;
}

// Spherical normal encoding of the G-buffer (ptRaygen.rgen:54-56) decoded as in
// accumulator.comp:40 / bmfrPre.comp:255-257.
__device__ __forceinline__ vec3 decode_normal(float theta, float phi) {
    float st, ct, sp, cp;
    sincosf(theta, &st, &ct);  // one range reduction for both
    sincosf(phi, &sp, &cp);
    return make_vec3(cp * st, sp * st, ct);
}

}  // namespace pbrt
