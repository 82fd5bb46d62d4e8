// Small GLSL-flavoured vector library shared by host and gfx950 device code.
// Semantics follow GLSL 4.60 where the reference shaders rely on them:
// mix(a, b, t) = a * (1 - t) + b * t, clamp(x, lo, hi) = min(max(x, lo), hi)
// with NaN-suppressing min/max, normalize(v) = v / length(v).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define PBRT_HD __host__ __device__ __forceinline__

namespace pbrt {

struct vec2 {
    float x, y;
};
struct vec3 {
    float x, y, z;
};
struct vec4 {
    float x, y, z, w;
};

PBRT_HD vec2 make_vec2(float x, float y) { return vec2{x, y}; }
PBRT_HD vec3 make_vec3(float x, float y, float z) { return vec3{x, y, z}; }
PBRT_HD vec3 make_vec3(float s) { return vec3{s, s, s}; }
PBRT_HD vec4 make_vec4(float x, float y, float z, float w) { return vec4{x, y, z, w}; }
PBRT_HD vec4 make_vec4(vec3 v, float w) { return vec4{v.x, v.y, v.z, w}; }
PBRT_HD vec3 xyz(vec4 v) { return vec3{v.x, v.y, v.z}; }

PBRT_HD vec2 operator+(vec2 a, vec2 b) { return {a.x + b.x, a.y + b.y}; }
PBRT_HD vec2 operator-(vec2 a, vec2 b) { return {a.x - b.x, a.y - b.y}; }
PBRT_HD vec2 operator*(vec2 a, float s) { return {a.x * s, a.y * s}; }
PBRT_HD vec2 operator*(float s, vec2 a) { return {a.x * s, a.y * s}; }
PBRT_HD bool operator==(vec2 a, vec2 b) { return a.x == b.x && a.y == b.y; }

PBRT_HD vec3 operator+(vec3 a, vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
PBRT_HD vec3 operator-(vec3 a, vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
PBRT_HD vec3 operator-(vec3 a) { return {-a.x, -a.y, -a.z}; }
PBRT_HD vec3 operator*(vec3 a, vec3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
PBRT_HD vec3 operator/(vec3 a, vec3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
PBRT_HD vec3 operator*(vec3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
PBRT_HD vec3 operator*(float s, vec3 a) { return {a.x * s, a.y * s, a.z * s}; }
PBRT_HD vec3 operator/(vec3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
PBRT_HD vec3 operator+(vec3 a, float s) { return {a.x + s, a.y + s, a.z + s}; }
PBRT_HD vec3 operator-(vec3 a, float s) { return {a.x - s, a.y - s, a.z - s}; }
PBRT_HD vec3 operator-(float s, vec3 a) { return {s - a.x, s - a.y, s - a.z}; }
PBRT_HD vec3& operator+=(vec3& a, vec3 b) { a = a + b; return a; }
PBRT_HD vec3& operator-=(vec3& a, vec3 b) { a = a - b; return a; }
PBRT_HD vec3& operator*=(vec3& a, vec3 b) { a = a * b; return a; }
PBRT_HD vec3& operator*=(vec3& a, float s) { a = a * s; return a; }
PBRT_HD vec3& operator/=(vec3& a, float s) { a = a / s; return a; }
PBRT_HD bool operator==(vec3 a, vec3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
PBRT_HD bool operator!=(vec3 a, vec3 b) { return !(a == b); }

PBRT_HD float dot(vec3 a, vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
PBRT_HD float dot(vec2 a, vec2 b) { return a.x * b.x + a.y * b.y; }
PBRT_HD vec3 cross(vec3 a, vec3 b) {
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
PBRT_HD float length(vec3 a) { return sqrtf(dot(a, a)); }
PBRT_HD vec3 normalize(vec3 a) { return a * (1.0f / sqrtf(dot(a, a))); }
PBRT_HD float distance(vec3 a, vec3 b) { return length(a - b); }
PBRT_HD float gmin(float a, float b) { return fminf(a, b); }
PBRT_HD float gmax(float a, float b) { return fmaxf(a, b); }
PBRT_HD float gclamp(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
PBRT_HD vec3 vmin(vec3 a, vec3 b) { return {fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)}; }
PBRT_HD vec3 vmax(vec3 a, vec3 b) { return {fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)}; }
PBRT_HD vec3 vmin(vec3 a, float b) { return vmin(a, make_vec3(b)); }
PBRT_HD vec3 vclamp(vec3 x, float lo, float hi) {
    return {gclamp(x.x, lo, hi), gclamp(x.y, lo, hi), gclamp(x.z, lo, hi)};
}
PBRT_HD float gmix(float a, float b, float t) { return a * (1.0f - t) + b * t; }
PBRT_HD vec3 vmix(vec3 a, vec3 b, float t) { return a * (1.0f - t) + b * t; }
PBRT_HD vec3 reflect(vec3 i, vec3 n) { return i - 2.0f * dot(n, i) * n; }
PBRT_HD float max3(vec3 v) { return fmaxf(fmaxf(v.x, v.y), v.z); }
PBRT_HD vec3 vpow(vec3 v, float e) { return {powf(v.x, e), powf(v.y, e), powf(v.z, e)}; }
PBRT_HD float smoothstep(float e0, float e1, float x) {
    float t = gclamp((x - e0) / (e1 - e0), 0.0f, 1.0f);
    return t * t * (3.0f - 2.0f * t);
}

// Column-major 3x3 (GLSL mat3): c[i] is column i.
struct mat3 {
    vec3 c[3];
};
PBRT_HD vec3 operator*(const mat3& m, vec3 v) { return m.c[0] * v.x + m.c[1] * v.y + m.c[2] * v.z; }

// Column-major 4x4 (GLSL / vsg mat4): m[col * 4 + row].
PBRT_HD vec4 mat4_mul(const float* m, vec4 v) {
    return {m[0] * v.x + m[4] * v.y + m[8] * v.z + m[12] * v.w, m[1] * v.x + m[5] * v.y + m[9] * v.z + m[13] * v.w,
            m[2] * v.x + m[6] * v.y + m[10] * v.z + m[14] * v.w, m[3] * v.x + m[7] * v.y + m[11] * v.z + m[15] * v.w};
}

PBRT_HD uint32_t float_as_uint(float f) { return __builtin_bit_cast(uint32_t, f); }
PBRT_HD float uint_as_float(uint32_t u) { return __builtin_bit_cast(float, u); }

}  // namespace pbrt
