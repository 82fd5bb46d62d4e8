// Shared helpers of the CPU oracle (test infrastructure only: the product never
// links or calls anything under oracle/).
//
// Every function is a plain restatement of the reference GLSL, templated on the
// real type T: T = float is the operation-order FP32 mirror used for
// per-pixel parity, T = double is the FP64 oracle used for converged images and
// least-squares ground truth.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace orc {

template <class T>
struct V3 {
    T x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(T a, T b, T c) : x(a), y(b), z(c) {}
    explicit V3(T s) : x(s), y(s), z(s) {}
};
template <class T>
struct V2 {
    T x, y;
};
template <class T>
struct V4 {
    T x, y, z, w;
};

template <class T> inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline V3<T> operator-(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T> inline V3<T> operator*(V3<T> a, V3<T> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class T> inline V3<T> operator/(V3<T> a, V3<T> b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
template <class T> inline V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> inline V3<T> operator*(T s, V3<T> a) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> inline V3<T> operator/(V3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <class T> inline V3<T> operator+(V3<T> a, T s) { return {a.x + s, a.y + s, a.z + s}; }
template <class T> inline V3<T> operator-(T s, V3<T> a) { return {s - a.x, s - a.y, s - a.z}; }
template <class T> inline bool operator==(V3<T> a, V3<T> b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
template <class T> inline bool operator==(V2<T> a, V2<T> b) { return a.x == b.x && a.y == b.y; }
template <class T> inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> inline V3<T> cross(V3<T> a, V3<T> b) {
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
template <class T> inline T length(V3<T> a) { return std::sqrt(dot(a, a)); }
template <class T> inline V3<T> normalize(V3<T> a) { return a * (T(1) / std::sqrt(dot(a, a))); }
// GLSL min/max/clamp: NaN-suppressing (fmin/fmax), as the GPU build uses.
template <class T> inline T gmin(T a, T b) { return std::fmin(a, b); }
template <class T> inline T gmax(T a, T b) { return std::fmax(a, b); }
template <class T> inline T gclamp(T x, T lo, T hi) { return std::fmin(std::fmax(x, lo), hi); }
template <class T> inline T gmix(T a, T b, T t) { return a * (T(1) - t) + b * t; }
template <class T> inline V3<T> vmix(V3<T> a, V3<T> b, T t) { return a * (T(1) - t) + b * t; }
template <class T> inline V3<T> vmin(V3<T> a, T b) { return {gmin(a.x, b), gmin(a.y, b), gmin(a.z, b)}; }
template <class T> inline V3<T> vclamp(V3<T> a, T lo, T hi) {
    return {gclamp(a.x, lo, hi), gclamp(a.y, lo, hi), gclamp(a.z, lo, hi)};
}
template <class T> inline T max3(V3<T> v) { return gmax(gmax(v.x, v.y), v.z); }
template <class T> inline V3<T> vpow(V3<T> v, T e) { return {std::pow(v.x, e), std::pow(v.y, e), std::pow(v.z, e)}; }
template <class T> inline V3<T> reflect(V3<T> i, V3<T> n) { return i - T(2) * dot(n, i) * n; }

// Column-major 4x4 product (GLSL mat4 * vec4).
template <class T>
inline V4<T> mat4_mul(const float* m, V4<T> v) {
    return {T(m[0]) * v.x + T(m[4]) * v.y + T(m[8]) * v.z + T(m[12]) * v.w,
            T(m[1]) * v.x + T(m[5]) * v.y + T(m[9]) * v.z + T(m[13]) * v.w,
            T(m[2]) * v.x + T(m[6]) * v.y + T(m[10]) * v.z + T(m[14]) * v.w,
            T(m[3]) * v.x + T(m[7]) * v.y + T(m[11]) * v.z + T(m[15]) * v.w};
}

inline uint32_t f2u(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}
inline float u2f(uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// IEEE binary16 conversions, round to nearest even (R16F / RGBA16F stores).
inline uint16_t float_to_half(float f) {
    uint32_t x = f2u(f);
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t mant = x & 0x007FFFFFu;
    int32_t exp = (int32_t)((x >> 23) & 0xFF);
    if (exp == 0xFF) return (uint16_t)(sign | 0x7C00u | (mant ? 0x200u | (mant >> 13) : 0u));
    int32_t e = exp - 127 + 15;
    if (e >= 0x1F) return (uint16_t)(sign | 0x7C00u);
    if (e <= 0) {
        if (e < -10) return (uint16_t)sign;
        mant |= 0x00800000u;
        uint32_t shift = (uint32_t)(14 - e);
        uint32_t half = mant >> shift;
        uint32_t rem = mant & ((1u << shift) - 1u);
        uint32_t mid = 1u << (shift - 1);
        if (rem > mid || (rem == mid && (half & 1u))) ++half;
        return (uint16_t)(sign | half);
    }
    uint32_t half = sign | ((uint32_t)e << 10) | (mant >> 13);
    uint32_t rem = mant & 0x1FFFu;
    if (rem > 0x1000u || (rem == 0x1000u && (half & 1u))) ++half;
    return (uint16_t)half;
}

inline float half_to_float(uint16_t h) {
    uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    uint32_t exp = (h >> 10) & 0x1Fu;
    uint32_t mant = h & 0x3FFu;
    if (exp == 0) {
        if (mant == 0) return u2f(sign);
        float v = std::ldexp((float)mant, -24);
        return sign ? -v : v;
    }
    if (exp == 0x1F) return u2f(sign | 0x7F800000u | (mant << 13));
    return u2f(sign | ((exp - 15 + 127) << 23) | (mant << 13));
}

// UNORM8 store: clamp to [0, 1], scale by 255, round to nearest even.
inline uint8_t unorm8(float x) {
    x = std::fmin(std::fmax(x, 0.0f), 1.0f);
    return (uint8_t)std::nearbyint(x * 255.0f);
}

}  // namespace orc
