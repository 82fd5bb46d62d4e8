// Transcendental functions of the tracer's shading arithmetic, defined by the
// project: fixed polynomials evaluated with explicit fused multiply-adds and
// exact bit and exponent operations, so the GPU and any C++ host compiler that
// evaluates them without contraction compute every result bit for bit alike.
// The shaders' exp2 / log2 / pow / sin / cos (brdf.glsl, color.glsl,
// sampling.glsl) have driver-defined precision; these stay within a few fp32 ulps
// of the correctly rounded values on the arguments the path tracer passes.
//
//   log2   x = m 2^e with m in [sqrt(1/2), sqrt(2)); log2(1 + f) = f P(f), f = m - 1,
//          P of degree 8 fitted by least squares (relative error) on Chebyshev nodes
//          of [sqrt(1/2) - 1, sqrt(2) - 1]: within 1.3e-7 relative
//   exp2   y = n + f, n = rint(y), |f| <= 1/2; 2^f = Q(f), Q of degree 6 fitted the same
//          way on [-1/2, 1/2] (within 8e-8 relative), scaled by 2^n with ldexp (exact)
//   pow    exp2(y log2 x) for x >= 0
//   pow5   x^5 as (x^2)^2 x
//   sincos Cody-Waite reduction by pi/2 (pi/2 in three floats, the first two with 12
//          significant bits so k * part is exact for |k| < 2^12), sin and cos of the
//          remainder as polynomials in r^2 fitted the same way on [0, pi/4]: within
//          1.5 fp32 ulp for |x| <= 4
#pragma once

#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#define PBRT_FN __host__ __device__ inline
#else
#define PBRT_FN inline
#endif

namespace pbrt_fn {

PBRT_FN uint32_t f2u(float f) { return __builtin_bit_cast(uint32_t, f); }
PBRT_FN float u2f(uint32_t u) { return __builtin_bit_cast(float, u); }

PBRT_FN float log2(float x) {
    if (!(x > 0.0f)) return x == 0.0f ? -INFINITY : NAN;
    if (x == INFINITY) return x;
    int e = 0;
    if (x < 1.17549435e-38f) {  // subnormal: scale into the normal range
        x *= 16777216.0f;
        e = -24;
    }
    uint32_t b = f2u(x);
    e += (int)((b >> 23) & 0xFFu) - 127;
    b = (b & 0x007FFFFFu) | 0x3F800000u;  // m in [1, 2)
    if (b > 0x3FB504F3u) {                // m > sqrt(2): halve it
        b -= 0x00800000u;
        ++e;
    }
    const float f = u2f(b) - 1.0f;  // exact
    float p = 1.228052825e-01f;
    p = fmaf(p, f, -2.057056129e-01f);
    p = fmaf(p, f, 2.161191255e-01f);
    p = fmaf(p, f, -2.391948998e-01f);
    p = fmaf(p, f, 2.879017591e-01f);
    p = fmaf(p, f, -3.606921136e-01f);
    p = fmaf(p, f, 4.809107780e-01f);
    p = fmaf(p, f, -7.213475108e-01f);
    p = fmaf(p, f, 1.442695022e+00f);
    return fmaf(f, p, (float)e);
}

PBRT_FN float exp2(float y) {
    if (y != y) return y;
    if (y >= 128.0f) return INFINITY;
    if (y < -150.0f) return 0.0f;
    const float n = rintf(y);
    const float f = y - n;  // exact
    float q = 1.533757750e-04f;
    q = fmaf(q, f, 1.339985989e-03f);
    q = fmaf(q, f, 9.618519805e-03f);
    q = fmaf(q, f, 5.550329015e-02f);
    q = fmaf(q, f, 2.402264625e-01f);
    q = fmaf(q, f, 6.931471825e-01f);
    q = fmaf(q, f, 1.0f);
    return ldexpf(q, (int)n);
}

// x^y for x >= 0 (pow(0, y > 0) = 0)
PBRT_FN float pow(float x, float y) { return exp2(y * log2(x)); }

PBRT_FN float pow5(float x) {
    const float x2 = x * x;
    const float x4 = x2 * x2;
    return x4 * x;
}

PBRT_FN void sincos(float x, float& sn, float& cs) {
    const float k = rintf(x * 0.636619772f);
    float r = fmaf(-k, 1.5703125f, x);
    r = fmaf(-k, 4.83751297e-4f, r);
    r = fmaf(-k, 7.54979013e-8f, r);
    const float r2 = r * r;
    const float ps = fmaf(fmaf(-1.95202127e-4f, r2, 8.33220407e-3f), r2, -0.166666552f);
    const float s = fmaf(r * r2, ps, r);
    const float pc = fmaf(fmaf(fmaf(2.43941031e-5f, r2, -1.38868089e-3f), r2, 4.16666232e-2f), r2, -0.5f);
    const float c = fmaf(r2, pc, 1.0f);
    const int q = (int)k & 3;
    sn = (q & 1) ? c : s;
    cs = (q & 1) ? s : c;
    if (q == 1 || q == 2) cs = -cs;
    if (q >= 2) sn = -sn;
}

}  // namespace pbrt_fn
