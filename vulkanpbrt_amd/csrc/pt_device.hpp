// Device-side path tracing building blocks: RNG, camera rays, surface
// reconstruction at a hit, GGX/Disney BRDF, BRDF/light sampling and next-event
// estimation.  Each function states the reference shader it follows; the
// arithmetic order is kept where the sample stream or parity depends on it.
#pragma once

#include "pbrt_fn.hpp"
#include "pbrt_math.hpp"
#include "scene_types.hpp"

namespace pbrt {

// ptConstants.glsl:4-11
constexpr float kPi = 3.14159265359f;
constexpr float kReciprocalPi = 0.31830988618f;
constexpr float kEpsilon = 1e-6f;
constexpr float kMinRoughness = 0.04f;
constexpr float kMaxRadiance = 1e1f;
constexpr float kMinTermination = 0.05f;

// ---------------------------------------------------------------------------
// RNG: combined Tausworthe + LCG (random.glsl:13-71), seeded by Wang hashes
// (random.glsl:74-96).  All integer arithmetic wraps mod 2^32.
// ---------------------------------------------------------------------------
struct Rng {
    uint32_t x, y, z, w;
};

PBRT_HD uint32_t wang_hash(uint32_t seed) {
    seed = (seed ^ 61u) ^ (seed >> 16);
    seed *= 9u;
    seed = seed ^ (seed >> 4);
    seed *= 0x27d4eb2du;
    seed = seed ^ (seed >> 15);
    return seed;
}

PBRT_HD uint32_t taus_step(uint32_t z, int s1, int s2, int s3, uint32_t m) {
    uint32_t b = ((z << s1) ^ z) >> s2;
    return ((z & m) << s3) ^ b;
}

PBRT_HD Rng rng_init(uint32_t px, uint32_t py, uint32_t frame) {
    Rng r;
    r.x = wang_hash(px);
    r.y = wang_hash(py);
    r.z = wang_hash(frame);
    r.w = wang_hash((px + py) * frame);
    return r;
}

// randomFloat: float(x ^ y ^ z ^ w) * 2^-32.  The literal 2.3283064365387e-10
// rounds to exactly 2^-32 in fp32, so the product is an exact power-of-two
// scaling of the round-to-nearest u32 -> f32 conversion (values >= 2^32 - 128
// give exactly 1.0).
PBRT_HD float rng_float(Rng& r) {
    r.x = taus_step(r.x, 13, 19, 12, 4294967294u);
    r.y = taus_step(r.y, 2, 25, 4, 4294967288u);
    r.z = taus_step(r.z, 3, 11, 17, 4294967280u);
    r.w = 1664525u * r.w + 1013904223u;
    return (float)(r.x ^ r.y ^ r.z ^ r.w) * 2.3283064365386963e-10f;
}

PBRT_HD vec2 rng_vec2(Rng& r) {
    float a = rng_float(r);
    float b = rng_float(r);
    return {a, b};
}
PBRT_HD vec3 rng_vec3(Rng& r) {
    float a = rng_float(r);
    float b = rng_float(r);
    float c = rng_float(r);
    return {a, b, c};
}

// ---------------------------------------------------------------------------
// Colour helpers (color.glsl:4-47)
// ---------------------------------------------------------------------------
PBRT_HD float luminance(vec3 c) { return dot(c, make_vec3(0.2126f, 0.7152f, 0.0722f)); }
PBRT_HD vec3 srgb_to_linear(vec3 c) {
    return make_vec3(pbrt_fn::pow(c.x, 2.2f), pbrt_fn::pow(c.y, 2.2f), pbrt_fn::pow(c.z, 2.2f));
}
// pow(c, 1/2.2) of the final-image output (ptRaygen.rgen:100)
PBRT_HD vec3 linear_to_srgb(vec3 c) {
    const float e = 1.0f / 2.2f;
    return make_vec3(pbrt_fn::pow(c.x, e), pbrt_fn::pow(c.y, e), pbrt_fn::pow(c.z, e));
}

PBRT_HD float convert_metallic(vec3 diffuse, vec3 specular, float max_specular) {
    float perceived_diffuse =
        sqrtf(0.299f * diffuse.x * diffuse.x + 0.587f * diffuse.y * diffuse.y + 0.114f * diffuse.z * diffuse.z);
    float perceived_specular =
        sqrtf(0.299f * specular.x * specular.x + 0.587f * specular.y * specular.y + 0.114f * specular.z * specular.z);
    if (perceived_specular < kMinRoughness) return 0.0f;
    float a = kMinRoughness;
    float b = perceived_diffuse * (1.0f - max_specular) / (1.0f - kMinRoughness) + perceived_specular -
              2.0f * kMinRoughness;
    float c = kMinRoughness - perceived_specular;
    float d = fmaxf(b * b - 4.0f * a * c, 0.0f);
    return gclamp((-b + sqrtf(d)) / (2.0f * a), 0.0f, 1.0f);
}

// ---------------------------------------------------------------------------
// Surface description handed from the hit stage to shading
// (struct SurfaceInfo, ptStructures.glsl:66-80).
// ---------------------------------------------------------------------------
struct SurfaceInfo {
    float perceptual_roughness;
    float metalness;
    float alpha_roughness;
    int illumination_type;
    vec3 reflectance0;
    vec3 reflectance90;
    vec3 diffuse_color;
    vec3 specular_color;
    vec3 emissive_color;
    vec3 transmissive_color;
    vec3 normal;
    mat3 basis;
    float index_of_refraction;
};

// Miss payload (ptMiss.rmiss:8-25).
PBRT_HD void miss_surface(SurfaceInfo& s, vec3& position, uint32_t& category) {
    position = make_vec3(1.0e10f);
    category = 0xFFFFFFFFu;
    s.normal = make_vec3(1.0f);
    s.emissive_color = make_vec3(0.0f);
    s.diffuse_color = make_vec3(0.0f);
    s.specular_color = make_vec3(0.0f);
    s.perceptual_roughness = 0.0f;
    s.metalness = 0.0f;
    s.alpha_roughness = 0.0f;
    s.illumination_type = 0;
    s.reflectance0 = make_vec3(0.0f);
    s.reflectance90 = make_vec3(0.0f);
    s.transmissive_color = make_vec3(0.0f);
    s.basis.c[0] = s.basis.c[1] = s.basis.c[2] = make_vec3(0.0f);
    s.index_of_refraction = 1.0f;
}

PBRT_HD bool is_miss(const SurfaceInfo& s) { return s.normal == make_vec3(1.0f); }

// ---------------------------------------------------------------------------
// BRDF (brdf.glsl:61-164)
// ---------------------------------------------------------------------------
PBRT_HD float specular_sample_weight(const SurfaceInfo& s) {
    float wd = gmix(luminance(s.diffuse_color), 0.0f, s.metalness);
    float ws = luminance(s.specular_color);
    return fminf(1.0f, ws / (wd + wd));
}

PBRT_HD float pdf_brdf(const SurfaceInfo& s, vec3 v, vec3 l, vec3 h) {
    (void)v;
    float pdf_diffuse = kReciprocalPi * dot(l, s.normal);
    float alpha2 = s.alpha_roughness * s.alpha_roughness;
    float nh = dot(h, s.normal);
    float denom = (nh * nh) * (alpha2 - 1.0f) + 1.0f;
    float pdf_specular = alpha2 / (kPi * (denom * denom));
    pdf_specular *= dot(h, s.normal);
    float sw = specular_sample_weight(s);
    return gmix(pdf_diffuse, pdf_specular, sw);
}

PBRT_HD vec3 brdf_eval(vec3 v, vec3 l, vec3 h, const SurfaceInfo& s) {
    vec3 n = s.normal;
    float ndotl = gclamp(dot(n, l), 0.001f, 1.0f);
    float ndotv = gclamp(fabsf(dot(n, v)), 0.001f, 1.0f);
    float ndoth = gclamp(dot(n, h), 0.0f, 1.0f);
    float vdoth = gclamp(dot(v, h), 0.0f, 1.0f);
    // specularReflection(PBRInfo): Schlick with the exp2 spherical-gaussian fit
    float fexp = pbrt_fn::exp2((-5.55473f * vdoth - 6.98316f) * vdoth);
    vec3 f = s.reflectance0 + (s.reflectance90 - s.reflectance90 * s.reflectance0) * fexp;
    // geometricOcclusion: r = alphaRoughness^2
    float r = s.alpha_roughness * s.alpha_roughness;
    float att_l = 2.0f * ndotl / (ndotl + sqrtf(r + (1.0f - r) * (ndotl * ndotl)));
    float att_v = 2.0f * ndotv / (ndotv + sqrtf(r + (1.0f - r) * (ndotv * ndotv)));
    float g = att_l * att_v;
    // microfacetDistribution (GGX / Trowbridge-Reitz)
    float rsq = s.alpha_roughness * s.alpha_roughness;
    float fd = (ndoth * rsq - ndoth) * ndoth + 1.0f;
    float d = rsq / (kPi * fd * fd);
    // BRDF_Diffuse_Disney; f0 = 0.1 => dim = 0.9
    float fd90 = 0.5f + 2.0f * s.perceptual_roughness * vdoth * vdoth;
    const float dim = 0.9f;
    float result = ((1.0f + (fd90 - 1.0f) * pbrt_fn::pow5(1.0f - ndotl)) *
                    (1.0f + (fd90 - 1.0f) * pbrt_fn::pow5(1.0f - ndotv))) *
                   kReciprocalPi * dim;
    vec3 diffuse_contrib = (1.0f - f) * (s.diffuse_color * result);
    vec3 spec_contrib = f * g * d / (4.0f * ndotl * ndotv);
    if (dot(n, l) <= 0.0f) spec_contrib = make_vec3(0.0f);
    return diffuse_contrib + spec_contrib;
}

// specularReflection(vec3, vec3, float) (brdf.glsl:94-98)
// with reflectance0 = 1, reflectance90 = 0 as sampleBRDF calls it.
PBRT_HD float fresnel_schlick_scalar(float vdoth) {
    return 1.0f + (0.0f - 1.0f) * pbrt_fn::pow5(gclamp(1.0f - vdoth, 0.0f, 1.0f));
}

// ---------------------------------------------------------------------------
// Sampling (sampling.glsl:11-78)
// ---------------------------------------------------------------------------
PBRT_HD vec3 sample_ggx(vec2 r, float alpha2) {
    float phi = 2.0f * kPi * r.x;
    float cos_h = sqrtf((1.0f - r.y) / (1.0f + (alpha2 - 1.0f) * r.y));
    float sin_h = sqrtf(1.0f - cos_h * cos_h);
    float sp, cp;
    pbrt_fn::sincos(phi, sp, cp);
    return {sin_h * cp, sin_h * sp, cos_h};
}

PBRT_HD vec3 sample_hemisphere(vec2 r) {
    float t = 2.0f * kPi * r.y;
    float sr = sqrtf(r.x);
    float st, ct;
    pbrt_fn::sincos(t, st, ct);
    float dx = sr * ct;
    float dy = sr * st;
    return {dx, dy, sqrtf(fmaxf(0.0f, 1.0f - dx * dx - dy * dy))};
}

PBRT_HD vec2 sample_triangle(vec2 u) {
    float s = sqrtf(u.x);
    return {1.0f - s, u.y * s};
}

// sampleBRDF receives the RNG *by value* (sampling.glsl:29): the caller's
// stream is not advanced, so its next draws repeat u.x, u.y, ...
PBRT_HD vec3 sample_brdf(const SurfaceInfo& s, Rng re, vec3 v, vec3& l, float& pdf) {
    vec3 h;
    vec3 u = rng_vec3(re);
    float sw = specular_sample_weight(s);
    mat3 basis = s.basis;
    bool entering = dot(v, s.normal) >= 0.0f;
    if (s.illumination_type == 7 && !entering) basis.c[2] = basis.c[2] * -1.0f;
    if (u.z < sw) {
        h = sample_ggx(make_vec2(u.x, u.y), s.alpha_roughness * s.alpha_roughness);
        h = basis * h;
        l = -reflect(v, h);
    } else {
        l = sample_hemisphere(make_vec2(u.x, u.y));
        l = basis * l;
        h = normalize(v + l);
    }
    pdf = pdf_brdf(s, v, l, h);
    vec3 brdf = brdf_eval(v, l, h, s);
    if (s.illumination_type == 7) {
        float f = fresnel_schlick_scalar(dot(v, h));
        if (rng_float(re) < f || !entering) {
            float t = !entering ? s.index_of_refraction : 1.0f / s.index_of_refraction;
            float cos_i = dot(basis.c[2], l);
            float sin_t2 = t * t * (1.0f - cos_i * cos_i);
            if (sin_t2 <= 1.0f) {
                l = l - 2.0f * cos_i * basis.c[2];
                float cos_t = sqrtf(1.0f - sin_t2);
                l = t * l + (t * cos_i - cos_t) * basis.c[2];
            }
            brdf = s.transmissive_color;
            pdf = 1.0f;
        } else {
            pdf = 1.0f;
            brdf = make_vec3(1.0f);
        }
    }
    return brdf;
}

// Sky colour (lighting.glsl:233-240); SRGBtoLINEAR(0.3, 0.5, 0.92) folded at
// run time to keep the pow() identical to the shader.
PBRT_HD vec3 sky_color(vec3 dir) {
    vec3 upper = srgb_to_linear(make_vec3(0.3f, 0.5f, 0.92f));
    upper = vmix(make_vec3(1.0f), upper, fmaxf(dir.z, 0.0f));
    vec3 lower = make_vec3(0.2f, 0.2f, 0.2f);
    float w = smoothstep(-0.02f, 0.02f, dir.z);
    return vmix(lower, upper, w);
}

PBRT_HD float power_heuristic(float a, float b) {
    float f = a * a;
    float g = b * b;
    return f / (f + g);
}

// ---------------------------------------------------------------------------
// Light sampling.  The shadow ray is returned instead of traced so that the
// caller decides whether to trace it inline or defer it to a shadow queue;
// the RNG consumption does not depend on visibility.
// ---------------------------------------------------------------------------
struct LightSample {
    vec3 radiance;  // pickedLightStrength (before the visibility factor)
    vec3 l;         // direction to the light
    float pdf;
    float tmax;
    bool valid;     // false => sampleLight returned vec3(0) without a shadow ray
};

PBRT_HD float light_attenuation(const DeviceLight& li, float d) {
    return 1.0f / (li.strengths.x + li.strengths.y * d + li.strengths.z * d * d);
}

// LIGHT_SAMPLE_SURFACE_STRENGTH (lighting.glsl:18-96): single-pass weighted
// reservoir over all lights.
PBRT_HD LightSample sample_light_surface_strength(const DeviceLight* lights, uint32_t light_count, vec3 pos,
                                                  vec3 n, Rng& re) {
    LightSample out;
    out.l = make_vec3(0.0f);
    out.pdf = 0.0f;
    out.tmax = 1000.0f;
    out.radiance = make_vec3(0.0f);
    out.valid = false;
    const float tmin = 0.001f;
    float strength_sum = 0.0f;
    float picked_strength = 0.0f;
    vec3 picked_light_strength = make_vec3(0.0f);
    for (uint32_t i = 0; i < light_count; ++i) {
        const DeviceLight li = lights[i];
        vec4 ctot = make_vec4(li.col_ambient.x + li.col_diffuse.x + li.col_specular.x,
                              li.col_ambient.y + li.col_diffuse.y + li.col_specular.y,
                              li.col_ambient.z + li.col_diffuse.z + li.col_specular.z,
                              li.col_ambient.w + li.col_diffuse.w + li.col_specular.w);
        float light_power = ctot.x + ctot.y + ctot.z + ctot.w;
        float strength = 0.0f;
        vec3 light_strength = make_vec3(ctot.x, ctot.y, ctot.z);
        float cur_tmax = 1000.0f;
        vec3 cur_l = make_vec3(0.0f);
        int type = (int)li.v0_type.w;
        if (type == 1) {  // directional
            float d = distance(pos, xyz(li.v0_type));
            float att = light_attenuation(li, d);
            vec3 ndir = -xyz(li.dir_angle2);
            strength = fmaxf(dot(n, ndir), 0.0f) * light_power * att;
            light_strength *= dot(n, ndir) * att;
            cur_l = normalize(ndir);
        } else if (type == 2) {  // point
            float d = distance(pos, xyz(li.v0_type));
            float att = light_attenuation(li, d);
            vec3 to = normalize(xyz(li.v0_type) - pos);
            strength = fmaxf(dot(n, to), 0.0f) * light_power * att;
            light_strength *= dot(n, to) * att;
            cur_l = to;
        } else if (type == 5) {  // area triangle
            vec2 bc = sample_triangle(rng_vec2(re));
            vec3 p1 = xyz(li.v0_type), p2 = xyz(li.v1_strength), p3 = xyz(li.v2_angle);
            vec3 light_p = (1.0f - bc.x - bc.y) * p1 + bc.x * p2 + bc.y * p3;
            vec3 light_dir = light_p - pos;
            vec3 light_normal = normalize(cross(p2 - p1, p3 - p1));
            float d = length(light_dir);
            light_dir /= d;
            float att = light_attenuation(li, d);
            float cos_s = fmaxf(dot(n, light_dir), 0.0f);
            float cos_l = fmaxf(dot(-light_dir, light_normal), 0.0f);
            strength = cos_s * cos_l * light_power * att;
            light_strength *= cos_s * cos_l * att;
            cur_l = light_dir;
            cur_tmax = d - tmin;
        }
        strength_sum += strength;
        if (rng_float(re) < strength / strength_sum) {
            picked_strength = strength;
            picked_light_strength = light_strength;
            out.tmax = cur_tmax;
            out.l = cur_l;
        }
    }
    if (strength_sum < 1e-6f || picked_strength < 1e-6f) {
        out.pdf = 0.0f;
        out.l = make_vec3(0.0f);
        return out;
    }
    out.pdf = picked_strength / strength_sum;
    out.radiance = picked_light_strength;
    out.valid = true;
    return out;
}

// LIGHT_SAMPLE_LIGHT_STRENGTH (lighting.glsl:98-168): binary search over the
// inclusive strength prefix sum; the result is scaled by sum / strength.
PBRT_HD LightSample sample_light_light_strength(const DeviceLight* lights, uint32_t light_count,
                                                float strength_sum_total, vec3 pos, vec3 n, Rng& re) {
    LightSample out;
    out.pdf = 0.0f;
    out.l = make_vec3(0.0f);
    out.tmax = 1000.0f;
    out.valid = false;
    out.radiance = make_vec3(0.0f);
    float rnd = rng_float(re);
    float picked = rnd * strength_sum_total;
    int begin = 0, end = (int)light_count, mid = (begin + end) / 2;
    int c = 0;
    while (end - begin > 1 && ++c < 100) {
        if (picked <= lights[mid].strengths.w)
            end = mid + 1;
        else
            begin = mid;
        mid = (begin + end) / 2;
    }
    float lstrength = lights[begin].strengths.w;
    if (begin > 0) lstrength -= lights[begin - 1].strengths.w;
    const DeviceLight li = lights[begin];
    vec3 light_strength = xyz(li.col_ambient) + xyz(li.col_diffuse) + xyz(li.col_specular);
    const float tmin = 0.001f;
    int type = (int)li.v0_type.w;
    vec3 l = make_vec3(0.0f);
    float tmax = 1000.0f;
    if (type == 1) {
        float d = distance(pos, xyz(li.v0_type));
        float att = light_attenuation(li, d);
        light_strength *= fmaxf(dot(n, -xyz(li.dir_angle2)), 0.0f) * att;
        l = normalize(-xyz(li.dir_angle2));
    } else if (type == 2) {
        float d = distance(pos, xyz(li.v0_type));
        float att = light_attenuation(li, d);
        light_strength *= fmaxf(dot(n, normalize(xyz(li.v0_type) - pos)), 0.0f) * att;
        l = normalize(xyz(li.v0_type) - pos);
    } else if (type == 5) {
        vec2 bc = sample_triangle(rng_vec2(re));
        vec3 p1 = xyz(li.v0_type), p2 = xyz(li.v1_strength), p3 = xyz(li.v2_angle);
        vec3 light_p = (1.0f - bc.x - bc.y) * p1 + bc.x * p2 + bc.y * p3;
        vec3 light_dir = light_p - pos;
        vec3 ln = cross(p2 - p1, p3 - p1);
        float area = 0.5f * length(ln);
        ln = normalize(ln);
        float d = length(light_dir);
        light_dir /= d;
        float att = light_attenuation(li, d);
        light_strength *= fmaxf(dot(n, light_dir), 0.0f) * fmaxf(dot(-light_dir, ln), 0.0f) * att * area;
        l = light_dir;
        tmax = d - tmin;
    }
    if (length(light_strength) < 1e-6f) return out;
    out.pdf = 1.0f;
    out.l = l;
    out.tmax = tmax;
    out.radiance = light_strength * strength_sum_total / lstrength;
    out.valid = true;
    return out;
}

// Uniform light selection (lighting.glsl:170-227).
// The light UNIFORM sampling picks for the random number rnd in [0, 1).
PBRT_HD int uniform_light_index(float rnd, uint32_t light_count) {
    return (int)(rnd * (float)light_count) - (int)rnd;
}

PBRT_HD LightSample sample_light_uniform(const DeviceLight* lights, uint32_t light_count, vec3 pos, vec3 n,
                                         Rng& re) {
    LightSample out;
    out.pdf = 0.0f;
    out.l = make_vec3(0.0f);
    out.tmax = 1000.0f;
    out.valid = false;
    out.radiance = make_vec3(0.0f);
    const int i = uniform_light_index(rng_float(re), light_count);
    const DeviceLight li = lights[i];
    vec3 light_strength = xyz(li.col_ambient) + xyz(li.col_diffuse) + xyz(li.col_specular);
    const float tmin = 0.001f;
    int type = (int)li.v0_type.w;
    vec3 l = make_vec3(0.0f);
    float tmax = 1000.0f;
    if (type == 1) {
        float d = distance(pos, xyz(li.v0_type));
        float att = light_attenuation(li, d);
        light_strength *= fmaxf(dot(n, -xyz(li.dir_angle2)), 0.0f) * att;
        l = normalize(-xyz(li.dir_angle2));
    } else if (type == 2) {
        float d = distance(pos, xyz(li.v0_type));
        float att = light_attenuation(li, d);
        light_strength *= fmaxf(dot(n, normalize(xyz(li.v0_type) - pos)), 0.0f) * att;
        l = normalize(xyz(li.v0_type) - pos);
    } else if (type == 5) {
        vec2 bc = sample_triangle(rng_vec2(re));
        vec3 p1 = xyz(li.v0_type), p2 = xyz(li.v1_strength), p3 = xyz(li.v2_angle);
        vec3 light_p = (1.0f - bc.x - bc.y) * p1 + bc.x * p2 + bc.y * p3;
        vec3 light_dir = light_p - pos;
        vec3 ln = cross(p2 - p1, p3 - p1);
        float area = 0.5f * length(ln);
        ln = normalize(ln);
        float d = length(light_dir);
        light_dir /= d;
        float att = light_attenuation(li, d);
        light_strength *= fmaxf(dot(n, light_dir), 0.0f) * fmaxf(dot(-light_dir, ln), 0.0f) * att * area;
        l = light_dir;
        tmax = d - tmin;
    }
    if (length(light_strength) < 1e-6f) return out;
    out.pdf = 1.0f;
    out.l = l;
    out.tmax = tmax;
    out.radiance = light_strength * (float)light_count;
    out.valid = true;
    return out;
}

// Unshadowed part of nextEventEsitmation (lighting.glsl:246-261): the value
// added when the shadow ray is unoccluded, min(throughput * L, 10).  A zero
// picked strength returns vec3(0) before the BRDF is evaluated, exactly as
// `if(lightCol == vec3(0)) return lightCol;` does for occluded samples.
PBRT_HD vec3 nee_unshadowed(const LightSample& ls, vec3 o, const SurfaceInfo& s, vec3 throughput) {
    if (ls.radiance == make_vec3(0.0f)) return make_vec3(0.0f);
    vec3 lh = normalize(ls.l + o);
    vec3 brdf = brdf_eval(o, ls.l, lh, s);
    float brdf_pdf = pdf_brdf(s, o, ls.l, lh);
    float w = power_heuristic(ls.pdf, brdf_pdf);
    vec3 light = ls.radiance * brdf * w / ls.pdf;
    return vmin(throughput * light, kMaxRadiance);
}

}  // namespace pbrt
