// BFR (blockwise feature regression fitted by Adam) and the three-size blender
// on gfx950.
//
// bfr.comp fits, per b x b block, a 7-feature linear model of the accumulated
// demodulated illumination,
//   features [1, sx, sy, 2 z - 1, n.x, n.y, n.z]   (sx, sy in [-0.5, 0.5] over the
//   block, z = per-block normalised depth), alpha: 7 x rgb
// by 40 steps of Adam on the block-summed residual gradient (plain residual, or its
// sign once a pixel has >= 10 samples; the optimiser constants blend between the
// two regimes by the fraction of such pixels), then reconstructs, blends with the
// reprojected previous output and remodulates (bfr.comp:202-308).
//
// The reference runs one b x b workgroup per block (up to 1024 threads) with two
// barriers per Adam step.  Here a block is 64 (b = 8) or 256 threads holding
// b*b/threads pixels each; a step reduces the 21 gradient sums inside each wave
// (xor butterfly), folds the waves through a double-buffered LDS slot behind one
// barrier, and every thread then applies the same Adam update to its own copy of
// alpha, m and v, so no second barrier or broadcast is needed.
//
// Project rules where bfr.comp depends on the hardware:
//   * the stop test `gradient_rest >= GRADIENT_THRESH (0)` is reduced over all 7
//     alpha components (bfr.comp takes a subgroup-dependent subset); it only ends
//     the loop early when the gradient is NaN.
//   * sums run in a fixed wave / lane tree order (the reference's subgroup order is
//     vendor-specific); results agree with the oracle to fp32 rounding.
#include <hip/hip_runtime.h>

#include "capi_common.hpp"
#include "image_ops.hpp"

namespace pbrt {

namespace {

constexpr int kBfrFeatures = 7;
constexpr int kBfrSums = 3 * kBfrFeatures;  // gradient components per step
constexpr int kBfrIterations = 40;          // GRADIENT_ITERATION_AMT

// Adam constants of bfr.comp:77-88
constexpr float kAlpha = 0.863f, kK = 0.116f, kBeta1 = 0.3f, kBeta2 = 0.7314f, kEps = 1e-8f;
constexpr float kAlphaL = 1.1f, kBeta1L = 0.45f, kBeta2L = 0.75f;
constexpr float kSppThreshold = 10.0f;

// pixelOffsets of bfr.comp:92 (integer, added to the invocation position)
const int kBfrOffsetsHost[16][2] = {{-7, -11}, {-14, -8}, {-5, -12}, {-15, -1}, {-5, -9},  {-1, -4},
                                    {-14, -7}, {0, -13},  {-5, -1},  {-1, 0},   {-15, -2}, {-14, -10},
                                    {-1, -1},  {-6, -3},  {0, -8},   {-10, -4}};

struct BfrGeom {
    int width, height, bs, blocks_x, blocks_y, offx, offy;
    uint32_t frame;
};

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}
__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fminf(v, __shfl_xor(v, off));
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off));
    return v;
}

__device__ __forceinline__ float sign_of(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }

}  // namespace

// kB: block edge; kThreads: 64 or 256; each thread owns kB*kB/kThreads pixels,
// pixel r = tid + p * kThreads at local (r / kB, r % kB) (bfr.comp's ID order).
template <int kB, int kThreads>
__global__ void __launch_bounds__(kThreads) k_bfr(BfrGeom g, pbrt_bfr_images im) {
    constexpr int kPix = kB * kB / kThreads;
    constexpr int kWaves = kThreads / 64;
    __shared__ float part[2][kWaves][kBfrSums];
    __shared__ float red[2][kWaves];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

    vec3 color[kPix];
    float feat[kPix][kBfrFeatures];
    bool l1px[kPix];
    int img_x[kPix], img_y[kPix];
    bool inside[kPix];
    float depth[kPix];
    float dmin = 3.4e38f, dmax = -3.4e38f;
    int l1_local = 0;
#pragma unroll
    for (int p = 0; p < kPix; ++p) {
        const int r = tid + p * kThreads;
        const int lx = r / kB, ly = r % kB;
        const int ax = blockIdx.x * kB + lx + g.offx, ay = blockIdx.y * kB + ly + g.offy;
        img_x[p] = mirror_coord(ax, g.width);
        img_y[p] = mirror_coord(ay, g.height);
        inside[p] = ax == img_x[p] && ay == img_y[p];
        const size_t pix = (size_t)img_y[p] * g.width + img_x[p];
        color[p] = xyz(load_half4(im.noisy, pix));
        depth[p] = im.depth[pix];
        const float2 n = reinterpret_cast<const float2*>(im.normal)[pix];
        const vec3 nn = decode_normal(n.x, n.y);
        const float spp = unorm8_to_float(im.spp[pix]) * 256.0f;
        l1px[p] = spp >= kSppThreshold;
        l1_local += l1px[p] ? 1 : 0;
        dmin = fminf(dmin, depth[p]);
        dmax = fmaxf(dmax, depth[p]);
        feat[p][0] = 1.0f;
        feat[p][1] = (float)lx / (float)(kB - 1) - 0.5f;
        feat[p][2] = (float)ly / (float)(kB - 1) - 0.5f;
        feat[p][4] = nn.x;
        feat[p][5] = nn.y;
        feat[p][6] = nn.z;
    }
    // block min / max depth and the count of pixels in the L1 regime
    dmin = wave_min(dmin);
    dmax = wave_max(dmax);
    float l1f = wave_sum((float)l1_local);
    if (kWaves > 1) {
        if (lane == 0) {
            red[0][wave] = dmin;
            red[1][wave] = dmax;
            part[1][wave][0] = l1f;
        }
        __syncthreads();
        dmin = red[0][0];
        dmax = red[1][0];
        l1f = part[1][0][0];
#pragma unroll
        for (int w = 1; w < kWaves; ++w) {
            dmin = fminf(dmin, red[0][w]);
            dmax = fmaxf(dmax, red[1][w]);
            l1f += part[1][w][0];
        }
        __syncthreads();
    }
#pragma unroll
    for (int p = 0; p < kPix; ++p) {
        float z = depth[p] - dmin;
        z /= dmax - dmin + kEps;
        feat[p][3] = z * 2.0f - 1.0f;
    }
    const float l1_ratio = l1f * 1.0f / (float)(kB * kB);
    const float a_mix = l1_ratio * kAlphaL + (1.0f - l1_ratio) * kAlpha;
    const float b1_mix = l1_ratio * kBeta1L + (1.0f - l1_ratio) * kBeta1;
    const float b2_mix = l1_ratio * kBeta2L + (1.0f - l1_ratio) * kBeta2;

    float alpha[kBfrSums], m[kBfrSums], v[kBfrSums];
#pragma unroll
    for (int j = 0; j < kBfrSums; ++j) alpha[j] = m[j] = v[j] = 0.0f;

    for (int it = 0; it < kBfrIterations; ++it) {
        float grad[kBfrSums];
#pragma unroll
        for (int j = 0; j < kBfrSums; ++j) grad[j] = 0.0f;
#pragma unroll
        for (int p = 0; p < kPix; ++p) {
            vec3 pred = make_vec3(0.0f);
#pragma unroll
            for (int j = 0; j < kBfrFeatures; ++j)
                pred += make_vec3(alpha[3 * j], alpha[3 * j + 1], alpha[3 * j + 2]) * feat[p][j];
            vec3 d = color[p] - pred;
            if (l1px[p]) d = make_vec3(sign_of(d.x), sign_of(d.y), sign_of(d.z));
#pragma unroll
            for (int j = 0; j < kBfrFeatures; ++j) {
                grad[3 * j] += feat[p][j] * d.x;
                grad[3 * j + 1] += feat[p][j] * d.y;
                grad[3 * j + 2] += feat[p][j] * d.z;
            }
        }
#pragma unroll
        for (int j = 0; j < kBfrSums; ++j) grad[j] = wave_sum(grad[j]);
        if (kWaves > 1) {
            const int buf = it & 1;
            if (lane == 0) {
#pragma unroll
                for (int j = 0; j < kBfrSums; ++j) part[buf][wave][j] = grad[j];
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < kBfrSums; ++j) {
                float s = part[buf][0][j];
#pragma unroll
                for (int w = 1; w < kWaves; ++w) s += part[buf][w][j];
                grad[j] = s;
            }
        }
        // Adam with exponential learning-rate decay (bfr.comp:124-131), t = it + 1
        const float t = (float)(it + 1);
        const float step =
            a_mix * expf(-kK * t) * sqrtf(1.0f - powf(kBeta2, t)) / (1.0f - powf(kBeta1, t));
        float rest = 0.0f;
#pragma unroll
        for (int j = 0; j < kBfrSums; ++j) {
            const float dl = grad[j];
            m[j] = b1_mix * m[j] + (1.0f - kBeta1) * dl;
            v[j] = b2_mix * v[j] + (1.0f - kBeta2) * (dl * dl);
            alpha[j] += step * (m[j] / (sqrtf(v[j]) + kEps));
            rest += fabsf(dl);
        }
        if (!(rest >= 0.0f)) break;  // NaN gradient (GRADIENT_THRESH = 0)
    }

    if (im.weights && tid < kBfrSums) {
        float* out = im.weights + ((size_t)blockIdx.y * g.blocks_x + blockIdx.x) * kBfrSums;
#pragma unroll
        for (int j = 0; j < kBfrSums; ++j)
            if (tid == j) out[j] = alpha[j];
    }

    const size_t plane = (size_t)g.width * g.height;
#pragma unroll
    for (int p = 0; p < kPix; ++p) {
        if (!inside[p]) continue;
        const size_t pix = (size_t)img_y[p] * g.width + img_x[p];
        vec3 den = make_vec3(0.0f);
#pragma unroll
        for (int j = 0; j < kBfrFeatures; ++j)
            den += make_vec3(alpha[3 * j], alpha[3 * j + 1], alpha[3 * j + 2]) * feat[p][j];
        den = vclamp(den, 0.0f, 10.0f);
        const ushort2 mv = reinterpret_cast<const ushort2*>(im.motion)[pix];
        const float mux = h2f(mv.x), muy = h2f(mv.y);
        vec3 prev = make_vec3(0.0f);
        float blend = 1.0f;
        if (g.frame > 0 && mux >= 0.0f) {
            const uint16_t* layer = im.denoised + 4 * plane * (g.frame & 1u);
            const int w = g.width;
            prev = xyz(sample_bilinear([&](int sx, int sy) { return load_half4(layer, (size_t)sy * w + sx); },
                                       g.width, g.height, mux, muy));
            const float spp = unorm8_to_float(im.spp[pix]) * 256.0f;
            blend = fmaxf(1.0f / spp, 0.1f);
        }
        den = blend * den + (1.0f - blend) * prev;
        store_half4(im.denoised + 4 * plane * ((g.frame & 1u) ^ 1u), pix, make_vec4(den, 1.0f));
        const vec3 alb = xyz(load_rgba8(im.albedo, pix)) + make_vec3(1e-6f);
        vec3 tm = alb * den;
        tm = make_vec3(fmaxf(tm.x, 0.0f), fmaxf(tm.y, 0.0f), fmaxf(tm.z, 0.0f));
        tm = vclamp(vpow(tm, 0.454545f), 0.0f, 1.0f);
        store_bgra8(im.final_image, pix, make_vec4(tm, 1.0f));
    }
}

// mix3 of bfrBlender.comp:19-27
__device__ __forceinline__ float mix3(float a, float b, float c, float t, float mid, float max_dev) {
    t /= max_dev;
    t = fminf(t, 1.0f);  // a NaN deviation selects t = 1 (IEEE minNum)
    const float fa = fmaxf(1.0f - (t / mid), 0.0f);
    const float fb = (t < mid) ? t / mid : 1.0f - (t - mid) / (1.0f - mid);
    const float fc = 1.0f - fa - fb;
    return fa * a + fb * b + fc * c;
}

__device__ __forceinline__ float glsl_mix(float x, float y, float a) { return x * (1.0f - a) + y * a; }

__global__ void __launch_bounds__(256) k_bfr_blend(int w, int h, int radius, pbrt_bfr_blend_images im) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= w || y >= h) return;
    float sq = 0.0f, av = 0.0f;
    int count = 0;
    for (int dy = -radius; dy <= radius; ++dy)
        for (int dx = -radius; dx <= radius; ++dx) {
            const int sx = x + dx, sy = y + dy;
            vec3 a = make_vec3(0.0f);
            if (sx >= 0 && sy >= 0 && sx < w && sy < h) a = xyz(load_half4(im.average, (size_t)sy * w + sx));
            const float cur = dot(a, make_vec3(1.0f / 3.0f));
            ++count;
            sq = glsl_mix(sq, cur * cur, 1.0f / (float)count);
            av = glsl_mix(av, cur, 1.0f / (float)count);
        }
    const size_t pix = (size_t)y * w + x;
    const vec3 ave = xyz(load_half4(im.average, pix));
    const vec3 ave2 = xyz(load_half4(im.average_squared, pix));
    av = glsl_mix(av, dot(ave, make_vec3(1.0f / 3.0f)), 0.5f);
    sq = glsl_mix(sq, dot(ave2, make_vec3(1.0f / 3.0f)), 0.5f);
    // bfrBlender.comp loads denoised0 (8x8) into den2 and denoised2 (32x32) into den0
    const vec3 den2 = xyz(load_bgra8(im.denoised8, pix));
    const vec3 den1 = xyz(load_bgra8(im.denoised16, pix));
    const vec3 den0 = xyz(load_bgra8(im.denoised32, pix));
    const float sd = sqrtf(sq - av * av);
    const vec3 c = make_vec3(mix3(den0.x, den1.x, den2.x, sd, 0.5f, 1.0f), mix3(den0.y, den1.y, den2.y, sd, 0.5f, 1.0f),
                             mix3(den0.z, den1.z, den2.z, sd, 0.5f, 1.0f));
    store_bgra8(im.final_image, pix, make_vec4(c, 1.0f));
}

}  // namespace pbrt

using namespace pbrt;

extern "C" int32_t pbrt_bfr(const pbrt_bfr_params* p, const pbrt_bfr_images* im, void* stream) {
    return guard([&] {
        PBRT_REQUIRE(p && im, "pbrt_bfr: NULL argument");
        PBRT_REQUIRE(p->block_size == 8 || p->block_size == 16 || p->block_size == 32,
                     "pbrt_bfr: block_size must be 8, 16 or 32");
        PBRT_REQUIRE(p->width >= 2 * p->block_size && p->height >= 2 * p->block_size && p->width >= 16 &&
                         p->height >= 16,
                     "pbrt_bfr: image must be at least two blocks (and 16 pixels) wide and high (mirror addressing)");
        PBRT_REQUIRE(im->depth && im->normal && im->albedo && im->motion && im->spp && im->noisy && im->denoised &&
                         im->final_image,
                     "pbrt_bfr: NULL image");
        BfrGeom g{};
        g.width = (int)p->width;
        g.height = (int)p->height;
        g.bs = (int)p->block_size;
        g.blocks_x = g.width / g.bs + 2;
        g.blocks_y = g.height / g.bs + 2;
        g.offx = kBfrOffsetsHost[p->frame_number % 16][0];
        g.offy = kBfrOffsetsHost[p->frame_number % 16][1];
        g.frame = p->frame_number;
        const dim3 grid(g.blocks_x, g.blocks_y);
        hipStream_t st = (hipStream_t)stream;
        if (g.bs == 8)
            hipLaunchKernelGGL((k_bfr<8, 64>), grid, dim3(64), 0, st, g, *im);
        else if (g.bs == 16)
            hipLaunchKernelGGL((k_bfr<16, 256>), grid, dim3(256), 0, st, g, *im);
        else
            hipLaunchKernelGGL((k_bfr<32, 256>), grid, dim3(256), 0, st, g, *im);
        PBRT_HIP(hipGetLastError());
    });
}

extern "C" int32_t pbrt_bfr_blend(uint32_t width, uint32_t height, uint32_t filter_radius,
                                  const pbrt_bfr_blend_images* im, void* stream) {
    return guard([&] {
        PBRT_REQUIRE(im && im->average && im->average_squared && im->denoised8 && im->denoised16 &&
                         im->denoised32 && im->final_image,
                     "pbrt_bfr_blend: NULL image");
        PBRT_REQUIRE(filter_radius <= 16, "pbrt_bfr_blend: filter_radius above 16");
        if (width == 0 || height == 0) return;
        const dim3 grid((width + 15) / 16, (height + 15) / 16);
        hipLaunchKernelGGL(k_bfr_blend, grid, dim3(256), 0, (hipStream_t)stream, (int)width, (int)height,
                           (int)filter_radius, *im);
        PBRT_HIP(hipGetLastError());
    });
}
