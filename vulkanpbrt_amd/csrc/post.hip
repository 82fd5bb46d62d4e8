// Screen-space passes: temporal accumulation, copy-back, TAA and format
// conversion.  One thread per pixel; 16x16 tiles like the reference's
// dispatches.  All are HBM-bound streaming kernels.
//
// Project semantics where the reference is undefined:
//   * the motion image is RG16F but accumulator.comp declares it r16f
//     (accumulator.comp:14); both components are written, which is what the
//     passes reading .xy (bmfrPost.comp:114, taa.comp:269) expect.
//   * images keep the byte order of their VkFormat: the TAA history copy is a
//     raw B8G8R8A8 -> R8G8B8A8 image copy (Taa.cpp:140-175), so the history is
//     read back with red and blue exchanged, as on the reference.
#include <hip/hip_runtime.h>

#include "capi_common.hpp"
#include "image_ops.hpp"

namespace pbrt {

struct AccumConsts {
    float inv_proj[16];
    float inv_view[16];
    float proj_prev_view[16];  // proj * prevView, proj = inverse(inv_proj)
    float prev_origin[4];
    uint32_t width, height, frame;
};

constexpr int kTile = 16;

// accumulator.comp:33-105 (SEPARATE_MATRICES)
__global__ void __launch_bounds__(kTile* kTile) k_accumulate(AccumConsts c, pbrt_accumulate_images im) {
    const int x = blockIdx.x * kTile + threadIdx.x % kTile;
    const int y = blockIdx.y * kTile + threadIdx.x / kTile;
    if (x >= (int)c.width || y >= (int)c.height) return;
    const size_t pix = (size_t)y * c.width + x;
    const float w = (float)c.width, h = (float)c.height;
    float pixel_spp = 1.0f / 256.0f;
    float depth = im.depth[pix];
    vec4 pos = make_vec4(c.inv_view[12], c.inv_view[13], c.inv_view[14], 1.0f);
    vec2 clip = make_vec2(((float)x + 0.5f) / w * 2.0f - 1.0f, ((float)y + 0.5f) / h * 2.0f - 1.0f);
    vec4 dir = mat4_mul(c.inv_proj, make_vec4(clip.x, clip.y, 1.0f, 1.0f));
    dir = mat4_mul(c.inv_view, make_vec4(normalize(xyz(dir)), 0.0f));
    vec4 p = make_vec4(pos.x + depth * dir.x, pos.y + depth * dir.y, pos.z + depth * dir.z, pos.w + depth * dir.w);
    vec4 prev = mat4_mul(c.proj_prev_view, p);
    float pre_depth = length(xyz(p) - make_vec3(c.prev_origin[0], c.prev_origin[1], c.prev_origin[2]));
    float pw = prev.w;
    prev = make_vec4(prev.x / pw, prev.y / pw, prev.z / pw, prev.w / pw);
    float ux = (prev.x + 1.0f) * 0.5f * (w / (w - 0.5f));
    float uy = (prev.y + 1.0f) * 0.5f * (h / (h - 0.5f));
    bool reprojected = false;
    vec3 prev_color = make_vec3(0.0f);
    if (c.frame > 0) {
        const int iw = (int)c.width, ih = (int)c.height;
        float true_prev_depth = sample_bilinear(
            [&](int sx, int sy) { return make_vec4(im.prev_depth[(size_t)sy * iw + sx], 0, 0, 0); }, iw, ih, ux, uy).x;
        float dissim = true_prev_depth / pre_depth - 1.0f;
        if (ux >= 0.0f && uy >= 0.0f && ux <= 1.0f && uy <= 1.0f && fabsf(dissim) <= 0.01f) {
            reprojected = true;
            prev_color = xyz(sample_bilinear(
                [&](int sx, int sy) { return load_half4(im.prev_accumulated, (size_t)sy * iw + sx); }, iw, ih, ux, uy));
            pixel_spp += sample_bilinear(
                             [&](int sx, int sy) {
                                 return make_vec4(unorm8_to_float(im.prev_spp[(size_t)sy * iw + sx]), 0, 0, 0);
                             },
                             iw, ih, ux, uy)
                             .x;
        }
    }
    ushort2 m;
    m.x = f2h(reprojected ? ux : -1.0f);
    m.y = f2h(reprojected ? uy : -1.0f);
    reinterpret_cast<ushort2*>(im.motion)[pix] = m;
    im.spp[pix] = unorm8_rne(pixel_spp);
    const float4 src = reinterpret_cast<const float4*>(im.illumination)[pix];
    vec3 demod = make_vec3(src.x, src.y, src.z);
    if (reprojected) {
        float a = fmaxf(1.0f / (pixel_spp * 256.0f), 0.1f);
        demod = vmix(prev_color, demod, a);
    }
    store_half4(im.accumulated, pix, make_vec4(demod, 1.0f));
}

__global__ void k_copy_back(uint32_t n, const float* depth, const uint8_t* spp, const uint16_t* acc, float* pdepth,
                            uint8_t* pspp, uint16_t* pacc) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (pdepth) pdepth[i] = depth[i];
    if (pspp) pspp[i] = spp[i];
    if (pacc) reinterpret_cast<uint2*>(pacc)[i] = reinterpret_cast<const uint2*>(acc)[i];
}

__device__ __forceinline__ vec3 rgb_to_ycocg(vec3 c) {
    return make_vec3(dot(c, make_vec3(1.0f, 2.0f, 1.0f)), dot(c, make_vec3(2.0f, 0.0f, -2.0f)),
                     dot(c, make_vec3(-1.0f, 2.0f, -1.0f)));
}

// taa.comp:48-104.  vec3(1/0) folds to the integer 0x7FFFFFFF (2^31 as float).
// The tile's denoised pixels and their one-pixel halo are converted to YCoCg once,
// into LDS, instead of once per neighbourhood they fall in (nine times).
__global__ void __launch_bounds__(kTile* kTile) k_taa(uint32_t width, uint32_t height, uint32_t frame,
                                                     pbrt_taa_images im) {
    constexpr int kHalo = kTile + 2;
    __shared__ float ycc[3][kHalo * kHalo];
    const int x0 = (int)blockIdx.x * kTile - 1, y0 = (int)blockIdx.y * kTile - 1;
    for (int i = threadIdx.x; i < kHalo * kHalo; i += kTile * kTile) {
        const int sx = x0 + i % kHalo, sy = y0 + i / kHalo;
        if (sx >= 0 && sy >= 0 && sx < (int)width && sy < (int)height) {
            const vec3 c = rgb_to_ycocg(xyz(load_bgra8(im.denoised, (size_t)sy * width + sx)));
            ycc[0][i] = c.x;
            ycc[1][i] = c.y;
            ycc[2][i] = c.z;
        }
    }
    __syncthreads();
    const int x = blockIdx.x * kTile + threadIdx.x % kTile;
    const int y = blockIdx.y * kTile + threadIdx.x / kTile;
    if (x >= (int)width || y >= (int)height) return;
    const size_t pix = (size_t)y * width + x;
    vec3 cur = xyz(load_bgra8(im.denoised, pix));
    const ushort2 mv = reinterpret_cast<const ushort2*>(im.motion)[pix];
    float mx = h2f(mv.x), my = h2f(mv.y);
    if (frame == 0 || mx < 0.0f || my < 0.0f || mx > 1.0f || my > 1.0f) {
        store_bgra8(im.final_image, pix, make_vec4(cur, 1.0f));
        return;
    }
    const float big = 2147483648.0f;
    vec3 min_box = make_vec3(big), min_cross = make_vec3(big);
    vec3 max_box = make_vec3(-big), max_cross = make_vec3(-big);
    for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
            int sx = x + dx, sy = y + dy;
            if (sx >= 0 && sy >= 0 && sx < (int)width && sy < (int)height) {
                const int h = (sy - y0) * kHalo + (sx - x0);
                const vec3 sc = make_vec3(ycc[0][h], ycc[1][h], ycc[2][h]);
                if (dx == 0 || dy == 0) {
                    min_cross = vmin(min_cross, sc);
                    max_cross = vmax(max_cross, sc);
                }
                min_box = vmin(min_box, sc);
                max_box = vmax(max_box, sc);
            }
        }
    const int iw = (int)width, ih = (int)height;
    vec3 prev = xyz(sample_bilinear([&](int sx, int sy) { return load_rgba8(im.history, (size_t)sy * iw + sx); }, iw,
                                    ih, mx, my));
    vec3 py = rgb_to_ycocg(prev);
    vec3 mn = (min_box + min_cross) * 0.5f;
    vec3 mxv = (max_box + max_cross) * 0.5f;
    bool inside = py.x >= mn.x && py.y >= mn.y && py.z >= mn.z && py.x <= mxv.x && py.y <= mxv.y && py.z <= mxv.z;
    vec3 res = inside ? 0.4f * cur + (1.0f - 0.4f) * prev : cur;
    store_bgra8(im.final_image, pix, make_vec4(res, 1.0f));
}

__global__ void k_copy_u32(uint32_t n, const uint32_t* src, uint32_t* dst) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

__global__ void __launch_bounds__(256) k_convert(uint32_t n, const float4* src, uint8_t* dst) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 v = src[i];
    store_bgra8(dst, i, make_vec4(v.x, v.y, v.z, v.w));
}

// Host-side fp32 inverse of a column-major 4x4 (cofactor expansion) in double.
static void inverse4(const float* m, float* out) {
    double a[16], inv[16];
    for (int i = 0; i < 16; ++i) a[i] = m[i];
    inv[0] = a[5] * a[10] * a[15] - a[5] * a[11] * a[14] - a[9] * a[6] * a[15] + a[9] * a[7] * a[14] +
             a[13] * a[6] * a[11] - a[13] * a[7] * a[10];
    inv[4] = -a[4] * a[10] * a[15] + a[4] * a[11] * a[14] + a[8] * a[6] * a[15] - a[8] * a[7] * a[14] -
             a[12] * a[6] * a[11] + a[12] * a[7] * a[10];
    inv[8] = a[4] * a[9] * a[15] - a[4] * a[11] * a[13] - a[8] * a[5] * a[15] + a[8] * a[7] * a[13] +
             a[12] * a[5] * a[11] - a[12] * a[7] * a[9];
    inv[12] = -a[4] * a[9] * a[14] + a[4] * a[10] * a[13] + a[8] * a[5] * a[14] - a[8] * a[6] * a[13] -
              a[12] * a[5] * a[10] + a[12] * a[6] * a[9];
    inv[1] = -a[1] * a[10] * a[15] + a[1] * a[11] * a[14] + a[9] * a[2] * a[15] - a[9] * a[3] * a[14] -
             a[13] * a[2] * a[11] + a[13] * a[3] * a[10];
    inv[5] = a[0] * a[10] * a[15] - a[0] * a[11] * a[14] - a[8] * a[2] * a[15] + a[8] * a[3] * a[14] +
             a[12] * a[2] * a[11] - a[12] * a[3] * a[10];
    inv[9] = -a[0] * a[9] * a[15] + a[0] * a[11] * a[13] + a[8] * a[1] * a[15] - a[8] * a[3] * a[13] -
             a[12] * a[1] * a[11] + a[12] * a[3] * a[9];
    inv[13] = a[0] * a[9] * a[14] - a[0] * a[10] * a[13] - a[8] * a[1] * a[14] + a[8] * a[2] * a[13] +
              a[12] * a[1] * a[10] - a[12] * a[2] * a[9];
    inv[2] = a[1] * a[6] * a[15] - a[1] * a[7] * a[14] - a[5] * a[2] * a[15] + a[5] * a[3] * a[14] +
             a[13] * a[2] * a[7] - a[13] * a[3] * a[6];
    inv[6] = -a[0] * a[6] * a[15] + a[0] * a[7] * a[14] + a[4] * a[2] * a[15] - a[4] * a[3] * a[14] -
             a[12] * a[2] * a[7] + a[12] * a[3] * a[6];
    inv[10] = a[0] * a[5] * a[15] - a[0] * a[7] * a[13] - a[4] * a[1] * a[15] + a[4] * a[3] * a[13] +
              a[12] * a[1] * a[7] - a[12] * a[3] * a[5];
    inv[14] = -a[0] * a[5] * a[14] + a[0] * a[6] * a[13] + a[4] * a[1] * a[14] - a[4] * a[2] * a[13] -
              a[12] * a[1] * a[6] + a[12] * a[2] * a[5];
    inv[3] = -a[1] * a[6] * a[11] + a[1] * a[7] * a[10] + a[5] * a[2] * a[11] - a[5] * a[3] * a[10] -
             a[9] * a[2] * a[7] + a[9] * a[3] * a[6];
    inv[7] = a[0] * a[6] * a[11] - a[0] * a[7] * a[10] - a[4] * a[2] * a[11] + a[4] * a[3] * a[10] +
             a[8] * a[2] * a[7] - a[8] * a[3] * a[6];
    inv[11] = -a[0] * a[5] * a[11] + a[0] * a[7] * a[9] + a[4] * a[1] * a[11] - a[4] * a[3] * a[9] -
              a[8] * a[1] * a[7] + a[8] * a[3] * a[5];
    inv[15] = a[0] * a[5] * a[10] - a[0] * a[6] * a[9] - a[4] * a[1] * a[10] + a[4] * a[2] * a[9] +
              a[8] * a[1] * a[6] - a[8] * a[2] * a[5];
    double det = a[0] * inv[0] + a[1] * inv[4] + a[2] * inv[8] + a[3] * inv[12];
    for (int i = 0; i < 16; ++i) out[i] = (float)(inv[i] / det);
}

// column-major C = A * B in fp32 (sum over k in increasing order)
static void matmul4(const float* a, const float* b, float* c) {
    for (int col = 0; col < 4; ++col)
        for (int row = 0; row < 4; ++row) {
            float s = 0.0f;
            for (int k = 0; k < 4; ++k) s += a[4 * k + row] * b[4 * col + k];
            c[4 * col + row] = s;
        }
}

}  // namespace pbrt

using namespace pbrt;

extern "C" int32_t pbrt_accumulate(const pbrt_accumulate_params* p, const pbrt_accumulate_images* im, void* stream) {
    return guard([&] {
        PBRT_REQUIRE(p && im, "pbrt_accumulate: NULL argument");
        PBRT_REQUIRE(p->width > 0 && p->height > 0, "pbrt_accumulate: zero-size image");
        PBRT_REQUIRE(im->illumination && im->depth && im->motion && im->spp && im->accumulated,
                     "pbrt_accumulate: missing image");
        PBRT_REQUIRE(p->frame_number == 0 || (im->prev_depth && im->prev_spp && im->prev_accumulated),
                     "pbrt_accumulate: previous-frame images required after frame 0");
        AccumConsts c{};
        std::memcpy(c.inv_proj, p->inv_proj, sizeof(c.inv_proj));
        std::memcpy(c.inv_view, p->inv_view, sizeof(c.inv_view));
        float proj[16];
        inverse4(p->inv_proj, proj);
        matmul4(proj, p->prev_view, c.proj_prev_view);
        std::memcpy(c.prev_origin, p->prev_origin, sizeof(c.prev_origin));
        c.width = p->width;
        c.height = p->height;
        c.frame = p->frame_number;
        dim3 grid((p->width + kTile - 1) / kTile, (p->height + kTile - 1) / kTile);
        hipLaunchKernelGGL(k_accumulate, grid, dim3(kTile * kTile), 0, (hipStream_t)stream, c, *im);
        PBRT_HIP(hipGetLastError());
    });
}

extern "C" int32_t pbrt_copy_back(uint32_t width, uint32_t height, const float* depth, const uint8_t* spp,
                                  const uint16_t* accumulated, float* prev_depth, uint8_t* prev_spp,
                                  uint16_t* prev_accumulated, void* stream) {
    return guard([&] {
        PBRT_REQUIRE((!prev_depth || depth) && (!prev_spp || spp) && (!prev_accumulated || accumulated),
                     "pbrt_copy_back: source missing for a requested copy");
        uint32_t n = width * height;
        if (n == 0) return;
        hipLaunchKernelGGL(k_copy_back, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, n, depth, spp,
                           accumulated, prev_depth, prev_spp, prev_accumulated);
        PBRT_HIP(hipGetLastError());
    });
}

extern "C" int32_t pbrt_taa(uint32_t width, uint32_t height, uint32_t frame_number, const pbrt_taa_images* im,
                            void* stream) {
    return guard([&] {
        PBRT_REQUIRE(im && im->motion && im->denoised && im->history && im->final_image, "pbrt_taa: missing image");
        PBRT_REQUIRE(width > 0 && height > 0, "pbrt_taa: zero-size image");
        dim3 grid((width + kTile - 1) / kTile, (height + kTile - 1) / kTile);
        hipStream_t st = (hipStream_t)stream;
        hipLaunchKernelGGL(k_taa, grid, dim3(kTile * kTile), 0, st, width, height, frame_number, *im);
        uint32_t n = width * height;
        hipLaunchKernelGGL(k_copy_u32, dim3((n + 255) / 256), dim3(256), 0, st, n,
                           reinterpret_cast<const uint32_t*>(im->final_image),
                           reinterpret_cast<uint32_t*>(im->history));
        PBRT_HIP(hipGetLastError());
    });
}

extern "C" int32_t pbrt_convert(uint32_t width, uint32_t height, const float* source, uint8_t* bgra8, void* stream) {
    return guard([&] {
        PBRT_REQUIRE(source && bgra8, "pbrt_convert: NULL image");
        uint32_t n = width * height;
        if (n == 0) return;
        hipLaunchKernelGGL(k_convert, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, n,
                           reinterpret_cast<const float4*>(source), bgra8);
        PBRT_HIP(hipGetLastError());
    });
}

namespace pbrt {
__global__ void k_running_mean(size_t n4, const float4* src, float4* dst, float inv) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n4) return;
    const float4 s = src[i], d = dst[i];
    dst[i] = make_float4(d.x + (s.x - d.x) * inv, d.y + (s.y - d.y) * inv, d.z + (s.z - d.z) * inv,
                         d.w + (s.w - d.w) * inv);
}
}  // namespace pbrt

extern "C" int32_t pbrt_running_mean(uint32_t width, uint32_t height, const float* src, float* dst, uint32_t n,
                                     void* stream) {
    return guard([&] {
        PBRT_REQUIRE(src && dst, "pbrt_running_mean: NULL image");
        const size_t n4 = (size_t)width * height;
        if (n4 == 0) return;
        hipLaunchKernelGGL(k_running_mean, dim3((unsigned)((n4 + 255) / 256)), dim3(256), 0, (hipStream_t)stream, n4,
                           reinterpret_cast<const float4*>(src), reinterpret_cast<float4*>(dst),
                           1.0f / (float)((double)n + 1.0));
        PBRT_HIP(hipGetLastError());
    });
}
