// Learned kernel-predicting denoiser on gfx950: BF16 matrix-core convolutions,
// forward and backward, and an Adam trainer.
//
// Data layout (per image, row-major pixels):
//   features       NHWC bf16, 16 channels (12 used)
//   activations    NHWC bf16, 32 channels; logits: 32 channels (25 used)
//   master weights fp32, per layer W[tap 9][cout 32][cin_pad] then b[32]
//                  (cin_pad 16 for the first layer, 32 after; padded slots stay 0)
//   forward bf16   same W layout; backward bf16 Wt[tap][cin][cout] with the taps
//                  flipped (the transposed convolution as a forward convolution)
//
// k_conv3x3       implicit GEMM per 16x16 pixel tile: the 18x18 input halo is
//                 staged in LDS; each wave owns four 16-pixel rows and computes
//                 D[cout][pixel] = sum_tap W_tap^T X_tap on v_mfma_f32_16x16x32_bf16
//                 (16x16x16 for the 16-channel input layer), so each lane ends
//                 with four consecutive output channels of one pixel.  The same
//                 kernel computes the input gradient from the flipped weights,
//                 masking the upstream gradient by the ReLU of the activation.
// k_conv3x3_wgrad per tile: dZ^T (channel-major) and three x-shifted,
//                 channel-major copies of the input halo in LDS; every tap's
//                 32 x cin weight gradient is a GEMM over the tile's 256 pixels;
//                 per-tile partials are summed in a fixed order by k_reduce.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "capi_common.hpp"
#include "image_ops.hpp"

namespace pbrt {
namespace {

constexpr int kMaxLayers = 6;  // the config-4 network; a network has 2..6 layers
constexpr int kTaps = 9;
constexpr int kCh = 32;
constexpr int kFeatCh = 16;
constexpr int kFeatUsed = 12;
constexpr int kKernelTaps = 25;
constexpr int kTile = 16;
constexpr float kRelEps = 0.01f;

constexpr int cin_pad(int l) { return l == 0 ? kFeatCh : kCh; }
constexpr int cin_real(int l) { return l == 0 ? kFeatUsed : kCh; }
constexpr int cout_real(int l, int layers) { return l == layers - 1 ? kKernelTaps : kCh; }
constexpr int layer_weights(int l) { return kTaps * kCh * cin_pad(l); }
constexpr int layer_size(int l) { return layer_weights(l) + kCh; }
constexpr int layer_offset(int l) { return l == 0 ? 0 : layer_offset(l - 1) + layer_size(l - 1); }
constexpr int master_size(int layers) { return layer_offset(layers); }
constexpr int weight_offset(int l) { return l == 0 ? 0 : weight_offset(l - 1) + layer_weights(l - 1); }
constexpr int weight_size(int layers) { return weight_offset(layers); }

int canonical_count(int layers) {
    int n = 0;
    for (int l = 0; l < layers; ++l) n += cout_real(l, layers) * cin_real(l) * kTaps + cout_real(l, layers);
    return n;
}

// canonical index -> padded master index
std::vector<int> canonical_map(int layers) {
    std::vector<int> map;
    for (int l = 0; l < layers; ++l) {
        for (int co = 0; co < cout_real(l, layers); ++co)
            for (int ci = 0; ci < cin_real(l); ++ci)
                for (int t = 0; t < kTaps; ++t) map.push_back(layer_offset(l) + (t * kCh + co) * cin_pad(l) + ci);
        for (int co = 0; co < cout_real(l, layers); ++co) map.push_back(layer_offset(l) + layer_weights(l) + co);
    }
    return map;
}

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short short4v __attribute__((ext_vector_type(4)));
typedef float float4v __attribute__((ext_vector_type(4)));
typedef _Float16 half2v __attribute__((ext_vector_type(2)));
typedef float float2v __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2v __attribute__((ext_vector_type(2)));

__device__ __forceinline__ float bf2f(uint32_t h) { return __uint_as_float(h << 16); }
// round to nearest even (v_cvt_pk_bf16_f32)
__device__ __forceinline__ uint32_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
// two floats to packed bf16 (lo = a), rounded to nearest even in one v_cvt_pk_bf16_f32
__device__ __forceinline__ uint32_t pack_bf16(float a, float b) {
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((float2v){a, b}, bf16x2v));
}
// min(x^(1/2.2), 1) for x >= 0 as exp2(log2(x) / 2.2) on v_log_f32 / v_exp_f32:
// the result is rounded to 8 bits, far above their error
__device__ __forceinline__ float fast_gamma(float x) {
    return fminf(__builtin_amdgcn_exp2f(0.454545f * __builtin_amdgcn_logf(x)), 1.0f);
}

__device__ __forceinline__ bool bf_positive(uint32_t h) { return (h & 0x8000u) == 0 && (h & 0x7FFFu) != 0; }

// zero the 16-bit lanes of v whose mask value is not > 0
__device__ __forceinline__ uint32_t relu_mask2(uint32_t v, uint32_t m) {
    uint32_t lo = bf_positive(m & 0xFFFFu) ? (v & 0xFFFFu) : 0u;
    uint32_t hi = bf_positive(m >> 16) ? (v & 0xFFFF0000u) : 0u;
    return lo | hi;
}
__device__ __forceinline__ uint4 relu_mask8(uint4 v, uint4 m) {
    return make_uint4(relu_mask2(v.x, m.x), relu_mask2(v.y, m.y), relu_mask2(v.z, m.z), relu_mask2(v.w, m.w));
}

struct CnnGeom {
    int width, height, count;
};

// ---------------------------------------------------------------------------
// Features
// ---------------------------------------------------------------------------
// The G-buffer inputs of one pixel's features, as loaded.
struct RawPixel {
    uint2 noisy;      // RGBA16F
    uint32_t albedo;  // RGBA8
    float2 normal;    // spherical
    float depth;
    uint32_t spp;
    uint32_t motion_x;  // FP16 bits
};

__device__ __forceinline__ RawPixel load_raw_pixel(const pbrt_cnn_images& im, size_t i) {
    RawPixel r;
    r.noisy = reinterpret_cast<const uint2*>(im.noisy)[i];
    r.albedo = reinterpret_cast<const uint32_t*>(im.albedo)[i];
    r.normal = reinterpret_cast<const float2*>(im.normal)[i];
    r.depth = im.depth[i];
    r.spp = im.spp[i];
    r.motion_x = reinterpret_cast<const ushort2*>(im.motion)[i].x;
    return r;
}

// the 16 bf16 feature channels of a pixel (12 used), packed in two 16-byte words
__device__ __forceinline__ void raw_features(const RawPixel& r, uint4& lo, uint4& hi) {
    const half2v rg = __builtin_bit_cast(half2v, r.noisy.x), ba = __builtin_bit_cast(half2v, r.noisy.y);
    const vec3 nz = make_vec3((float)rg.x, (float)rg.y, (float)ba.x);
    const uchar4 ac = __builtin_bit_cast(uchar4, r.albedo);
    const float k255 = 1.0f / 255.0f;
    const vec3 alb = make_vec3((float)ac.x * k255, (float)ac.y * k255, (float)ac.z * k255);
    const float2 sn = r.normal;
    // fast sine / cosine / log / reciprocal: the features are rounded to bf16 (8
    // significand bits)
    float st, ct, sp, cp;
    __sincosf(sn.x, &st, &ct);
    __sincosf(sn.y, &sp, &cp);
    const vec3 nn = make_vec3(cp * st, sp * st, ct);
    const float d = r.depth;
    const float spp = (float)r.spp * (256.0f / 255.0f);
    const float mvx = h2f((uint16_t)r.motion_x);
    // log(1 + x) as ln2 * v_log_f32 (its argument is >= 1: no denormal scaling) and
    // d / (1 + d) on v_rcp_f32 (1 ulp)
    const float ln2 = 0.69314718f;
    float f[kFeatCh] = {ln2 * __builtin_amdgcn_logf(1.0f + fmaxf(nz.x, 0.0f)),
                        ln2 * __builtin_amdgcn_logf(1.0f + fmaxf(nz.y, 0.0f)),
                        ln2 * __builtin_amdgcn_logf(1.0f + fmaxf(nz.z, 0.0f)),
                        alb.x, alb.y, alb.z, nn.x, nn.y, nn.z, d * __builtin_amdgcn_rcpf(1.0f + d),
                        fminf(spp, 16.0f) * 0.0625f,
                        mvx >= 0.0f ? 1.0f : 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    uint32_t w[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) w[k] = pack_bf16(f[2 * k], f[2 * k + 1]);
    lo = make_uint4(w[0], w[1], w[2], w[3]);
    hi = make_uint4(w[4], w[5], w[6], w[7]);
}

__device__ __forceinline__ void pixel_features(const pbrt_cnn_images& im, size_t i, uint4& lo, uint4& hi) {
    raw_features(load_raw_pixel(im, i), lo, hi);
}

__global__ void __launch_bounds__(256) k_features(CnnGeom g, pbrt_cnn_images im, uint16_t* feat) {
    const size_t n = (size_t)g.width * g.height * g.count;
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    uint4* o = reinterpret_cast<uint4*>(feat + i * kFeatCh);
    pixel_features(im, i, o[0], o[1]);
}

// ---------------------------------------------------------------------------
// 3x3 convolution (forward, or input gradient with flipped weights)
// ---------------------------------------------------------------------------
// kConvKpn: last layer of inference; the logits stay on chip (LDS) and the softmax
// 5x5 kernel is applied to the noisy taps staged with the tile.
enum ConvMode { kConvRelu = 0, kConvLinear = 1, kConvDgrad = 2, kConvKpn = 4 };

// Persistent: each workgroup keeps the layer's weights in registers and walks a
// strided sequence of 16x16 tiles inside its XCD's contiguous share of the tile
// range (neighbouring tiles, whose halos overlap, stay in one L2); the next tile's
// halo is fetched into registers while the current one is computed and is stored
// into the other LDS buffer afterwards.
template <int CIN, int MODE>
__global__ void __launch_bounds__(256) k_conv3x3(CnnGeom g, const uint16_t* __restrict__ in,
                                                 const uint16_t* __restrict__ mask, const uint16_t* __restrict__ wgt,
                                                 const float* __restrict__ bias, uint16_t* __restrict__ out,
                                                 pbrt_cnn_images im) {
    constexpr int kStride = CIN + 8;  // bf16 per LDS pixel: conflict-free 16 / 8 byte lane reads
    constexpr int kChunks = CIN / 8;
    constexpr int kTilePix = 18 * 18;
    constexpr int kLoads = (kTilePix * kChunks + 255) / 256;
    constexpr int kNoisyPix = 20 * 20;  // kernel-prediction taps of a tile (kConvKpn)
    constexpr int kNoisyLoads = MODE == kConvKpn ? (kNoisyPix + 255) / 256 : 1;
    __shared__ __attribute__((aligned(16))) uint16_t tile[2][kTilePix * kStride];
    __shared__ __attribute__((aligned(16))) uint2 ntile[MODE == kConvKpn ? 2 : 1][MODE == kConvKpn ? kNoisyPix : 1];
    __shared__ __attribute__((aligned(16))) uint16_t logit[MODE == kConvKpn ? kTile * kTile * kCh : 8];
    uint2 nregs[kNoisyLoads];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int j = lane & 15, grp = lane >> 4;
    const int tiles_x = (g.width + kTile - 1) / kTile, tiles_y = (g.height + kTile - 1) / kTile;
    const int per_image = tiles_x * tiles_y, ntiles = per_image * g.count;
    const int slots = gridDim.x / 8, xcd = blockIdx.x % 8, slot = blockIdx.x / 8;
    const int chunk = (ntiles + 7) / 8, begin = xcd * chunk, end = min(ntiles, begin + chunk);
    int t = begin + slot;
    if (t >= end) return;
    const size_t plane = (size_t)g.width * g.height;

    auto fetch = [&](int tt, uint4 (&r)[kLoads]) {
        const int img = tt / per_image, rem = tt % per_image;
        const int x0 = (rem % tiles_x) * kTile, y0 = (rem / tiles_x) * kTile;
        if (MODE == kConvKpn) {
#pragma unroll
            for (int k = 0; k < kNoisyLoads; ++k) {
                const int p = tid + k * 256;
                if (p < kNoisyPix) {
                    const int yy = min(max(y0 + p / 20 - 2, 0), g.height - 1);
                    const int xx = min(max(x0 + p % 20 - 2, 0), g.width - 1);
                    nregs[k] = reinterpret_cast<const uint2*>(im.noisy)[plane * img + (size_t)yy * g.width + xx];
                }
            }
        }
        const uint16_t* src = in + plane * img * CIN;
        const uint16_t* msk = mask ? mask + plane * img * CIN : nullptr;
#pragma unroll
        for (int k = 0; k < kLoads; ++k) {
            const int e = tid + k * 256;
            r[k] = make_uint4(0, 0, 0, 0);
            if (e < kTilePix * kChunks) {
                const int p = e / kChunks, c = e % kChunks;
                const int gy = y0 + p / 18 - 1, gx = x0 + p % 18 - 1;
                if (gx >= 0 && gy >= 0 && gx < g.width && gy < g.height) {
                    const size_t off = ((size_t)gy * g.width + gx) * CIN + c * 8;
                    r[k] = *reinterpret_cast<const uint4*>(src + off);
                    if (MODE == kConvDgrad && msk) r[k] = relu_mask8(r[k], *reinterpret_cast<const uint4*>(msk + off));
                }
            }
        }
    };
    auto stash = [&](int b, const uint4 (&r)[kLoads]) {
        if (MODE == kConvKpn) {
#pragma unroll
            for (int k = 0; k < kNoisyLoads; ++k) {
                const int p = tid + k * 256;
                if (p < kNoisyPix) ntile[b][p] = nregs[k];
            }
        }
#pragma unroll
        for (int k = 0; k < kLoads; ++k) {
            const int e = tid + k * 256;
            if (e < kTilePix * kChunks)
                *reinterpret_cast<uint4*>(&tile[b][(e / kChunks) * kStride + (e % kChunks) * 8]) = r[k];
        }
    };

    // weights (A operand, lane holds output channel ct*16 + j), resident for every tile
    constexpr int kAVec = CIN == 32 ? 8 : 4;  // bf16 per lane per MFMA
    uint32_t wreg[kTaps][2][kAVec / 2];
#pragma unroll
    for (int tp = 0; tp < kTaps; ++tp)
#pragma unroll
        for (int ct = 0; ct < 2; ++ct) {
            const uint16_t* w = wgt + (tp * kCh + ct * 16 + j) * CIN + kAVec * grp;
            if (CIN == 32) {
                const uint4 v = *reinterpret_cast<const uint4*>(w);
                wreg[tp][ct][0] = v.x;
                wreg[tp][ct][1] = v.y;
                wreg[tp][ct][2 % (kAVec / 2)] = v.z;
                wreg[tp][ct][3 % (kAVec / 2)] = v.w;
            } else {
                const uint2 v = *reinterpret_cast<const uint2*>(w);
                wreg[tp][ct][0] = v.x;
                wreg[tp][ct][1] = v.y;
            }
        }
    float bv[2][4];
#pragma unroll
    for (int ct = 0; ct < 2; ++ct)
#pragma unroll
        for (int i = 0; i < 4; ++i) bv[ct][i] = MODE != kConvDgrad ? bias[ct * 16 + 4 * grp + i] : 0.0f;

    uint4 regs[kLoads];
    fetch(t, regs);
    stash(0, regs);
    __syncthreads();
    int buf = 0;
    for (; t < end; t += slots) {
        const int tn = t + slots;
        if (tn < end) fetch(tn, regs);
        float4v acc[4][2];
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[r][0] = acc[r][1] = (float4v){0.0f, 0.0f, 0.0f, 0.0f};
        const uint16_t* tb = tile[buf];
#pragma unroll
        for (int tp = 0; tp < kTaps; ++tp) {
            const int dy = tp / 3, dx = tp % 3;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int ty = wave * 4 + r;
                const uint16_t* bp = &tb[((ty + dy) * 18 + j + dx) * kStride + kAVec * grp];
                if (CIN == 32) {
                    const bf16x8 b = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(bp));
#pragma unroll
                    for (int ct = 0; ct < 2; ++ct) {
                        const bf16x8 a = __builtin_bit_cast(
                            bf16x8, make_uint4(wreg[tp][ct][0], wreg[tp][ct][1], wreg[tp][ct][2 % (kAVec / 2)],
                                               wreg[tp][ct][3 % (kAVec / 2)]));
                        acc[r][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[r][ct], 0, 0, 0);
                    }
                } else {
                    const short4v b = __builtin_bit_cast(short4v, *reinterpret_cast<const uint2*>(bp));
#pragma unroll
                    for (int ct = 0; ct < 2; ++ct) {
                        const short4v a = __builtin_bit_cast(short4v, make_uint2(wreg[tp][ct][0], wreg[tp][ct][1]));
                        acc[r][ct] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[r][ct], 0, 0, 0);
                    }
                }
            }
        }
        // D[cout][pixel]: lane holds pixel j, output channels ct*16 + 4*grp + i
        const int img = t / per_image, rem = t % per_image;
        const int x0 = (rem % tiles_x) * kTile, y0 = (rem / tiles_x) * kTile;
        const int gx = x0 + j;
        if (MODE == kConvKpn) {
            // bf16 logits of the tile to LDS, then one thread per pixel: softmax over the
            // 25 logits and the 5x5 weighted sum of the staged noisy taps
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int p = (wave * 4 + r) * kTile + j;
#pragma unroll
                for (int ct = 0; ct < 2; ++ct) {
                    float v[4];
#pragma unroll
                    for (int i = 0; i < 4; ++i) v[i] = acc[r][ct][i] + bv[ct][i];
                    *reinterpret_cast<uint2*>(&logit[p * kCh + ct * 16 + 4 * grp]) =
                        make_uint2(pack_bf16(v[0], v[1]), pack_bf16(v[2], v[3]));
                }
            }
            __syncthreads();
            const int py = tid / kTile, px = tid % kTile;
            const int gy = y0 + py, gxp = x0 + px;
            if (gxp < g.width && gy < g.height) {
                float z[kKernelTaps];
                float zmax = -3.4e38f;
#pragma unroll
                for (int c = 0; c < kKernelTaps; ++c) {
                    z[c] = bf2f(logit[tid * kCh + c]);
                    zmax = fmaxf(zmax, z[c]);
                }
                float sum = 0.0f;
#pragma unroll
                for (int c = 0; c < kKernelTaps; ++c) {
                    z[c] = expf(z[c] - zmax);
                    sum += z[c];
                }
                const float inv = 1.0f / sum;
                vec3 o = make_vec3(0.0f);
#pragma unroll
                for (int c = 0; c < kKernelTaps; ++c) {
                    // staged taps: tile-local row/col of the clamped image position
                    const int yy = min(max(gy + c / 5 - 2, 0), g.height - 1) - (y0 - 2);
                    const int xx = min(max(gxp + c % 5 - 2, 0), g.width - 1) - (x0 - 2);
                    const uint2 hv = ntile[buf][yy * 20 + xx];
                    o += make_vec3(h2f(hv.x & 0xFFFFu), h2f(hv.x >> 16), h2f(hv.y & 0xFFFFu)) * (z[c] * inv);
                }
                const size_t i = plane * img + (size_t)gy * g.width + gxp;
                if (im.denoised) reinterpret_cast<float4*>(im.denoised)[i] = make_float4(o.x, o.y, o.z, 1.0f);
                if (im.final_image) {
                    const vec3 alb = xyz(load_rgba8(im.albedo, i)) + make_vec3(1e-6f);
                    vec3 tm = alb * o;
                    tm = make_vec3(fmaxf(tm.x, 0.0f), fmaxf(tm.y, 0.0f), fmaxf(tm.z, 0.0f));
                    store_bgra8(im.final_image, i,
                                make_vec4(fast_gamma(tm.x), fast_gamma(tm.y), fast_gamma(tm.z), 1.0f));
                }
            }
        } else {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int gy = y0 + wave * 4 + r;
            if (gx >= g.width || gy >= g.height) continue;
            uint16_t* o = out + (plane * img + (size_t)gy * g.width + gx) * kCh;
#pragma unroll
            for (int ct = 0; ct < 2; ++ct) {
                float v[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    v[i] = acc[r][ct][i] + bv[ct][i];
                    if (MODE == kConvRelu) v[i] = fmaxf(v[i], 0.0f);
                }
                *reinterpret_cast<uint2*>(o + ct * 16 + 4 * grp) =
                    make_uint2(pack_bf16(v[0], v[1]), pack_bf16(v[2], v[3]));
            }
        }
        }
        if (tn < end) stash(buf ^ 1, regs);
        __syncthreads();
        buf ^= 1;
    }
}

// ---------------------------------------------------------------------------
// Weight gradient: per-tile partial dW[tap][cout][cin] and db[cout]
// ---------------------------------------------------------------------------
template <int CIN>
__global__ void __launch_bounds__(256) k_conv3x3_wgrad(CnnGeom g, const uint16_t* __restrict__ dsrc,
                                                       const uint16_t* __restrict__ mask,
                                                       const uint16_t* __restrict__ in, float* __restrict__ partial) {
    constexpr int kPart = kTaps * kCh * CIN + kCh;
    __shared__ __attribute__((aligned(16))) uint16_t dz[kCh][256];              // [cout][pixel]
    __shared__ __attribute__((aligned(16))) uint16_t xs[3][CIN][18][kTile];     // [dx][cin][row][col]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int x0 = blockIdx.x * kTile, y0 = blockIdx.y * kTile, img = blockIdx.z;
    const size_t plane = (size_t)g.width * g.height;
    {
        // upstream gradient of this tile, masked by the ReLU, transposed to channel-major
        const int p = tid;  // one pixel per thread
        const int gy = y0 + p / kTile, gx = x0 + p % kTile;
        const bool inside = gx < g.width && gy < g.height;
        const size_t off = (plane * img + (size_t)gy * g.width + gx) * kCh;
#pragma unroll
        for (int c = 0; c < kCh / 8; ++c) {
            uint4 v = make_uint4(0, 0, 0, 0);
            if (inside) {
                v = *reinterpret_cast<const uint4*>(dsrc + off + c * 8);
                if (mask) v = relu_mask8(v, *reinterpret_cast<const uint4*>(mask + off + c * 8));
            }
            const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                dz[c * 8 + 2 * k][p] = (uint16_t)(w[k] & 0xFFFFu);
                dz[c * 8 + 2 * k + 1][p] = (uint16_t)(w[k] >> 16);
            }
        }
    }
    for (int e = tid; e < 3 * 18 * kTile; e += 256) {
        const int dx = e / (18 * kTile), rem = e % (18 * kTile);
        const int row = rem / kTile, col = rem % kTile;
        const int gy = y0 + row - 1, gx = x0 + col + dx - 1;
        const bool inside = gx >= 0 && gy >= 0 && gx < g.width && gy < g.height;
        const uint16_t* s = in + (plane * img + (size_t)gy * g.width + gx) * CIN;
#pragma unroll
        for (int c = 0; c < CIN / 8; ++c) {
            uint4 v = make_uint4(0, 0, 0, 0);
            if (inside) v = *reinterpret_cast<const uint4*>(s + c * 8);
            const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                xs[dx][c * 8 + 2 * k][row][col] = (uint16_t)(w[k] & 0xFFFFu);
                xs[dx][c * 8 + 2 * k + 1][row][col] = (uint16_t)(w[k] >> 16);
            }
        }
    }
    __syncthreads();
    float* part = partial + (((size_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x) * kPart;
    const int j = lane & 15, grp = lane >> 4;
    constexpr int kCiTiles = CIN / 16;
    constexpr int kCombos = kTaps * 2 * kCiTiles;
    for (int cb = wave; cb < kCombos; cb += 4) {
        const int t = cb / (2 * kCiTiles), rest = cb % (2 * kCiTiles);
        const int ct = rest / kCiTiles, it = rest % kCiTiles;
        const int dy = t / 3, dx = t % 3;
        float4v acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int ks = 0; ks < 8; ++ks) {
            const int p0 = ks * 32 + 8 * grp;  // eight pixels of one row
            const int row = p0 / kTile, col = p0 % kTile;
            const bf16x8 a = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(&dz[ct * 16 + j][p0]));
            const bf16x8 b =
                __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(&xs[dx][it * 16 + j][row + dy][col]));
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
        }
        // D[cout][cin]: lane holds cin it*16 + j, couts ct*16 + 4*grp + i
#pragma unroll
        for (int i = 0; i < 4; ++i) part[(t * kCh + ct * 16 + 4 * grp + i) * CIN + it * 16 + j] = acc[i];
    }
    if (tid < kCh) {
        float s = 0.0f;
        for (int p = 0; p < 256; ++p) s += bf2f(dz[tid][p]);
        part[kTaps * kCh * CIN + tid] = s;
    }
}

// grad[i] = sum over tiles of partial[tile][i], in tile order
__global__ void __launch_bounds__(256) k_reduce_partials(const float* partial, int tiles, int size, float* grad) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= size) return;
    float s = 0.0f;
    for (int t = 0; t < tiles; ++t) s += partial[(size_t)t * size + i];
    grad[i] = s;
}

// ---------------------------------------------------------------------------
// Kernel prediction: softmax over 25 logits, 5x5 weighted sum of the noisy
// illumination (taps clamped to the image); loss and its logit gradient.
// ---------------------------------------------------------------------------
template <bool kTrain>
__global__ void __launch_bounds__(256) k_kpn(CnnGeom g, pbrt_cnn_images im, const uint16_t* logits, uint16_t* dlogits,
                                            float* loss_part) {
    __shared__ float red[4];
    const size_t plane = (size_t)g.width * g.height;
    const size_t n = plane * g.count;
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    float lsum = 0.0f;
    if (i < n) {
        const int img = (int)(i / plane);
        const int rem = (int)(i % plane);
        const int y = rem / g.width, x = rem % g.width;
        const uint16_t* lz = logits + i * kCh;
        float z[kKernelTaps];
        float zmax = -3.4e38f;
#pragma unroll
        for (int k = 0; k < kKernelTaps; ++k) {
            z[k] = bf2f(lz[k]);
            zmax = fmaxf(zmax, z[k]);
        }
        float s = 0.0f;
#pragma unroll
        for (int k = 0; k < kKernelTaps; ++k) {
            z[k] = expf(z[k] - zmax);
            s += z[k];
        }
        const float inv = 1.0f / s;
        vec3 tap[kKernelTaps];
        vec3 o = make_vec3(0.0f);
#pragma unroll
        for (int k = 0; k < kKernelTaps; ++k) {
            z[k] *= inv;
            const int yy = min(max(y + k / 5 - 2, 0), g.height - 1), xx = min(max(x + k % 5 - 2, 0), g.width - 1);
            tap[k] = xyz(load_half4(im.noisy, plane * img + (size_t)yy * g.width + xx));
            o += tap[k] * z[k];
        }
        if (!kTrain) {
            if (im.denoised) reinterpret_cast<float4*>(im.denoised)[i] = make_float4(o.x, o.y, o.z, 1.0f);
            if (im.final_image) {
                const vec3 alb = xyz(load_rgba8(im.albedo, i)) + make_vec3(1e-6f);
                vec3 t = alb * o;
                t = make_vec3(fmaxf(t.x, 0.0f), fmaxf(t.y, 0.0f), fmaxf(t.z, 0.0f));
                store_bgra8(im.final_image, i, make_vec4(vclamp(vpow(t, 0.454545f), 0.0f, 1.0f), 1.0f));
            }
        } else {
            // relative squared error of the remodulated radiance (albedo + 1e-6 as in bmfrPost)
            const float4 t4 = reinterpret_cast<const float4*>(im.target)[i];
            const vec3 alb = xyz(load_rgba8(im.albedo, i)) + make_vec3(1e-6f);
            const float t[3] = {t4.x, t4.y, t4.z}, ov[3] = {o.x, o.y, o.z}, a[3] = {alb.x, alb.y, alb.z};
            const float norm = 1.0f / (float)(n * 3);
            float gch[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float at = a[c] * t[c];
                const float d = a[c] * ov[c] - at, den = at * at + kRelEps;
                lsum += d * d / den;
                gch[c] = 2.0f * a[c] * d / den * norm;
            }
            float dw[kKernelTaps], wdw = 0.0f;
#pragma unroll
            for (int k = 0; k < kKernelTaps; ++k) {
                dw[k] = gch[0] * tap[k].x + gch[1] * tap[k].y + gch[2] * tap[k].z;
                wdw += z[k] * dw[k];
            }
            uint32_t packed[kCh / 2];
#pragma unroll
            for (int k = 0; k < kCh / 2; ++k) {
                const float a = 2 * k < kKernelTaps ? z[2 * k] * (dw[2 * k] - wdw) : 0.0f;
                const float b = 2 * k + 1 < kKernelTaps ? z[2 * k + 1] * (dw[2 * k + 1] - wdw) : 0.0f;
                packed[k] = pack_bf16(a, b);
            }
            uint4* d4 = reinterpret_cast<uint4*>(dlogits + i * kCh);
#pragma unroll
            for (int q = 0; q < 4; ++q) d4[q] = make_uint4(packed[4 * q], packed[4 * q + 1], packed[4 * q + 2], packed[4 * q + 3]);
        }
    }
    if (kTrain) {
        for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = lsum;
        __syncthreads();
        if (threadIdx.x == 0) loss_part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
    }
}

__global__ void k_loss_total(const float* part, int blocks, float scale, float* loss) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    float s = 0.0f;
    for (int b = 0; b < blocks; ++b) s += part[b];
    *loss = s * scale;
}

// ---------------------------------------------------------------------------
// Parameters: bf16 refresh, Adam, canonical gather
// ---------------------------------------------------------------------------
struct LayerTable {
    int offset[kMaxLayers + 1];
    int woffset[kMaxLayers + 1];
    int cin[kMaxLayers];
    int layers;
};

__device__ __forceinline__ int layer_of(const LayerTable& lt, int i) {
    int l = 0;
    while (l + 1 < lt.layers && i >= lt.offset[l + 1]) ++l;
    return l;
}

// writes forward bf16 W and the flipped / transposed backward copy for master slot i
__device__ __forceinline__ void refresh_weight(const LayerTable& lt, int i, float p, uint16_t* wf, uint16_t* wb) {
    const int l = layer_of(lt, i);
    const int local = i - lt.offset[l];
    const int ci_pad = lt.cin[l];
    if (local >= kTaps * kCh * ci_pad) return;  // bias
    const int t = local / (kCh * ci_pad), co = (local / ci_pad) % kCh, ci = local % ci_pad;
    const uint16_t h = (uint16_t)f2bf(p);
    wf[lt.woffset[l] + local] = h;
    if (l > 0) wb[lt.woffset[l] + ((kTaps - 1 - t) * ci_pad + ci) * kCh + co] = h;
}

__global__ void __launch_bounds__(256) k_refresh(LayerTable lt, const float* master, uint16_t* wf, uint16_t* wb) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < lt.offset[lt.layers]) refresh_weight(lt, i, master[i], wf, wb);
}

__global__ void __launch_bounds__(256) k_adam(LayerTable lt, float* master, float* m, float* v, const float* grad,
                                             float lr, float c1, float c2, uint16_t* wf, uint16_t* wb) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= lt.offset[lt.layers]) return;
    const float gi = grad[i];
    const float mi = 0.9f * m[i] + 0.1f * gi;
    const float vi = 0.999f * v[i] + 0.001f * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float p = master[i] - lr * (mi * c1) / (sqrtf(vi * c2) + 1e-8f);
    master[i] = p;
    refresh_weight(lt, i, p, wf, wb);
}

__global__ void __launch_bounds__(256) k_gather(const float* src, const int* map, int n, float* dst) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) dst[i] = src[map[i]];
}

LayerTable layer_table(int layers) {
    LayerTable lt{};
    for (int l = 0; l <= layers; ++l) {
        lt.offset[l] = layer_offset(l);
        lt.woffset[l] = weight_offset(l);
    }
    for (int l = 0; l < layers; ++l) lt.cin[l] = cin_pad(l);
    lt.layers = layers;
    return lt;
}

// Softmax over one output pixel's 25 bf16 logits (LDS, `logits`: 32 channels, 25
// used) and the 5x5 weighted sum of its noisy taps; thread `tid` of the 16x16 tile at
// (x0, y0).  `taps` is the tile's noisy halo (raw FP16 RGBA, clamped to the image)
// with row length `halo`, offset so that tap (ky, kx) of tile pixel (py, px) sits at
// (py + ky, px + kx).  Writes the denoised radiance and / or the final image.
__device__ __forceinline__ void kpn_apply_pixel(const CnnGeom& g, const pbrt_cnn_images& im, const uint16_t* logits,
                                                const uint2* ntile, int halo, size_t image_base, int x0, int y0,
                                                int tid, const uint32_t* albedo_lds = nullptr) {
    const int py = tid / kTile, px = tid % kTile;
    const int gy = y0 + py, gx = x0 + px;
    if (gx >= g.width || gy >= g.height) return;
    // the 25 bf16 logits (and 7 padding channels): four 16-byte reads, twice
    // (maximum, then the weighted sum) to keep few registers live
    const uint4* lz = reinterpret_cast<const uint4*>(logits);
    float zmax = -3.4e38f;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const uint4 v = lz[q];
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int c = 8 * q + k;
            const float zc = __uint_as_float((k & 1) ? (w[k / 2] & 0xFFFF0000u) : (w[k / 2] << 16));
            if (c < kKernelTaps) zmax = fmaxf(zmax, zc);
        }
    }
    // exp(z - max) = exp2(z log2(e) - max log2(e)); the normalisation is applied once
    const float l2e = 1.4426950408889634f, zm = zmax * l2e;
    // the halo holds the clamped noisy pixel at every position, so tap (ky, kx) of
    // this pixel sits at (py + ky, px + kx) also on border tiles: the 25 reads are
    // constant offsets from one address
    const uint2* taps = &ntile[py * halo + px];
    float sum = 0.0f, ox = 0.0f, oy = 0.0f, oz = 0.0f;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const uint4 v = lz[q];
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int c = 8 * q + k;
            if (c < kKernelTaps) {
                const float zc = __uint_as_float((k & 1) ? (w[k / 2] & 0xFFFF0000u) : (w[k / 2] << 16));
                const float e = __builtin_amdgcn_exp2f(fmaf(zc, l2e, -zm));
                // FP16 taps widen exactly inside the multiply-add (v_fma_mix_f32)
                const uint2 nv = taps[(c / 5) * halo + c % 5];
                const half2v rg = __builtin_bit_cast(half2v, nv.x), ba = __builtin_bit_cast(half2v, nv.y);
                sum += e;
                ox = fmaf((float)rg.x, e, ox);
                oy = fmaf((float)rg.y, e, oy);
                oz = fmaf((float)ba.x, e, oz);
            }
        }
    }
    const float inv = __builtin_amdgcn_rcpf(sum);  // 1 ulp
    ox *= inv;
    oy *= inv;
    oz *= inv;
    const size_t i = image_base + (size_t)gy * g.width + gx;
    if (im.denoised) reinterpret_cast<float4*>(im.denoised)[i] = make_float4(ox, oy, oz, 1.0f);
    if (im.final_image) {
        // remodulated, gamma 1/2.2 (the albedo from LDS when the caller staged it)
        const uchar4 ac = albedo_lds ? __builtin_bit_cast(uchar4, *albedo_lds)
                                     : reinterpret_cast<const uchar4*>(im.albedo)[i];
        const float k = 1.0f / 255.0f;
        const float tx = fmaxf(((float)ac.x * k + 1e-6f) * ox, 0.0f);
        const float ty = fmaxf(((float)ac.y * k + 1e-6f) * oy, 0.0f);
        const float tz = fmaxf(((float)ac.z * k + 1e-6f) * oz, 0.0f);
        store_bgra8(im.final_image, i, make_vec4(fast_gamma(tx), fast_gamma(ty), fast_gamma(tz), 1.0f));
    }
}

// ---------------------------------------------------------------------------
// Two-layer network, fused: features -> 3x3 conv 16->32 (ReLU) -> 3x3 conv
// 32->25 logits -> softmax 5x5 kernel -> filtered radiance, one workgroup per
// 16x16 output tile with every intermediate in LDS (inference only).
// ---------------------------------------------------------------------------
// The tile's 20x20 feature halo is computed from the G-buffer into LDS; the
// first convolution turns it into the 18x18 hidden region, the second into the
// 16x16 logits.  Output pixels of a region are processed as a linear sequence of
// 16-pixel chunks: one MFMA per tap and 16 output channels multiplies W_tap
// (registers, A operand) by the chunk pixels' shifted inputs (LDS, B operand).
// Hidden positions outside the image hold zero (the second convolution's zero
// padding).  Weights, biases and the bf16 rounding points are those of the
// layer-by-layer path (the biases enter as the accumulators' initial values, so
// the FP32 sums round in another order).  The grid is persistent:
// each workgroup keeps both layers' weights in registers across its tiles (three
// waves per SIMD, with a few spilled registers: reloading the second layer's
// weights per tile instead costs more).
constexpr int kTwoHalo = kTile + 4, kTwoHid = kTile + 2;
constexpr int kFeatStride = kFeatCh + 8;  // bf16 per LDS pixel (16 channels)
constexpr int kHidStride = kCh + 8;       // bf16 per LDS pixel (32 channels)

__global__ void __launch_bounds__(256, 3) k_two_layer(CnnGeom g, pbrt_cnn_images im, const uint16_t* __restrict__ wf,
                                                   const float* __restrict__ master) {
    __shared__ __attribute__((aligned(16))) uint16_t buf_f[kTile * kTile * kHidStride];  // features, then logits
    __shared__ __attribute__((aligned(16))) uint16_t buf_h[kTwoHid * kTwoHid * kHidStride];
    __shared__ __attribute__((aligned(16))) uint2 ntile[kTwoHalo * kTwoHalo];  // noisy taps, raw FP16
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int j = lane & 15, grp = lane >> 4;
    const int tiles_x = (g.width + kTile - 1) / kTile, tiles_y = (g.height + kTile - 1) / kTile;
    const int per_image = tiles_x * tiles_y, ntiles = per_image * g.count;
    // persistent: workgroup b runs on XCD b % 8 and walks that XCD's contiguous
    // share of the tiles (neighbouring tiles, whose halos overlap, share an L2);
    // the weights are loaded once per workgroup
    const int slots = gridDim.x / 8, xcd = blockIdx.x % 8, slot = blockIdx.x / 8;
    const int chunk = (ntiles + 7) / 8, begin = xcd * chunk, end = min(ntiles, begin + chunk);
    if (begin + slot >= end) return;
    const size_t plane = (size_t)g.width * g.height;

    // A operands (row = output channel ct*16 + j): 16 input channels per lane for the
    // first layer, 8 of 32 for the second
    uint2 w0[kTaps][2];
    uint4 w1[kTaps][2];
#pragma unroll
    for (int tp = 0; tp < kTaps; ++tp)
#pragma unroll
        for (int ct = 0; ct < 2; ++ct) {
            w0[tp][ct] =
                *reinterpret_cast<const uint2*>(wf + weight_offset(0) + (tp * kCh + ct * 16 + j) * kFeatCh + 4 * grp);
            w1[tp][ct] =
                *reinterpret_cast<const uint4*>(wf + weight_offset(1) + (tp * kCh + ct * 16 + j) * kCh + 8 * grp);
        }
    const float* b0 = master + layer_offset(0) + layer_weights(0);
    const float* b1 = master + layer_offset(1) + layer_weights(1);
    // the biases of this lane's four output rows: the accumulators' initial values
    float4v bv0[2], bv1[2];
#pragma unroll
    for (int ct = 0; ct < 2; ++ct)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            bv0[ct][i] = b0[ct * 16 + 4 * grp + i];
            bv1[ct][i] = b1[ct * 16 + 4 * grp + i];
        }

    for (int t = begin + slot; t < end; t += slots) {
        const int img = t / per_image, rem = t % per_image;
        const int x0 = (rem % tiles_x) * kTile, y0 = (rem / tiles_x) * kTile;

        for (int p = tid; p < kTwoHalo * kTwoHalo; p += 256) {
            const int gx = x0 - 2 + p % kTwoHalo, gy = y0 - 2 + p / kTwoHalo;
            uint4 lo = make_uint4(0, 0, 0, 0), hi = lo;
            if (gx >= 0 && gy >= 0 && gx < g.width && gy < g.height)
                pixel_features(im, plane * img + (size_t)gy * g.width + gx, lo, hi);
            *reinterpret_cast<uint4*>(&buf_f[p * kFeatStride]) = lo;
            *reinterpret_cast<uint4*>(&buf_f[p * kFeatStride + 8]) = hi;
            // the clamped noisy taps share the 20x20 footprint
            const int yy = min(max(gy, 0), g.height - 1), xx = min(max(gx, 0), g.width - 1);
            ntile[p] = reinterpret_cast<const uint2*>(im.noisy)[plane * img + (size_t)yy * g.width + xx];
        }
        __syncthreads();
        {
            constexpr int npix = kTwoHid * kTwoHid, nchunks = (npix + 15) / 16;
            for (int c = wave; c < nchunks; c += 4) {
                const int idx = c * 16 + j;
                const int ic = idx < npix ? idx : npix - 1;
                const int oy = ic / kTwoHid, ox = ic % kTwoHid;
                const uint16_t* base = buf_f + (oy * kTwoHalo + ox) * kFeatStride + 4 * grp;
                float4v acc[2] = {bv0[0], bv0[1]};
#pragma unroll
                for (int tp = 0; tp < kTaps; ++tp) {
                    const short4v bx = __builtin_bit_cast(
                        short4v, *reinterpret_cast<const uint2*>(base + ((tp / 3) * kTwoHalo + tp % 3) * kFeatStride));
#pragma unroll
                    for (int ct = 0; ct < 2; ++ct)
                        acc[ct] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(__builtin_bit_cast(short4v, w0[tp][ct]),
                                                                            bx, acc[ct], 0, 0, 0);
                }
                if (idx < npix) {
                    const int gx = x0 - 1 + ox, gy = y0 - 1 + oy;
                    // ReLU inside the image, 0 outside (zero padding): median of (x, 0, hi)
                    // with hi = +inf inside, 0 outside
                    const float hi = (gx >= 0 && gy >= 0 && gx < g.width && gy < g.height) ? INFINITY : 0.0f;
#pragma unroll
                    for (int ct = 0; ct < 2; ++ct) {
                        float v[4];
#pragma unroll
                        for (int i = 0; i < 4; ++i) v[i] = __builtin_amdgcn_fmed3f(acc[ct][i], 0.0f, hi);
                        *reinterpret_cast<uint2*>(&buf_h[idx * kHidStride + ct * 16 + 4 * grp]) =
                            make_uint2(pack_bf16(v[0], v[1]), pack_bf16(v[2], v[3]));
                    }
                }
            }
        }
        __syncthreads();
        {
            constexpr int nchunks = kTile * kTile / 16;
            for (int c = wave; c < nchunks; c += 4) {
                const int idx = c * 16 + j;
                const int oy = idx / kTile, ox = idx % kTile;
                const uint16_t* base = buf_h + (oy * kTwoHid + ox) * kHidStride + 8 * grp;
                float4v acc[2] = {bv1[0], bv1[1]};
#pragma unroll
                for (int tp = 0; tp < kTaps; ++tp) {
                    const bf16x8 bx = __builtin_bit_cast(
                        bf16x8, *reinterpret_cast<const uint4*>(base + ((tp / 3) * kTwoHid + tp % 3) * kHidStride));
#pragma unroll
                    for (int ct = 0; ct < 2; ++ct)
                        acc[ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, w1[tp][ct]), bx,
                                                                          acc[ct], 0, 0, 0);
                }
#pragma unroll
                for (int ct = 0; ct < 2; ++ct) {
                    *reinterpret_cast<uint2*>(&buf_f[idx * kHidStride + ct * 16 + 4 * grp]) =
                        make_uint2(pack_bf16(acc[ct][0], acc[ct][1]), pack_bf16(acc[ct][2], acc[ct][3]));
                }
            }
        }
        __syncthreads();

        // softmax over this thread's pixel's 25 logits and the 5x5 weighted sum
        kpn_apply_pixel(g, im, buf_f + tid * kHidStride, ntile, kTwoHalo, plane * img, x0, y0, tid);
        __syncthreads();  // the next tile's features overwrite the logits and taps
    }
}

// ---------------------------------------------------------------------------
// Compact network, fused: features -> 3x3 conv 16->32 (ReLU) -> 1x1 conv 32->25
// logits -> softmax 5x5 kernel -> filtered radiance (inference only).
// ---------------------------------------------------------------------------
// The two-layer network with its second convolution pointwise (pbrt_cnn_create_compact:
// the eight outer taps of that layer are held at zero in training), which needs no
// hidden halo: per 16x16 tile the 18x18 feature halo goes to LDS and each 16-pixel
// row of the tile runs both layers in registers.  The first convolution pairs two
// taps per v_mfma_f32_16x16x32_bf16 (K = 16 channels of tap 2p, then of tap 2p + 1;
// 5 instructions per 16 output channels); its ReLU'd, bf16-rounded output is already
// the B operand of the pointwise layer, whose K order (each lane's four channels of
// the first 16, then its four of the second 16) the second layer's A operand follows.
// The logits go to LDS for the per-pixel softmax (kpn_apply_pixel), which each wave
// runs on the 64 pixels it computed.  Rounding points and biases are those of the
// two-layer network.  A tile's G-buffer inputs are loaded while the previous tile is
// computed (three waves per SIMD: four, at most 128 registers, spilled and ran 23 %
// slower; without the look-ahead loads, four waves ran 5 % slower).
constexpr int kCompactHalo = kTile + 2;  // features of the 3x3 convolution
constexpr int kCompactTaps = kTile + 4;  // noisy taps of the 5x5 filter

__global__ void __launch_bounds__(256, 3) k_compact(CnnGeom g, pbrt_cnn_images im, const uint16_t* __restrict__ wf,
                                                    const float* __restrict__ master) {
    __shared__ __attribute__((aligned(16))) uint16_t feat[kCompactHalo * kCompactHalo * kFeatStride];
    __shared__ __attribute__((aligned(16))) uint16_t logits[kTile * kTile * kHidStride];
    __shared__ __attribute__((aligned(16))) uint2 ntile[kCompactTaps * kCompactTaps];  // noisy taps, raw FP16
    __shared__ uint32_t albedo[kCompactHalo * kCompactHalo];                            // RGBA8 of the halo
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int j = lane & 15, grp = lane >> 4;
    const int tiles_x = (g.width + kTile - 1) / kTile, tiles_y = (g.height + kTile - 1) / kTile;
    const int per_image = tiles_x * tiles_y, ntiles = per_image * g.count;
    // persistent, XCD-aware tile ranges as in k_two_layer
    const int slots = gridDim.x / 8, xcd = blockIdx.x % 8, slot = blockIdx.x / 8;
    const int chunk = (ntiles + 7) / 8, begin = xcd * chunk, end = min(ntiles, begin + chunk);
    if (begin + slot >= end) return;
    const size_t plane = (size_t)g.width * g.height;

    // first layer, A operand of tap pair p (row = output channel ct * 16 + j): lanes of
    // groups 0 / 1 hold tap 2p's input channels 0-7 / 8-15, groups 2 / 3 tap 2p + 1's
    // (zero past tap 8)
    const int pair_tap = grp >> 1, chan = 8 * (grp & 1);
    uint4 w0[5][2];
#pragma unroll
    for (int pr = 0; pr < 5; ++pr)
#pragma unroll
        for (int ct = 0; ct < 2; ++ct) {
            const int tap = 2 * pr + pair_tap;
            w0[pr][ct] = tap < kTaps ? *reinterpret_cast<const uint4*>(wf + weight_offset(0) +
                                                                       (tap * kCh + ct * 16 + j) * kFeatCh + chan)
                                     : make_uint4(0, 0, 0, 0);
        }
    // pointwise layer (centre tap), A operand (row = logit ct * 16 + j): K slot
    // 8 * grp + k is hidden channel 4 * grp + k (k < 4) or 16 + 4 * grp + k - 4
    uint4 w1[2];
#pragma unroll
    for (int ct = 0; ct < 2; ++ct) {
        const uint16_t* row = wf + weight_offset(1) + (4 * kCh + ct * 16 + j) * kCh;
        const uint2 lo = *reinterpret_cast<const uint2*>(row + 4 * grp);
        const uint2 hi = *reinterpret_cast<const uint2*>(row + 16 + 4 * grp);
        w1[ct] = make_uint4(lo.x, lo.y, hi.x, hi.y);
    }
    const float* b0 = master + layer_offset(0) + layer_weights(0);
    const float* b1 = master + layer_offset(1) + layer_weights(1);
    float4v bv0[2], bv1[2];
#pragma unroll
    for (int ct = 0; ct < 2; ++ct)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            bv0[ct][i] = b0[ct * 16 + 4 * grp + i];
            bv1[ct][i] = b1[ct * 16 + 4 * grp + i];
        }

    // This thread's share of a tile's inputs: feature halo positions tid and tid + 256
    // (of 18 x 18) and tap positions tid and tid + 256 (of 20 x 20).  They are loaded
    // one tile ahead, so the loads of the next tile are in flight while this one runs.
    constexpr int kFeatPix = kCompactHalo * kCompactHalo, kTapPix = kCompactTaps * kCompactTaps;
    RawPixel raw[2];
    bool inside[2];
    uint2 tapv[2];
    auto fetch = [&](int t) {
        const int img = t / per_image, rem = t % per_image;
        const int x0 = (rem % tiles_x) * kTile, y0 = (rem / tiles_x) * kTile;
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int p = tid + 256 * k;
            const int gx = x0 - 1 + p % kCompactHalo, gy = y0 - 1 + p / kCompactHalo;
            inside[k] = p < kFeatPix && gx >= 0 && gy >= 0 && gx < g.width && gy < g.height;
            if (inside[k]) raw[k] = load_raw_pixel(im, plane * img + (size_t)gy * g.width + gx);
            const int q = tid + 256 * k;
            if (q < kTapPix) {
                const int tx = x0 - 2 + q % kCompactTaps, ty = y0 - 2 + q / kCompactTaps;
                const int yy = min(max(ty, 0), g.height - 1), xx = min(max(tx, 0), g.width - 1);
                tapv[k] = reinterpret_cast<const uint2*>(im.noisy)[plane * img + (size_t)yy * g.width + xx];
            }
        }
    };
    // the feature tile's padding (channels 16..23 of every pixel) stays zero: the B
    // operand past the last tap reads it
    for (int p = tid; p < kCompactHalo * kCompactHalo; p += 256)
        *reinterpret_cast<uint4*>(&feat[p * kFeatStride + kFeatCh]) = make_uint4(0, 0, 0, 0);
    fetch(begin + slot);

    for (int t = begin + slot; t < end; t += slots) {
        const int img = t / per_image, rem = t % per_image;
        const int x0 = (rem % tiles_x) * kTile, y0 = (rem / tiles_x) * kTile;
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int p = tid + 256 * k;
            if (p < kFeatPix) {
                uint4 lo = make_uint4(0, 0, 0, 0), hi = lo;
                if (inside[k]) raw_features(raw[k], lo, hi);
                *reinterpret_cast<uint4*>(&feat[p * kFeatStride]) = lo;
                *reinterpret_cast<uint4*>(&feat[p * kFeatStride + 8]) = hi;
                albedo[p] = inside[k] ? raw[k].albedo : 0u;
            }
            if (p < kTapPix) ntile[p] = tapv[k];
        }
        if (t + slots < end) fetch(t + slots);
        __syncthreads();
        for (int oy = 4 * wave; oy < 4 * wave + 4; ++oy) {
            // tile row oy, pixel j: the first layer over tap pairs
            float4v acc[2] = {bv0[0], bv0[1]};
#pragma unroll
            for (int pr = 0; pr < 5; ++pr) {
                // past tap 8 the lane reads its pixel's zero padding (LDS bytes 32..47)
                const int tap = 2 * pr + pair_tap;
                const int off = tap < kTaps ? ((oy + tap / 3) * kCompactHalo + j + tap % 3) * kFeatStride + chan
                                            : ((oy + 1) * kCompactHalo + j + 1) * kFeatStride + kFeatCh;
                const bf16x8 bx = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(&feat[off]));
#pragma unroll
                for (int ct = 0; ct < 2; ++ct)
                    acc[ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, w0[pr][ct]), bx,
                                                                      acc[ct], 0, 0, 0);
            }
            // ReLU (one v_med3_f32 each), bf16: this lane's hidden channels 4 grp .. + 3 and
            // 16 + 4 grp .. + 3 of pixel j
            auto relu = [](float x) { return __builtin_amdgcn_fmed3f(x, 0.0f, INFINITY); };
            const uint4 hid = make_uint4(pack_bf16(relu(acc[0][0]), relu(acc[0][1])),
                                         pack_bf16(relu(acc[0][2]), relu(acc[0][3])),
                                         pack_bf16(relu(acc[1][0]), relu(acc[1][1])),
                                         pack_bf16(relu(acc[1][2]), relu(acc[1][3])));
            const bf16x8 hx = __builtin_bit_cast(bf16x8, hid);
            float4v z[2] = {bv1[0], bv1[1]};
#pragma unroll
            for (int ct = 0; ct < 2; ++ct)
                z[ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, w1[ct]), hx, z[ct], 0, 0,
                                                                0);
            const int idx = oy * kTile + j;
#pragma unroll
            for (int ct = 0; ct < 2; ++ct)
                *reinterpret_cast<uint2*>(&logits[idx * kHidStride + ct * 16 + 4 * grp]) =
                    make_uint2(pack_bf16(z[ct][0], z[ct][1]), pack_bf16(z[ct][2], z[ct][3]));
        }
        // each wave filters the 64 pixels whose logits it wrote (rows 4 wave .. 4 wave + 3,
        // lane = pixel: the noisy-tap reads of its four rows fall in distinct LDS banks):
        // LDS operations of one wave complete in order, so no workgroup barrier here
        const int own = 64 * wave + lane;
        kpn_apply_pixel(g, im, logits + own * kHidStride, ntile, kCompactTaps, plane * img, x0, y0, own,
                        &albedo[(own / kTile + 1) * kCompactHalo + own % kTile + 1]);
        __syncthreads();  // the next tile's features overwrite the logits and taps
    }
}

// ---------------------------------------------------------------------------
// Workspace
// ---------------------------------------------------------------------------
size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

struct CnnWork {
    uint16_t* feat;
    uint16_t* act[kMaxLayers + 1];  // act[0] = feat; act[1..L-1] hidden, act[L] logits
    uint16_t* grad_act[2];       // training: upstream gradients (ping-pong)
    float* partial;              // training: per-tile weight-gradient partials
    float* loss_part;
    float* grad;                 // training: padded gradient
    size_t bytes;
};

CnnWork layout(uint32_t w, uint32_t h, uint32_t count, bool training, int layers, char* base) {
    CnnWork k{};
    const size_t px = (size_t)w * h * count;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        char* p = base ? base + off : nullptr;
        off = align256(off + bytes);
        return p;
    };
    k.feat = (uint16_t*)take(px * kFeatCh * 2);
    k.act[0] = k.feat;
    if (training) {
        for (int l = 1; l <= layers; ++l) k.act[l] = (uint16_t*)take(px * kCh * 2);
        k.grad_act[0] = (uint16_t*)take(px * kCh * 2);
        k.grad_act[1] = (uint16_t*)take(px * kCh * 2);
        const size_t tiles = (size_t)((w + kTile - 1) / kTile) * ((h + kTile - 1) / kTile) * count;
        k.partial = (float*)take(tiles * (kTaps * kCh * kCh + kCh) * 4);
        k.loss_part = (float*)take(((px + 255) / 256) * 4);
        k.grad = (float*)take(master_size(layers) * 4);
    } else {
        // ping-pong hidden activations; logits in the second buffer
        uint16_t* a = (uint16_t*)take(px * kCh * 2);
        uint16_t* b = (uint16_t*)take(px * kCh * 2);
        for (int l = 1; l <= layers; ++l) k.act[l] = (l & 1) ? a : b;
    }
    k.bytes = off;
    return k;
}

}  // namespace
}  // namespace pbrt

using namespace pbrt;

struct pbrt_cnn {
    DeviceBuffer master, m, v, wf, wb, grad_map;
    uint64_t step = 0;
    int layers = kMaxLayers;
    bool compact = false;  // two layers, the second pointwise (its outer taps held at zero)
};

namespace pbrt {
namespace {

void check_images(const pbrt_cnn_images* im, bool training, int layers, const char* fn) {
    PBRT_REQUIRE(im, std::string(fn) + ": NULL images");
    PBRT_REQUIRE(im->width > 0 && im->height > 0 && im->count > 0, std::string(fn) + ": empty batch");
    PBRT_REQUIRE(im->noisy && im->depth && im->normal && im->albedo && im->spp && im->motion,
                 std::string(fn) + ": missing input image");
    PBRT_REQUIRE(!training || im->target, std::string(fn) + ": missing target");
    if (layers == 0) {  // the fused two-layer network needs no workspace
        PBRT_REQUIRE(im->denoised || im->final_image, std::string(fn) + ": no output image");
        return;
    }
    const size_t need = layout(im->width, im->height, im->count, training, layers, nullptr).bytes;
    PBRT_REQUIRE(im->workspace && im->workspace_bytes >= need, std::string(fn) + ": workspace too small");
}

// Workgroups for a persistent convolution: every CU's resident share, a multiple of
// the 8 XCDs, no more than the tiles.
template <class K>
dim3 persistent_grid(K kernel, int tiles) {
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        PBRT_HIP(hipGetDevice(&dev));
        PBRT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    }
    int per_cu = 0;
    PBRT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, 0));
    int blocks = std::max(1, cus * std::max(per_cu, 1));
    blocks = std::min(blocks, ((tiles + 7) / 8) * 8);
    return dim3((unsigned)(((blocks + 7) / 8) * 8));
}

template <int CIN, int MODE>
void launch_conv(const CnnGeom& g, const uint16_t* in, const uint16_t* mask, const uint16_t* w, const float* b,
                 uint16_t* out, hipStream_t st, const pbrt_cnn_images& im = pbrt_cnn_images{}) {
    const int tiles = ((g.width + kTile - 1) / kTile) * ((g.height + kTile - 1) / kTile) * g.count;
    const dim3 grid = persistent_grid(k_conv3x3<CIN, MODE>, tiles);
    hipLaunchKernelGGL((k_conv3x3<CIN, MODE>), grid, dim3(256), 0, st, g, in, mask, w, b, out, im);
}

// inference of a two-layer network: one fused kernel, no workspace
void forward_two_layer(const pbrt_cnn* net, const pbrt_cnn_images& im, hipStream_t st) {
    const CnnGeom g{(int)im.width, (int)im.height, (int)im.count};
    const size_t tiles = (size_t)((im.width + kTile - 1) / kTile) * ((im.height + kTile - 1) / kTile) * im.count;
    PBRT_REQUIRE(tiles < (1u << 30), "pbrt_cnn_denoise: batch too large");
    const dim3 grid = persistent_grid(k_two_layer, (int)tiles);
    hipLaunchKernelGGL(k_two_layer, grid, dim3(256), 0, st, g, im, net->wf.as<uint16_t>(), net->master.as<float>());
}

// inference of the compact network: one fused kernel, no workspace
void forward_compact(const pbrt_cnn* net, const pbrt_cnn_images& im, hipStream_t st) {
    const CnnGeom g{(int)im.width, (int)im.height, (int)im.count};
    const size_t tiles = (size_t)((im.width + kTile - 1) / kTile) * ((im.height + kTile - 1) / kTile) * im.count;
    const dim3 grid = persistent_grid(k_compact, (int)tiles);
    hipLaunchKernelGGL(k_compact, grid, dim3(256), 0, st, g, im, net->wf.as<uint16_t>(), net->master.as<float>());
}

// zero the padded gradient of every tap but the centre one of layer l (the compact
// network's pointwise layer): with zero gradient and moments, Adam leaves those
// weights at zero
__global__ void __launch_bounds__(256) k_mask_outer_taps(float* grad, int cin_pad) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= kTaps * kCh * cin_pad) return;
    if (i / (kCh * cin_pad) != kTaps / 2) grad[i] = 0.0f;
}

void mask_pointwise_gradient(const pbrt_cnn* net, float* padded_grad, hipStream_t st) {
    if (!net->compact) return;
    const int l = net->layers - 1;
    hipLaunchKernelGGL(k_mask_outer_taps, dim3((layer_weights(l) + 255) / 256), dim3(256), 0, st,
                       padded_grad + layer_offset(l), cin_pad(l));
}

// inference: kernel prediction fused into the last layer
void forward_fused(const pbrt_cnn* net, const pbrt_cnn_images& im, const CnnWork& k, hipStream_t st) {
    const CnnGeom g{(int)im.width, (int)im.height, (int)im.count};
    const size_t px = (size_t)im.width * im.height * im.count;
    hipLaunchKernelGGL(k_features, dim3((unsigned)((px + 255) / 256)), dim3(256), 0, st, g, im, k.feat);
    const uint16_t* wf = net->wf.as<uint16_t>();
    const float* master = net->master.as<float>();
    const int layers = net->layers;
    for (int l = 0; l < layers; ++l) {
        const uint16_t* w = wf + weight_offset(l);
        const float* b = master + layer_offset(l) + layer_weights(l);
        if (l == 0)
            launch_conv<kFeatCh, kConvRelu>(g, k.act[0], nullptr, w, b, k.act[1], st);
        else if (l < layers - 1)
            launch_conv<kCh, kConvRelu>(g, k.act[l], nullptr, w, b, k.act[l + 1], st);
        else
            launch_conv<kCh, kConvKpn>(g, k.act[l], nullptr, w, b, nullptr, st, im);
    }
}

void forward(const pbrt_cnn* net, const pbrt_cnn_images& im, const CnnWork& k, hipStream_t st) {
    const CnnGeom g{(int)im.width, (int)im.height, (int)im.count};
    const size_t px = (size_t)im.width * im.height * im.count;
    hipLaunchKernelGGL(k_features, dim3((unsigned)((px + 255) / 256)), dim3(256), 0, st, g, im, k.feat);
    const uint16_t* wf = net->wf.as<uint16_t>();
    const float* master = net->master.as<float>();
    const int layers = net->layers;
    for (int l = 0; l < layers; ++l) {
        const uint16_t* w = wf + weight_offset(l);
        const float* b = master + layer_offset(l) + layer_weights(l);
        if (l == 0)
            launch_conv<kFeatCh, kConvRelu>(g, k.act[0], nullptr, w, b, k.act[1], st);
        else if (l < layers - 1)
            launch_conv<kCh, kConvRelu>(g, k.act[l], nullptr, w, b, k.act[l + 1], st);
        else
            launch_conv<kCh, kConvLinear>(g, k.act[l], nullptr, w, b, k.act[l + 1], st);
    }
}

// loss and padded gradient of the batch into k.grad / loss (device)
void backward(const pbrt_cnn* net, const pbrt_cnn_images& im, const CnnWork& k, float* loss, hipStream_t st) {
    const CnnGeom g{(int)im.width, (int)im.height, (int)im.count};
    const size_t px = (size_t)im.width * im.height * im.count;
    const int blocks = (int)((px + 255) / 256);
    const int layers = net->layers;
    hipLaunchKernelGGL(k_kpn<true>, dim3(blocks), dim3(256), 0, st, g, im, k.act[layers], k.grad_act[0],
                       k.loss_part);
    if (loss) hipLaunchKernelGGL(k_loss_total, dim3(1), dim3(64), 0, st, k.loss_part, blocks, 1.0f / (float)(px * 3), loss);
    const dim3 grid((im.width + kTile - 1) / kTile, (im.height + kTile - 1) / kTile, im.count);
    const int tiles = (int)(grid.x * grid.y * grid.z);
    const uint16_t* wb = net->wb.as<uint16_t>();
    int cur = 0;  // grad_act[cur] holds dA of layer l+1 (dLogits for the last layer)
    for (int l = layers - 1; l >= 0; --l) {
        const uint16_t* dsrc = k.grad_act[cur];
        const uint16_t* mask = l == layers - 1 ? nullptr : k.act[l + 1];
        const int size = layer_size(l);
        if (l == 0)
            hipLaunchKernelGGL(k_conv3x3_wgrad<kFeatCh>, grid, dim3(256), 0, st, g, dsrc, mask, k.act[0], k.partial);
        else
            hipLaunchKernelGGL(k_conv3x3_wgrad<kCh>, grid, dim3(256), 0, st, g, dsrc, mask, k.act[l], k.partial);
        hipLaunchKernelGGL(k_reduce_partials, dim3((size + 255) / 256), dim3(256), 0, st, k.partial, tiles, size,
                           k.grad + layer_offset(l));
        if (l > 0) {
            launch_conv<kCh, kConvDgrad>(g, dsrc, mask, wb + weight_offset(l), nullptr, k.grad_act[cur ^ 1], st);
            cur ^= 1;
        }
    }
}

}  // namespace
}  // namespace pbrt

extern "C" int32_t pbrt_cnn_parameter_count(uint32_t* out) {
    return guard([&] {
        PBRT_REQUIRE(out, "pbrt_cnn_parameter_count: NULL output");
        *out = (uint32_t)canonical_count(kMaxLayers);
    });
}

extern "C" int32_t pbrt_cnn_layers_parameter_count(uint32_t layers, uint32_t* out) {
    return guard([&] {
        PBRT_REQUIRE(out, "pbrt_cnn_layers_parameter_count: NULL output");
        PBRT_REQUIRE(layers >= 2 && layers <= (uint32_t)kMaxLayers,
                     "pbrt_cnn_layers_parameter_count: layers must be 2..6");
        *out = (uint32_t)canonical_count((int)layers);
    });
}

extern "C" int32_t pbrt_cnn_layers(const pbrt_cnn* net, uint32_t* out) {
    return guard([&] {
        PBRT_REQUIRE(net && out, "pbrt_cnn_layers: NULL argument");
        *out = (uint32_t)net->layers;
    });
}

extern "C" int32_t pbrt_cnn_create(uint64_t seed, pbrt_cnn** out) {
    return pbrt_cnn_create_layers((uint32_t)kMaxLayers, seed, out);
}

extern "C" int32_t pbrt_cnn_create_layers(uint32_t layers, uint64_t seed, pbrt_cnn** out) {
    return guard([&] {
        PBRT_REQUIRE(out, "pbrt_cnn_create: NULL output");
        *out = nullptr;
        PBRT_REQUIRE(layers >= 2 && layers <= (uint32_t)kMaxLayers, "pbrt_cnn_create: layers must be 2..6");
        const int L = (int)layers;
        auto net = std::make_unique<pbrt_cnn>();
        net->layers = L;
        // He-normal weights from splitmix64 + Box-Muller; last layer scaled by 0.1 so the
        // initial kernel is close to a 5x5 box filter; zero biases
        std::vector<float> canon(canonical_count(L));
        uint64_t state = seed;
        auto next = [&]() {
            uint64_t z = (state += 0x9E3779B97F4A7C15ull);
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            return z ^ (z >> 31);
        };
        auto uniform = [&]() { return ((double)(next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); };
        size_t o = 0;
        for (int l = 0; l < L; ++l) {
            const double std = std::sqrt(2.0 / (kTaps * cin_real(l))) * (l == L - 1 ? 0.1 : 1.0);
            const int nw = cout_real(l, L) * cin_real(l) * kTaps;
            for (int i = 0; i < nw; ++i) {
                const double u1 = uniform(), u2 = uniform();
                canon[o++] = (float)(std * std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2));
            }
            for (int i = 0; i < cout_real(l, L); ++i) canon[o++] = 0.0f;
        }
        const std::vector<int> map = canonical_map(L);
        net->grad_map.upload(map.data(), map.size() * sizeof(int));
        std::vector<float> zeros(master_size(L), 0.0f);
        net->master.upload(zeros.data(), zeros.size() * 4);
        net->m.upload(zeros.data(), zeros.size() * 4);
        net->v.upload(zeros.data(), zeros.size() * 4);
        std::vector<uint16_t> zw(weight_size(L), 0);
        net->wf.upload(zw.data(), zw.size() * 2);
        net->wb.upload(zw.data(), zw.size() * 2);
        *out = net.release();
        int32_t rc = pbrt_cnn_set_parameters(*out, canon.data());
        if (rc != PBRT_OK) {
            delete *out;
            *out = nullptr;
            fail(rc, "pbrt_cnn_create: initial parameters");
        }
    });
}

extern "C" int32_t pbrt_cnn_create_compact(uint64_t seed, pbrt_cnn** out) {
    return guard([&] {
        PBRT_REQUIRE(out, "pbrt_cnn_create_compact: NULL output");
        *out = nullptr;
        pbrt_cnn* net = nullptr;
        int32_t rc = pbrt_cnn_create_layers(2, seed, &net);
        if (rc != PBRT_OK) fail(rc, std::string("pbrt_cnn_create_compact: ") + pbrt_last_error());
        std::unique_ptr<pbrt_cnn> hold(net);
        net->compact = true;
        // the two-layer initialisation with the second layer's fan-in of one tap: its
        // centre weights scaled by 3 (sqrt of the 9 taps), the outer ones zeroed
        std::vector<float> canon(canonical_count(2));
        rc = pbrt_cnn_get_parameters(net, canon.data());
        if (rc != PBRT_OK) fail(rc, std::string("pbrt_cnn_create_compact: ") + pbrt_last_error());
        const size_t o = (size_t)cout_real(0, 2) * cin_real(0) * kTaps + cout_real(0, 2);
        for (size_t i = 0; i < (size_t)cout_real(1, 2) * cin_real(1) * kTaps; ++i)
            canon[o + i] = i % kTaps == kTaps / 2 ? 3.0f * canon[o + i] : 0.0f;
        rc = pbrt_cnn_set_parameters(net, canon.data());
        if (rc != PBRT_OK) fail(rc, std::string("pbrt_cnn_create_compact: ") + pbrt_last_error());
        *out = hold.release();
    });
}

extern "C" int32_t pbrt_cnn_is_compact(const pbrt_cnn* net, uint32_t* out) {
    return guard([&] {
        PBRT_REQUIRE(net && out, "pbrt_cnn_is_compact: NULL argument");
        *out = net->compact ? 1u : 0u;
    });
}

extern "C" int32_t pbrt_cnn_destroy(pbrt_cnn* net) {
    return guard([&] { delete net; });
}

extern "C" int32_t pbrt_cnn_get_parameters(const pbrt_cnn* net, float* host) {
    return guard([&] {
        PBRT_REQUIRE(net && host, "pbrt_cnn_get_parameters: NULL argument");
        const int n = master_size(net->layers);
        std::vector<float> padded(n);
        PBRT_HIP(hipMemcpy(padded.data(), net->master.ptr, n * 4, hipMemcpyDeviceToHost));
        const std::vector<int> map = canonical_map(net->layers);
        for (size_t i = 0; i < map.size(); ++i) host[i] = padded[map[i]];
    });
}

extern "C" int32_t pbrt_cnn_set_parameters(pbrt_cnn* net, const float* host) {
    return guard([&] {
        PBRT_REQUIRE(net && host, "pbrt_cnn_set_parameters: NULL argument");
        const int n = master_size(net->layers);
        std::vector<float> padded(n, 0.0f);
        const std::vector<int> map = canonical_map(net->layers);
        for (size_t i = 0; i < map.size(); ++i) padded[map[i]] = host[i];
        if (net->compact) {
            // the pointwise layer keeps only its centre tap
            const int l = net->layers - 1;
            for (int i = 0; i < layer_weights(l); ++i)
                if (i / (kCh * cin_pad(l)) != kTaps / 2) padded[layer_offset(l) + i] = 0.0f;
        }
        PBRT_HIP(hipMemcpy(net->master.ptr, padded.data(), n * 4, hipMemcpyHostToDevice));
        PBRT_HIP(hipMemset(net->m.ptr, 0, n * 4));
        PBRT_HIP(hipMemset(net->v.ptr, 0, n * 4));
        net->step = 0;
        hipLaunchKernelGGL(k_refresh, dim3((n + 255) / 256), dim3(256), 0, 0, layer_table(net->layers),
                           net->master.as<float>(), net->wf.as<uint16_t>(), net->wb.as<uint16_t>());
        PBRT_HIP(hipGetLastError());
        PBRT_HIP(hipDeviceSynchronize());
    });
}

extern "C" int32_t pbrt_cnn_workspace_size(uint32_t w, uint32_t h, uint32_t count, uint32_t training, size_t* out) {
    return guard([&] {
        PBRT_REQUIRE(out, "pbrt_cnn_workspace_size: NULL output");
        PBRT_REQUIRE(w > 0 && h > 0 && count > 0, "pbrt_cnn_workspace_size: empty batch");
        *out = layout(w, h, count, training != 0, kMaxLayers, nullptr).bytes;
    });
}

extern "C" int32_t pbrt_cnn_denoise(const pbrt_cnn* net, const pbrt_cnn_images* im, void* stream) {
    return guard([&] {
        PBRT_REQUIRE(net, "pbrt_cnn_denoise: NULL network");
        hipStream_t st = (hipStream_t)stream;
        if (net->compact) {
            check_images(im, false, 0, "pbrt_cnn_denoise");
            forward_compact(net, *im, st);
        } else if (net->layers == 2) {
            check_images(im, false, 0, "pbrt_cnn_denoise");
            forward_two_layer(net, *im, st);
        } else {
            check_images(im, false, net->layers, "pbrt_cnn_denoise");
            const CnnWork k =
                layout(im->width, im->height, im->count, false, net->layers, static_cast<char*>(im->workspace));
            forward_fused(net, *im, k, st);
        }
        PBRT_HIP(hipGetLastError());
    });
}

extern "C" int32_t pbrt_cnn_gradients(const pbrt_cnn* net, const pbrt_cnn_images* im, float* loss, float* grad,
                                      void* stream) {
    return guard([&] {
        PBRT_REQUIRE(net && grad, "pbrt_cnn_gradients: NULL argument");
        check_images(im, true, net->layers, "pbrt_cnn_gradients");
        hipStream_t st = (hipStream_t)stream;
        const CnnWork k =
            layout(im->width, im->height, im->count, true, net->layers, static_cast<char*>(im->workspace));
        forward(net, *im, k, st);
        backward(net, *im, k, loss, st);
        mask_pointwise_gradient(net, k.grad, st);
        const int n = canonical_count(net->layers);
        hipLaunchKernelGGL(k_gather, dim3((n + 255) / 256), dim3(256), 0, st, k.grad, net->grad_map.as<int>(), n,
                           grad);
        PBRT_HIP(hipGetLastError());
    });
}

extern "C" int32_t pbrt_cnn_train_step(pbrt_cnn* net, const pbrt_cnn_images* im, float lr, float* loss,
                                       void* stream) {
    return guard([&] {
        PBRT_REQUIRE(net, "pbrt_cnn_train_step: NULL network");
        PBRT_REQUIRE(lr > 0.0f && std::isfinite(lr), "pbrt_cnn_train_step: learning rate must be positive");
        check_images(im, true, net->layers, "pbrt_cnn_train_step");
        hipStream_t st = (hipStream_t)stream;
        const CnnWork k =
            layout(im->width, im->height, im->count, true, net->layers, static_cast<char*>(im->workspace));
        forward(net, *im, k, st);
        backward(net, *im, k, loss, st);
        mask_pointwise_gradient(net, k.grad, st);
        net->step += 1;
        const float c1 = 1.0f / (1.0f - std::pow(0.9f, (float)net->step));
        const float c2 = 1.0f / (1.0f - std::pow(0.999f, (float)net->step));
        const int n = master_size(net->layers);
        hipLaunchKernelGGL(k_adam, dim3((n + 255) / 256), dim3(256), 0, st, layer_table(net->layers),
                           net->master.as<float>(), net->m.as<float>(), net->v.as<float>(), k.grad, lr, c1, c2,
                           net->wf.as<uint16_t>(), net->wb.as<uint16_t>());
        PBRT_HIP(hipGetLastError());
    });
}
