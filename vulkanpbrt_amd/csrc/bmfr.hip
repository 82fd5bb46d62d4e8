// BMFR (blockwise multi-order feature regression) on gfx950.
//
// Block geometry, features, noise and the post pass follow bmfrGeneral.comp,
// bmfrPre.comp, bmfrFit.comp and bmfrPost.comp (POSITION_TYPE 0: screen-space
// block coordinates plus per-block normalised depth):
//   features  [1, n.x, n.y, n.z, x, y, z, x^2, y^2, z^2 | r, g, b]
//   fit rows  row i of block (bx, by) is padded pixel (bx*b + i / b, by*b + i % b)
//   noise     feature j < 10 gets + 2e-4 * (hash(i + j*B^2 + frame*13*B^2) - 0.5),
//             B = b*b, integer wrap, hash = bmfrGeneral.comp random()
//
// Fits:
//   k_bmfr_fit_householder   the reference's FP32 Householder QR over the fp16
//                            feature buffer, subgroup reductions on wave64
//   k_bmfr_fit_gram<F64>     pre + fit fused, features in registers: the Gram matrix
//                            of the raw columns on v_mfma_f64_16x16x4f64, the
//                            projection against the noised constant column a0 as a
//                            rank-1 downdate in FP64, equilibrated FP64 elimination
//   k_bmfr_fit_gram<F16>     the columns projected against a0 in FP64 per row, scaled
//                            by powers of two, split into fp16 hi + lo parts whose
//                            Gram v_mfma_f32_16x16x32_f16 accumulates (FP32 inside a
//                            32-row chunk, FP64 across chunks); FP64 elimination
//   The BMFR features are nearly collinear on planar blocks (depth is almost a
//   linear function of the block coordinates; only the 1e-4 noise separates
//   them), so cond(A^T A) reaches ~1e10: an FP32-accumulated Gram loses the fit
//   there, the FP64-accumulated one does not.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "capi_common.hpp"
#include "image_ops.hpp"
#include "pbrt_fn.hpp"

namespace pbrt {

constexpr int kAlphaSize = 13;  // features + rgb
constexpr int kFeatures = 10;

struct BmfrGeom {
    int width, height, bs, wp, hp, blocks_x, blocks_y;
    int offx, offy;
    uint32_t frame;
    int fit_threads;
    uint16_t* features;  // [13][hp][wp] fp16
    float* weights;      // [blocks_y][blocks_x][30]
    float* ranges;       // [blocks_y][blocks_x][2]: the block's depth min and max, from the fit's pass
    float* noise;        // [10][bs*bs]: the noise of feature c for each block pixel (Gram fits), in
                         // the fit's load order: entry c * bs^2 + t * kSub + s for thread t's s-th pixel
};

// Spherical G-buffer normal (theta, phi) as (cos phi sin theta, sin phi sin theta,
// cos theta), bmfrPre.comp:26-29.
__device__ __forceinline__ vec3 bmfr_decode_normal(float theta, float phi) {
    float st, ct, sp, cp;
    pbrt_fn::sincos(theta, st, ct);  // bounded-argument polynomial: the angles lie in [-pi, pi]
    pbrt_fn::sincos(phi, sp, cp);
    return make_vec3(cp * st, sp * st, ct);
}

// random(uint) of bmfrGeneral.comp:83-93 (Jenkins' 6-step integer hash)
__device__ __forceinline__ float bmfr_random(uint32_t a) {
    a = (a + 0x7ed55d16u) + (a << 12);
    a = (a ^ 0xc761c23cu) ^ (a >> 19);
    a = (a + 0x165667b1u) + (a << 5);
    a = (a + 0xd3a2646cu) ^ (a << 9);
    a = (a + 0xfd7046c5u) + (a << 3);
    a = (a ^ 0xb55a4f09u) ^ (a >> 16);
    // float(a) / 2^32: dividing by a power of two is exact, so is this product
    return __fmul_rn((float)a, 0x1p-32f);
}

// addRandom (bmfrGeneral.comp:95-97) without contraction
__device__ __forceinline__ float add_noise(float v, uint32_t key) {
    float r = __fsub_rn(bmfr_random(key), 0.5f);
    return __fadd_rn(v, __fmul_rn(2e-4f, r));
}

// The noise of the fused Gram fits.  Its key depends on the row within the block,
// the feature and the frame only, so every block of a frame adds the same values:
// they are computed once per frame, in the order the fit's threads load the block's
// pixels.  With `threads` threads, thread t loads pixels q = t + s * threads
// (s < rows / threads), and entry c * rows + t * (rows / threads) + s holds the
// noise of feature c for pixel q = block-local (q % bs, q / bs) = fit row
// (q % bs) * bs + q / bs, so a thread reads its pixels' values as one vector.
// The fit adds them with __fadd_rn, as add_noise does.
__device__ __forceinline__ int fit_rows_per_thread(int rows, int threads) { return rows > threads ? rows / threads : 1; }

__device__ __forceinline__ void noise_table_entry(uint32_t frame, int bs, int threads, int e, float* table) {
    const int rows = bs * bs;
    const int sub = fit_rows_per_thread(rows, threads);
    const int c = e / rows, k = e % rows;
    const int q = (k / sub) + (k % sub) * threads;
    const uint32_t pb = (uint32_t)rows;
    const uint32_t r = (uint32_t)((q % bs) * bs + q / bs);
    const uint32_t key = r + (uint32_t)c * pb * pb + frame * (uint32_t)kAlphaSize * pb * pb;
    table[e] = __fmul_rn(2e-4f, __fsub_rn(bmfr_random(key), 0.5f));
}

__global__ void __launch_bounds__(256) k_bmfr_noise(uint32_t frame, int bs, int threads, float* table) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e < kFeatures * bs * bs) noise_table_entry(frame, bs, threads, e, table);
}

// ---------------------------------------------------------------------------
// Workgroup reductions (wave64 butterfly, then thread 0 folds the waves in order)
// ---------------------------------------------------------------------------
template <class T, class Op>
__device__ __forceinline__ T wave_reduce(T v, Op op) {
    for (int off = 32; off > 0; off >>= 1) v = op(v, __shfl_xor(v, off));
    return v;
}

template <class T, class Op>
__device__ T block_reduce(T v, Op op, T* scratch) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nwaves = (blockDim.x + 63) >> 6;
    v = wave_reduce(v, op);
    if (lane == 0) scratch[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        T t = scratch[0];
        for (int i = 1; i < nwaves; ++i) t = op(t, scratch[i]);
        scratch[0] = t;
    }
    __syncthreads();
    T r = scratch[0];
    __syncthreads();
    return r;
}

// Wave64 reduction without LDS: DPP swaps within each row of 16 lanes (xor 1,
// xor 2, half-row mirror, row mirror leave every lane of a row with the row's
// result), then the four rows combined from lanes 0, 16, 32, 48.  Every lane
// returns the same value, combined in the same order.
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v, int ctrl) {
    switch (ctrl) {
        case 0xB1: return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, false);   // quad_perm 1,0,3,2
        case 0x4E: return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xF, 0xF, false);   // quad_perm 2,3,0,1
        case 0x141: return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x141, 0xF, 0xF, false); // row_half_mirror
        default: return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x140, 0xF, 0xF, false);   // row_mirror
    }
}
__device__ __forceinline__ float dpp_move(float v, int ctrl) { return __uint_as_float(dpp_u32(__float_as_uint(v), ctrl)); }
__device__ __forceinline__ double dpp_move(double v, int ctrl) {
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    const uint64_t lo = dpp_u32((uint32_t)b, ctrl), hi = dpp_u32((uint32_t)(b >> 32), ctrl);
    return __longlong_as_double((long long)(hi << 32 | lo));
}
__device__ __forceinline__ float read_lane(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double read_lane(double v, int l) {
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), l);
    return __longlong_as_double((long long)((uint64_t)hi << 32 | lo));
}
template <class T, class Op>
__device__ __forceinline__ T wave_reduce_dpp(T v, Op op) {
    v = op(v, dpp_move(v, 0xB1));
    v = op(v, dpp_move(v, 0x4E));
    v = op(v, dpp_move(v, 0x141));
    v = op(v, dpp_move(v, 0x140));
    return op(op(read_lane(v, 0), read_lane(v, 16)), op(read_lane(v, 32), read_lane(v, 48)));
}

struct OpMin {
    __device__ float operator()(float a, float b) const { return fminf(a, b); }
};
struct OpMax {
    __device__ float operator()(float a, float b) const { return fmaxf(a, b); }
};
struct OpAdd {
    template <class T>
    __device__ T operator()(T a, T b) const { return a + b; }
};

struct PixelIn {
    float depth;
    vec3 normal;
    vec3 noisy;
    int img_x, img_y;
    bool inside;  // curAbsolutPos == curImagePos
};

__device__ __forceinline__ PixelIn load_pixel(const BmfrGeom& g, const pbrt_bmfr_images& im, int gx, int gy) {
    PixelIn p;
    int ax = gx - g.offx, ay = gy - g.offy;
    p.img_x = mirror_coord(ax, g.width);
    p.img_y = mirror_coord(ay, g.height);
    p.inside = ax == p.img_x && ay == p.img_y;
    size_t pix = (size_t)p.img_y * g.width + p.img_x;
    p.depth = im.depth[pix];
    float2 n = reinterpret_cast<const float2*>(im.normal)[pix];
    p.normal = bmfr_decode_normal(n.x, n.y);
    p.noisy = xyz(load_half4(im.noisy, pix));
    return p;
}

// The block's depth range, for the post pass (it normalises the depth feature again).
__device__ __forceinline__ void store_range(const BmfrGeom& g, int bx, int by, float mn, float mx) {
    float* r = g.ranges + 2 * ((size_t)by * g.blocks_x + bx);
    r[0] = mn;
    r[1] = mx;
}

// ---------------------------------------------------------------------------
// bmfrPre: 13 features of every padded pixel to the fp16 feature buffer
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(1024) k_bmfr_pre(BmfrGeom g, pbrt_bmfr_images im) {
    __shared__ float scratch[16];
    const int lx = threadIdx.x % g.bs, ly = threadIdx.x / g.bs;
    const int gx = blockIdx.x * g.bs + lx, gy = blockIdx.y * g.bs + ly;
    PixelIn p = load_pixel(g, im, gx, gy);
    float mn = block_reduce(p.depth, OpMin(), scratch);
    float mx = block_reduce(p.depth, OpMax(), scratch);
    if (threadIdx.x == 0) store_range(g, blockIdx.x, blockIdx.y, mn, mx);
    float z = (p.depth - mn) / (mx - mn + 1e-6f);
    float px = (float)lx / (float)(g.bs - 1), py = (float)ly / (float)(g.bs - 1);
    float f[kAlphaSize] = {1.0f, p.normal.x, p.normal.y, p.normal.z, px, py, z, px * px, py * py, z * z,
                           p.noisy.x, p.noisy.y, p.noisy.z};
    const size_t plane = (size_t)g.wp * g.hp;
    const size_t at = (size_t)gy * g.wp + gx;
#pragma unroll
    for (int i = 0; i < kAlphaSize; ++i) g.features[i * plane + at] = f2h(f[i]);
}

// ---------------------------------------------------------------------------
// bmfrFit: FP32 Householder QR of the (bs^2 x 13) block, reference algorithm
// ---------------------------------------------------------------------------
template <int kSub>
__global__ void __launch_bounds__(256) k_bmfr_fit_householder(BmfrGeom g) {
    __shared__ float scratch[16];
    __shared__ float u_length_squared;
    __shared__ vec3 ws[kFeatures];
    const int id = threadIdx.x;
    const int T = blockDim.x;
    const int bs = g.bs;
    const uint32_t pb = (uint32_t)(bs * bs);
    const int basex = blockIdx.x * bs, basey = blockIdx.y * bs;
    const size_t plane = (size_t)g.wp * g.hp;
    float feat[kSub][kAlphaSize];
    for (int col = 0; col < kAlphaSize; ++col)
        for (int s = 0; s < kSub; ++s) {
            const int index = id + s * T;
            const int px = index / bs + basex, py = index % bs + basey;
            float v = h2f(g.features[col * plane + (size_t)py * g.wp + px]);
            if (col < kFeatures)
                v = add_noise(v, (uint32_t)index + (uint32_t)col * pb * pb + g.frame * (uint32_t)kAlphaSize * pb * pb);
            feat[s][col] = v;
        }
    for (int col = 0; col < kFeatures; ++col) {
        float u[kSub];
        float val2 = 0.0f;
        for (int s = 0; s < kSub; ++s) {
            const int index = id + s * T;
            u[s] = feat[s][col];
            if (index > col) val2 += u[s] * u[s];
        }
        float vec_len_squ = block_reduce(val2, OpAdd(), scratch);
        if (id < col) {
            u[0] = 0.0f;
        } else if (id == col) {
            float vec_len = sqrtf(vec_len_squ + u[0] * u[0]);
            u[0] -= vec_len;
            vec_len_squ += u[0] * u[0];
            u_length_squared = vec_len_squ;
            feat[0][col] = vec_len;
        } else {
            feat[0][col] = 0.0f;
        }
        __syncthreads();
        const float uls = u_length_squared;
        for (int f = col + 1; f < kAlphaSize; ++f) {
            float v = 0.0f;
            for (int s = 0; s < kSub; ++s) {
                const int index = id + s * T;
                if (index >= col) v += feat[s][f] * u[s];
            }
            v = block_reduce(v, OpAdd(), scratch);
            for (int s = 0; s < kSub; ++s) {
                const int index = id + s * T;
                if (index >= col) feat[s][f] -= 2.0f * u[s] * v / uls;
            }
        }
    }
    for (int i = kFeatures - 1; i >= 0; --i) {
        if (id == i) {
            vec3 w = make_vec3(feat[0][10], feat[0][11], feat[0][12]);
            for (int x = i + 1; x < kFeatures; ++x) w -= ws[x] * feat[0][x];
            w /= feat[0][i];
            ws[i] = w;
        }
        __syncthreads();
    }
    if (id < kFeatures) {
        vec3 w = ws[id];
        if (u_length_squared == 0.0f) w = make_vec3(0.2f);
        float* out = g.weights + ((size_t)blockIdx.y * g.blocks_x + blockIdx.x) * 30;
        out[id * 3 + 0] = w.x;
        out[id * 3 + 1] = w.y;
        out[id * 3 + 2] = w.z;
    }
}

// ---------------------------------------------------------------------------
// Fused pre + fit on the matrix cores: FP64 (raw Gram, below) and FP16 (projected
// features split into fp16 hi + lo parts)
// ---------------------------------------------------------------------------
typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float float4v __attribute__((ext_vector_type(4)));
typedef double double4v __attribute__((ext_vector_type(4)));

constexpr int kGramCols = 16;  // 13 columns (FP64: a0, 9 features, rgb; FP16: 9 projected features, rgb) + zeros
constexpr int kFitThreads = 256;

// FP16 fit: wave-private staging of 64 rows (one per lane) as column-major fp16 hi / lo
// planes, the per-wave partial Gram matrices, and the per-wave partial sums of the
// block reductions (each reduction costs one workgroup barrier).
// The 9 x 9 normal equations of the features with the three colours as right-hand
// sides, solved in one wave with column j of the system [G_ff | G_fc] in lane j's
// m[0..8] (lanes 0..8 the features, 9..11 the colours; other lanes carry copies).
// Gaussian elimination without row exchanges: on a symmetric positive semidefinite
// matrix its pivots are the diagonals of the successive Schur complements (the
// squares of the Cholesky diagonal) and it is backward stable as Cholesky is.  A
// pivot at or below 1e-15 x maxdiag drops its feature: it eliminates nothing and its
// weight is 0, which solves the system without that feature.  Every pivot step is
// one division and independent multiply-adds across the lanes (a Cholesky column
// step is a square root, a division and a chain of j dependent products), and the
// back-substitution runs in the three colour lanes at once.  Returns in lanes 9..11
// the weights w[0..8] of colour lane - 9.
__device__ __forceinline__ void solve_feature_block(const double (&m_in)[kFeatures - 1], double maxdiag,
                                                    double (&w)[kFeatures - 1]) {
    constexpr int n = kFeatures - 1;
    double m[n], inv_piv[n];
#pragma unroll
    for (int i = 0; i < n; ++i) m[i] = m_in[i];
#pragma unroll
    for (int p = 0; p < n; ++p) {
        const double piv = read_lane(m[p], p);
        const bool kp = piv > 1e-15 * maxdiag && piv > 0.0;
        const double inv = kp ? 1.0 / piv : 0.0;
        inv_piv[p] = inv;
#pragma unroll
        for (int i = p + 1; i < n; ++i) m[i] = fma(-(read_lane(m[i], p) * inv), m[p], m[i]);
    }
    // m is now the upper triangle U (lane j: U[i][j] for i <= j) and, in the colour
    // lanes, the eliminated right-hand sides; r takes the substitution's updates
    double r[n];
#pragma unroll
    for (int i = 0; i < n; ++i) r[i] = m[i];
#pragma unroll
    for (int p = n - 1; p >= 0; --p) {
        w[p] = r[p] * inv_piv[p];
#pragma unroll
        for (int i = 0; i < p; ++i) r[i] = fma(-read_lane(m[i], p), w[p], r[i]);
    }
}

template <int kWaves>
struct GramShared {
    struct {
        _Float16 hi[kWaves][kGramCols][64];
        _Float16 lo[kWaves][kGramCols][64];
    } h;
    double gram[kWaves][kGramCols * kGramCols];
    float depth_min[kWaves], depth_max[kWaves];
    float coord[32];
    double dots_part[kWaves][kAlphaSize];
    float amax_part[kWaves][12];
    double G[12][12];
};

// LDS written by some lanes of a wave and then read by others of the same wave.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// kThreads: workgroup size (256, 512 or 1024); its waves split the block's rows,
// so the thread count sets how the rows are tiled over the matrix cores
// (pbrt_autotune_bmfr picks it).  Thread t loads pixels t, t + kThreads, ... of
// the block in row-major order (x fastest: coalesced); the pixel at block-local
// (x, y) is fit row i = x * bs + y, whose index keys the noise.
// The FP16 fit: the columns are projected against the noised constant column a0 in
// FP64, rounded to fp32, scaled by powers of two and split into fp16 hi + lo parts
// whose Gram matrix v_mfma_f32_16x16x32_f16 accumulates (FP32 inside a 32-row chunk,
// FP64 across chunks); FP64 elimination then solves for w1..w9 and w0 = a0.(y - A'w') / a0.a0.
template <int kRows, int kThreads>
__device__ __forceinline__ void bmfr_fit_gram_f16_block(const BmfrGeom& g, const pbrt_bmfr_images& im, int bx,
                                                        int by) {
    constexpr int kSub = (kRows + kThreads - 1) / kThreads;
    constexpr int kWaves = kThreads / 64;
    __shared__ GramShared<kWaves> S;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int bs = g.bs;
    const int basex = bx * bs, basey = by * bs;

    // 1. load pixels (normal and noisy colour straight into their feature slots);
    //    block min/max of depth
    float a[kSub][kAlphaSize];
    float depth[kSub];
    int row[kSub];
    float dmin = 3.4e38f, dmax = -3.4e38f;
#pragma unroll
    for (int s = 0; s < kSub; ++s) {
        const int q = tid + s * kThreads;
        row[s] = -1;
        if (q < kRows) {
            const int qx = q % bs, qy = q / bs;
            row[s] = qx * bs + qy;
            const PixelIn p = load_pixel(g, im, basex + qx, basey + qy);
            depth[s] = p.depth;
            a[s][1] = p.normal.x;
            a[s][2] = p.normal.y;
            a[s][3] = p.normal.z;
            a[s][10] = p.noisy.x;
            a[s][11] = p.noisy.y;
            a[s][12] = p.noisy.z;
            dmin = fminf(dmin, p.depth);
            dmax = fmaxf(dmax, p.depth);
        }
    }
    dmin = wave_reduce_dpp(dmin, OpMin());
    dmax = wave_reduce_dpp(dmax, OpMax());
    if (lane == 0) {
        S.depth_min[wave] = dmin;
        S.depth_max[wave] = dmax;
    }
    // block coordinates i / (bs - 1) (bmfrPre.comp:36-43), one IEEE division each
    if (tid < bs) S.coord[tid] = (float)tid / (float)(bs - 1);
    __syncthreads();
    float mn = S.depth_min[0], mx = S.depth_max[0];
#pragma unroll
    for (int w = 1; w < kWaves; ++w) {
        mn = fminf(mn, S.depth_min[w]);
        mx = fmaxf(mx, S.depth_max[w]);
    }
    if (tid == 0) store_range(g, bx, by, mn, mx);
    // this thread's noise values, kSub consecutive floats per feature
    float noise[kFeatures][kSub];
#pragma unroll
    for (int c = 0; c < kFeatures; ++c) {
        const float* src = g.noise + c * kRows + tid * kSub;
        if (tid * kSub < kRows) {
            if constexpr (kSub == 4) {
                const float4 v = *reinterpret_cast<const float4*>(src);
                noise[c][0] = v.x;
                noise[c][1] = v.y;
                noise[c][2] = v.z;
                noise[c][3] = v.w;
            } else {
#pragma unroll
                for (int s = 0; s < kSub; ++s) noise[c][s] = src[s];
            }
        }
    }
    // 2. fp16-quantised features (the reference's R16F feature buffer) + noise;
    //    FP64 sums a0.a_c with the noised constant column a0
    double d_local[kAlphaSize];
#pragma unroll
    for (int c = 0; c < kAlphaSize; ++c) d_local[c] = 0.0;
#pragma unroll
    for (int s = 0; s < kSub; ++s) {
        if (row[s] < 0) continue;
        const int r = row[s];
        const int lx = r / bs, ly = r % bs;
        float z = (depth[s] - mn) / (mx - mn + 1e-6f);
        const float fx = S.coord[lx], fy = S.coord[ly];
        float f[kAlphaSize] = {1.0f, a[s][1], a[s][2], a[s][3], fx, fy, z, fx * fx, fy * fy,
                               z * z, a[s][10], a[s][11], a[s][12]};
#pragma unroll
        for (int c = 0; c < kAlphaSize; ++c) {
            float v = h2f(f2h(f[c]));
            if (c < kFeatures) v = __fadd_rn(v, noise[c][s]);
            a[s][c] = v;
        }
        const double a0 = a[s][0];
#pragma unroll
        for (int c = 0; c < kAlphaSize; ++c) d_local[c] += a0 * (double)a[s][c];
    }
#pragma unroll
    for (int c = 0; c < kAlphaSize; ++c) d_local[c] = wave_reduce_dpp(d_local[c], OpAdd());
    if (lane == 0) {
#pragma unroll
        for (int c = 0; c < kAlphaSize; ++c) S.dots_part[wave][c] = d_local[c];
    }
    __syncthreads();
    // block sum c, folded in wave order (every thread gets the same value)
    auto dot = [&](int c) {
        double t = S.dots_part[0][c];
#pragma unroll
        for (int w = 1; w < kWaves; ++w) t += S.dots_part[w][c];
        return t;
    };
    // 3. project the columns against a0 (FP64), keep fp32; block max-abs per column
    const double inv_a0 = 1.0 / dot(0);
    float amax[12];
#pragma unroll
    for (int c = 0; c < 12; ++c) {
        const double coef = dot(c + 1) * inv_a0;
        amax[c] = 0.0f;
#pragma unroll
        for (int s = 0; s < kSub; ++s) {
            if (row[s] < 0) continue;
            const float b = (float)((double)a[s][c + 1] - coef * (double)a[s][0]);
            a[s][c + 1] = b;
            amax[c] = fmaxf(amax[c], fabsf(b));
        }
    }
#pragma unroll
    for (int c = 0; c < 12; ++c) amax[c] = wave_reduce_dpp(amax[c], OpMax());
    if (lane == 0) {
#pragma unroll
        for (int c = 0; c < 12; ++c) S.amax_part[wave][c] = amax[c];
    }
    __syncthreads();
    float scale[12];
#pragma unroll
    for (int c = 0; c < 12; ++c) {
        float m = S.amax_part[0][c];
#pragma unroll
        for (int w = 1; w < kWaves; ++w) m = fmaxf(m, S.amax_part[w][c]);
        int e = 0;
        frexpf(m, &e);
        scale[c] = m > 0.0f ? ldexpf(1.0f, -e) : 1.0f;  // |scaled| < 1, exact power of two
    }
    // 4. Gram on the matrix cores, 64 rows at a time through the wave's own LDS
    //    slice (lane l stages its row of chunk s).  For both instructions the A and
    //    B operands of a Gram product hold the same values (lane l: column l % 16
    //    of a row group).
    {
        const int col = lane & 15, grp = lane >> 4;
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s = 0; s < kSub; ++s) {
            if (s > 0) wave_lds_sync();  // the previous chunk has been read
#pragma unroll
            for (int c = 0; c < kGramCols; ++c) {
                const float v = (c < 12 && row[s] >= 0) ? a[s][c + 1] * scale[c] : 0.0f;
                const _Float16 hi = (_Float16)v;
                S.h.hi[wave][c][lane] = hi;
                S.h.lo[wave][c][lane] = (_Float16)(v - (float)hi);
            }
            wave_lds_sync();
#pragma unroll
            for (int ch = 0; ch < 2; ++ch) {
                const int row0 = ch * 32 + grp * 8;
                half8 hi, lo;
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    hi[j] = S.h.hi[wave][col][row0 + j];
                    lo[j] = S.h.lo[wave][col][row0 + j];
                }
                float4v e = {0.0f, 0.0f, 0.0f, 0.0f};
                e = __builtin_amdgcn_mfma_f32_16x16x32_f16(hi, hi, e, 0, 0, 0);
                e = __builtin_amdgcn_mfma_f32_16x16x32_f16(hi, lo, e, 0, 0, 0);
                e = __builtin_amdgcn_mfma_f32_16x16x32_f16(lo, hi, e, 0, 0, 0);
#pragma unroll
                for (int i = 0; i < 4; ++i) acc[i] += (double)e[i];
            }
        }
        // register i of v_mfma_f32_16x16x32_f16's result is row 4 * grp + i, column `col`
#pragma unroll
        for (int i = 0; i < 4; ++i) S.gram[wave][(4 * grp + i) * kGramCols + col] = acc[i];
    }
    __syncthreads();
    // 5. the symmetrised, unscaled 12 x 12 Gram matrix (9 projected features, 3 colours)
    constexpr int pw = kGramCols * kGramCols;
    for (int e = tid; e < 12 * 12; e += kThreads) {
        const int i = e / 12, j = e % 12;
        const int ia = i * kGramCols + j, ib = j * kGramCols + i;
        const double* gr = &S.gram[0][0];
        double gij = 0.0, gji = 0.0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            gij += gr[w * pw + ia];
            gji += gr[w * pw + ib];
        }
        S.G[i][j] = 0.5 * (gij + gji) / ((double)scale[i] * (double)scale[j]);
    }
    __syncthreads();
    if (wave != 0) return;
    // 6. FP64 solve of the 9 x 9 feature block in wave 0 (a pivot below 1e-15 of
    //    the largest diagonal drops its feature), then in the colour lanes
    //    w0 = a0.(y - A'w') / a0.a0.
    constexpr int n = kFeatures - 1;
    double maxdiag = 0.0;
#pragma unroll
    for (int i = 0; i < n; ++i) maxdiag = fmax(maxdiag, S.G[i][i]);
    const int lj = lane < 12 ? lane : 0;
    double m[n], w[n];
#pragma unroll
    for (int i = 0; i < n; ++i) m[i] = S.G[i][lj];
    solve_feature_block(m, maxdiag, w);
    if (lane >= n && lane < n + 3) {
        const int c = lane - n;
        float* out = g.weights + ((size_t)by * g.blocks_x + bx) * 30;
        double w0 = dot(10 + c);
#pragma unroll
        for (int i = 0; i < n; ++i) w0 -= w[i] * dot(i + 1);
        w0 *= inv_a0;
        out[c] = (float)w0;
#pragma unroll
        for (int i = 0; i < n; ++i) out[(i + 1) * 3 + c] = (float)w[i];
    }
}

// Occupancy: the F64 fit of 256-thread workgroups runs five waves per SIMD (at most
// 96 VGPRs, five blocks per CU: 27 KB of LDS each), the other fits four (at most
// 128); three blocks of 256 threads per CU left the per-block latency (load,
// reductions, solve) uncovered, and with four a block whose wave 0 is still solving
// keeps the CU from starting the next block.
// The FP64 fit.  The Gram matrix R of the raw columns [a0 | a1..a9 | r g b] (fp32
// values: their products are exact in FP64) is accumulated on v_mfma_f64_16x16x4f64;
// the normal equations of the features projected against the noised constant column
// a0 follow in FP64 by the rank-1 downdate G = R[1:,1:] - R[0,1:] R[0,1:]^T / R[0,0].
// The 9 x 9 feature block of G is equilibrated by its diagonal and solved for the three
// colours by elimination (a pivot below 1e-15 drops its feature);
// w0 = (R[0,10+c] - sum_i w_i R[0,1+i]) / R[0,0].  Against FP64 least squares on the
// same noised blocks this is more accurate than projecting each row first and
// rounding the projected columns to fp32 (fitted values within ~1e-8 instead of
// ~5e-6 on near-collinear planar blocks), and it needs no per-row FP64 work and no
// block reductions beyond the depth range.
template <int kWaves>
struct RawGramShared {
    float st[kWaves][64][kGramCols + 1];  // wave-private rows (odd stride: no bank conflicts)
    double gram[kWaves][kGramCols * kGramCols];
    float depth_min[kWaves], depth_max[kWaves];
    float coord[32];
    double r0[kAlphaSize];  // row 0 of R: a0 . a_c
    double G[12][12];       // projected normal equations (9 features | 3 colours)
    double eq[kFeatures - 1];  // equilibration 1 / sqrt(G_ii)
};

template <int kRows, int kThreads>
__device__ __forceinline__ void bmfr_fit_raw_gram_block(const BmfrGeom& g, const pbrt_bmfr_images& im, int bx,
                                                        int by) {
    constexpr int kSub = (kRows + kThreads - 1) / kThreads;
    constexpr int kWaves = kThreads / 64;
    __shared__ RawGramShared<kWaves> S;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int bs = g.bs;
    const int basex = bx * bs, basey = by * bs;

    // 1. load pixels; block min/max of depth
    float a[kSub][kAlphaSize];
    float depth[kSub];
    int row[kSub];
    float dmin = 3.4e38f, dmax = -3.4e38f;
#pragma unroll
    for (int s = 0; s < kSub; ++s) {
        const int q = tid + s * kThreads;
        row[s] = -1;
        if (q < kRows) {
            const int qx = q % bs, qy = q / bs;
            row[s] = qx * bs + qy;
            const PixelIn p = load_pixel(g, im, basex + qx, basey + qy);
            depth[s] = p.depth;
            a[s][1] = p.normal.x;
            a[s][2] = p.normal.y;
            a[s][3] = p.normal.z;
            a[s][10] = p.noisy.x;
            a[s][11] = p.noisy.y;
            a[s][12] = p.noisy.z;
            dmin = fminf(dmin, p.depth);
            dmax = fmaxf(dmax, p.depth);
        }
    }
    dmin = wave_reduce_dpp(dmin, OpMin());
    dmax = wave_reduce_dpp(dmax, OpMax());
    if (lane == 0) {
        S.depth_min[wave] = dmin;
        S.depth_max[wave] = dmax;
    }
    if (tid < bs) S.coord[tid] = (float)tid / (float)(bs - 1);  // bmfrPre.comp:36-43
    __syncthreads();
    float mn = S.depth_min[0], mx = S.depth_max[0];
#pragma unroll
    for (int w = 1; w < kWaves; ++w) {
        mn = fminf(mn, S.depth_min[w]);
        mx = fmaxf(mx, S.depth_max[w]);
    }
    if (tid == 0) store_range(g, bx, by, mn, mx);
    // 2. features, fp16-quantised (the reference's R16F feature buffer), + noise; the
    //    noise of feature c for this thread's pixels is one vector load
#pragma unroll
    for (int s = 0; s < kSub; ++s) {
        const int r = row[s] < 0 ? 0 : row[s];
        const float z = (depth[s] - mn) / (mx - mn + 1e-6f);
        const float fx = S.coord[r / bs], fy = S.coord[r % bs];
        a[s][0] = 1.0f;
        a[s][4] = fx;
        a[s][5] = fy;
        a[s][6] = z;
        a[s][7] = fx * fx;
        a[s][8] = fy * fy;
        a[s][9] = z * z;
    }
#pragma unroll
    for (int c = 0; c < kAlphaSize; ++c) {
        float nz[kSub];
        if (c < kFeatures && tid * kSub < kRows) {
            const float* src = g.noise + c * kRows + tid * kSub;
            if constexpr (kSub == 4) {
                const float4 v = *reinterpret_cast<const float4*>(src);
                nz[0] = v.x;
                nz[1] = v.y;
                nz[2] = v.z;
                nz[3] = v.w;
            } else {
#pragma unroll
                for (int s = 0; s < kSub; ++s) nz[s] = src[s];
            }
        }
#pragma unroll
        for (int s = 0; s < kSub; ++s) {
            float v = h2f(f2h(a[s][c]));
            if (c < kFeatures) v = __fadd_rn(v, nz[s]);
            a[s][c] = v;
        }
    }
    // 3. R = A^T A on the matrix cores, 64 rows at a time through the wave's own LDS
    //    slice (lane l stages its row of chunk s); both operands of the product hold
    //    the same values (lane l: column l % 16 of a group of four rows)
    {
        const int col = lane & 15, grp = lane >> 4;
        // two accumulators: even and odd row groups form independent MFMA chains
        double4v d = {0.0, 0.0, 0.0, 0.0}, d1 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s = 0; s < kSub; ++s) {
            if (s > 0) wave_lds_sync();  // the previous chunk has been read
#pragma unroll
            for (int c = 0; c < kGramCols; ++c) S.st[wave][lane][c] = (c < kAlphaSize && row[s] >= 0) ? a[s][c] : 0.0f;
            wave_lds_sync();
#pragma unroll
            for (int ch = 0; ch < 16; ch += 2) {
                const double x0 = (double)S.st[wave][ch * 4 + grp][col];
                const double x1 = (double)S.st[wave][ch * 4 + 4 + grp][col];
                d = __builtin_amdgcn_mfma_f64_16x16x4f64(x0, x0, d, 0, 0, 0);
                d1 = __builtin_amdgcn_mfma_f64_16x16x4f64(x1, x1, d1, 0, 0, 0);
            }
        }
        d += d1;
        // register i of v_mfma_f64_16x16x4f64's result is row grp + 4 * i, column `col`
#pragma unroll
        for (int i = 0; i < 4; ++i) S.gram[wave][(grp + 4 * i) * kGramCols + col] = d[i];
    }
    __syncthreads();
    // 4. R summed over the waves and symmetrised; G by the rank-1 downdate
    constexpr int pw = kGramCols * kGramCols;
    const double* gr = &S.gram[0][0];
    auto rsum = [&](int i, int j) {
        double a_ij = 0.0, a_ji = 0.0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            a_ij += gr[w * pw + i * kGramCols + j];
            a_ji += gr[w * pw + j * kGramCols + i];
        }
        return 0.5 * (a_ij + a_ji);
    };
    const double r00 = rsum(0, 0), inv_r00 = 1.0 / r00;
    for (int e = tid; e < 12 * 12 + kAlphaSize; e += kThreads) {
        if (e < 12 * 12) {
            const int i = e / 12 + 1, j = e % 12 + 1;
            S.G[i - 1][j - 1] = rsum(i, j) - rsum(0, i) * rsum(0, j) * inv_r00;
        } else {
            S.r0[e - 12 * 12] = rsum(0, e - 12 * 12);
        }
    }
    __syncthreads();
    if (wave != 0) return;
    // 5. the 9 x 9 feature block equilibrated by its diagonal (the colour columns
    //    by the feature rows' factor only), solved in wave 0
    constexpr int n = kFeatures - 1;
    const int lj = lane < 12 ? lane : 0;
    const double gjj = S.G[lj][lj];
    const double eqj = lj < n ? (gjj > 0.0 ? 1.0 / sqrt(gjj) : 0.0) : 1.0;
    if (lane < n) S.eq[lane] = eqj;
    wave_lds_sync();
    double maxdiag = 0.0, m[n], w[n];
#pragma unroll
    for (int i = 0; i < n; ++i) {
        const double e = S.eq[i];
        maxdiag = fmax(maxdiag, S.G[i][i] * e * e);
        m[i] = S.G[i][lj] * e * eqj;
    }
    solve_feature_block(m, maxdiag, w);
    // 6. the colour lanes: unequilibrate, w0
    if (lane >= n && lane < n + 3) {
        const int c = lane - n;
        float* out = g.weights + ((size_t)by * g.blocks_x + bx) * 30;
        double w0 = S.r0[10 + c];
#pragma unroll
        for (int i = 0; i < n; ++i) {
            w[i] *= S.eq[i];
            w0 -= w[i] * S.r0[i + 1];
        }
        w0 /= r00;
        out[c] = (float)w0;
#pragma unroll
        for (int i = 0; i < n; ++i) out[(i + 1) * 3 + c] = (float)w[i];
    }
}

// Workgroup b of a launch of n runs on XCD b % 8 (the dispatcher deals workgroups
// out round-robin).  The workgroups of XCD x take the x-th contiguous share of the
// block order instead, so horizontally neighbouring blocks, whose grid is shifted by
// the frame's pixel offset and whose rows therefore share cache lines at their
// ends, are read through one XCD's L2.  A permutation of [0, n).
__device__ __forceinline__ uint32_t xcd_contiguous(uint32_t b, uint32_t n) {
    const uint32_t x = b % 8u, s = b / 8u, q = n / 8u, r = n % 8u;
    return x * q + min(x, r) + s;
}

// block (bx, by) of the 2D launch grid over the blocks_x x blocks_y BMFR blocks
__device__ __forceinline__ void grid_block(const BmfrGeom& g, int& bx, int& by) {
    const uint32_t n = gridDim.x * gridDim.y;
    const uint32_t b = xcd_contiguous(blockIdx.x + blockIdx.y * gridDim.x, n);
    bx = (int)(b % (uint32_t)g.blocks_x);
    by = (int)(b / (uint32_t)g.blocks_x);
}

template <int kRows, bool kF64, int kThreads>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(kF64 && kThreads == 256 ? 5 : 4)))
k_bmfr_fit_gram(BmfrGeom g, pbrt_bmfr_images im) {
    int bx, by;
    grid_block(g, bx, by);
    if constexpr (kF64)
        bmfr_fit_raw_gram_block<kRows, kThreads>(g, im, bx, by);
    else
        bmfr_fit_gram_f16_block<kRows, kThreads>(g, im, bx, by);
}

// ---------------------------------------------------------------------------
// bmfrPost: weighted feature sum, temporal blend, remodulation, gamma
// ---------------------------------------------------------------------------
// x^0.454545 for the gamma of bmfrPost.comp:122 as exp2(0.454545 * log2 x) on the
// hardware log2 / exp2: within a few fp32 ulps of powf, and the result is stored as
// unorm8.  x = 0 gives exp2(-inf) = 0.
__device__ __forceinline__ float gamma_454545(float x) {
    return __builtin_amdgcn_exp2f(0.454545f * __builtin_amdgcn_logf(x));
}


// Rows per workgroup of k_bmfr_post: a block's rows are cut into workgroups of at
// most 256 threads.
__host__ __device__ constexpr int post_rows(int bs) { return bs * bs > 256 ? 256 / bs : bs; }

// bmfrPost per pixel: the block's depth range is the one its fit recorded, so a
// workgroup needs no reduction and no barrier; workgroup w of the launch takes
// rows [row0, row0 + post_rows) of one block (XCD-contiguous order, as the fit),
// one thread per padded pixel, and the padding pixels leave at once.  The weights
// are uniform over the workgroup: scalar loads, and a non-finite weight (an
// exponent of all ones) becomes 0 with integer operations.
// Pixel (lx, ly) of block blk of a frame (the per-thread part of both post launches).
template <int kBs>
__device__ __forceinline__ void bmfr_post_pixel(const BmfrGeom& g, const pbrt_bmfr_images& im, uint32_t blk, int lx,
                                                int ly) {
    const int bx = (int)(blk % (uint32_t)g.blocks_x), by = (int)(blk / (uint32_t)g.blocks_x);
    const int gx = bx * kBs + lx, gy = by * kBs + ly;
    const int ax = gx - g.offx, ay = gy - g.offy;
    if (ax != mirror_coord(ax, g.width) || ay != mirror_coord(ay, g.height)) return;
    const PixelIn p = load_pixel(g, im, gx, gy);
    const size_t pix = (size_t)p.img_y * g.width + p.img_x;
    const float mn = g.ranges[2 * blk], mx = g.ranges[2 * blk + 1];
    float wts[30];
#pragma unroll
    for (int i = 0; i < 30; ++i) {
        const uint32_t u = __float_as_uint(g.weights[(size_t)blk * 30 + i]);
        wts[i] = __uint_as_float((u & 0x7F800000u) == 0x7F800000u ? 0u : u);
    }
    const ushort2 mv = reinterpret_cast<const ushort2*>(im.motion)[pix];
    const float mux = h2f(mv.x), muy = h2f(mv.y);
    const bool accept = mux >= 0.0f;
    const float pixel_spp = unorm8_to_float(im.spp[pix]) * 256.0f;
    float z = (p.depth - mn) / (mx - mn + 1e-6f);
    float fx = div_small<kBs - 1>((float)lx), fy = div_small<kBs - 1>((float)ly);
    float f[kFeatures] = {1.0f, p.normal.x, p.normal.y, p.normal.z, fx, fy, z, fx * fx, fy * fy, z * z};
    vec3 den = make_vec3(0.0f);
#pragma unroll
    for (int i = 0; i < kFeatures; ++i) den += make_vec3(wts[3 * i], wts[3 * i + 1], wts[3 * i + 2]) * f[i];
    den = vclamp(den, 0.0f, 10.0f);
    vec3 prev = make_vec3(0.0f);
    float alpha = 1.0f;
    const size_t plane = (size_t)g.width * g.height;
    if (g.frame > 0 && accept) {
        const uint16_t* layer = im.denoised + 4 * plane * (g.frame & 1u);
        const int wd = g.width;
        prev = xyz(sample_bilinear([&](int sx, int sy) { return load_half4(layer, (size_t)sy * wd + sx); }, g.width,
                                   g.height, mux, muy));
        alpha = fmaxf(1.0f / pixel_spp, 0.1f);
    }
    den = alpha * den + (1.0f - alpha) * prev;
    store_half4(im.denoised + 4 * plane * ((g.frame & 1u) ^ 1u), pix, make_vec4(den, 1.0f));
    vec3 alb = xyz(load_rgba8(im.albedo, pix)) + make_vec3(1e-6f);
    vec3 t = alb * den;
    t = make_vec3(fmaxf(t.x, 0.0f), fmaxf(t.y, 0.0f), fmaxf(t.z, 0.0f));
    t = vclamp(make_vec3(gamma_454545(t.x), gamma_454545(t.y), gamma_454545(t.z)), 0.0f, 1.0f);
    store_bgra8(im.final_image, pix, make_vec4(t, 1.0f));
}

// kBs: the block size (8, 16 or 32), so that the pixel's place in its block is a
// shift and a mask of the thread index
template <int kBs>
__global__ void __launch_bounds__(256) k_bmfr_post(BmfrGeom g, pbrt_bmfr_images im) {
    constexpr uint32_t rows = post_rows(kBs), tiles = kBs / rows;
    const uint32_t w = xcd_contiguous(blockIdx.x, gridDim.x);
    bmfr_post_pixel<kBs>(g, im, w / tiles, threadIdx.x % kBs, (int)((w % tiles) * rows + threadIdx.x / kBs));
}

// ---------------------------------------------------------------------------
// Batched BMFR: one fit launch and one post launch over the blocks of many frames.
// Block b of the launch belongs to the last frame whose first_block <= b.
// ---------------------------------------------------------------------------
struct BmfrBatchFrame {
    BmfrGeom g;
    pbrt_bmfr_images im;
    uint32_t first_block;
};

__device__ __forceinline__ uint32_t batch_frame_of(const BmfrBatchFrame* frames, uint32_t n, uint32_t b) {
    uint32_t lo = 0, hi = n - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (frames[mid].first_block <= b)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

template <int kRows, bool kF64>
__global__ void __launch_bounds__(kFitThreads) __attribute__((amdgpu_waves_per_eu(kF64 ? 5 : 4)))
k_bmfr_fit_gram_batched(const BmfrBatchFrame* frames, uint32_t n) {
    const uint32_t b = xcd_contiguous(blockIdx.x, gridDim.x);
    const uint32_t f = batch_frame_of(frames, n, b);
    const BmfrGeom g = frames[f].g;
    const pbrt_bmfr_images im = frames[f].im;
    const uint32_t local = b - frames[f].first_block;
    const int bx = (int)(local % (uint32_t)g.blocks_x), by = (int)(local / (uint32_t)g.blocks_x);
    if constexpr (kF64)
        bmfr_fit_raw_gram_block<kRows, kFitThreads>(g, im, bx, by);
    else
        bmfr_fit_gram_f16_block<kRows, kFitThreads>(g, im, bx, by);
}

// the noise table of every frame of a batch (blockIdx.y: frame)
__global__ void __launch_bounds__(256) k_bmfr_noise_batched(const BmfrBatchFrame* frames) {
    const BmfrGeom& g = frames[blockIdx.y].g;
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e < kFeatures * g.bs * g.bs) noise_table_entry(g.frame, g.bs, kFitThreads, e, g.noise);
}

template <int kBs>  // one block size per batch
__global__ void __launch_bounds__(256) k_bmfr_post_batched(const BmfrBatchFrame* frames, uint32_t n) {
    constexpr uint32_t rows = post_rows(kBs), tiles = kBs / rows;
    const uint32_t w = xcd_contiguous(blockIdx.x, gridDim.x);
    const uint32_t b = w / tiles;
    const uint32_t f = batch_frame_of(frames, n, b);
    bmfr_post_pixel<kBs>(frames[f].g, frames[f].im, b - frames[f].first_block, threadIdx.x % kBs,
                         (int)((w % tiles) * rows + threadIdx.x / kBs));
}

// pixelOffsets of bmfrGeneral.comp:16
static const float kPixelOffsets[16][2] = {{.7f, .85f},  {.95f, .5f},  {.43f, .76f}, {.97f, .03f},
                                           {.37f, .58f}, {.03f, .36f}, {.81f, .46f}, {0.f, .78f},
                                           {.36f, -.08f}, {-.06f, 0.f}, {.95f, .1f},  {.85f, .61f},
                                           {.06f, .1f},  {.43f, .16f}, {0.f, .5f},   {.73f, .38f}};

static BmfrGeom make_geom(const pbrt_bmfr_params& p) {
    BmfrGeom g{};
    PBRT_REQUIRE(p.block_size == 8 || p.block_size == 16 || p.block_size == 32, "pbrt_bmfr: block_size must be 8, 16 or 32");
    PBRT_REQUIRE(p.width >= 2 * p.block_size && p.height >= 2 * p.block_size,
                 "pbrt_bmfr: image must be at least two blocks wide and high (mirror addressing)");
    PBRT_REQUIRE(p.fit == PBRT_BMFR_FIT_HOUSEHOLDER || p.fit == PBRT_BMFR_FIT_GRAM_F16 ||
                     p.fit == PBRT_BMFR_FIT_GRAM_F64,
                 "pbrt_bmfr: unknown fit");
    PBRT_REQUIRE(p.fit_threads == 0 || p.fit_threads == 256 || p.fit_threads == 512 || p.fit_threads == 1024,
                 "pbrt_bmfr: fit_threads must be 0, 256, 512 or 1024");
    g.width = (int)p.width;
    g.height = (int)p.height;
    g.bs = (int)p.block_size;
    g.wp = (g.width / g.bs + 2) * g.bs;
    g.hp = (g.height / g.bs + 2) * g.bs;
    // mirror() reflects once: the padded grid shifted by any frame's pixelOffsets
    // (the most negative shift is trunc(-0.06 bs) in x, trunc(-0.08 bs) in y) must
    // land within one image size of [0, size).  A side of exactly two 16- or
    // 32-pixel blocks would reach 2 * size and read outside the image.
    int min_offx = 0, min_offy = 0;
    for (const auto& o : kPixelOffsets) {
        min_offx = std::min(min_offx, (int)((float)g.bs * o[0]));
        min_offy = std::min(min_offy, (int)((float)g.bs * o[1]));
    }
    PBRT_REQUIRE(g.wp - min_offx <= 2 * g.width && g.hp - min_offy <= 2 * g.height,
                 "pbrt_bmfr: a side of exactly two blocks is too small for the shifted, mirrored block grid "
                 "(use at least two blocks and one pixel)");
    g.blocks_x = g.wp / g.bs;
    g.blocks_y = g.hp / g.bs;
    const float* o = kPixelOffsets[p.frame_number % 16];
    g.offx = (int)((float)g.bs * o[0]);
    g.offy = (int)((float)g.bs * o[1]);
    g.frame = p.frame_number;
    g.fit_threads = g.bs == 8 ? 64 : 256;
    return g;
}

// The Gram fit for a block size and workgroup size (256, 512 or 1024 threads).
template <bool kF64, int kThreads>
static void launch_gram_rows(int bs, dim3 grid, hipStream_t st, const BmfrGeom& g, const pbrt_bmfr_images& im) {
    if (bs == 32)
        hipLaunchKernelGGL((k_bmfr_fit_gram<1024, kF64, kThreads>), grid, dim3(kThreads), 0, st, g, im);
    else if (bs == 16)
        hipLaunchKernelGGL((k_bmfr_fit_gram<256, kF64, kThreads>), grid, dim3(kThreads), 0, st, g, im);
    else
        hipLaunchKernelGGL((k_bmfr_fit_gram<64, kF64, kThreads>), grid, dim3(kThreads), 0, st, g, im);
}

template <bool kF64>
static void launch_gram_threads(int bs, uint32_t threads, dim3 grid, hipStream_t st, const BmfrGeom& g,
                                const pbrt_bmfr_images& im) {
    if (threads == 1024)
        launch_gram_rows<kF64, 1024>(bs, grid, st, g, im);
    else if (threads == 512)
        launch_gram_rows<kF64, 512>(bs, grid, st, g, im);
    else
        launch_gram_rows<kF64, 256>(bs, grid, st, g, im);
}

static void launch_gram_fit(bool f64, int bs, uint32_t threads, dim3 grid, hipStream_t st, const BmfrGeom& g,
                            const pbrt_bmfr_images& im) {
    if (f64)
        launch_gram_threads<true>(bs, threads, grid, st, g, im);
    else
        launch_gram_threads<false>(bs, threads, grid, st, g, im);
}

static size_t feature_bytes(const BmfrGeom& g) { return (size_t)kAlphaSize * g.wp * g.hp * sizeof(uint16_t); }
static size_t weight_bytes(const BmfrGeom& g) { return (size_t)g.blocks_x * g.blocks_y * 30 * sizeof(float); }
static size_t noise_bytes(const BmfrGeom& g) { return (size_t)kFeatures * g.bs * g.bs * sizeof(float); }
static size_t range_bytes(const BmfrGeom& g) { return (size_t)g.blocks_x * g.blocks_y * 2 * sizeof(float); }
static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

static void require_bmfr_images(const pbrt_bmfr_images* im, const char* who) {
    PBRT_REQUIRE(im->depth && im->normal && im->albedo && im->motion && im->spp && im->noisy && im->denoised &&
                     im->final_image,
                 std::string(who) + ": missing image");
}

// pre + fit + post of one frame with the feature array (Householder fit only), the
// weights and the noise table (Gram fits only) at the given device addresses.
static void launch_bmfr(const pbrt_bmfr_params& p, const pbrt_bmfr_images& im, BmfrGeom g, uint16_t* features,
                        float* weights, float* noise, float* ranges, hipStream_t st) {
    g.features = features;
    g.weights = weights;
    g.noise = noise;
    g.ranges = ranges;
    dim3 grid(g.blocks_x, g.blocks_y);
    const int threads = g.bs * g.bs;
    if (p.fit == PBRT_BMFR_FIT_HOUSEHOLDER) {
        hipLaunchKernelGGL(k_bmfr_pre, grid, dim3(threads), 0, st, g, im);
        if (g.bs == 32)
            hipLaunchKernelGGL(k_bmfr_fit_householder<4>, grid, dim3(256), 0, st, g);
        else if (g.bs == 16)
            hipLaunchKernelGGL(k_bmfr_fit_householder<1>, grid, dim3(256), 0, st, g);
        else
            hipLaunchKernelGGL(k_bmfr_fit_householder<1>, grid, dim3(64), 0, st, g);
    } else {
        const int entries = kFeatures * g.bs * g.bs;
        const int threads = (int)(p.fit_threads ? p.fit_threads : kFitThreads);
        hipLaunchKernelGGL(k_bmfr_noise, dim3((entries + 255) / 256), dim3(256), 0, st, g.frame, g.bs, threads,
                           g.noise);
        launch_gram_fit(p.fit == PBRT_BMFR_FIT_GRAM_F64, g.bs, (uint32_t)threads, grid, st, g, im);
    }
    const int rows = post_rows(g.bs);
    const dim3 post_grid(g.blocks_x * g.blocks_y * (g.bs / rows)), post_block(g.bs * rows);
    if (g.bs == 32)
        hipLaunchKernelGGL(k_bmfr_post<32>, post_grid, post_block, 0, st, g, im);
    else if (g.bs == 16)
        hipLaunchKernelGGL(k_bmfr_post<16>, post_grid, post_block, 0, st, g, im);
    else
        hipLaunchKernelGGL(k_bmfr_post<8>, post_grid, post_block, 0, st, g, im);
    PBRT_HIP(hipGetLastError());
}

// Scratch of pbrt_bmfr_noalloc: a file-scope device array, one copy per GPU, held
// until the process ends.  Host code never passes `scratch` itself to a kernel (on
// the host it names a stand-in, not device memory): it uses the device address
// hipGetSymbolAddress returns for the current device, looked up once per device.
constexpr size_t kNoallocScratchBytes = PBRT_BMFR_NOALLOC_SCRATCH_BYTES;
__device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

static char* device_scratch() {
    int dev = 0;
    PBRT_HIP(hipGetDevice(&dev));
    PBRT_REQUIRE(dev >= 0 && dev < kMaxDevices, "pbrt_bmfr_noalloc: device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        PBRT_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        PBRT_REQUIRE(ptr != nullptr, "pbrt_bmfr_noalloc: no device address for the scratch array");
        g_scratch_ptr[dev] = ptr;
    }
    return static_cast<char*>(g_scratch_ptr[dev]);
}

}  // namespace pbrt

using namespace pbrt;

extern "C" int32_t pbrt_bmfr_workspace_size(const pbrt_bmfr_params* p, size_t* out) {
    return guard([&] {
        PBRT_REQUIRE(p && out, "pbrt_bmfr_workspace_size: NULL argument");
        BmfrGeom g = make_geom(*p);
        *out = align256(feature_bytes(g)) + align256(weight_bytes(g)) + align256(noise_bytes(g)) + range_bytes(g);
    });
}

extern "C" int32_t pbrt_bmfr(const pbrt_bmfr_params* p, const pbrt_bmfr_images* im, void* stream) {
    return guard([&] {
        PBRT_REQUIRE(p && im, "pbrt_bmfr: NULL argument");
        require_bmfr_images(im, "pbrt_bmfr");
        PBRT_REQUIRE(im->workspace, "pbrt_bmfr: missing workspace");
        BmfrGeom g = make_geom(*p);
        const size_t fb = align256(feature_bytes(g)), wb = align256(weight_bytes(g)), nb = align256(noise_bytes(g));
        PBRT_REQUIRE(im->workspace_bytes >= fb + wb + nb + range_bytes(g), "pbrt_bmfr: workspace too small");
        char* ws = static_cast<char*>(im->workspace);
        launch_bmfr(*p, *im, g, reinterpret_cast<uint16_t*>(ws), reinterpret_cast<float*>(ws + fb),
                    reinterpret_cast<float*>(ws + fb + wb), reinterpret_cast<float*>(ws + fb + wb + nb),
                    (hipStream_t)stream);
    });
}

extern "C" int32_t pbrt_bmfr_noalloc(const pbrt_bmfr_params* p, const pbrt_bmfr_images* im, void* stream) {
    return guard([&] {
        PBRT_REQUIRE(p && im, "pbrt_bmfr_noalloc: NULL argument");
        require_bmfr_images(im, "pbrt_bmfr_noalloc");
        BmfrGeom g = make_geom(*p);
        // the Gram fits keep the weights, the frame's noise table and the blocks'
        // depth ranges; the Householder fit the fp16 feature array, the weights and
        // the ranges.  Every byte used is written before it is read.
        const bool hh = p->fit == PBRT_BMFR_FIT_HOUSEHOLDER;
        const size_t fb = hh ? align256(feature_bytes(g)) : 0, wb = align256(weight_bytes(g));
        const size_t nb = hh ? 0 : align256(noise_bytes(g));
        const size_t need = fb + wb + nb + range_bytes(g);
        if (need > kNoallocScratchBytes)
            fail(PBRT_ERR_INVALID_ARG, "pbrt_bmfr_noalloc: this frame needs " + std::to_string(need) +
                                           " bytes of scratch, more than the " + std::to_string(kNoallocScratchBytes) +
                                           " the library holds per device (use pbrt_bmfr with a workspace)");
        char* base = device_scratch();
        launch_bmfr(*p, *im, g, reinterpret_cast<uint16_t*>(base), reinterpret_cast<float*>(base + fb),
                    reinterpret_cast<float*>(base + fb + wb), reinterpret_cast<float*>(base + fb + wb + nb),
                    (hipStream_t)stream);
    });
}

extern "C" int32_t pbrt_autotune_bmfr(const pbrt_bmfr_params* p, const pbrt_bmfr_images* im, uint32_t repeats,
                                      pbrt_bmfr_tuning* out, void* stream) {
    return guard([&] {
        PBRT_REQUIRE(p && im && out, "pbrt_autotune_bmfr: NULL argument");
        PBRT_REQUIRE(p->fit == PBRT_BMFR_FIT_GRAM_F16 || p->fit == PBRT_BMFR_FIT_GRAM_F64,
                     "pbrt_autotune_bmfr: the fit must be a Gram fit");
        PBRT_REQUIRE(repeats >= 1 && repeats <= 100, "pbrt_autotune_bmfr: repeats must be 1..100");
        hipStream_t st = (hipStream_t)stream;
        std::memset(out, 0, sizeof(*out));
        hipEvent_t e0, e1;
        PBRT_HIP(hipEventCreate(&e0));
        PBRT_HIP(hipEventCreate(&e1));
        float best_ms = 3.4e38f;
        const uint32_t threads[3] = {256, 512, 1024};
        int32_t status = PBRT_OK;
        for (uint32_t t : threads) {
            pbrt_bmfr_params c = *p;
            c.fit_threads = t;
            status = pbrt_bmfr(&c, im, stream);  // warm-up (and argument check)
            if (status != PBRT_OK) break;
            (void)hipEventRecord(e0, st);
            for (uint32_t r = 0; r < repeats; ++r) (void)pbrt_bmfr(&c, im, stream);
            (void)hipEventRecord(e1, st);
            (void)hipEventSynchronize(e1);
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            ms /= (float)repeats;
            out->fit_threads[out->candidates] = t;
            out->ms[out->candidates] = ms;
            ++out->candidates;
            if (ms < best_ms) {
                best_ms = ms;
                out->best = c;
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (status != PBRT_OK) fail(status, std::string("pbrt_autotune_bmfr: ") + pbrt_last_error());
        PBRT_HIP(hipGetLastError());
    });
}

namespace {

// Frame table of a batch and the workspace offsets of every frame's weights and
// noise table.
struct BatchPlan {
    std::vector<BmfrBatchFrame> table;
    size_t table_bytes = 0, total_bytes = 0;
    uint32_t blocks = 0;
    int bs = 0;
    uint32_t fit = 0;
};

BatchPlan plan_batch(const pbrt_bmfr_batch_frame* frames, uint32_t count, char* workspace) {
    PBRT_REQUIRE(frames && count > 0, "pbrt_bmfr_batched: no frames");
    BatchPlan plan;
    plan.table.resize(count);
    plan.table_bytes = align256(sizeof(BmfrBatchFrame) * count);
    size_t off = plan.table_bytes;
    uint64_t blocks = 0;
    for (uint32_t i = 0; i < count; ++i) {
        const pbrt_bmfr_params& p = frames[i].params;
        PBRT_REQUIRE(p.fit == PBRT_BMFR_FIT_GRAM_F16 || p.fit == PBRT_BMFR_FIT_GRAM_F64,
                     "pbrt_bmfr_batched: the batched fit is PBRT_BMFR_FIT_GRAM_F16 or PBRT_BMFR_FIT_GRAM_F64");
        if (i == 0) {
            plan.bs = (int)p.block_size;
            plan.fit = p.fit;
        }
        PBRT_REQUIRE((int)p.block_size == plan.bs && p.fit == plan.fit,
                     "pbrt_bmfr_batched: every frame of a batch needs the same block_size and fit");
        BmfrBatchFrame& e = plan.table[i];
        e.g = make_geom(p);
        e.im = frames[i].images;
        e.im.workspace = nullptr;
        e.im.workspace_bytes = 0;
        e.g.features = nullptr;
        e.g.weights = workspace ? reinterpret_cast<float*>(workspace + off) : nullptr;
        off = align256(off + weight_bytes(e.g));
        e.g.noise = workspace ? reinterpret_cast<float*>(workspace + off) : nullptr;
        off = align256(off + noise_bytes(e.g));
        e.g.ranges = workspace ? reinterpret_cast<float*>(workspace + off) : nullptr;
        off = align256(off + range_bytes(e.g));
        e.first_block = (uint32_t)blocks;
        blocks += (uint64_t)e.g.blocks_x * e.g.blocks_y;
    }
    // up to four post workgroups per block: the launch grid stays below 2^31
    PBRT_REQUIRE(blocks < (1ull << 29), "pbrt_bmfr_batched: too many blocks for one launch");
    plan.blocks = (uint32_t)blocks;
    plan.total_bytes = off;
    return plan;
}

// Pinned staging buffer for frame tables; reused once the previous upload is done.
struct TableStaging {
    std::mutex mu;
    void* host = nullptr;
    size_t capacity = 0;
    hipEvent_t uploaded = nullptr;
    bool pending = false;
};

TableStaging& table_staging() {
    static TableStaging* t = new TableStaging();  // lives until the process ends
    return *t;
}

}  // namespace

extern "C" int32_t pbrt_bmfr_batched_workspace_size(const pbrt_bmfr_batch_frame* frames, uint32_t count,
                                                    size_t* out) {
    return guard([&] {
        PBRT_REQUIRE(out, "pbrt_bmfr_batched_workspace_size: NULL output");
        *out = plan_batch(frames, count, nullptr).total_bytes;
    });
}

extern "C" int32_t pbrt_bmfr_batched(const pbrt_bmfr_batch_frame* frames, uint32_t count, void* workspace,
                                     size_t workspace_bytes, void* stream) {
    return guard([&] {
        PBRT_REQUIRE(workspace, "pbrt_bmfr_batched: NULL workspace");
        for (uint32_t i = 0; frames && i < count; ++i) {
            const pbrt_bmfr_images* im = &frames[i].images;
            PBRT_REQUIRE(im->depth && im->normal && im->albedo && im->motion && im->spp && im->noisy &&
                             im->denoised && im->final_image,
                         "pbrt_bmfr_batched: missing image");
        }
        BatchPlan plan = plan_batch(frames, count, static_cast<char*>(workspace));
        PBRT_REQUIRE(workspace_bytes >= plan.total_bytes, "pbrt_bmfr_batched: workspace too small");
        hipStream_t st = (hipStream_t)stream;
        const size_t bytes = sizeof(BmfrBatchFrame) * count;
        {
            TableStaging& ts = table_staging();
            std::lock_guard<std::mutex> lock(ts.mu);
            if (ts.pending) PBRT_HIP(hipEventSynchronize(ts.uploaded));
            ts.pending = false;
            if (ts.capacity < bytes) {
                if (ts.host) PBRT_HIP(hipHostFree(ts.host));
                ts.host = nullptr;
                ts.capacity = 0;
                PBRT_HIP(hipHostMalloc(&ts.host, bytes, hipHostMallocDefault));
                ts.capacity = bytes;
            }
            if (!ts.uploaded) PBRT_HIP(hipEventCreateWithFlags(&ts.uploaded, hipEventDisableTiming));
            std::memcpy(ts.host, plan.table.data(), bytes);
            PBRT_HIP(hipMemcpyAsync(workspace, ts.host, bytes, hipMemcpyHostToDevice, st));
            PBRT_HIP(hipEventRecord(ts.uploaded, st));
            ts.pending = true;
        }
        const BmfrBatchFrame* table = static_cast<const BmfrBatchFrame*>(workspace);
        const dim3 grid(plan.blocks);
        const bool f64 = plan.fit == PBRT_BMFR_FIT_GRAM_F64;
        const int entries = kFeatures * plan.bs * plan.bs;
        hipLaunchKernelGGL(k_bmfr_noise_batched, dim3((entries + 255) / 256, count), dim3(256), 0, st, table);
        if (plan.bs == 32) {
            if (f64)
                hipLaunchKernelGGL((k_bmfr_fit_gram_batched<1024, true>), grid, dim3(kFitThreads), 0, st, table, count);
            else
                hipLaunchKernelGGL((k_bmfr_fit_gram_batched<1024, false>), grid, dim3(kFitThreads), 0, st, table, count);
        } else if (plan.bs == 16) {
            if (f64)
                hipLaunchKernelGGL((k_bmfr_fit_gram_batched<256, true>), grid, dim3(kFitThreads), 0, st, table, count);
            else
                hipLaunchKernelGGL((k_bmfr_fit_gram_batched<256, false>), grid, dim3(kFitThreads), 0, st, table, count);
        } else {
            if (f64)
                hipLaunchKernelGGL((k_bmfr_fit_gram_batched<64, true>), grid, dim3(kFitThreads), 0, st, table, count);
            else
                hipLaunchKernelGGL((k_bmfr_fit_gram_batched<64, false>), grid, dim3(kFitThreads), 0, st, table, count);
        }
        const int rows = post_rows(plan.bs);
        const dim3 post_grid(plan.blocks * (uint32_t)(plan.bs / rows)), post_block(plan.bs * rows);
        if (plan.bs == 32)
            hipLaunchKernelGGL(k_bmfr_post_batched<32>, post_grid, post_block, 0, st, table, count);
        else if (plan.bs == 16)
            hipLaunchKernelGGL(k_bmfr_post_batched<16>, post_grid, post_block, 0, st, table, count);
        else
            hipLaunchKernelGGL(k_bmfr_post_batched<8>, post_grid, post_block, 0, st, table, count);
        PBRT_HIP(hipGetLastError());
    });
}
