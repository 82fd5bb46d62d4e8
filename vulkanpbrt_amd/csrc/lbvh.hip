// Morton-code LBVH built on gfx950 (PBRT_BVH_LBVH_GPU), emitting the same node and
// triangle-record layout as the host SAH builder (scene_types.hpp), so traversal
// and shading are unchanged.
//
//   1. centroid bounds          per-block min/max, one folding block
//   2. Morton codes             30 bits (10 per axis) of the centroid in its bounds
//   3. sort                     stable radix sort of (code, triangle) pairs (gpu_sort.hpp)
//   4. hierarchy                Karras' binary radix tree: internal node i finds its
//                               key range and split from longest common prefixes
//                               (equal codes are told apart by their sorted index)
//   5. refit                    bottom-up from every leaf; the second child to
//                               arrive at a node (atomic counter) merges the boxes
//   6. emit                     children whose key range holds <= max_leaf
//                               triangles become leaves; depth = emitted levels
//   7. triangle records         gathered into the sorted (BVH) order
// The reference gets its acceleration structures from the Vulkan driver
// (AccelerationStructure.cpp:44); this builder is the project's own.
#include <hip/hip_runtime.h>

#include <vector>

#include "capi_common.hpp"
#include "gpu_sort.hpp"
#include "lbvh.hpp"
#include "scene_types.hpp"

namespace pbrt {
namespace {

struct Aabb {
    float lo[3], hi[3];
};

__device__ __forceinline__ uint32_t expand_bits(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

__device__ __forceinline__ void tri_box(const float* t, Aabb& b) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        b.lo[a] = fminf(t[a], fminf(t[3 + a], t[6 + a]));
        b.hi[a] = fmaxf(t[a], fmaxf(t[3 + a], t[6 + a]));
    }
}

__global__ void __launch_bounds__(256) k_centroid_bounds(const float* tris, uint32_t n, float* partial) {
    __shared__ float red[6][4];
    float mn[3] = {3.4e38f, 3.4e38f, 3.4e38f}, mx[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const float* t = tris + 9 * (size_t)i;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const float c = (t[a] + t[3 + a] + t[6 + a]) * (1.0f / 3.0f);
            mn[a] = fminf(mn[a], c);
            mx[a] = fmaxf(mx[a], c);
        }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a)
        for (int off = 32; off > 0; off >>= 1) {
            mn[a] = fminf(mn[a], __shfl_xor(mn[a], off));
            mx[a] = fmaxf(mx[a], __shfl_xor(mx[a], off));
        }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
        for (int a = 0; a < 3; ++a) {
            red[a][wave] = mn[a];
            red[3 + a][wave] = mx[a];
        }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int a = threadIdx.x;
        float v = red[a][0];
        for (int w = 1; w < 4; ++w) v = a < 3 ? fminf(v, red[a][w]) : fmaxf(v, red[a][w]);
        partial[blockIdx.x * 6 + a] = v;
    }
}

__global__ void k_fold_bounds(float* partial, int blocks) {
    const int a = threadIdx.x;
    if (a >= 6) return;
    float v = partial[a];
    for (int b = 1; b < blocks; ++b) v = a < 3 ? fminf(v, partial[b * 6 + a]) : fmaxf(v, partial[b * 6 + a]);
    partial[a] = v;
}

__global__ void __launch_bounds__(256) k_morton(const float* tris, uint32_t n, const float* bounds, uint32_t* keys,
                                               uint32_t* vals) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float* t = tris + 9 * (size_t)i;
    uint32_t q[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float c = (t[a] + t[3 + a] + t[6 + a]) * (1.0f / 3.0f);
        const float ext = bounds[3 + a] - bounds[a];
        const float u = ext > 0.0f ? (c - bounds[a]) / ext : 0.5f;
        q[a] = (uint32_t)fminf(fmaxf(u * 1024.0f, 0.0f), 1023.0f);
    }
    keys[i] = (expand_bits(q[0]) << 2) | (expand_bits(q[1]) << 1) | expand_bits(q[2]);
    vals[i] = i;
}

// longest common prefix of sorted keys i and j (equal keys: 32 + that of the indices)
__device__ __forceinline__ int delta(const uint32_t* keys, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    const uint32_t a = keys[i], b = keys[j];
    if (a != b) return __clz(a ^ b);
    return 32 + __clz((uint32_t)i ^ (uint32_t)j);
}

// child code: >= 0 internal node index, < 0: ~leaf (sorted triangle position)
__global__ void __launch_bounds__(256) k_karras(const uint32_t* keys, int n, int2* child, int2* range, int* parent) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n - 1) return;
    const int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = delta(keys, n, i, i - d);
    int lmax = 2;
    while (delta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = delta(keys, n, i, j);
    int s = 0;
    for (int t = (l + 1) >> 1;; t = (t + 1) >> 1) {
        if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
        if (t == 1) break;
    }
    const int gamma = i + s * d + min(d, 0);
    const int lo = min(i, j), hi = max(i, j);
    const int left = lo == gamma ? ~gamma : gamma;
    const int right = hi == gamma + 1 ? ~(gamma + 1) : gamma + 1;
    child[i] = make_int2(left, right);
    range[i] = make_int2(lo, hi);
    // parent of internal children; leaves store theirs at n - 1 + leaf
    parent[left >= 0 ? left : n - 1 + ~left] = i;
    parent[right >= 0 ? right : n - 1 + ~right] = i;
}

__global__ void __launch_bounds__(256) k_refit(const float* tris, const uint32_t* vals, int n, const int2* child,
                                              const int* parent, int* arrivals, Aabb* boxes) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    Aabb b;
    tri_box(tris + 9 * (size_t)vals[k], b);
    boxes[n - 1 + k] = b;  // leaf boxes after the internal nodes
    int node = parent[n - 1 + k];
    // a radix tree over 30-bit codes with index tie-breaks is at most 62 levels deep;
    // the bound only guards the walk
    for (int level = 0; node >= 0 && level < 64; ++level) {
        __threadfence();
        if (atomicAdd(&arrivals[node], 1) == 0) return;  // the sibling is not done yet
        __threadfence();
        const int2 c = child[node];
        const Aabb l = boxes[c.x >= 0 ? c.x : n - 1 + ~c.x];
        const Aabb r = boxes[c.y >= 0 ? c.y : n - 1 + ~c.y];
        Aabb u;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            u.lo[a] = fminf(l.lo[a], r.lo[a]);
            u.hi[a] = fmaxf(l.hi[a], r.hi[a]);
        }
        boxes[node] = u;
        node = node == 0 ? -1 : parent[node];
    }
}

// depth of the emitted tree: internal ancestors whose range exceeds max_leaf, + 1
__global__ void __launch_bounds__(256) k_depth(int n, const int2* range, const int* parent, int max_leaf,
                                              int* depth) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    int d = 1;
    int node = parent[n - 1 + k];
    for (int level = 0; node >= 0 && level < 64; ++level) {
        const int2 r = range[node];
        if (node == 0 || r.y - r.x + 1 > max_leaf) ++d;
        node = node == 0 ? -1 : parent[node];
    }
    atomicMax(depth, d);
}

__device__ __forceinline__ int32_t encode_child(int c, int n, const int2* range, int max_leaf) {
    if (c < 0) return ~(int32_t)(((uint32_t)~c << kLeafCountBits) | 0u);
    const int2 r = range[c];
    const int count = r.y - r.x + 1;
    if (count <= max_leaf) return ~(int32_t)(((uint32_t)r.x << kLeafCountBits) | (uint32_t)(count - 1));
    return c;
}

__global__ void __launch_bounds__(256) k_emit(int n, const int2* child, const int2* range, const Aabb* boxes,
                                             int max_leaf, float4* nodes) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n - 1) return;
    const int2 c = child[i];
    const Aabb a = boxes[c.x >= 0 ? c.x : n - 1 + ~c.x];
    const Aabb b = boxes[c.y >= 0 ? c.y : n - 1 + ~c.y];
    float4* o = nodes + 4 * (size_t)i;
    o[0] = make_float4(a.lo[0], a.hi[0], a.lo[1], a.hi[1]);
    o[1] = make_float4(b.lo[0], b.hi[0], b.lo[1], b.hi[1]);
    o[2] = make_float4(a.lo[2], a.hi[2], b.lo[2], b.hi[2]);
    o[3] = make_float4(__int_as_float(encode_child(c.x, n, range, max_leaf)),
                       __int_as_float(encode_child(c.y, n, range, max_leaf)), 0.0f, 0.0f);
}

__global__ void __launch_bounds__(256) k_gather_tris(const float* tris, const uint4* meta, const uint32_t* vals,
                                                    uint32_t n, float4* out) {
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    const uint32_t t = vals[k];
    const float* p = tris + 9 * (size_t)t;
    const uint4 m = meta[t];  // instance, primitive, flags
    out[3 * (size_t)k + 0] = make_float4(p[0], p[1], p[2], __uint_as_float(m.x));
    out[3 * (size_t)k + 1] = make_float4(p[3] - p[0], p[4] - p[1], p[5] - p[2], __uint_as_float(m.y));
    out[3 * (size_t)k + 2] = make_float4(p[6] - p[0], p[7] - p[1], p[8] - p[2], __uint_as_float(m.z));
}

unsigned blocks_for(size_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

LbvhResult build_lbvh(const std::vector<float>& world_tris, const std::vector<uint32_t>& inst,
                      const std::vector<uint32_t>& prim, const std::vector<uint32_t>& flags, int max_depth,
                      DeviceBuffer& tris_out, DeviceBuffer& nodes_out) {
    const uint32_t n = (uint32_t)inst.size();
    PBRT_REQUIRE(n >= 2, "LBVH: needs at least two triangles");
    hipStream_t st = nullptr;
    hipEvent_t e0, e1;
    PBRT_HIP(hipEventCreate(&e0));
    PBRT_HIP(hipEventCreate(&e1));
    DeviceBuffer tris, meta, keys, vals, keys2, vals2, partial, child, range, parent, arrivals, boxes, depth, temp;
    tris.upload(world_tris.data(), world_tris.size() * sizeof(float));
    std::vector<uint4> m(n);
    for (uint32_t i = 0; i < n; ++i) m[i] = make_uint4(inst[i], prim[i], flags[i], 0);
    meta.upload(m.data(), m.size() * sizeof(uint4));
    keys.allocate(n * 4);
    vals.allocate(n * 4);
    keys2.allocate(n * 4);
    vals2.allocate(n * 4);
    const int rb = 512;
    partial.allocate(rb * 6 * 4);
    child.allocate((size_t)(n - 1) * sizeof(int2));
    range.allocate((size_t)(n - 1) * sizeof(int2));
    parent.allocate((size_t)(2 * n - 1) * sizeof(int));
    arrivals.allocate((size_t)(n - 1) * sizeof(int));
    boxes.allocate((size_t)(2 * n - 1) * sizeof(Aabb));
    depth.allocate(sizeof(int));
    temp.allocate(gpusort::sort_scratch_bytes(n));
    PBRT_HIP(hipMemset(parent.ptr, 0xFF, parent.bytes));  // root's parent: -1
    PBRT_HIP(hipDeviceSynchronize());

    PBRT_HIP(hipEventRecord(e0, st));
    hipLaunchKernelGGL(k_centroid_bounds, dim3(rb), dim3(256), 0, st, tris.as<float>(), n, partial.as<float>());
    hipLaunchKernelGGL(k_fold_bounds, dim3(1), dim3(64), 0, st, partial.as<float>(), rb);
    hipLaunchKernelGGL(k_morton, dim3(blocks_for(n)), dim3(256), 0, st, tris.as<float>(), n, partial.as<float>(),
                       keys.as<uint32_t>(), vals.as<uint32_t>());
    const bool in2 = gpusort::sort_pairs(keys.as<uint32_t>(), vals.as<uint32_t>(), keys2.as<uint32_t>(),
                                         vals2.as<uint32_t>(), n, 30, temp.ptr, st);
    const uint32_t* skeys = in2 ? keys2.as<uint32_t>() : keys.as<uint32_t>();
    const uint32_t* svals = in2 ? vals2.as<uint32_t>() : vals.as<uint32_t>();
    hipLaunchKernelGGL(k_karras, dim3(blocks_for(n - 1)), dim3(256), 0, st, skeys, (int)n,
                       child.as<int2>(), range.as<int2>(), parent.as<int>());
    PBRT_HIP(hipMemsetAsync(arrivals.ptr, 0, arrivals.bytes, st));
    hipLaunchKernelGGL(k_refit, dim3(blocks_for(n)), dim3(256), 0, st, tris.as<float>(), svals, (int)n,
                       child.as<int2>(), parent.as<int>(), arrivals.as<int>(), boxes.as<Aabb>());
    nodes_out.allocate((size_t)(n - 1) * 4 * sizeof(float4));
    tris_out.allocate((size_t)n * 3 * sizeof(float4));
    LbvhResult res{};
    // smallest leaf size (up to the 16 of the node format) whose tree fits the stack
    int max_leaf = 4;
    int d = 0;
    for (;; max_leaf *= 2) {
        PBRT_HIP(hipMemsetAsync(depth.ptr, 0, sizeof(int), st));
        hipLaunchKernelGGL(k_depth, dim3(blocks_for(n)), dim3(256), 0, st, (int)n, range.as<int2>(), parent.as<int>(),
                           max_leaf, depth.as<int>());
        PBRT_HIP(hipMemcpyAsync(&d, depth.ptr, sizeof(int), hipMemcpyDeviceToHost, st));
        PBRT_HIP(hipStreamSynchronize(st));
        if (d <= max_depth || max_leaf >= kMaxLeafSize) break;
    }
    if (d > max_depth)
        fail(PBRT_ERR_UNSUPPORTED, "pbrt_scene_create: LBVH deeper than the traversal stack (" + std::to_string(d) +
                                       " levels); use PBRT_BVH_SAH_HOST for this scene");
    hipLaunchKernelGGL(k_emit, dim3(blocks_for(n - 1)), dim3(256), 0, st, (int)n, child.as<int2>(), range.as<int2>(),
                       boxes.as<Aabb>(), max_leaf, nodes_out.as<float4>());
    hipLaunchKernelGGL(k_gather_tris, dim3(blocks_for(n)), dim3(256), 0, st, tris.as<float>(), meta.as<uint4>(),
                       svals, n, tris_out.as<float4>());
    PBRT_HIP(hipEventRecord(e1, st));
    PBRT_HIP(hipGetLastError());
    PBRT_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    PBRT_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    res.node_count = n - 1;
    res.depth = (uint32_t)d;
    res.max_leaf = (uint32_t)max_leaf;
    res.build_ms = ms;
    return res;
}

}  // namespace pbrt
