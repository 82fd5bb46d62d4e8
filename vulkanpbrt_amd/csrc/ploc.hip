// PLOC BVH builder on gfx950 (PBRT_BVH_PLOC_GPU): parallel locally-ordered
// clustering (Meister & Bittner, "Parallel Locally-Ordered Clustering for
// Bounding Volume Hierarchy Construction", IEEE TVCG 2018), then a SAH leaf
// collapse, emitting the node and triangle-record layout of the other builders
// (scene_types.hpp), so traversal and shading are unchanged.
//
//   1. Morton codes of the triangle centroids (30 bits, cubic cells), stable radix sort
//      (gpu_sort.hpp); every sorted triangle starts as a cluster
//   2. per iteration, while more than one cluster is left:
//        nearest neighbour   each cluster scans the kRadius clusters on either
//                            side in the sorted array for the smallest surface
//                            area of the union box (ties: the lower index; the
//                            relation then always holds a mutual pair)
//        merge               mutual pairs become a new inner node at the lower
//                            index; one scan of (merge, survive) flags numbers the
//                            new nodes and compacts the surviving clusters
//      inner nodes are numbered in creation order, so a node's children always
//      have smaller numbers and each iteration's nodes are one contiguous range
//   3. SAH collapse, at creation: a node over <= max_leaf triangles becomes a
//      leaf when count * C_tri * SA <= C_node * SA + C(left) + C(right)
//   4. emit, root first (iterations in reverse): each emitted inner node places
//      its children (triangle offsets in depth-first leaf order, depth); an
//      emitted leaf writes its subtree's triangle records contiguously.  Output
//      node index = (inner nodes - 1) - creation number, so the root is node 0;
//      nodes inside collapsed leaves stay allocated and unreferenced.
// The reference gets its acceleration structures from the Vulkan driver
// (AccelerationStructure.cpp:44); this builder is the project's own.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <utility>
#include <string>
#include <vector>

#include "capi_common.hpp"
#include "gpu_sort.hpp"
#include "bvh_host.hpp"
#include "ploc.hpp"
#include "scene_types.hpp"

namespace pbrt {
namespace {

// Neighbours searched on either side (on the 4.9M-triangle city, 8 bounces at 2160p:
// 11.6 ms per trace with 32, 11.8 with 16, 12.6 with 128; the LDS window holds kMaxRadius).
constexpr int kRadius = 32;
// clusters a host SAH top tree starts from on large scenes (city, 8 bounces at 2160p:
// 9.6 ms per trace with 16384, 9.9 with 4096, 10.4 with 1024; clustering to the root: 11.6)
constexpr uint32_t kTopClusters = 16384;
constexpr int kMaxRadius = 128;
constexpr float kNodeCost = 1.5f;  // an inner-node visit priced in triangle tests (as the host SAH)

struct Box {
    float lo[3], hi[3];
};

__device__ __forceinline__ Box box_union(const Box& a, const Box& b) {
    Box u;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        u.lo[k] = fminf(a.lo[k], b.lo[k]);
        u.hi[k] = fmaxf(a.hi[k], b.hi[k]);
    }
    return u;
}

// half the surface area
__device__ __forceinline__ float half_area(const Box& b) {
    const float dx = b.hi[0] - b.lo[0], dy = b.hi[1] - b.lo[1], dz = b.hi[2] - b.lo[2];
    return dx * dy + dy * dz + dz * dx;
}

__device__ __forceinline__ Box triangle_box(const float* t) {
    Box b;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        b.lo[k] = fminf(t[k], fminf(t[3 + k], t[6 + k]));
        b.hi[k] = fmaxf(t[k], fmaxf(t[3 + k], t[6 + k]));
    }
    return b;
}

__device__ __forceinline__ uint32_t spread_bits(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

__global__ void __launch_bounds__(256) k_centroid_box(const float* __restrict__ tris, uint32_t n, float* partial) {
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

__global__ void k_fold_box(float* partial, int blocks) {
    const int a = threadIdx.x;
    if (a >= 6) return;
    float v = partial[a];
    for (int b = 1; b < blocks; ++b) v = a < 3 ? fminf(v, partial[b * 6 + a]) : fmaxf(v, partial[b * 6 + a]);
    partial[a] = v;
}

// 30-bit Morton code of each centroid over cubic cells (the centroid bounds' largest
// extent on every axis): an axis of small extent, such as a city's height, then does
// not take a third of the code's bits and split every tall object across the order
__global__ void __launch_bounds__(256) k_codes(const float* __restrict__ tris, uint32_t n,
                                              const float* __restrict__ bounds, uint32_t* keys, uint32_t* vals) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float* t = tris + 9 * (size_t)i;
    uint32_t q[3];
    const float ext = fmaxf(fmaxf(bounds[3] - bounds[0], bounds[4] - bounds[1]), bounds[5] - bounds[2]);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float c = (t[a] + t[3 + a] + t[6 + a]) * (1.0f / 3.0f);
        const float u = ext > 0.0f ? (c - bounds[a]) / ext : 0.5f;
        q[a] = (uint32_t)fminf(fmaxf(u * 1024.0f, 0.0f), 1023.0f);
    }
    keys[i] = (spread_bits(q[0]) << 2) | (spread_bits(q[1]) << 1) | spread_bits(q[2]);
    vals[i] = i;
}

// clusters of the sorted triangles: leaf references ~k, their boxes
__global__ void __launch_bounds__(256) k_leaf_clusters(const float* __restrict__ tris,
                                                       const uint32_t* __restrict__ order, uint32_t n, int* cnode,
                                                       Box* cbox) {
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    cnode[k] = ~(int)k;
    cbox[k] = triangle_box(tris + 9 * (size_t)order[k]);
}

// nearest neighbour of every cluster within `radius` (<= kMaxRadius) positions
// (boxes staged in LDS)
__global__ void __launch_bounds__(256) k_nearest(const Box* __restrict__ cbox, int n, int radius,
                                                int* __restrict__ nn) {
    __shared__ Box win[256 + 2 * kMaxRadius];
    const int first = blockIdx.x * 256 - radius;
    for (int k = threadIdx.x; k < 256 + 2 * radius; k += 256) {
        const int j = first + k;
        if (j >= 0 && j < n) win[k] = cbox[j];
    }
    __syncthreads();
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const Box b = win[threadIdx.x + radius];
    // a neighbour even when no area compares (infinite or NaN boxes)
    float best = INFINITY;
    int arg = i + 1 < n ? i + 1 : i - 1;
    const int lo = max(0, i - radius), hi = min(n - 1, i + radius);
    for (int j = lo; j <= hi; ++j) {
        if (j == i) continue;
        const float d = half_area(box_union(b, win[j - first]));
        if (d < best) {  // ascending j: ties keep the lower index
            best = d;
            arg = j;
        }
    }
    nn[i] = arg;
}

// (merges << 32) | survives: a mutual pair merges at its lower index and drops the higher
__global__ void __launch_bounds__(256) k_merge_flags(const int* __restrict__ nn, int n, uint64_t* flags) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int j = nn[i];
    const bool mutual = nn[j] == i;
    const uint64_t merge = mutual && i < j ? 1u : 0u;
    const uint64_t survive = mutual && i > j ? 0u : 1u;
    flags[i] = merge << 32 | survive;
}

struct NodeArrays {
    int2* child;     // creation number -> children (>= 0 inner creation number, < 0 ~sorted triangle)
    Box* box;        // the node's box
    uint32_t* info;  // triangle count << 1 | collapse
    float* cost;     // SAH cost times the node's area, as a subtree of leaves <= max_leaf
    uint8_t* levels; // levels of the emitted subtree, its leaf level included
};

__device__ __forceinline__ void child_stats(const NodeArrays& nd, int c, const Box& b, uint32_t& count, float& cost) {
    if (c < 0) {
        count = 1;
        cost = half_area(b);
    } else {
        count = nd.info[c] >> 1;
        cost = nd.cost[c];
    }
}

__global__ void __launch_bounds__(256) k_merge(const int* __restrict__ nn, const uint64_t* __restrict__ flags,
                                              const uint64_t* __restrict__ prefix, int n, const int* __restrict__ cnode,
                                              const Box* __restrict__ cbox, int node_base, int max_leaf, NodeArrays nd,
                                              int* __restrict__ cnode2, Box* __restrict__ cbox2) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint64_t f = flags[i];
    if (!(f & 1u)) return;
    const uint64_t p = prefix[i];
    const uint32_t pos = (uint32_t)p;
    if (f >> 32) {
        const int j = nn[i];
        const int id = node_base + (int)(p >> 32);
        const Box a = cbox[i], b = cbox[j];
        const Box u = box_union(a, b);
        uint32_t ca, cb;
        float sa, sb;
        child_stats(nd, cnode[i], a, ca, sa);
        child_stats(nd, cnode[j], b, cb, sb);
        const uint32_t count = ca + cb;
        const float area = half_area(u);
        const float inner = kNodeCost * area + sa + sb;
        const float leaf = (float)count * area;
        const bool collapse = (int)count <= max_leaf && leaf <= inner;
        const int la = cnode[i] < 0 ? 1 : nd.levels[cnode[i]], lb = cnode[j] < 0 ? 1 : nd.levels[cnode[j]];
        nd.levels[id] = (uint8_t)min(255, collapse ? 1 : 1 + max(la, lb));
        nd.child[id] = make_int2(cnode[i], cnode[j]);
        nd.box[id] = u;
        nd.info[id] = count << 1 | (collapse ? 1u : 0u);
        nd.cost[id] = collapse ? leaf : inner;
        cnode2[pos] = id;
        cbox2[pos] = u;
    } else {
        cnode2[pos] = cnode[i];
        cbox2[pos] = cbox[i];
    }
}

struct EmitArrays {
    uint32_t* offset;  // per creation number: first triangle record of the subtree
    uint32_t* depth;   // levels from the root (root 1)
    uint32_t* state;   // 0 not reached, 1 emitted inner node, 2 emitted leaf
};

__device__ __forceinline__ void write_record(const float* __restrict__ tris, const uint4* __restrict__ meta,
                                             const uint32_t* __restrict__ order, uint32_t sorted, uint32_t slot,
                                             float4* out) {
    const uint32_t t = order[sorted];
    const float* p = tris + 9 * (size_t)t;
    const uint4 m = meta[t];  // instance, primitive, flags
    out[3 * (size_t)slot + 0] = make_float4(p[0], p[1], p[2], __uint_as_float(m.x));
    out[3 * (size_t)slot + 1] = make_float4(p[3] - p[0], p[4] - p[1], p[5] - p[2], __uint_as_float(m.y));
    out[3 * (size_t)slot + 2] = make_float4(p[6] - p[0], p[7] - p[1], p[8] - p[2], __uint_as_float(m.z));
}

__device__ __forceinline__ uint32_t leaf_code(uint32_t first, uint32_t count) {
    return ~((first << kLeafCountBits) | (count - 1u));
}

__device__ __forceinline__ Box ch_box(const NodeArrays& nd, int c, const float* __restrict__ tris,
                                     const uint32_t* __restrict__ order) {
    return c < 0 ? triangle_box(tris + 9 * (size_t)order[~c]) : nd.box[c];
}

// The clusters left under a tree built on the host: inner clusters get their place
// (triangle offset, depth, leaf or inner node) like children of an emitted node;
// single-triangle clusters are written as records.
struct ClusterSeed {
    int node;         // creation number, or ~sorted triangle
    uint32_t offset;  // first triangle record
    uint32_t depth;   // level of the cluster's node (root 1)
    uint32_t pad;
};

__global__ void __launch_bounds__(256) k_seed_clusters(const ClusterSeed* __restrict__ seeds, int count,
                                                      NodeArrays nd, EmitArrays em, const float* __restrict__ tris,
                                                      const uint4* __restrict__ meta,
                                                      const uint32_t* __restrict__ order, float4* __restrict__ records) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= count) return;
    const ClusterSeed c = seeds[k];
    if (c.node < 0) {
        write_record(tris, meta, order, (uint32_t)~c.node, c.offset, records);
        return;
    }
    em.offset[c.node] = c.offset;
    em.depth[c.node] = c.depth;
    em.state[c.node] = (nd.info[c.node] & 1u) ? 2u : 1u;
}

// (triangle count << 1 | collapse, emitted levels, SAH cost of the subtree per unit of its
// box's area) of each remaining cluster, for the host
__global__ void __launch_bounds__(256) k_cluster_stats(const int* __restrict__ cnode, const Box* __restrict__ cbox,
                                                      int live, NodeArrays nd, uint4* out) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= live) return;
    const int c = cnode[k];
    if (c < 0) {
        out[k] = make_uint4(2u, 1u, __float_as_uint(1.0f), 0u);
        return;
    }
    const float area = half_area(cbox[k]);
    const float per_area = area > 0.0f ? nd.cost[c] / area : 1.0f;
    out[k] = make_uint4(nd.info[c], nd.levels[c], __float_as_uint(per_area), 0u);
}

__global__ void k_emit_root(EmitArrays em, int root) {
    em.offset[root] = 0;
    em.depth[root] = 1;
    em.state[root] = 1;  // the root stays an inner node: traversal starts at node 0
}

// One iteration's nodes, root side first: places the children of emitted inner
// nodes and writes the triangle records of emitted leaves.
__global__ void __launch_bounds__(256) k_emit_range(int begin, int count, int inner_total, NodeArrays nd,
                                                   EmitArrays em, const float* __restrict__ tris,
                                                   const uint4* __restrict__ meta, const uint32_t* __restrict__ order,
                                                   float4* __restrict__ nodes, float4* __restrict__ records,
                                                   uint32_t* __restrict__ max_depth) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= count) return;
    const int id = begin + k;
    const uint32_t state = em.state[id];
    if (state == 0) return;
    const uint32_t off = em.offset[id], dep = em.depth[id];
    const int2 c = nd.child[id];
    if (state == 2) {
        // leaf: the subtree's triangles in depth-first order (at most max_leaf of them)
        int stack[kMaxLeafSize];
        int sp = 0;
        stack[sp++] = id;
        uint32_t slot = off;
        while (sp > 0 && slot < off + kMaxLeafSize) {
            const int node = stack[--sp];
            if (node < 0) {
                write_record(tris, meta, order, (uint32_t)~node, slot++, records);
            } else {
                const int2 cc = nd.child[node];
                if (sp + 2 > kMaxLeafSize) break;  // unreachable: a subtree of <= 16 leaves
                stack[sp++] = cc.y;
                stack[sp++] = cc.x;
            }
        }
        atomicMax(max_depth, dep);
        return;
    }
    // inner node: both children, left first in the leaf order
    const int kids[2] = {c.x, c.y};
    uint32_t codes[2];
    uint32_t child_off = off;
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const int ch = kids[s];
        if (ch < 0) {
            write_record(tris, meta, order, (uint32_t)~ch, child_off, records);
            codes[s] = leaf_code(child_off, 1);
            child_off += 1;
            atomicMax(max_depth, dep + 1);
        } else {
            const uint32_t info = nd.info[ch];
            const bool leaf = info & 1u;
            em.offset[ch] = child_off;
            em.depth[ch] = dep + 1;
            em.state[ch] = leaf ? 2u : 1u;
            codes[s] = leaf ? leaf_code(child_off, info >> 1) : (uint32_t)(inner_total - 1 - ch);
            child_off += info >> 1;
        }
    }
    const Box a = ch_box(nd, kids[0], tris, order), b = ch_box(nd, kids[1], tris, order);
    float4* o = nodes + 4 * (size_t)(inner_total - 1 - id);
    o[0] = make_float4(a.lo[0], a.hi[0], a.lo[1], a.hi[1]);
    o[1] = make_float4(b.lo[0], b.hi[0], b.lo[1], b.hi[1]);
    o[2] = make_float4(a.lo[2], a.hi[2], b.lo[2], b.hi[2]);
    o[3] = make_float4(__uint_as_float(codes[0]), __uint_as_float(codes[1]), 0.0f, 0.0f);
}

unsigned blocks_for(size_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

namespace {

// The top of the tree over `live` clusters as the host's binned SAH would build it,
// with single-cluster leaves, each cluster priced at its subtree's SAH cost per unit of
// its box's area (depth budget: the stack less the deepest cluster subtree), and the
// SAH cost of the whole tree that results.
struct TopPlan {
    uint32_t live = 0;
    std::vector<int> node;
    std::vector<uint4> stats;  // count << 1 | collapse, levels, cost per area (bits)
    BvhHost top;
    double cost = 0.0;  // SAH cost (triangle tests x half area) of the tree with this top
};

TopPlan plan_top(uint32_t live, const DeviceBuffer& cnode, const DeviceBuffer& cbox, const NodeArrays& nd,
                 int max_depth, hipStream_t st) {
    TopPlan p;
    p.live = live;
    DeviceBuffer stats_d;
    stats_d.allocate(live * sizeof(uint4));
    hipLaunchKernelGGL(k_cluster_stats, dim3(blocks_for(live)), dim3(256), 0, st, cnode.as<int>(), cbox.as<Box>(),
                       (int)live, nd, stats_d.as<uint4>());
    p.node.resize(live);
    std::vector<Box> box(live);
    p.stats.resize(live);
    PBRT_HIP(hipMemcpyAsync(p.node.data(), cnode.ptr, live * sizeof(int), hipMemcpyDeviceToHost, st));
    PBRT_HIP(hipMemcpyAsync(box.data(), cbox.ptr, live * sizeof(Box), hipMemcpyDeviceToHost, st));
    PBRT_HIP(hipMemcpyAsync(p.stats.data(), stats_d.ptr, live * sizeof(uint4), hipMemcpyDeviceToHost, st));
    PBRT_HIP(hipStreamSynchronize(st));
    auto area = [](const float* lo, const float* hi) {
        const double dx = (double)hi[0] - lo[0], dy = (double)hi[1] - lo[1], dz = (double)hi[2] - lo[2];
        return dx * dy + dy * dz + dz * dx;
    };
    std::vector<float> lo(3 * (size_t)live), hi(3 * (size_t)live), weight(live);
    uint32_t deepest = 1;
    for (uint32_t k = 0; k < live; ++k) {
        std::memcpy(&weight[k], &p.stats[k].z, 4);
        for (int a = 0; a < 3; ++a) {
            lo[3 * (size_t)k + a] = box[k].lo[a];
            hi[3 * (size_t)k + a] = box[k].hi[a];
        }
        deepest = std::max(deepest, p.stats[k].y);
        p.cost += (double)weight[k] * area(box[k].lo, box[k].hi);
    }
    // a top leaf at level L puts its cluster's leaves at level L - 1 + levels
    const int budget = max_depth + 1 - (int)deepest;
    if (budget < 2) {
        p.cost = INFINITY;
        return p;
    }
    p.top = build_bvh_sah(lo.data(), hi.data(), live, 1, budget, weight.data());
    PBRT_REQUIRE(p.top.nodes.size() / 16 == live - 1, "PLOC: top tree node count mismatch");
    for (size_t h = 0; h + 1 < live; ++h) {
        const float* r = &p.top.nodes[16 * h];  // {c0 x, c1 x, (c0 z, c1 z)}: the union of both child boxes
        const float ulo[3] = {std::min(r[0], r[4]), std::min(r[2], r[6]), std::min(r[8], r[10])};
        const float uhi[3] = {std::max(r[1], r[5]), std::max(r[3], r[7]), std::max(r[9], r[11])};
        p.cost += kNodeCost * area(ulo, uhi);
    }
    return p;
}

// Writes a planned top as the first live - 1 output nodes (root = node 0; the clustered
// nodes take the indices after them), with each cluster's triangle range in
// depth-first order.  Returns the deepest level among the single-triangle clusters
// placed here.
uint32_t apply_top(TopPlan& p, const NodeArrays& nd, const EmitArrays& em, DeviceBuffer& tris, DeviceBuffer& meta,
                   const uint32_t* order, DeviceBuffer& nodes_out, DeviceBuffer& tris_out, hipStream_t st) {
    const uint32_t live = p.live;
    BvhHost& top = p.top;
    // depth-first triangle offsets of the clusters (leaf order = order of top.tri_order)
    std::vector<uint32_t> offset(live);
    uint32_t run = 0;
    for (uint32_t q = 0; q < live; ++q) {
        const uint32_t c = top.tri_order[q];
        offset[c] = run;
        run += p.stats[c].x >> 1;
    }
    const uint32_t inner_total = (uint32_t)(nodes_out.bytes / (4 * sizeof(float4)));
    std::vector<ClusterSeed> seeds;
    seeds.reserve(live);
    uint32_t tri_depth = 0;
    // walk the top tree from the root: translate leaf children to the clusters' references
    std::vector<std::pair<uint32_t, uint32_t>> stack{{0u, 1u}};  // (node, level)
    while (!stack.empty()) {
        const auto [h, level] = stack.back();
        stack.pop_back();
        float* rec = &top.nodes[16 * (size_t)h];
        for (int s = 0; s < 2; ++s) {
            int32_t code;
            std::memcpy(&code, &rec[12 + s], 4);
            if (code >= 0) {
                stack.push_back({(uint32_t)code, level + 1});
                continue;
            }
            const uint32_t leaf = ~(uint32_t)code;
            const uint32_t c = top.tri_order[leaf >> kLeafCountBits];
            const int cn = p.node[c];
            const uint32_t count = p.stats[c].x >> 1;
            uint32_t ref;
            if (cn < 0) {
                ref = ~(offset[c] << kLeafCountBits);
                tri_depth = std::max(tri_depth, level + 1);
            } else if (p.stats[c].x & 1u) {
                ref = ~((offset[c] << kLeafCountBits) | (count - 1u));
            } else {
                ref = inner_total - 1 - (uint32_t)cn;
            }
            std::memcpy(&rec[12 + s], &ref, 4);
            seeds.push_back(ClusterSeed{cn, offset[c], level + 1, 0});
        }
    }
    PBRT_REQUIRE(seeds.size() == live, "PLOC: top tree leaves do not cover the clusters");
    DeviceBuffer seeds_d;
    PBRT_HIP(hipMemcpyAsync(nodes_out.ptr, top.nodes.data(), top.nodes.size() * sizeof(float), hipMemcpyHostToDevice,
                            st));
    seeds_d.upload(seeds.data(), seeds.size() * sizeof(ClusterSeed));
    hipLaunchKernelGGL(k_seed_clusters, dim3(blocks_for(live)), dim3(256), 0, st, seeds_d.as<ClusterSeed>(),
                       (int)live, nd, em, tris.as<float>(), meta.as<uint4>(), order, tris_out.as<float4>());
    PBRT_HIP(hipStreamSynchronize(st));  // the host vectors and seed buffer go out of scope
    return std::max(tri_depth, top.depth);
}

}  // namespace

PlocResult build_ploc(const std::vector<float>& world_tris, const std::vector<uint32_t>& inst,
                      const std::vector<uint32_t>& prim, const std::vector<uint32_t>& flags, int max_leaf,
                      int max_depth, DeviceBuffer& tris_out, DeviceBuffer& nodes_out) {
    const uint32_t n = (uint32_t)inst.size();
    PBRT_REQUIRE(n >= 2, "PLOC: needs at least two triangles");
    PBRT_REQUIRE(max_leaf >= 1 && max_leaf <= kMaxLeafSize, "PLOC: leaf size out of range");
    hipStream_t st = nullptr;
    struct Events {
        hipEvent_t e[2] = {nullptr, nullptr};
        ~Events() {
            for (hipEvent_t ev : e)
                if (ev) (void)hipEventDestroy(ev);
        }
    } ev;
    PBRT_HIP(hipEventCreate(&ev.e[0]));
    PBRT_HIP(hipEventCreate(&ev.e[1]));
    const uint32_t inner = n - 1;
    DeviceBuffer tris, meta, keys, vals, keys2, vals2, partial, sort_tmp, cn[2], cb[2], nnb, fl, pre, scan_tmp;
    DeviceBuffer child, box, info, cost, levels, offset, depth, state, maxd, total;
    tris.upload(world_tris.data(), world_tris.size() * sizeof(float));
    std::vector<uint4> m(n);
    for (uint32_t i = 0; i < n; ++i) m[i] = make_uint4(inst[i], prim[i], flags[i], 0);
    meta.upload(m.data(), m.size() * sizeof(uint4));
    keys.allocate(n * 4ull);
    vals.allocate(n * 4ull);
    keys2.allocate(n * 4ull);
    vals2.allocate(n * 4ull);
    const int rb = 512;
    partial.allocate(rb * 6 * 4);
    sort_tmp.allocate(gpusort::sort_scratch_bytes(n));
    for (int k = 0; k < 2; ++k) {
        cn[k].allocate(n * sizeof(int));
        cb[k].allocate(n * sizeof(Box));
    }
    nnb.allocate(n * sizeof(int));
    fl.allocate(n * sizeof(uint64_t));
    pre.allocate(n * sizeof(uint64_t));
    scan_tmp.allocate(gpusort::scan_scratch_bytes<uint64_t>(n));
    child.allocate(inner * sizeof(int2));
    box.allocate(inner * sizeof(Box));
    info.allocate(inner * sizeof(uint32_t));
    cost.allocate(inner * sizeof(float));
    levels.allocate(inner);
    offset.allocate(inner * sizeof(uint32_t));
    depth.allocate(inner * sizeof(uint32_t));
    state.allocate(inner * sizeof(uint32_t));
    maxd.allocate(sizeof(uint32_t));
    PBRT_HIP(hipMemset(state.ptr, 0, state.bytes));
    PBRT_HIP(hipMemset(maxd.ptr, 0, maxd.bytes));
    nodes_out.allocate((size_t)inner * 4 * sizeof(float4));
    tris_out.allocate((size_t)n * 3 * sizeof(float4));
    // nodes inside collapsed leaves are never written: zero boxes for the quantization pass
    PBRT_HIP(hipMemset(nodes_out.ptr, 0, nodes_out.bytes));
    PBRT_HIP(hipDeviceSynchronize());

    PBRT_HIP(hipEventRecord(ev.e[0], st));
    hipLaunchKernelGGL(k_centroid_box, dim3(rb), dim3(256), 0, st, tris.as<float>(), n, partial.as<float>());
    hipLaunchKernelGGL(k_fold_box, dim3(1), dim3(64), 0, st, partial.as<float>(), rb);
    hipLaunchKernelGGL(k_codes, dim3(blocks_for(n)), dim3(256), 0, st, tris.as<float>(), n, partial.as<float>(),
                       keys.as<uint32_t>(), vals.as<uint32_t>());
    const bool in2 = gpusort::sort_pairs(keys.as<uint32_t>(), vals.as<uint32_t>(), keys2.as<uint32_t>(),
                                         vals2.as<uint32_t>(), n, 30, sort_tmp.ptr, st);
    const uint32_t* order = in2 ? vals2.as<uint32_t>() : vals.as<uint32_t>();
    hipLaunchKernelGGL(k_leaf_clusters, dim3(blocks_for(n)), dim3(256), 0, st, tris.as<float>(), order, n,
                       cn[0].as<int>(), cb[0].as<Box>());

    NodeArrays nd{child.as<int2>(), box.as<Box>(), info.as<uint32_t>(), cost.as<float>(), levels.as<uint8_t>()};
    struct Range {
        int begin, count;
    };
    std::vector<Range> ranges;
    uint32_t live = n;
    int node_base = 0, cur = 0;
    uint64_t host_total[2] = {0, 0};
    // large scenes: clustering stops at kTopClusters clusters, whose tree is built by the
    // host's binned SAH (the top levels, which every ray visits, are where a 1-D window
    // over the Morton order chooses worst)
    DeviceBuffer snap_node, snap_box;
    uint32_t snap_live = 0;
    size_t snap_ranges = 0;
    while (live > 1) {
        if (!snap_live && live <= kTopClusters && n > 2 * kTopClusters) {
            // the clusters a host SAH top would start from
            snap_live = live;
            snap_ranges = ranges.size();
            snap_node.allocate(live * sizeof(int));
            snap_box.allocate(live * sizeof(Box));
            PBRT_HIP(hipMemcpyAsync(snap_node.ptr, cn[cur].ptr, live * sizeof(int), hipMemcpyDeviceToDevice, st));
            PBRT_HIP(hipMemcpyAsync(snap_box.ptr, cb[cur].ptr, live * sizeof(Box), hipMemcpyDeviceToDevice, st));
        }
        PBRT_REQUIRE(ranges.size() < 4096, "PLOC: clustering did not converge");
        const int nl = (int)live;
        hipLaunchKernelGGL(k_nearest, dim3(blocks_for(live)), dim3(256), 0, st, cb[cur].as<Box>(), nl, kRadius,
                           nnb.as<int>());
        hipLaunchKernelGGL(k_merge_flags, dim3(blocks_for(live)), dim3(256), 0, st, nnb.as<int>(), nl,
                           fl.as<uint64_t>());
        uint64_t* tot = gpusort::exclusive_scan<uint64_t>(fl.as<uint64_t>(), pre.as<uint64_t>(), live,
                                                          scan_tmp.as<uint64_t>(), st);
        hipLaunchKernelGGL(k_merge, dim3(blocks_for(live)), dim3(256), 0, st, nnb.as<int>(), fl.as<uint64_t>(),
                           pre.as<uint64_t>(), nl, cn[cur].as<int>(), cb[cur].as<Box>(), node_base, max_leaf, nd,
                           cn[1 - cur].as<int>(), cb[1 - cur].as<Box>());
        PBRT_HIP(hipMemcpyAsync(host_total, tot, sizeof(uint64_t), hipMemcpyDeviceToHost, st));
        PBRT_HIP(hipStreamSynchronize(st));
        const uint32_t merges = (uint32_t)(host_total[0] >> 32), survivors = (uint32_t)host_total[0];
        PBRT_REQUIRE(merges > 0 && survivors + merges == live, "PLOC: an iteration merged no clusters");
        ranges.push_back({node_base, (int)merges});
        node_base += (int)merges;
        live = survivors;
        cur = 1 - cur;
    }
    PBRT_REQUIRE(node_base == (int)inner, "PLOC: inner node count mismatch");
    EmitArrays em{offset.as<uint32_t>(), depth.as<uint32_t>(), state.as<uint32_t>()};
    // The tree's top: the clustering's own or, on large scenes, the host SAH's over the
    // clusters of the snapshot, whichever has the lower SAH cost (the host's wins on the
    // city, where the last clustering iterations pair big clusters through a narrow
    // window of the Morton order; the clustering's on the atrium, whose instances of
    // very different sizes a top-down SAH over cluster boxes splits badly).
    uint32_t host_depth = 0;
    size_t emit_ranges = ranges.size();
    bool host_top = false;
    if (snap_live > 1) {
        TopPlan plan = plan_top(snap_live, snap_node, snap_box, nd, max_depth, st);
        float root_cost = 0.0f;
        PBRT_HIP(hipMemcpy(&root_cost, nd.cost + (inner - 1), sizeof(float), hipMemcpyDeviceToHost));
        if (plan.cost < (double)root_cost) {
            host_depth = apply_top(plan, nd, em, tris, meta, order, nodes_out, tris_out, st);
            emit_ranges = snap_ranges;  // the nodes clustered after the snapshot are not used
            host_top = true;
        }
    }
    if (!host_top) hipLaunchKernelGGL(k_emit_root, dim3(1), dim3(1), 0, st, em, (int)inner - 1);
    for (size_t r = emit_ranges; r-- > 0;) {
        hipLaunchKernelGGL(k_emit_range, dim3(blocks_for(ranges[r].count)), dim3(256), 0, st, ranges[r].begin,
                           ranges[r].count, (int)inner, nd, em, tris.as<float>(), meta.as<uint4>(), order,
                           nodes_out.as<float4>(), tris_out.as<float4>(), maxd.as<uint32_t>());
    }
    PBRT_HIP(hipEventRecord(ev.e[1], st));
    PBRT_HIP(hipGetLastError());
    PBRT_HIP(hipEventSynchronize(ev.e[1]));
    float ms = 0.0f;
    PBRT_HIP(hipEventElapsedTime(&ms, ev.e[0], ev.e[1]));
    uint32_t d = 0;
    PBRT_HIP(hipMemcpy(&d, maxd.ptr, sizeof(uint32_t), hipMemcpyDeviceToHost));
    d = std::max(d, host_depth);
    if ((int)d > max_depth)
        fail(PBRT_ERR_UNSUPPORTED, "pbrt_scene_create: PLOC tree deeper than the traversal stack (" +
                                       std::to_string(d) + " levels); use PBRT_BVH_SAH_HOST for this scene");
    PlocResult res{};
    res.node_count = inner;
    res.depth = d;
    res.iterations = (uint32_t)ranges.size();
    res.build_ms = ms;
    return res;
}

}  // namespace pbrt

using namespace pbrt;

extern "C" int32_t pbrt_radix_sort_scratch_size(uint32_t n, size_t* out_bytes) {
    return guard([&] {
        PBRT_REQUIRE(out_bytes, "pbrt_radix_sort_scratch_size: NULL argument");
        *out_bytes = gpusort::sort_scratch_bytes(n > 0 ? n : 1);
    });
}

extern "C" int32_t pbrt_radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t* keys_tmp, uint32_t* vals_tmp,
                                         uint32_t n, uint32_t bits, void* scratch, size_t scratch_bytes,
                                         void* stream) {
    return guard([&] {
        if (n == 0 || bits == 0) return;
        PBRT_REQUIRE(keys && vals && keys_tmp && vals_tmp && scratch, "pbrt_radix_sort_pairs: NULL argument");
        PBRT_REQUIRE(bits <= 32, "pbrt_radix_sort_pairs: more than 32 key bits");
        PBRT_REQUIRE(scratch_bytes >= gpusort::sort_scratch_bytes(n), "pbrt_radix_sort_pairs: scratch too small");
        hipStream_t st = (hipStream_t)stream;
        if (gpusort::sort_pairs(keys, vals, keys_tmp, vals_tmp, n, (int)bits, scratch, st)) {
            PBRT_HIP(hipMemcpyAsync(keys, keys_tmp, (size_t)n * 4, hipMemcpyDeviceToDevice, st));
            PBRT_HIP(hipMemcpyAsync(vals, vals_tmp, (size_t)n * 4, hipMemcpyDeviceToDevice, st));
        }
        PBRT_HIP(hipGetLastError());
    });
}
