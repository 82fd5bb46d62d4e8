#pragma once

#include <cstdint>
#include <vector>

namespace pbrt {

constexpr int kLeafBits = 4;             // leaf triangle count stored as count - 1 in 4 bits
constexpr int kMaxLeaf = 1 << kLeafBits;

struct BvhHost {
    std::vector<float> nodes;         // 16 floats per node (see scene_types.hpp)
    std::vector<uint32_t> tri_order;  // BVH position -> input triangle index (a triangle may
                                      // appear more than once after spatial splits)
    uint32_t depth = 0;               // levels of inner nodes + leaf level
};

// tri_lo / tri_hi: 3 floats per triangle (axis-aligned bounds).  `weight` (optional):
// each primitive's intersection cost in triangle tests, for primitives that are
// whole subtrees (the GPU PLOC builder's clusters); unit cost when NULL.
BvhHost build_bvh_sah(const float* tri_lo, const float* tri_hi, uint32_t n, int max_leaf, int max_depth,
                      const float* weight = nullptr);
// tris: 9 floats per triangle (three vertices); spatial splits may reference a
// triangle from several leaves.
BvhHost build_bvh_sbvh(const float* tris, uint32_t n, int max_leaf, int max_depth);

}  // namespace pbrt
