// Device LBVH builder (lbvh.hip).
#pragma once

#include <cstdint>
#include <vector>

#include "capi_common.hpp"

namespace pbrt {

struct LbvhResult {
    uint32_t node_count;  // n - 1 inner nodes (subtrees folded into leaves stay allocated)
    uint32_t depth;       // levels of the emitted tree, leaves included
    uint32_t max_leaf;    // largest leaf: 4, or 8 / 16 when needed to fit max_depth
    float build_ms;       // device time from bounds to triangle records
};

// world_tris: 9 floats per triangle; inst / prim / flags per triangle.  Writes
// the triangle records (3 float4 each, BVH order) and nodes (4 float4 each).
LbvhResult build_lbvh(const std::vector<float>& world_tris, const std::vector<uint32_t>& inst,
                      const std::vector<uint32_t>& prim, const std::vector<uint32_t>& flags, int max_depth,
                      DeviceBuffer& tris_out, DeviceBuffer& nodes_out);

}  // namespace pbrt
