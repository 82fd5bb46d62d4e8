// Device PLOC builder (ploc.hip).
#pragma once

#include <cstdint>
#include <vector>

#include "capi_common.hpp"

namespace pbrt {

struct PlocResult {
    uint32_t node_count;  // n - 1 inner nodes (nodes inside collapsed leaves stay allocated)
    uint32_t depth;       // levels of the emitted tree, leaves included
    uint32_t iterations;  // clustering iterations
    float build_ms;       // device time from the centroid bounds to the last record
};

// world_tris: 9 floats per triangle; inst / prim / flags per triangle.  Writes the
// triangle records (3 float4 each, BVH order) and nodes (4 float4 each, root first).
// Leaves hold up to max_leaf triangles where the SAH prefers them.
PlocResult build_ploc(const std::vector<float>& world_tris, const std::vector<uint32_t>& inst,
                      const std::vector<uint32_t>& prim, const std::vector<uint32_t>& flags, int max_leaf,
                      int max_depth, DeviceBuffer& tris_out, DeviceBuffer& nodes_out);

}  // namespace pbrt
