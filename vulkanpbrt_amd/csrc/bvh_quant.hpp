// Quantized BVH2 nodes (bvh_quant.hip).
#pragma once

#include <cstdint>

#include "capi_common.hpp"

namespace pbrt {

// Rewrites `nodes` (node_count full-precision nodes of 4 x float4, as the
// builders emit them) into the 2 x uint4 quantized layout the traversal reads.
void quantize_bvh_nodes(DeviceBuffer& nodes, uint32_t node_count);

}  // namespace pbrt
