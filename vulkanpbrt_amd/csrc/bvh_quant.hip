// Quantized BVH2 nodes: 32 bytes per node instead of 64.
//
// Traversal of incoherent rays on gfx950 is bound by the vector memory path's
// per-lane cache-line lookups: every 16-byte load of a lane is one lookup, so a
// node's cost is its size in 16-byte pieces.  The builders (host SAH, device
// LBVH) emit full-precision nodes; this pass rewrites each node in place of
// its index as
//
//   word 0..2  node origin x, y, z: fp32 rounded towards -inf to 16 significant
//              bits, whose low 8 bits carry the axis' biased scale exponent
//              e + 127 (scale s = 2^e)
//   word 3     child 0 reference (inner node index, or ~leaf code)
//   word 4..6  12 bytes: child 0 lo.xyz, hi.xyz, child 1 lo.xyz, hi.xyz as
//              integers q in [0, 255]; the bound is fma(q, s, origin)
//   word 7     child 1 reference
//
// Each child's bounds are rounded outwards and checked with the same fmaf the
// traversal evaluates, so every dequantized box contains its child's exact box.
// A node is a function of its own two child boxes only, so the pass is one
// thread per node.
#include <hip/hip_runtime.h>

#include <cmath>

#include "bvh_quant.hpp"

namespace pbrt {

namespace {

// fp32 bits rounded towards -inf with the low 8 significand bits cleared
__device__ uint32_t origin_bits_down(float f) {
    uint32_t b = __float_as_uint(f);
    uint32_t t = b & ~0xFFu;
    if ((b & 0x80000000u) && (b & 0xFFu)) t += 0x100u;  // negative: round the magnitude up
    return t;
}

__device__ __forceinline__ float pow2(int e) { return __uint_as_float((uint32_t)(e + 127) << 23); }

// Quantizes [lo, hi] against origin o and scale 2^e; false if 255 steps do not reach hi.
__device__ bool quantize_range(float lo, float hi, float o, int e, uint32_t& qlo, uint32_t& qhi) {
    const float s = pow2(e);
    double a = floor(((double)lo - (double)o) / (double)s);
    double b = ceil(((double)hi - (double)o) / (double)s);
    int ql = (int)fmin(fmax(a, 0.0), 255.0);
    int qh = (int)fmin(fmax(b, 0.0), 256.0);
    while (ql > 0 && fmaf((float)ql, s, o) > lo) --ql;
    while (qh <= 255 && fmaf((float)qh, s, o) < hi) ++qh;
    if (qh > 255) return false;
    qlo = (uint32_t)ql;
    qhi = (uint32_t)qh;
    return true;
}

__global__ void k_quantize_nodes(const float4* __restrict__ in, uint4* __restrict__ out, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 n0 = in[4 * (size_t)i + 0], n1 = in[4 * (size_t)i + 1], n2 = in[4 * (size_t)i + 2],
                 n3 = in[4 * (size_t)i + 3];
    float lo[2][3] = {{n0.x, n0.z, n2.x}, {n1.x, n1.z, n2.z}};
    float hi[2][3] = {{n0.y, n0.w, n2.y}, {n1.y, n1.w, n2.w}};
    bool empty[2];
    for (int c = 0; c < 2; ++c) empty[c] = !(lo[c][0] <= hi[c][0] && lo[c][1] <= hi[c][1] && lo[c][2] <= hi[c][2]);
    uint32_t w[3], q[2][2][3] = {};
    for (int a = 0; a < 3; ++a) {
        float nlo = INFINITY, nhi = -INFINITY;
        for (int c = 0; c < 2; ++c)
            if (!empty[c]) {
                nlo = fminf(nlo, lo[c][a]);
                nhi = fmaxf(nhi, hi[c][a]);
            }
        if (!(nlo <= nhi)) nlo = nhi = 0.0f;  // both children empty: a point box at 0
        const uint32_t ob = origin_bits_down(nlo);
        const float o = __uint_as_float(ob);
        const double ext = (double)nhi - (double)o;
        int e = ext > 0.0 ? (int)ceil(log2(ext / 254.0)) : -126;
        e = e < -126 ? -126 : (e > 127 ? 127 : e);
        for (;; ++e) {
            bool ok = true;
            for (int c = 0; c < 2 && ok; ++c) {
                if (empty[c]) continue;  // stays the point box (0, 0)
                ok = quantize_range(lo[c][a], hi[c][a], o, e, q[c][0][a], q[c][1][a]);
            }
            if (ok || e >= 127) break;
        }
        w[a] = ob | (uint32_t)(e + 127);
    }
    auto pack = [](uint32_t b0, uint32_t b1, uint32_t b2, uint32_t b3) { return b0 | (b1 << 8) | (b2 << 16) | (b3 << 24); };
    out[2 * (size_t)i + 0] = make_uint4(w[0], w[1], w[2], __float_as_uint(n3.x));
    out[2 * (size_t)i + 1] =
        make_uint4(pack(q[0][0][0], q[0][0][1], q[0][0][2], q[0][1][0]), pack(q[0][1][1], q[0][1][2], q[1][0][0], q[1][0][1]),
                   pack(q[1][0][2], q[1][1][0], q[1][1][1], q[1][1][2]), __float_as_uint(n3.y));
}

}  // namespace

void quantize_bvh_nodes(DeviceBuffer& nodes, uint32_t node_count) {
    DeviceBuffer q;
    q.allocate((size_t)node_count * 32);
    hipLaunchKernelGGL(k_quantize_nodes, dim3((node_count + 255) / 256), dim3(256), 0, nullptr, nodes.as<float4>(),
                       q.as<uint4>(), node_count);
    PBRT_HIP(hipGetLastError());
    PBRT_HIP(hipDeviceSynchronize());
    nodes.swap(q);
}

}  // namespace pbrt
