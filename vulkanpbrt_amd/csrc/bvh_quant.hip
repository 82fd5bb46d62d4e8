// Quantized BVH2 nodes: 32 bytes per node instead of 64.
//
// Traversal of incoherent rays on gfx950 is bound by the vector memory path's
// per-lane cache-line lookups: every 16-byte load of a lane is one lookup, so a
// node's cost is its size in 16-byte pieces.  The builders (host SAH, device
// LBVH) emit full-precision nodes; this pass rewrites each node in place of
// its index as
//
//   word 0..2  node origin x, y, z: an fp32 value at or below the node's lower
//              bound whose low 8 significand bits hold e + 24, signed, for the
//              axis' scale 2^e (e in [-126, 103]); the word is the origin as it
//              stands, byte included
//   word 3     child 0 reference (inner node index, or ~leaf code)
//   word 4..6  12 bytes of integers q in [0, 255], grouped by axis: word 4 =
//              child 0 lo x, hi x, lo y, hi y; word 5 = child 0 lo z, hi z,
//              child 1 lo z, hi z; word 6 = child 1 lo x, hi x, lo y, hi y; the
//              bound is origin + q 2^e
//   word 7     child 1 reference
//
// Each child's bounds are rounded outwards and checked in FP64 (where origin + q s
// is exact unless the scale is below 2^-29 of the origin's own last bit, i.e. for
// boxes flat to within a fraction of an fp32 ulp), so every box origin + [qlo, qhi] s
// contains its child's box.  The traversal never forms the bound itself: it meets
// the plane at fma(q 2^-24, 2^(e + 24) / d, fma(origin, 1/d, -o/d)), q 2^-24 an
// fp16 subnormal.  A node is a function of
// its own two child boxes only, so the pass is one thread per node.
#include <hip/hip_runtime.h>

#include <cmath>

#include "bvh_quant.hpp"

namespace pbrt {

namespace {

// Largest scale exponent: e + 24 fits the origin word's signed byte.
constexpr int kMaxScaleExp = 103;

// The origin word of an axis whose lower bound is f and whose scale exponent is e:
// the fp32 value with the byte e + 24 as its low 8 significand bits that is
// closest to f from below (the byte adds to a negative value's magnitude, to a
// positive one's value); zero, denormals and values next to them take -2^-126 -
// byte ulps (normal, so no denormal origin reaches the traversal).
__device__ uint32_t origin_word(float f, int e) {
    const uint32_t byte = (uint32_t)(e + 24) & 0xFFu, b = __float_as_uint(f), low = b & 0xFFu;
    uint32_t t = b & ~0xFFu;
    if (b & 0x80000000u) {
        if (byte < low) t += 0x100u;
    } else if (byte > low) {
        t -= 0x100u;  // 0 - 0x100 wraps to a NaN pattern, caught below
    }
    const bool normal = (t & 0x7F800000u) != 0u && (t & 0x7F800000u) != 0x7F800000u;
    return (normal ? t : 0x80800000u) | byte;
}

// Quantizes [lo, hi] against origin o and scale 2^e: o + qlo 2^e <= lo and
// o + qhi 2^e >= hi in exact arithmetic; false if 255 steps do not reach hi.
__device__ bool quantize_range(float lo, float hi, float o, int e, uint32_t& qlo, uint32_t& qhi) {
    // scaling by powers of two is exact: no FP64 division
    const double s = ldexp(1.0, e), inv_s = ldexp(1.0, -e), od = (double)o;
    const double a = floor(((double)lo - od) * inv_s);
    const double b = ceil(((double)hi - od) * inv_s);
    int ql = (int)fmin(fmax(a, 0.0), 255.0);
    int qh = (int)fmin(fmax(b, 0.0), 256.0);
    while (ql > 0 && od + ql * s > (double)lo) --ql;
    while (qh <= 255 && od + qh * s < (double)hi) ++qh;
    if (qh > 255) return false;
    qlo = (uint32_t)ql;
    qhi = (uint32_t)qh;
    return true;
}

// ceil(log2(x)) for x > 0, from the binary exponent (no transcendental)
__device__ int ceil_log2(double x) {
    const int e = ilogb(x);
    return x > ldexp(1.0, e) ? e + 1 : e;
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
        // The first exponent tried (the loop below moves up until both children
        // fit).  The origin lies up to ~512 ulps below nlo, so the span to cover is
        // at least that much even on a flat axis (whose own extent, 0, would start
        // the search a hundred exponents too low).
        const double ext = fmax((double)nhi - (double)nlo, fabs((double)nlo) * 0x1p-14 + 0x1p-120);
        int e = ceil_log2(ext * (1.0 / 254.0));
        e = e < -126 ? -126 : (e > kMaxScaleExp ? kMaxScaleExp : e);
        for (;; ++e) {
            const float o = __uint_as_float(origin_word(nlo, e));
            bool ok = true;
            for (int c = 0; c < 2 && ok; ++c) {
                if (empty[c]) continue;  // stays the point box (0, 0)
                ok = quantize_range(lo[c][a], hi[c][a], o, e, q[c][0][a], q[c][1][a]);
            }
            if (ok || e >= kMaxScaleExp) break;
        }
        w[a] = origin_word(nlo, e);
    }
    auto pack = [](uint32_t b0, uint32_t b1, uint32_t b2, uint32_t b3) { return b0 | (b1 << 8) | (b2 << 16) | (b3 << 24); };
    out[2 * (size_t)i + 0] = make_uint4(w[0], w[1], w[2], __float_as_uint(n3.x));
    out[2 * (size_t)i + 1] =
        make_uint4(pack(q[0][0][0], q[0][1][0], q[0][0][1], q[0][1][1]), pack(q[0][0][2], q[0][1][2], q[1][0][2], q[1][1][2]),
                   pack(q[1][0][0], q[1][1][0], q[1][0][1], q[1][1][1]), __float_as_uint(n3.y));
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
