// Device-wide exclusive scan and stable LSD radix sort of (key, value) pairs on
// gfx950, used by the GPU BVH builders (Morton-code sort, cluster compaction).
//
// Scan: tiles of 4096 items (256 threads x 16 consecutive items); per-tile sums,
// one workgroup scans the sums, each tile then scans itself from its offset.
// Radix sort: 8-bit digits.  Per pass, each 4096-key tile counts its digits
// (k_sort_hist, digit-major so that one exclusive scan gives every tile its
// output base per digit), then k_sort_scatter ranks the tile's keys in 16 rounds
// of 256 in input order: within a wave, lanes holding the same digit find each
// other with eight ballots (one per digit bit) and take their rank from the
// lanes below them; across the four waves a per-digit running offset in LDS
// keeps input order, so the sort is stable.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "capi_common.hpp"

namespace pbrt {
namespace gpusort {
// internal linkage: every translation unit that builds trees gets its own kernels
namespace {

constexpr int kThreads = 256;
constexpr int kItems = 16;
constexpr int kTileItems = kThreads * kItems;

inline unsigned tiles_for(size_t n) { return (unsigned)((n + kTileItems - 1) / kTileItems); }

// inclusive scan of one value per lane over a 256-thread workgroup; `wsum` holds
// 4 entries of LDS; returns the workgroup total in `total`
template <class T>
__device__ __forceinline__ T block_inclusive_scan(T v, T* wsum, T& total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const T o = __shfl_up(v, off);
        if (lane >= off) v += o;
    }
    if (lane == 63) wsum[wave] = v;
    __syncthreads();
    T before = 0;
    for (int w = 0; w < wave; ++w) before += wsum[w];
    total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    __syncthreads();
    return v + before;
}

template <class T>
__global__ void __launch_bounds__(kThreads) k_scan_reduce(const T* __restrict__ in, size_t n, T* __restrict__ sums) {
    __shared__ T wsum[4];
    const size_t base = (size_t)blockIdx.x * kTileItems + (size_t)threadIdx.x * kItems;
    T s = 0;
#pragma unroll
    for (int k = 0; k < kItems; ++k)
        if (base + k < n) s += in[base + k];
    T total;
    (void)block_inclusive_scan(s, wsum, total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// exclusive scan of `count` tile sums in place by one workgroup; total -> sums[count]
template <class T>
__global__ void __launch_bounds__(kThreads) k_scan_top(T* sums, unsigned count) {
    __shared__ T wsum[4];
    const unsigned per = (count + kThreads - 1) / kThreads;
    const unsigned lo = threadIdx.x * per, hi = min(count, lo + per);
    T s = 0;
    for (unsigned i = lo; i < hi; ++i) s += sums[i];
    T total;
    const T incl = block_inclusive_scan(s, wsum, total);
    T run = incl - s;
    for (unsigned i = lo; i < hi; ++i) {
        const T v = sums[i];
        sums[i] = run;
        run += v;
    }
    if (threadIdx.x == 0) sums[count] = total;
}

template <class T>
__global__ void __launch_bounds__(kThreads) k_scan_down(const T* __restrict__ in, size_t n, const T* __restrict__ sums,
                                                        T* __restrict__ out) {
    __shared__ T wsum[4];
    const size_t base = (size_t)blockIdx.x * kTileItems + (size_t)threadIdx.x * kItems;
    T v[kItems];
    T s = 0;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        v[k] = base + k < n ? in[base + k] : T(0);
        s += v[k];
    }
    T total;
    T run = block_inclusive_scan(s, wsum, total) - s + sums[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kItems; ++k)
        if (base + k < n) {
            out[base + k] = run;
            run += v[k];
        }
}

// Scratch of an exclusive scan of n items: (tiles + 1) sums.
template <class T>
inline size_t scan_scratch_bytes(size_t n) {
    return ((size_t)tiles_for(n) + 1) * sizeof(T);
}

// out[i] = in[0] + ... + in[i - 1]; the total lands in sums[tiles_for(n)] (device).
// `in` and `out` may be the same array.
template <class T>
inline T* exclusive_scan(const T* in, T* out, size_t n, T* sums, hipStream_t st) {
    const unsigned tiles = tiles_for(n);
    hipLaunchKernelGGL(k_scan_reduce<T>, dim3(tiles), dim3(kThreads), 0, st, in, n, sums);
    hipLaunchKernelGGL(k_scan_top<T>, dim3(1), dim3(kThreads), 0, st, sums, tiles);
    hipLaunchKernelGGL(k_scan_down<T>, dim3(tiles), dim3(kThreads), 0, st, in, n, (const T*)sums, out);
    return sums + tiles;
}

__global__ void __launch_bounds__(kThreads) k_sort_hist(const uint32_t* __restrict__ keys, uint32_t n, int shift,
                                                        uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * kTileItems;
#pragma unroll 4
    for (int k = 0; k < kItems; ++k) {
        const uint32_t i = base + k * kThreads + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * gridDim.x + blockIdx.x] = h[threadIdx.x];
}

__global__ void __launch_bounds__(kThreads) k_sort_scatter(const uint32_t* __restrict__ keys,
                                                           const uint32_t* __restrict__ vals, uint32_t n, int shift,
                                                           const uint32_t* __restrict__ offsets,
                                                           uint32_t* __restrict__ okeys, uint32_t* __restrict__ ovals) {
    __shared__ uint32_t run[256];      // next output position per digit
    __shared__ uint32_t wcnt[4][256];  // this round: digit counts per wave, then their output bases
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    run[tid] = offsets[tid * gridDim.x + blockIdx.x];
    const unsigned long long below = (1ull << lane) - 1ull;
    const uint32_t base = blockIdx.x * kTileItems;
    for (int round = 0; round < kItems; ++round) {
        const uint32_t i = base + round * kThreads + tid;
        const bool valid = i < n;
        const uint32_t key = valid ? keys[i] : 0u;
        const uint32_t d = (key >> shift) & 255u;
#pragma unroll
        for (int w = 0; w < 4; ++w) wcnt[w][tid] = 0;
        __syncthreads();
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const unsigned long long bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        const uint32_t rank = (uint32_t)__popcll(peers & below);
        if (valid && rank == 0) wcnt[wave][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        {
            uint32_t r = run[tid];
#pragma unroll
            for (int w = 0; w < 4; ++w) {
                const uint32_t c = wcnt[w][tid];
                wcnt[w][tid] = r;
                r += c;
            }
            run[tid] = r;
        }
        __syncthreads();
        if (valid) {
            const uint32_t pos = wcnt[wave][d] + rank;
            okeys[pos] = key;
            ovals[pos] = vals[i];
        }
        __syncthreads();
    }
}

// Scratch of sort_pairs for n keys: the digit histograms of every tile and their scan sums.
inline size_t sort_scratch_bytes(size_t n) {
    const size_t h = (size_t)256 * tiles_for(n);
    return h * sizeof(uint32_t) + scan_scratch_bytes<uint32_t>(h) + 256;
}

// Stable sort of n (key, value) pairs by key bits [0, bits); the pairs ping-pong
// between (keys, vals) and (keys2, vals2); returns true when the result is in
// (keys2, vals2), false when it is back in (keys, vals).
inline bool sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t* keys2, uint32_t* vals2, uint32_t n, int bits,
                       void* scratch, hipStream_t st) {
    const unsigned tiles = tiles_for(n);
    uint32_t* hist = static_cast<uint32_t*>(scratch);
    const size_t hn = (size_t)256 * tiles;
    uint32_t* sums = hist + ((hn + 63) & ~(size_t)63);
    bool flipped = false;
    for (int shift = 0; shift < bits; shift += 8) {
        uint32_t *ik = flipped ? keys2 : keys, *iv = flipped ? vals2 : vals;
        uint32_t *ok = flipped ? keys : keys2, *ov = flipped ? vals : vals2;
        hipLaunchKernelGGL(k_sort_hist, dim3(tiles), dim3(kThreads), 0, st, ik, n, shift, hist);
        exclusive_scan<uint32_t>(hist, hist, hn, sums, st);
        hipLaunchKernelGGL(k_sort_scatter, dim3(tiles), dim3(kThreads), 0, st, ik, iv, n, shift, hist, ok, ov);
        flipped = !flipped;
    }
    return flipped;
}

}  // namespace
}  // namespace gpusort
}  // namespace pbrt
