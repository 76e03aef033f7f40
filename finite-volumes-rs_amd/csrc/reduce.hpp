// reduce.hpp -- deterministic wave64 / workgroup reductions shared by the kernels.
//
// Every reduction here has a fixed summation tree for a given launch shape, so a
// dot product of the same vectors gives the same bits on every run and does not
// depend on what else is resident on the GPU.
#pragma once

#include <hip/hip_runtime.h>

namespace fvg {

// The dispatcher hands workgroup k to XCD k mod 8.  xcd_tile maps workgroup k to the
// tile it works on so that each XCD gets a contiguous run of tiles: neighbouring tiles
// (which read overlapping parts of a gathered vector) then share one L2.  A bijection
// on [0, g); only the schedule changes, never the arithmetic.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_tile(int k, int g) {
    const int x = k % kXcds, local = k / kXcds;
    const int per = g / kXcds, rem = g % kXcds;
    return x * per + min(x, rem) + local;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;  // lane 0 holds the wave total
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// Sum over a workgroup of NT threads; the total is valid in thread 0.
template <int NT, class T>
__device__ __forceinline__ T block_sum(T v, T* lds) {
    static_assert(NT % 64 == 0, "wave64 multiple");
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) lds[w] = v;
    __syncthreads();
    T s = T(0);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NT / 64; ++i) s += lds[i];
    }
    __syncthreads();
    return s;
}

// Fixed-order sum of `count` partials: thread t owns elements t, t+NT, t+2NT, ...
// summed with four independent accumulators (so the loads overlap), then the block tree.
template <int NT>
__device__ __forceinline__ double sum_partials(const double* __restrict__ p, int count, double* lds) {
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int i = threadIdx.x;
    for (; i + 3 * NT < count; i += 4 * NT) {
        a0 += p[i];
        a1 += p[i + NT];
        a2 += p[i + 2 * NT];
        a3 += p[i + 3 * NT];
    }
    for (; i < count; i += NT) a0 += p[i];
    return block_sum<NT>((a0 + a1) + (a2 + a3), lds);
}

}  // namespace fvg
