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

// ---- double-double accumulation of FP64 dot products ------------------------------------
// The FP64 solvers sum their dot products' (rounded) products in double-double: an
// unevaluated pair hi + lo carries the rounding error of every addition, so the final
// rounding to double is that of the exact sum of the products for sums without heavy
// cancellation (a dot of a vector with itself, r.z with a positive definite
// preconditioner).  The result then no longer depends on how the terms were grouped (launch
// shape, fused or separate kernels), which keeps a long, ill-conditioned iteration on the
// exact-sum trajectory.  Error-free transformations: TwoSum (Knuth, TAOCP vol. 2, 4.2.2),
// FastTwoSum (Dekker, Numer. Math. 18, 1971); pair addition: Joldes, Muller and Popescu,
// ACM TOMS 44 (2017), algorithms 4 and 6.  Built without contraction, as the whole library.
struct DD {
    double hi, lo;
};
__device__ __forceinline__ DD two_sum(double a, double b) {
    const double s = a + b;
    const double bb = s - a;
    return {s, (a - (s - bb)) + (b - bb)};
}
__device__ __forceinline__ DD fast_two_sum(double a, double b) {   // |a| >= |b| or a == 0
    const double s = a + b;
    return {s, b - (s - a)};
}
__device__ __forceinline__ DD dd_add(DD x, DD y) {
    const DD s = two_sum(x.hi, y.hi);
    const DD t = two_sum(x.lo, y.lo);
    const DD v = fast_two_sum(s.hi, s.lo + t.hi);
    return fast_two_sum(v.hi, t.lo + v.lo);
}
__device__ __forceinline__ DD dd_add(DD x, double y) {
    const DD s = two_sum(x.hi, y);
    return fast_two_sum(s.hi, x.lo + s.lo);
}
__device__ __forceinline__ DD wave_sum(DD v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = dd_add(v, DD{__shfl_down(v.hi, off, 64), __shfl_down(v.lo, off, 64)});
    return v;   // lane 0 holds the wave total
}
// Sum over a workgroup of NT threads (lds: 2 * NT / 64 doubles); valid in thread 0.
template <int NT>
__device__ __forceinline__ DD block_sum(DD v, double* lds) {
    static_assert(NT % 64 == 0, "wave64 multiple");
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
        lds[2 * w] = v.hi;
        lds[2 * w + 1] = v.lo;
    }
    __syncthreads();
    DD s{0.0, 0.0};
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NT / 64; ++i) s = dd_add(s, DD{lds[2 * i], lds[2 * i + 1]});
    }
    __syncthreads();
    return s;
}
// Double-double partials: the pair of workgroup b in slot k sits at p[2k * stride + 2b] (hi)
// and the double after it (lo), so slot k spans 2 * stride doubles.
__device__ __forceinline__ void put_dd(double* partials, int64_t stride, int k, int b, DD v) {
    reinterpret_cast<double2*>(partials + 2 * k * stride)[b] = double2{v.hi, v.lo};
}
// Fixed-order double-double sum of `count` pairs from p (a slot's start); valid in thread 0.
template <int NT>
__device__ __forceinline__ DD sum_partials_dd(const double* __restrict__ p, int count, double* lds) {
    DD a{0.0, 0.0};
    for (int i = threadIdx.x; i < count; i += NT) {
        const double2 q = reinterpret_cast<const double2*>(p)[i];
        a = dd_add(a, DD{q.x, q.y});
    }
    return block_sum<NT>(a, lds);
}

}  // namespace fvg
