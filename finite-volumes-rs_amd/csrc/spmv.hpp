// spmv.hpp -- SpMV launch interface shared by the solvers (internal).
#pragma once

#include "internal.hpp"
#include "kstate.hpp"

namespace fvg {

// Optional work fused after y = A x, per row r:
//   if b:        y_r = (y_r - b[r]) * -1      (residual r = b - A x, solvers' setup)
//   copy1/copy2: also store y_r there
//   partials slot 0: sum of dotw[r] * y_r     (if dotw)
//   partials slot 1: sum of y_r * y_r         (if dot_self)
//   partials slot 2: sum of dotw2[r] * y_r    (if dotw2)
// (slot k at partials[k * stride + block], or with `dd` the pair of put_dd)
struct SpmvEpilogue {
    const double* b = nullptr;
    double* copy1 = nullptr;
    double* copy2 = nullptr;
    const double* dotw = nullptr;
    const double* dotw2 = nullptr;
    bool dot_self = false;
    // if set, a workgroup does nothing when skip[member * skip_stride] != 0, member being
    // blk_inst[workgroup] (or 0 without blk_inst)
    const int32_t* skip = nullptr;
    int32_t skip_stride = 0;
    const int32_t* blk_inst = nullptr;
    // partial sums into partials[slot * pstride + block] instead of the context's
    // (caller-provided storage of at least 3 slots of spmv_blocks(n); nothing is allocated)
    double* partials = nullptr;
    int64_t pstride = 0;
    // the dots as double-double pairs (put_dd / sum_partials_dd layout) rather than doubles
    bool dd = false;
    // x read in reverse: x[xrev - 1 - column] (a vector kept reversed; 0: as stored)
    int64_t xrev = 0;
};

// Number of partial sums one fused SpMV writes per dot (one per workgroup).
int spmv_blocks(int64_t nrows);
// Ensures ctx->partials holds kPartialSlots slots of `count` partials each.
int ensure_partials(fvg_ctx* ctx, int64_t count);
// y = A x with the epilogue; partials go to ctx->partials[slot * stride + block].
int launch_spmv(fvg_ctx* ctx, const fvg_csr* a, const double* x, double* y, const SpmvEpilogue& epi);
int64_t partial_stride(const fvg_ctx* ctx);

}  // namespace fvg
