// csr.hip -- DistributedMatrix<f64> on the device and its SpMV (dmatrix.rs:136-146).
//
// SpMV is a row-block CSR-stream kernel: a workgroup owns kSpmvRows consecutive rows,
// streams their contiguous (column, value) range from HBM with fully coalesced
// loads, forms each product v*x[col] once into LDS, and then every thread sums its
// own row from LDS in ascending column order starting from 0.0 -- the reference's
// sequential `result += value * x` loop, so the result is bit-identical to it.
#include <cstring>

#include "reduce.hpp"
#include "spmv.hpp"

namespace fvg {

constexpr int kSpmvRows = 256;     // rows (and threads) per workgroup
constexpr int kSpmvStage = 2048;   // staged products per workgroup (16 KiB of LDS)
constexpr int kSpmvUnroll = 6;     // products per thread whose loads are in flight together

int spmv_blocks(int64_t nrows) { return grid_for(nrows, kSpmvRows); }

int64_t partial_stride(const fvg_ctx* ctx) { return ctx->partials_cap; }

int ensure_partials(fvg_ctx* ctx, int64_t count) {
    if (count <= ctx->partials_cap) return FVG_OK;
    FVG_HIP(hipStreamSynchronize(ctx->stream));
    dev_free_n(ctx, ctx->partials, ctx->partials_cap * kPartialSlots);
    ctx->partials_cap = 0;
    FVG_TRY(dev_alloc_n(ctx, count * kPartialSlots, &ctx->partials));
    ctx->partials_cap = count;
    return FVG_OK;
}

template <bool EPI>
__global__ void __launch_bounds__(kSpmvRows) k_spmv(int n, const int32_t* __restrict__ rp,
                                                    const int32_t* __restrict__ ci, const double* __restrict__ v,
                                                    const double* __restrict__ x, double* __restrict__ y,
                                                    SpmvEpilogue epi, double* __restrict__ partials,
                                                    int64_t pstride) {
    __shared__ double prod[kSpmvStage];
    __shared__ double red[2 * kSpmvRows / 64];
    const int blk = xcd_tile(blockIdx.x, gridDim.x);
    const int r0 = blk * kSpmvRows;
    const int r1 = min(r0 + kSpmvRows, n);
    // the tile's and this thread's row bounds are loaded together with the skip flag
    const int base = rp[r0];
    const int cnt = rp[r1] - base;
    const int r = r0 + (int)threadIdx.x;
    const int rc = min(r, r1 - 1);
    const int rb = rp[rc], re = rp[rc + 1];
    if (EPI && epi.skip && epi.skip[(epi.blk_inst ? epi.blk_inst[blk] : 0) * epi.skip_stride]) return;
    double sum = 0.0;
    if (cnt <= kSpmvStage) {
        // every thread issues all its entry loads, then all its gathers, before using any
        // (indices clamped into the block's range, so no load is conditional)
        for (int k0 = threadIdx.x; k0 < cnt; k0 += kSpmvRows * kSpmvUnroll) {
            int c[kSpmvUnroll];
            double a[kSpmvUnroll], xv[kSpmvUnroll];
#pragma unroll
            for (int u = 0; u < kSpmvUnroll; ++u) {
                const int k = min(k0 + u * kSpmvRows, cnt - 1);
                c[u] = __builtin_nontemporal_load(ci + base + k);   // streamed once: keep x in L2
                a[u] = __builtin_nontemporal_load(v + base + k);
            }
#pragma unroll
            for (int u = 0; u < kSpmvUnroll; ++u) xv[u] = x[EPI && epi.xrev ? epi.xrev - 1 - c[u] : c[u]];
#pragma unroll
            for (int u = 0; u < kSpmvUnroll; ++u)
                if (k0 + u * kSpmvRows < cnt) prod[k0 + u * kSpmvRows] = a[u] * xv[u];
        }
        __syncthreads();
        if (r < r1) {
            const int e = re - base;
            for (int k = rb - base; k < e; ++k) sum += prod[k];
        }
    } else if (r < r1) {
        const int e = rp[r + 1];
        for (int k = rp[r]; k < e; ++k) sum += v[k] * x[EPI && epi.xrev ? epi.xrev - 1 - ci[k] : ci[k]];
    }
    if (!EPI) {
        if (r < r1) y[r] = sum;
        return;
    }
    double d0 = 0.0, d1 = 0.0, d2 = 0.0;
    if (r < r1) {
        if (epi.b) sum = (sum - epi.b[r]) * -1.0;
        y[r] = sum;
        if (epi.copy1) epi.copy1[r] = sum;
        if (epi.copy2) epi.copy2[r] = sum;
        if (epi.dotw) d0 = epi.dotw[r] * sum;
        if (epi.dot_self) d1 = sum * sum;
        if (epi.dotw2) d2 = epi.dotw2[r] * sum;
    }
    if (epi.dd) {
        if (epi.dotw) {
            const DD v = block_sum<kSpmvRows>(DD{d0, 0.0}, red);
            if (threadIdx.x == 0) put_dd(partials, pstride, 0, blk, v);
        }
        if (epi.dot_self) {
            const DD v = block_sum<kSpmvRows>(DD{d1, 0.0}, red);
            if (threadIdx.x == 0) put_dd(partials, pstride, 1, blk, v);
        }
        if (epi.dotw2) {
            const DD v = block_sum<kSpmvRows>(DD{d2, 0.0}, red);
            if (threadIdx.x == 0) put_dd(partials, pstride, 2, blk, v);
        }
    } else {
        if (epi.dotw) {
            d0 = block_sum<kSpmvRows>(d0, red);
            if (threadIdx.x == 0) partials[blk] = d0;
        }
        if (epi.dot_self) {
            d1 = block_sum<kSpmvRows>(d1, red);
            if (threadIdx.x == 0) partials[pstride + blk] = d1;
        }
        if (epi.dotw2) {
            d2 = block_sum<kSpmvRows>(d2, red);
            if (threadIdx.x == 0) partials[2 * pstride + blk] = d2;
        }
    }
}

int launch_spmv(fvg_ctx* ctx, const fvg_csr* a, const double* x, double* y, const SpmvEpilogue& epi) {
    const int64_t n = a->nrows;
    if (n <= 0) return FVG_OK;
    const int g = spmv_blocks(n);
    const bool has_epi = epi.b || epi.copy1 || epi.copy2 || epi.dotw || epi.dotw2 || epi.dot_self || epi.skip;
    if (has_epi && !epi.partials) FVG_TRY(ensure_partials(ctx, g));
    double* part = epi.partials ? epi.partials : ctx->partials;
    const int64_t pstride = epi.partials ? epi.pstride : ctx->partials_cap;
    const int tok = timing_begin(ctx);
    if (has_epi)
        hipLaunchKernelGGL(k_spmv<true>, dim3(g), dim3(kSpmvRows), 0, ctx->stream, (int)n, a->row_ptr, a->col, a->val,
                           x, y, epi, part, pstride);
    else
        hipLaunchKernelGGL(k_spmv<false>, dim3(g), dim3(kSpmvRows), 0, ctx->stream, (int)n, a->row_ptr, a->col,
                           a->val, x, y, epi, part, pstride);
    timing_end(ctx, tok, kSpanSpmv, 1);
    FVG_HIP(hipGetLastError());
    return FVG_OK;
}

static void free_csr(fvg_ctx* c, fvg_csr* a) {
    dev_free_n(c, a->val32, a->nnz);
    if (a->dcol32) dev_free_n(c, a->dcol32, a->nnz);
    dev_free_n(c, a->row_ptr, a->nrows + 1);
    dev_free_n(c, a->col, a->nnz);
    dev_free_n(c, a->val, a->nnz);
}

// A CSR whose (valid, sorted) pattern already lives on the device: copied on the
// device, zero values, and a host copy of the pattern for the preconditioner analysis.
int csr_from_device_pattern(fvg_ctx* c, int64_t nrows, int64_t ncols, int64_t nnz, const int32_t* d_row_ptr,
                            const int32_t* d_col, fvg_csr** out) {
    *out = nullptr;
    auto* a = new fvg_csr();
    a->nrows = nrows;
    a->ncols = ncols;
    a->nnz = nnz;
    a->h_row_ptr.resize((size_t)nrows + 1);
    a->h_col.resize((size_t)nnz);
    int s;
    if ((s = dev_alloc_n(c, nrows + 1, &a->row_ptr)) || (s = dev_alloc_n(c, nnz, &a->col)) ||
        (s = dev_alloc_n(c, nnz, &a->val)) || (s = launch_fill(c, nnz, a->val, 0.0))) {
        (void)hipStreamSynchronize(c->stream);
        free_csr(c, a);
        delete a;
        return s;
    }
    if (hipMemcpyAsync(a->row_ptr, d_row_ptr, sizeof(int32_t) * (size_t)(nrows + 1), hipMemcpyDeviceToDevice,
                       c->stream) != hipSuccess ||
        (nnz > 0 && hipMemcpyAsync(a->col, d_col, sizeof(int32_t) * (size_t)nnz, hipMemcpyDeviceToDevice,
                                   c->stream) != hipSuccess))
        s = fail(FVG_E_HIP, "CSR pattern copy failed");
    if (s == FVG_OK) s = download(c, a->h_row_ptr.data(), d_row_ptr, nrows + 1);
    if (s == FVG_OK) s = download(c, a->h_col.data(), d_col, nnz);
    if (s != FVG_OK) {
        (void)hipStreamSynchronize(c->stream);
        free_csr(c, a);
        delete a;
        return s;
    }
    *out = a;
    return FVG_OK;
}

}  // namespace fvg

using namespace fvg;

extern "C" {

int fvg_csr_create(fvg_ctx* c, int64_t nrows, int64_t ncols, int64_t nnz, const int32_t* row_ptr,
                   const int32_t* col_idx, const double* values, fvg_csr** out) {
    if (!c || !out || !row_ptr || (nnz > 0 && !col_idx)) return fail(FVG_E_INVALID, "null argument");
    if (nrows < 0 || ncols < 0 || nnz < 0 || nnz >= INT32_MAX || nrows >= INT32_MAX)
        return fail(FVG_E_INVALID, "bad CSR sizes");
    if (row_ptr[0] != 0 || row_ptr[nrows] != nnz) return fail(FVG_E_INVALID, "row_ptr must run from 0 to nnz");
    for (int64_t i = 0; i < nrows; ++i) {
        if (row_ptr[i + 1] < row_ptr[i]) return fail(FVG_E_INVALID, "row_ptr must be non-decreasing");
        for (int64_t k = row_ptr[i]; k < row_ptr[i + 1]; ++k) {
            if (col_idx[k] < 0 || col_idx[k] >= ncols) return fail(FVG_E_INVALID, "column index out of range");
            if (k > row_ptr[i] && col_idx[k] <= col_idx[k - 1])
                return fail(FVG_E_INVALID, "columns must be strictly increasing within a row");
        }
    }
    *out = nullptr;
    auto* a = new fvg_csr();
    a->nrows = nrows;
    a->ncols = ncols;
    a->nnz = nnz;
    a->h_row_ptr.assign(row_ptr, row_ptr + nrows + 1);
    a->h_col.assign(col_idx, col_idx + nnz);
    int s;
    if ((s = dev_alloc_n(c, nrows + 1, &a->row_ptr)) || (s = dev_alloc_n(c, nnz, &a->col)) ||
        (s = dev_alloc_n(c, nnz, &a->val)) || (s = upload(c, a->row_ptr, row_ptr, nrows + 1)) ||
        (s = upload(c, a->col, col_idx, nnz))) {
        free_csr(c, a);
        delete a;
        return s;
    }
    if (values) s = upload(c, a->val, values, nnz);
    else if (nnz > 0) s = launch_fill(c, nnz, a->val, 0.0);
    if (s == FVG_OK && hipStreamSynchronize(c->stream) != hipSuccess) s = fail(FVG_E_HIP, "CSR upload failed");
    if (s != FVG_OK) {
        free_csr(c, a);
        delete a;
        return s;
    }
    *out = a;
    return FVG_OK;
}

int fvg_csr_destroy(fvg_ctx* c, fvg_csr* a) {
    if (!a) return FVG_OK;
    if (c) {
        (void)hipStreamSynchronize(c->stream);
        free_csr(c, a);
    }
    delete a;
    return FVG_OK;
}

int fvg_csr_sizes(const fvg_csr* a, int64_t* nrows, int64_t* ncols, int64_t* nnz) {
    if (!a) return fail(FVG_E_INVALID, "null matrix");
    if (nrows) *nrows = a->nrows;
    if (ncols) *ncols = a->ncols;
    if (nnz) *nnz = a->nnz;
    return FVG_OK;
}

int fvg_csr_download(fvg_ctx* c, const fvg_csr* a, int32_t* row_ptr, int32_t* col_idx, double* values) {
    if (!c || !a) return fail(FVG_E_INVALID, "null argument");
    if (row_ptr) std::memcpy(row_ptr, a->h_row_ptr.data(), sizeof(int32_t) * (size_t)(a->nrows + 1));
    if (col_idx) std::memcpy(col_idx, a->h_col.data(), sizeof(int32_t) * (size_t)a->nnz);
    if (values) FVG_TRY(download(c, values, a->val, a->nnz));
    return FVG_OK;
}

int fvg_csr_upload_values(fvg_ctx* c, fvg_csr* a, const double* values) {
    if (!c || !a || (!values && a->nnz)) return fail(FVG_E_INVALID, "null argument");
    FVG_TRY(upload(c, a->val, values, a->nnz));
    FVG_HIP(hipStreamSynchronize(c->stream));
    a->version += 1;
    return FVG_OK;
}

int fvg_csr_set_row_coords(fvg_csr* a, int32_t dim, const double* coords) {
    if (!a || dim < 0 || dim > 3 || (dim > 0 && !coords && a->nrows > 0)) return fail(FVG_E_INVALID, "bad argument");
    const size_t count = (size_t)a->nrows * (size_t)dim;
    a->h_coords.assign(count, 0.f);
    for (size_t k = 0; k < count; ++k) a->h_coords[k] = (float)coords[k];
    a->coord_dim = dim;
    a->coord_mesh = nullptr;
    a->coord_mesh_id = 0;
    if (dim == 0) a->h_coords.shrink_to_fit();
    return FVG_OK;
}

int fvg_spmv(fvg_ctx* c, const fvg_csr* a, const fvg_vec* x, fvg_vec* y) {
    if (!c || !a || !x || !y) return fail(FVG_E_INVALID, "null argument");
    if (x->n < a->ncols || y->n != a->nrows) return fail(FVG_E_INVALID, "vector lengths do not match the matrix");
    if (x->d == y->d) return fail(FVG_E_INVALID, "x and y must not alias");
    return launch_spmv(c, a, x->d, y->d, SpmvEpilogue{});
}

int fvg_spmv_time(fvg_ctx* c, const fvg_csr* a, const fvg_vec* x, fvg_vec* y, int32_t reps, double* avg_ms) {
    if (!c || !a || !x || !y || !avg_ms || reps <= 0) return fail(FVG_E_INVALID, "bad argument");
    if (x->n < a->ncols || y->n < a->nrows) return fail(FVG_E_INVALID, "vector too short for the matrix");
    hipEvent_t e0, e1;
    FVG_HIP(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return fail(FVG_E_HIP, "event creation failed");
    }
    SpmvEpilogue none;
    int s = FVG_OK;
    if (hipEventRecord(e0, c->stream) != hipSuccess) s = fail(FVG_E_HIP, "event record failed");
    for (int k = 0; k < reps && s == FVG_OK; ++k) s = launch_spmv(c, a, x->d, y->d, none);
    if (s == FVG_OK && hipEventRecord(e1, c->stream) != hipSuccess) s = fail(FVG_E_HIP, "event record failed");
    float ms = 0.f;
    if (s == FVG_OK && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
        s = fail(FVG_E_HIP, "event timing failed");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (s == FVG_OK) *avg_ms = (double)ms / reps;
    return s;
}

}  // extern "C"
