// mixed.hip -- FP32 inner Krylov iterations inside FP64 iterative refinement.
//
//   x_0 given;  loop:  r = b - A x        (FP64 SpMV, FP64 vectors)
//                      stop if |r|/|b| meets the caller's tolerances (same tests as the
//                      reference solvers, applied to the true residual)
//                      solve A d = r approximately in FP32: preconditioned CG or BiCGStab
//                      with FP32 matrix values, FP32 IC(0)/ILU(0) factors (factorised in
//                      FP64, pivots included, then rounded), FP32 vectors and dots
//                      x += d              (FP64 update)
//
// An inner solve runs until its residual has dropped by the factor still needed to meet
// the caller's tolerance (with a margin of 0.3), by at most 0.1 and at least
// `inner_tolerance` (default 1e-3 for CG, 1e-4 for BiCGStab), or until the total inner
// iteration budget (max_iterations) is spent.  Everything scalar lives on the device, as
// in the FP64 solvers.
#include <cmath>
#include <cstring>

#include "precond.hpp"
#include "reduce.hpp"
#include "spmv.hpp"

namespace fvg {

constexpr int kScalarBlock = 1024;   // threads of the one-workgroup scalar kernels

constexpr int kRows32 = 256;
constexpr int kStage32 = 3072;
constexpr int kUnroll32 = 6;   // products per thread whose loads are in flight together
constexpr int kGraphIters = 4; // inner iterations between host checks (one captured graph)

struct MState {
    double bnorm, init_res, true_res, atol, rtol, inner_tol;
    double inner_red;    // reduction asked of the current inner solve
    double rin0, rin, rz, alpha, beta, rho, omega;
    int64_t inner_iter, total_inner, max_inner, outer, restarts;
    int32_t done;        // outer loop finished
    int32_t inner_done;  // current inner solve finished
    int32_t mode;        // BiCGStab: 0 direction update, 1 restart
    int32_t pad;
};

// ---- FP32 SpMV with optional fused dots (partials written as doubles) ---------------
// y = A32 x; slot 0: sum dotw[r]*y[r]; slot 1: sum y[r]^2.  DELTA: the columns are read as
// 16-bit offsets from the tile's first row (dci) instead of ci.
template <bool DELTA>
__global__ void __launch_bounds__(kRows32) k_spmv32(int n, const int32_t* __restrict__ rp,
                                                    const int32_t* __restrict__ ci,
                                                    const int16_t* __restrict__ dci, const float* __restrict__ v,
                                                    const float* __restrict__ x, float* __restrict__ y,
                                                    const float* __restrict__ dotw, int dot_self,
                                                    double* __restrict__ partials, int64_t pstride,
                                                    const int32_t* __restrict__ skip) {
    __shared__ float prod[kStage32];
    __shared__ float red[kRows32 / 64];
    if (skip && *skip) return;
    const int blk = xcd_tile(blockIdx.x, gridDim.x);
    const int r0 = blk * kRows32;
    const int r1 = min(r0 + kRows32, n);
    const int base = rp[r0];
    const int cnt = rp[r1] - base;
    const int r = r0 + (int)threadIdx.x;
    const int rc = min(r, r1 - 1);
    const int rb = rp[rc], re = rp[rc + 1];   // this thread's row bounds, loaded before the products
    float sum = 0.f;
    if (cnt <= kStage32) {
        // all of a thread's entry loads, then all its gathers, before any use (clamped indices)
        for (int k0 = threadIdx.x; k0 < cnt; k0 += kRows32 * kUnroll32) {
            int c[kUnroll32];
            float a[kUnroll32], xv[kUnroll32];
#pragma unroll
            for (int u = 0; u < kUnroll32; ++u) {
                const int k = min(k0 + u * kRows32, cnt - 1);
                c[u] = DELTA ? r0 + (int)__builtin_nontemporal_load(dci + base + k)
                             : __builtin_nontemporal_load(ci + base + k);   // streamed once: keep x in L2
                a[u] = __builtin_nontemporal_load(v + base + k);
            }
#pragma unroll
            for (int u = 0; u < kUnroll32; ++u) xv[u] = x[c[u]];
#pragma unroll
            for (int u = 0; u < kUnroll32; ++u)
                if (k0 + u * kRows32 < cnt) prod[k0 + u * kRows32] = a[u] * xv[u];
        }
        __syncthreads();
        if (r < r1) {
            const int e = re - base;
            for (int k = rb - base; k < e; ++k) sum += prod[k];
        }
    } else if (r < r1) {
        const int e = rp[r + 1];
        for (int k = rp[r]; k < e; ++k) sum += v[k] * x[DELTA ? r0 + (int)dci[k] : ci[k]];
    }
    float d0 = 0.f, d1 = 0.f;
    if (r < r1) {
        y[r] = sum;
        if (dotw) d0 = dotw[r] * sum;
        if (dot_self) d1 = sum * sum;
    }
    if (dotw) {
        d0 = block_sum<kRows32>(d0, red);
        if (threadIdx.x == 0) partials[blk] = (double)d0;
    }
    if (dot_self) {
        d1 = block_sum<kRows32>(d1, red);
        if (threadIdx.x == 0) partials[pstride + blk] = (double)d1;
    }
}

__global__ void k_vals_f32(int64_t n, const double* __restrict__ s, float* __restrict__ d) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = (float)s[i];
}

// ---- scalar steps ----------------------------------------------------------------------
enum MStep {
    M_OUTER = 0,   // s1 = r.r (FP64 true residual): refinement test
    M_IN_INIT,     // s0 = r32.r32 (CG) -> rin0 ; rz from slot 2 ; BiCGStab rho
    M_CG_ALPHA,    // s0 = p.Ap
    M_CG_ERR,      // s0 = r.r -> inner test
    M_CG_BETA,     // s0 = r.z
    M_BI_ALPHA,    // s0 = rh.v
    M_BI_S,        // s0 = s.s -> inner test
    M_BI_OMEGA,    // s0 = t.s, s1 = t.t
    M_BI_R,        // s0 = rh.r, s1 = r.r -> inner test, rho, beta, restart
};

template <int STEP>
__global__ void __launch_bounds__(kScalarBlock) k_mscalar(MState* __restrict__ st, const double* __restrict__ partials,
                                                    int64_t stride, int count) {
    __shared__ double lds[kScalarBlock / 64];
    __shared__ int skip;
    if (threadIdx.x == 0) skip = STEP == M_OUTER ? st->done : (st->done | st->inner_done);
    __syncthreads();
    if (skip) return;
    constexpr bool need0 = STEP != M_OUTER;
    constexpr bool need1 = STEP == M_OUTER || STEP == M_BI_OMEGA || STEP == M_BI_R;
    constexpr bool need2 = STEP == M_IN_INIT;
    const double s0 = need0 ? sum_partials<kScalarBlock>(partials, count, lds) : 0.0;
    const double s1 = need1 ? sum_partials<kScalarBlock>(partials + stride, count, lds) : 0.0;
    const double s2 = need2 ? sum_partials<kScalarBlock>(partials + 2 * stride, count, lds) : 0.0;
    if (threadIdx.x != 0) return;
    switch (STEP) {
    case M_OUTER: {
        const double e = sqrt(s1) / st->bnorm;
        if (st->outer == 0) st->init_res = e;
        st->true_res = e;
        st->outer += 1;
        if ((e < st->atol) || (e / st->init_res < st->rtol) || st->total_inner >= st->max_inner) st->done = 1;
        st->inner_done = st->done;
        st->inner_iter = 0;
        // the remaining factor to the caller's tolerance with a margin of 0.3, at most 0.1
        // and at least inner_tol (an inner solve asked for more only overshoots the goal)
        const double goal = fmax(st->rtol * st->init_res, st->atol);
        st->inner_red = fmax(fmin(0.3 * goal / e, 0.1), st->inner_tol);
        break;
    }
    case M_IN_INIT:
        st->rin0 = sqrt(s0);
        st->rin = st->rin0;
        st->rz = s2;
        st->rho = s0;
        if (!(st->rin0 > 0.0)) st->inner_done = 1;
        break;
    case M_CG_ALPHA:   // an FP32 breakdown ends the inner solve; refinement goes on from d
        st->alpha = st->rz / s0;
        if (!isfinite(st->alpha)) st->inner_done = 1;
        break;
    case M_CG_ERR:
        st->rin = sqrt(s0);
        st->inner_iter += 1;
        st->total_inner += 1;
        if (st->rin <= st->inner_red * st->rin0 || st->total_inner >= st->max_inner) st->inner_done = 1;
        break;
    case M_CG_BETA:
        st->beta = s0 / st->rz;
        st->rz = s0;
        break;
    case M_BI_ALPHA:
        st->alpha = st->rho / s0;
        if (!isfinite(st->alpha)) st->inner_done = 1;
        break;
    case M_BI_S: {
        const double ss = sqrt(s0);
        if (ss <= st->inner_red * st->rin0) {
            st->rin = ss;
            st->inner_iter += 1;
            st->total_inner += 1;
            st->inner_done = 1;
        }
        break;
    }
    case M_BI_OMEGA:   // breakdown after h = d + alpha y: keep h
        st->omega = s0 / s1;
        if (!isfinite(st->omega) || st->omega == 0.0) {
            st->inner_iter += 1;
            st->total_inner += 1;
            st->inner_done = 1;
        }
        break;
    case M_BI_R: {
        st->rin = sqrt(s1);
        st->inner_iter += 1;
        st->total_inner += 1;
        if (st->rin <= st->inner_red * st->rin0 || st->total_inner >= st->max_inner) {
            st->inner_done = 1;
            break;
        }
        const double rho_new = s0;
        st->beta = (rho_new / st->rho) * (st->alpha / st->omega);
        st->rho = rho_new;
        if (fabs(st->rho) > 1e-30 * st->rin0 * st->rin0) {
            st->mode = 0;
        } else {
            st->mode = 1;
            st->restarts += 1;
            st->rho = s1;
        }
        break;
    }
    }
}

// ---- FP32 vector kernels ---------------------------------------------------------------
#define M_GUARD(st) \
    if ((st)->done | (st)->inner_done) return

// r32 = (float) r64; d = 0; slot 0: r32.r32.  For BiCGStab also rh = p = r32.
__global__ void __launch_bounds__(kBlock) k_m_start(int64_t n, const MState* __restrict__ st,
                                                    const double* __restrict__ r64, float* __restrict__ r,
                                                    float* __restrict__ d, float* __restrict__ rh,
                                                    float* __restrict__ p, double* __restrict__ partials) {
    __shared__ float lds[kBlock / 64];
    if (st->done) return;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    float q = 0.f;
    if (i < n) {
        const float ri = (float)r64[i];
        r[i] = ri;
        d[i] = 0.f;
        if (rh) rh[i] = ri;
        if (p) p[i] = ri;
        q = ri * ri;
    }
    q = block_sum<kBlock>(q, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = (double)q;
}

// slot `slot`: a.b (FP32)
__global__ void __launch_bounds__(kBlock) k_m_dot(int64_t n, const MState* __restrict__ st, int guard_inner,
                                                  const float* __restrict__ a, const float* __restrict__ b,
                                                  double* __restrict__ partials) {
    __shared__ float lds[kBlock / 64];
    if (st->done || (guard_inner && st->inner_done)) return;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    float q = i < n ? a[i] * b[i] : 0.f;
    q = block_sum<kBlock>(q, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = (double)q;
}

__global__ void k_m_copy(int64_t n, const MState* __restrict__ st, const float* __restrict__ s,
                         float* __restrict__ d) {
    M_GUARD(st);
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) d[i] = s[i];
}

// CG: d += alpha p; r -= alpha ap; slot 0: r.r
__global__ void __launch_bounds__(kBlock) k_m_cg_dr(int64_t n, const MState* __restrict__ st,
                                                    const float* __restrict__ p, const float* __restrict__ ap,
                                                    float* __restrict__ d, float* __restrict__ r,
                                                    double* __restrict__ partials) {
    __shared__ float lds[kBlock / 64];
    M_GUARD(st);
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const float alpha = (float)st->alpha;
    float q = 0.f;
    if (i < n) {
        d[i] += alpha * p[i];
        const float ri = r[i] - alpha * ap[i];
        r[i] = ri;
        q = ri * ri;
    }
    q = block_sum<kBlock>(q, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = (double)q;
}

__global__ void k_m_cg_p(int64_t n, const MState* __restrict__ st, const float* __restrict__ z,
                         float* __restrict__ p) {
    M_GUARD(st);
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) p[i] = z[i] + (float)st->beta * p[i];
}

// BiCGStab: d += alpha y (d holds h); s = r - alpha v; slot 0: s.s
__global__ void __launch_bounds__(kBlock) k_m_bi_hs(int64_t n, const MState* __restrict__ st,
                                                    const float* __restrict__ y, const float* __restrict__ v,
                                                    const float* __restrict__ r, float* __restrict__ d,
                                                    float* __restrict__ s, double* __restrict__ partials) {
    __shared__ float lds[kBlock / 64];
    M_GUARD(st);
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const float alpha = (float)st->alpha;
    float q = 0.f;
    if (i < n) {
        d[i] += alpha * y[i];
        const float si = r[i] - alpha * v[i];
        s[i] = si;
        q = si * si;
    }
    q = block_sum<kBlock>(q, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = (double)q;
}

// BiCGStab: d += omega z; r = s - omega t; slot 0: rh.r, slot 1: r.r
__global__ void __launch_bounds__(kBlock) k_m_bi_xr(int64_t n, const MState* __restrict__ st,
                                                    const float* __restrict__ z, const float* __restrict__ t,
                                                    const float* __restrict__ s, const float* __restrict__ rh,
                                                    float* __restrict__ d, float* __restrict__ r,
                                                    double* __restrict__ partials, int64_t stride) {
    __shared__ float lds[kBlock / 64];
    M_GUARD(st);
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const float omega = (float)st->omega;
    float q0 = 0.f, q1 = 0.f;
    if (i < n) {
        d[i] += omega * z[i];
        const float ri = s[i] - omega * t[i];
        r[i] = ri;
        q0 = rh[i] * ri;
        q1 = ri * ri;
    }
    q0 = block_sum<kBlock>(q0, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = (double)q0;
    q1 = block_sum<kBlock>(q1, lds);
    if (threadIdx.x == 0) partials[stride + blockIdx.x] = (double)q1;
}

__global__ void k_m_bi_p(int64_t n, const MState* __restrict__ st, const float* __restrict__ r,
                         const float* __restrict__ v, float* __restrict__ p, float* __restrict__ rh) {
    M_GUARD(st);
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    if (st->mode == 0) {
        p[i] = r[i] + (float)st->beta * (p[i] - (float)st->omega * v[i]);
    } else {
        rh[i] = r[i];
        p[i] = r[i];
    }
}

// x += (double) d  (FP64 solution update)
__global__ void k_m_update(int64_t n, const MState* __restrict__ st, const float* __restrict__ d,
                           double* __restrict__ x) {
    if (st->done) return;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) x[i] += (double)d[i];
}

// ---- host driver --------------------------------------------------------------------------

static int ensure_val32(fvg_ctx* c, fvg_csr* a) {
    if (a->val32 && a->val32_version == a->version) return FVG_OK;
    if (!a->val32) FVG_TRY(dev_alloc_n(c, a->nnz, &a->val32));
    if (a->nnz > 0)
        hipLaunchKernelGGL(k_vals_f32, dim3(grid_for(a->nnz)), dim3(kBlock), 0, c->stream, a->nnz, a->val, a->val32);
    FVG_HIP(hipGetLastError());
    a->val32_version = a->version;
    if (a->dcol32_state == 0 && a->nnz > 0 && (int64_t)a->h_row_ptr.size() == a->nrows + 1) {
        // 16-bit column offsets from each tile's first row, where they all fit: the FP32
        // SpMV then streams 2 bytes per entry for the columns instead of 4
        std::vector<int16_t> d((size_t)a->nnz);
        bool fits = true;
        for (int64_t i = 0; i < a->nrows && fits; ++i) {
            const int64_t r0 = i / kRows32 * kRows32;
            for (int32_t k = a->h_row_ptr[(size_t)i]; k < a->h_row_ptr[(size_t)i + 1]; ++k) {
                const int64_t off = (int64_t)a->h_col[(size_t)k] - r0;
                if (off < INT16_MIN || off > INT16_MAX) {
                    fits = false;
                    break;
                }
                d[(size_t)k] = (int16_t)off;
            }
        }
        a->dcol32_state = fits ? 1 : -1;
        if (fits) {
            FVG_TRY(dev_alloc_n(c, a->nnz, &a->dcol32));
            FVG_TRY(upload(c, a->dcol32, d.data(), a->nnz));
            FVG_HIP(hipStreamSynchronize(c->stream));   // d goes out of scope
        }
    }
    return FVG_OK;
}

struct FBuf {
    fvg_ctx* c;
    std::vector<std::pair<float*, int64_t>> f;
    std::vector<std::pair<double*, int64_t>> d;
    ~FBuf() {
        (void)hipStreamSynchronize(c->stream);
        for (auto& b : f) dev_free_n(c, b.first, b.second);
        for (auto& b : d) dev_free_n(c, b.first, b.second);
    }
    int get(int64_t n, float** out) {
        FVG_TRY(dev_alloc_n(c, n, out));
        f.emplace_back(*out, n);
        return FVG_OK;
    }
    int get(int64_t n, double** out) {
        FVG_TRY(dev_alloc_n(c, n, out));
        d.emplace_back(*out, n);
        return FVG_OK;
    }
};

static int launch_spmv32(fvg_ctx* c, const fvg_csr* a, const float* x, float* y, const float* dotw, int dot_self,
                         const int32_t* skip) {
    const int64_t n = a->nrows;
    const int g = grid_for(n, kRows32);
    const int tok = timing_begin(c);
    if (a->dcol32)
        hipLaunchKernelGGL(k_spmv32<true>, dim3(g), dim3(kRows32), 0, c->stream, (int)n, a->row_ptr, a->col,
                           a->dcol32, a->val32, x, y, dotw, dot_self, c->partials, c->partials_cap, skip);
    else
        hipLaunchKernelGGL(k_spmv32<false>, dim3(g), dim3(kRows32), 0, c->stream, (int)n, a->row_ptr, a->col,
                           a->dcol32, a->val32, x, y, dotw, dot_self, c->partials, c->partials_cap, skip);
    timing_end(c, tok, kSpanSpmv, 1);
    FVG_HIP(hipGetLastError());
    return FVG_OK;
}

int run_mixed(fvg_ctx* c, fvg_csr* a, fvg_prec* p, const fvg_vec* b, fvg_vec* x, const fvg_solver_options* o,
              fvg_solver_info* info, double* history, int64_t cap, bool bicgstab) {
    const int64_t n = a->nrows;
    FVG_TRY(launch_dot(c, n, b->d, b->d, c->scalar));
    double bb;
    FVG_TRY(read_scalars(c, c->scalar, &bb, 1));
    const double bnorm = std::sqrt(bb);
    if (bnorm < 1e-14) {
        FVG_TRY(launch_fill(c, n, x->d, 0.0));
        FVG_HIP(hipStreamSynchronize(c->stream));
        if (info) {
            std::memset(info, 0, sizeof(*info));
            info->initial_residual = info->final_residual = info->true_residual = bnorm;
            info->history_len = 1;
        }
        if (history && cap > 0) history[0] = bnorm;
        return FVG_OK;
    }
    FVG_TRY(ensure_val32(c, a));
    FVG_TRY(ensure_prec32(c, p));
    const int nb = spmv_blocks(n);
    FVG_TRY(ensure_partials(c, nb));
    const int64_t stride = c->partials_cap;
    double* part = c->partials;
    FBuf B{c};
    double* r64;
    float *r, *d, *z, *pp, *ap, *rh = nullptr, *y = nullptr, *v = nullptr, *s = nullptr, *t = nullptr;
    FVG_TRY(B.get(n, &r64));
    FVG_TRY(B.get(n, &r));
    FVG_TRY(B.get(n, &d));
    FVG_TRY(B.get(n, &z));
    FVG_TRY(B.get(n, &pp));
    FVG_TRY(B.get(n, &ap));
    if (bicgstab) {
        FVG_TRY(B.get(n, &rh));
        FVG_TRY(B.get(n, &y));
        FVG_TRY(B.get(n, &v));
        FVG_TRY(B.get(n, &s));
        FVG_TRY(B.get(n, &t));
    }
    MState* st;
    FVG_TRY(dev_alloc_n(c, 1, &st));
    struct StFree {
        fvg_ctx* c;
        MState* st;
        ~StFree() {
            (void)hipStreamSynchronize(c->stream);
            dev_free_n(c, st, 1);
        }
    } stfree{c, st};
    MState h{};
    h.bnorm = bnorm;
    h.atol = o->absolute_tolerance;
    h.rtol = o->relative_tolerance;
    h.inner_tol = o->inner_tolerance > 0.0 ? o->inner_tolerance : (bicgstab ? 1e-4 : 1e-3);
    h.max_inner = o->max_iterations > 0 ? o->max_iterations : 1;
    FVG_HIP(hipMemcpyAsync(st, &h, sizeof(MState), hipMemcpyHostToDevice, c->stream));
    const int32_t* inner_skip = &st->inner_done;
    std::vector<double> hist;

    // one inner iteration (kernels only): launched directly, or captured into a graph
    auto iteration = [&]() -> int {
        if (!bicgstab) {
            FVG_TRY(launch_spmv32(c, a, pp, ap, pp, 0, inner_skip));
            hipLaunchKernelGGL(k_mscalar<M_CG_ALPHA>, dim3(1), dim3(kScalarBlock), 0, c->stream, st, part, stride, nb);
            hipLaunchKernelGGL(k_m_cg_dr, dim3(nb), dim3(kBlock), 0, c->stream, n, st, pp, ap, d, r, part);
            hipLaunchKernelGGL(k_mscalar<M_CG_ERR>, dim3(1), dim3(kScalarBlock), 0, c->stream, st, part, stride, nb);
            // r.z: fused into the preconditioner's last sweep where it can be
            bool dot_done = false;
            FVG_TRY(apply_prec32(c, p, r, z, inner_skip, DotFuse{r, part, nb, &dot_done}));
            if (!dot_done) hipLaunchKernelGGL(k_m_dot, dim3(nb), dim3(kBlock), 0, c->stream, n, st, 1, r, z, part);
            hipLaunchKernelGGL(k_mscalar<M_CG_BETA>, dim3(1), dim3(kScalarBlock), 0, c->stream, st, part, stride, nb);
            hipLaunchKernelGGL(k_m_cg_p, dim3(nb), dim3(kBlock), 0, c->stream, n, st, z, pp);
        } else {
            FVG_TRY(apply_prec32(c, p, pp, y, inner_skip));
            FVG_TRY(launch_spmv32(c, a, y, v, rh, 0, inner_skip));
            hipLaunchKernelGGL(k_mscalar<M_BI_ALPHA>, dim3(1), dim3(kScalarBlock), 0, c->stream, st, part, stride, nb);
            hipLaunchKernelGGL(k_m_bi_hs, dim3(nb), dim3(kBlock), 0, c->stream, n, st, y, v, r, d, s, part);
            hipLaunchKernelGGL(k_mscalar<M_BI_S>, dim3(1), dim3(kScalarBlock), 0, c->stream, st, part, stride, nb);
            FVG_TRY(apply_prec32(c, p, s, z, inner_skip));
            FVG_TRY(launch_spmv32(c, a, z, t, s, 1, inner_skip));
            hipLaunchKernelGGL(k_mscalar<M_BI_OMEGA>, dim3(1), dim3(kScalarBlock), 0, c->stream, st, part, stride, nb);
            hipLaunchKernelGGL(k_m_bi_xr, dim3(nb), dim3(kBlock), 0, c->stream, n, st, z, t, s, rh, d, r, part,
                               stride);
            hipLaunchKernelGGL(k_mscalar<M_BI_R>, dim3(1), dim3(kScalarBlock), 0, c->stream, st, part, stride, nb);
            hipLaunchKernelGGL(k_m_bi_p, dim3(nb), dim3(kBlock), 0, c->stream, n, st, r, v, pp, rh);
        }
        FVG_HIP(hipGetLastError());
        return FVG_OK;
    };
    // inner iterations from a captured graph unless timing spans (event records), the
    // exact FP32 sweeps (cooperative launches) or direct launches are asked for
    struct Graph {
        bool usable = false;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        ~Graph() {
            if (exec) (void)hipGraphExecDestroy(exec);
            if (graph) (void)hipGraphDestroy(graph);
        }
    } graph;
    graph.usable = !c->timing && !c->tune_inner_exact && !c->tune_graphs;
    for (int outer = 0; outer < 64; ++outer) {
        // r = b - A x (FP64) and the refinement test
        SpmvEpilogue e;
        e.b = b->d;
        e.dot_self = true;
        e.skip = &st->done;
        FVG_TRY(launch_spmv(c, a, x->d, r64, e));
        hipLaunchKernelGGL(k_mscalar<M_OUTER>, dim3(1), dim3(kScalarBlock), 0, c->stream, st, part, stride, nb);
        FVG_HIP(hipMemcpyAsync(c->pinned, st, sizeof(MState), hipMemcpyDeviceToHost, c->stream));
        FVG_HIP(hipStreamSynchronize(c->stream));
        std::memcpy(&h, c->pinned, sizeof(MState));
        hist.push_back(h.true_res);
        if (h.done) break;
        // FP32 iterations barely progressing on this system: leave it to the FP64 iterations
        const size_t m = hist.size();
        if (m >= 4 && hist[m - 1] > 0.5 * hist[m - 2] && hist[m - 2] > 0.5 * hist[m - 3] &&
            hist[m - 3] > 0.5 * hist[m - 4])
            break;
        // inner start: r32 = r64, d = 0, |r32|; CG: z = M r, p = z, r.z; BiCGStab: rh = p = r
        hipLaunchKernelGGL(k_m_start, dim3(nb), dim3(kBlock), 0, c->stream, n, st, r64, r, d, rh, bicgstab ? pp : nullptr,
                           part);
        if (!bicgstab) {
            FVG_TRY(apply_prec32(c, p, r, z, &st->done));
            hipLaunchKernelGGL(k_m_copy, dim3(nb), dim3(kBlock), 0, c->stream, n, st, z, pp);
            hipLaunchKernelGGL(k_m_dot, dim3(nb), dim3(kBlock), 0, c->stream, n, st, 0, r, z, part + 2 * stride);
        }
        hipLaunchKernelGGL(k_mscalar<M_IN_INIT>, dim3(1), dim3(kScalarBlock), 0, c->stream, st, part, stride, nb);
        FVG_HIP(hipGetLastError());
        if (graph.usable) {
            // the inner iterations replayed from one captured graph of kGraphIters iterations
            // (kernels only: the same launches every time, each a no-op once the solve is done)
            if (!graph.exec) {
                FVG_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
                int cs = FVG_OK;
                for (int k = 0; k < kGraphIters && cs == FVG_OK; ++k) cs = iteration();
                hipGraph_t g = nullptr;
                const hipError_t ce = hipStreamEndCapture(c->stream, &g);
                if (cs != FVG_OK) {
                    if (g) (void)hipGraphDestroy(g);
                    return cs;
                }
                FVG_HIP(ce);
                graph.graph = g;
                FVG_HIP(hipGraphInstantiate(&graph.exec, g, nullptr, nullptr, 0));
            }
            for (;;) {
                FVG_HIP(hipGraphLaunch(graph.exec, c->stream));
                FVG_HIP(hipMemcpyAsync(c->pinned, st, sizeof(MState), hipMemcpyDeviceToHost, c->stream));
                FVG_HIP(hipStreamSynchronize(c->stream));
                std::memcpy(&h, c->pinned, sizeof(MState));
                if (h.inner_done || h.done) break;
            }
        } else {
            for (int64_t it = 0;; ++it) {
                FVG_TRY(iteration());
                if ((it + 1) % kGraphIters == 0) {
                    FVG_HIP(hipMemcpyAsync(c->pinned, st, sizeof(MState), hipMemcpyDeviceToHost, c->stream));
                    FVG_HIP(hipStreamSynchronize(c->stream));
                    std::memcpy(&h, c->pinned, sizeof(MState));
                    if (h.inner_done || h.done) break;
                }
            }
        }
        hipLaunchKernelGGL(k_m_update, dim3(nb), dim3(kBlock), 0, c->stream, n, st, d, x->d);
        FVG_HIP(hipGetLastError());
    }
    FVG_TRY(check_sweep_status(c, p));
    FVG_HIP(hipMemcpyAsync(c->pinned, st, sizeof(MState), hipMemcpyDeviceToHost, c->stream));
    FVG_HIP(hipStreamSynchronize(c->stream));
    std::memcpy(&h, c->pinned, sizeof(MState));
    if (info) {
        std::memset(info, 0, sizeof(*info));
        info->iterations = h.total_inner;
        info->initial_residual = h.init_res;
        info->final_residual = h.true_res;
        info->true_residual = h.true_res;
        info->outer_iterations = h.outer;
        info->restarts = h.restarts;
        info->history_len = (int64_t)hist.size();
    }
    if (history)
        for (size_t k = 0; k < hist.size() && (int64_t)k < cap; ++k) history[k] = hist[k];
    return FVG_OK;
}

}  // namespace fvg
