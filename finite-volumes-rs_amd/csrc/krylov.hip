// krylov.hip -- conjugate_gradient and bi_conjugate_gradient_stab on the device.
//
// The iteration is the reference's, step for step (conjugate_gradient.rs:12-112,
// bi_conjugate_gradient_stab.rs:13-179): same updates in the same operation order,
// same stopping tests on the recursive residual, same iteration counting and
// history.  All scalars live on the device: each reduction ends in a one-workgroup
// "scalar" kernel that sums the per-workgroup partials in a fixed order and runs the
// scalar part of the algorithm, so the host only enqueues work and polls a flag.
// Vector updates are fused with the dot products that follow them.
#include <cmath>
#include <cstring>

#include "precond.hpp"
#include "reduce.hpp"
#include "spmv.hpp"

namespace fvg {

constexpr int kScalarBlock = 1024;   // threads of the one-workgroup scalar kernels

struct KState {
    double bnorm, init_res, error, atol, rtol;
    double rz, alpha, beta, rho, omega;
    double ts, tt;
    int64_t iter, max_iter, hist_len, hist_cap, restarts;
    int32_t done;   // converged or stopped inside an iteration: the rest of it is skipped
    int32_t halt;   // the loop test failed before an iteration began
    int32_t mode;   // BiCGStab direction update: 0 update, 1 restart
    int32_t pad;
};

enum ScalarStep {
    CG_INIT = 0,    // initial residual, loop test
    CG_RZ0,         // r.z of the first iteration
    CG_ALPHA,       // alpha = rz / p.Ap
    CG_ERROR,       // error = |r| / |b| and the convergence test
    CG_BETA,        // beta = rz_new / rz, history, loop test
    BI_INIT,        // initial residual, rho = rh.r
    BI_ALPHA,       // alpha = rho / rh.v
    BI_S,           // |s| test
    BI_OMEGA,       // omega = t.s / t.t
    BI_R,           // |r| test, rho, beta, restart decision, history, loop test
};

__device__ __forceinline__ void push_history(KState* st, double* hist, double e) {
    if (st->hist_len < st->hist_cap) hist[st->hist_len] = e;
    st->hist_len += 1;
}

__device__ __forceinline__ void loop_test(KState* st) {
    // while error > tolerance { if iter > max_iter { break } ... }
    if (!(st->error > st->atol) || st->iter > st->max_iter) st->halt = 1;
}

template <int STEP>
__global__ void __launch_bounds__(kScalarBlock) k_scalar(KState* __restrict__ st, const double* __restrict__ partials,
                                                   int64_t stride, int count, double* __restrict__ hist) {
    __shared__ double lds[kScalarBlock / 64];
    __shared__ int skip;
    if (threadIdx.x == 0) {
        skip = 0;
        if (STEP == CG_ALPHA || STEP == BI_ALPHA) {
            if (st->halt) st->done = 1;
            skip = st->done;
        } else if (STEP != CG_INIT && STEP != BI_INIT) {
            skip = st->done;
        }
    }
    __syncthreads();
    if (skip) return;
    const double s0 = sum_partials<kScalarBlock>(partials, count, lds);
    double s1 = 0.0, s2 = 0.0;
    if (STEP == BI_OMEGA || STEP == BI_R) {
        s1 = sum_partials<kScalarBlock>(partials + stride, count, lds);
    }
    if (STEP == BI_R) s2 = sum_partials<kScalarBlock>(partials + 2 * stride, count, lds);
    if (threadIdx.x != 0) return;
    switch (STEP) {
    case CG_INIT:   // s0 = r.r   (conjugate_gradient.rs:52, 62-66)
        st->init_res = sqrt(s0) / st->bnorm;
        st->error = 1.0;
        st->iter = 0;
        st->done = 0;
        st->halt = 0;
        loop_test(st);
        break;
    case CG_RZ0:    // s0 = r.z (conjugate_gradient.rs:71)
        st->rz = s0;
        break;
    case CG_ALPHA:  // s0 = p.Ap (:72-74)
        st->alpha = st->rz / s0;
        break;
    case CG_ERROR: {   // s0 = r.r (:83-89)
        st->error = sqrt(s0) / st->bnorm;
        if ((st->error < st->atol) || (st->error / st->init_res < st->rtol)) {
            push_history(st, hist, st->error);
            st->iter += 1;
            st->done = 1;
            st->halt = 1;
        }
        break;
    }
    case CG_BETA:   // s0 = r.z_new (:94-105)
        st->beta = s0 / st->rz;
        st->rz = s0;
        push_history(st, hist, st->error);
        st->iter += 1;
        loop_test(st);
        break;
    case BI_INIT:   // s0 = r.r; rho = rh.r with rh == r (bi_conjugate_gradient_stab.rs:174, 190)
        st->init_res = sqrt(s0) / st->bnorm;
        st->rho = s0;
        st->error = 1.0;
        st->iter = 0;
        st->done = 0;
        st->halt = 0;
        st->restarts = 0;
        loop_test(st);
        break;
    case BI_ALPHA:  // s0 = rh.v (:208-209)
        st->alpha = st->rho / s0;
        break;
    case BI_S: {    // s0 = s.s (:219-231)
        const double ss = sqrt(s0) / st->bnorm;
        if ((ss < st->atol) || (ss / st->init_res < st->rtol)) {
            st->error = ss;
            push_history(st, hist, ss);
            st->iter += 1;
            st->done = 1;
            st->halt = 1;
        }
        break;
    }
    case BI_OMEGA:  // s0 = t.s, s1 = t.t (:239-242)
        st->ts = s0;
        st->tt = s1;
        st->omega = s0 / s1;
        break;
    case BI_R: {    // s1 = r.r, s0 = rh.r (:252-287)
        const double rr = sqrt(s1) / st->bnorm;
        if ((rr < st->atol) || (rr / st->init_res < st->rtol)) {
            st->error = rr;
            push_history(st, hist, rr);
            st->iter += 1;
            st->done = 1;
            st->halt = 1;
            break;
        }
        const double rho_new = s0;
        st->beta = (rho_new / st->rho) * (st->alpha / st->omega);
        st->rho = rho_new;
        if (fabs(st->rho) > (1e-15 * st->bnorm)) {
            st->mode = 0;
        } else {   // restart: rh = r, p = r, rho = rh.r == r.r
            st->mode = 1;
            st->restarts += 1;
            st->rho = s1;
        }
        st->error = rr / st->bnorm;
        push_history(st, hist, st->error);
        st->iter += 1;
        loop_test(st);
        (void)s2;
        break;
    }
    }
}

// ---- fused vector kernels (one element per thread, kBlock per workgroup) --------------

// CG (conjugate_gradient.rs:77-80): x += p*alpha; r -= ap*alpha; partial r.r
__global__ void __launch_bounds__(kBlock) k_cg_xr(int64_t n, const KState* __restrict__ st, const double* __restrict__ p,
                                                  const double* __restrict__ ap, double* __restrict__ x,
                                                  double* __restrict__ r, double* __restrict__ partials) {
    __shared__ double lds[kBlock / 64];
    if (st->done) return;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const double alpha = st->alpha;
    double d = 0.0;
    if (i < n) {
        x[i] += p[i] * alpha;
        const double ri = r[i] - ap[i] * alpha;
        r[i] = ri;
        d = ri * ri;
    }
    d = block_sum<kBlock>(d, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = d;
}

// partial a.b, skipped when done
__global__ void __launch_bounds__(kBlock) k_dot_g(int64_t n, const KState* __restrict__ st, const double* __restrict__ a,
                                                  const double* __restrict__ b, double* __restrict__ partials) {
    __shared__ double lds[kBlock / 64];
    if (st && st->done) return;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    double d = i < n ? a[i] * b[i] : 0.0;
    d = block_sum<kBlock>(d, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = d;
}

// CG (conjugate_gradient.rs:98-100): p = z + p*beta
__global__ void __launch_bounds__(kBlock) k_cg_p(int64_t n, const KState* __restrict__ st, const double* __restrict__ z,
                                                 double* __restrict__ p) {
    if (st->done) return;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) p[i] = z[i] + p[i] * st->beta;
}

// BiCGStab (bi_conjugate_gradient_stab.rs:213-216): h = x + y*alpha (kept in x); s = r - v*alpha; partial s.s
__global__ void __launch_bounds__(kBlock) k_bi_hs(int64_t n, const KState* __restrict__ st, const double* __restrict__ y,
                                                  const double* __restrict__ v, const double* __restrict__ r,
                                                  double* __restrict__ x, double* __restrict__ s,
                                                  double* __restrict__ partials) {
    __shared__ double lds[kBlock / 64];
    if (st->done) return;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const double alpha = st->alpha;
    double d = 0.0;
    if (i < n) {
        x[i] = x[i] + y[i] * alpha;
        const double si = r[i] - v[i] * alpha;
        s[i] = si;
        d = si * si;
    }
    d = block_sum<kBlock>(d, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = d;
}

// BiCGStab (:246-249): x = h + z*omega; r = s - t*omega; partials rh.r (slot 0) and r.r (slot 1)
__global__ void __launch_bounds__(kBlock) k_bi_xr(int64_t n, const KState* __restrict__ st, const double* __restrict__ z,
                                                  const double* __restrict__ t, const double* __restrict__ s,
                                                  const double* __restrict__ rh, double* __restrict__ x,
                                                  double* __restrict__ r, double* __restrict__ partials,
                                                  int64_t stride) {
    __shared__ double lds[kBlock / 64];
    if (st->done) return;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const double omega = st->omega;
    double d0 = 0.0, d1 = 0.0;
    if (i < n) {
        x[i] = x[i] + z[i] * omega;
        const double ri = s[i] - t[i] * omega;
        r[i] = ri;
        d0 = rh[i] * ri;
        d1 = ri * ri;
    }
    d0 = block_sum<kBlock>(d0, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = d0;
    d1 = block_sum<kBlock>(d1, lds);
    if (threadIdx.x == 0) partials[stride + blockIdx.x] = d1;
}

// BiCGStab (:268-282): p = r + (p - v*omega)*beta, or restart rh = r, p = r
__global__ void __launch_bounds__(kBlock) k_bi_p(int64_t n, const KState* __restrict__ st, const double* __restrict__ r,
                                                 const double* __restrict__ v, double* __restrict__ p,
                                                 double* __restrict__ rh) {
    if (st->done) return;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    if (st->mode == 0) {
        p[i] = r[i] + (p[i] - v[i] * st->omega) * st->beta;
    } else {
        rh[i] = r[i];
        p[i] = r[i];
    }
}

// ---- host drivers ----------------------------------------------------------------------

struct Workspace {
    fvg_ctx* c;
    std::vector<std::pair<double*, int64_t>> bufs;
    ~Workspace() {
        (void)hipStreamSynchronize(c->stream);
        for (auto& b : bufs) dev_free_n(c, b.first, b.second);
    }
    int get(int64_t n, double** out) {
        FVG_TRY(dev_alloc_n(c, n, out));
        bufs.emplace_back(*out, n);
        return FVG_OK;
    }
};

static int check_solve_args(fvg_ctx* c, const fvg_csr* a, const fvg_prec* p, const fvg_vec* b, const fvg_vec* x,
                            const fvg_solver_options* o) {
    if (!c || !a || !p || !b || !x || !o) return fail(FVG_E_INVALID, "null argument");
    if (a->nrows != a->ncols) return fail(FVG_E_INVALID, "solver needs a square matrix");
    if (b->n != a->nrows || x->n != a->nrows || p->n != a->nrows)
        return fail(FVG_E_INVALID, "vector / preconditioner size does not match the matrix");
    if (o->max_iterations < 0) return fail(FVG_E_INVALID, "max_iterations must be non-negative");
    return FVG_OK;
}

struct KHost {
    fvg_ctx* c;
    KState* st = nullptr;
    double* hist = nullptr;
    int64_t hist_cap = 0;
    ~KHost() {
        if (st) {
            (void)hipStreamSynchronize(c->stream);
            dev_free_n(c, st, 1);
            dev_free_n(c, hist, hist_cap);
        }
    }
    int init(double bnorm, const fvg_solver_options* o) {
        hist_cap = o->max_iterations + 2;
        FVG_TRY(dev_alloc_n(c, 1, &st));
        FVG_TRY(dev_alloc_n(c, hist_cap, &hist));
        KState h{};
        h.bnorm = bnorm;
        h.atol = o->absolute_tolerance;
        h.rtol = o->relative_tolerance;
        h.max_iter = o->max_iterations;
        h.hist_cap = hist_cap;
        FVG_HIP(hipMemcpyAsync(st, &h, sizeof(KState), hipMemcpyHostToDevice, c->stream));
        FVG_HIP(hipStreamSynchronize(c->stream));
        return FVG_OK;
    }
    int read(KState* h) {
        FVG_HIP(hipMemcpyAsync(c->pinned, st, sizeof(KState), hipMemcpyDeviceToHost, c->stream));
        FVG_HIP(hipStreamSynchronize(c->stream));
        std::memcpy(h, c->pinned, sizeof(KState));
        return FVG_OK;
    }
    int finish(fvg_solver_info* info, double* history, int64_t cap) {
        KState h;
        FVG_TRY(read(&h));
        if (info) {
            std::memset(info, 0, sizeof(*info));
            info->iterations = h.iter;
            info->initial_residual = h.init_res;
            info->final_residual = h.error;
            info->history_len = h.hist_len;
            info->restarts = h.restarts;
        }
        if (history && cap > 0) {
            int64_t m = h.hist_len < cap ? h.hist_len : cap;
            m = m < hist_cap ? m : hist_cap;
            FVG_TRY(download(c, history, hist, m));
        }
        return FVG_OK;
    }
};

// zero right-hand side early exit shared by both solvers (conjugate_gradient.rs:31-36)
static int zero_rhs_exit(fvg_ctx* c, fvg_vec* x, double bnorm, fvg_solver_info* info, double* history, int64_t cap) {
    FVG_TRY(launch_fill(c, x->n, x->d, 0.0));
    FVG_HIP(hipStreamSynchronize(c->stream));
    if (info) {
        std::memset(info, 0, sizeof(*info));
        info->initial_residual = bnorm;
        info->final_residual = bnorm;
        info->history_len = 1;
    }
    if (history && cap > 0) history[0] = bnorm;
    return FVG_OK;
}

static int rhs_norm(fvg_ctx* c, const fvg_vec* b, double* bnorm) {
    FVG_TRY(launch_dot(c, b->n, b->d, b->d, c->scalar));
    double bb;
    FVG_TRY(read_scalars(c, c->scalar, &bb, 1));
    *bnorm = std::sqrt(bb);
    return FVG_OK;
}

// iterations enqueued between two polls of the device state
static int poll_batch(const fvg_prec* p) {
    (void)p;
    return 4;
}

int run_cg(fvg_ctx* c, const fvg_csr* a, const fvg_prec* p, const fvg_vec* b, fvg_vec* x,
           const fvg_solver_options* o, fvg_solver_info* info, double* history, int64_t cap) {
    const int64_t n = a->nrows;
    double bnorm;
    FVG_TRY(rhs_norm(c, b, &bnorm));
    if (bnorm < 1e-14) return zero_rhs_exit(c, x, bnorm, info, history, cap);
    Workspace ws{c};
    double *r, *z, *pp, *ap;
    FVG_TRY(ws.get(n, &r));
    FVG_TRY(ws.get(n, &z));
    FVG_TRY(ws.get(n, &pp));
    FVG_TRY(ws.get(n, &ap));
    KHost K{c};
    FVG_TRY(K.init(bnorm, o));
    const int nb = spmv_blocks(n);
    FVG_TRY(ensure_partials(c, nb));
    const int64_t stride = c->partials_cap;
    double* part = c->partials;
    const int32_t* done = &K.st->done;
    const int32_t* halt = &K.st->halt;

    // r = (A x - b) * -1 and r.r (conjugate_gradient.rs:47-52)
    SpmvEpilogue e0;
    e0.b = b->d;
    e0.dot_self = true;
    FVG_TRY(launch_spmv(c, a, x->d, r, e0));
    hipLaunchKernelGGL(k_scalar<CG_INIT>, dim3(1), dim3(kScalarBlock), 0, c->stream, K.st, part + stride, stride, nb, K.hist);
    // z = M^-1 r; p = z; r.z
    FVG_TRY(apply_prec(c, p, r, z, nullptr));
    FVG_TRY(launch_copy(c, n, z, pp));
    hipLaunchKernelGGL(k_dot_g, dim3(nb), dim3(kBlock), 0, c->stream, n, (const KState*)nullptr, r, z, part);
    hipLaunchKernelGGL(k_scalar<CG_RZ0>, dim3(1), dim3(kScalarBlock), 0, c->stream, K.st, part, stride, nb, K.hist);
    FVG_HIP(hipGetLastError());

    const int batch = poll_batch(p);
    for (int64_t it = 0;; ++it) {
        // ap = A p with p.ap
        SpmvEpilogue e1;
        e1.dotw = pp;
        e1.skip = halt;
        FVG_TRY(launch_spmv(c, a, pp, ap, e1));
        hipLaunchKernelGGL(k_scalar<CG_ALPHA>, dim3(1), dim3(kScalarBlock), 0, c->stream, K.st, part, stride, nb, K.hist);
        hipLaunchKernelGGL(k_cg_xr, dim3(nb), dim3(kBlock), 0, c->stream, n, K.st, pp, ap, x->d, r, part);
        hipLaunchKernelGGL(k_scalar<CG_ERROR>, dim3(1), dim3(kScalarBlock), 0, c->stream, K.st, part, stride, nb, K.hist);
        FVG_TRY(apply_prec(c, p, r, z, done));
        hipLaunchKernelGGL(k_dot_g, dim3(nb), dim3(kBlock), 0, c->stream, n, K.st, r, z, part);
        hipLaunchKernelGGL(k_scalar<CG_BETA>, dim3(1), dim3(kScalarBlock), 0, c->stream, K.st, part, stride, nb, K.hist);
        hipLaunchKernelGGL(k_cg_p, dim3(nb), dim3(kBlock), 0, c->stream, n, K.st, z, pp);
        FVG_HIP(hipGetLastError());
        if ((it + 1) % batch == 0) {
            KState h;
            FVG_TRY(K.read(&h));
            if (h.done || h.halt) break;
        }
    }
    FVG_TRY(check_sweep_status(c, p));
    return K.finish(info, history, cap);
}

int run_bicgstab(fvg_ctx* c, const fvg_csr* a, const fvg_prec* p, const fvg_vec* b, fvg_vec* x,
                 const fvg_solver_options* o, fvg_solver_info* info, double* history, int64_t cap) {
    const int64_t n = a->nrows;
    double bnorm;
    FVG_TRY(rhs_norm(c, b, &bnorm));
    if (bnorm < 1e-14) return zero_rhs_exit(c, x, bnorm, info, history, cap);
    Workspace ws{c};
    double *r, *rh, *pp, *y, *v, *s, *z, *t;
    FVG_TRY(ws.get(n, &r));
    FVG_TRY(ws.get(n, &rh));
    FVG_TRY(ws.get(n, &pp));
    FVG_TRY(ws.get(n, &y));
    FVG_TRY(ws.get(n, &v));
    FVG_TRY(ws.get(n, &s));
    FVG_TRY(ws.get(n, &z));
    FVG_TRY(ws.get(n, &t));
    KHost K{c};
    FVG_TRY(K.init(bnorm, o));
    const int nb = spmv_blocks(n);
    FVG_TRY(ensure_partials(c, nb));
    const int64_t stride = c->partials_cap;
    double* part = c->partials;
    const int32_t* done = &K.st->done;
    const int32_t* halt = &K.st->halt;

    // r = b - A x, rh = r, p = r, r.r (bi_conjugate_gradient_stab.rs:165-190)
    SpmvEpilogue e0;
    e0.b = b->d;
    e0.copy1 = rh;
    e0.copy2 = pp;
    e0.dot_self = true;
    FVG_TRY(launch_spmv(c, a, x->d, r, e0));
    hipLaunchKernelGGL(k_scalar<BI_INIT>, dim3(1), dim3(kScalarBlock), 0, c->stream, K.st, part + stride, stride, nb, K.hist);
    FVG_HIP(hipGetLastError());
    {
        KState h;
        FVG_TRY(K.read(&h));
        if (h.init_res < h.atol) {   // :176-178
            if (info) {
                std::memset(info, 0, sizeof(*info));
                info->initial_residual = h.init_res;
                info->final_residual = h.init_res;
                info->history_len = 1;
            }
            if (history && cap > 0) history[0] = h.init_res;
            return FVG_OK;
        }
    }

    const int batch = poll_batch(p);
    for (int64_t it = 0;; ++it) {
        FVG_TRY(apply_prec(c, p, pp, y, halt));                  // y = M^-1 p
        SpmvEpilogue e1;                                         // v = A y, rh.v
        e1.dotw = rh;
        e1.skip = halt;
        FVG_TRY(launch_spmv(c, a, y, v, e1));
        hipLaunchKernelGGL(k_scalar<BI_ALPHA>, dim3(1), dim3(kScalarBlock), 0, c->stream, K.st, part, stride, nb, K.hist);
        hipLaunchKernelGGL(k_bi_hs, dim3(nb), dim3(kBlock), 0, c->stream, n, K.st, y, v, r, x->d, s, part);
        hipLaunchKernelGGL(k_scalar<BI_S>, dim3(1), dim3(kScalarBlock), 0, c->stream, K.st, part, stride, nb, K.hist);
        FVG_TRY(apply_prec(c, p, s, z, done));                   // z = M^-1 s
        SpmvEpilogue e2;                                         // t = A z, t.s, t.t
        e2.dotw = s;
        e2.dot_self = true;
        e2.skip = done;
        FVG_TRY(launch_spmv(c, a, z, t, e2));
        hipLaunchKernelGGL(k_scalar<BI_OMEGA>, dim3(1), dim3(kScalarBlock), 0, c->stream, K.st, part, stride, nb, K.hist);
        hipLaunchKernelGGL(k_bi_xr, dim3(nb), dim3(kBlock), 0, c->stream, n, K.st, z, t, s, rh, x->d, r, part, stride);
        hipLaunchKernelGGL(k_scalar<BI_R>, dim3(1), dim3(kScalarBlock), 0, c->stream, K.st, part, stride, nb, K.hist);
        hipLaunchKernelGGL(k_bi_p, dim3(nb), dim3(kBlock), 0, c->stream, n, K.st, r, v, pp, rh);
        FVG_HIP(hipGetLastError());
        if ((it + 1) % batch == 0) {
            KState h;
            FVG_TRY(K.read(&h));
            if (h.done || h.halt) break;
        }
    }
    FVG_TRY(check_sweep_status(c, p));
    return K.finish(info, history, cap);
}

// mixed.hip
int run_mixed(fvg_ctx* c, fvg_csr* a, fvg_prec* p, const fvg_vec* b, fvg_vec* x, const fvg_solver_options* o,
              fvg_solver_info* info, double* history, int64_t cap, bool bicgstab);

}  // namespace fvg

using namespace fvg;

extern "C" {

int fvg_cg(fvg_ctx* c, const fvg_csr* a, const fvg_prec* p, const fvg_vec* b, fvg_vec* x,
           const fvg_solver_options* o, fvg_solver_info* info, double* history, int64_t cap) {
    FVG_TRY(check_solve_args(c, a, p, b, x, o));
    // the FP32 copies of the matrix values and factors are caches owned by a and p
    if (o->precision == FVG_PRECISION_MIXED)
        return run_mixed(c, const_cast<fvg_csr*>(a), const_cast<fvg_prec*>(p), b, x, o, info, history, cap, false);
    return run_cg(c, a, p, b, x, o, info, history, cap);
}

int fvg_bicgstab(fvg_ctx* c, const fvg_csr* a, const fvg_prec* p, const fvg_vec* b, fvg_vec* x,
                 const fvg_solver_options* o, fvg_solver_info* info, double* history, int64_t cap) {
    FVG_TRY(check_solve_args(c, a, p, b, x, o));
    if (o->precision == FVG_PRECISION_MIXED)
        return run_mixed(c, const_cast<fvg_csr*>(a), const_cast<fvg_prec*>(p), b, x, o, info, history, cap, true);
    return run_bicgstab(c, a, p, b, x, o, info, history, cap);
}

}  // extern "C"
