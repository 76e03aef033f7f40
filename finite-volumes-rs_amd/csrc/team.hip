// team.hip -- batched solves: one workgroup runs one system's whole Krylov solve.
//
// A batch holds many independent systems A_k x_k = b_k (each with its own matrix and
// IC(0) / ILU(0) / Jacobi / identity preconditioner).  The solve of one system is a
// single persistent workgroup: SpMV, triangular sweeps, vector updates and dot products
// all run inside it, separated by workgroup barriers, so a system never waits on another
// workgroup and the whole batch is one kernel launch (replayed from a captured hipGraph).
// That makes each result independent of what else is in the batch: a system gives the
// same bits alone and in any mix, and batches on different streams overlap freely.
//
// Storage is permuted into "positions": rows sorted by the level of the lower triangular
// sweep (dependency depth), so a level is a contiguous range.  The upper sweep runs the
// same levels in reverse (valid for structurally symmetric patterns, which mesh matrices
// are), so both sweeps, the SpMV and every vector pass stream contiguous ranges.  Entries
// keep the original ascending column order, so per-row sums are the reference's
// (dmatrix.rs:136-146, incomplete_cholesky.rs:217-254, incomplete_lower_upper.rs:287-328).
//
// Solver variants (the algorithm steps are those of krylov.hip / mixed.hip):
//   CG64 / BI64   conjugate_gradient / bi_conjugate_gradient_stab in FP64, with the
//                 reference's stopping tests, iteration counting and history;
//   CG32 / BI32   FP32 inner iterations (FP32 values, factors, vectors and accumulation)
//                 inside FP64 iterative refinement on the true residual b - A x.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include "precond.hpp"
#include "reduce.hpp"

namespace fvg {

enum TeamVariant { kCG64 = 0, kBI64 = 1, kCG32 = 2, kBI32 = 3 };

// device-side description of one system (uploaded once per batch, updated when b/x move)
struct TeamDesc {
    int32_t n, nlev, kind, pad;
    const int32_t* perm;    // natural row of each position
    const int32_t* lvl;     // lower-sweep level l holds positions [lvl[l], lvl[l+1])
    const int32_t* a_ptr;   // A by position: entries [a_ptr[p], a_ptr[p+1])
    const int32_t* a_dep;   // position of each entry's column
    const double* a_val;
    const float* a_val32;
    const int32_t* l_ptr;   // strict lower factor by position
    const int32_t* l_dep;
    const double* l_val;
    const float* l_val32;
    const int32_t* u_ptr;   // strict upper factor by processing index q (position n-1-q)
    const int32_t* u_dep;
    const double* u_val;
    const float* u_val32;
    const double* sc;       // per position: IC 1/l_ii, ILU 1/u_ii, Jacobi 1/a_ii
    const float* sc32;
    const double* b;        // caller vectors, natural order
    double* x;
    double* w64;            // workspace
    float* w32;
    double* hist;
    int64_t hcap;
    fvg_solver_info* info;
    double atol, rtol, inner_tol;
    int64_t max_iter;
};

// ---- workgroup reductions (result in every thread, fixed order) -------------------------
template <int NT, class T>
struct Reducer {
    T* buf;   // 2 parities x 2 values x NT/64 waves
    int par;
    __device__ T sum(T v) {
        v = wave_sum(v);
        T* r = buf + par * 2 * (NT / 64);
        par ^= 1;
        if ((threadIdx.x & 63) == 0) r[threadIdx.x >> 6] = v;
        __syncthreads();
        T s = T(0);
#pragma unroll
        for (int i = 0; i < NT / 64; ++i) s += r[i];
        return s;
    }
    __device__ void sum2(T& a, T& b) {
        a = wave_sum(a);
        b = wave_sum(b);
        T* r = buf + par * 2 * (NT / 64);
        par ^= 1;
        if ((threadIdx.x & 63) == 0) {
            r[threadIdx.x >> 6] = a;
            r[NT / 64 + (threadIdx.x >> 6)] = b;
        }
        __syncthreads();
        T s = T(0), t = T(0);
#pragma unroll
        for (int i = 0; i < NT / 64; ++i) {
            s += r[i];
            t += r[NT / 64 + i];
        }
        a = s;
        b = t;
    }
};

template <class T>
__device__ __forceinline__ const T* pick(const double* d, const float* f);
template <>
__device__ __forceinline__ const double* pick<double>(const double* d, const float*) { return d; }
template <>
__device__ __forceinline__ const float* pick<float>(const double*, const float* f) { return f; }

// ---- preconditioner apply: z = M^-1 r (ends with a barrier) -----------------------------
template <int NT, class T>
__device__ void team_lower(const TeamDesc& D, const T* __restrict__ val, const T* sc, const T* b, T* x) {
    for (int l = 0; l < D.nlev; ++l) {
        const int p0 = D.lvl[l], p1 = D.lvl[l + 1];
        for (int p = p0 + (int)threadIdx.x; p < p1; p += NT) {
            T s = b[p];
            const int ke = D.l_ptr[p + 1];
            for (int k = D.l_ptr[p]; k < ke; ++k) s -= val[k] * x[D.l_dep[k]];
            x[p] = sc ? sc[p] * s : s;
        }
        __syncthreads();
    }
}

template <int NT, class T>
__device__ void team_upper(const TeamDesc& D, const T* __restrict__ val, const T* sc, const T* b, T* x) {
    const int n = D.n;
    for (int l = D.nlev - 1; l >= 0; --l) {
        const int p0 = D.lvl[l], p1 = D.lvl[l + 1];
        for (int t = (int)threadIdx.x; t < p1 - p0; t += NT) {
            const int p = p1 - 1 - t;
            const int q = n - 1 - p;
            T s = b[p];
            const int ke = D.u_ptr[q + 1];
            for (int k = D.u_ptr[q]; k < ke; ++k) s -= val[k] * x[D.u_dep[k]];
            x[p] = sc[p] * s;
        }
        __syncthreads();
    }
}

template <int NT, class T>
__device__ void team_prec(const TeamDesc& D, const T* r, T* z, T* tmp) {
    const T* sc = pick<T>(D.sc, D.sc32);
    switch (D.kind) {
    case kIdentity:
        for (int p = threadIdx.x; p < D.n; p += NT) z[p] = r[p];
        __syncthreads();
        break;
    case kDiagonal:
        for (int p = threadIdx.x; p < D.n; p += NT) z[p] = sc[p] * r[p];
        __syncthreads();
        break;
    case kIC:
        team_lower<NT, T>(D, pick<T>(D.l_val, D.l_val32), sc, r, tmp);
        team_upper<NT, T>(D, pick<T>(D.u_val, D.u_val32), sc, tmp, z);
        break;
    default:   // kILU: unit lower factor
        team_lower<NT, T>(D, pick<T>(D.l_val, D.l_val32), (const T*)nullptr, r, tmp);
        team_upper<NT, T>(D, pick<T>(D.u_val, D.u_val32), sc, tmp, z);
        break;
    }
}

// y_p = sum_k a_k x_{dep_k}, in stored (ascending column) order
template <class T>
__device__ __forceinline__ T row_product(const TeamDesc& D, const T* __restrict__ val, const T* x, int p) {
    T s = T(0);
    const int ke = D.a_ptr[p + 1];
    for (int k = D.a_ptr[p]; k < ke; ++k) s += val[k] * x[D.a_dep[k]];
    return s;
}

__device__ __forceinline__ void team_push(const TeamDesc& D, int64_t& len, double e) {
    if (threadIdx.x == 0 && len < D.hcap) D.hist[len] = e;
    len += 1;
}

// gather b and x into positions; returns |b|^2
template <int NT>
__device__ double team_gather(const TeamDesc& D, double* b, double* x) {
    double q = 0.0;
    for (int p = threadIdx.x; p < D.n; p += NT) {
        const int row = D.perm[p];
        const double bi = D.b[row];
        b[p] = bi;
        x[p] = D.x[row];
        q += bi * bi;
    }
    return q;
}

template <int NT>
__device__ void team_scatter(const TeamDesc& D, const double* x) {
    for (int p = threadIdx.x; p < D.n; p += NT) D.x[D.perm[p]] = x[p];
}

__device__ void team_info(const TeamDesc& D, int64_t iters, double init, double fin, int64_t hlen, int64_t restarts,
                          int64_t outer, double true_res) {
    if (threadIdx.x != 0) return;
    fvg_solver_info I;
    I.iterations = iters;
    I.initial_residual = init;
    I.final_residual = fin;
    I.history_len = hlen;
    I.restarts = restarts;
    I.outer_iterations = outer;
    I.true_residual = true_res;
    *D.info = I;
}

// |b| below 1e-14: x = 0 and one history entry (conjugate_gradient.rs:29-36)
template <int NT>
__device__ bool team_zero_rhs(const TeamDesc& D, double bnorm, bool mixed) {
    if (!(bnorm < 1e-14)) return false;
    for (int p = threadIdx.x; p < D.n; p += NT) D.x[D.perm[p]] = 0.0;
    int64_t len = 0;
    team_push(D, len, bnorm);
    team_info(D, 0, bnorm, bnorm, len, 0, 0, mixed ? bnorm : 0.0);
    return true;
}

// ---- FP64 conjugate_gradient (conjugate_gradient.rs:12-112) ------------------------------
template <int NT>
__device__ void team_cg64(const TeamDesc& D, Reducer<NT, double>& R) {
    const int n = D.n, tid = threadIdx.x;
    double* b = D.w64;
    double* x = b + n;
    double* r = x + n;
    double* z = r + n;
    double* p = z + n;
    double* ap = p + n;   // also the intermediate vector of the two sweeps
    const double bnorm = sqrt(R.sum(team_gather<NT>(D, b, x)));
    if (team_zero_rhs<NT>(D, bnorm, false)) return;
    // r = (A x - b) * -1 and |r| (:47-52)
    double q = 0.0;
    for (int i = tid; i < n; i += NT) {
        const double ri = (row_product<double>(D, D.a_val, x, i) - b[i]) * -1.0;
        r[i] = ri;
        q += ri * ri;
    }
    const double init_res = sqrt(R.sum(q)) / bnorm;
    double error = 1.0;
    int64_t iter = 0, hlen = 0;
    if ((error > D.atol) && !(iter > D.max_iter)) {
        team_prec<NT, double>(D, r, z, ap);
        q = 0.0;
        for (int i = tid; i < n; i += NT) {
            p[i] = z[i];
            q += r[i] * z[i];
        }
        double rz = R.sum(q);
        for (;;) {
            q = 0.0;   // ap = A p, p.ap (:72-74)
            for (int i = tid; i < n; i += NT) {
                const double s = row_product<double>(D, D.a_val, p, i);
                ap[i] = s;
                q += p[i] * s;
            }
            const double alpha = rz / R.sum(q);
            q = 0.0;   // (:77-80)
            for (int i = tid; i < n; i += NT) {
                x[i] += p[i] * alpha;
                const double ri = r[i] - ap[i] * alpha;
                r[i] = ri;
                q += ri * ri;
            }
            error = sqrt(R.sum(q)) / bnorm;   // (:83-89)
            if ((error < D.atol) || (error / init_res < D.rtol)) {
                team_push(D, hlen, error);
                iter += 1;
                break;
            }
            team_prec<NT, double>(D, r, z, ap);
            q = 0.0;
            for (int i = tid; i < n; i += NT) q += r[i] * z[i];
            const double rz_new = R.sum(q);
            const double beta = rz_new / rz;   // (:94-105)
            rz = rz_new;
            team_push(D, hlen, error);
            iter += 1;
            if (!(error > D.atol) || iter > D.max_iter) break;
            for (int i = tid; i < n; i += NT) p[i] = z[i] + p[i] * beta;
            __syncthreads();
        }
    }
    __syncthreads();
    team_scatter<NT>(D, x);
    team_info(D, iter, init_res, error, hlen, 0, 0, 0.0);
}

// ---- FP64 bi_conjugate_gradient_stab (bi_conjugate_gradient_stab.rs:13-179) --------------
template <int NT>
__device__ void team_bi64(const TeamDesc& D, Reducer<NT, double>& R) {
    const int n = D.n, tid = threadIdx.x;
    double* b = D.w64;
    double* x = b + n;
    double* r = x + n;
    double* rh = r + n;
    double* p = rh + n;
    double* y = p + n;
    double* v = y + n;
    double* s = v + n;
    double* z = s + n;
    double* t = z + n;   // also the intermediate vector of the sweeps
    const double bnorm = sqrt(R.sum(team_gather<NT>(D, b, x)));
    if (team_zero_rhs<NT>(D, bnorm, false)) return;
    double q = 0.0;   // r = b - A x, rh = p = r (:165-190)
    for (int i = tid; i < n; i += NT) {
        const double ri = (row_product<double>(D, D.a_val, x, i) - b[i]) * -1.0;
        r[i] = ri;
        rh[i] = ri;
        p[i] = ri;
        q += ri * ri;
    }
    const double rr0 = R.sum(q);
    const double init_res = sqrt(rr0) / bnorm;
    double rho = rr0, error = 1.0, alpha = 0.0, omega = 0.0;
    int64_t iter = 0, hlen = 0, restarts = 0;
    bool run = true;
    if (init_res < D.atol) {   // (:176-178)
        error = init_res;
        team_push(D, hlen, init_res);
        run = false;
    } else if (!(error > D.atol) || iter > D.max_iter) {
        run = false;
    }
    while (run) {
        team_prec<NT, double>(D, p, y, t);   // y = M^-1 p
        q = 0.0;                              // v = A y, rh.v (:208-209)
        for (int i = tid; i < n; i += NT) {
            const double vi = row_product<double>(D, D.a_val, y, i);
            v[i] = vi;
            q += rh[i] * vi;
        }
        alpha = rho / R.sum(q);
        q = 0.0;   // h = x + y alpha; s = r - v alpha (:213-216)
        for (int i = tid; i < n; i += NT) {
            x[i] = x[i] + y[i] * alpha;
            const double si = r[i] - v[i] * alpha;
            s[i] = si;
            q += si * si;
        }
        const double ss = sqrt(R.sum(q)) / bnorm;   // (:219-231)
        if ((ss < D.atol) || (ss / init_res < D.rtol)) {
            error = ss;
            team_push(D, hlen, ss);
            iter += 1;
            break;
        }
        team_prec<NT, double>(D, s, z, t);   // z = M^-1 s
        double q0 = 0.0, q1 = 0.0;            // t = A z; t.s, t.t (:239-242)
        for (int i = tid; i < n; i += NT) {
            const double ti = row_product<double>(D, D.a_val, z, i);
            t[i] = ti;
            q0 += s[i] * ti;
            q1 += ti * ti;
        }
        R.sum2(q0, q1);
        omega = q0 / q1;
        q0 = 0.0;
        q1 = 0.0;   // x = h + z omega; r = s - t omega; rh.r, r.r (:246-249)
        for (int i = tid; i < n; i += NT) {
            x[i] = x[i] + z[i] * omega;
            const double ri = s[i] - t[i] * omega;
            r[i] = ri;
            q0 += rh[i] * ri;
            q1 += ri * ri;
        }
        R.sum2(q0, q1);
        const double rr = sqrt(q1) / bnorm;   // (:252-287)
        if ((rr < D.atol) || (rr / init_res < D.rtol)) {
            error = rr;
            team_push(D, hlen, rr);
            iter += 1;
            break;
        }
        const double beta = (q0 / rho) * (alpha / omega);
        rho = q0;
        int mode = 0;
        if (!(fabs(rho) > (1e-15 * bnorm))) {
            mode = 1;
            restarts += 1;
            rho = q1;
        }
        error = rr / bnorm;
        team_push(D, hlen, error);
        iter += 1;
        if (!(error > D.atol) || iter > D.max_iter) break;
        for (int i = tid; i < n; i += NT) {
            if (mode == 0) {
                p[i] = r[i] + (p[i] - v[i] * omega) * beta;
            } else {
                rh[i] = r[i];
                p[i] = r[i];
            }
        }
        __syncthreads();
    }
    __syncthreads();
    team_scatter<NT>(D, x);
    team_info(D, iter, init_res, error, hlen, restarts, 0, 0.0);
}

// ---- mixed precision: FP32 inner solves inside FP64 refinement ---------------------------
// r64 = b - A x in FP64, |r64|, r32 = (float) r64, d = 0; returns |r64|^2 (every thread)
template <int NT>
__device__ double team_outer_residual(const TeamDesc& D, Reducer<NT, double>& R, const double* b, const double* x,
                                      float* r, float* d) {
    double q = 0.0;
    for (int i = threadIdx.x; i < D.n; i += NT) {
        const double ri = (row_product<double>(D, D.a_val, x, i) - b[i]) * -1.0;
        r[i] = (float)ri;
        d[i] = 0.f;
        q += ri * ri;
    }
    return R.sum(q);
}

template <int NT>
__device__ void team_mixed(const TeamDesc& D, Reducer<NT, double>& R, Reducer<NT, float>& F, bool bicg) {
    const int n = D.n, tid = threadIdx.x;
    double* b = D.w64;
    double* x = b + n;
    float* r = D.w32;
    float* d = r + n;
    float* z = d + n;
    float* p = z + n;
    float* ap = p + n;   // CG: A p and the sweeps' intermediate
    float* rh = ap + n;  // BiCGStab only
    float* y = rh + n;
    float* v = y + n;
    float* s = v + n;
    float* t = s + n;    // BiCGStab: A z and the sweeps' intermediate
    const double bnorm = sqrt(R.sum(team_gather<NT>(D, b, x)));
    if (team_zero_rhs<NT>(D, bnorm, true)) return;
    const double inner_tol = D.inner_tol;
    const int64_t max_inner = D.max_iter > 0 ? D.max_iter : 1;
    double init_res = 0.0, true_res = 0.0;
    int64_t total = 0, outer = 0, restarts = 0, hlen = 0;
    for (int k = 0; k < 64; ++k) {
        const double e = sqrt(team_outer_residual<NT>(D, R, b, x, r, d)) / bnorm;
        if (outer == 0) init_res = e;
        true_res = e;
        outer += 1;
        team_push(D, hlen, e);
        if ((e < D.atol) || (e / init_res < D.rtol) || total >= max_inner) break;
        if (!bicg) {
            // z = M r; p = z; rin0 = |r|; rz = r.z
            team_prec<NT, float>(D, r, z, ap);
            float q0 = 0.f, q1 = 0.f;
            for (int i = tid; i < n; i += NT) {
                p[i] = z[i];
                q0 += r[i] * r[i];
                q1 += r[i] * z[i];
            }
            F.sum2(q0, q1);
            const double rin0 = sqrt((double)q0);
            double rz = (double)q1;
            if (!(rin0 > 0.0)) break;   // the FP32 residual underflowed: refinement cannot progress
            {
                for (;;) {
                    float qa = 0.f;
                    for (int i = tid; i < n; i += NT) {
                        const float si = row_product<float>(D, D.a_val32, p, i);
                        ap[i] = si;
                        qa += p[i] * si;
                    }
                    const float alpha = (float)(rz / (double)F.sum(qa));
                    float qr = 0.f;
                    for (int i = tid; i < n; i += NT) {
                        d[i] += alpha * p[i];
                        const float ri = r[i] - alpha * ap[i];
                        r[i] = ri;
                        qr += ri * ri;
                    }
                    const double rin = sqrt((double)F.sum(qr));
                    total += 1;
                    if (rin <= inner_tol * rin0 || total >= max_inner) break;
                    team_prec<NT, float>(D, r, z, ap);
                    float qz = 0.f;
                    for (int i = tid; i < n; i += NT) qz += r[i] * z[i];
                    const double rz_new = (double)F.sum(qz);
                    const float beta = (float)(rz_new / rz);
                    rz = rz_new;
                    for (int i = tid; i < n; i += NT) p[i] = z[i] + beta * p[i];
                    __syncthreads();
                }
            }
        } else {
            float q0 = 0.f;
            for (int i = tid; i < n; i += NT) {
                rh[i] = r[i];
                p[i] = r[i];
                q0 += r[i] * r[i];
            }
            q0 = F.sum(q0);
            const double rin0 = sqrt((double)q0);
            double rho = (double)q0, alpha = 0.0, omega = 0.0;
            if (!(rin0 > 0.0)) break;
            {
                for (;;) {
                    team_prec<NT, float>(D, p, y, t);
                    float qv = 0.f;
                    for (int i = tid; i < n; i += NT) {
                        const float vi = row_product<float>(D, D.a_val32, y, i);
                        v[i] = vi;
                        qv += rh[i] * vi;
                    }
                    alpha = rho / (double)F.sum(qv);
                    const float af = (float)alpha;
                    float qs = 0.f;
                    for (int i = tid; i < n; i += NT) {
                        d[i] += af * y[i];
                        const float si = r[i] - af * v[i];
                        s[i] = si;
                        qs += si * si;
                    }
                    const double sn = sqrt((double)F.sum(qs));
                    if (sn <= inner_tol * rin0) {
                        total += 1;
                        break;
                    }
                    team_prec<NT, float>(D, s, z, t);
                    float qts = 0.f, qtt = 0.f;
                    for (int i = tid; i < n; i += NT) {
                        const float ti = row_product<float>(D, D.a_val32, z, i);
                        t[i] = ti;
                        qts += ti * s[i];
                        qtt += ti * ti;
                    }
                    F.sum2(qts, qtt);
                    omega = (double)qts / (double)qtt;
                    const float of = (float)omega;
                    float qhr = 0.f, qrr = 0.f;
                    for (int i = tid; i < n; i += NT) {
                        d[i] += of * z[i];
                        const float ri = s[i] - of * t[i];
                        r[i] = ri;
                        qhr += rh[i] * ri;
                        qrr += ri * ri;
                    }
                    F.sum2(qhr, qrr);
                    const double rin = sqrt((double)qrr);
                    total += 1;
                    if (rin <= inner_tol * rin0 || total >= max_inner) break;
                    const double rho_new = (double)qhr;
                    const double beta = (rho_new / rho) * (alpha / omega);
                    rho = rho_new;
                    int mode = 0;
                    if (!(fabs(rho) > 1e-30 * rin0 * rin0)) {
                        mode = 1;
                        restarts += 1;
                        rho = (double)qrr;
                    }
                    const float bf = (float)beta;
                    for (int i = tid; i < n; i += NT) {
                        if (mode == 0) {
                            p[i] = r[i] + bf * (p[i] - of * v[i]);
                        } else {
                            rh[i] = r[i];
                            p[i] = r[i];
                        }
                    }
                    __syncthreads();
                }
            }
        }
        for (int i = tid; i < n; i += NT) x[i] += (double)d[i];   // FP64 solution update
        __syncthreads();
    }
    __syncthreads();
    team_scatter<NT>(D, x);
    team_info(D, total, init_res, true_res, hlen, restarts, outer, true_res);
}

template <int NT, int VAR>
__global__ void __launch_bounds__(NT) k_team(const TeamDesc* __restrict__ descs, int n_inst) {
    __shared__ double red64[4 * (NT / 64)];
    __shared__ float red32[4 * (NT / 64)];
    Reducer<NT, double> R{red64, 0};
    Reducer<NT, float> F{red32, 0};
    for (int k = blockIdx.x; k < n_inst; k += gridDim.x) {
        const TeamDesc D = descs[k];
        if (VAR == kCG64) team_cg64<NT>(D, R);
        if (VAR == kBI64) team_bi64<NT>(D, R);
        if (VAR == kCG32) team_mixed<NT>(D, R, F, false);
        if (VAR == kBI32) team_mixed<NT>(D, R, F, true);
        __syncthreads();
    }
}

// out64[k] = in[src[k]], out32[k] = (float) in[src[k]]
__global__ void k_team_gather(int64_t n, const int32_t* __restrict__ src, const double* __restrict__ in,
                              double* __restrict__ out64, float* __restrict__ out32) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double v = in[src[k]];
    out64[k] = v;
    out32[k] = (float)v;
}

// ---- host side ---------------------------------------------------------------------------

template <int NT>
static const void* team_kernel(int var) {
    switch (var) {
    case kCG64: return (const void*)k_team<NT, kCG64>;
    case kBI64: return (const void*)k_team<NT, kBI64>;
    case kCG32: return (const void*)k_team<NT, kCG32>;
    default: return (const void*)k_team<NT, kBI32>;
    }
}

static const void* team_kernel_for(int nt, int var) {
    switch (nt) {
    case 256: return team_kernel<256>(var);
    case 512: return team_kernel<512>(var);
    default: return team_kernel<1024>(var);
    }
}

}  // namespace fvg

// arrays of one system, as offsets into the batch arenas
struct TeamLayout {
    int64_t n = 0, nlev = 0, nnz_a = 0, nnz_l = 0, nnz_u = 0;
    int64_t perm = 0, lvl = 0, a_ptr = 0, a_dep = 0, l_ptr = 0, l_dep = 0, u_ptr = 0, u_dep = 0;   // int arena
    int64_t a_src = 0, l_src = 0, u_src = 0;                                                    // int arena
    int64_t a_val = 0, l_val = 0, u_val = 0, sc = 0;                                            // value arenas
    int64_t w64 = 0, w32 = 0, hist = 0;
};

struct fvg_batch {
    int n_inst = 0;
    int solver = 0;      // 0 CG, 1 BiCGStab
    int precision = 0;
    int variant = 0;
    int nt = 1024;
    fvg_solver_options opts{};
    std::vector<const fvg_csr*> A;
    std::vector<const fvg_prec*> P;
    std::vector<uint64_t> a_ver, p_ver;
    std::vector<TeamLayout> lay;
    std::vector<int> order;           // launch slot -> system (largest first)
    int32_t* iarena = nullptr;
    int64_t ni = 0;
    double* darena = nullptr;
    float* farena = nullptr;
    int64_t nv = 0;
    double* w64 = nullptr;
    int64_t nw64 = 0;
    float* w32 = nullptr;
    int64_t nw32 = 0;
    double* hist = nullptr;
    int64_t nhist = 0, hcap = 0;
    fvg_solver_info* d_info = nullptr;
    fvg::TeamDesc* d_desc = nullptr;
    std::vector<fvg::TeamDesc> desc;  // by launch slot
    hipStream_t stream = nullptr;
    hipEvent_t ready = nullptr;       // recorded on the context stream before each launch
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int graph_nt = 0;
    bool launched = false;
};

namespace fvg {

static void batch_free(fvg_ctx* c, fvg_batch* B) {
    if (B->stream) (void)hipStreamSynchronize(B->stream);
    if (B->exec) (void)hipGraphExecDestroy(B->exec);
    if (B->graph) (void)hipGraphDestroy(B->graph);
    dev_free_n(c, B->iarena, B->ni);
    dev_free_n(c, B->darena, B->nv);
    dev_free_n(c, B->farena, B->nv);
    dev_free_n(c, B->w64, B->nw64);
    dev_free_n(c, B->w32, B->nw32);
    dev_free_n(c, B->hist, B->nhist);
    dev_free_n(c, B->d_info, B->n_inst);
    dev_free_n(c, B->d_desc, B->n_inst);
    if (B->ready) (void)hipEventDestroy(B->ready);
    if (B->stream) (void)hipStreamDestroy(B->stream);
}

// Host analysis of one system: level order, permuted A, lower and upper factor
// entries with their source slots.  Appends index data to `ints`.
static int team_analyse(const fvg_csr* a, const fvg_prec* p, TeamLayout& L, std::vector<int32_t>& ints) {
    const int64_t n = a->nrows;
    const auto& rp = a->h_row_ptr;
    const auto& col = a->h_col;
    const bool fac = p->kind == kIC || p->kind == kILU;
    // structural symmetry (the upper sweep runs the lower levels in reverse)
    if (fac) {
        for (int64_t i = 0; i < n; ++i)
            for (int32_t k = rp[(size_t)i]; k < rp[(size_t)i + 1]; ++k) {
                const int32_t j = col[(size_t)k];
                auto b = col.begin() + rp[(size_t)j], e = col.begin() + rp[(size_t)j + 1];
                if (!std::binary_search(b, e, (int32_t)i))
                    return fail(FVG_E_UNSUPPORTED, "batched solves need a structurally symmetric matrix pattern");
            }
    }
    std::vector<int32_t> lev((size_t)n, 0);
    int32_t maxlev = 0;
    if (fac) {
        for (int64_t i = 0; i < n; ++i) {
            int32_t l = 0;
            for (int32_t k = rp[(size_t)i]; k < rp[(size_t)i + 1]; ++k) {
                const int32_t j = col[(size_t)k];
                if (j < i) l = std::max(l, lev[(size_t)j] + 1);
            }
            lev[(size_t)i] = l;
            maxlev = std::max(maxlev, l);
        }
    }
    const int32_t nlev = maxlev + 1;
    std::vector<int32_t> lvl((size_t)nlev + 1, 0);
    for (int64_t i = 0; i < n; ++i) lvl[(size_t)lev[(size_t)i] + 1]++;
    for (int32_t l = 0; l < nlev; ++l) lvl[(size_t)l + 1] += lvl[(size_t)l];
    std::vector<int32_t> perm((size_t)n), pos((size_t)n);
    {
        std::vector<int32_t> cur(lvl.begin(), lvl.end() - 1);
        for (int64_t i = 0; i < n; ++i) {
            const int32_t q = cur[(size_t)lev[(size_t)i]]++;
            perm[(size_t)q] = (int32_t)i;
            pos[(size_t)i] = q;
        }
    }
    std::vector<int32_t> a_ptr((size_t)n + 1, 0), a_dep, a_src, l_ptr((size_t)n + 1, 0), l_dep, l_src,
        u_ptr((size_t)n + 1, 0), u_dep, u_src;
    a_dep.reserve(col.size());
    a_src.reserve(col.size());
    for (int64_t q = 0; q < n; ++q) {
        const int32_t i = perm[(size_t)q];
        for (int32_t k = rp[(size_t)i]; k < rp[(size_t)i + 1]; ++k) {
            const int32_t j = col[(size_t)k];
            a_dep.push_back(pos[(size_t)j]);
            a_src.push_back(k);
            if (fac && j < i) {
                l_dep.push_back(pos[(size_t)j]);
                l_src.push_back(k);
            }
        }
        a_ptr[(size_t)q + 1] = (int32_t)a_dep.size();
        l_ptr[(size_t)q + 1] = (int32_t)l_dep.size();
    }
    if (fac) {
        for (int64_t q = 0; q < n; ++q) {   // processing index q -> position n-1-q
            const int32_t i = perm[(size_t)(n - 1 - q)];
            for (int32_t k = rp[(size_t)i]; k < rp[(size_t)i + 1]; ++k) {
                const int32_t j = col[(size_t)k];
                if (j > i) {
                    u_dep.push_back(pos[(size_t)j]);
                    u_src.push_back(p->kind == kIC ? p->h_tpos[(size_t)k] : k);
                }
            }
            u_ptr[(size_t)q + 1] = (int32_t)u_dep.size();
        }
    }
    L.n = n;
    L.nlev = nlev;
    L.nnz_a = (int64_t)a_dep.size();
    L.nnz_l = (int64_t)l_dep.size();
    L.nnz_u = (int64_t)u_dep.size();
    auto put = [&](const std::vector<int32_t>& v) {
        const int64_t off = (int64_t)ints.size();
        ints.insert(ints.end(), v.begin(), v.end());
        return off;
    };
    L.perm = put(perm);
    L.lvl = put(lvl);
    L.a_ptr = put(a_ptr);
    L.a_dep = put(a_dep);
    L.a_src = put(a_src);
    L.l_ptr = put(l_ptr);
    L.l_dep = put(l_dep);
    L.l_src = put(l_src);
    L.u_ptr = put(u_ptr);
    L.u_dep = put(u_dep);
    L.u_src = put(u_src);
    return FVG_OK;
}

// copy matrix and factor values into the batch's permuted arrays (FP64 and FP32)
static int team_refresh(fvg_ctx* c, fvg_batch* B, int k, hipStream_t s) {
    const TeamLayout& L = B->lay[(size_t)k];
    const fvg_csr* a = B->A[(size_t)k];
    const fvg_prec* p = B->P[(size_t)k];
    auto gather = [&](int64_t cnt, int64_t src_off, const double* in, int64_t dst) {
        if (cnt <= 0 || !in) return;
        hipLaunchKernelGGL(k_team_gather, dim3(grid_for(cnt)), dim3(kBlock), 0, s, cnt, B->iarena + src_off, in,
                           B->darena + dst, B->farena + dst);
    };
    gather(L.nnz_a, L.a_src, a->val, L.a_val);
    if (p->kind == kIC || p->kind == kILU) {
        gather(L.nnz_l, L.l_src, p->d_fac, L.l_val);
        gather(L.nnz_u, L.u_src, p->d_fac, L.u_val);
    }
    if (p->kind != kIdentity) gather(L.n, L.perm, p->d_inv_diag, L.sc);
    FVG_HIP(hipGetLastError());
    B->a_ver[(size_t)k] = a->version;
    B->p_ver[(size_t)k] = p->fac_version;
    return FVG_OK;
}

static int batch_capture(fvg_batch* B) {
    if (B->exec) {
        (void)hipGraphExecDestroy(B->exec);
        B->exec = nullptr;
    }
    if (B->graph) {
        (void)hipGraphDestroy(B->graph);
        B->graph = nullptr;
    }
    const void* fn = team_kernel_for(B->nt, B->variant);
    int n_inst = B->n_inst;
    void* args[] = {(void*)&B->d_desc, (void*)&n_inst};
    FVG_HIP(hipStreamBeginCapture(B->stream, hipStreamCaptureModeThreadLocal));
    const hipError_t le = hipLaunchKernel(fn, dim3((unsigned)B->n_inst), dim3((unsigned)B->nt), args, 0, B->stream);
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(B->stream, &g);
    if (le != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return fail(FVG_E_HIP, std::string("batch kernel launch: ") + hipGetErrorString(le));
    }
    FVG_HIP(ce);
    B->graph = g;
    FVG_HIP(hipGraphInstantiate(&B->exec, B->graph, nullptr, nullptr, 0));
    B->graph_nt = B->nt;
    return FVG_OK;
}

}  // namespace fvg

using namespace fvg;

extern "C" {

int fvg_batch_create(fvg_ctx* c, int32_t n, const fvg_csr* const* a, const fvg_prec* const* p, int32_t solver,
                     const fvg_solver_options* o, fvg_batch** out) {
    if (!c || !a || !p || !o || !out || n <= 0) return fail(FVG_E_INVALID, "bad argument");
    if (solver != FVG_SOLVER_CG && solver != FVG_SOLVER_BICGSTAB) return fail(FVG_E_INVALID, "unknown solver");
    if (o->max_iterations < 0) return fail(FVG_E_INVALID, "max_iterations must be non-negative");
    *out = nullptr;
    for (int k = 0; k < n; ++k) {
        if (!a[k] || !p[k]) return fail(FVG_E_INVALID, "null matrix or preconditioner");
        if (a[k]->nrows != a[k]->ncols || a[k]->nrows <= 0) return fail(FVG_E_INVALID, "batched systems must be square");
        if (a[k]->nrows > INT32_MAX / 2) return fail(FVG_E_INVALID, "system too large for a batch");
        if (p[k]->n != a[k]->nrows) return fail(FVG_E_INVALID, "preconditioner size does not match its matrix");
        if ((p[k]->kind == kIC || p[k]->kind == kILU) && p[k]->h_col != a[k]->h_col)
            return fail(FVG_E_INVALID, "preconditioner pattern differs from its matrix");
    }
    auto* B = new fvg_batch();
    B->n_inst = n;
    B->solver = solver;
    B->precision = o->precision;
    B->opts = *o;
    B->variant = (solver == FVG_SOLVER_CG ? 0 : 1) + (o->precision == FVG_PRECISION_MIXED ? 2 : 0);
    B->nt = 1024;
    B->A.assign(a, a + n);
    B->P.assign(p, p + n);
    B->a_ver.assign((size_t)n, ~0ull);
    B->p_ver.assign((size_t)n, ~0ull);
    B->lay.resize((size_t)n);
    std::vector<int32_t> ints;
    int s = FVG_OK;
    for (int k = 0; k < n && s == FVG_OK; ++k) s = team_analyse(a[k], p[k], B->lay[(size_t)k], ints);
    const bool mixed = o->precision == FVG_PRECISION_MIXED;
    const int w64_per = mixed ? 2 : (solver == FVG_SOLVER_CG ? 6 : 10);
    const int w32_per = mixed ? (solver == FVG_SOLVER_CG ? 5 : 10) : 0;
    B->hcap = mixed ? 64 : o->max_iterations + 2;
    for (int k = 0; k < n && s == FVG_OK; ++k) {
        TeamLayout& L = B->lay[(size_t)k];
        L.a_val = B->nv;
        B->nv += L.nnz_a;
        L.l_val = B->nv;
        B->nv += L.nnz_l;
        L.u_val = B->nv;
        B->nv += L.nnz_u;
        L.sc = B->nv;
        B->nv += L.n;
        L.w64 = B->nw64;
        B->nw64 += w64_per * L.n;
        L.w32 = B->nw32;
        B->nw32 += w32_per * L.n;
        L.hist = B->nhist;
        B->nhist += B->hcap;
    }
    B->ni = (int64_t)ints.size();
    if (s == FVG_OK) s = dev_alloc_n(c, B->ni, &B->iarena);
    if (s == FVG_OK) s = dev_alloc_n(c, B->nv, &B->darena);
    if (s == FVG_OK) s = dev_alloc_n(c, B->nv, &B->farena);
    if (s == FVG_OK) s = dev_alloc_n(c, B->nw64, &B->w64);
    if (s == FVG_OK) s = dev_alloc_n(c, B->nw32, &B->w32);
    if (s == FVG_OK) s = dev_alloc_n(c, B->nhist, &B->hist);
    if (s == FVG_OK) s = dev_alloc_n(c, n, &B->d_info);
    if (s == FVG_OK) s = dev_alloc_n(c, n, &B->d_desc);
    if (s == FVG_OK) s = upload(c, B->iarena, ints.data(), B->ni);
    if (s == FVG_OK && hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking) != hipSuccess)
        s = fail(FVG_E_HIP, "batch stream creation failed");
    if (s == FVG_OK && hipEventCreateWithFlags(&B->ready, hipEventDisableTiming) != hipSuccess)
        s = fail(FVG_E_HIP, "batch event creation failed");
    for (int k = 0; k < n && s == FVG_OK; ++k) s = team_refresh(c, B, k, c->stream);
    if (s == FVG_OK) {
        // largest systems first, so they start first and spread over the XCDs
        B->order.resize((size_t)n);
        std::iota(B->order.begin(), B->order.end(), 0);
        std::stable_sort(B->order.begin(), B->order.end(),
                         [&](int x, int y) { return B->lay[(size_t)x].n > B->lay[(size_t)y].n; });
        B->desc.resize((size_t)n);
        for (int slot = 0; slot < n; ++slot) {
            const int k = B->order[(size_t)slot];
            const TeamLayout& L = B->lay[(size_t)k];
            TeamDesc& D = B->desc[(size_t)slot];
            std::memset(&D, 0, sizeof(D));
            D.n = (int32_t)L.n;
            D.nlev = (int32_t)L.nlev;
            D.kind = p[k]->kind;
            int32_t* I = B->iarena;
            D.perm = I + L.perm;
            D.lvl = I + L.lvl;
            D.a_ptr = I + L.a_ptr;
            D.a_dep = I + L.a_dep;
            D.a_val = B->darena + L.a_val;
            D.a_val32 = B->farena + L.a_val;
            D.l_ptr = I + L.l_ptr;
            D.l_dep = I + L.l_dep;
            D.l_val = B->darena + L.l_val;
            D.l_val32 = B->farena + L.l_val;
            D.u_ptr = I + L.u_ptr;
            D.u_dep = I + L.u_dep;
            D.u_val = B->darena + L.u_val;
            D.u_val32 = B->farena + L.u_val;
            D.sc = B->darena + L.sc;
            D.sc32 = B->farena + L.sc;
            D.w64 = B->w64 + L.w64;
            D.w32 = B->w32 + L.w32;
            D.hist = B->hist + L.hist;
            D.hcap = B->hcap;
            D.info = B->d_info + k;
            D.atol = o->absolute_tolerance;
            D.rtol = o->relative_tolerance;
            D.inner_tol = o->inner_tolerance > 0.0 ? o->inner_tolerance : 1e-4;
            D.max_iter = o->max_iterations;
        }
        if (hipMemcpyAsync(B->d_desc, B->desc.data(), sizeof(TeamDesc) * B->desc.size(), hipMemcpyHostToDevice,
                           c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            s = fail(FVG_E_HIP, "batch descriptor upload failed");
    }
    if (s != FVG_OK) {
        (void)hipStreamSynchronize(c->stream);
        batch_free(c, B);
        delete B;
        return s;
    }
    *out = B;
    return FVG_OK;
}

int fvg_batch_destroy(fvg_ctx* c, fvg_batch* B) {
    if (!c || !B) return fail(FVG_E_INVALID, "null argument");
    batch_free(c, B);
    delete B;
    return FVG_OK;
}

int fvg_batch_size(const fvg_batch* B, int32_t* n_systems, int64_t* total_rows) {
    if (!B) return fail(FVG_E_INVALID, "null argument");
    if (n_systems) *n_systems = B->n_inst;
    if (total_rows) {
        int64_t t = 0;
        for (const auto& L : B->lay) t += L.n;
        *total_rows = t;
    }
    return FVG_OK;
}

int fvg_batch_launch(fvg_ctx* c, fvg_batch* B, const fvg_vec* const* b, fvg_vec* const* x) {
    if (!c || !B || !b || !x) return fail(FVG_E_INVALID, "null argument");
    if (B->launched) return fail(FVG_E_INVALID, "batch already in flight: finish it first");
    bool moved = false;
    for (int slot = 0; slot < B->n_inst; ++slot) {
        const int k = B->order[(size_t)slot];
        const fvg_vec* bk = b[k];
        fvg_vec* xk = x[k];
        if (!bk || !xk) return fail(FVG_E_INVALID, "null vector");
        if (bk->n != B->lay[(size_t)k].n || xk->n != B->lay[(size_t)k].n)
            return fail(FVG_E_INVALID, "vector size does not match its system");
        if (bk->d == xk->d) return fail(FVG_E_INVALID, "b and x must not alias");
        TeamDesc& D = B->desc[(size_t)slot];
        if (D.b != bk->d || D.x != xk->d) {
            D.b = bk->d;
            D.x = xk->d;
            moved = true;
        }
    }
    // refresh values assembled or refactored since the last launch
    for (int k = 0; k < B->n_inst; ++k)
        if (B->a_ver[(size_t)k] != B->A[(size_t)k]->version || B->p_ver[(size_t)k] != B->P[(size_t)k]->fac_version)
            FVG_TRY(team_refresh(c, B, k, c->stream));
    if (moved) FVG_HIP(hipMemcpyAsync(B->d_desc, B->desc.data(), sizeof(TeamDesc) * B->desc.size(),
                                      hipMemcpyHostToDevice, c->stream));
    // the batch stream starts after everything enqueued on the context stream so far
    FVG_HIP(hipEventRecord(B->ready, c->stream));
    FVG_HIP(hipStreamWaitEvent(B->stream, B->ready, 0));
    if (!B->exec || B->graph_nt != B->nt) FVG_TRY(batch_capture(B));
    FVG_HIP(hipGraphLaunch(B->exec, B->stream));
    B->launched = true;
    return FVG_OK;
}

int fvg_batch_finish(fvg_ctx* c, fvg_batch* B, fvg_solver_info* infos, double* const* histories, int64_t cap) {
    if (!c || !B) return fail(FVG_E_INVALID, "null argument");
    if (!B->launched) return fail(FVG_E_INVALID, "batch not launched");
    B->launched = false;
    FVG_HIP(hipStreamSynchronize(B->stream));
    std::vector<fvg_solver_info> tmp;
    if (!infos && histories && cap > 0) {
        tmp.resize((size_t)B->n_inst);
        infos = tmp.data();
    }
    if (infos)
        FVG_HIP(hipMemcpyAsync(infos, B->d_info, sizeof(fvg_solver_info) * (size_t)B->n_inst, hipMemcpyDeviceToHost,
                               B->stream));
    FVG_HIP(hipStreamSynchronize(B->stream));
    if (histories && cap > 0) {
        for (int k = 0; k < B->n_inst; ++k) {
            if (!histories[k]) continue;
            const int64_t m = std::min(infos[k].history_len, std::min(cap, B->hcap));
            if (m > 0)
                FVG_HIP(hipMemcpyAsync(histories[k], B->hist + B->lay[(size_t)k].hist, sizeof(double) * (size_t)m,
                                       hipMemcpyDeviceToHost, B->stream));
        }
        FVG_HIP(hipStreamSynchronize(B->stream));
    }
    return FVG_OK;
}

int fvg_batch_query(fvg_ctx* c, fvg_batch* B, int32_t* done) {
    if (!c || !B || !done) return fail(FVG_E_INVALID, "null argument");
    if (!B->launched) {
        *done = 1;
        return FVG_OK;
    }
    const hipError_t e = hipStreamQuery(B->stream);
    if (e == hipSuccess) {
        *done = 1;
        return FVG_OK;
    }
    if (e == hipErrorNotReady) {
        *done = 0;
        return FVG_OK;
    }
    return fail(FVG_E_HIP, std::string("batch stream: ") + hipGetErrorString(e));
}

int fvg_batch_solve(fvg_ctx* c, fvg_batch* B, const fvg_vec* const* b, fvg_vec* const* x, fvg_solver_info* infos,
                    double* const* histories, int64_t cap) {
    FVG_TRY(fvg_batch_launch(c, B, b, x));
    return fvg_batch_finish(c, B, infos, histories, cap);
}

int fvg_batch_set_threads(fvg_ctx* c, fvg_batch* B, int32_t threads) {
    if (!c || !B) return fail(FVG_E_INVALID, "null argument");
    if (threads != 256 && threads != 512 && threads != 1024)
        return fail(FVG_E_INVALID, "workgroup size must be 256, 512 or 1024");
    if (B->launched) return fail(FVG_E_INVALID, "batch in flight");
    B->nt = threads;
    return FVG_OK;
}

int fvg_batch_get_threads(const fvg_batch* B, int32_t* threads) {
    if (!B || !threads) return fail(FVG_E_INVALID, "null argument");
    *threads = B->nt;
    return FVG_OK;
}

}  // extern "C"
