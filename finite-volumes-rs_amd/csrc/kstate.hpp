// kstate.hpp -- Krylov solver state and its scalar steps (internal).
//
// The scalar steps are the reference's scalar algebra, in its order:
// conjugate_gradient.rs:12-112 and bi_conjugate_gradient_stab.rs:13-179.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "reduce.hpp"

namespace fvg {

// Per-system solver state, resident on the device.
struct KState {
    double bnorm, init_res, error, atol, rtol;
    double rz, alpha, beta, rho, omega;
    int64_t iter, max_iter, hist_len, hist_cap, restarts;
    int32_t done;      // converged or stopped inside an iteration: the rest of it is skipped
    int32_t halt;      // the system no longer iterates
    int32_t mode;      // BiCGStab direction update: 0 update, 1 restart
    int32_t counted;   // already retired from the batch's active count
    int32_t zero_rhs;  // |b| < 1e-14: x = 0 answered without iterating
    int32_t pad;
};

enum ScalarStep {
    ST_BNORM = 0,   // |b|, zero right-hand side exit (conjugate_gradient.rs:29-36)
    CG_INIT,        // initial residual, loop test
    CG_RZ0,         // r.z of the first iteration
    CG_ALPHA,       // alpha = rz / p.Ap
    CG_ERROR,       // error = |r| / |b| and the convergence test
    CG_BETA,        // beta = rz_new / rz, history, loop test
    BI_INIT,        // initial residual, rho = rh.r, early exit (bi_conjugate_gradient_stab.rs:174-190)
    BI_ALPHA,       // alpha = rho / rh.v
    BI_S,           // |s| test
    BI_OMEGA,       // omega = t.s / t.t
    BI_R,           // |r| test, rho, beta, restart decision, history, loop test
};

__device__ __forceinline__ void push_history(KState* st, double* hist, double e) {
    if (st->hist_len < st->hist_cap) hist[st->hist_len] = e;
    st->hist_len += 1;
}

// first time a system stops iterating: retire it from the batch
__device__ __forceinline__ void retire(KState* st, int32_t* active, int32_t* all_done) {
    st->halt = 1;
    if (!st->counted) {
        st->counted = 1;
        if (atomicSub(active, 1) == 1) *all_done = 1;
    }
}

__device__ __forceinline__ void loop_test(KState* st, int32_t* active, int32_t* all_done) {
    // while error > tolerance { if iter > max_iter { break } ... }
    if (!(st->error > st->atol) || st->iter > st->max_iter) retire(st, active, all_done);
}

// The scalar step `step` on the sums s0 (partial slot 0) and s1 (slot 1).  One thread.
__device__ inline void scalar_step(int step, KState* st, double* hist, double s0, double s1, int32_t* active,
                                   int32_t* all_done) {
    switch (step) {
    case ST_BNORM:
        st->bnorm = sqrt(s0);
        if (st->bnorm < 1e-14) {
            st->zero_rhs = 1;
            st->init_res = st->bnorm;
            st->error = st->bnorm;
            push_history(st, hist, st->bnorm);
            st->done = 1;
            retire(st, active, all_done);
        }
        break;
    case CG_INIT:   // s1 = r.r   (conjugate_gradient.rs:52, 62-66)
        st->init_res = sqrt(s1) / st->bnorm;
        st->error = 1.0;
        st->iter = 0;
        loop_test(st, active, all_done);
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
            retire(st, active, all_done);
        }
        break;
    }
    case CG_BETA:   // s0 = r.z_new (:94-105)
        st->beta = s0 / st->rz;
        st->rz = s0;
        push_history(st, hist, st->error);
        st->iter += 1;
        loop_test(st, active, all_done);
        break;
    case BI_INIT:   // s1 = r.r; rho = rh.r with rh == r (bi_conjugate_gradient_stab.rs:174-190)
        st->init_res = sqrt(s1) / st->bnorm;
        st->rho = s1;
        st->error = 1.0;
        st->iter = 0;
        if (st->init_res < st->atol) {   // :176-178
            st->error = st->init_res;
            push_history(st, hist, st->init_res);
            st->done = 1;
            retire(st, active, all_done);
            break;
        }
        loop_test(st, active, all_done);
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
            retire(st, active, all_done);
        }
        break;
    }
    case BI_OMEGA:  // s0 = t.s, s1 = t.t (:239-242)
        st->omega = s0 / s1;
        break;
    case BI_R: {    // s1 = r.r, s0 = rh.r (:252-287)
        const double rr = sqrt(s1) / st->bnorm;
        if ((rr < st->atol) || (rr / st->init_res < st->rtol)) {
            st->error = rr;
            push_history(st, hist, rr);
            st->iter += 1;
            st->done = 1;
            retire(st, active, all_done);
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
        loop_test(st, active, all_done);
        break;
    }
    }
}

}  // namespace fvg
