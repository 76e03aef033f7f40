// tools.hip -- fvm::tools on the device: compute_gradients (src/fvm/tools/gradients.rs:41-79)
// and compute_limiters (src/fvm/tools/limiters.rs:47-126).
//
// Gradients: one thread per cell, summing its listings in the order the cell lists its
// faces, with the least-squares coefficients the mesh build made.  Limiters: one
// thread per face.  Both repeat the reference's operation order, so results are
// bit-identical to its sequential loops.  A field with NC components is NC scalar
// fields for the gradient (Vector::outer and f64 * Vector give the same products); the
// limiter ratio of a vector field uses the vector form of to_limiter_r.
#include <string>

#include "schemes.hpp"

namespace fvg {

namespace {

constexpr int kGradUnroll = 8;   // listings up to this long take the unrolled path

template <int D>
__global__ void __launch_bounds__(kBlock) k_gradients(int n, int nc, const int32_t* __restrict__ lptr,
                                                      const int32_t* __restrict__ lother,
                                                      const double* __restrict__ coef,
                                                      const double* __restrict__ diag,
                                                      const double* __restrict__ values,
                                                      const double* __restrict__ bci,
                                                      const double* __restrict__ bce, double* __restrict__ out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const int s = lptr[c], e = lptr[c + 1], m = e - s;
    if (m >= 1 && m <= kGradUnroll) {
        // all of the listing's entries and coefficients in flight at once (indices clamped
        // into the listing), then all the neighbour values; the products are added in
        // listing order, as in the loop below
        int32_t o[kGradUnroll];
        double cf[kGradUnroll][D];
#pragma unroll
        for (int u = 0; u < kGradUnroll; ++u) {
            const int k = s + min(u, m - 1);
            o[u] = lother[k];
#pragma unroll
            for (int d = 0; d < D; ++d) cf[u][d] = coef[(int64_t)k * D + d];
        }
        for (int q = 0; q < nc; ++q) {
            const double vc = values[(int64_t)c * nc + q];
            double v[kGradUnroll];
#pragma unroll
            for (int u = 0; u < kGradUnroll; ++u) {
                if (o[u] >= 0) {
                    v[u] = values[(int64_t)o[u] * nc + q];
                } else {   // boundary: value_j = bci * value_i + bce
                    const int b = -o[u] - 1;
                    v[u] = bci[b] * vc + bce[(int64_t)b * nc + q];
                }
            }
            double gr[D];
#pragma unroll
            for (int d = 0; d < D; ++d) gr[d] = diag[(int64_t)c * D + d] * vc;
#pragma unroll
            for (int u = 0; u < kGradUnroll; ++u)
                if (u < m) {
#pragma unroll
                    for (int d = 0; d < D; ++d) gr[d] += cf[u][d] * v[u];
                }
#pragma unroll
            for (int d = 0; d < D; ++d) out[((int64_t)c * nc + q) * D + d] = gr[d];
        }
        return;
    }
    for (int q = 0; q < nc; ++q) {
        const double vc = values[(int64_t)c * nc + q];
        double gr[D];
#pragma unroll
        for (int d = 0; d < D; ++d) gr[d] = diag[(int64_t)c * D + d] * vc;   // mul_to_gradient(own_grad)
        for (int k = s; k < e; ++k) {
            const int32_t o = lother[k];
            double v;
            if (o >= 0) {
                v = values[(int64_t)o * nc + q];
            } else {   // boundary: value_j = bci * value_i + bce
                const int b = -o - 1;
                v = bci[b] * vc + bce[(int64_t)b * nc + q];
            }
#pragma unroll
            for (int d = 0; d < D; ++d) gr[d] += coef[(int64_t)k * D + d] * v;
        }
#pragma unroll
        for (int d = 0; d < D; ++d) out[((int64_t)c * nc + q) * D + d] = gr[d];
    }
}

// schemes/limiters/mod.rs:17-82
__device__ __forceinline__ double limiter_value(int kind, double k, double r) {
    switch (kind) {
    case FVG_LIMITER_LIMITED_LINEAR: return fmin(2.0 * r * k, 1.0);
    case FVG_LIMITER_MINMOD: return fmin(1.0 * r, 1.0);
    case FVG_LIMITER_VANLEER: return 2.0 * r / (1.0 + r);
    case FVG_LIMITER_VANALBADA: return (r * r + r) / (r * r + 1.0);
    default: {   // Sweby(k); SuperBee is Sweby(2)
        const double beta = kind == FVG_LIMITER_SUPERBEE ? 2.0 : k;
        const double a = fmin(r, beta), b = fmin(beta * r, 1.0);
        return fmax(fmax(a, b), 0.0);
    }
    }
}

struct LimArgs {
    int limiter, interp, fngrad;
    double k, reduction;
    const double* flux;      // Upwind interpolation of the gradients
    const double* fgrads;    // Corrected face-normal gradient
};

template <int D, int NC>
__global__ void __launch_bounds__(kBlock) k_limiters(int n_faces, MeshView g, const int32_t* __restrict__ bnd,
                                                     const double* __restrict__ tinv,
                                                     const double* __restrict__ values,
                                                     const double* __restrict__ grads, LimArgs a,
                                                     const double* __restrict__ bci,
                                                     const double* __restrict__ bce, double* __restrict__ out) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= n_faces) return;
    const int i = g.owner[f], j = g.neigh[f];
    // FaceNormalGradientScheme::terms: Orthogonal (orthogonal.rs:26-44) or Corrected
    double gi, gj, grhs[NC];
    if (a.fngrad == FVG_FNGRAD_CORRECTED) {
        corrected_terms<D, NC>(g, f, i, j, a.fgrads, a.reduction, gi, gj, grhs);
    } else {
        const double t = tinv[f];
        gi = 1.0 * (-t);
        gj = 1.0 * t;
#pragma unroll
        for (int q = 0; q < NC; ++q) grhs[q] = 0.0;
    }
    const double* xi = g.ccenter + (int64_t)i * D;
    double delta[D], fng[NC], gdd[NC];
    if (j >= 0) {
        const double* xj = g.ccenter + (int64_t)j * D;
#pragma unroll
        for (int k = 0; k < D; ++k) delta[k] = xj[k] - xi[k];
        double ii, ij;   // FaceInterpolationScheme::terms for the gradients
        if (a.interp == FVG_INTERP_LINEAR) {
            const double w = linear_weight<D>(g, f, i, j);
            ii = 1.0 * w;
            ij = 1.0 * (1.0 - w);
        } else if (a.flux[f] > 0.0) {
            ii = 1.0;
            ij = 0.0;
        } else {
            ii = 0.0;
            ij = 1.0;
        }
#pragma unroll
        for (int q = 0; q < NC; ++q) {
            double v = 0.0;
            v += gi * values[(int64_t)i * NC + q];
            v += gj * values[(int64_t)j * NC + q];
            v += grhs[q];
            fng[q] = v;
            const double* g0 = grads + ((int64_t)i * NC + q) * D;
            const double* g1 = grads + ((int64_t)j * NC + q) * D;
            double gf[D];
#pragma unroll
            for (int k = 0; k < D; ++k) {
                gf[k] = 0.0;
                gf[k] += g0[k] * ii;
                gf[k] += g1[k] * ij;
            }
            gdd[q] = dotv<D>(gf, delta);
        }
    } else {
        const int b = bnd[f];
        const double* xf = g.fcenter + (int64_t)f * D;
#pragma unroll
        for (int k = 0; k < D; ++k) delta[k] = xf[k] - xi[k];
#pragma unroll
        for (int q = 0; q < NC; ++q) {
            const double vi = values[(int64_t)i * NC + q];
            const double fval = bci[b] * vi + bce[(int64_t)b * NC + q];
            double v = 0.0;
            v += gi * vi;
            v += gj * fval;
            v += grhs[q];
            fng[q] = v;
            gdd[q] = dotv<D>(grads + ((int64_t)i * NC + q) * D, delta);
        }
    }
    // LimiterFrom::to_limiter_r (limiters.rs:17-44)
    const double dn = sqrt(dotv<D>(delta, delta));
    double r;
    if (NC == 1) {
        r = fabs(fng[0]) > 1e-12 ? fmax(2.0 * gdd[0] / (dn * fng[0]) - 1.0, 0.0) : 1.0;
    } else {
        double ff = 0.0, fg = 0.0;
#pragma unroll
        for (int q = 0; q < NC; ++q) ff += fng[q] * fng[q];
#pragma unroll
        for (int q = 0; q < NC; ++q) fg += fng[q] * gdd[q];
        r = sqrt(ff) > 1e-12 ? fmax(2.0 * fg / (dn * ff) - 1.0, 0.0) : 1.0;
    }
    out[f] = limiter_value(a.limiter, a.k, r);
}

template <int D, int NC>
void launch_limiters(fvg_ctx* c, const fvg_mesh* m, const double* values, const double* grads, const LimArgs& a,
                     const double* bci, const double* bce, double* out) {
    const int F = (int)m->n_faces;
    hipLaunchKernelGGL((k_limiters<D, NC>), dim3(grid_for(F)), dim3(kBlock), 0, c->stream, F, mesh_view(m),
                       m->d_bnd, m->d_tinv, values, grads, a, bci, bce, out);
}

template <int D>
void launch_limiters_nc(fvg_ctx* c, const fvg_mesh* m, int nc, const double* values, const double* grads,
                        const LimArgs& a, const double* bci, const double* bce, double* out) {
    if (nc == 1) launch_limiters<D, 1>(c, m, values, grads, a, bci, bce, out);
    else if (nc == 2) launch_limiters<D, 2>(c, m, values, grads, a, bci, bce, out);
    else launch_limiters<D, 3>(c, m, values, grads, a, bci, bce, out);
}

int check_bc(const fvg_mesh* m, int nc, const fvg_vec* bc_i, const fvg_vec* bc_e) {
    if (m->n_bnd == 0) return FVG_OK;
    if (!bc_i || !bc_e || bc_i->n != m->n_bnd || bc_e->n != m->n_bnd * nc)
        return fail(FVG_E_INVALID, "boundary closure: bc_i needs one entry per boundary face and bc_e one per "
                                   "boundary face and component");
    return FVG_OK;
}

}  // namespace
}  // namespace fvg

using namespace fvg;

extern "C" {

int fvg_compute_gradients(fvg_ctx* c, const fvg_mesh* m, int32_t nc, const fvg_vec* values, const fvg_vec* bc_i,
                          const fvg_vec* bc_e, fvg_vec* gradients) {
    if (!c || !m || !values || !gradients) return fail(FVG_E_INVALID, "null argument");
    if (nc < 1 || nc > 3) return fail(FVG_E_INVALID, "fields have 1 to 3 components");
    const int64_t n = m->n_cells, D = m->dim;
    if (values->n != n * nc) return fail(FVG_E_INVALID, "values must hold n_cells x ncomp entries");
    if (gradients->n != n * nc * D) return fail(FVG_E_INVALID, "gradients must hold n_cells x ncomp x dim entries");
    FVG_TRY(check_bc(m, nc, bc_i, bc_e));
    const double* bi = m->n_bnd ? bc_i->d : nullptr;
    const double* be = m->n_bnd ? bc_e->d : nullptr;
    if (D == 2)
        hipLaunchKernelGGL(k_gradients<2>, dim3(grid_for(n)), dim3(kBlock), 0, c->stream, (int)n, (int)nc, m->d_lptr,
                           m->d_lother, m->d_gcoef, m->d_gdiag, values->d, bi, be, gradients->d);
    else
        hipLaunchKernelGGL(k_gradients<3>, dim3(grid_for(n)), dim3(kBlock), 0, c->stream, (int)n, (int)nc, m->d_lptr,
                           m->d_lother, m->d_gcoef, m->d_gdiag, values->d, bi, be, gradients->d);
    FVG_HIP(hipGetLastError());
    return FVG_OK;
}

int fvg_compute_limiters(fvg_ctx* c, const fvg_mesh* m, int32_t nc, const fvg_vec* values, const fvg_vec* gradients,
                         const fvg_limiter_spec* spec, const fvg_vec* bc_i, const fvg_vec* bc_e, fvg_vec* limiters) {
    if (!c || !m || !values || !gradients || !spec || !limiters) return fail(FVG_E_INVALID, "null argument");
    if (nc < 1 || nc > 3) return fail(FVG_E_INVALID, "fields have 1 to 3 components");
    const int64_t n = m->n_cells, F = m->n_faces, D = m->dim;
    if (values->n != n * nc) return fail(FVG_E_INVALID, "values must hold n_cells x ncomp entries");
    if (gradients->n != n * nc * D) return fail(FVG_E_INVALID, "gradients must hold n_cells x ncomp x dim entries");
    if (limiters->n != F) return fail(FVG_E_INVALID, "limiters must hold one entry per face");
    if (spec->limiter < FVG_LIMITER_LIMITED_LINEAR || spec->limiter > FVG_LIMITER_SUPERBEE)
        return fail(FVG_E_INVALID, "unknown limiter scheme");
    if (spec->interp != FVG_INTERP_LINEAR && spec->interp != FVG_INTERP_UPWIND)
        return fail(FVG_E_INVALID, "unknown gradient interpolation scheme");
    if (spec->fngrad != FVG_FNGRAD_ORTHOGONAL && spec->fngrad != FVG_FNGRAD_CORRECTED)
        return fail(FVG_E_INVALID, "unknown face-normal gradient scheme");
    if (spec->interp == FVG_INTERP_UPWIND && (!spec->flux || spec->flux->n != F))
        return fail(FVG_E_INVALID, "Upwind needs a face flux with one entry per face");
    if (spec->fngrad == FVG_FNGRAD_CORRECTED && (!spec->fn_grads || spec->fn_grads->n != n * nc * D))
        return fail(FVG_E_INVALID, "Corrected needs a gradient field of n_cells x ncomp x dim entries");
    FVG_TRY(check_bc(m, nc, bc_i, bc_e));
    LimArgs a{spec->limiter, spec->interp, spec->fngrad, spec->k, spec->reduction,
              spec->interp == FVG_INTERP_UPWIND ? spec->flux->d : nullptr,
              spec->fngrad == FVG_FNGRAD_CORRECTED ? spec->fn_grads->d : nullptr};
    const double* bi = m->n_bnd ? bc_i->d : nullptr;
    const double* be = m->n_bnd ? bc_e->d : nullptr;
    if (D == 2) launch_limiters_nc<2>(c, m, nc, values->d, gradients->d, a, bi, be, limiters->d);
    else launch_limiters_nc<3>(c, m, nc, values->d, gradients->d, a, bi, be, limiters->d);
    FVG_HIP(hipGetLastError());
    return FVG_OK;
}

}  // extern "C"
