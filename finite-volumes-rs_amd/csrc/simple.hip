// simple.hip -- the fvm tools of pressure-based flow solvers (src/fvm/tools/hbya.rs) and the
// p-Poisson wall distance (src/fvm/tools/wall_distance.rs) on the device.
//
// compute_hbya_ainv is one thread per matrix row, walking the row in column order;
// intepolate_hbya_ainv_faces and correct_phi are one thread per face; correct_velocity
// one thread per cell.  Each repeats the reference's operation order (Vector arithmetic
// component by component, Vector::dot summed from 0.0 in component order), so results
// are bit-identical to its sequential loops.  compute_wall_distance is the reference's
// loop of assembly, IC(1) and CG driven from the host over the library's own entry
// points, with the diffusivity update and the final normalisation as kernels.
#include <cmath>
#include <string>
#include <vector>

#include "schemes.hpp"

namespace fvg {

namespace {

struct InterpArgs {
    int kind;              // FVG_INTERP_*
    const double* flux;    // Upwind, LimitedLinear
    const double* lim;     // LimitedLinear
};

// FaceInterpolationScheme::terms with its zero right-hand side: Linear (linear.rs:25-43),
// Upwind (upwind.rs:26-33), LimitedLinear (limitedlinear.rs:30-76)
template <int D>
__device__ __forceinline__ void interp_terms(const MeshView& g, int f, const InterpArgs& s, double& t0, double& t1) {
    if (s.kind == FVG_INTERP_LINEAR) {
        const double w = linear_weight<D>(g, f, g.owner[f], g.neigh[f]);
        t0 = 1.0 * w;
        t1 = 1.0 * (1.0 - w);
    } else if (s.kind == FVG_INTERP_UPWIND) {
        const bool up = s.flux[f] > 0.0;
        t0 = up ? 1.0 : 0.0;
        t1 = up ? 0.0 : 1.0;
    } else {
        const int j = g.neigh[f];
        const double fl = s.flux[f], lim = s.lim[f];
        const double w = linear_weight<D>(g, f, g.owner[f], j);
        const bool up0 = fl >= 0.0;   // t[u] = wu lim + (1 - lim), t[d] = (1 - wu) lim
        const double wu = up0 ? w : 1.0 - w;
        const double tu = wu * lim + (1.0 - lim), td = (1.0 - wu) * lim;
        if (j >= 0) {
            t0 = 1.0 * (up0 ? tu : td);
            t1 = 1.0 * (up0 ? td : tu);
        } else if (fl > 0.0) {   // outlet
            t0 = 1.0;
            t1 = 0.0;
        } else {                 // inlet
            t0 = 0.0;
            t1 = 1.0;
        }
    }
}

// hbya.rs:24-48
template <int D>
__global__ void __launch_bounds__(kBlock) k_hbya_ainv(int n, const int32_t* __restrict__ row_ptr,
                                                      const int32_t* __restrict__ col, const double* __restrict__ val,
                                                      const double* __restrict__ vol, const double* __restrict__ mrhs,
                                                      const double* __restrict__ vel, double* __restrict__ hbya,
                                                      double* __restrict__ ainv) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int rs = row_ptr[i], re = row_ptr[i + 1];
    double aii = 0.0;   // every row holds its diagonal (checked on the host)
    for (int k = rs; k < re; ++k)
        if (col[k] == i) aii = val[k];
    ainv[i] = vol[i] / aii;
    double h[D];
#pragma unroll
    for (int d = 0; d < D; ++d) h[d] = mrhs[(int64_t)i * D + d];
    for (int k = rs; k < re; ++k) {
        const int j = col[k];
        if (j != i) {
            const double a = val[k];
#pragma unroll
            for (int d = 0; d < D; ++d) h[d] -= vel[(int64_t)j * D + d] * a;
        }
    }
#pragma unroll
    for (int d = 0; d < D; ++d) hbya[(int64_t)i * D + d] = h[d] / aii;
}

// hbya.rs:69-99
template <int D>
__global__ void __launch_bounds__(kBlock) k_interp_hbya_ainv(int F, MeshView g, const int32_t* __restrict__ bnd,
                                                             const double* __restrict__ hbya,
                                                             const double* __restrict__ ainv, InterpArgs hs,
                                                             InterpArgs as, const double* __restrict__ ubci,
                                                             const double* __restrict__ ubce,
                                                             const double* __restrict__ pbci,
                                                             const double* __restrict__ pbce,
                                                             double* __restrict__ hbyan_face,
                                                             double* __restrict__ ainv_face) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    const int i = g.owner[f], j = g.neigh[f];
    double hi, hj, ai, aj;
    interp_terms<D>(g, f, hs, hi, hj);
    interp_terms<D>(g, f, as, ai, aj);
    const double* n = g.fnormal + (int64_t)f * D;
    if (j >= 0) {
        double hf[D];
#pragma unroll
        for (int d = 0; d < D; ++d)   // hi * hbya_i + hj * hbya_j + (the schemes' zero rhs)
            hf[d] = (hbya[(int64_t)i * D + d] * hi + hbya[(int64_t)j * D + d] * hj) + 0.0;
        hbyan_face[f] = dotv<D>(hf, n);
        ainv_face[f] = (ai * ainv[i] + aj * ainv[j]) + 0.0;
    } else {
        const int b = bnd[f];
        const double blhsu = ubci[b], blhsp = pbci[b], bvp = pbce[b];
        const bool p_wall = fabs(blhsp - 1.0) < 1e-10 && fabs(bvp) < 1e-10;
        const bool u_wall = fabs(blhsu) < 1e-10 && fabs(dotv<D>(ubce + (int64_t)b * D, n)) < 1e-10;
        hbyan_face[f] = p_wall && u_wall ? 0.0 : dotv<D>(hbya + (int64_t)i * D, n);
        ainv_face[f] = ainv[i];
    }
}

// hbya.rs:117-152
template <int D>
__global__ void __launch_bounds__(kBlock) k_correct_phi(int F, MeshView g, const double* __restrict__ tinv,
                                                        const int32_t* __restrict__ bnd,
                                                        const double* __restrict__ hbyan_face,
                                                        const double* __restrict__ ainv_face,
                                                        const double* __restrict__ p, int corrected,
                                                        const double* __restrict__ grads, double reduction,
                                                        const double* __restrict__ pbci,
                                                        const double* __restrict__ pbce, double* __restrict__ phi) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    const int i = g.owner[f], j = g.neigh[f];
    double gi, gj, grhs;
    if (corrected) {
        corrected_terms<D, 1>(g, f, i, j, grads, reduction, gi, gj, &grhs);
    } else {   // orthogonal.rs:26-44
        const double t = tinv[f];
        gi = 1.0 * (-t);
        gj = 1.0 * t;
        grhs = 0.0;
    }
    double pgrad_n;
    if (j >= 0) {
        pgrad_n = gi * p[i] + gj * p[j] + grhs;
    } else {
        const int b = bnd[f];
        const double p_face = pbci[b] * p[i] + pbce[b];
        pgrad_n = gi * p[i] + gj * p_face + grhs;
    }
    phi[f] = hbyan_face[f] - ainv_face[f] * pgrad_n;
}

// hbya.rs:169-176
__global__ void __launch_bounds__(kBlock) k_correct_velocity(int64_t total, int D, const double* __restrict__ hbya,
                                                             const double* __restrict__ ainv,
                                                             const double* __restrict__ pgrad,
                                                             double* __restrict__ vel) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= total) return;
    vel[k] = hbya[k] - pgrad[k] * ainv[k / D];
}

// f64::powi for a small positive exponent: binary exponentiation, the expansion the
// reference's compiler emits for a constant exponent (each product's two factors
// commute, so their order changes no bit)
__device__ __forceinline__ double powi_pos(double x, int n) {
    double r = 1.0;
    bool first = true;
    while (n) {
        if (n & 1) {
            r = first ? x : r * x;
            first = false;
        }
        n >>= 1;
        if (n) x = x * x;
    }
    return r;
}

constexpr int kWallP = 3;   // the reference's p (wall_distance.rs:30)

// wall_distance.rs:92-103: |grad_f|^(p-2) with grad_f the Linear interpolation of the cell
// gradients, the owner's on a boundary face
template <int D>
__global__ void __launch_bounds__(kBlock) k_wall_visc(int F, MeshView g, const double* __restrict__ grads,
                                                      double* __restrict__ visc) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    const int i = g.owner[f], j = g.neigh[f];
    double gf[D];
    if (j >= 0) {
        const double w = linear_weight<D>(g, f, i, j);
        const double t0 = 1.0 * w, t1 = 1.0 * (1.0 - w);
#pragma unroll
        for (int d = 0; d < D; ++d) gf[d] = grads[(int64_t)i * D + d] * t0 + grads[(int64_t)j * D + d] * t1;
    } else {
#pragma unroll
        for (int d = 0; d < D; ++d) gf[d] = grads[(int64_t)i * D + d];
    }
    visc[f] = powi_pos(sqrt(dotv<D>(gf, gf)), kWallP - 2);
}

// wall_distance.rs:108-116
template <int D>
__global__ void __launch_bounds__(kBlock) k_wall_normalize(int n, const double* __restrict__ grads,
                                                           double* __restrict__ dist) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const double* gc = grads + (int64_t)c * D;
    const double gnorm = sqrt(dotv<D>(gc, gc)), u = dist[c], pf = (double)kWallP;
    dist[c] = -powi_pos(gnorm, kWallP - 1) + pow((pf / (pf - 1.)) * u + powi_pos(gnorm, kWallP), (pf - 1.) / pf);
}

int check_closure(const fvg_mesh* m, const fvg_vec* bi, const fvg_vec* be, int64_t per_face, const char* what) {
    if (m->n_bnd == 0) return FVG_OK;
    if (!bi || !be || bi->n != m->n_bnd || be->n != m->n_bnd * per_face)
        return fail(FVG_E_INVALID, std::string(what) + " closure: one bc_i per boundary face and " +
                                       (per_face == 1 ? "one bc_e per boundary face"
                                                      : "one bc_e per boundary face and component"));
    return FVG_OK;
}

int interp_args(const fvg_mesh* m, const fvg_interp_spec* s, InterpArgs* out) {
    if (!s) return fail(FVG_E_INVALID, "null interpolation scheme");
    out->kind = s->interp;
    out->flux = out->lim = nullptr;
    if (s->interp == FVG_INTERP_LINEAR) return FVG_OK;
    if (s->interp != FVG_INTERP_UPWIND && s->interp != FVG_INTERP_LIMITED_LINEAR)
        return fail(FVG_E_INVALID, "interpolation schemes: Linear, Upwind, LimitedLinear");
    if (!s->flux || s->flux->n != m->n_faces) return fail(FVG_E_INVALID, "the scheme's flux needs one entry per face");
    out->flux = s->flux->d;
    if (s->interp == FVG_INTERP_LIMITED_LINEAR) {
        if (!s->limiters || s->limiters->n != m->n_faces)
            return fail(FVG_E_INVALID, "LimitedLinear needs one limiter per face");
        out->lim = s->limiters->d;
    }
    return FVG_OK;
}

// host check, once per pattern: does every row hold its diagonal entry?
bool rows_hold_diagonal(const fvg_csr* a) {
    for (int64_t i = 0; i < a->nrows; ++i) {
        const int32_t* b = a->h_col.data() + a->h_row_ptr[i];
        const int32_t* e = a->h_col.data() + a->h_row_ptr[i + 1];
        bool found = false;
        for (const int32_t* q = b; q < e; ++q) found |= *q == (int32_t)i;
        if (!found) return false;
    }
    return true;
}

// owns the work objects of one wall-distance computation
struct WallWork {
    fvg_ctx* c;
    fvg_vec *bci = nullptr, *bce = nullptr, *grad = nullptr, *visc = nullptr, *src = nullptr, *rhs = nullptr;
    fvg_csr* lhs = nullptr;
    fvg_prec* prec = nullptr;
    explicit WallWork(fvg_ctx* ctx) : c(ctx) {}
    ~WallWork() {
        if (prec) fvg_prec_destroy(c, prec);
        if (lhs) fvg_csr_destroy(c, lhs);
        for (fvg_vec* v : {bci, bce, grad, visc, src, rhs})
            if (v) fvg_vec_destroy(c, v);
    }
};

}  // namespace
}  // namespace fvg

using namespace fvg;

extern "C" {

int fvg_compute_hbya_ainv(fvg_ctx* c, const fvg_mesh* m, const fvg_csr* a, const fvg_vec* mrhs,
                          const fvg_vec* velocity, fvg_vec* hbya, fvg_vec* ainv) {
    if (!c || !m || !a || !mrhs || !velocity || !hbya || !ainv) return fail(FVG_E_INVALID, "null argument");
    const int64_t n = m->n_cells, D = m->dim;
    if (a->nrows != n) return fail(FVG_E_INVALID, "mlhs needs one row per cell");
    if (a->ncols > n) return fail(FVG_E_INVALID, "mlhs has columns beyond the cells");
    if (mrhs->n != n * D || velocity->n != n * D || hbya->n != n * D)
        return fail(FVG_E_INVALID, "mrhs, velocity and hbya hold n_cells x dim entries");
    if (ainv->n != n) return fail(FVG_E_INVALID, "ainv holds one entry per cell");
    if (!rows_hold_diagonal(a)) return fail(FVG_E_MISSING_ENTRY, "a row of mlhs has no diagonal entry");
    auto k = D == 2 ? k_hbya_ainv<2> : k_hbya_ainv<3>;
    hipLaunchKernelGGL(k, dim3(grid_for(n)), dim3(kBlock), 0, c->stream, (int)n, a->row_ptr, a->col, a->val,
                       m->d_cell_vol, mrhs->d, velocity->d, hbya->d, ainv->d);
    FVG_HIP(hipGetLastError());
    return FVG_OK;
}

int fvg_interpolate_hbya_ainv_faces(fvg_ctx* c, const fvg_mesh* m, const fvg_vec* hbya, const fvg_vec* ainv,
                                    const fvg_interp_spec* hbya_interp, const fvg_interp_spec* ainv_interp,
                                    const fvg_vec* u_bc_i, const fvg_vec* u_bc_e, const fvg_vec* p_bc_i,
                                    const fvg_vec* p_bc_e, fvg_vec* hbyan_face, fvg_vec* ainv_face) {
    if (!c || !m || !hbya || !ainv || !hbyan_face || !ainv_face) return fail(FVG_E_INVALID, "null argument");
    const int64_t n = m->n_cells, F = m->n_faces, D = m->dim;
    if (hbya->n != n * D) return fail(FVG_E_INVALID, "hbya holds n_cells x dim entries");
    if (ainv->n != n) return fail(FVG_E_INVALID, "ainv holds one entry per cell");
    if (hbyan_face->n != F || ainv_face->n != F) return fail(FVG_E_INVALID, "face results hold one entry per face");
    InterpArgs hs, as;
    FVG_TRY(interp_args(m, hbya_interp, &hs));
    FVG_TRY(interp_args(m, ainv_interp, &as));
    FVG_TRY(check_closure(m, u_bc_i, u_bc_e, D, "velocity"));
    FVG_TRY(check_closure(m, p_bc_i, p_bc_e, 1, "pressure"));
    const bool nb = m->n_bnd > 0;
    auto k = D == 2 ? k_interp_hbya_ainv<2> : k_interp_hbya_ainv<3>;
    hipLaunchKernelGGL(k, dim3(grid_for(F)), dim3(kBlock), 0, c->stream, (int)F, mesh_view(m), m->d_bnd, hbya->d,
                       ainv->d, hs, as, nb ? u_bc_i->d : nullptr, nb ? u_bc_e->d : nullptr, nb ? p_bc_i->d : nullptr,
                       nb ? p_bc_e->d : nullptr, hbyan_face->d, ainv_face->d);
    FVG_HIP(hipGetLastError());
    return FVG_OK;
}

int fvg_correct_phi(fvg_ctx* c, const fvg_mesh* m, const fvg_vec* hbyan_face, const fvg_vec* ainv_face,
                    const fvg_vec* pressure, int32_t fngrad, const fvg_vec* fn_grads, double reduction,
                    const fvg_vec* p_bc_i, const fvg_vec* p_bc_e, fvg_vec* phi) {
    if (!c || !m || !hbyan_face || !ainv_face || !pressure || !phi) return fail(FVG_E_INVALID, "null argument");
    const int64_t n = m->n_cells, F = m->n_faces, D = m->dim;
    if (hbyan_face->n != F || ainv_face->n != F || phi->n != F)
        return fail(FVG_E_INVALID, "face fields hold one entry per face");
    if (pressure->n != n) return fail(FVG_E_INVALID, "pressure holds one entry per cell");
    if (fngrad != FVG_FNGRAD_ORTHOGONAL && fngrad != FVG_FNGRAD_CORRECTED)
        return fail(FVG_E_INVALID, "face-normal gradient schemes: Orthogonal, Corrected");
    const bool corr = fngrad == FVG_FNGRAD_CORRECTED;
    if (corr && (!fn_grads || fn_grads->n != n * D))
        return fail(FVG_E_INVALID, "Corrected needs a gradient field of n_cells x dim entries");
    FVG_TRY(check_closure(m, p_bc_i, p_bc_e, 1, "pressure"));
    const bool nb = m->n_bnd > 0;
    auto k = D == 2 ? k_correct_phi<2> : k_correct_phi<3>;
    hipLaunchKernelGGL(k, dim3(grid_for(F)), dim3(kBlock), 0, c->stream, (int)F, mesh_view(m), m->d_tinv, m->d_bnd,
                       hbyan_face->d, ainv_face->d, pressure->d, corr ? 1 : 0, corr ? fn_grads->d : nullptr,
                       reduction, nb ? p_bc_i->d : nullptr, nb ? p_bc_e->d : nullptr, phi->d);
    FVG_HIP(hipGetLastError());
    return FVG_OK;
}

int fvg_correct_velocity(fvg_ctx* c, const fvg_mesh* m, const fvg_vec* hbya, const fvg_vec* ainv,
                         const fvg_vec* pressure_gradient, fvg_vec* velocity) {
    if (!c || !m || !hbya || !ainv || !pressure_gradient || !velocity) return fail(FVG_E_INVALID, "null argument");
    const int64_t n = m->n_cells, D = m->dim;
    if (hbya->n != n * D || pressure_gradient->n != n * D || velocity->n != n * D)
        return fail(FVG_E_INVALID, "hbya, pressure_gradient and velocity hold n_cells x dim entries");
    if (ainv->n != n) return fail(FVG_E_INVALID, "ainv holds one entry per cell");
    hipLaunchKernelGGL(k_correct_velocity, dim3(grid_for(n * D)), dim3(kBlock), 0, c->stream, n * D, (int)D, hbya->d,
                       ainv->d, pressure_gradient->d, velocity->d);
    FVG_HIP(hipGetLastError());
    return FVG_OK;
}

int fvg_compute_wall_distance(fvg_ctx* c, const fvg_mesh* m, const uint8_t* wall, int32_t max_iterations,
                              double tolerance, const fvg_solver_options* opts, fvg_vec* distance,
                              fvg_solver_info* info, int32_t* solves) {
    if (!c || !m || !distance) return fail(FVG_E_INVALID, "null argument");
    if (m->n_bnd > 0 && !wall) return fail(FVG_E_INVALID, "wall flags need one byte per boundary face");
    if (max_iterations < 0) return fail(FVG_E_INVALID, "max_iterations must not be negative");
    const int64_t n = m->n_cells, F = m->n_faces, D = m->dim, nb = m->n_bnd;
    if (distance->n != n) return fail(FVG_E_INVALID, "distance holds one entry per cell");
    fvg_solver_options o{1e-6, 1e-6, 1000, FVG_PRECISION_FP64, 0, 0.0};   // LinearSolverOptions::default()
    if (opts) o = *opts;
    if (solves) *solves = 0;
    WallWork w(c);
    // closure: a wall (0, 0), any other boundary face (1, 0) (wall_distance.rs:44-55)
    std::vector<double> h_bci((size_t)nb);
    for (int64_t b = 0; b < nb; ++b) h_bci[(size_t)b] = wall[b] ? 0.0 : 1.0;
    FVG_TRY(fvg_vec_create(c, nb, &w.bci));
    FVG_TRY(fvg_vec_create(c, nb, &w.bce));
    if (nb) FVG_TRY(fvg_vec_upload(c, w.bci, h_bci.data()));
    FVG_TRY(fvg_vec_create(c, n * D, &w.grad));   // zero, as Field::from(mesh)
    FVG_TRY(fvg_vec_create(c, F, &w.visc));
    FVG_TRY(fvg_vec_fill(c, w.visc, 1.0));
    FVG_TRY(fvg_vec_create(c, n, &w.src));
    FVG_TRY(fvg_vec_fill(c, w.src, -1.0));
    FVG_TRY(fvg_vec_create(c, n, &w.rhs));
    FVG_TRY(fvg_csr_from_mesh(c, m, &w.lhs));
    // - laplacian(Corrected(&distance_grad, 1.0), &viscosity) + source(&source)
    const fvg_term_op prog[4] = {{FVG_OP_LAPLACIAN_CORRECTED, 0, 1, 0, 1.0},
                                 {FVG_OP_NEG, 0, 0, 0, 0.0},
                                 {FVG_OP_SOURCE, 0, 0, 0, 0.0},
                                 {FVG_OP_ADD, 0, 0, 0, 0.0}};
    const fvg_vec* ff[1] = {w.visc};
    const fvg_vec* cf[2] = {w.src, w.grad};
    const fvg_vec* bi = nb ? w.bci : nullptr;
    const fvg_vec* be = nb ? w.bce : nullptr;
    const MeshView g = mesh_view(m);
    for (int32_t iter = 1; iter <= max_iterations; ++iter) {
        FVG_TRY(fvg_assemble(c, m, prog, 4, ff, 1, cf, 2, bi, be, w.lhs, w.rhs));
        // IncompleteCholesky::from_matrix(&lhs, 1): the same pattern every round, so the
        // factor is built once and refactorised after
        if (!w.prec) FVG_TRY(fvg_ic_create(c, w.lhs, 1, &w.prec));
        else FVG_TRY(fvg_prec_refactor(c, w.prec, w.lhs));
        fvg_solver_info si{};
        FVG_TRY(fvg_cg(c, w.lhs, w.prec, w.rhs, distance, &o, &si, nullptr, 0));
        if (info) *info = si;
        if (solves) *solves = iter;
        if (si.initial_residual < tolerance) break;
        FVG_TRY(fvg_compute_gradients(c, m, 1, distance, bi, be, w.grad));
        auto kv = D == 2 ? k_wall_visc<2> : k_wall_visc<3>;
        hipLaunchKernelGGL(kv, dim3(grid_for(F)), dim3(kBlock), 0, c->stream, (int)F, g, w.grad->d, w.visc->d);
        FVG_HIP(hipGetLastError());
    }
    auto kn = D == 2 ? k_wall_normalize<2> : k_wall_normalize<3>;
    hipLaunchKernelGGL(kn, dim3(grid_for(n)), dim3(kBlock), 0, c->stream, (int)n, w.grad->d, distance->d);
    FVG_HIP(hipGetLastError());
    FVG_HIP(hipStreamSynchronize(c->stream));
    return FVG_OK;
}

}  // extern "C"
