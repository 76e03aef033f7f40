// schemes.hpp -- face schemes shared by assembly and the fvm tools (device side).
//
// Each function repeats the reference scheme's arithmetic in its order; the library is
// built without contraction, so results are bit-identical to the reference's.
#pragma once

#include "mesh.hpp"

namespace fvg {

// the mesh geometry a face scheme reads, by original face / cell id
struct MeshView {
    const int32_t* owner;
    const int32_t* neigh;      // -1 on boundary faces
    const double* fcenter;     // [F * D]
    const double* fnormal;     // [F * D]
    const double* ccenter;     // [N * D]
};

inline MeshView mesh_view(const fvg_mesh* m) {
    return MeshView{m->d_owner, m->d_neigh, m->d_face_center, m->d_face_normal, m->d_cell_center};
}

// Vector::dot (vector.rs:37-43)
template <int D>
__device__ __forceinline__ double dotv(const double* a, const double* b) {
    double x = 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k) x += a[k] * b[k];
    return x;
}

// Linear interpolation weight of the owner (faceinterp/linear.rs:22-44, also
// limitedlinear.rs:46-55): n.(x_n - x_f) / max(n.(x_n - x_o), 1e-14), x_n the
// neighbour's centre or, on a boundary, the face centre.
template <int D>
__device__ __forceinline__ double linear_weight(const MeshView& g, int f, int i, int j) {
    const double* n = g.fnormal + (int64_t)f * D;
    const double* xf = g.fcenter + (int64_t)f * D;
    const double* xo = g.ccenter + (int64_t)i * D;
    const double* xn = j >= 0 ? g.ccenter + (int64_t)j * D : xf;
    double a[D], b[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
        a[k] = xn[k] - xf[k];
        b[k] = xn[k] - xo[k];
    }
    return dotv<D>(n, a) / fmax(dotv<D>(n, b), 1e-14);
}

// facengrad::Corrected::terms (facengrad/corrected.rs:34-67) for a value with NC
// components: (gi, gj, grhs[NC]).  G holds NC x D gradient entries per cell; for a
// vector value every operation acts on one row of the Matrix<NC, D> gradient.
template <int D, int NC>
__device__ __forceinline__ void corrected_terms(const MeshView& g, int f, int i, int j, const double* __restrict__ G,
                                                double reduction, double& gi, double& gj, double* grhs) {
    const double* n = g.fnormal + (int64_t)f * D;
    const double* xf = g.fcenter + (int64_t)f * D;
    const double* xi = g.ccenter + (int64_t)i * D;
    double d[D], m[D];
    double c_corr;
    if (j >= 0) {
        const double* xj = g.ccenter + (int64_t)j * D;
        double fi[D];
#pragma unroll
        for (int k = 0; k < D; ++k) {
            d[k] = xj[k] - xi[k];
            fi[k] = xf[k] - xi[k];
        }
        c_corr = 1.0 / dotv<D>(n, d);
        const double w = dotv<D>(fi, n) / dotv<D>(d, n);
#pragma unroll
        for (int k = 0; k < D; ++k) m[k] = n[k] - d[k] * c_corr;
#pragma unroll
        for (int q = 0; q < NC; ++q) {
            const double* g0 = G + ((int64_t)i * NC + q) * D;
            const double* g1 = G + ((int64_t)j * NC + q) * D;
            double gf[D];
#pragma unroll
            for (int k = 0; k < D; ++k) gf[k] = g0[k] * (1.0 - w) + g1[k] * w;
            grhs[q] = dotv<D>(gf, m) * reduction;
        }
    } else {
#pragma unroll
        for (int k = 0; k < D; ++k) d[k] = xf[k] - xi[k];
        c_corr = 1.0 / dotv<D>(n, d);
#pragma unroll
        for (int k = 0; k < D; ++k) m[k] = n[k] - d[k] * c_corr;
#pragma unroll
        for (int q = 0; q < NC; ++q) grhs[q] = dotv<D>(G + ((int64_t)i * NC + q) * D, m) * reduction;
    }
    gi = 1.0 * (-c_corr);
    gj = 1.0 * c_corr;
}

}  // namespace fvg
