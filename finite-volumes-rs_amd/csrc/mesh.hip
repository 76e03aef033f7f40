// mesh.hip -- Mesh::new + add_face/add_cell + Mesh::compute on the device
// (src/core/mesh/mod.rs:444-527, src/core/mesh/compute.rs:12-152, 267-375), the
// cell-to-cell pattern as DistributedMatrix::from_cut_sparsity leaves it
// (src/linalg/dmatrix.rs:38-50), and the per-cell lists that assembly gathers.
//
// The host only checks the two offset arrays and uploads the mesh; everything else
// runs as face- or cell-parallel kernels:
//   owners/neighbours  the first cell that lists a face owns it and the second, if the
//                      face is interior, is its neighbour (mod.rs:511-525).  Each
//                      listing has the key (cell << 32 | position in the cell's list),
//                      which orders listings as the sequential add_cell calls meet
//                      them; the owner is the smallest key (64-bit atomicMin) and the
//                      neighbour the smallest of the rest.
//   face geometry      one thread per face (compute.rs:35-45, 267-321)
//   cell geometry      one thread per cell (compute.rs:52-69, 325-375)
//   orientation        one thread per face (compute.rs:97-131), then the Orthogonal
//                      distance factor 1/|x_j - x_i| (orthogonal.rs:26-44)
//   pattern            one thread per cell writes its row candidates (itself and
//                      FaceRef::other_cell of every listed face with a neighbour,
//                      compute.rs:135-152) into a private stretch of a scratch array,
//                      sorts them and drops repeats; an exclusive scan of the row
//                      lengths gives row_ptr and a second pass compacts the columns.
//   assembly lists     per cell, the faces it owns or neighbours in ascending face
//                      id, with the CSR slot each one updates.
//   gradient coeffs    one thread per cell: the least-squares coefficients of every
//                      listing (compute.rs:154-200), with Matrix::inv restated.
//
// The arithmetic repeats the reference's operation for operation (component sums in
// index order, division by the count, the sub-triangle loop in 3D).  The library is
// compiled without floating-point contraction and the device's f64 division and
// square root are correctly rounded, so the geometry is bit-identical to the
// reference's.
#include <atomic>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>

#include "mesh.hpp"

namespace fvg {

namespace {

constexpr unsigned long long kUnsetKey = ~0ull;

// A check that fails reports the key (level << 48 | index << 3 | kind); the smallest
// key wins, so the error named is the one a sequential build meets first.
enum : int {
    kLevelCellFace = 1,    // index = cell
    kLevelTopology = 2,    // index = face
    kLevelGeometry = 3,    // index = face
    kLevelSlot = 4,        // index = cell
};
enum : int { kErrFaceRange = 0, kErrNoOwner = 1, kErrNoNeighbour = 2, kErrSelf = 3, kErrNodeRange = 4,
             kErrDegenerate = 5, kErrSlot = 6, kErrSingular = 7 };

__device__ __forceinline__ void report(unsigned long long* err, int level, int64_t index, int kind) {
    atomicMin(err, ((unsigned long long)level << 48) | ((unsigned long long)index << 3) | (unsigned long long)kind);
}

// dot product of two D-vectors summed from 0.0 in index order (vector.rs:37-43)
template <int D>
__device__ __forceinline__ double vdot(const double* a, const double* b) {
    double x = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d) x += a[d] * b[d];
    return x;
}

// ---- owners and neighbours (mod.rs:511-525) -----------------------------------------

__global__ void __launch_bounds__(kBlock) k_mesh_owner(int64_t n_cells, int64_t n_faces,
                                                       const int64_t* __restrict__ cptr,
                                                       const int32_t* __restrict__ cfaces,
                                                       unsigned long long* __restrict__ own_key,
                                                       unsigned long long* __restrict__ err) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n_cells) return;
    const int64_t s = cptr[c], e = cptr[c + 1];
    for (int64_t k = s; k < e; ++k) {
        const int32_t f = cfaces[k];
        if (f < 0 || f >= n_faces) {
            report(err, kLevelCellFace, c, kErrFaceRange);
            continue;
        }
        atomicMin(&own_key[f], ((unsigned long long)c << 32) | (unsigned long long)(k - s));
    }
}

__global__ void __launch_bounds__(kBlock) k_mesh_neighbour(int64_t n_cells, int64_t n_faces,
                                                           const int64_t* __restrict__ cptr,
                                                           const int32_t* __restrict__ cfaces,
                                                           const int32_t* __restrict__ patch,
                                                           const unsigned long long* __restrict__ own_key,
                                                           unsigned long long* __restrict__ nb_key) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n_cells) return;
    const int64_t s = cptr[c], e = cptr[c + 1];
    for (int64_t k = s; k < e; ++k) {
        const int32_t f = cfaces[k];
        if (f < 0 || f >= n_faces || patch[f] >= 0) continue;   // boundary faces keep no neighbour
        const unsigned long long key = ((unsigned long long)c << 32) | (unsigned long long)(k - s);
        if (key != own_key[f]) atomicMin(&nb_key[f], key);
    }
}

// ---- face data (compute.rs:35-45, 267-321) -----------------------------------------

// Geometry of one face: area, centre and the last sub-triangle's normal divided by
// the area (compute.rs:305), before orientation.  False for a degenerate face.
template <int D>
__device__ bool face_geometry(const double* __restrict__ nodes, const int32_t* __restrict__ fn, int64_t len,
                              double& area_out, double* center, double* normal) {
    if (D == 2) {
        if (len != 2) return false;
        const double* n0 = nodes + 2 * (int64_t)fn[0];
        const double* n1 = nodes + 2 * (int64_t)fn[1];
        const double dn[2] = {n1[0] - n0[0], n1[1] - n0[1]};
        const double dsn[2] = {-dn[1], dn[0]};   // Normal for N == 2 (vector.rs:305-313)
        const double area = sqrt(vdot<2>(dsn, dsn));
        area_out = area;
        normal[0] = dsn[0] / area;
        normal[1] = dsn[1] / area;
        center[0] = (n0[0] + n1[0]) * 0.5;
        center[1] = (n0[1] + n1[1]) * 0.5;
        return true;
    } else {
        double ave[3] = {0.0, 0.0, 0.0};   // node average (compute.rs:36-43)
        for (int64_t k = 0; k < len; ++k)
            for (int d = 0; d < 3; ++d) ave[d] += nodes[3 * (int64_t)fn[k] + d];
        for (int d = 0; d < 3; ++d) ave[d] /= (double)len;
        double c[3] = {0.0, 0.0, 0.0}, nrm[3] = {0.0, 0.0, 0.0}, area = 0.0;
        for (int64_t i = 0; i < len; ++i) {
            const double* n0 = nodes + 3 * (int64_t)fn[i];
            const double* n1 = nodes + 3 * (int64_t)fn[i == len - 1 ? 0 : i + 1];
            double a[3], b[3];
            for (int d = 0; d < 3; ++d) {
                a[d] = n0[d] - ave[d];
                b[d] = n1[d] - ave[d];
            }
            double dsn[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
            for (int d = 0; d < 3; ++d) dsn[d] *= 0.5;   // Normal for N == 3 scales by 0.5 (vector.rs:314-324)
            const double area_i = sqrt(vdot<3>(dsn, dsn));
            area += area_i;
            for (int d = 0; d < 3; ++d) nrm[d] = dsn[d];
            for (int d = 0; d < 3; ++d) c[d] += (((n0[d] + n1[d]) + ave[d]) * 1.0 / 3.0) * area_i;
        }
        if (!(area > 2.220446049250313e-16)) return false;   // compute.rs:308
        for (int d = 0; d < 3; ++d) {
            nrm[d] /= area;
            c[d] /= area;
        }
        area_out = area;
        for (int d = 0; d < 3; ++d) {
            center[d] = c[d];
            normal[d] = nrm[d];
        }
        return true;
    }
}

template <int D>
__global__ void __launch_bounds__(kBlock) k_mesh_faces(int64_t n_faces, int64_t n_nodes,
                                                       const double* __restrict__ nodes,
                                                       const int64_t* __restrict__ fptr,
                                                       const int32_t* __restrict__ fnodes,
                                                       const int32_t* __restrict__ patch,
                                                       const unsigned long long* __restrict__ own_key,
                                                       const unsigned long long* __restrict__ nb_key,
                                                       int32_t* __restrict__ owner, int32_t* __restrict__ neigh,
                                                       int32_t* __restrict__ is_bnd, double* __restrict__ area,
                                                       double* __restrict__ center, double* __restrict__ normal,
                                                       unsigned long long* __restrict__ err) {
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= n_faces) return;
    // compute.rs:15-31: every face has an owner; an interior face also a neighbour
    const bool bnd = patch[f] >= 0;
    const unsigned long long ok = own_key[f];
    int32_t o = -1, j = -1;
    if (ok == kUnsetKey) {
        report(err, kLevelTopology, f, kErrNoOwner);
    } else {
        o = (int32_t)(ok >> 32);
        if (!bnd) {
            const unsigned long long nk = nb_key[f];
            if (nk == kUnsetKey) report(err, kLevelTopology, f, kErrNoNeighbour);
            else if ((j = (int32_t)(nk >> 32)) == o) report(err, kLevelTopology, f, kErrSelf);
        }
    }
    owner[f] = o;
    neigh[f] = j;
    is_bnd[f] = bnd ? 1 : 0;
    const int64_t s = fptr[f], len = fptr[f + 1] - s;
    for (int64_t k = 0; k < len; ++k) {
        const int32_t v = fnodes[s + k];
        if (v < 0 || v >= n_nodes) {
            report(err, kLevelGeometry, f, kErrNodeRange);
            return;
        }
    }
    double a = 0.0, c[D], nrm[D];
    if (!face_geometry<D>(nodes, fnodes + s, len, a, c, nrm)) {
        report(err, kLevelGeometry, f, kErrDegenerate);
        return;
    }
    area[f] = a;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        center[f * D + d] = c[d];
        normal[f * D + d] = nrm[d];
    }
}

// ---- cell data (compute.rs:52-69, 325-375), pattern rows and assembly lists -----------

// insertion sort of a short private list, then repeats dropped; returns the new length
__device__ int sort_unique(int32_t* a, int m) {
    for (int i = 1; i < m; ++i) {
        const int32_t v = a[i];
        int k = i - 1;
        while (k >= 0 && a[k] > v) {
            a[k + 1] = a[k];
            --k;
        }
        a[k + 1] = v;
    }
    int u = 0;
    for (int i = 0; i < m; ++i)
        if (u == 0 || a[i] != a[u - 1]) a[u++] = a[i];
    return u;
}

template <int D>
__global__ void __launch_bounds__(kBlock) k_mesh_cells(int64_t n_cells, const int64_t* __restrict__ cptr,
                                                       const int32_t* __restrict__ cfaces,
                                                       const double* __restrict__ farea,
                                                       const double* __restrict__ fcenter,
                                                       const double* __restrict__ fnormal,
                                                       const int32_t* __restrict__ owner,
                                                       const int32_t* __restrict__ neigh,
                                                       double* __restrict__ vol, double* __restrict__ ccenter,
                                                       int32_t* __restrict__ cand, int32_t* __restrict__ rowcnt,
                                                       int32_t* __restrict__ entf, int32_t* __restrict__ entcnt) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n_cells) return;
    const int64_t s = cptr[c], e = cptr[c + 1];
    // node-free centre estimate: the average of the face centres
    double ave[D];
#pragma unroll
    for (int d = 0; d < D; ++d) ave[d] = 0.0;
    for (int64_t k = s; k < e; ++k) {
        const int64_t f = cfaces[k];
#pragma unroll
        for (int d = 0; d < D; ++d) ave[d] += fcenter[f * D + d];
    }
#pragma unroll
    for (int d = 0; d < D; ++d) ave[d] /= (double)(e - s);
    // pyramid decomposition over the faces
    double center[D], volume = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d) center[d] = 0.0;
    for (int64_t k = s; k < e; ++k) {
        const int64_t f = cfaces[k];
        double fc[D], fnrm[D], dfc[D];
#pragma unroll
        for (int d = 0; d < D; ++d) {
            fc[d] = fcenter[f * D + d];
            fnrm[d] = fnormal[f * D + d];
            dfc[d] = fc[d] - ave[d];
        }
        const double height = fabs(vdot<D>(dfc, fnrm));
        double volume_i, ci[D];
        if (D == 2) {
            volume_i = height * farea[f] / 2.0;
#pragma unroll
            for (int d = 0; d < D; ++d) ci[d] = (fc[d] * 2.0 + ave[d]) / 3.0;
        } else {
            volume_i = height * farea[f] / 3.0;
#pragma unroll
            for (int d = 0; d < D; ++d) ci[d] = (fc[d] * 3.0 + ave[d]) / 4.0;
        }
#pragma unroll
        for (int d = 0; d < D; ++d) center[d] += ci[d] * volume_i;
        volume += volume_i;
    }
    vol[c] = volume;
#pragma unroll
    for (int d = 0; d < D; ++d) ccenter[c * D + d] = center[d] / volume;

    // row c of cell_to_cell: the cell, then FaceRef::other_cell (mod.rs:317-329) of
    // each listed face with a neighbour; from_cut_sparsity keeps it sorted
    int32_t* row = cand + s + c;
    int m = 0;
    row[m++] = (int32_t)c;
    for (int64_t k = s; k < e; ++k) {
        const int32_t f = cfaces[k], j = neigh[f];
        if (j >= 0) row[m++] = j == (int32_t)c ? owner[f] : j;
    }
    rowcnt[c] = sort_unique(row, m);

    // the faces this cell updates in assembly (assembly.rs:35-76), in face order
    int32_t* ent = entf + s;
    int q = 0;
    for (int64_t k = s; k < e; ++k) {
        const int32_t f = cfaces[k];
        if (owner[f] == (int32_t)c || neigh[f] == (int32_t)c) ent[q++] = f;
    }
    entcnt[c] = sort_unique(ent, q);
}

// normal orientation owner -> neighbour / outward (compute.rs:97-131), then the
// Orthogonal scheme's 1/|x_j - x_i|, x_j the face centre on the boundary
template <int D>
__global__ void __launch_bounds__(kBlock) k_mesh_orient(int64_t n_faces, const int32_t* __restrict__ owner,
                                                        const int32_t* __restrict__ neigh,
                                                        const double* __restrict__ fcenter,
                                                        const double* __restrict__ ccenter,
                                                        double* __restrict__ normal, double* __restrict__ tinv,
                                                        const double* __restrict__ area,
                                                        double2* __restrict__ tarea) {
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= n_faces) return;
    const int64_t i = owner[f], j = neigh[f];
    double fc[D], ci[D], nrm[D], d[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
        fc[k] = fcenter[f * D + k];
        ci[k] = ccenter[i * D + k];
        nrm[k] = normal[f * D + k];
        d[k] = fc[k] - ci[k];
    }
    if (vdot<D>(nrm, d) < 0.0) {
#pragma unroll
        for (int k = 0; k < D; ++k) normal[f * D + k] = nrm[k] * -1.0;
    }
#pragma unroll
    for (int k = 0; k < D; ++k) d[k] = (j >= 0 ? ccenter[j * D + k] : fc[k]) - ci[k];
    const double dx = sqrt(vdot<D>(d, d));
    tinv[f] = 1.0 / dx;
    tarea[f] = make_double2(1.0 / dx, area[f]);
}

// compacts the sorted rows into the CSR pattern and turns each cell's face list into
// assembly entries {face | role << 30, slot or boundary ordinal}
__global__ void __launch_bounds__(kBlock) k_mesh_rows(int64_t n_cells, const int64_t* __restrict__ cptr,
                                                      const int32_t* __restrict__ cand,
                                                      const int32_t* __restrict__ row_ptr,
                                                      int32_t* __restrict__ col, int32_t* __restrict__ diagpos,
                                                      const int32_t* __restrict__ entf,
                                                      const int32_t* __restrict__ ent_ptr,
                                                      const int32_t* __restrict__ owner,
                                                      const int32_t* __restrict__ neigh,
                                                      const int32_t* __restrict__ bnd, int2* __restrict__ ent,
                                                      unsigned long long* __restrict__ err) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n_cells) return;
    const int64_t s = cptr[c];
    const int32_t* row = cand + s + c;
    const int32_t rb = row_ptr[c], m = row_ptr[c + 1] - rb;
    int32_t dpos = -1;
    for (int t = 0; t < m; ++t) {
        col[rb + t] = row[t];
        if (row[t] == (int32_t)c) dpos = rb + t;
    }
    diagpos[c] = dpos;
    const int32_t eb = ent_ptr[c], q = ent_ptr[c + 1] - eb;
    for (int t = 0; t < q; ++t) {
        const int32_t f = entf[s + t], i = owner[f], j = neigh[f];
        int32_t role, aux;
        if (j < 0) {
            role = kRoleBoundary;
            aux = bnd[f];
        } else {
            role = i == (int32_t)c ? kRoleOwner : kRoleNeighbour;
            const int32_t other = i == (int32_t)c ? j : i;
            int lo = 0, hi = m;   // lower bound of `other` in the sorted row
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (row[mid] < other) lo = mid + 1;
                else hi = mid;
            }
            if (lo < m && row[lo] == other) {
                aux = rb + lo;
            } else {
                report(err, kLevelSlot, c, kErrSlot);
                aux = dpos;
            }
        }
        ent[eb + t] = make_int2(f | (role << kRoleShift), aux);
    }
}

// ---- least-squares gradient coefficients (compute.rs:154-200) -----------------------

// Matrix::inv (matrix.rs:320-383) on rows a[D][D]: Gauss-Jordan without pivoting unless
// a pivot is below 1e-14, then a swap with row `maxi`, maxi being the largest entry of
// row k right of the diagonal (the reference's search as written).  False where the
// reference returns Err(SingularMatrix).
template <int D>
__device__ bool invert(double (&a)[D][D], double (&out)[D][D]) {
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) out[i][j] = i == j ? 1.0 : 0.0;
#pragma unroll
    for (int k = 0; k < D - 1; ++k) {
        double piv = a[k][k];
        if (fabs(piv) < 1e-14) {
            int maxi = k;
            double maxp = fabs(piv);
            for (int j = k + 1; j < D; ++j) {
                const double pj = fabs(a[k][j]);
                if (pj > maxp) {
                    maxi = j;
                    maxp = pj;
                }
            }
            if (maxi == k) return false;
#pragma unroll
            for (int j = 0; j < D; ++j) {   // row swap, written without dynamic indexing
                double t0 = a[k][j], t1 = out[k][j];
#pragma unroll
                for (int m = k + 1; m < D; ++m)
                    if (m == maxi) {
                        a[k][j] = a[m][j];
                        a[m][j] = t0;
                        out[k][j] = out[m][j];
                        out[m][j] = t1;
                    }
            }
        }
        piv = a[k][k];
#pragma unroll
        for (int i = k + 1; i < D; ++i) {
            const double f = -a[i][k] / piv;
#pragma unroll
            for (int j = 0; j < D; ++j) {
                a[i][j] = a[i][j] + a[k][j] * f;
                out[i][j] = out[i][j] + out[k][j] * f;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < D; ++k) {
        const double piv = a[k][k];
        if (fabs(piv) < 1e-14) return false;
#pragma unroll
        for (int j = 0; j < D; ++j) {
            a[k][j] = a[k][j] / piv;
            out[k][j] = out[k][j] / piv;
        }
    }
#pragma unroll
    for (int k = D - 1; k >= 1; --k) {
        const double piv = a[k][k];
#pragma unroll
        for (int i = 0; i < k; ++i) {
            const double f = -a[i][k] / piv;
#pragma unroll
            for (int j = 0; j < D; ++j) {
                a[i][j] = a[i][j] + a[k][j] * f;
                out[i][j] = out[i][j] + out[k][j] * f;
            }
        }
    }
    return true;
}

// One thread per cell: dgc_k to the other cell's centre (or the face centre on a
// boundary) for every listing k, G = sum dgc dgc^T, coef_k = G^-1 dgc_k, and
// diag = -sum coef_k.  Also records, per listing, the other cell or the boundary
// ordinal that compute_gradients reads.
template <int D>
__global__ void __launch_bounds__(kBlock) k_mesh_grad(int64_t n_cells, const int64_t* __restrict__ cptr,
                                                      const int32_t* __restrict__ cfaces,
                                                      const int32_t* __restrict__ owner,
                                                      const int32_t* __restrict__ neigh,
                                                      const int32_t* __restrict__ bnd,
                                                      const double* __restrict__ fcenter,
                                                      const double* __restrict__ ccenter,
                                                      int32_t* __restrict__ lptr, int32_t* __restrict__ lother,
                                                      double* __restrict__ coef, double* __restrict__ diag,
                                                      unsigned long long* __restrict__ err) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n_cells) return;
    const int64_t s = cptr[c], e = cptr[c + 1];
    lptr[c] = (int32_t)s;
    if (c == n_cells - 1) lptr[n_cells] = (int32_t)e;
    double xc[D];
#pragma unroll
    for (int d = 0; d < D; ++d) xc[d] = ccenter[c * D + d];
    auto dgc_of = [&](int64_t k, double (&v)[D]) {
        const int32_t f = cfaces[k], j = neigh[f];
        int32_t code;
        const double* x;
        if (j < 0) {
            code = -bnd[f] - 1;
            x = fcenter + (int64_t)f * D;
        } else {
            const int32_t o = j == (int32_t)c ? owner[f] : j;   // FaceRef::other_cell
            code = o;
            x = ccenter + (int64_t)o * D;
        }
#pragma unroll
        for (int d = 0; d < D; ++d) v[d] = x[d] - xc[d];
        return code;
    };
    double g[D][D];
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) g[i][j] = 0.0;
    for (int64_t k = s; k < e; ++k) {
        double v[D];
        lother[k] = dgc_of(k, v);
#pragma unroll
        for (int i = 0; i < D; ++i)
#pragma unroll
            for (int j = 0; j < D; ++j) g[i][j] += v[i] * v[j];
    }
    double ginv[D][D];
    if (!invert<D>(g, ginv)) {
        report(err, kLevelSlot + 1, c, kErrSingular);
        return;
    }
    double dg[D];
#pragma unroll
    for (int d = 0; d < D; ++d) dg[d] = 0.0;
    for (int64_t k = s; k < e; ++k) {
        double v[D];
        (void)dgc_of(k, v);
#pragma unroll
        for (int i = 0; i < D; ++i) {
            double x = 0.0;   // Matrix * Vector: one dot per row (matrix.rs:221-230)
#pragma unroll
            for (int j = 0; j < D; ++j) x += ginv[i][j] * v[j];
            coef[k * D + i] = x;
            dg[i] -= x;
        }
    }
#pragma unroll
    for (int d = 0; d < D; ++d) diag[c * D + d] = dg[d];
}

// ---- exclusive scan of int32 counts (three passes over tiles of 1024) ---------------

constexpr int kScanThreads = 256, kScanItems = 4, kScanTile = kScanThreads * kScanItems;

// inclusive scan over the NT threads of a workgroup; lds holds NT/64 entries
template <int NT>
__device__ __forceinline__ int64_t block_incl_scan(int64_t v, int64_t* lds) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int64_t t = __shfl_up(v, off, 64);
        if (lane >= off) v += t;
    }
    if (lane == 63) lds[w] = v;
    __syncthreads();
    int64_t base = 0;
    for (int k = 0; k < w; ++k) base += lds[k];
    __syncthreads();
    return v + base;
}

__global__ void __launch_bounds__(kScanThreads) k_scan_tiles(int64_t n, const int32_t* __restrict__ cnt,
                                                             int64_t* __restrict__ tile_sum) {
    __shared__ int64_t lds[kScanThreads / 64];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    int64_t v = 0;
#pragma unroll
    for (int q = 0; q < kScanItems; ++q)
        if (base + q < n) v += cnt[base + q];
    v = block_incl_scan<kScanThreads>(v, lds);
    if (threadIdx.x == kScanThreads - 1) tile_sum[blockIdx.x] = v;
}

// one workgroup: tile sums -> exclusive tile offsets, grand total into *total
__global__ void __launch_bounds__(1024) k_scan_tile_sums(int64_t n_tiles, int64_t* __restrict__ tile_sum,
                                                         int64_t* __restrict__ total) {
    __shared__ int64_t lds[1024 / 64];
    __shared__ int64_t chunk;
    int64_t carry = 0;
    for (int64_t b = 0; b < n_tiles; b += 1024) {
        const int64_t i = b + threadIdx.x;
        const int64_t v = i < n_tiles ? tile_sum[i] : 0;
        const int64_t incl = block_incl_scan<1024>(v, lds);
        if (i < n_tiles) tile_sum[i] = carry + incl - v;
        if (threadIdx.x == 1023) chunk = incl;
        __syncthreads();
        carry += chunk;
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

// out[i] = sum of cnt[0..i), out[n] = the total
__global__ void __launch_bounds__(kScanThreads) k_scan_apply(int64_t n, const int32_t* __restrict__ cnt,
                                                             const int64_t* __restrict__ tile_off,
                                                             int32_t* __restrict__ out) {
    __shared__ int64_t lds[kScanThreads / 64];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    int32_t v[kScanItems];
    int64_t sum = 0;
#pragma unroll
    for (int q = 0; q < kScanItems; ++q) {
        v[q] = base + q < n ? cnt[base + q] : 0;
        sum += v[q];
    }
    const int64_t incl = block_incl_scan<kScanThreads>(sum, lds);
    int64_t run = tile_off[blockIdx.x] + incl - sum;
#pragma unroll
    for (int q = 0; q < kScanItems; ++q) {
        const int64_t i = base + q;
        if (i < n) {
            out[i] = (int32_t)run;
            run += v[q];
            if (i == n - 1) out[n] = (int32_t)run;
        }
    }
}

int64_t scan_tiles(int64_t n) { return (n + kScanTile - 1) / kScanTile; }

// first two passes: tile offsets and the total (left on the device)
void scan_prepare(fvg_ctx* c, int64_t n, const int32_t* cnt, int64_t* tile_off, int64_t* total) {
    const int64_t nt = scan_tiles(n);
    hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)nt), dim3(kScanThreads), 0, c->stream, n, cnt, tile_off);
    hipLaunchKernelGGL(k_scan_tile_sums, dim3(1), dim3(1024), 0, c->stream, nt, tile_off, total);
}

void scan_finish(fvg_ctx* c, int64_t n, const int32_t* cnt, const int64_t* tile_off, int32_t* out) {
    hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)scan_tiles(n)), dim3(kScanThreads), 0, c->stream, n, cnt,
                       tile_off, out);
}

// Device scratch of one build, released (after the stream drains) when it goes out
// of scope.
struct Scratch {
    fvg_ctx* c;
    std::vector<std::pair<void*, size_t>> held;
    explicit Scratch(fvg_ctx* ctx) : c(ctx) {}
    template <class T>
    int get(int64_t n, T** p) {
        const size_t bytes = (size_t)(n > 0 ? n : 1) * sizeof(T);
        FVG_TRY(dev_alloc(c, bytes, (void**)p));
        held.emplace_back((void*)*p, bytes);
        return FVG_OK;
    }
    ~Scratch() {
        (void)hipStreamSynchronize(c->stream);
        for (auto& h : held) dev_free(c, h.first, h.second);
    }
};

std::string mesh_error(unsigned long long key) {
    const int level = (int)(key >> 48), kind = (int)(key & 7);
    const long long idx = (long long)((key >> 3) & ((1ull << 45) - 1));
    switch (kind) {
    case kErrFaceRange: return "cell " + std::to_string(idx) + " names a face out of range";
    case kErrNoOwner: return "face " + std::to_string(idx) + " has no owner cell";
    case kErrNoNeighbour: return "interior face " + std::to_string(idx) + " has no neighbour cell";
    case kErrSelf: return "interior face " + std::to_string(idx) + " has one cell as both owner and neighbour";
    case kErrNodeRange: return "face " + std::to_string(idx) + " names a node out of range";
    case kErrDegenerate: return "degenerate face " + std::to_string(idx);
    case kErrSlot: return "cell " + std::to_string(idx) + ": face slot missing from the CSR pattern";
    case kErrSingular: return "cell " + std::to_string(idx) + ": singular gradient matrix (compute.rs:180)";
    default: return "mesh check " + std::to_string(level) + " failed";
    }
}

std::atomic<uint64_t> g_pattern_counter{1};

// Meshes alive in this process, with their pattern ids: a matrix made from a mesh reads
// the cell centres from it later only while that mesh (not another at the same address)
// is alive.  A reader holds the lock for its whole copy, so a destroy waits for it.
std::mutex g_live_mu;
std::unordered_map<const fvg_mesh*, uint64_t> g_live;


void free_mesh(fvg_ctx* c, fvg_mesh* m) {
    const int64_t F = m->n_faces, N = m->n_cells, D = m->dim;
    dev_free_n(c, m->d_area, F);
    dev_free_n(c, m->d_face_center, F * D);
    dev_free_n(c, m->d_face_normal, F * D);
    dev_free_n(c, m->d_tinv, F);
    dev_free_n(c, m->d_tarea, F);
    dev_free_n(c, m->d_cell_vol, N);
    dev_free_n(c, m->d_cell_center, N * D);
    dev_free_n(c, m->d_owner, F);
    dev_free_n(c, m->d_neigh, F);
    dev_free_n(c, m->d_bnd, F + 1);
    dev_free_n(c, m->d_row_ptr, N + 1);
    dev_free_n(c, m->d_col, m->nnz);
    dev_free_n(c, m->d_diagpos, N);
    dev_free_n(c, m->d_ent_ptr, N + 1);
    dev_free_n(c, m->d_ent, m->n_ent);
    dev_free_n(c, m->d_lptr, N + 1);
    dev_free_n(c, m->d_lother, m->n_list);
    dev_free_n(c, m->d_gcoef, m->n_list * D);
    dev_free_n(c, m->d_gdiag, N * D);
}

// offsets must start at 0 and never decrease (then every range lies inside the list)
bool offsets_valid(const int64_t* p, int64_t n) {
    if (p[0] != 0) return false;
    for (int64_t i = 0; i < n; ++i)
        if (p[i + 1] < p[i]) return false;
    return true;
}

template <int D>
int build_mesh(fvg_ctx* c, fvg_mesh* m, int64_t n_nodes, const double* nodes, const int64_t* face_ptr,
               const int32_t* face_nodes, const int32_t* face_patch, const int64_t* cell_ptr,
               const int32_t* cell_faces) {
    const int64_t F = m->n_faces, N = m->n_cells;
    const int64_t Lf = face_ptr[F], Lc = cell_ptr[N];
    if (Lc >= INT32_MAX - N) return fail(FVG_E_INVALID, "mesh too large for int32 cell-face lists");
    int64_t maxf = 0;
    for (int64_t i = 0; i < N; ++i) maxf = std::max(maxf, cell_ptr[i + 1] - cell_ptr[i]);
    m->max_cell_faces = (int32_t)std::min<int64_t>(maxf, INT32_MAX);

    Scratch t(c);
    double* d_nodes;
    int64_t *d_fptr, *d_cptr, *d_meta, *d_toff_b, *d_toff_r, *d_toff_e;
    int32_t *d_fnodes, *d_patch, *d_cfaces, *d_isb, *d_cand, *d_rowcnt, *d_entf, *d_entcnt;
    unsigned long long *d_own, *d_nb;
    FVG_TRY(t.get(n_nodes * D, &d_nodes));
    FVG_TRY(t.get(F + 1, &d_fptr));
    FVG_TRY(t.get(Lf, &d_fnodes));
    FVG_TRY(t.get(F, &d_patch));
    FVG_TRY(t.get(N + 1, &d_cptr));
    FVG_TRY(t.get(Lc, &d_cfaces));
    FVG_TRY(t.get(2 * F, &d_own));
    d_nb = d_own + F;
    FVG_TRY(t.get(8, &d_meta));   // [0] error key, [1] boundary faces, [2] nnz, [3] entries
    FVG_TRY(upload(c, d_nodes, nodes, n_nodes * D));
    FVG_TRY(upload(c, d_fptr, face_ptr, F + 1));
    FVG_TRY(upload(c, d_fnodes, face_nodes, Lf));
    FVG_TRY(upload(c, d_patch, face_patch, F));
    FVG_TRY(upload(c, d_cptr, cell_ptr, N + 1));
    FVG_TRY(upload(c, d_cfaces, cell_faces, Lc));
    FVG_HIP(hipMemsetAsync(d_own, 0xff, sizeof(unsigned long long) * (size_t)(2 * F), c->stream));
    FVG_HIP(hipMemsetAsync(d_meta, 0xff, sizeof(int64_t), c->stream));
    auto* err = reinterpret_cast<unsigned long long*>(d_meta);

    // topology and face data
    FVG_TRY(dev_alloc_n(c, F, &m->d_owner));
    FVG_TRY(dev_alloc_n(c, F, &m->d_neigh));
    FVG_TRY(dev_alloc_n(c, F, &m->d_area));
    FVG_TRY(dev_alloc_n(c, F * D, &m->d_face_center));
    FVG_TRY(dev_alloc_n(c, F * D, &m->d_face_normal));
    FVG_TRY(dev_alloc_n(c, F, &m->d_tinv));
    FVG_TRY(dev_alloc_n(c, F, &m->d_tarea));
    FVG_TRY(t.get(F, &d_isb));
    FVG_TRY(dev_alloc_n(c, F + 1, &m->d_bnd));
    int32_t* d_bnd = m->d_bnd;
    FVG_TRY(t.get(scan_tiles(F), &d_toff_b));
    hipLaunchKernelGGL(k_mesh_owner, dim3(grid_for(N)), dim3(kBlock), 0, c->stream, N, F, d_cptr, d_cfaces, d_own,
                       err);
    hipLaunchKernelGGL(k_mesh_neighbour, dim3(grid_for(N)), dim3(kBlock), 0, c->stream, N, F, d_cptr, d_cfaces,
                       d_patch, d_own, d_nb);
    hipLaunchKernelGGL(k_mesh_faces<D>, dim3(grid_for(F)), dim3(kBlock), 0, c->stream, F, n_nodes, d_nodes, d_fptr,
                       d_fnodes, d_patch, d_own, d_nb, m->d_owner, m->d_neigh, d_isb, m->d_area, m->d_face_center,
                       m->d_face_normal, err);
    scan_prepare(c, F, d_isb, d_toff_b, d_meta + 1);
    FVG_HIP(hipGetLastError());
    int64_t meta[4];
    FVG_TRY(download(c, meta, d_meta, 2));
    if ((unsigned long long)meta[0] != kUnsetKey) return fail(FVG_E_MESH, mesh_error((unsigned long long)meta[0]));
    m->n_bnd = meta[1];
    scan_finish(c, F, d_isb, d_toff_b, d_bnd);

    // cell data, orientation, pattern rows and assembly lists
    FVG_TRY(dev_alloc_n(c, N, &m->d_cell_vol));
    FVG_TRY(dev_alloc_n(c, N * D, &m->d_cell_center));
    FVG_TRY(t.get(Lc + N, &d_cand));
    FVG_TRY(t.get(N, &d_rowcnt));
    FVG_TRY(t.get(Lc, &d_entf));
    FVG_TRY(t.get(N, &d_entcnt));
    FVG_TRY(t.get(scan_tiles(N), &d_toff_r));
    FVG_TRY(t.get(scan_tiles(N), &d_toff_e));
    hipLaunchKernelGGL(k_mesh_cells<D>, dim3(grid_for(N)), dim3(kBlock), 0, c->stream, N, d_cptr, d_cfaces,
                       m->d_area, m->d_face_center, m->d_face_normal, m->d_owner, m->d_neigh, m->d_cell_vol,
                       m->d_cell_center, d_cand, d_rowcnt, d_entf, d_entcnt);
    hipLaunchKernelGGL(k_mesh_orient<D>, dim3(grid_for(F)), dim3(kBlock), 0, c->stream, F, m->d_owner, m->d_neigh,
                       m->d_face_center, m->d_cell_center, m->d_face_normal, m->d_tinv, m->d_area, m->d_tarea);
    scan_prepare(c, N, d_rowcnt, d_toff_r, d_meta + 2);
    scan_prepare(c, N, d_entcnt, d_toff_e, d_meta + 3);
    FVG_HIP(hipGetLastError());
    FVG_TRY(download(c, meta, d_meta, 4));
    if (meta[2] >= INT32_MAX) return fail(FVG_E_INVALID, "cell-to-cell pattern too large for int32 indices");
    m->nnz = meta[2];
    m->n_ent = meta[3];
    // each off-diagonal slot is met by at least one interior entry, so equal counts mean
    // exactly one each
    m->unique_slots = m->n_ent - m->n_bnd == m->nnz - N;
    FVG_TRY(dev_alloc_n(c, N + 1, &m->d_row_ptr));
    FVG_TRY(dev_alloc_n(c, m->nnz, &m->d_col));
    FVG_TRY(dev_alloc_n(c, N, &m->d_diagpos));
    FVG_TRY(dev_alloc_n(c, N + 1, &m->d_ent_ptr));
    FVG_TRY(dev_alloc_n(c, m->n_ent, &m->d_ent));
    scan_finish(c, N, d_rowcnt, d_toff_r, m->d_row_ptr);
    scan_finish(c, N, d_entcnt, d_toff_e, m->d_ent_ptr);
    hipLaunchKernelGGL(k_mesh_rows, dim3(grid_for(N)), dim3(kBlock), 0, c->stream, N, d_cptr, d_cand,
                       m->d_row_ptr, m->d_col, m->d_diagpos, d_entf, m->d_ent_ptr, m->d_owner, m->d_neigh, d_bnd,
                       m->d_ent, err);
    m->n_list = Lc;
    FVG_TRY(dev_alloc_n(c, N + 1, &m->d_lptr));
    FVG_TRY(dev_alloc_n(c, Lc, &m->d_lother));
    FVG_TRY(dev_alloc_n(c, Lc * D, &m->d_gcoef));
    FVG_TRY(dev_alloc_n(c, N * D, &m->d_gdiag));
    hipLaunchKernelGGL(k_mesh_grad<D>, dim3(grid_for(N)), dim3(kBlock), 0, c->stream, N, d_cptr, d_cfaces,
                       m->d_owner, m->d_neigh, d_bnd, m->d_face_center, m->d_cell_center, m->d_lptr, m->d_lother,
                       m->d_gcoef, m->d_gdiag, err);
    FVG_HIP(hipGetLastError());
    FVG_TRY(download(c, meta, d_meta, 1));
    if ((unsigned long long)meta[0] != kUnsetKey) {
        const unsigned long long key = (unsigned long long)meta[0];
        return fail((key & 7) == kErrSlot ? FVG_E_MISSING_ENTRY : FVG_E_MESH, mesh_error(key));
    }
    return FVG_OK;
}

}  // namespace

int mesh_cell_centres(fvg_ctx* c, const fvg_mesh* m, uint64_t id, std::vector<float>& out, int& dim) {
    out.clear();
    dim = 0;
    std::lock_guard<std::mutex> lock(g_live_mu);
    const auto it = g_live.find(m);
    if (it == g_live.end() || it->second != id) return FVG_OK;
    std::vector<double> centre((size_t)m->n_cells * (size_t)m->dim);
    FVG_HIP(hipMemcpyAsync(centre.data(), m->d_cell_center, centre.size() * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
    FVG_HIP(hipStreamSynchronize(c->stream));
    out.assign(centre.begin(), centre.end());
    dim = m->dim;
    return FVG_OK;
}

int csr_from_device_pattern(fvg_ctx* c, int64_t nrows, int64_t ncols, int64_t nnz, const int32_t* d_row_ptr,
                            const int32_t* d_col, fvg_csr** out);

}  // namespace fvg

using namespace fvg;

extern "C" {

int fvg_mesh_create(fvg_ctx* c, int dim, int64_t n_nodes, const double* nodes, int64_t n_faces,
                    const int64_t* face_ptr, const int32_t* face_nodes, const int32_t* face_patch, int64_t n_cells,
                    const int64_t* cell_ptr, const int32_t* cell_faces, fvg_mesh** out) {
    if (!c || !out || !nodes || !face_ptr || !face_nodes || !face_patch || !cell_ptr || !cell_faces)
        return fail(FVG_E_INVALID, "null argument");
    *out = nullptr;
    if (dim != 2 && dim != 3) return fail(FVG_E_INVALID, "dimension must be 2 or 3");
    if (n_faces <= 0 || n_cells <= 0 || n_nodes <= 0) return fail(FVG_E_INVALID, "empty mesh");
    if (n_cells >= INT32_MAX || n_faces > kFaceMask) return fail(FVG_E_INVALID, "mesh too large for int32 indices");
    if (!offsets_valid(face_ptr, n_faces)) return fail(FVG_E_INVALID, "face_ptr must start at 0 and not decrease");
    if (!offsets_valid(cell_ptr, n_cells)) return fail(FVG_E_INVALID, "cell_ptr must start at 0 and not decrease");
    auto* m = new fvg_mesh();
    m->dim = dim;
    m->n_nodes = n_nodes;
    m->n_faces = n_faces;
    m->n_cells = n_cells;
    int s = dim == 2 ? build_mesh<2>(c, m, n_nodes, nodes, face_ptr, face_nodes, face_patch, cell_ptr, cell_faces)
                     : build_mesh<3>(c, m, n_nodes, nodes, face_ptr, face_nodes, face_patch, cell_ptr, cell_faces);
    if (s == FVG_OK && hipStreamSynchronize(c->stream) != hipSuccess) s = fail(FVG_E_HIP, "mesh build failed");
    if (s != FVG_OK) {
        (void)hipStreamSynchronize(c->stream);
        free_mesh(c, m);
        delete m;
        return s;
    }
    m->pattern_id = g_pattern_counter.fetch_add(1);
    {
        std::lock_guard<std::mutex> lock(g_live_mu);
        g_live[m] = m->pattern_id;
    }
    *out = m;
    return FVG_OK;
}

int fvg_mesh_destroy(fvg_ctx* c, fvg_mesh* m) {
    if (!m) return FVG_OK;
    {
        std::lock_guard<std::mutex> lock(g_live_mu);
        g_live.erase(m);
    }
    if (c) {
        (void)hipStreamSynchronize(c->stream);
        free_mesh(c, m);
    }
    delete m;
    return FVG_OK;
}

int fvg_mesh_sizes(const fvg_mesh* m, int64_t* n_cells, int64_t* n_faces, int64_t* n_bnd, int64_t* nnz,
                   int32_t* max_cell_faces) {
    if (!m) return fail(FVG_E_INVALID, "null mesh");
    if (n_cells) *n_cells = m->n_cells;
    if (n_faces) *n_faces = m->n_faces;
    if (n_bnd) *n_bnd = m->n_bnd;
    if (nnz) *nnz = m->nnz;
    if (max_cell_faces) *max_cell_faces = m->max_cell_faces;
    return FVG_OK;
}

int fvg_mesh_geometry(const fvg_mesh* m, double* face_area, double* face_center, double* face_normal,
                      int32_t* owner, int32_t* neighbor, double* cell_volume, double* cell_center) {
    if (!m) return fail(FVG_E_INVALID, "null mesh");
    const size_t F = (size_t)m->n_faces, N = (size_t)m->n_cells, D = (size_t)m->dim;
    if (face_area) FVG_HIP(hipMemcpy(face_area, m->d_area, F * sizeof(double), hipMemcpyDeviceToHost));
    if (face_center) FVG_HIP(hipMemcpy(face_center, m->d_face_center, F * D * sizeof(double), hipMemcpyDeviceToHost));
    if (face_normal) FVG_HIP(hipMemcpy(face_normal, m->d_face_normal, F * D * sizeof(double), hipMemcpyDeviceToHost));
    if (owner) FVG_HIP(hipMemcpy(owner, m->d_owner, F * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (neighbor) FVG_HIP(hipMemcpy(neighbor, m->d_neigh, F * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (cell_volume) FVG_HIP(hipMemcpy(cell_volume, m->d_cell_vol, N * sizeof(double), hipMemcpyDeviceToHost));
    if (cell_center) FVG_HIP(hipMemcpy(cell_center, m->d_cell_center, N * D * sizeof(double), hipMemcpyDeviceToHost));
    return FVG_OK;
}

int fvg_csr_from_mesh(fvg_ctx* c, const fvg_mesh* m, fvg_csr** out) {
    if (!c || !m || !out) return fail(FVG_E_INVALID, "null argument");
    // every row holds its diagonal, so max_minor + 1 (dmatrix.rs:47) is the cell count
    FVG_TRY(csr_from_device_pattern(c, m->n_cells, m->n_cells, m->nnz, m->d_row_ptr, m->d_col, out));
    (*out)->pattern_id = m->pattern_id;
    // the cell centres as row coordinates: read from the mesh when a preconditioner
    // analysis asks for them (mesh_cell_centres)
    (*out)->coord_mesh = m;
    (*out)->coord_mesh_id = m->pattern_id;
    return FVG_OK;
}

}  // extern "C"
