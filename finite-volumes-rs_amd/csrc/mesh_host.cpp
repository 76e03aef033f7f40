// mesh_host.cpp -- host side of the mesh: Mesh::compute geometry, the cell-to-cell
// CSR pattern, per-face CSR slots and the face colouring for assembly.
//
// Mesh I/O and geometry stay on the host.  The arithmetic below repeats, operation
// for operation, what src/core/mesh/compute.rs does (sequential component sums,
// division by the count, sub-triangle loop in 3D), so the geometry is bit-identical
// to the reference's when compiled without floating-point contraction.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

#include "mesh.hpp"

namespace fvg {

namespace {

constexpr int32_t kUnset = -1;

// dot product of two DIM-vectors, summed from 0.0 in index order (vector.rs:37-43)
inline double vdot(const double* a, const double* b, int dim) {
    double x = 0.0;
    for (int d = 0; d < dim; ++d) x += a[d] * b[d];
    return x;
}

// face data: compute.rs:267-321
bool face_geometry(int dim, const double* nodes, const int32_t* fn, int64_t len, double* area_out, double* center,
                   double* normal) {
    if (dim == 2) {
        if (len != 2) return false;
        const double* n0 = nodes + 2 * (int64_t)fn[0];
        const double* n1 = nodes + 2 * (int64_t)fn[1];
        double dn[2] = {n1[0] - n0[0], n1[1] - n0[1]};
        double dsn[2] = {-dn[1], dn[0]};            // Normal for N == 2 (vector.rs:305-313)
        double area = std::sqrt(vdot(dsn, dsn, 2));
        *area_out = area;
        normal[0] = dsn[0] / area;
        normal[1] = dsn[1] / area;
        center[0] = (n0[0] + n1[0]) * 0.5;
        center[1] = (n0[1] + n1[1]) * 0.5;
        return true;
    }
    // node average (compute.rs:36-43)
    double ave[3] = {0.0, 0.0, 0.0};
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
        double area_i = std::sqrt(vdot(dsn, dsn, 3));
        area += area_i;
        for (int d = 0; d < 3; ++d) nrm[d] = dsn[d];  // the last sub-triangle's normal (compute.rs:305)
        for (int d = 0; d < 3; ++d) c[d] += (((n0[d] + n1[d]) + ave[d]) * 1.0 / 3.0) * area_i;
    }
    if (!(area > 2.220446049250313e-16)) return false;  // compute.rs:308
    for (int d = 0; d < 3; ++d) {
        nrm[d] /= area;
        c[d] /= area;
    }
    *area_out = area;
    for (int d = 0; d < 3; ++d) {
        center[d] = c[d];
        normal[d] = nrm[d];
    }
    return true;
}

}  // namespace

int build_mesh_host(fvg_mesh* m, int dim, int64_t n_nodes, const double* nodes, int64_t n_faces,
                    const int64_t* face_ptr, const int32_t* face_nodes, const int32_t* face_patch, int64_t n_cells,
                    const int64_t* cell_ptr, const int32_t* cell_faces, std::vector<int32_t>& slots,
                    std::vector<int32_t>& order) {
    if (dim != 2 && dim != 3) return fail(FVG_E_INVALID, "dimension must be 2 or 3");
    if (n_faces <= 0 || n_cells <= 0 || n_nodes <= 0) return fail(FVG_E_INVALID, "empty mesh");
    if (n_cells >= INT32_MAX || n_faces >= INT32_MAX) return fail(FVG_E_INVALID, "mesh too large for int32 indices");
    m->dim = dim;
    m->n_nodes = n_nodes;
    m->n_faces = n_faces;
    m->n_cells = n_cells;
    const int D = dim;

    // ---- owners and neighbours, as add_cell assigns them (mod.rs:511-525) ----
    m->owner.assign((size_t)n_faces, kUnset);
    m->neighbor.assign((size_t)n_faces, kUnset);
    for (int64_t c = 0; c < n_cells; ++c) {
        for (int64_t k = cell_ptr[c]; k < cell_ptr[c + 1]; ++k) {
            int32_t f = cell_faces[k];
            if (f < 0 || f >= n_faces) return fail(FVG_E_MESH, "cell names a face out of range");
            if (m->owner[(size_t)f] == kUnset) m->owner[(size_t)f] = (int32_t)c;
            else if (face_patch[f] < 0 && m->neighbor[(size_t)f] == kUnset) m->neighbor[(size_t)f] = (int32_t)c;
        }
    }
    m->bnd_ordinal.assign((size_t)n_faces, -1);
    int64_t nb = 0;
    for (int64_t f = 0; f < n_faces; ++f) {
        if (m->owner[(size_t)f] == kUnset) return fail(FVG_E_MESH, "face " + std::to_string(f) + " has no owner cell");
        if (face_patch[f] >= 0) m->bnd_ordinal[(size_t)f] = (int32_t)nb++;
        else if (m->neighbor[(size_t)f] == kUnset)
            return fail(FVG_E_MESH, "interior face " + std::to_string(f) + " has no neighbour cell");   // compute.rs:21-31
    }
    m->n_bnd = nb;

    // ---- face geometry (compute.rs:35-45) ----
    m->face_area.resize((size_t)n_faces);
    m->face_center.resize((size_t)(n_faces * D));
    m->face_normal.resize((size_t)(n_faces * D));
    for (int64_t f = 0; f < n_faces; ++f) {
        for (int64_t k = face_ptr[f]; k < face_ptr[f + 1]; ++k)
            if (face_nodes[k] < 0 || face_nodes[k] >= n_nodes) return fail(FVG_E_MESH, "face node out of range");
        if (!face_geometry(D, nodes, face_nodes + face_ptr[f], face_ptr[f + 1] - face_ptr[f], &m->face_area[(size_t)f],
                           &m->face_center[(size_t)(f * D)], &m->face_normal[(size_t)(f * D)]))
            return fail(FVG_E_MESH, "degenerate face " + std::to_string(f));
    }

    // ---- cell geometry (compute.rs:52-69, 325-375) ----
    m->cell_volume.resize((size_t)n_cells);
    m->cell_center.resize((size_t)(n_cells * D));
    for (int64_t c = 0; c < n_cells; ++c) {
        const int64_t s = cell_ptr[c], e = cell_ptr[c + 1];
        double ave[3] = {0.0, 0.0, 0.0};
        for (int64_t k = s; k < e; ++k)
            for (int d = 0; d < D; ++d) ave[d] += m->face_center[(size_t)(cell_faces[k] * (int64_t)D + d)];
        for (int d = 0; d < D; ++d) ave[d] /= (double)(e - s);
        double center[3] = {0.0, 0.0, 0.0}, volume = 0.0;
        for (int64_t k = s; k < e; ++k) {
            const int64_t f = cell_faces[k];
            const double* fc = &m->face_center[(size_t)(f * D)];
            const double* fnrm = &m->face_normal[(size_t)(f * D)];
            double dfc[3];
            for (int d = 0; d < D; ++d) dfc[d] = fc[d] - ave[d];
            double height = std::fabs(vdot(dfc, fnrm, D));
            double volume_i;
            double ci[3];
            if (D == 2) {
                volume_i = height * m->face_area[(size_t)f] / 2.0;
                for (int d = 0; d < 2; ++d) ci[d] = (fc[d] * 2.0 + ave[d]) / 3.0;
            } else {
                volume_i = height * m->face_area[(size_t)f] / 3.0;
                for (int d = 0; d < 3; ++d) ci[d] = (fc[d] * 3.0 + ave[d]) / 4.0;
            }
            for (int d = 0; d < D; ++d) center[d] += ci[d] * volume_i;
            volume += volume_i;
        }
        m->cell_volume[(size_t)c] = volume;
        for (int d = 0; d < D; ++d) m->cell_center[(size_t)(c * D + d)] = center[d] / volume;
    }

    // ---- normal orientation, owner -> neighbour / outward (compute.rs:97-131) ----
    for (int64_t f = 0; f < n_faces; ++f) {
        const double* inside = &m->cell_center[(size_t)((int64_t)m->owner[(size_t)f] * D)];
        double* nrm = &m->face_normal[(size_t)(f * D)];
        const double* fc = &m->face_center[(size_t)(f * D)];
        double d[3];
        for (int k = 0; k < D; ++k) d[k] = fc[k] - inside[k];
        if (vdot(nrm, d, D) < 0.0)
            for (int k = 0; k < D; ++k) nrm[k] *= -1.0;
    }

    // ---- Orthogonal scheme distance factor (orthogonal.rs:26-44) ----
    m->face_tinv.resize((size_t)n_faces);
    for (int64_t f = 0; f < n_faces; ++f) {
        const double* ci = &m->cell_center[(size_t)((int64_t)m->owner[(size_t)f] * D)];
        const double* cj = m->neighbor[(size_t)f] >= 0 ? &m->cell_center[(size_t)((int64_t)m->neighbor[(size_t)f] * D)]
                                                       : &m->face_center[(size_t)(f * D)];
        double d[3];
        for (int k = 0; k < D; ++k) d[k] = cj[k] - ci[k];
        double dx = std::sqrt(vdot(d, d, D));
        m->face_tinv[(size_t)f] = 1.0 / dx;
    }

    // ---- cell_to_cell (compute.rs:135-152) -> from_cut_sparsity sorted rows (dmatrix.rs:38-50) ----
    m->row_ptr.assign((size_t)n_cells + 1, 0);
    std::vector<int32_t> tmp;
    for (int64_t c = 0; c < n_cells; ++c) {
        int64_t deg = 1;
        for (int64_t k = cell_ptr[c]; k < cell_ptr[c + 1]; ++k)
            if (m->neighbor[(size_t)cell_faces[k]] >= 0) ++deg;
        m->row_ptr[(size_t)c + 1] = (int32_t)deg;   // upper bound, compacted below
    }
    m->col.clear();
    m->col.reserve((size_t)n_cells * 6);
    m->row_ptr[0] = 0;
    for (int64_t c = 0; c < n_cells; ++c) {
        tmp.clear();
        tmp.push_back((int32_t)c);
        for (int64_t k = cell_ptr[c]; k < cell_ptr[c + 1]; ++k) {
            const int32_t f = cell_faces[k];
            if (m->neighbor[(size_t)f] < 0) continue;
            tmp.push_back(m->owner[(size_t)f] == (int32_t)c ? m->neighbor[(size_t)f] : m->owner[(size_t)f]);
        }
        std::sort(tmp.begin(), tmp.end());
        tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
        m->col.insert(m->col.end(), tmp.begin(), tmp.end());
        m->row_ptr[(size_t)c + 1] = (int32_t)m->col.size();
    }
    m->nnz = (int64_t)m->col.size();
    m->diagpos.resize((size_t)n_cells);
    auto find_slot = [&](int32_t i, int32_t j) -> int32_t {
        auto b = m->col.begin() + m->row_ptr[(size_t)i], e = m->col.begin() + m->row_ptr[(size_t)i + 1];
        auto it = std::lower_bound(b, e, j);
        if (it == e || *it != j) return -1;
        return (int32_t)(it - m->col.begin());
    };
    for (int64_t c = 0; c < n_cells; ++c) m->diagpos[(size_t)c] = find_slot((int32_t)c, (int32_t)c);

    // ---- greedy face colouring: faces sharing a cell get different colours ----
    std::vector<uint64_t> used((size_t)n_cells, 0);
    std::vector<int32_t> colour((size_t)n_faces);
    int ncol = 0;
    for (int64_t f = 0; f < n_faces; ++f) {
        const int32_t i = m->owner[(size_t)f], j = m->neighbor[(size_t)f];
        uint64_t busy = used[(size_t)i] | (j >= 0 ? used[(size_t)j] : 0);
        if (busy == ~0ull) return fail(FVG_E_MESH, "face colouring needs more than 64 colours");
        int c = __builtin_ctzll(~busy);
        colour[(size_t)f] = c;
        used[(size_t)i] |= 1ull << c;
        if (j >= 0) used[(size_t)j] |= 1ull << c;
        ncol = std::max(ncol, c + 1);
    }
    m->n_colors = ncol;
    m->color_ptr.assign((size_t)ncol + 1, 0);
    for (int64_t f = 0; f < n_faces; ++f) m->color_ptr[(size_t)colour[(size_t)f] + 1]++;
    for (int c = 0; c < ncol; ++c) m->color_ptr[(size_t)c + 1] += m->color_ptr[(size_t)c];
    order.assign((size_t)n_faces, 0);
    {
        std::vector<int64_t> cursor(m->color_ptr.begin(), m->color_ptr.end() - 1);
        for (int64_t f = 0; f < n_faces; ++f) order[(size_t)cursor[(size_t)colour[(size_t)f]]++] = (int32_t)f;
    }
    slots.assign((size_t)n_faces * 4, -1);
    for (int64_t q = 0; q < n_faces; ++q) {
        const int32_t f = order[(size_t)q];
        const int32_t i = m->owner[(size_t)f], j = m->neighbor[(size_t)f];
        slots[(size_t)(4 * q + 0)] = m->diagpos[(size_t)i];
        if (j >= 0) {
            slots[(size_t)(4 * q + 1)] = find_slot(i, j);
            slots[(size_t)(4 * q + 2)] = find_slot(j, i);
            slots[(size_t)(4 * q + 3)] = m->diagpos[(size_t)j];
            if (slots[(size_t)(4 * q + 1)] < 0 || slots[(size_t)(4 * q + 2)] < 0)
                return fail(FVG_E_MISSING_ENTRY, "face slot missing from the CSR pattern");
        }
    }
    return FVG_OK;
}

}  // namespace fvg
