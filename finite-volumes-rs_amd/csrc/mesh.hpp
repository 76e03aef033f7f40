// mesh.hpp -- device-resident mesh used by assembly (internal).
//
// Everything Mesh::compute produces lives on the device, indexed by original face and
// cell ids.  Assembly gathers per cell: `ent` lists, for every cell, the faces it owns
// or neighbours, in ascending face id, each with its role and the CSR slot (or
// boundary ordinal) it updates.
#pragma once

#include "internal.hpp"

struct fvg_mesh {
    int dim = 0;
    int64_t n_nodes = 0, n_faces = 0, n_cells = 0, n_bnd = 0, nnz = 0, n_ent = 0;
    int32_t max_cell_faces = 0;
    uint64_t pattern_id = 0;
    // every off-diagonal CSR slot is updated by exactly one cell-face entry (no two faces
    // join the same pair of cells), so assembly may store those slots instead of adding
    bool unique_slots = false;

    // geometry, exactly as Mesh::compute leaves it
    double* d_area = nullptr;          // [F]
    double* d_face_center = nullptr;   // [F * dim]
    double* d_face_normal = nullptr;   // [F * dim], oriented owner -> neighbour / outward
    double* d_tinv = nullptr;          // [F] 1/|x_j - x_i| or 1/|x_f - x_i| (Orthogonal)
    double2* d_tarea = nullptr;        // [F] {tinv, area}: one 16-byte gather per face in assembly
    double* d_cell_vol = nullptr;      // [N]
    double* d_cell_center = nullptr;   // [N * dim]
    int32_t* d_owner = nullptr;        // [F]
    int32_t* d_neigh = nullptr;        // [F], -1 on boundary faces
    int32_t* d_bnd = nullptr;          // [F + 1] boundary ordinal (meaningful where d_neigh < 0)
    // cell-to-cell CSR pattern (from_cut_sparsity), int32
    int32_t* d_row_ptr = nullptr;      // [N + 1]
    int32_t* d_col = nullptr;          // [nnz]
    int32_t* d_diagpos = nullptr;      // [N]
    // per-cell assembly entries: ent[k] = {face | role << 30, slot or boundary ordinal}
    int32_t* d_ent_ptr = nullptr;      // [N + 1]
    int2* d_ent = nullptr;             // [n_ent]
    // least-squares gradient coefficients (compute.rs:154-200), per cell-face listing in
    // the order the cells list their faces
    int64_t n_list = 0;
    int32_t* d_lptr = nullptr;         // [N + 1] listing offsets
    int32_t* d_lother = nullptr;       // [n_list] other cell, or -(boundary ordinal + 1)
    double* d_gcoef = nullptr;         // [n_list * dim]
    double* d_gdiag = nullptr;         // [N * dim]
};

namespace fvg {
// The cell centres of mesh m (dim floats per cell) into out, if m is still alive with
// pattern id `id`; out is left empty otherwise.
int mesh_cell_centres(fvg_ctx* c, const fvg_mesh* m, uint64_t id, std::vector<float>& out, int& dim);
// entry roles of the per-cell assembly list
constexpr int kRoleOwner = 0;       // interior face, this cell owns it: .y = slot (i, j)
constexpr int kRoleNeighbour = 1;   // interior face, this cell is its neighbour: .y = slot (j, i)
constexpr int kRoleBoundary = 2;    // boundary face owned by this cell: .y = boundary ordinal
constexpr int kRoleShift = 30;
constexpr int32_t kFaceMask = (1 << kRoleShift) - 1;
}  // namespace fvg
