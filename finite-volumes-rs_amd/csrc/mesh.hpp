// mesh.hpp -- device-resident mesh used by assembly (internal).
#pragma once

#include "internal.hpp"

struct fvg_mesh {
    int dim = 0;
    int64_t n_nodes = 0, n_faces = 0, n_cells = 0, n_bnd = 0, nnz = 0;
    int n_colors = 0;
    uint64_t pattern_id = 0;

    // host geometry, exactly as Mesh::compute leaves it
    std::vector<double> face_area, face_center, face_normal, cell_volume, cell_center;
    std::vector<int32_t> owner, neighbor, bnd_ordinal;   // neighbor -1 / bnd_ordinal -1 as applicable
    std::vector<double> face_tinv;                        // 1/|x_j - x_i| or 1/|x_f - x_i| (Orthogonal)
    // cell-to-cell CSR pattern (from_cut_sparsity), int32
    std::vector<int32_t> row_ptr, col, diagpos;
    std::vector<int64_t> color_ptr;                       // faces of colour c: [color_ptr[c], color_ptr[c+1])

    // device arrays, faces in colour order
    int32_t* d_face = nullptr;     // original face id
    int32_t* d_owner = nullptr;
    int32_t* d_neigh = nullptr;    // -1 on boundary faces
    int32_t* d_bnd = nullptr;      // boundary ordinal, -1 interior
    int32_t* d_slot = nullptr;     // 4 per face: (i,i) (i,j) (j,i) (j,j)
    double* d_area = nullptr;
    double* d_tinv = nullptr;
    double* d_cell_vol = nullptr;
    int32_t* d_diagpos = nullptr;
    int32_t* d_row_ptr = nullptr;
};

namespace fvg {
// Host part of fvg_mesh_create: geometry, pattern, slots, colouring.  Returns status.
int build_mesh_host(fvg_mesh* m, int dim, int64_t n_nodes, const double* nodes, int64_t n_faces,
                    const int64_t* face_ptr, const int32_t* face_nodes, const int32_t* face_patch,
                    int64_t n_cells, const int64_t* cell_ptr, const int32_t* cell_faces,
                    std::vector<int32_t>& slots_colour_order, std::vector<int32_t>& face_colour_order);
}
