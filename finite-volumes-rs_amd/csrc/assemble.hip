// assemble.hip -- device mesh and fvm::assembly::assemble (src/fvm/assembly.rs:11-79).
//
// Assembly is face-parallel.  Faces are greedily coloured on the host so that no two
// faces of one colour touch the same cell; each colour is one launch whose threads
// scatter straight into their precomputed CSR slots with plain read-modify-write, no
// atomics.  The equation arrives as a postfix program of reference terms evaluated
// per face / per cell in the reference's operation order.
#include <atomic>
#include <cstring>

#include "mesh.hpp"

namespace fvg {

constexpr int kMaxFields = 8;
constexpr int kMaxStack = 16;

struct TermProgram {
    int n_ops;
    fvg_term_op ops[FVG_MAX_TERM_OPS];
    const double* face_fields[kMaxFields];
    const double* cell_fields[kMaxFields];
};

struct Triple {
    double li, lj, r;
};

// Term::face_terms of the whole expression for face f (original id), face geometry
// supplied.  laplacian.rs:35-45 + orthogonal.rs:26-44, convection.rs:36-44 +
// upwind.rs:26-33, divergence.rs:36-45, algebra operations.rs.
__device__ __forceinline__ Triple face_terms(const TermProgram& P, int f, double tinv) {
    Triple st[kMaxStack];
    int sp = 0;
    for (int k = 0; k < P.n_ops; ++k) {
        const fvg_term_op o = P.ops[k];
        switch (o.op) {
        case FVG_OP_LAPLACIAN_ORTHOGONAL: {
            const double t = tinv;
            const double l0 = 1.0 * (-t), l1 = 1.0 * t, r = 0.0;
            const double diff = P.face_fields[o.arg0][f];
            st[sp++] = Triple{diff * l0, diff * l1, -(diff * r)};
            break;
        }
        case FVG_OP_CONVECTION_UPWIND: {
            const double w = P.face_fields[o.arg1][f];
            const double l0 = w > 0.0 ? 1.0 : 0.0, l1 = w > 0.0 ? 0.0 : 1.0, r = 0.0;
            const double flux = P.face_fields[o.arg0][f];
            st[sp++] = Triple{l0 * flux, l1 * flux, (-r) * flux};
            break;
        }
        case FVG_OP_DIVERGENCE:
            st[sp++] = Triple{0.0, 0.0, -P.face_fields[o.arg0][f]};
            break;
        case FVG_OP_ADD: {
            Triple b = st[--sp], a = st[--sp];
            st[sp++] = Triple{a.li + b.li, a.lj + b.lj, a.r + b.r};
            break;
        }
        case FVG_OP_SUB: {
            Triple b = st[--sp], a = st[--sp];
            st[sp++] = Triple{a.li - b.li, a.lj - b.lj, a.r - b.r};
            break;
        }
        case FVG_OP_NEG: {
            Triple a = st[--sp];
            st[sp++] = Triple{-a.li, -a.lj, -a.r};
            break;
        }
        default:   // cell-only terms contribute the Term::face_terms default
            st[sp++] = Triple{0.0, 0.0, 0.0};
            break;
        }
    }
    return st[0];
}

// Term::cell_terms of the whole expression for cell c.  source.rs:34-39, 72-77,
// terms/time.rs:31-34 + schemes/time/euler.rs:27-37.
__device__ __forceinline__ void cell_terms(const TermProgram& P, int c, double& l_out, double& r_out) {
    double sl[kMaxStack], sr[kMaxStack];
    int sp = 0;
    for (int k = 0; k < P.n_ops; ++k) {
        const fvg_term_op o = P.ops[k];
        switch (o.op) {
        case FVG_OP_SOURCE:
            sl[sp] = 0.0;
            sr[sp++] = -P.cell_fields[o.arg0][c];
            break;
        case FVG_OP_LINEAR_SOURCE:
            sl[sp] = P.cell_fields[o.arg0][c];
            sr[sp++] = 0.0;
            break;
        case FVG_OP_TIME_EULER: {
            double r0 = 1.0, r1 = 1.0;
            if (o.arg1 >= 0) {
                r0 = P.cell_fields[o.arg1][c];
                r1 = P.cell_fields[o.arg2][c];
            }
            const double dt = o.scalar;
            sl[sp] = 1.0 * (1.0 / dt) * r0;
            sr[sp++] = P.cell_fields[o.arg0][c] / dt * r1;
            break;
        }
        case FVG_OP_ADD: {
            --sp;
            sl[sp - 1] = sl[sp - 1] + sl[sp];
            sr[sp - 1] = sr[sp - 1] + sr[sp];
            break;
        }
        case FVG_OP_SUB: {
            --sp;
            sl[sp - 1] = sl[sp - 1] - sl[sp];
            sr[sp - 1] = sr[sp - 1] - sr[sp];
            break;
        }
        case FVG_OP_NEG:
            sl[sp - 1] = -sl[sp - 1];
            sr[sp - 1] = -sr[sp - 1];
            break;
        default:   // face-only terms contribute the Term::cell_terms default
            sl[sp] = 0.0;
            sr[sp++] = 0.0;
            break;
        }
    }
    l_out = sl[0];
    r_out = sr[0];
}

// Cell loop (assembly.rs:26-32): A_cc = l*vol, b_c = r*vol, every other slot of the
// row cleared (the matrix starts from zero values in the reference).
__global__ void __launch_bounds__(kBlock) k_assemble_cells(TermProgram P, int n, const int32_t* __restrict__ row_ptr,
                                                           const int32_t* __restrict__ diagpos,
                                                           const double* __restrict__ vol, double* __restrict__ val,
                                                           double* __restrict__ rhs) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    double l, r;
    cell_terms(P, c, l, r);
    const double v = vol[c];
    const int d = diagpos[c];
    for (int k = row_ptr[c]; k < row_ptr[c + 1]; ++k) val[k] = k == d ? l * v : 0.0;
    rhs[c] = r * v;
}

// Face loop (assembly.rs:35-76) for the faces of one colour.
__global__ void __launch_bounds__(kBlock) k_assemble_faces(TermProgram P, int64_t begin, int64_t end,
                                                           const int32_t* __restrict__ face,
                                                           const int32_t* __restrict__ owner,
                                                           const int32_t* __restrict__ neigh,
                                                           const int32_t* __restrict__ bnd,
                                                           const int32_t* __restrict__ slot,
                                                           const double* __restrict__ area,
                                                           const double* __restrict__ tinv,
                                                           const double* __restrict__ bc_i,
                                                           const double* __restrict__ bc_e, double* __restrict__ val,
                                                           double* __restrict__ rhs) {
    const int64_t q = begin + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= end) return;
    const int f = face[q];
    const Triple t = face_terms(P, f, tinv[q]);
    const double A = area[q];
    const double li = t.li * A, lj = t.lj * A, r = t.r * A;
    const int i = owner[q], j = neigh[q];
    const int4 s = reinterpret_cast<const int4*>(slot)[q];
    if (j >= 0) {
        val[s.x] += li;
        val[s.y] += lj;
        rhs[i] += r;
        val[s.z] -= li;
        val[s.w] -= lj;
        rhs[j] -= r;
    } else {
        const int b = bnd[q];
        const double bndi = bc_i[b], bnde = bc_e[b];
        val[s.x] += li;
        val[s.x] += lj * bndi;
        rhs[i] += r - lj * bnde;
    }
}

static std::atomic<uint64_t> g_pattern_counter{1};

static void free_mesh(fvg_ctx* c, fvg_mesh* m) {
    const int64_t F = m->n_faces, N = m->n_cells;
    dev_free_n(c, m->d_face, F);
    dev_free_n(c, m->d_owner, F);
    dev_free_n(c, m->d_neigh, F);
    dev_free_n(c, m->d_bnd, F);
    dev_free_n(c, m->d_slot, 4 * F);
    dev_free_n(c, m->d_area, F);
    dev_free_n(c, m->d_tinv, F);
    dev_free_n(c, m->d_cell_vol, N);
    dev_free_n(c, m->d_diagpos, N);
    dev_free_n(c, m->d_row_ptr, N + 1);
}

// stack-depth check so the device evaluator never over- or under-runs
static int validate_program(const fvg_term_op* prog, int n, int nff, int ncf) {
    if (n <= 0 || n > FVG_MAX_TERM_OPS) return fail(FVG_E_INVALID, "term program length out of range");
    int depth = 0;
    for (int k = 0; k < n; ++k) {
        const fvg_term_op& o = prog[k];
        switch (o.op) {
        case FVG_OP_LAPLACIAN_ORTHOGONAL:
        case FVG_OP_DIVERGENCE:
            if (o.arg0 < 0 || o.arg0 >= nff) return fail(FVG_E_INVALID, "face field index out of range");
            ++depth;
            break;
        case FVG_OP_CONVECTION_UPWIND:
            if (o.arg0 < 0 || o.arg0 >= nff || o.arg1 < 0 || o.arg1 >= nff)
                return fail(FVG_E_INVALID, "face field index out of range");
            ++depth;
            break;
        case FVG_OP_SOURCE:
        case FVG_OP_LINEAR_SOURCE:
            if (o.arg0 < 0 || o.arg0 >= ncf) return fail(FVG_E_INVALID, "cell field index out of range");
            ++depth;
            break;
        case FVG_OP_TIME_EULER:
            if (o.arg0 < 0 || o.arg0 >= ncf) return fail(FVG_E_INVALID, "cell field index out of range");
            if (o.arg1 >= 0 && (o.arg1 >= ncf || o.arg2 < 0 || o.arg2 >= ncf))
                return fail(FVG_E_INVALID, "density field index out of range");
            ++depth;
            break;
        case FVG_OP_ADD:
        case FVG_OP_SUB:
            if (depth < 2) return fail(FVG_E_INVALID, "binary term operator needs two operands");
            --depth;
            break;
        case FVG_OP_NEG:
            if (depth < 1) return fail(FVG_E_INVALID, "negation needs an operand");
            break;
        default:
            return fail(FVG_E_INVALID, "unknown term opcode " + std::to_string(o.op));
        }
        if (depth > kMaxStack) return fail(FVG_E_INVALID, "term program too deep");
    }
    if (depth != 1) return fail(FVG_E_INVALID, "term program must leave exactly one term");
    return FVG_OK;
}

}  // namespace fvg

using namespace fvg;

extern "C" {

int fvg_mesh_create(fvg_ctx* c, int dim, int64_t n_nodes, const double* nodes, int64_t n_faces,
                    const int64_t* face_ptr, const int32_t* face_nodes, const int32_t* face_patch, int64_t n_cells,
                    const int64_t* cell_ptr, const int32_t* cell_faces, fvg_mesh** out) {
    if (!c || !out || !nodes || !face_ptr || !face_nodes || !face_patch || !cell_ptr || !cell_faces)
        return fail(FVG_E_INVALID, "null argument");
    *out = nullptr;
    auto* m = new fvg_mesh();
    std::vector<int32_t> slots, order;
    int s = build_mesh_host(m, dim, n_nodes, nodes, n_faces, face_ptr, face_nodes, face_patch, n_cells, cell_ptr,
                            cell_faces, slots, order);
    if (s != FVG_OK) {
        delete m;
        return s;
    }
    m->pattern_id = g_pattern_counter.fetch_add(1);
    const int64_t F = n_faces, N = n_cells;
    std::vector<int32_t> h_owner(F), h_neigh(F), h_bnd(F);
    std::vector<double> h_area(F), h_tinv(F);
    for (int64_t q = 0; q < F; ++q) {
        const int32_t f = order[(size_t)q];
        h_owner[(size_t)q] = m->owner[(size_t)f];
        h_neigh[(size_t)q] = m->neighbor[(size_t)f];
        h_bnd[(size_t)q] = m->bnd_ordinal[(size_t)f];
        h_area[(size_t)q] = m->face_area[(size_t)f];
        h_tinv[(size_t)q] = m->face_tinv[(size_t)f];
    }
    if ((s = dev_alloc_n(c, F, &m->d_face)) || (s = dev_alloc_n(c, F, &m->d_owner)) ||
        (s = dev_alloc_n(c, F, &m->d_neigh)) || (s = dev_alloc_n(c, F, &m->d_bnd)) ||
        (s = dev_alloc_n(c, 4 * F, &m->d_slot)) || (s = dev_alloc_n(c, F, &m->d_area)) ||
        (s = dev_alloc_n(c, F, &m->d_tinv)) || (s = dev_alloc_n(c, N, &m->d_cell_vol)) ||
        (s = dev_alloc_n(c, N, &m->d_diagpos)) || (s = dev_alloc_n(c, N + 1, &m->d_row_ptr)) ||
        (s = upload(c, m->d_face, order.data(), F)) || (s = upload(c, m->d_owner, h_owner.data(), F)) ||
        (s = upload(c, m->d_neigh, h_neigh.data(), F)) || (s = upload(c, m->d_bnd, h_bnd.data(), F)) ||
        (s = upload(c, m->d_slot, slots.data(), 4 * F)) || (s = upload(c, m->d_area, h_area.data(), F)) ||
        (s = upload(c, m->d_tinv, h_tinv.data(), F)) || (s = upload(c, m->d_cell_vol, m->cell_volume.data(), N)) ||
        (s = upload(c, m->d_diagpos, m->diagpos.data(), N)) ||
        (s = upload(c, m->d_row_ptr, m->row_ptr.data(), N + 1))) {
        free_mesh(c, m);
        delete m;
        return s;
    }
    if (hipStreamSynchronize(c->stream) != hipSuccess) {
        free_mesh(c, m);
        delete m;
        return fail(FVG_E_HIP, "mesh upload failed");
    }
    *out = m;
    return FVG_OK;
}

int fvg_mesh_destroy(fvg_ctx* c, fvg_mesh* m) {
    if (!m) return FVG_OK;
    if (c) {
        (void)hipStreamSynchronize(c->stream);
        free_mesh(c, m);
    }
    delete m;
    return FVG_OK;
}

int fvg_mesh_sizes(const fvg_mesh* m, int64_t* n_cells, int64_t* n_faces, int64_t* n_bnd, int64_t* nnz,
                   int32_t* n_colors) {
    if (!m) return fail(FVG_E_INVALID, "null mesh");
    if (n_cells) *n_cells = m->n_cells;
    if (n_faces) *n_faces = m->n_faces;
    if (n_bnd) *n_bnd = m->n_bnd;
    if (nnz) *nnz = m->nnz;
    if (n_colors) *n_colors = m->n_colors;
    return FVG_OK;
}

int fvg_mesh_geometry(const fvg_mesh* m, double* face_area, double* face_center, double* face_normal,
                      int32_t* owner, int32_t* neighbor, double* cell_volume, double* cell_center) {
    if (!m) return fail(FVG_E_INVALID, "null mesh");
    const size_t F = (size_t)m->n_faces, N = (size_t)m->n_cells, D = (size_t)m->dim;
    if (face_area) std::memcpy(face_area, m->face_area.data(), F * sizeof(double));
    if (face_center) std::memcpy(face_center, m->face_center.data(), F * D * sizeof(double));
    if (face_normal) std::memcpy(face_normal, m->face_normal.data(), F * D * sizeof(double));
    if (owner) std::memcpy(owner, m->owner.data(), F * sizeof(int32_t));
    if (neighbor) std::memcpy(neighbor, m->neighbor.data(), F * sizeof(int32_t));
    if (cell_volume) std::memcpy(cell_volume, m->cell_volume.data(), N * sizeof(double));
    if (cell_center) std::memcpy(cell_center, m->cell_center.data(), N * D * sizeof(double));
    return FVG_OK;
}

int fvg_csr_from_mesh(fvg_ctx* c, const fvg_mesh* m, fvg_csr** out) {
    if (!c || !m || !out) return fail(FVG_E_INVALID, "null argument");
    int64_t ncols = 0;
    for (int32_t j : m->col) ncols = j + 1 > ncols ? j + 1 : ncols;   // max_minor + 1 (dmatrix.rs:47)
    FVG_TRY(fvg_csr_create(c, m->n_cells, ncols, m->nnz, m->row_ptr.data(), m->col.data(), nullptr, out));
    (*out)->pattern_id = m->pattern_id;
    return FVG_OK;
}

int fvg_assemble(fvg_ctx* c, const fvg_mesh* m, const fvg_term_op* program, int32_t n_ops,
                 const fvg_vec* const* face_fields, int32_t n_face_fields, const fvg_vec* const* cell_fields,
                 int32_t n_cell_fields, const fvg_vec* bc_i, const fvg_vec* bc_e, fvg_csr* lhs, fvg_vec* rhs) {
    if (!c || !m || !program || !lhs || !rhs) return fail(FVG_E_INVALID, "null argument");
    if (n_face_fields < 0 || n_face_fields > kMaxFields || n_cell_fields < 0 || n_cell_fields > kMaxFields)
        return fail(FVG_E_INVALID, "too many fields");
    FVG_TRY(validate_program(program, n_ops, n_face_fields, n_cell_fields));
    if (lhs->pattern_id != m->pattern_id) return fail(FVG_E_INVALID, "lhs was not created from this mesh");
    if (rhs->n != m->n_cells) return fail(FVG_E_INVALID, "rhs length must equal the number of cells");
    if (m->n_bnd > 0 && (!bc_i || !bc_e || bc_i->n != m->n_bnd || bc_e->n != m->n_bnd))
        return fail(FVG_E_INVALID, "boundary closure arrays must have one entry per boundary face");
    TermProgram P{};
    P.n_ops = n_ops;
    for (int k = 0; k < n_ops; ++k) P.ops[k] = program[k];
    for (int k = 0; k < n_face_fields; ++k) {
        if (!face_fields[k] || face_fields[k]->n != m->n_faces)
            return fail(FVG_E_INVALID, "face field length must equal the number of faces");
        P.face_fields[k] = face_fields[k]->d;
    }
    for (int k = 0; k < n_cell_fields; ++k) {
        if (!cell_fields[k] || cell_fields[k]->n != m->n_cells)
            return fail(FVG_E_INVALID, "cell field length must equal the number of cells");
        P.cell_fields[k] = cell_fields[k]->d;
    }
    const int n = (int)m->n_cells;
    hipLaunchKernelGGL(k_assemble_cells, dim3(grid_for(n)), dim3(kBlock), 0, c->stream, P, n, m->d_row_ptr,
                       m->d_diagpos, m->d_cell_vol, lhs->val, rhs->d);
    const double* bi = m->n_bnd ? bc_i->d : nullptr;
    const double* be = m->n_bnd ? bc_e->d : nullptr;
    for (int col = 0; col < m->n_colors; ++col) {
        const int64_t b = m->color_ptr[(size_t)col], e = m->color_ptr[(size_t)col + 1];
        if (e <= b) continue;
        hipLaunchKernelGGL(k_assemble_faces, dim3(grid_for(e - b)), dim3(kBlock), 0, c->stream, P, b, e, m->d_face,
                           m->d_owner, m->d_neigh, m->d_bnd, m->d_slot, m->d_area, m->d_tinv, bi, be, lhs->val,
                           rhs->d);
    }
    FVG_HIP(hipGetLastError());
    lhs->version += 1;
    return FVG_OK;
}

}  // extern "C"
