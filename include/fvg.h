/*
 * fvg.h -- C ABI of the MI355X assemble-and-solve library (libfvg.so).
 *
 * The library replaces, behind the same call shapes, the hot path of the
 * finite-volumes crate: cell-centred face-flux assembly into CSR and the
 * preconditioned Krylov solves.  Each entry point names the reference item it
 * stands for (paths relative to the crate root).  Everything is plain C: POD
 * structs, raw pointers and sizes; no exceptions cross the boundary.
 *
 * Conventions
 *   - Every function returns an int status (FVG_OK == 0).  fvg_last_error()
 *     returns a message for the most recent failure on the calling thread.
 *   - Handles are owned by the library until the matching *_destroy call.
 *   - Host pointers passed to *_create / *_upload are read during the call
 *     and not retained.
 *   - A context owns one HIP stream on one device.  It is not thread safe:
 *     use one context per host thread.  All work on a context is ordered.
 *   - Indices are int32 (rows, columns, faces, cells < 2^31).
 */
#ifndef FVG_H
#define FVG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
enum {
    FVG_OK = 0,
    FVG_E_INVALID = 1,        /* bad argument, shape or handle */
    FVG_E_HIP = 2,            /* HIP runtime failure */
    FVG_E_BUDGET = 3,         /* allocation would exceed the context's HBM budget */
    FVG_E_ZERO_PIVOT = 4,     /* ILU |u_kk| < 1e-16; the reference panics
                                 (src/linalg/preconditioners/incomplete_lower_upper.rs:234-236) */
    FVG_E_NOT_SYMMETRIC = 5,  /* IC on a matrix failing |a_ij - a_ji| < 1e-14; the reference asserts
                                 (src/linalg/preconditioners/incomplete_cholesky.rs:153,
                                  src/linalg/dmatrix.rs:76-88) */
    FVG_E_MISSING_ENTRY = 6,  /* a needed CSR slot is absent; the reference panics
                                 (src/linalg/dmatrix.rs:262-286) */
    FVG_E_MESH = 7,           /* invalid mesh topology (src/core/mesh/compute.rs:15-31) */
    FVG_E_TIMEOUT = 8,        /* a device-side wait exceeded its bounded spin limit */
    FVG_E_UNSUPPORTED = 9     /* valid request outside what this build implements */
};

const char* fvg_last_error(void);
const char* fvg_status_string(int status);

/* ---- context ------------------------------------------------------------ */
typedef struct fvg_ctx fvg_ctx;

/* hbm_budget_bytes == 0 means unlimited.  Allocations made through the
 * context are counted against the budget; exceeding it returns FVG_E_BUDGET. */
int fvg_ctx_create(int device, uint64_t hbm_budget_bytes, fvg_ctx** out);
int fvg_ctx_destroy(fvg_ctx* ctx);
int fvg_ctx_sync(fvg_ctx* ctx);
int fvg_ctx_bytes_in_use(const fvg_ctx* ctx, uint64_t* bytes);
/* The hipStream_t of the context, as void*, for callers that interleave work. */
int fvg_ctx_stream(const fvg_ctx* ctx, void** stream);

/* Launch-shape knobs (no counterpart in the reference; used by the autotuner).
 * Values of 0 restore the built-in choice. */
enum {
    FVG_TUNE_SWEEP_MODE = 1,    /* 0: the blocked sweep where its layout was built, else the persistent
                                   one; 1: one launch per level; 2: the persistent sweep */
    FVG_TUNE_SWEEP_WAVES = 2,   /* waves of the persistent sweep (0: about two and a half levels of chunks) */
    FVG_TUNE_SWEEP_POLL = 3,    /* sweep wait loop: 0 re-poll at once, 1 short sleep between polls, 2 two polls in flight */
    /* FP32 inner iterations of FVG_PRECISION_MIXED (device-wide solves): the incomplete
     * factor is applied by Jacobi-iterated triangular solves (0, the default) or by the
     * exact level-scheduled sweeps (1); FVG_TUNE_INNER_SWEEPS Jacobi sweeps per factor
     * (0: 2).  The FP64 refinement reaches the requested tolerance either way. */
    FVG_TUNE_INNER_EXACT = 4,
    FVG_TUNE_INNER_SWEEPS = 5,
    /* Launch mode of those FP32 inner iterations: 0 (default) replays them from a captured
     * hipGraph, 1 launches every kernel directly (the same kernels and work; for kernel
     * tracers that cannot follow graph launches). */
    FVG_TUNE_GRAPHS = 6,
    /* About how many blocks the blocked sweep layout of a preconditioner created
     * afterwards is cut into (0: one per compute unit for packed components, two for
     * blocks of consecutive rows; at most 4096). */
    FVG_TUNE_SWEEP_BLOCKS = 7,
    /* Polls a triangular-sweep wait makes before it gives up and the call returns
     * FVG_E_TIMEOUT (0: 2^18).  A tiny value forces the time-out path, for tests. */
    FVG_TUNE_SPIN_LIMIT = 8,
    /* Launch of the triangular sweeps whose workgroups wait on one another (sized so
     * that all of them fit on the device at once): 1 an ordinary launch, 2 a cooperative
     * launch, which also guarantees that they are resident together while other streams
     * run kernels that hold compute units; 0 (default) as fvg_sweep_launch_rule decides:
     * ordinary while this context is the only one alive on its device in the process, no
     * batch of it is in flight and no sweep wait of it has timed out, else cooperative.
     * The ordinary launch saves about 19 us of launch latency per sweep.  Should a wait
     * still find a workgroup that never became resident (kernels of another process can
     * hold compute units), it gives up after the spin limit, the call returns
     * FVG_E_TIMEOUT, and the context launches cooperatively from then on. */
    FVG_TUNE_SWEEP_LAUNCH = 9,
    /* Test-only: 1 makes the triangular sweeps keep every solved value from the other
     * workgroups and waves that wait on it, so every such wait runs to its spin limit and
     * the call returns FVG_E_TIMEOUT, while all kernels still exit normally. */
    FVG_TUNE_SWEEP_WITHHOLD = 10,
    /* FP64 fvg_cg / fvg_bicgstab with IC(0) / ILU(0) whose two sweeps are blocked and
     * mirrored (structurally symmetric patterns): 0 (default) iterates with every vector in
     * the lower sweep's position order, so no kernel reorders a vector for the sweeps; 1
     * iterates in natural row order.  The same iteration arithmetic element by element. */
    FVG_TUNE_SOLVE_ORDER = 11,
    /* fvg_assemble's kernel: 0 (default) one lane per cell-face entry where each
     * off-diagonal slot belongs to one face (a workgroup's rows built in LDS and written
     * out whole), else one lane per row; 1 always one lane per row; 2 the entry kernel
     * storing each value straight to its slot.  The same values every way. */
    FVG_TUNE_ASSEMBLE = 12,
    /* FP64 CG in position order: 0 (default) its residual update and r.r are done by the
     * lower sweep and r.z by the upper sweep of each preconditioner apply; 1 separate
     * kernels.  The same values either way. */
    FVG_TUNE_CG_FUSION = 13,
    /* Least number of levels between a value of one workgroup of the blocked triangular
     * sweep and its use in another, in the layouts of preconditioners created afterwards:
     * 0 (default) 3, so a workgroup fetches each such value a whole level before it needs
     * it; 2 to 8 that many (2: no delay beyond what the dependency components give).  More
     * levels for fewer waits on the fetches.  The same values every way. */
    FVG_TUNE_SWEEP_GAP = 14,
    /* Blocked layouts of matrices whose rows carry coordinates (fvg_csr_from_mesh,
     * fvg_csr_set_row_coords) where the dependencies form one large component (extruded
     * and 2D meshes), for preconditioners created afterwards: 0 (default) boxes of the
     * coordinates in a grid of about CUs/2 boxes, kept where their levels (and widths)
     * cost less than those of blocks of consecutive rows; 1 never; 2 to 1024 a grid of that
     * many boxes (rounded down to a power of two).  The same values every way. */
    FVG_TUNE_SWEEP_BOXES = 15
};
int fvg_ctx_set_tuning(fvg_ctx* ctx, int32_t knob, int64_t value);
int fvg_ctx_get_tuning(const fvg_ctx* ctx, int32_t knob, int64_t* value);
/* The launch mode the last co-resident sweep of this context used (1 ordinary,
 * 2 cooperative, 0 none yet) and how many of its batches are launched and not yet
 * finished (no counterpart in the reference). */
int fvg_ctx_sweep_launch_state(const fvg_ctx* ctx, int32_t* last_launch, int32_t* batches_in_flight);
/* The rule of FVG_TUNE_SWEEP_LAUNCH: the mode (1 ordinary, 2 cooperative) for the knob's
 * value, the live contexts of the device in the process, the context's batches in flight
 * and whether a sweep wait of the context timed out before; 0 for arguments out of range.
 * Pure host function (no device needed; no counterpart in the reference). */
int fvg_sweep_launch_rule(int64_t tune, int32_t live_contexts, int32_t batches_in_flight, int32_t timed_out_before);

/* ---- dense vectors: DistributedVector<f64> (src/linalg/dvector.rs:10-179) ---- */
typedef struct fvg_vec fvg_vec;

/* from_size (dvector.rs:19-21): zero-filled device vector of length n. */
int fvg_vec_create(fvg_ctx* ctx, int64_t n, fvg_vec** out);
/* Non-owning view of caller-owned device memory (n doubles). */
int fvg_vec_wrap(fvg_ctx* ctx, int64_t n, double* device_ptr, fvg_vec** out);
int fvg_vec_destroy(fvg_ctx* ctx, fvg_vec* v);
int fvg_vec_len(const fvg_vec* v, int64_t* n);
int fvg_vec_device_ptr(const fvg_vec* v, double** device_ptr);
int fvg_vec_upload(fvg_ctx* ctx, fvg_vec* v, const double* host);     /* from_data (dvector.rs:23-25) */
/* Uploads several vectors through one staging buffer without waiting for the copies
 * (they complete in stream order before any later work on the context stream). */
int fvg_vec_upload_many(fvg_ctx* ctx, int32_t count, fvg_vec* const* vecs, const double* const* hosts);
int fvg_vec_download(fvg_ctx* ctx, const fvg_vec* v, double* host);
/* Set every entry to value on the device (asynchronous on the context stream). */
int fvg_vec_fill(fvg_ctx* ctx, fvg_vec* v, double value);
/* DistributedVector::dot (dvector.rs:63-71); deterministic fixed-order tree sum. */
int fvg_vec_dot(fvg_ctx* ctx, const fvg_vec* a, const fvg_vec* b, double* out);

/* ---- mesh: Mesh<DIM> built with add_node / add_face / add_cell and
 *      Mesh::compute (src/core/mesh/mod.rs:444-527, src/core/mesh/compute.rs:12-152).
 *      The topology is uploaded once; owners and neighbours, geometry, normal
 *      orientation, the cell-to-cell CSR pattern and the per-cell assembly lists
 *      are computed on the device, bit-for-bit as the reference computes them,
 *      including its face-normal scaling in 3D (compute.rs:305) and the cell
 *      volume that follows from it (compute.rs:359-361). ----------------------- */
typedef struct fvg_mesh fvg_mesh;

/* face_patch[f] >= 0 marks a boundary face of that patch, -1 an interior face.
 * Cells name their faces; the first cell naming a face owns it, the second is
 * its neighbour (mod.rs:511-525).  face_ptr and cell_ptr must start at 0 and not
 * decrease (FVG_E_INVALID otherwise).  FVG_E_MESH names the first failing face or
 * cell: a face out of range, a face without owner, an interior face without
 * neighbour (compute.rs:15-31), a node out of range, a degenerate face
 * (compute.rs:308), or an interior face whose owner and neighbour are one cell
 * (the reference would divide by a zero distance there). */
int fvg_mesh_create(fvg_ctx* ctx, int dim, int64_t n_nodes, const double* nodes, int64_t n_faces,
                    const int64_t* face_ptr, const int32_t* face_nodes, const int32_t* face_patch,
                    int64_t n_cells, const int64_t* cell_ptr, const int32_t* cell_faces, fvg_mesh** out);
int fvg_mesh_destroy(fvg_ctx* ctx, fvg_mesh* mesh);
/* Sizes: cells, faces, boundary faces, nnz of the cell-to-cell CSR pattern, and the
 * largest number of faces one cell lists. */
int fvg_mesh_sizes(const fvg_mesh* mesh, int64_t* n_cells, int64_t* n_faces, int64_t* n_boundary_faces,
                   int64_t* nnz, int32_t* max_cell_faces);
/* Host copies of Mesh::compute results; any output may be NULL.
 * face_center/face_normal/cell_center are (count x dim) row-major.
 * neighbor[f] is -1 on boundary faces (FaceRef::owner/neighbor mod.rs:339-351). */
int fvg_mesh_geometry(const fvg_mesh* mesh, double* face_area, double* face_center, double* face_normal,
                      int32_t* owner, int32_t* neighbor, double* cell_volume, double* cell_center);

/* ---- mesh files (src/core/mesh/io.rs) and results (src/post/write_pvtu.rs) ----------
 * A topology is what Mesh::add_node / add_face / add_cell receive: coordinates, face
 * node lists with their boundary patch (-1 interior), cell face lists.  It lives on the
 * host; fvg_mesh_create_from_topology hands it to the device build. */
typedef struct fvg_topology fvg_topology;

/* Mesh::read_su2 (io.rs:489-533): NDIME, NPOIN, NELEM (VTK kinds 3, 5, 9, 10, 12),
 * markers numbered 0.. in file order; faces numbered as compute_vtk_faces does
 * (io.rs:364-486): interior faces first, then boundary faces by marker, each in order of
 * first appearance.  Tokens may be separated by any whitespace. */
int fvg_topology_read_su2(const char* path, fvg_topology** out);
/* Mesh::read (io.rs:68-211): the crate's own format written by Mesh::write
 * (io.rs:10-64).  A partition of a distributed mesh (remote-owned faces or cells) is
 * refused with FVG_E_UNSUPPORTED. */
int fvg_topology_read_native(const char* path, fvg_topology** out);
/* A topology from arrays (copied), laid out as for fvg_mesh_create. */
int fvg_topology_create(int dim, int64_t n_nodes, const double* nodes, int64_t n_faces, const int64_t* face_ptr,
                        const int32_t* face_nodes, const int32_t* face_patch, int64_t n_cells,
                        const int64_t* cell_ptr, const int32_t* cell_faces, fvg_topology** out);
int fvg_topology_destroy(fvg_topology* topology);
int fvg_topology_sizes(const fvg_topology* topology, int32_t* dim, int64_t* n_nodes, int64_t* n_faces,
                       int64_t* n_face_nodes, int64_t* n_cells, int64_t* n_cell_faces, int32_t* n_patches);
/* Pointers into the topology's own arrays (valid until it is destroyed); any may be NULL. */
int fvg_topology_arrays(const fvg_topology* topology, const double** nodes, const int64_t** face_ptr,
                        const int32_t** face_nodes, const int32_t** face_patch, const int64_t** cell_ptr,
                        const int32_t** cell_faces);
/* The marker / patch name of a patch id ("" when the file names none). */
int fvg_topology_patch_name(const fvg_topology* topology, int32_t patch, const char** name);
int fvg_mesh_create_from_topology(fvg_ctx* ctx, const fvg_topology* topology, fvg_mesh** out);

/* PvtuWriter (write_pvtu.rs:30-466) for one rank: <prefix>.pvtu and <prefix>_0.vtu, the
 * prefix being the path without its extension (file_without_extension, :416-424);
 * polygons in 2D, polyhedra in 3D; cell fields of 1 to 9 components (n_cells x ncomp,
 * row-major); numbers printed as Rust prints f64.  The mesh part is built once per
 * writer.  The polyhedra connectivity lists a cell's nodes in order of first
 * appearance, where the reference's HashSet gives an arbitrary order. */
typedef struct fvg_pvtu_writer fvg_pvtu_writer;
int fvg_pvtu_writer_create(const fvg_topology* topology, fvg_pvtu_writer** out);
int fvg_pvtu_write(fvg_pvtu_writer* writer, const char* path, int32_t n_fields, const char* const* names,
                   const int32_t* ncomps, const double* const* values);
/* The same from device vectors, off the critical path (no counterpart in the reference):
 * the fields are copied to pinned host memory on the context stream and the files are
 * written by a host thread once the copies land; the call returns at once.  One write is
 * in flight per writer: the next call, fvg_pvtu_wait or destroy waits for it. */
int fvg_pvtu_write_async(fvg_pvtu_writer* writer, fvg_ctx* ctx, const char* path, int32_t n_fields,
                         const char* const* names, const int32_t* ncomps, const fvg_vec* const* values);
/* Waits for the writer's asynchronous write and returns its status. */
int fvg_pvtu_wait(fvg_pvtu_writer* writer);
int fvg_pvtu_writer_destroy(fvg_pvtu_writer* writer);

/* ---- sparse matrix: DistributedMatrix<f64> in CSR (src/linalg/dmatrix.rs:14-301) ---- */
typedef struct fvg_csr fvg_csr;

/* DistributedMatrix::from_cut_sparsity(mesh.cell_to_cell_sparsity(), n_cells)
 * (dmatrix.rs:38-50): sorted pattern, zero values.  The matrix remembers the mesh: while
 * the mesh is alive, a preconditioner analysis may read its cell centres as the row
 * coordinates (see fvg_csr_set_row_coords, which replaces them). */
int fvg_csr_from_mesh(fvg_ctx* ctx, const fvg_mesh* mesh, fvg_csr** out);
/* General CSR (columns sorted per row).  values may be NULL (zeros). */
int fvg_csr_create(fvg_ctx* ctx, int64_t nrows, int64_t ncols, int64_t nnz, const int32_t* row_ptr,
                   const int32_t* col_idx, const double* values, fvg_csr** out);
int fvg_csr_destroy(fvg_ctx* ctx, fvg_csr* a);
int fvg_csr_sizes(const fvg_csr* a, int64_t* nrows, int64_t* ncols, int64_t* nnz);
int fvg_csr_download(fvg_ctx* ctx, const fvg_csr* a, int32_t* row_ptr, int32_t* col_idx, double* values);
int fvg_csr_upload_values(fvg_ctx* ctx, fvg_csr* a, const double* values);
/* Coordinates of the rows (no counterpart in the reference): `dim` (1 to 3) values per
 * row, row-major, e.g. the cell centres of the mesh the matrix was assembled on; dim 0
 * (coords NULL) drops them.  Only the layout analysis of preconditioners created
 * afterwards reads them (FVG_TUNE_SWEEP_BOXES): where the sweep dependencies form one
 * large component, boxes of the coordinates can give a blocked sweep fewer levels.
 * Solutions are the same bits with or without them. */
int fvg_csr_set_row_coords(fvg_csr* a, int32_t dim, const double* coords);
/* DistributedMatrix::imul (dmatrix.rs:136-146): y = A x, each row summed in
 * ascending column order, bit-identical to the sequential loop. */
int fvg_spmv(fvg_ctx* ctx, const fvg_csr* a, const fvg_vec* x, fvg_vec* y);
/* Measurement: `reps` back-to-back launches of y = A x between two HIP events on the
 * context stream; *avg_ms = elapsed / reps (no reference counterpart). */
int fvg_spmv_time(fvg_ctx* ctx, const fvg_csr* a, const fvg_vec* x, fvg_vec* y, int32_t reps, double* avg_ms);

/* ---- assembly: fvm::assembly::assemble (src/fvm/assembly.rs:11-79) -------------
 * The equation (an `impl Term` expression built from laplacian, convection,
 * source, ... and + - neg, src/fvm/terms/operations.rs) is passed as a postfix
 * program.  Field arguments index the face_fields / cell_fields arrays. */
enum {
    FVG_OP_LAPLACIAN_ORTHOGONAL = 1,  /* laplacian(Orthogonal, k_f): terms/laplacian.rs:35-45,
                                         schemes/facengrad/orthogonal.rs:26-44.  arg0 = face field k */
    FVG_OP_CONVECTION_UPWIND = 2,     /* convection(Upwind(&w), &flux): terms/convection.rs:36-44,
                                         schemes/faceinterp/upwind.rs:26-33.  arg0 = face field flux,
                                         arg1 = face field w deciding the upwind side */
    FVG_OP_SOURCE = 3,                /* source(&f): terms/source.rs:34-39.  arg0 = cell field f */
    FVG_OP_LINEAR_SOURCE = 4,         /* linear_source(&c): terms/source.rs:72-77.  arg0 = cell field c */
    FVG_OP_TIME_EULER = 5,            /* time(Euler): terms/time.rs:31-34, schemes/time/euler.rs:27-37.
                                         arg0 = cell field previous, scalar = dt, arg1/arg2 = cell
                                         fields density/density_previous or -1 */
    FVG_OP_DIVERGENCE = 6,            /* divergence(&flux): terms/divergence.rs:36-45.  arg0 = face field */
    FVG_OP_LAPLACIAN_CORRECTED = 7,   /* laplacian(Corrected(&grad, reduction), k_f): terms/laplacian.rs:35-45,
                                         schemes/facengrad/corrected.rs:34-67.  arg0 = face field k,
                                         arg1 = cell field holding the gradient (n_cells x dim, row-major),
                                         scalar = corrector_reduction */
    FVG_OP_CONVECTION_LINEAR = 8,     /* convection(Linear, &flux): schemes/faceinterp/linear.rs:22-44.
                                         arg0 = face field flux */
    FVG_OP_CONVECTION_LIMITED_LINEAR = 9, /* convection(LimitedLinear(&w, &lim), &flux):
                                         schemes/faceinterp/limitedlinear.rs:24-72.  arg0 = face field
                                         flux, arg1 = face field w (upwind side), arg2 = face limiters */
    FVG_OP_TIME_BACKWARD = 10,        /* time(Backward): terms/time.rs:31-34, schemes/time/backward.rs:21-62.
                                         arg0 / arg1 = cell fields previous / previous_2, arg2 = the first of
                                         three consecutive cell fields density, density_last,
                                         density_last2 or -1, scalar = dt; the next word must be
                                         FVG_OP_PARAM with scalar = dt_2 */
    FVG_OP_USER_TERM = 11,            /* any other `impl Term` (terms/mod.rs:24-40) with its values given:
                                         arg0 = the first of three consecutive face fields li, lj, r
                                         (its face_terms) or -1 for the trait's zero default, arg1 =
                                         the first of two consecutive cell fields l, r (its
                                         cell_terms) or -1 likewise */
    FVG_OP_PARAM = 15,                /* a parameter of the word before it (scalar); no term */
    FVG_OP_ADD = 16,                  /* a + b (operations.rs:86-120) */
    FVG_OP_SUB = 17,                  /* a - b (operations.rs:123-157) */
    FVG_OP_NEG = 18                   /* -a    (operations.rs:190-199) */
};

typedef struct fvg_term_op {
    int32_t op;
    int32_t arg0;
    int32_t arg1;
    int32_t arg2;
    double scalar;
} fvg_term_op;

#define FVG_MAX_TERM_OPS 32

/* Writes every value of `lhs` (whose pattern must be fvg_csr_from_mesh(mesh))
 * and every entry of `rhs`, one row per thread: each slot receives its updates in
 * the reference's face order, so the values are bit-identical to the reference's
 * sequential loops.  bc_i / bc_e carry the boundary closure
 * (value_outside = bc_i * value_inside + bc_e, assembly.rs:66-72) for each
 * boundary face in face order; they may be NULL when the mesh has none. */
int fvg_assemble(fvg_ctx* ctx, const fvg_mesh* mesh, const fvg_term_op* program, int32_t n_ops,
                 const fvg_vec* const* face_fields, int32_t n_face_fields, const fvg_vec* const* cell_fields,
                 int32_t n_cell_fields, const fvg_vec* bc_i, const fvg_vec* bc_e, fvg_csr* lhs, fvg_vec* rhs);

/* ---- fvm tools (src/fvm/tools) ---------------------------------------------------
 * Fields with ncomp components (1: Field<f64>; 2 or 3: Field<Vector<ncomp>>) are
 * n_cells x ncomp row-major; their gradients (Vector<dim> or Matrix<ncomp, dim>) are
 * n_cells x ncomp x dim.  The boundary closure is bc_i (one per boundary face) and
 * bc_e (one per boundary face and component), in boundary-face order; both may be
 * NULL on a mesh without boundary faces.  Results are bit-identical to the reference's
 * sequential loops. */

/* compute_gradients (tools/gradients.rs:41-79): least-squares cell gradients from the
 * coefficients Mesh::compute made (compute.rs:154-200). */
int fvg_compute_gradients(fvg_ctx* ctx, const fvg_mesh* mesh, int32_t ncomp, const fvg_vec* values,
                          const fvg_vec* bc_i, const fvg_vec* bc_e, fvg_vec* gradients);

enum {   /* schemes/limiters/mod.rs */
    FVG_LIMITER_LIMITED_LINEAR = 0,   /* LimitedLinear(k) */
    FVG_LIMITER_MINMOD = 1,
    FVG_LIMITER_VANLEER = 2,
    FVG_LIMITER_VANALBADA = 3,
    FVG_LIMITER_SWEBY = 4,            /* Sweby(k) */
    FVG_LIMITER_SUPERBEE = 5
};
enum { FVG_INTERP_LINEAR = 0, FVG_INTERP_UPWIND = 1 };          /* gradient interpolation scheme */
enum { FVG_FNGRAD_ORTHOGONAL = 0, FVG_FNGRAD_CORRECTED = 1 };  /* face-normal gradient scheme */

typedef struct fvg_limiter_spec {
    int32_t limiter;          /* FVG_LIMITER_* */
    int32_t interp;           /* FVG_INTERP_*: how the cell gradients are taken to the face */
    int32_t fngrad;           /* FVG_FNGRAD_* */
    int32_t reserved;
    double k;                 /* the limiter's parameter (LimitedLinear, Sweby) */
    double reduction;         /* Corrected: corrector_reduction */
    const fvg_vec* flux;      /* Upwind: face flux deciding the upwind side */
    const fvg_vec* fn_grads;  /* Corrected: the gradient field it corrects with */
} fvg_limiter_spec;

/* compute_limiters (tools/limiters.rs:47-126): one limiter per face. */
int fvg_compute_limiters(fvg_ctx* ctx, const fvg_mesh* mesh, int32_t ncomp, const fvg_vec* values,
                         const fvg_vec* gradients, const fvg_limiter_spec* spec, const fvg_vec* bc_i,
                         const fvg_vec* bc_e, fvg_vec* limiters);

/* ---- SIMPLE helpers (tools/hbya.rs) -----------------------------------------------
 * Vector<dim> cell fields (velocity, hbya, pressure gradient) are n_cells x dim; the
 * velocity closure is u_bc_i (one per boundary face) and u_bc_e (boundary face x dim),
 * the pressure closure p_bc_i and p_bc_e (one each per boundary face).  Results are
 * bit-identical to the reference's sequential loops. */

/* The face interpolation schemes the helpers take (schemes/faceinterp): Linear,
 * Upwind(flux) or LimitedLinear(flux, limiters). */
enum { FVG_INTERP_LIMITED_LINEAR = 2 };
typedef struct fvg_interp_spec {
    int32_t interp;            /* FVG_INTERP_LINEAR, FVG_INTERP_UPWIND or FVG_INTERP_LIMITED_LINEAR */
    int32_t reserved;
    const fvg_vec* flux;       /* Upwind, LimitedLinear: one entry per face */
    const fvg_vec* limiters;   /* LimitedLinear: one entry per face */
} fvg_interp_spec;

/* compute_hbya_ainv (hbya.rs:15-50): ainv_i = vol_i / a_ii and hbya_i = (mrhs_i -
 * sum_{j != i} u_j a_ij) / a_ii over row i in column order.  mlhs has one row per cell;
 * FVG_E_MISSING_ENTRY where a row has no diagonal entry (the reference's index panics). */
int fvg_compute_hbya_ainv(fvg_ctx* ctx, const fvg_mesh* mesh, const fvg_csr* mlhs, const fvg_vec* mrhs,
                          const fvg_vec* velocity, fvg_vec* hbya, fvg_vec* ainv);
/* intepolate_hbya_ainv_faces (hbya.rs:56-102): hbya interpolated to the faces and dotted
 * with the face normal, ainv interpolated; on a boundary face the owner's values, with
 * hbyan = 0 where both closures describe a wall. */
int fvg_interpolate_hbya_ainv_faces(fvg_ctx* ctx, const fvg_mesh* mesh, const fvg_vec* hbya, const fvg_vec* ainv,
                                    const fvg_interp_spec* hbya_interp, const fvg_interp_spec* ainv_interp,
                                    const fvg_vec* u_bc_i, const fvg_vec* u_bc_e, const fvg_vec* p_bc_i,
                                    const fvg_vec* p_bc_e, fvg_vec* hbyan_face, fvg_vec* ainv_face);
/* correct_phi (hbya.rs:108-157): phi = hbyan - ainv * (the pressure's face-normal gradient)
 * by FVG_FNGRAD_ORTHOGONAL, or FVG_FNGRAD_CORRECTED with fn_grads (n_cells x dim) and
 * reduction. */
int fvg_correct_phi(fvg_ctx* ctx, const fvg_mesh* mesh, const fvg_vec* hbyan_face, const fvg_vec* ainv_face,
                    const fvg_vec* pressure, int32_t fngrad, const fvg_vec* fn_grads, double reduction,
                    const fvg_vec* p_bc_i, const fvg_vec* p_bc_e, fvg_vec* phi);
/* correct_velocity (hbya.rs:161-179): velocity = hbya - ainv * pressure_gradient. */
int fvg_correct_velocity(fvg_ctx* ctx, const fvg_mesh* mesh, const fvg_vec* hbya, const fvg_vec* ainv,
                         const fvg_vec* pressure_gradient, fvg_vec* velocity);

/* ---- preconditioners: Preconditioner<T> (src/linalg/preconditioners/mod.rs:13-17) ---- */
typedef struct fvg_prec fvg_prec;

/* IdentityPreconditioner (mod.rs:22-33) */
int fvg_identity_create(fvg_ctx* ctx, int64_t n, fvg_prec** out);
/* DiagonalPreconditioner::from_matrix (mod.rs:38-68) */
int fvg_diagonal_create(fvg_ctx* ctx, const fvg_csr* a, fvg_prec** out);
/* IncompleteCholesky::from_matrix(&A, level) (incomplete_cholesky.rs:205-211).
 * Factorised on the device; level > 0 adds the reference's fill (incomplete_cholesky.rs:
 * 35-85, 148-202: a depth-first walk of the matrix graph per row).  FVG_E_NOT_SYMMETRIC
 * mirrors the reference's assert; FVG_E_MISSING_ENTRY where its fill pattern breaks the
 * transpose lookup (:138-143) that panics in the reference. */
int fvg_ic_create(fvg_ctx* ctx, const fvg_csr* a, int32_t level, fvg_prec** out);
/* IncompleteLowerUpper::from_matrix(&A, level) (incomplete_lower_upper.rs:266-282),
 * with its drop rule |v| <= 1e-6 |u_ii| (:255-258); level > 0 adds the reference's fill
 * (:30-137).  Preconditioners with level > 0 serve fvg_cg / fvg_bicgstab /
 * fvg_precondition, not batches (FVG_E_UNSUPPORTED there). */
int fvg_ilu_create(fvg_ctx* ctx, const fvg_csr* a, int32_t level, fvg_prec** out);
/* Diagnostics (no reference counterpart): the shape of the blocked sweep layouts, the
 * first `count` (at most 12) of: blocks of the lower and of the upper factor (0: not
 * built, the persistent sweep serves), dependency levels of the lower sweep, the lower
 * layout's most rows of one block in one level and most values one block fetches from
 * outside for one level, its total number of such fetches per sweep, the upper layout's
 * most rows and most fetches, the dependency slots per row of the lower and of the
 * upper layout (4 or 8), 1 if the upper layout is the lower one mirrored (then FP64
 * solves iterate in the lower layout's position order, FVG_TUNE_SOLVE_ORDER), and the
 * levels of the blocked lower layout (more than the dependency levels where it delays
 * the dependencies between workgroups, FVG_TUNE_SWEEP_GAP; 0 when not built). */
int fvg_prec_sweep_layout(const fvg_prec* prec, int32_t* info, int32_t count);
/* Re-run the numeric factorisation for new values on the same pattern. */
int fvg_prec_refactor(fvg_ctx* ctx, fvg_prec* p, const fvg_csr* a);
/* Enqueues the refactorisation without waiting (no symmetry / pivot checks yet):
 * fvg_prec_status reports them later (and synchronises). */
int fvg_prec_refactor_async(fvg_ctx* ctx, fvg_prec* p, const fvg_csr* a);
int fvg_prec_status(fvg_ctx* ctx, const fvg_prec* p);
/* Refactorises count preconditioners (each with its matrix) in one launch, one workgroup
 * per system, without the checks; fvg_prec_status reports them per preconditioner. */
int fvg_prec_refactor_many(fvg_ctx* ctx, int32_t count, fvg_prec* const* precs, const fvg_csr* const* matrices);
int fvg_prec_destroy(fvg_ctx* ctx, fvg_prec* p);
/* Preconditioner::precondition(solution, rhs): solution = M^-1 rhs. */
int fvg_precondition(fvg_ctx* ctx, const fvg_prec* p, fvg_vec* solution, const fvg_vec* rhs);
/* The factor pattern (A's at level 0; A's plus the fill above): n, its entries, and (each
 * may be NULL) row pointers (n + 1) and columns (nnz), rows sorted. */
int fvg_prec_factor_pattern(const fvg_prec* p, int64_t* n, int64_t* nnz, int32_t* row_ptr, int32_t* col);
/* Factor download (for checks): IC -> L on the lower part of the factor pattern (with the
 * diagonal) and inv_diag; ILU -> LU values on the factor pattern (dropped entries read 0). */
int fvg_prec_factor_download(fvg_ctx* ctx, const fvg_prec* p, double* values, double* inv_diag);
/* Number of levels of the lower and upper triangular dependency graphs. */
int fvg_prec_levels(const fvg_prec* p, int32_t* lower_levels, int32_t* upper_levels);
/* Count of IC pivots that were not positive (the reference yields NaN silently). */
int fvg_prec_nonpositive_pivots(fvg_ctx* ctx, const fvg_prec* p, int64_t* count);

/* ---- Krylov solvers (src/linalg/solvers/) ------------------------------------ */
enum { FVG_PRECISION_FP64 = 0, FVG_PRECISION_MIXED = 1 };

/* LinearSolverOptions (solvers/mod.rs:20-36); defaults rel 1e-6, abs 1e-6, 1000. */
typedef struct fvg_solver_options {
    double relative_tolerance;
    double absolute_tolerance;
    int64_t max_iterations;
    int32_t precision;          /* FVG_PRECISION_FP64: the reference's algorithm in FP64.
                                   FVG_PRECISION_MIXED: FP32 inner preconditioned iterations
                                   inside an FP64 iterative-refinement loop. */
    int32_t reserved;
    double inner_tolerance;     /* mixed only: the deepest relative reduction asked of one inner
                                   solve; an inner solve asks for less once the caller's
                                   tolerance is within reach (0 = default: 1e-3 for the
                                   single-system CG, 1e-4 for its BiCGStab and for batches) */
} fvg_solver_options;

/* LinearSolverInfo (solvers/mod.rs:11-17) plus counters. */
typedef struct fvg_solver_info {
    int64_t iterations;         /* FP64: reference iteration count; mixed: total inner iterations */
    double initial_residual;
    double final_residual;
    int64_t history_len;
    int64_t restarts;           /* BiCGStab restarts (bi_conjugate_gradient_stab.rs:155-169) */
    int64_t outer_iterations;   /* mixed only: refinement steps */
    double true_residual;       /* mixed only: ||b - A x|| / ||b|| in FP64 at exit */
    int32_t status;             /* batched solves: FVG_OK, or why the system was not solved
                                   (FVG_E_ZERO_PIVOT / FVG_E_NOT_SYMMETRIC from its factorisation,
                                   FVG_E_TIMEOUT when a multi-workgroup barrier wait expired) */
    int32_t flags;              /* FVG_INFO_FP64_FALLBACK: a mixed solve whose FP32 iterations stalled
                                   needed FP64 iterations (the FP64 solver from the refined solution,
                                   or FP64 inner iterations in multi-workgroup batches);
                                   FVG_INFO_POSITION_ORDER: an FP64 solve iterated with its vectors in
                                   the lower sweep's position order (FVG_TUNE_SOLVE_ORDER);
                                   FVG_INFO_FUSED_SWEEPS: an FP64 CG did its residual update and dot
                                   products inside the triangular sweeps (FVG_TUNE_CG_FUSION) */
} fvg_solver_info;
enum { FVG_INFO_FP64_FALLBACK = 1, FVG_INFO_POSITION_ORDER = 2, FVG_INFO_FUSED_SWEEPS = 4 };

/* conjugate_gradient (solvers/conjugate_gradient.rs:12-112).  x is the initial
 * guess on entry and the solution on exit.  history (may be NULL) receives up to
 * history_capacity residuals. */
int fvg_cg(fvg_ctx* ctx, const fvg_csr* a, const fvg_prec* p, const fvg_vec* b, fvg_vec* x,
           const fvg_solver_options* opts, fvg_solver_info* info, double* history, int64_t history_capacity);
/* bi_conjugate_gradient_stab (solvers/bi_conjugate_gradient_stab.rs:13-179), right
 * preconditioned, including its residual scaling in the loop test (:139, :171). */
int fvg_bicgstab(fvg_ctx* ctx, const fvg_csr* a, const fvg_prec* p, const fvg_vec* b, fvg_vec* x,
                 const fvg_solver_options* opts, fvg_solver_info* info, double* history,
                 int64_t history_capacity);

/* compute_wall_distance (tools/wall_distance.rs:8-119): the p-Poisson (p = 3) wall
 * distance.  wall (host, one byte per boundary face in boundary-face order, may be NULL
 * when the mesh has none) marks the wall faces: closure (0, 0); every other boundary face
 * is zero-gradient, (1, 0).  distance holds the first solve's initial guess on entry and
 * the normalised distance on exit.  Each of up to max_iterations rounds assembles
 * -laplacian(Corrected(grad, 1), visc) + source(-1), factorises IC(1) and runs fvg_cg with
 * opts (NULL: the reference's defaults); it stops after a solve whose initial residual is
 * below tolerance, else takes the new gradients and the face diffusivity |grad_f|^(p-2).
 * info (may be NULL) receives the last solve's record, solves (may be NULL) their count;
 * a solver or factorisation error ends the loop with that status. */
int fvg_compute_wall_distance(fvg_ctx* ctx, const fvg_mesh* mesh, const uint8_t* wall, int32_t max_iterations,
                              double tolerance, const fvg_solver_options* opts, fvg_vec* distance,
                              fvg_solver_info* info, int32_t* solves);

/* FP64 fvg_cg / fvg_bicgstab that allocate nothing: no counterpart in the reference,
 * for callers that cannot allocate device memory (code under stream capture, tight
 * serving loops).  Same arguments, kernels, iterations and results as fvg_cg /
 * fvg_bicgstab with precision FVG_PRECISION_FP64 (other precisions give
 * FVG_E_UNSUPPORTED).  Work vectors, partial sums and solver state live in one static
 * device array of fvg_noalloc_scratch_bytes() bytes (160 MiB), which each device the
 * library runs them on holds from the first call there until the process ends: a CG on
 * up to about 5.2M rows, a BiCGStab on up to about 2.6M rows; a larger system (or a
 * history beyond the array) gives FVG_E_INVALID.  Every call on a device shares that
 * device's array: calls that overlap on one device race.  The preconditioner, matrix and
 * vectors are the caller's, set up beforehand; kernel timing is off during the call. */
int fvg_cg_noalloc(fvg_ctx* ctx, const fvg_csr* a, const fvg_prec* p, const fvg_vec* b, fvg_vec* x,
                   const fvg_solver_options* opts, fvg_solver_info* info, double* history, int64_t history_capacity);
int fvg_bicgstab_noalloc(fvg_ctx* ctx, const fvg_csr* a, const fvg_prec* p, const fvg_vec* b, fvg_vec* x,
                         const fvg_solver_options* opts, fvg_solver_info* info, double* history,
                         int64_t history_capacity);
int fvg_noalloc_scratch_bytes(uint64_t* bytes);

/* ---- batched solves (no reference counterpart: many independent systems) ----- */
/* A batch holds n independent systems A_k x_k = b_k, each with its own matrix and
 * preconditioner (identity, Jacobi, IC(0) or ILU(0)), solved by one kernel launch in
 * which one workgroup runs one system's whole solve.  FP64 batches run the reference
 * algorithm of fvg_cg / fvg_bicgstab per system; mixed batches run FP32 inner
 * iterations inside FP64 refinement.  Each system's result is independent of the rest
 * of the batch.  Matrices and preconditioners stay owned by the caller and must
 * outlive the batch; values assembled or refactored later are picked up at the next
 * launch.  A batch launches on a stream of its own, after all work enqueued on the
 * context stream so far, so batches overlap on the device. */
enum { FVG_SOLVER_CG = 0, FVG_SOLVER_BICGSTAB = 1 };
typedef struct fvg_batch fvg_batch;
int fvg_batch_create(fvg_ctx* ctx, int32_t n, const fvg_csr* const* a, const fvg_prec* const* p, int32_t solver,
                     const fvg_solver_options* opts, fvg_batch** out);
int fvg_batch_destroy(fvg_ctx* ctx, fvg_batch* batch);
/* Points a finished (or never launched) batch at another set of n systems, keeping its
 * stream, captured launch and storage (grown when n exceeds what it held).  A server
 * reuses one batch per lane this way. */
int fvg_batch_rebind(fvg_ctx* ctx, fvg_batch* batch, int32_t n, const fvg_csr* const* a, const fvg_prec* const* p);
/* Sizes a batch's per-system storage for up to capacity systems now, so later rebinds
 * never reallocate (reallocation synchronises the device). */
int fvg_batch_reserve(fvg_ctx* ctx, fvg_batch* batch, int32_t capacity);
int fvg_batch_size(const fvg_batch* batch, int32_t* n_systems, int64_t* total_rows);
/* Enqueues the solve of every system (x_k: initial guess in, solution out). */
int fvg_batch_launch(fvg_ctx* ctx, fvg_batch* batch, const fvg_vec* const* b, fvg_vec* const* x);
/* 1 in *done once a launched batch has finished (does not block). */
int fvg_batch_query(fvg_ctx* ctx, fvg_batch* batch, int32_t* done);
/* Completion of each system of a launched batch, without waiting: done[k] = 1 once system
 * k's solution is in its x vector (readable from any stream) and, if infos is not NULL,
 * infos[k] holds its result.  Systems finish independently of the rest of the batch. */
int fvg_batch_poll(fvg_ctx* ctx, fvg_batch* batch, int32_t* done, fvg_solver_info* infos);
/* Waits for the launched batch; infos[k] and histories[k] (each may be NULL) per system. */
int fvg_batch_finish(fvg_ctx* ctx, fvg_batch* batch, fvg_solver_info* infos, double* const* histories,
                     int64_t history_capacity);
/* launch + finish */
int fvg_batch_solve(fvg_ctx* ctx, fvg_batch* batch, const fvg_vec* const* b, fvg_vec* const* x,
                    fvg_solver_info* infos, double* const* histories, int64_t history_capacity);
/* Threads per system (256, 512 or 1024; default 512) of the one-workgroup-per-system
 * kernel (FP64 batches, and mixed batches with rows_per_group 0). */
int fvg_batch_set_threads(fvg_ctx* ctx, fvg_batch* batch, int32_t threads);
int fvg_batch_get_threads(const fvg_batch* batch, int32_t* threads);
/* Mixed-precision batches: each system is solved by ceil(n / rows_per_group) workgroups
 * (at most 64) that meet at device-wide barriers (default 4096 rows per workgroup;
 * 0 = one workgroup per system, as FP64 batches always are).  The inner iterations then
 * apply IC(0) / ILU(0) by Jacobi-iterated triangular solves (FVG_TUNE_INNER_SWEEPS of
 * the context), and an FP32 inner solve that stops progressing is continued in FP64
 * (FVG_INFO_FP64_FALLBACK).  max_workers bounds the workgroups the launch keeps
 * resident (0: as many as fit the device; callers running several batches at once give
 * each a share).  A result depends on its system and its workgroup count only.
 * No counterpart in the reference. */
int fvg_batch_set_cluster(fvg_ctx* ctx, fvg_batch* batch, int32_t rows_per_group, int32_t max_workers);
/* Per size bucket (mixed-precision clusters; the autotuner's choice per system size): a
 * system of at most max_rows[i] rows (ascending) is solved by ceil(n / rows_per_group[i])
 * workgroups, the first bucket that holds it deciding; larger systems keep the batch's
 * rows_per_group (fvg_batch_set_cluster).  count 0 removes the buckets.  No counterpart in
 * the reference. */
int fvg_batch_set_cluster_buckets(fvg_ctx* ctx, fvg_batch* batch, int32_t count, const int64_t* max_rows,
                                  const int32_t* rows_per_group);
/* The rows per workgroup in effect (0: one workgroup per system) and the worker pool a
 * launch uses (a system never gets more workgroups than the pool). */
int fvg_batch_get_cluster(const fvg_batch* batch, int32_t* rows_per_group, int32_t* workers);
/* Phase timing inside each system's workgroup (diagnostic; off by default).  After a
 * finished launch, us[4k + j] holds system k's microseconds in SpMV (j=0), lower sweeps
 * (1), upper sweeps (2) and its whole solve (3). */
int fvg_batch_set_phase_timing(fvg_ctx* ctx, fvg_batch* batch, int32_t enable);
int fvg_batch_phase_times(fvg_ctx* ctx, fvg_batch* batch, double* us);

/* ---- device timing of the most recent solve ------------------------------- */
typedef struct fvg_kernel_stats {
    double spmv_ms;             /* total time of the SpMV kernels */
    double sweep_ms;            /* total time of the triangular solves: each one's set-up kernel (pending
                                   fill, right-hand side in sweep order), its sweep (all levels) and
                                   the scatter of its values to their rows */
    int64_t spmv_launches;
    int64_t sweep_launches;     /* kernel launches spent in sweeps */
    int64_t spmv_calls;         /* SpMV products timed */
    int64_t sweep_calls;        /* whole triangular sweeps timed */
} fvg_kernel_stats;
/* HIP-event timing of SpMV / sweep kernels on this context: 0 off; N > 0 times one of
 * every N launch groups (an SpMV, a triangular solve) with an event pair on the context
 * stream, so that sampled timing barely slows the stream (1: every group). */
int fvg_ctx_enable_timing(fvg_ctx* ctx, int enable);
int fvg_ctx_kernel_stats(fvg_ctx* ctx, fvg_kernel_stats* out, int reset);

#ifdef __cplusplus
}
#endif
#endif /* FVG_H */
