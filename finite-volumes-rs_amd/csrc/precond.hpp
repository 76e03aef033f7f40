// precond.hpp -- preconditioner objects (internal).
#pragma once

#include "internal.hpp"

namespace fvg {

// One triangular sweep x = scale .* (b - T x) over a strictly triangular T, stored
// in level order: positions are rows sorted by (dependency level, row index), so
// each level is a contiguous range and its rows' entries are contiguous in HBM.
struct Sweep {
    int64_t n = 0, nnz = 0;
    int nlevels = 0;
    std::vector<int64_t> lvl_ptr;       // host: level l holds positions [lvl_ptr[l], lvl_ptr[l+1])
    int32_t* rows = nullptr;            // row of each position
    int32_t* ptr = nullptr;             // entry range of each position (n+1)
    int32_t* col = nullptr;             // column of each entry, ascending within a row
    int32_t* src = nullptr;             // slot of each entry in the factor value array
    double* val = nullptr;              // entry values gathered from the factor
};

enum PrecKind { kIdentity = 0, kDiagonal = 1, kIC = 2, kILU = 3 };

}  // namespace fvg

struct fvg_prec {
    int kind = 0;
    int64_t n = 0, nnz = 0;
    // pattern of the matrix the preconditioner was built from
    std::vector<int32_t> h_rp, h_col, h_diag, h_tpos;
    int32_t* d_rp = nullptr;
    int32_t* d_col = nullptr;
    int32_t* d_diag = nullptr;
    int32_t* d_tpos = nullptr;          // IC: slot of (j,i) for each slot (i,j), -1 if j >= n
    double* d_fac = nullptr;            // factor values on the matrix pattern
    double* d_inv_diag = nullptr;       // IC 1/l_ii, ILU 1/u_ii, Jacobi 1/a_ii
    double* d_tmp = nullptr;            // intermediate vector of the two sweeps
    int32_t* d_flags = nullptr;         // [0] zero pivots, [1] non-positive IC pivots, [2] asymmetry
    fvg::Sweep lower, upper;
};

namespace fvg {
// solution = M^-1 rhs on the context stream (no host synchronisation).
// skip: optional device flag; when *skip != 0 the launches do nothing.
int apply_prec(fvg_ctx* ctx, const fvg_prec* p, const double* rhs, double* solution, const int32_t* skip);
}  // namespace fvg
