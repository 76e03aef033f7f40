// precond.hpp -- preconditioner objects (internal).
#pragma once

#include <functional>
#include <vector>

#include "internal.hpp"

namespace fvg {

// One triangular sweep x = scale .* (b - T x) over a strictly triangular T, stored
// in level order: positions are rows sorted by (dependency level, row index), each
// level padded to a whole number of kChunk-row chunks (padding rows are -1), so a
// level is a contiguous range of chunks and its rows' entries are contiguous in HBM.
constexpr int kChunk = 64;              // rows per wave64 chunk

// Coordinates of the rows (cell centres, `dim` floats per row) for the layout analysis,
// and their source, called only where the analysis uses them (it may copy them from a
// mesh first).
struct RowCoords {
    const float* v = nullptr;
    int dim = 0;
};
using CoordSource = std::function<RowCoords()>;

// Block and level of every row of a blocked sweep layout (host).
struct BlockAssign {
    int32_t nb = 0;                     // blocks
    int nlevels = 0;                    // levels (more than the sweep's when delayed)
    int kmax = 0;                       // dependency slots per row (4 or 8)
    std::vector<int32_t> blk, lev;      // per row
};

struct Sweep {
    int64_t n = 0, npos = 0, nnz = 0;
    int nlevels = 0;
    std::vector<int64_t> lvl_ptr;       // host: level l holds positions [lvl_ptr[l], lvl_ptr[l+1])
    int64_t* d_lvl_ptr = nullptr;       // device copy
    int32_t* rows = nullptr;            // row of each position, -1 for padding (npos)
    int32_t* ptr = nullptr;             // entry range of each position (npos+1)
    // col, val and val32 hold nnz + 1 entries: a padding entry (column 0, value 0) follows
    // the last row, so clamped loads of a row with no entries stay in bounds
    int32_t* col = nullptr;             // column of each entry, ascending within a row
    int32_t* src = nullptr;             // slot of each entry in the factor value array (nnz)
    double* val = nullptr;              // entry values gathered from the factor
    float* val32 = nullptr;             // FP32 copy for the mixed-precision inner solver
    // Blocked layout (k_sweep_blocked; nblk == 0 when not built).  Rows are grouped into
    // blocks of whole gap-1 components (rows joined by dependencies one level apart), so
    // every dependency between blocks spans at least two levels.  Positions run block by
    // block, level by level; a block's rows of one level are consecutive.
    int32_t nblk = 0;
    int32_t bL = 0;                     // levels of the layout (more than nlevels when delayed)
    int32_t bW = 0;                     // most rows of one block in one level
    int32_t bX = 0;                     // most values one block fetches from outside for one level
    int32_t bK = 0;                     // most dependencies of one row
    int32_t bF = 0;                     // most fetched rows of one block (all levels)
    mutable int32_t bOK = -1;           // blocked launch fits co-resident: -1 unknown, 0 no, 1 yes
    mutable int32_t bOKfuse = -1;       // the same for the kernels with fused CG work
                                        // (decided on first use, cached)
    int32_t* brange = nullptr;          // per block: first level, one past the last (2 x nblk)
    int32_t* blvl = nullptr;            // nblk x (nlevels + 1): first position of each level
    int32_t* bxptr = nullptr;           // nblk x (nlevels + 1): offsets into bxidx
    int32_t* bxidx = nullptr;           // positions fetched from outside the block's value ring
    int64_t nbx = 0;                    // length of bxidx
    int32_t* brow = nullptr;            // row of each position (n)
    uint16_t* bdep = nullptr;           // n x bK: LDS slot of each dependency, ascending columns, then
                                        // kNoSlot (a row's bK slots are one or two 64-bit words)
    int32_t* bsrc = nullptr;            // n x bK: factor slot of each dependency (-1: none)
    double* bval = nullptr;             // n x bK: its coefficient (0 for none)
    double* spos = nullptr;             // scale (inverse diagonal) in position order, per refactor
    double* bpos = nullptr;             // right-hand side in position order, per call
    int32_t* bfill = nullptr;           // positions some block fetches from outside, ascending (nfill)
    int64_t nfill = 0;                  // only they need the pending fill before a call
    double* xpos = nullptr;             // the sweep's values in position order (n), scattered after
    // Upper layout of a factor whose lower layout is blocked too: the lower layout's
    // position of the row at each upper position, so the upper set-up gathers its
    // right-hand side straight from the lower sweep's xpos (no scatter in between).
    int32_t* bfrom = nullptr;           // n, or null
    // Lower layout paired with a blocked upper one: the position of each row (n), so a
    // solver can write the sweep's right-hand side in position order itself.
    int32_t* bposof = nullptr;
    std::vector<int32_t> h_brow;        // row of each position (host copy, freed once paired)
    BlockAssign h_assign;               // the blocked layout's assignment (host, freed once paired)
    // Mirrored upper layout: the lower layout's blocks, levels and rows reversed, so upper
    // position q holds the row of lower position n - 1 - q (structurally symmetric patterns).
    bool mirrored = false;
    // per position: 1 if some block polls it (it must hold the pending pattern before a
    // sweep); lower layouts of a mirrored pair only
    uint8_t* bpolled = nullptr;
};
constexpr int kBlockRing = 3;           // levels of values a block keeps in LDS
constexpr int kBlockGap = 3;            // least levels between blocks of a blocked layout (default)
constexpr uint16_t kNoSlot = 0xFFFF;    // an unused dependency slot of the blocked layout

// One strict triangular factor as a CSR in natural row order with FP32 values, for the
// mixed-precision inner solver's Jacobi-iterated apply.
struct Tri32 {
    int64_t n = 0, nnz = 0;
    int32_t* ptr = nullptr;    // n + 1
    int32_t* col = nullptr;    // nnz, ascending within a row
    int16_t* dcol = nullptr;   // nnz: col minus the first row of the row's 256-row tile, where
                               // every such offset fits 16 bits (else null)
    int32_t* src = nullptr;    // factor slot of each entry
    float* val = nullptr;      // nnz, gathered from the factor after each refactorisation
};

enum PrecKind { kIdentity = 0, kDiagonal = 1, kIC = 2, kILU = 3 };

// Jacobi sweeps per triangular factor when FP32 inner iterations apply an incomplete
// factor inexactly (FVG_TUNE_INNER_SWEEPS = 0)
constexpr int kInnerSweeps = 2;

}  // namespace fvg

struct fvg_prec {
    int kind = 0;
    int level = 0;                      // fill level of IC / ILU
    int64_t n = 0, nnz = 0;             // nnz: entries of the factor pattern (the matrix's at level 0)
    int64_t a_nnz = 0;                  // entries of the matrix pattern it was built from
    std::vector<int32_t> h_arp, h_acol; // level > 0: that matrix pattern (the factor's is h_rp / h_col)
    int32_t* d_amap = nullptr;          // level > 0: factor slot of each matrix slot (-1: none)
    double* d_afill = nullptr;          // level > 0: the matrix values on the factor pattern
    // pattern of the matrix the preconditioner was built from
    std::vector<int32_t> h_rp, h_col, h_diag, h_tpos;
    int32_t* d_rp = nullptr;
    int32_t* d_col = nullptr;
    int32_t* d_diag = nullptr;
    int32_t* d_tpos = nullptr;          // IC: slot of (j,i) for each slot (i,j), -1 if j >= n
    double* d_fac = nullptr;            // factor values on the matrix pattern
    double* d_inv_diag = nullptr;       // IC 1/l_ii, ILU 1/u_ii, Jacobi 1/a_ii
    double* d_tmp = nullptr;            // intermediate vector of the two sweeps
    int32_t* d_flags = nullptr;         // [0] zero pivots, [1] non-positive IC pivots, [2] asymmetry,
                                        // [3] sweep wait timed out
    fvg::Sweep lower, upper;
    // FP32 copies used by the mixed-precision inner iterations (made on first use)
    float* d_inv_diag32 = nullptr;
    float* d_tmp32 = nullptr;
    float* d_tmp32b = nullptr;          // second FP32 temporary (Jacobi-iterated sweeps)
    fvg::Tri32 jlo, jup;                // FP32 strict factors in natural row order (inexact apply)
    uint64_t fac_version = 0, fac32_version = 0;
    void* team = nullptr;               // batched-solve plan of this system (team.hip), made on use
    // the matrix in the lower layout's position order (position-order solves; made on use)
    // (a cache: made and refreshed by solves that take the preconditioner as const)
    mutable fvg_csr* pa = nullptr;
    mutable int32_t* pa_src = nullptr;          // slot of the original matrix of each of its entries
    mutable const fvg_csr* pa_of = nullptr;     // the matrix whose pattern was checked and whose values it holds
    mutable uint64_t pa_ver = ~0ull;            // version of those values
    int32_t max_row = -1;               // longest factor row (set on first factorisation)
    uint64_t pattern_id = 0;            // mesh pattern the analysis was made for (0: not a mesh pattern)
};

namespace fvg {
// solution = M^-1 rhs on the context stream (no host synchronisation).
// skip: optional device flag; when *skip != 0 the launches do nothing.
int apply_prec(fvg_ctx* ctx, const fvg_prec* p, const double* rhs, double* solution, const int32_t* skip);
// FVG_E_TIMEOUT if a sweep wait gave up since the last check (synchronises).
int check_sweep_status(fvg_ctx* ctx, const fvg_prec* p);
// FP32 factor copies for the mixed-precision inner solver (refreshed after a refactor).
int ensure_prec32(fvg_ctx* ctx, fvg_prec* p);
// solution = M^-1 rhs with FP32 factors, vectors and arithmetic.
// Optional dot product w . solution fused into an FP32 apply: `count` partials expected (one
// per 256 rows, at partials[0 .. count)); *done tells whether the apply left them.
struct DotFuse {
    const float* w = nullptr;
    double* part = nullptr;
    int64_t count = 0;
    bool* done = nullptr;
};
int apply_prec32(fvg_ctx* ctx, const fvg_prec* p, const float* rhs, float* solution, const int32_t* skip,
                 const DotFuse& dot = DotFuse());
// frees the batched-solve plan cached on p (team.hip)
void team_plan_release(fvg_ctx* ctx, fvg_prec* p);

// The FP64 solve in the lower sweep's position order (IC / ILU with a mirrored pair of
// blocked layouts): the solver's vectors are indexed by lower position; the lower sweep
// reads its right-hand side from a position-order vector, the upper sweep reads the lower
// one's values in reverse, and the result sits reversed in up_x (z at lower position q is
// up_x[n - 1 - q]).  Before each apply the caller writes the pending pattern where polled:
// lo_x[q] where lo_poll[q], up_x[q] where up_poll[q].
// Work of the CG iteration fused into the two sweeps of a position-order apply: the lower
// sweep forms its right-hand side r - ap*alpha itself (writing r, x += p*alpha and the
// partials of r.r), the upper sweep leaves the partials of r.z (conjugate_gradient.rs:
// 77-94).  One double-double partial per block of the sweep.
enum { kFuseNone = 0, kFuseXR = 1, kFuseRZ = 2 };
struct KState;
struct SweepFuse {
    const double* alpha = nullptr;   // kFuseXR: the step length in the solver state
    const double* ap = nullptr;      // kFuseXR: A p, by position
    double* x = nullptr;             // kFuseXR
    const double* pp = nullptr;      // kFuseXR: p
    double* r = nullptr;             // kFuseXR: read and written; kFuseRZ: read
    double* partials = nullptr;
    // with `counter`, the last block to finish sums the partials and runs the solver's
    // scalar step `step` (the convergence test after the lower sweep, beta after the upper)
    int32_t* counter = nullptr;      // zero before the sweep; left zero after it
    KState* st = nullptr;
    double* hist = nullptr;
    int32_t* active = nullptr;
    int32_t* all_done = nullptr;
    int32_t step = 0;
};

struct PosSolve {
    int64_t n = 0;
    const int32_t* row = nullptr;       // row of each lower position
    double* lo_x = nullptr;
    double* up_x = nullptr;
    const uint8_t* lo_poll = nullptr;
    const uint8_t* up_poll = nullptr;
    const fvg_csr* pa = nullptr;        // the matrix in position order
};
// *ok: whether p preconditions a this way (same pattern, layouts mirrored, blocked sweeps
// selected).  Makes the position-order matrix if needed (when allow_alloc) and refreshes its
// values when a's have changed.
int pos_solve(fvg_ctx* ctx, const fvg_prec* p, const fvg_csr* a, bool allow_alloc, PosSolve* out, bool* ok);
// z = M^-1 r, r in position order: both sweeps (no host synchronisation)
int apply_pos(fvg_ctx* ctx, const fvg_prec* p, const double* r_pos, const int32_t* skip);
// apply_pos with the CG iteration's vector work and scalar steps fused into the sweeps
int apply_pos_fused(fvg_ctx* ctx, const fvg_prec* p, double* r_pos, const int32_t* skip, const SweepFuse& lo_f,
                    const SweepFuse& up_f);
// whether the fused sweep kernels can run this preconditioner's layouts co-resident
bool pos_fusable(fvg_ctx* ctx, const fvg_prec* p);
// pending value bits of an FP64 sweep output
constexpr uint64_t kPendingBits = 0x7FF4DEADBEEF0F64ull;
}  // namespace fvg
