// krylov.hpp -- Krylov solver state and drivers shared by the single and batched paths (internal).
#pragma once

#include "internal.hpp"
#include "kstate.hpp"

struct fvg_prec;

namespace fvg {

// How the rows of the system split into members.  Default: one member owning all rows.
struct Layout {
    int n_inst = 1;
    const int32_t* blk_inst = nullptr;   // member of each 256-row workgroup (device)
    const int32_t* blk_ptr = nullptr;    // member i owns workgroups [blk_ptr[i], blk_ptr[i+1]) (device)
};

// Device state of one solve over one or more members.
struct KRun {
    fvg_ctx* c = nullptr;
    int n_inst = 0;
    KState* st = nullptr;
    double* hist = nullptr;
    int64_t hist_cap = 0;
    int32_t* flags = nullptr;   // [0] members still iterating, [1] all done
    bool owned = true;          // st, hist and flags were allocated by init()
    int32_t path = 0;           // FVG_INFO_POSITION_ORDER / FVG_INFO_FUSED_SWEEPS of the run
    int init(fvg_ctx* ctx, int ninst, const fvg_solver_options* o);
    // the same state in caller-provided device memory (nothing allocated, nothing freed):
    // ninst states, ninst * (max_iterations + 2) history slots, 4 flags
    int init_at(fvg_ctx* ctx, int ninst, const fvg_solver_options* o, KState* st, double* hist, int32_t* flags);
    int all_done(bool* out);
    // poll number `batch` of the loop (see krylov.hip): *done is the flag of the poll before
    int poll(int64_t batch, bool* done);
    int finish(fvg_solver_info* infos, double* const* histories, int64_t cap);
    ~KRun();
};

// Work vectors and partial sums of one solve, when the caller provides them (no
// allocation): vec[0..3] for CG, vec[0..7] for BiCGStab, each of n doubles; partials
// holds kPartialSlots slots of `pstride` >= spmv_blocks(n) doubles.
struct KBuffers {
    double* vec[8] = {};
    double* partials = nullptr;
    int64_t pstride = 0;
};

int run_cg_multi(fvg_ctx* c, const fvg_csr* a, const fvg_prec* p, const double* b, double* x, const Layout& L,
                 KRun& K, const KBuffers* kb = nullptr);
int run_bicgstab_multi(fvg_ctx* c, const fvg_csr* a, const fvg_prec* p, const double* b, double* x, const Layout& L,
                       KRun& K, const KBuffers* kb = nullptr);
// checks of the solver entry points' arguments
int check_solve_args(fvg_ctx* c, const fvg_csr* a, const fvg_prec* p, const fvg_vec* b, const fvg_vec* x,
                     const fvg_solver_options* o);

}  // namespace fvg
