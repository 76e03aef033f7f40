// internal.hpp -- shared definitions of libfvg (not part of the public ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/fvg.h"

namespace fvg {

// ---- error reporting ------------------------------------------------------------
void set_error(const std::string& msg);
int fail(int status, const std::string& msg);

#define FVG_HIP(expr)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return ::fvg::fail(FVG_E_HIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

#define FVG_TRY(expr)              \
    do {                           \
        int s_ = (expr);           \
        if (s_ != FVG_OK) return s_; \
    } while (0)

// ---- launch geometry ------------------------------------------------------------
constexpr int kBlock = 256;                 // threads per workgroup (4 wave64)
constexpr int kMaxReduceBlocks = 2048;      // partial sums per reduction
constexpr int kPartialSlots = 8;            // slots of partials (a double-double slot takes two)
inline int grid_for(int64_t n, int per_block = kBlock) {
    int64_t g = (n + per_block - 1) / per_block;
    return (int)(g < 1 ? 1 : g);
}

}  // namespace fvg

// ---- public handle definitions -----------------------------------------------------
struct fvg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    uint64_t budget = 0;       // 0 = unlimited
    uint64_t in_use = 0;
    double* partials = nullptr;        // kPartialSlots slots of partials_cap doubles of reduction scratch
    int64_t partials_cap = 0;
    double* scalar = nullptr;          // small device scalar area
    void* pinned = nullptr;            // pinned host staging (4 KiB)
    int32_t* poll_host = nullptr;      // pinned: the two slots of the solvers' done-flag polls
    hipEvent_t poll_evt[2] = {nullptr, nullptr};   // recorded after each slot's copy
    void* stage = nullptr;             // pinned host staging for vector transfers (grown on demand)
    size_t stage_cap = 0;
    void* jobs = nullptr;              // device table of multi-system launches (grown on demand)
    void* jobs_host = nullptr;         // its pinned host twin
    size_t jobs_cap = 0;
    hipEvent_t jobs_evt = nullptr;     // recorded after the last copy out of the twin
    bool timing = false;
    struct Span { hipEvent_t a, b; int kind; int64_t launches; };
    std::vector<Span> spans;           // event pairs recorded while timing is enabled
    std::vector<hipEvent_t> event_pool;   // events of collected spans, reused
    int timing_every = 1;              // time one launch group in this many
    int timing_skip = 0;               // groups to skip before the next timed one
    fvg_kernel_stats stats{};
    int64_t tune_sweep_mode = 0;       // FVG_TUNE_SWEEP_MODE
    int64_t tune_sweep_waves = 0;      // FVG_TUNE_SWEEP_WAVES
    int64_t tune_sweep_poll = 0;       // FVG_TUNE_SWEEP_POLL
    int64_t tune_inner_exact = 0;      // FVG_TUNE_INNER_EXACT
    int64_t tune_inner_sweeps = 0;     // FVG_TUNE_INNER_SWEEPS
    int64_t tune_graphs = 0;           // FVG_TUNE_GRAPHS
    int64_t tune_sweep_blocks = 0;     // FVG_TUNE_SWEEP_BLOCKS
    int64_t tune_spin_limit = 0;       // FVG_TUNE_SPIN_LIMIT
    int64_t tune_sweep_launch = 0;     // FVG_TUNE_SWEEP_LAUNCH
    int64_t tune_withhold = 0;         // FVG_TUNE_SWEEP_WITHHOLD
    int64_t tune_solve_order = 0;      // FVG_TUNE_SOLVE_ORDER
    int64_t tune_assemble = 0;         // FVG_TUNE_ASSEMBLE
    int64_t tune_cg_fusion = 0;        // FVG_TUNE_CG_FUSION
    int64_t tune_sweep_gap = 0;        // FVG_TUNE_SWEEP_GAP
    int64_t tune_sweep_boxes = 0;      // FVG_TUNE_SWEEP_BOXES
    // work vectors of finished solves, kept for the next solve on this context's stream
    // (length in doubles); released when an allocation would not fit otherwise
    std::vector<std::pair<double*, int64_t>> work_cache;
    int batches_in_flight = 0;         // batches of this context launched and not yet finished
    bool sweep_timed_out = false;      // a sweep wait gave up here: co-resident launches stay cooperative
    int last_sweep_launch = 0;         // 1 ordinary, 2 cooperative (0: none yet)
    bool counted = false;              // counted among the device's live contexts
    int num_cus = 0;
};

struct fvg_vec {
    int64_t n = 0;
    double* d = nullptr;
    bool owned = true;
};

struct fvg_csr {
    int64_t nrows = 0, ncols = 0, nnz = 0;
    int32_t* row_ptr = nullptr;
    int32_t* col = nullptr;
    double* val = nullptr;
    // host copy of the pattern, used by the preconditioner analysis
    std::vector<int32_t> h_row_ptr, h_col;
    uint64_t pattern_id = 0;           // identifies a mesh pattern (assembly target check)
    uint64_t version = 0;              // bumped whenever the values change
    float* val32 = nullptr;            // FP32 copy of the values for mixed precision (made on use)
    uint64_t val32_version = ~0ull;
    // the columns as 16-bit offsets from the first row of each 256-row tile, for the FP32
    // SpMV (made on use where every offset fits; state 0 not tried, 1 made, -1 does not fit)
    int16_t* dcol32 = nullptr;
    int32_t dcol32_state = 0;
    // coordinates of the rows (coord_dim per row, 0: none), used by the preconditioner
    // analysis to shape its blocked layouts; a matrix made from a mesh instead names the
    // mesh, whose cell centres the analysis reads while it is alive
    std::vector<float> h_coords;
    int32_t coord_dim = 0;
    const fvg_mesh* coord_mesh = nullptr;
    uint64_t coord_mesh_id = 0;
};

namespace fvg {

// ---- device memory accounted against the context budget ---------------------------
int dev_alloc(fvg_ctx* ctx, size_t bytes, void** out);
// frees the context's cached solver work vectors (after its stream has drained)
void release_work_cache(fvg_ctx* ctx);
void dev_free(fvg_ctx* ctx, void* p, size_t bytes);
template <class T>
int dev_alloc_n(fvg_ctx* ctx, int64_t n, T** out) {
    return dev_alloc(ctx, (size_t)(n > 0 ? n : 1) * sizeof(T), (void**)out);
}
template <class T>
void dev_free_n(fvg_ctx* ctx, T*& p, int64_t n) {
    if (p) dev_free(ctx, (void*)p, (size_t)(n > 0 ? n : 1) * sizeof(T));
    p = nullptr;
}
template <class T>
int upload(fvg_ctx* ctx, T* dst, const T* src, int64_t n) {
    if (n <= 0) return FVG_OK;
    FVG_HIP(hipMemcpyAsync(dst, src, (size_t)n * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
    return FVG_OK;
}
template <class T>
int download(fvg_ctx* ctx, T* dst, const T* src, int64_t n) {
    if (n <= 0) return FVG_OK;
    FVG_HIP(hipMemcpyAsync(dst, src, (size_t)n * sizeof(T), hipMemcpyDeviceToHost, ctx->stream));
    FVG_HIP(hipStreamSynchronize(ctx->stream));
    return FVG_OK;
}

// ---- shared kernels (defined in vector_ops.hip) ------------------------------------
// deterministic dot product: partials -> fixed-order final sum into *out_dev
int launch_dot(fvg_ctx* ctx, int64_t n, const double* a, const double* b, double* out_dev);
int launch_fill(fvg_ctx* ctx, int64_t n, double* x, double value);
int launch_copy(fvg_ctx* ctx, int64_t n, const double* src, double* dst);
int read_scalars(fvg_ctx* ctx, const double* dev, double* host, int count);

// timing of launch groups (kind: 0 SpMV, 1 sweep); no-ops unless timing is enabled
enum { kSpanSpmv = 0, kSpanSweep = 1 };
int timing_begin(fvg_ctx* ctx);
void timing_end(fvg_ctx* ctx, int token, int kind, int64_t launches);

// Launch of a kernel whose workgroups wait on one another and that the caller sized to
// be resident together: ordinary or cooperative as sweep_launch_rule decides.
hipError_t launch_coresident(fvg_ctx* ctx, const void* fn, unsigned groups, unsigned threads, void** args,
                             size_t lds);
// 1 (ordinary) or 2 (cooperative) for FVG_TUNE_SWEEP_LAUNCH = tune, the live contexts of the
// device in this process, this context's batches in flight and whether a wait of this
// context gave up before
int sweep_launch_rule(int64_t tune, int live_contexts, int batches_in_flight, bool timed_out_before);

}  // namespace fvg
