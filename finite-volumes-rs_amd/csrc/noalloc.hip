// noalloc.hip -- FP64 conjugate_gradient / bi_conjugate_gradient_stab that allocate nothing.
//
// The solvers of krylov.hip allocate their work vectors and solver state per call.  These
// entry points run the same kernels with every buffer carved out of one file-scope device
// array instead, for callers that cannot allocate (code under stream capture, tight
// serving loops).  The array is a per-device copy held by the process from the first call
// on a device until it ends; its address is looked up once per device (the array's name
// is a host-side stand-in that must never reach a kernel).  All calls on one device share
// it, so calls that overlap on one device race; every region is written before it is read,
// so nothing needs clearing between calls.
#include <mutex>
#include <string>

#include "krylov.hpp"
#include "precond.hpp"
#include "spmv.hpp"

namespace fvg {

// 160 MiB: the eight work vectors of a BiCGStab on up to 2.6M rows, or the four of a CG
// on up to 5.2M rows, with the partial sums, the solver state and the residual history.
constexpr size_t kScratchBytes = size_t(160) << 20;
static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

constexpr int kScratchDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch[kScratchDevices];   // the array's device address, per device

// The array's address on the context's device, looked up on the first call there with
// that device current.
static int scratch_base(fvg_ctx* c, double** out) {
    if (c->device < 0 || c->device >= kScratchDevices) return fail(FVG_E_INVALID, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[c->device]) {
        int prev = 0;
        FVG_HIP(hipGetDevice(&prev));
        if (prev != c->device) FVG_HIP(hipSetDevice(c->device));
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (prev != c->device) (void)hipSetDevice(prev);
        if (e != hipSuccess || !p)
            return fail(FVG_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch[c->device] = static_cast<double*>(p);
    }
    *out = g_scratch[c->device];
    return FVG_OK;
}

// doubles rounded up to whole 256-byte lines
static int64_t lines(int64_t doubles) { return (doubles + 31) / 32 * 32; }

static int solve_noalloc(fvg_ctx* c, const fvg_csr* a, const fvg_prec* p, const fvg_vec* b, fvg_vec* x,
                         const fvg_solver_options* o, fvg_solver_info* info, double* history, int64_t cap,
                         bool bicg) {
    FVG_TRY(check_solve_args(c, a, p, b, x, o));
    if (o->precision != FVG_PRECISION_FP64)
        return fail(FVG_E_UNSUPPORTED, "the no-allocation solvers run FP64 iterations only");
    const int64_t n = a->nrows;
    const int nvec = bicg ? 8 : 4;
    if (o->max_iterations > (int64_t)(kScratchBytes / sizeof(double)))
        return fail(FVG_E_INVALID, "max_iterations exceeds the no-allocation scratch's history");
    // layout: partials (kPartialSlots slots), solver state, flags, history, work vectors
    const int64_t pstride = lines(spmv_blocks(n));
    const int64_t o_state = kPartialSlots * pstride;
    const int64_t o_flags = o_state + lines((int64_t)((sizeof(KState) + 7) / 8));
    const int64_t o_hist = o_flags + 32;
    const int64_t hist = o->max_iterations + 2;   // KRun's history slots
    const int64_t o_vec = o_hist + lines(hist);
    const int64_t vstride = lines(n);
    const int64_t total = o_vec + nvec * vstride;
    if (total > (int64_t)(kScratchBytes / sizeof(double)))
        return fail(FVG_E_INVALID, "the no-allocation " + std::string(bicg ? "BiCGStab" : "CG") + " on " +
                                       std::to_string(n) + " rows with " + std::to_string(o->max_iterations) +
                                       " iterations needs " + std::to_string(total * 8) +
                                       " bytes of scratch; it holds " + std::to_string(kScratchBytes));
    double* base;
    FVG_TRY(scratch_base(c, &base));
    KBuffers kb;
    for (int k = 0; k < nvec; ++k) kb.vec[k] = base + o_vec + k * vstride;
    kb.partials = base;
    kb.pstride = pstride;
    KRun K;
    FVG_TRY(K.init_at(c, 1, o, reinterpret_cast<KState*>(base + o_state), base + o_hist,
                      reinterpret_cast<int32_t*>(base + o_flags)));
    // kernel timing records HIP events: off for the call
    const bool timing = c->timing;
    c->timing = false;
    Layout L;
    const int s = bicg ? run_bicgstab_multi(c, a, p, b->d, x->d, L, K, &kb) : run_cg_multi(c, a, p, b->d, x->d, L, K, &kb);
    c->timing = timing;
    FVG_TRY(s);
    double* hs[1] = {history};
    return K.finish(info, hs, cap);
}

}  // namespace fvg

using namespace fvg;

extern "C" {

int fvg_cg_noalloc(fvg_ctx* c, const fvg_csr* a, const fvg_prec* p, const fvg_vec* b, fvg_vec* x,
                   const fvg_solver_options* o, fvg_solver_info* info, double* history, int64_t cap) {
    return solve_noalloc(c, a, p, b, x, o, info, history, cap, false);
}

int fvg_bicgstab_noalloc(fvg_ctx* c, const fvg_csr* a, const fvg_prec* p, const fvg_vec* b, fvg_vec* x,
                         const fvg_solver_options* o, fvg_solver_info* info, double* history, int64_t cap) {
    return solve_noalloc(c, a, p, b, x, o, info, history, cap, true);
}

int fvg_noalloc_scratch_bytes(uint64_t* bytes) {
    if (!bytes) return fail(FVG_E_INVALID, "null argument");
    *bytes = kScratchBytes;
    return FVG_OK;
}

}  // extern "C"
