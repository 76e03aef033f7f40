// krylov.hip -- conjugate_gradient and bi_conjugate_gradient_stab on the device,
// for one system or a batch of independent systems.
//
// The iteration is the reference's, step for step (conjugate_gradient.rs:12-112,
// bi_conjugate_gradient_stab.rs:13-179): same updates in the same operation order,
// same stopping tests on the recursive residual, same iteration counting and
// history.  All scalars live on the device: each reduction ends in a "scalar" kernel
// with one workgroup per system that sums that system's per-workgroup partials in a
// fixed order and runs the scalar part of the algorithm, so the host only enqueues
// work and polls one flag.  Vector updates are fused with the dot products after them.
//
// A batch is a block-diagonal system whose members start on 256-row boundaries, so
// every 256-row workgroup belongs to exactly one member and each member's partial
// sums (and so every result bit) are the same as when that member is solved alone.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "krylov.hpp"
#include "precond.hpp"
#include "reduce.hpp"
#include "spmv.hpp"

namespace fvg {

constexpr int kScalarBlock = 1024;   // threads per scalar workgroup

template <int STEP>
__global__ void __launch_bounds__(kScalarBlock) k_scalar(KState* __restrict__ states, const double* __restrict__ partials,
                                                         int64_t stride, int count, const int32_t* __restrict__ blk_ptr,
                                                         double* __restrict__ hist_all, int64_t hcap,
                                                         int32_t* __restrict__ active, int32_t* __restrict__ all_done) {
    __shared__ double lds[2 * kScalarBlock / 64];
    __shared__ int skip;
    KState* st = states + blockIdx.x;
    double* hist = hist_all + (int64_t)blockIdx.x * hcap;
    if (threadIdx.x == 0) {
        skip = 0;
        if (STEP == CG_ALPHA || STEP == BI_ALPHA) {
            if (st->halt) st->done = 1;
            skip = st->done;
        } else if (STEP == CG_INIT || STEP == BI_INIT) {
            skip = st->halt;   // zero right-hand side already answered
        } else if (STEP != ST_BNORM) {
            skip = st->done;
        }
    }
    __syncthreads();
    if (skip) return;
    const int b0 = blk_ptr ? blk_ptr[blockIdx.x] : 0;
    const int nb = blk_ptr ? blk_ptr[blockIdx.x + 1] - b0 : count;
    constexpr bool need0 = !(STEP == CG_INIT || STEP == BI_INIT);
    constexpr bool need1 = STEP == CG_INIT || STEP == BI_INIT || STEP == BI_OMEGA || STEP == BI_R;
    const double s0 = need0 ? sum_partials_dd<kScalarBlock>(partials + 2 * b0, nb, lds).hi : 0.0;
    const double s1 = need1 ? sum_partials_dd<kScalarBlock>(partials + 2 * stride + 2 * b0, nb, lds).hi : 0.0;
    if (threadIdx.x != 0) return;
    scalar_step(STEP, st, hist, s0, s1, active, all_done);
}

// ---- fused vector kernels (one element per thread, kBlock rows per workgroup) -----------
#define K_STATE const KState& S = states[blk_inst ? blk_inst[blockIdx.x] : 0]

// partial a.b (slot 0); skipped when `skip_done` and the system is done
__global__ void __launch_bounds__(kBlock) k_dot_g(int64_t n, const KState* __restrict__ states,
                                                  const int32_t* __restrict__ blk_inst, int skip_done,
                                                  const double* __restrict__ a, const double* __restrict__ b,
                                                  double* __restrict__ partials) {
    __shared__ double lds[2 * kBlock / 64];
    K_STATE;
    if (skip_done && S.done) return;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    double d = i < n ? a[i] * b[i] : 0.0;
    const DD dd = block_sum<kBlock>(DD{d, 0.0}, lds);
    if (threadIdx.x == 0) put_dd(partials, 0, 0, blockIdx.x, dd);
}

// x = 0 where the right-hand side was zero (conjugate_gradient.rs:31-36)
__global__ void k_zero_rhs_x(int64_t n, const KState* __restrict__ states, const int32_t* __restrict__ blk_inst,
                             double* __restrict__ x) {
    K_STATE;
    if (!S.zero_rhs) return;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) x[i] = 0.0;
}

// CG (conjugate_gradient.rs:77-80): x += p*alpha; r -= ap*alpha; partial r.r
__global__ void __launch_bounds__(kBlock) k_cg_xr(int64_t n, const KState* __restrict__ states,
                                                  const int32_t* __restrict__ blk_inst, const double* __restrict__ p,
                                                  const double* __restrict__ ap, double* __restrict__ x,
                                                  double* __restrict__ r, double* __restrict__ partials) {
    __shared__ double lds[2 * kBlock / 64];
    K_STATE;
    if (S.done) return;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const double alpha = S.alpha;
    double d = 0.0;
    if (i < n) {
        x[i] += p[i] * alpha;
        const double ri = r[i] - ap[i] * alpha;
        r[i] = ri;
        d = ri * ri;
    }
    const DD dd = block_sum<kBlock>(DD{d, 0.0}, lds);
    if (threadIdx.x == 0) put_dd(partials, 0, 0, blockIdx.x, dd);
}

// CG (conjugate_gradient.rs:98-100): p = z + p*beta
__global__ void __launch_bounds__(kBlock) k_cg_p(int64_t n, const KState* __restrict__ states,
                                                 const int32_t* __restrict__ blk_inst, const double* __restrict__ z,
                                                 double* __restrict__ p) {
    K_STATE;
    if (S.done) return;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) p[i] = z[i] + p[i] * S.beta;
}

// CG start: p = z (conjugate_gradient.rs:59), partial r.z
__global__ void __launch_bounds__(kBlock) k_cg_start(int64_t n, const KState* __restrict__ states,
                                                     const int32_t* __restrict__ blk_inst, const double* __restrict__ r,
                                                     const double* __restrict__ z, double* __restrict__ p,
                                                     double* __restrict__ partials) {
    __shared__ double lds[2 * kBlock / 64];
    K_STATE;
    if (S.halt) return;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    double d = 0.0;
    if (i < n) {
        p[i] = z[i];
        d = r[i] * z[i];
    }
    const DD dd = block_sum<kBlock>(DD{d, 0.0}, lds);
    if (threadIdx.x == 0) put_dd(partials, 0, 0, blockIdx.x, dd);
}

// BiCGStab (bi_conjugate_gradient_stab.rs:213-216): h = x + y*alpha (kept in x); s = r - v*alpha; partial s.s
__global__ void __launch_bounds__(kBlock) k_bi_hs(int64_t n, const KState* __restrict__ states,
                                                  const int32_t* __restrict__ blk_inst, const double* __restrict__ y,
                                                  const double* __restrict__ v, const double* __restrict__ r,
                                                  double* __restrict__ x, double* __restrict__ s,
                                                  double* __restrict__ partials) {
    __shared__ double lds[2 * kBlock / 64];
    K_STATE;
    if (S.done) return;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const double alpha = S.alpha;
    double d = 0.0;
    if (i < n) {
        x[i] = x[i] + y[i] * alpha;
        const double si = r[i] - v[i] * alpha;
        s[i] = si;
        d = si * si;
    }
    const DD dd = block_sum<kBlock>(DD{d, 0.0}, lds);
    if (threadIdx.x == 0) put_dd(partials, 0, 0, blockIdx.x, dd);
}

// BiCGStab (:246-249): x = h + z*omega; r = s - t*omega; partials rh.r (slot 0) and r.r (slot 1)
__global__ void __launch_bounds__(kBlock) k_bi_xr(int64_t n, const KState* __restrict__ states,
                                                  const int32_t* __restrict__ blk_inst, const double* __restrict__ z,
                                                  const double* __restrict__ t, const double* __restrict__ s,
                                                  const double* __restrict__ rh, double* __restrict__ x,
                                                  double* __restrict__ r, double* __restrict__ partials,
                                                  int64_t stride) {
    __shared__ double lds[2 * kBlock / 64];
    K_STATE;
    if (S.done) return;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const double omega = S.omega;
    double d0 = 0.0, d1 = 0.0;
    if (i < n) {
        x[i] = x[i] + z[i] * omega;
        const double ri = s[i] - t[i] * omega;
        r[i] = ri;
        d0 = rh[i] * ri;
        d1 = ri * ri;
    }
    const DD e0 = block_sum<kBlock>(DD{d0, 0.0}, lds);
    if (threadIdx.x == 0) put_dd(partials, stride, 0, blockIdx.x, e0);
    const DD e1 = block_sum<kBlock>(DD{d1, 0.0}, lds);
    if (threadIdx.x == 0) put_dd(partials, stride, 1, blockIdx.x, e1);
}

// BiCGStab (:268-282): p = r + (p - v*omega)*beta, or restart rh = r, p = r
__global__ void __launch_bounds__(kBlock) k_bi_p(int64_t n, const KState* __restrict__ states,
                                                 const int32_t* __restrict__ blk_inst, const double* __restrict__ r,
                                                 const double* __restrict__ v, double* __restrict__ p,
                                                 double* __restrict__ rh) {
    K_STATE;
    if (S.done) return;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    if (S.mode == 0) {
        p[i] = r[i] + (p[i] - v[i] * S.omega) * S.beta;
    } else {
        rh[i] = r[i];
        p[i] = r[i];
    }
}

// ---- kernels of the position-order solve (one system; see PosSolve) ------------------------
// Every vector is indexed by the lower sweep's position; z (and BiCGStab's y) is read
// reversed from the upper sweep's output.  The vector kernels that precede an apply also
// write the pending pattern at the positions the two sweeps poll; where such a kernel reads
// the upper output, each thread reads its element before overwriting it, so no element is
// overwritten while another thread may still read it.
#define K_ONE const KState& S = *states

__device__ __forceinline__ void pending_at(const PosSolve& P, int64_t q, int64_t qu) {
    if (P.lo_poll[q]) reinterpret_cast<uint64_t*>(P.lo_x)[q] = kPendingBits;
    if (P.up_poll[qu]) reinterpret_cast<uint64_t*>(P.up_x)[qu] = kPendingBits;
}

// b and x into position order
__global__ void k_pos_in(int64_t n, const int32_t* __restrict__ row, const double* __restrict__ b,
                         const double* __restrict__ x, double* __restrict__ bp, double* __restrict__ xp) {
    const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (q < n) {
        const int32_t i = row[q];
        bp[q] = b[i];
        xp[q] = x[i];
    }
}

// the solution back to its rows
__global__ void k_pos_out(int64_t n, const int32_t* __restrict__ row, const double* __restrict__ xp,
                          double* __restrict__ x) {
    const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (q < n) x[row[q]] = xp[q];
}

// the pending pattern for the first apply
__global__ void k_pos_pending(int64_t n, const KState* __restrict__ states, PosSolve P) {
    K_ONE;
    if (S.halt) return;
    const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (q < n) pending_at(P, q, q);
}

// CG start (conjugate_gradient.rs:59): p = z, partial r.z; the next apply's pending pattern
__global__ void __launch_bounds__(kBlock) k_cg_start_pos(int64_t n, const KState* __restrict__ states,
                                                         const double* __restrict__ r, double* __restrict__ p,
                                                         double* __restrict__ partials, PosSolve P) {
    __shared__ double lds[2 * kBlock / 64];
    K_ONE;
    if (S.halt) return;
    const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    double d = 0.0;
    if (q < n) {
        const int64_t qu = n - 1 - q;
        const double zq = P.up_x[qu];
        p[q] = zq;
        d = r[q] * zq;
        pending_at(P, q, qu);
    }
    const DD dd = block_sum<kBlock>(DD{d, 0.0}, lds);
    if (threadIdx.x == 0) put_dd(partials, 0, 0, blockIdx.x, dd);
}

// CG (conjugate_gradient.rs:94): partial r.z_new
__global__ void __launch_bounds__(kBlock) k_cg_rz_pos(int64_t n, const KState* __restrict__ states,
                                                      const double* __restrict__ r, const double* __restrict__ up_x,
                                                      double* __restrict__ partials) {
    __shared__ double lds[2 * kBlock / 64];
    K_ONE;
    if (S.done) return;
    const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const double d = q < n ? r[q] * up_x[n - 1 - q] : 0.0;
    const DD dd = block_sum<kBlock>(DD{d, 0.0}, lds);
    if (threadIdx.x == 0) put_dd(partials, 0, 0, blockIdx.x, dd);
}

// CG (conjugate_gradient.rs:98-100): p = z + p*beta; the next apply's pending pattern
__global__ void __launch_bounds__(kBlock) k_cg_p_pos(int64_t n, const KState* __restrict__ states,
                                                     double* __restrict__ p, PosSolve P) {
    K_ONE;
    if (S.done) return;
    const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (q < n) {
        const int64_t qu = n - 1 - q;
        p[q] = P.up_x[qu] + p[q] * S.beta;
        pending_at(P, q, qu);
    }
}

// BiCGStab (bi_conjugate_gradient_stab.rs:213-216): h = x + y*alpha (kept in x); s = r - v*alpha;
// partial s.s; the second apply's pending pattern (y is read from the upper output first)
__global__ void __launch_bounds__(kBlock) k_bi_hs_pos(int64_t n, const KState* __restrict__ states,
                                                      const double* __restrict__ v, const double* __restrict__ r,
                                                      double* __restrict__ x, double* __restrict__ s,
                                                      double* __restrict__ partials, PosSolve P) {
    __shared__ double lds[2 * kBlock / 64];
    K_ONE;
    if (S.done) return;
    const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const double alpha = S.alpha;
    double d = 0.0;
    if (q < n) {
        const int64_t qu = n - 1 - q;
        x[q] = x[q] + P.up_x[qu] * alpha;
        const double sq = r[q] - v[q] * alpha;
        s[q] = sq;
        d = sq * sq;
        pending_at(P, q, qu);
    }
    const DD dd = block_sum<kBlock>(DD{d, 0.0}, lds);
    if (threadIdx.x == 0) put_dd(partials, 0, 0, blockIdx.x, dd);
}

// BiCGStab (:246-249): x = h + z*omega; r = s - t*omega; partials rh.r (slot 0) and r.r (slot 1)
__global__ void __launch_bounds__(kBlock) k_bi_xr_pos(int64_t n, const KState* __restrict__ states,
                                                      const double* __restrict__ up_x, const double* __restrict__ t,
                                                      const double* __restrict__ s, const double* __restrict__ rh,
                                                      double* __restrict__ x, double* __restrict__ r,
                                                      double* __restrict__ partials, int64_t stride) {
    __shared__ double lds[2 * kBlock / 64];
    K_ONE;
    if (S.done) return;
    const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const double omega = S.omega;
    double d0 = 0.0, d1 = 0.0;
    if (q < n) {
        x[q] = x[q] + up_x[n - 1 - q] * omega;
        const double rq = s[q] - t[q] * omega;
        r[q] = rq;
        d0 = rh[q] * rq;
        d1 = rq * rq;
    }
    const DD e0 = block_sum<kBlock>(DD{d0, 0.0}, lds);
    if (threadIdx.x == 0) put_dd(partials, stride, 0, blockIdx.x, e0);
    const DD e1 = block_sum<kBlock>(DD{d1, 0.0}, lds);
    if (threadIdx.x == 0) put_dd(partials, stride, 1, blockIdx.x, e1);
}

// BiCGStab (:268-282): p = r + (p - v*omega)*beta, or the restart rh = r, p = r; the next
// iteration's first apply's pending pattern
__global__ void __launch_bounds__(kBlock) k_bi_p_pos(int64_t n, const KState* __restrict__ states,
                                                     const double* __restrict__ r, const double* __restrict__ v,
                                                     double* __restrict__ p, double* __restrict__ rh, PosSolve P) {
    K_ONE;
    if (S.done) return;
    const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (q >= n) return;
    if (S.mode == 0) {
        p[q] = r[q] + (p[q] - v[q] * S.omega) * S.beta;
    } else {
        rh[q] = r[q];
        p[q] = r[q];
    }
    pending_at(P, q, q);
}

// ---- host drivers ----------------------------------------------------------------------

// Work vectors of one solve.  They come from the context's cache of earlier solves' work
// vectors when one is large enough (the smallest such), and go back to it afterwards: the
// next solve on the same stream reuses them in stream order, with no allocation or free
// (a free waits for the device).  The cache keeps at most kWorkCache vectors.
constexpr size_t kWorkCache = 16;
struct Workspace {
    fvg_ctx* c;
    std::vector<std::pair<double*, int64_t>> bufs;
    ~Workspace() {
        for (auto& b : bufs) c->work_cache.push_back(b);
        if (c->work_cache.size() > kWorkCache) {   // keep the largest
            (void)hipStreamSynchronize(c->stream);
            std::sort(c->work_cache.begin(), c->work_cache.end(),
                      [](const std::pair<double*, int64_t>& x, const std::pair<double*, int64_t>& y) {
                          return x.second > y.second;
                      });
            for (size_t k = kWorkCache; k < c->work_cache.size(); ++k)
                dev_free_n(c, c->work_cache[k].first, c->work_cache[k].second);
            c->work_cache.resize(kWorkCache);
        }
    }
    int get(int64_t n, double** out) {
        size_t best = c->work_cache.size();
        for (size_t k = 0; k < c->work_cache.size(); ++k)
            if (c->work_cache[k].second >= n &&
                (best == c->work_cache.size() || c->work_cache[k].second < c->work_cache[best].second))
                best = k;
        if (best < c->work_cache.size()) {
            bufs.push_back(c->work_cache[best]);
            c->work_cache.erase(c->work_cache.begin() + (ptrdiff_t)best);
        } else {
            double* p = nullptr;
            FVG_TRY(dev_alloc_n(c, n, &p));
            bufs.emplace_back(p, n);
        }
        *out = bufs.back().first;
        return FVG_OK;
    }
};

// zeroed states carrying the options, and the flags {members iterating, all done}
__global__ void k_kstate_init(int n_inst, KState* __restrict__ st, double atol, double rtol, int64_t max_iter,
                              int64_t hist_cap, int32_t* __restrict__ flags) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_inst) {
        KState s{};
        s.atol = atol;
        s.rtol = rtol;
        s.max_iter = max_iter;
        s.hist_cap = hist_cap;
        st[i] = s;
    }
    if (i == 0) {
        flags[0] = n_inst;
        flags[1] = 0;
        flags[2] = 0;   // unused
        flags[3] = 0;
    }
}

static int kstate_init(KRun& K, const fvg_solver_options* o) {
    hipLaunchKernelGGL(k_kstate_init, dim3(grid_for(K.n_inst)), dim3(kBlock), 0, K.c->stream, K.n_inst, K.st,
                       o->absolute_tolerance, o->relative_tolerance, (int64_t)o->max_iterations, K.hist_cap, K.flags);
    FVG_HIP(hipGetLastError());
    return FVG_OK;
}

int KRun::init_at(fvg_ctx* ctx, int ninst, const fvg_solver_options* o, KState* st_, double* hist_,
                  int32_t* flags_) {
    c = ctx;
    n_inst = ninst;
    hist_cap = o->max_iterations + 2;
    st = st_;
    hist = hist_;
    flags = flags_;
    owned = false;
    return kstate_init(*this, o);
}

int KRun::init(fvg_ctx* ctx, int ninst, const fvg_solver_options* o) {
    c = ctx;
    n_inst = ninst;
    hist_cap = o->max_iterations + 2;
    owned = true;
    FVG_TRY(dev_alloc_n(c, n_inst, &st));
    FVG_TRY(dev_alloc_n(c, hist_cap * n_inst, &hist));
    FVG_TRY(dev_alloc_n(c, 4, &flags));
    return kstate_init(*this, o);
}

KRun::~KRun() {
    if (c && owned) {
        (void)hipStreamSynchronize(c->stream);
        dev_free_n(c, st, n_inst);
        dev_free_n(c, hist, hist_cap * n_inst);
        dev_free_n(c, flags, 4);
    }
}

int KRun::all_done(bool* out) {
    FVG_HIP(hipMemcpyAsync(c->pinned, flags + 1, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    FVG_HIP(hipStreamSynchronize(c->stream));
    int32_t v;
    std::memcpy(&v, c->pinned, sizeof(v));
    *out = v != 0;
    return FVG_OK;
}

int KRun::finish(fvg_solver_info* infos, double* const* histories, int64_t cap) {
    // one state through the context's pinned buffer (nothing allocated); several through a host array
    std::vector<KState> h;
    const KState* hs;
    if (sizeof(KState) * (size_t)n_inst <= 4096) {
        FVG_HIP(hipMemcpyAsync(c->pinned, st, sizeof(KState) * (size_t)n_inst, hipMemcpyDeviceToHost, c->stream));
        FVG_HIP(hipStreamSynchronize(c->stream));
        hs = static_cast<const KState*>(c->pinned);
    } else {
        h.resize((size_t)n_inst);
        FVG_TRY(download(c, h.data(), st, n_inst));
        hs = h.data();
    }
    for (int i = 0; i < n_inst; ++i) {
        const KState& s = hs[i];
        if (infos) {
            fvg_solver_info& info = infos[i];
            std::memset(&info, 0, sizeof(info));
            info.iterations = s.iter;
            info.initial_residual = s.init_res;
            info.final_residual = s.error;
            info.history_len = s.hist_len;
            info.restarts = s.restarts;
            info.flags = path;
        }
        if (histories && histories[i] && cap > 0) {
            int64_t m = s.hist_len < cap ? s.hist_len : cap;
            m = m < hist_cap ? m : hist_cap;
            FVG_TRY(download(c, histories[i], hist + (int64_t)i * hist_cap, m));
        }
    }
    return FVG_OK;
}

#define SCALAR(step, part0)                                                                                \
    hipLaunchKernelGGL(k_scalar<step>, dim3(L.n_inst), dim3(kScalarBlock), 0, c->stream, K.st, part0, stride, \
                       nb, L.blk_ptr, K.hist, K.hist_cap, K.flags, K.flags + 1)

// iterations enqueued between two polls of the "all systems done" flag
constexpr int kPollBatch = 4;

// Poll number `batch` (1, 2, ...) of a solve: enqueues the copy of the done flag into the
// context's pinned slot batch & 1 and records the slot's event, then reads the flag of
// the poll before (the host waits for that one only).  So one batch of iterations is
// always queued behind the one the host waits for, and the device never idles on the
// host; a solve that converges runs at most one batch of iterations that skip at once.
int KRun::poll(int64_t batch, bool* done) {
    const int slot = (int)(batch & 1);
    FVG_HIP(hipMemcpyAsync(c->poll_host + slot, flags + 1, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    FVG_HIP(hipEventRecord(c->poll_evt[slot], c->stream));
    *done = false;
    if (batch < 2) return FVG_OK;
    FVG_HIP(hipEventSynchronize(c->poll_evt[slot ^ 1]));
    *done = __atomic_load_n(c->poll_host + (slot ^ 1), __ATOMIC_ACQUIRE) != 0;
    return FVG_OK;
}

// work vectors and partials: the caller's (kb), or allocated here and the context's
static int solve_buffers(fvg_ctx* c, int64_t n, int count, const KBuffers* kb, Workspace& ws, double** v,
                         double** part, int64_t* stride) {
    if (kb) {
        for (int k = 0; k < count; ++k) v[k] = kb->vec[k];
        *part = kb->partials;
        *stride = kb->pstride;
        return FVG_OK;
    }
    for (int k = 0; k < count; ++k) FVG_TRY(ws.get(n, &v[k]));
    FVG_TRY(ensure_partials(c, spmv_blocks(n)));
    *part = c->partials;
    *stride = c->partials_cap;
    return FVG_OK;
}

// CG in the lower sweep's position order (PosSolve): none of its kernels reorders a
// vector; with the sweeps fused, an iteration is five kernels (SpMV, alpha, the two
// sweeps, the direction update).
static int run_cg_pos(fvg_ctx* c, const fvg_prec* p, const double* b, double* x, const Layout& L, KRun& K,
                      const KBuffers* kb, const PosSolve& P) {
    const int64_t n = P.n;
    Workspace ws{c};
    double* v[4];
    double* part;
    int64_t stride;
    FVG_TRY(solve_buffers(c, n, 4, kb, ws, v, &part, &stride));
    double *xp = v[0], *r = v[1], *pp = v[2], *ap = v[3];
    double* bp = v[3];   // b in position order, dead once r is formed (then ap)
    const int nb = spmv_blocks(n);
    double* const epart = kb ? part : nullptr;
    const int32_t* all_done = K.flags + 1;
    hipLaunchKernelGGL(k_pos_in, dim3(nb), dim3(kBlock), 0, c->stream, n, P.row, b, (const double*)x, bp, xp);
    hipLaunchKernelGGL(k_dot_g, dim3(nb), dim3(kBlock), 0, c->stream, n, K.st, nullptr, 0, bp, bp, part);
    SCALAR(ST_BNORM, part);
    hipLaunchKernelGGL(k_zero_rhs_x, dim3(nb), dim3(kBlock), 0, c->stream, n, K.st, nullptr, xp);
    SpmvEpilogue e0;   // r = (A x - b) * -1 and r.r
    e0.b = bp;
    e0.dot_self = true;
    e0.skip = &K.st[0].halt;
    e0.partials = epart;
    e0.pstride = stride;
    e0.dd = true;
    FVG_TRY(launch_spmv(c, P.pa, xp, r, e0));
    SCALAR(CG_INIT, part);
    hipLaunchKernelGGL(k_pos_pending, dim3(nb), dim3(kBlock), 0, c->stream, n, K.st, P);
    FVG_TRY(apply_pos(c, p, r, all_done));
    hipLaunchKernelGGL(k_cg_start_pos, dim3(nb), dim3(kBlock), 0, c->stream, n, K.st, r, pp, part, P);
    SCALAR(CG_RZ0, part);
    FVG_HIP(hipGetLastError());
    // the residual update with r.r and the r.z product run inside the two sweeps
    const int64_t lo_nb = p->lower.nblk, up_nb = p->upper.nblk;
    const bool fuse = c->tune_cg_fusion == 0 && lo_nb <= stride && up_nb <= stride && pos_fusable(c, p);
    SweepFuse lf, uf;
    lf.alpha = &K.st[0].alpha;
    lf.ap = ap;
    lf.x = xp;
    lf.pp = pp;
    lf.r = r;
    uf.r = r;
    for (SweepFuse* f : {&lf, &uf}) {   // the sweeps' last blocks run the scalar steps
        f->partials = part;
        f->counter = K.flags + 2;
        f->st = K.st;
        f->hist = K.hist;
        f->active = K.flags;
        f->all_done = K.flags + 1;
    }
    lf.step = CG_ERROR;
    uf.step = CG_BETA;
    K.path = FVG_INFO_POSITION_ORDER | (fuse ? FVG_INFO_FUSED_SWEEPS : 0);
    for (int64_t it = 0;; ++it) {
        SpmvEpilogue e1;   // ap = A p with p.ap
        e1.dotw = pp;
        e1.skip = &K.st[0].halt;
        e1.partials = epart;
        e1.pstride = stride;
        e1.dd = true;
        FVG_TRY(launch_spmv(c, P.pa, pp, ap, e1));
        SCALAR(CG_ALPHA, part);
        if (fuse) {
            FVG_TRY(apply_pos_fused(c, p, r, all_done, lf, uf));
        } else {
            hipLaunchKernelGGL(k_cg_xr, dim3(nb), dim3(kBlock), 0, c->stream, n, K.st, nullptr, pp, ap, xp, r, part);
            SCALAR(CG_ERROR, part);
            FVG_TRY(apply_pos(c, p, r, all_done));
            hipLaunchKernelGGL(k_cg_rz_pos, dim3(nb), dim3(kBlock), 0, c->stream, n, K.st, r, (const double*)P.up_x,
                               part);
            SCALAR(CG_BETA, part);
        }
        hipLaunchKernelGGL(k_cg_p_pos, dim3(nb), dim3(kBlock), 0, c->stream, n, K.st, pp, P);
        FVG_HIP(hipGetLastError());
        if ((it + 1) % kPollBatch == 0) {
            bool d;
            FVG_TRY(K.poll((it + 1) / kPollBatch, &d));
            if (d) break;
        }
    }
    hipLaunchKernelGGL(k_pos_out, dim3(nb), dim3(kBlock), 0, c->stream, n, P.row, (const double*)xp, x);
    FVG_HIP(hipGetLastError());
    return check_sweep_status(c, p);
}

// BiCGStab in the lower sweep's position order (PosSolve); y and z are read reversed from
// the upper sweep's output.
static int run_bicgstab_pos(fvg_ctx* c, const fvg_prec* p, const double* b, double* x, const Layout& L, KRun& K,
                            const KBuffers* kb, const PosSolve& P) {
    const int64_t n = P.n;
    Workspace ws{c};
    double* w[7];
    double* part;
    int64_t stride;
    FVG_TRY(solve_buffers(c, n, 7, kb, ws, w, &part, &stride));
    double *xp = w[0], *r = w[1], *rh = w[2], *pp = w[3], *v = w[4], *s = w[5], *t = w[6];
    double* bp = w[6];   // b in position order, dead once r is formed (then t)
    const int nb = spmv_blocks(n);
    double* const epart = kb ? part : nullptr;
    const int32_t* all_done = K.flags + 1;
    hipLaunchKernelGGL(k_pos_in, dim3(nb), dim3(kBlock), 0, c->stream, n, P.row, b, (const double*)x, bp, xp);
    hipLaunchKernelGGL(k_dot_g, dim3(nb), dim3(kBlock), 0, c->stream, n, K.st, nullptr, 0, bp, bp, part);
    SCALAR(ST_BNORM, part);
    hipLaunchKernelGGL(k_zero_rhs_x, dim3(nb), dim3(kBlock), 0, c->stream, n, K.st, nullptr, xp);
    SpmvEpilogue e0;   // r = b - A x, rh = r, p = r, r.r
    e0.b = bp;
    e0.copy1 = rh;
    e0.copy2 = pp;
    e0.dot_self = true;
    e0.skip = &K.st[0].halt;
    e0.partials = epart;
    e0.pstride = stride;
    e0.dd = true;
    FVG_TRY(launch_spmv(c, P.pa, xp, r, e0));
    SCALAR(BI_INIT, part);
    hipLaunchKernelGGL(k_pos_pending, dim3(nb), dim3(kBlock), 0, c->stream, n, K.st, P);
    FVG_HIP(hipGetLastError());
    K.path = FVG_INFO_POSITION_ORDER;
    for (int64_t it = 0;; ++it) {
        FVG_TRY(apply_pos(c, p, pp, all_done));   // y = M^-1 p (reversed in up_x)
        SpmvEpilogue e1;                            // v = A y, rh.v
        e1.dotw = rh;
        e1.skip = &K.st[0].halt;
        e1.partials = epart;
        e1.pstride = stride;
        e1.dd = true;
        e1.xrev = n;
        FVG_TRY(launch_spmv(c, P.pa, P.up_x, v, e1));
        SCALAR(BI_ALPHA, part);
        hipLaunchKernelGGL(k_bi_hs_pos, dim3(nb), dim3(kBlock), 0, c->stream, n, K.st, v, r, xp, s, part, P);
        SCALAR(BI_S, part);
        FVG_TRY(apply_pos(c, p, s, all_done));    // z = M^-1 s (reversed in up_x)
        SpmvEpilogue e2;                            // t = A z, t.s, t.t
        e2.dotw = s;
        e2.dot_self = true;
        e2.skip = &K.st[0].done;
        e2.partials = epart;
        e2.pstride = stride;
        e2.dd = true;
        e2.xrev = n;
        FVG_TRY(launch_spmv(c, P.pa, P.up_x, t, e2));
        SCALAR(BI_OMEGA, part);
        hipLaunchKernelGGL(k_bi_xr_pos, dim3(nb), dim3(kBlock), 0, c->stream, n, K.st, (const double*)P.up_x, t, s, rh,
                           xp, r, part, stride);
        SCALAR(BI_R, part);
        hipLaunchKernelGGL(k_bi_p_pos, dim3(nb), dim3(kBlock), 0, c->stream, n, K.st, r, v, pp, rh, P);
        FVG_HIP(hipGetLastError());
        if ((it + 1) % kPollBatch == 0) {
            bool d;
            FVG_TRY(K.poll((it + 1) / kPollBatch, &d));
            if (d) break;
        }
    }
    hipLaunchKernelGGL(k_pos_out, dim3(nb), dim3(kBlock), 0, c->stream, n, P.row, (const double*)xp, x);
    FVG_HIP(hipGetLastError());
    return check_sweep_status(c, p);
}

int run_cg_multi(fvg_ctx* c, const fvg_csr* a, const fvg_prec* p, const double* b, double* x, const Layout& L,
                 KRun& K, const KBuffers* kb) {
    if (L.n_inst == 1) {
        PosSolve P;
        bool ok;
        FVG_TRY(pos_solve(c, p, a, kb == nullptr, &P, &ok));
        if (ok) return run_cg_pos(c, p, b, x, L, K, kb, P);
    }
    const int64_t n = a->nrows;
    Workspace ws{c};
    double* v[4];
    double* part;
    int64_t stride;
    FVG_TRY(solve_buffers(c, n, 4, kb, ws, v, &part, &stride));
    double *r = v[0], *z = v[1], *pp = v[2], *ap = v[3];
    const int nb = spmv_blocks(n);
    double* const epart = kb ? part : nullptr;   // the SpMV epilogues' partials (null: the context's)
    const int32_t* all_done = K.flags + 1;
    const int32_t sstride = (int32_t)(sizeof(KState) / sizeof(int32_t));

    // |b| and the zero right-hand side exit
    hipLaunchKernelGGL(k_dot_g, dim3(nb), dim3(kBlock), 0, c->stream, n, K.st, L.blk_inst, 0, b, b, part);
    SCALAR(ST_BNORM, part);
    hipLaunchKernelGGL(k_zero_rhs_x, dim3(nb), dim3(kBlock), 0, c->stream, n, K.st, L.blk_inst, x);
    // r = (A x - b) * -1 and r.r (conjugate_gradient.rs:47-52)
    SpmvEpilogue e0;
    e0.b = b;
    e0.dot_self = true;
    e0.skip = &K.st[0].halt;
    e0.skip_stride = sstride;
    e0.blk_inst = L.blk_inst;
    e0.partials = epart;
    e0.dd = true;
    e0.pstride = stride;
    FVG_TRY(launch_spmv(c, a, x, r, e0));
    SCALAR(CG_INIT, part);
    // z = M^-1 r; p = z; r.z
    FVG_TRY(apply_prec(c, p, r, z, all_done));
    hipLaunchKernelGGL(k_cg_start, dim3(nb), dim3(kBlock), 0, c->stream, n, K.st, L.blk_inst, r, z, pp, part);
    SCALAR(CG_RZ0, part);
    FVG_HIP(hipGetLastError());

    for (int64_t it = 0;; ++it) {
        SpmvEpilogue e1;   // ap = A p with p.ap
        e1.dotw = pp;
        e1.skip = &K.st[0].halt;
        e1.skip_stride = sstride;
        e1.blk_inst = L.blk_inst;
        e1.partials = epart;
        e1.dd = true;
        e1.pstride = stride;
        FVG_TRY(launch_spmv(c, a, pp, ap, e1));
        SCALAR(CG_ALPHA, part);
        hipLaunchKernelGGL(k_cg_xr, dim3(nb), dim3(kBlock), 0, c->stream, n, K.st, L.blk_inst, pp, ap, x, r, part);
        SCALAR(CG_ERROR, part);
        FVG_TRY(apply_prec(c, p, r, z, all_done));
        hipLaunchKernelGGL(k_dot_g, dim3(nb), dim3(kBlock), 0, c->stream, n, K.st, L.blk_inst, 1, r, z, part);
        SCALAR(CG_BETA, part);
        hipLaunchKernelGGL(k_cg_p, dim3(nb), dim3(kBlock), 0, c->stream, n, K.st, L.blk_inst, z, pp);
        FVG_HIP(hipGetLastError());
        if ((it + 1) % kPollBatch == 0) {
            bool d;
            FVG_TRY(K.poll((it + 1) / kPollBatch, &d));
            if (d) break;
        }
    }
    return check_sweep_status(c, p);
}

int run_bicgstab_multi(fvg_ctx* c, const fvg_csr* a, const fvg_prec* p, const double* b, double* x, const Layout& L,
                       KRun& K, const KBuffers* kb) {
    if (L.n_inst == 1) {
        PosSolve P;
        bool ok;
        FVG_TRY(pos_solve(c, p, a, kb == nullptr, &P, &ok));
        if (ok) return run_bicgstab_pos(c, p, b, x, L, K, kb, P);
    }
    const int64_t n = a->nrows;
    Workspace ws{c};
    double* w[8];
    double* part;
    int64_t stride;
    FVG_TRY(solve_buffers(c, n, 8, kb, ws, w, &part, &stride));
    double *r = w[0], *rh = w[1], *pp = w[2], *y = w[3], *v = w[4], *s = w[5], *z = w[6], *t = w[7];
    const int nb = spmv_blocks(n);
    double* const epart = kb ? part : nullptr;   // the SpMV epilogues' partials (null: the context's)
    const int32_t* all_done = K.flags + 1;
    const int32_t sstride = (int32_t)(sizeof(KState) / sizeof(int32_t));

    hipLaunchKernelGGL(k_dot_g, dim3(nb), dim3(kBlock), 0, c->stream, n, K.st, L.blk_inst, 0, b, b, part);
    SCALAR(ST_BNORM, part);
    hipLaunchKernelGGL(k_zero_rhs_x, dim3(nb), dim3(kBlock), 0, c->stream, n, K.st, L.blk_inst, x);
    // r = b - A x, rh = r, p = r, r.r (bi_conjugate_gradient_stab.rs:165-190)
    SpmvEpilogue e0;
    e0.b = b;
    e0.copy1 = rh;
    e0.copy2 = pp;
    e0.dot_self = true;
    e0.skip = &K.st[0].halt;
    e0.skip_stride = sstride;
    e0.blk_inst = L.blk_inst;
    e0.partials = epart;
    e0.dd = true;
    e0.pstride = stride;
    FVG_TRY(launch_spmv(c, a, x, r, e0));
    SCALAR(BI_INIT, part);
    FVG_HIP(hipGetLastError());

    for (int64_t it = 0;; ++it) {
        FVG_TRY(apply_prec(c, p, pp, y, all_done));   // y = M^-1 p
        SpmvEpilogue e1;                                // v = A y, rh.v
        e1.dotw = rh;
        e1.skip = &K.st[0].halt;
        e1.skip_stride = sstride;
        e1.blk_inst = L.blk_inst;
        e1.partials = epart;
        e1.dd = true;
        e1.pstride = stride;
        FVG_TRY(launch_spmv(c, a, y, v, e1));
        SCALAR(BI_ALPHA, part);
        hipLaunchKernelGGL(k_bi_hs, dim3(nb), dim3(kBlock), 0, c->stream, n, K.st, L.blk_inst, y, v, r, x, s, part);
        SCALAR(BI_S, part);
        FVG_TRY(apply_prec(c, p, s, z, all_done));    // z = M^-1 s
        SpmvEpilogue e2;                                // t = A z, t.s, t.t
        e2.dotw = s;
        e2.dot_self = true;
        e2.skip = &K.st[0].done;
        e2.skip_stride = sstride;
        e2.blk_inst = L.blk_inst;
        e2.partials = epart;
        e2.dd = true;
        e2.pstride = stride;
        FVG_TRY(launch_spmv(c, a, z, t, e2));
        SCALAR(BI_OMEGA, part);
        hipLaunchKernelGGL(k_bi_xr, dim3(nb), dim3(kBlock), 0, c->stream, n, K.st, L.blk_inst, z, t, s, rh, x, r, part,
                           stride);
        SCALAR(BI_R, part);
        hipLaunchKernelGGL(k_bi_p, dim3(nb), dim3(kBlock), 0, c->stream, n, K.st, L.blk_inst, r, v, pp, rh);
        FVG_HIP(hipGetLastError());
        if ((it + 1) % kPollBatch == 0) {
            bool d;
            FVG_TRY(K.poll((it + 1) / kPollBatch, &d));
            if (d) break;
        }
    }
    return check_sweep_status(c, p);
}

int check_solve_args(fvg_ctx* c, const fvg_csr* a, const fvg_prec* p, const fvg_vec* b, const fvg_vec* x,
                            const fvg_solver_options* o) {
    if (!c || !a || !p || !b || !x || !o) return fail(FVG_E_INVALID, "null argument");
    if (a->nrows != a->ncols) return fail(FVG_E_INVALID, "solver needs a square matrix");
    if (b->n != a->nrows || x->n != a->nrows || p->n != a->nrows)
        return fail(FVG_E_INVALID, "vector / preconditioner size does not match the matrix");
    if (o->max_iterations < 0) return fail(FVG_E_INVALID, "max_iterations must be non-negative");
    if (b->d == x->d) return fail(FVG_E_INVALID, "b and x must not alias");
    return FVG_OK;
}

// mixed.hip
int run_mixed(fvg_ctx* c, fvg_csr* a, fvg_prec* p, const fvg_vec* b, fvg_vec* x, const fvg_solver_options* o,
              fvg_solver_info* info, double* history, int64_t cap, bool bicgstab);

static int run_single(fvg_ctx* c, const fvg_csr* a, const fvg_prec* p, const fvg_vec* b, fvg_vec* x,
                      const fvg_solver_options* o, fvg_solver_info* info, double* history, int64_t cap, bool bicg) {
    KRun K;
    FVG_TRY(K.init(c, 1, o));
    Layout L;
    FVG_TRY(bicg ? run_bicgstab_multi(c, a, p, b->d, x->d, L, K) : run_cg_multi(c, a, p, b->d, x->d, L, K));
    double* hs[1] = {history};
    return K.finish(info, hs, cap);
}

}  // namespace fvg

using namespace fvg;

extern "C" {

// Mixed precision, and when the FP32 iterations cannot reach the tolerance on a system
// (refinement stalled), the reference FP64 iterations from the refined solution.
static int solve_mixed(fvg_ctx* c, const fvg_csr* a, const fvg_prec* p, const fvg_vec* b, fvg_vec* x,
                       const fvg_solver_options* o, fvg_solver_info* info, double* history, int64_t cap, bool bicg) {
    fvg_solver_info mi{};
    // the FP32 copies of the matrix values and factors are caches owned by a and p
    FVG_TRY(run_mixed(c, const_cast<fvg_csr*>(a), const_cast<fvg_prec*>(p), b, x, o, &mi, history, cap, bicg));
    const bool met = (mi.true_residual < o->absolute_tolerance) ||
                     (mi.initial_residual > 0.0 && mi.true_residual / mi.initial_residual < o->relative_tolerance) ||
                     mi.initial_residual == 0.0;
    if (!met && mi.iterations < o->max_iterations) {
        fvg_solver_info fi{};
        FVG_TRY(run_single(c, a, p, b, x, o, &fi, nullptr, 0, bicg));
        mi.iterations += fi.iterations;
        mi.final_residual = fi.final_residual;
        mi.true_residual = fi.final_residual;
        mi.restarts += fi.restarts;
        mi.flags |= FVG_INFO_FP64_FALLBACK;
    }
    if (info) *info = mi;
    return FVG_OK;
}

int fvg_cg(fvg_ctx* c, const fvg_csr* a, const fvg_prec* p, const fvg_vec* b, fvg_vec* x,
           const fvg_solver_options* o, fvg_solver_info* info, double* history, int64_t cap) {
    FVG_TRY(check_solve_args(c, a, p, b, x, o));
    if (o->precision == FVG_PRECISION_MIXED) return solve_mixed(c, a, p, b, x, o, info, history, cap, false);
    return run_single(c, a, p, b, x, o, info, history, cap, false);
}

int fvg_bicgstab(fvg_ctx* c, const fvg_csr* a, const fvg_prec* p, const fvg_vec* b, fvg_vec* x,
                 const fvg_solver_options* o, fvg_solver_info* info, double* history, int64_t cap) {
    FVG_TRY(check_solve_args(c, a, p, b, x, o));
    if (o->precision == FVG_PRECISION_MIXED) return solve_mixed(c, a, p, b, x, o, info, history, cap, true);
    return run_single(c, a, p, b, x, o, info, history, cap, true);
}

}  // extern "C"
