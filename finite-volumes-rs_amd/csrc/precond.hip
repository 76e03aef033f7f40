// precond.hip -- IncompleteCholesky / IncompleteLowerUpper / Jacobi / identity on the device.
//
// Analysis (host, once per pattern): dependency levels of the strictly lower and
// strictly upper triangles, rows sorted by level, and the slot maps that gather a
// factor into level-ordered sweep matrices.
//
// Factorisation (device): one launch per lower level, one thread per row, in the
// row-oriented (left-looking) form of each reference algorithm.  For every factor
// entry the sequence of floating-point operations is the reference's own, so the
// factors are bit-identical:
//   IC(0)  incomplete_cholesky.rs:87-145 -- l_ij = (1/l_jj) * (a_ij - sum_m l_im l_jm)
//          with the subtractions in ascending m, l_ii = sqrt(a_ii - sum_m l_im^2);
//   ILU(0) incomplete_lower_upper.rs:200-263 -- IKJ row elimination, then the drop
//          rule |v| <= 1e-6 |u_ii| (:255-258) applied as an explicit zero.
// Sweeps (device): one launch per level; each row subtracts its entries in ascending
// column order, as the reference's sequential loops do (incomplete_cholesky.rs:217-254,
// incomplete_lower_upper.rs:291-328).
#include <algorithm>
#include <cmath>
#include <numeric>
#include <cstdlib>
#include <cstring>

#include "kstate.hpp"
#include "mesh.hpp"
#include "precond.hpp"
#include "reduce.hpp"

namespace fvg {

// ---- kernels ------------------------------------------------------------------------

__global__ void k_sym_check(int64_t nnz, const int32_t* __restrict__ tpos, const double* __restrict__ a,
                            int32_t* __restrict__ flags) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nnz) return;
    const int32_t t = tpos[k];
    if (t < 0) return;
    if (!(fabs(a[k] - a[t]) < 1e-14)) atomicAdd(&flags[2], 1);   // ApproximateCmp for f64, linalg/mod.rs:73-77
}

// row i of IC(0): strictly lower entries, then the pivot (left-looking)
__device__ __forceinline__ void ic0_row(int i, const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                                        const int32_t* __restrict__ diag, const double* __restrict__ a, double* l,
                                        double* inv_diag, int32_t* flags) {
    const int s = rp[i], d = diag[i];
    for (int k = s; k < d; ++k) {
        const int j = col[k];
        double v = a[k];
        int x = s, y = rp[j];
        const int ye = diag[j];
        while (x < k && y < ye) {
            const int cx = col[x], cy = col[y];
            if (cx == cy) {
                v -= l[x] * l[y];
                ++x;
                ++y;
            } else if (cx < cy) {
                ++x;
            } else {
                ++y;
            }
        }
        l[k] = inv_diag[j] * v;
    }
    double dd = a[d];
    for (int k = s; k < d; ++k) dd -= l[k] * l[k];
    if (!(dd > 0.0)) atomicAdd(&flags[1], 1);
    const double lkk = sqrt(dd);
    l[d] = lkk;
    inv_diag[i] = 1.0 / lkk;
}

// row i of ILU(0): IKJ elimination against the rows above, then the drop rule
__device__ __forceinline__ void ilu0_row(int i, const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                                         const int32_t* __restrict__ diag, const double* __restrict__ a, double* lu,
                                         int32_t* flags) {
    const int s = rp[i], e = rp[i + 1], d = diag[i];
    for (int k = s; k < e; ++k) lu[k] = a[k];
    for (int k = s; k < d; ++k) {
        const int kk = col[k];
        const double ukk = lu[diag[kk]];
        if (fabs(ukk) < 1e-16) atomicAdd(&flags[0], 1);
        const double inv = 1.0 / ukk;
        lu[k] = inv * lu[k];
        const double lik = lu[k];
        int x = k + 1, y = diag[kk] + 1;
        const int ye = rp[kk + 1];
        while (x < e && y < ye) {
            const int cx = col[x], cy = col[y];
            if (cx == cy) {
                lu[x] -= lik * lu[y];
                ++x;
                ++y;
            } else if (cx < cy) {
                ++x;
            } else {
                ++y;
            }
        }
    }
    const double tol = fabs(lu[d]) * 1e-6;
    for (int k = s; k < e; ++k)
        if (k != d && !(fabs(lu[k]) > tol)) lu[k] = 0.0;
}

__global__ void k_ic0_level(int64_t begin, int64_t end, const int32_t* __restrict__ rows,
                            const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                            const int32_t* __restrict__ diag, const double* __restrict__ a, double* __restrict__ l,
                            double* __restrict__ inv_diag, int32_t* __restrict__ flags) {
    const int64_t p = begin + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= end) return;
    const int i = rows[p];
    if (i < 0) return;
    ic0_row(i, rp, col, diag, a, l, inv_diag, flags);
}

__global__ void k_ilu0_level(int64_t begin, int64_t end, const int32_t* __restrict__ rows,
                             const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                             const int32_t* __restrict__ diag, const double* __restrict__ a, double* __restrict__ lu,
                             int32_t* __restrict__ flags) {
    const int64_t p = begin + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= end) return;
    const int i = rows[p];
    if (i < 0) return;
    ilu0_row(i, rp, col, diag, a, lu, flags);
}

// IC(0) rows of one level, one wave per row (rows of at most 64 entries): lane t keeps
// l[s + t].  For each strictly lower entry k, the lanes of row i's earlier entries find
// their column among row j's lower entries (one wave-wide load of that row, shuffled) and
// form their products; every lane then subtracts the products in column order, so the
// sum has ic0_row's order and bits.
__global__ void __launch_bounds__(kBlock) k_ic0_level_wave(int64_t begin, int64_t end,
                                                           const int32_t* __restrict__ rows,
                                                           const int32_t* __restrict__ rp,
                                                           const int32_t* __restrict__ col,
                                                           const int32_t* __restrict__ diag,
                                                           const double* __restrict__ a, double* l,
                                                           double* inv_diag, int32_t* __restrict__ flags) {
    const int64_t p = begin + ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64;
    if (p >= end) return;   // the same for all lanes of a wave
    const int i = rows[p];
    if (i < 0) return;
    const int lane = threadIdx.x & 63;
    const int s = rp[i], d = diag[i];
    const bool mine = s + lane <= d;
    const int cx = mine ? col[s + lane] : -1;
    const double ax = mine ? a[s + lane] : 0.0;
    double lx = 0.0;
    for (int k = s; k < d; ++k) {
        const int j = __shfl(cx, k - s);
        const int yb = rp[j], ye = diag[j];
        const double inv = inv_diag[j];
        double prod = 0.0;
        bool hit = false;
        for (int y0 = yb; y0 < ye; y0 += 64) {
            const int m = min(64, ye - y0);
            const int cy = lane < m ? col[y0 + lane] : -2;
            const double ly = lane < m ? l[y0 + lane] : 0.0;
            for (int t = 0; t < m; ++t) {
                const int c = __shfl(cy, t);
                const double u = __shfl(ly, t);
                if (lane < k - s && cx == c) {
                    prod = lx * u;
                    hit = true;
                }
            }
        }
        double v = __shfl(ax, k - s);
        for (int x = 0; x < k - s; ++x) {
            const double px = __shfl(prod, x);
            if (__shfl((int)hit, x)) v -= px;
        }
        if (lane == k - s) lx = inv * v;
    }
    double dd = __shfl(ax, d - s);
    for (int x = 0; x < d - s; ++x) {
        const double q = __shfl(lx, x);
        dd -= q * q;
    }
    if (lane == 0 && !(dd > 0.0)) atomicAdd(&flags[1], 1);
    const double lkk = sqrt(dd);
    if (s + lane < d) l[s + lane] = lx;
    if (s + lane == d) {
        l[d] = lkk;
        inv_diag[i] = 1.0 / lkk;
    }
}

// ILU(0) rows of one level, one wave per row (rows of at most 64 entries): lane t keeps
// entry s + t of row i in a register, and for each lower entry k, in column order, the
// lanes read the pivot row's entries right of its diagonal as one wave-wide load and take
// them through shuffles.  Every entry receives the same updates in the same order as in
// ilu0_row, so the factor has the same bits; what changes is that no merge walks the two
// rows' columns one dependent load at a time.
__global__ void __launch_bounds__(kBlock) k_ilu0_level_wave(int64_t begin, int64_t end,
                                                            const int32_t* __restrict__ rows,
                                                            const int32_t* __restrict__ rp,
                                                            const int32_t* __restrict__ col,
                                                            const int32_t* __restrict__ diag,
                                                            const double* __restrict__ a, double* lu,
                                                            int32_t* __restrict__ flags) {
    const int64_t p = begin + ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64;
    if (p >= end) return;   // the same for all lanes of a wave
    const int i = rows[p];
    if (i < 0) return;
    const int lane = threadIdx.x & 63;
    const int s = rp[i], e = rp[i + 1], d = diag[i];
    const bool mine = s + lane < e;
    const int cx = mine ? col[s + lane] : -1;
    double v = mine ? a[s + lane] : 0.0;
    for (int k = s; k < d; ++k) {
        const int kk = __shfl(cx, k - s);
        const int dk = diag[kk], ek = rp[kk + 1];
        const double ukk = lu[dk];
        if (lane == 0 && fabs(ukk) < 1e-16) atomicAdd(&flags[0], 1);
        const double inv = 1.0 / ukk;
        if (lane == k - s) v = inv * v;
        const double lik = __shfl(v, k - s);
        for (int y0 = dk + 1; y0 < ek; y0 += 64) {
            const int m = min(64, ek - y0);
            const int cy = lane < m ? col[y0 + lane] : -2;
            const double uy = lane < m ? lu[y0 + lane] : 0.0;
            for (int t = 0; t < m; ++t) {
                const int c = __shfl(cy, t);
                const double u = __shfl(uy, t);
                if (mine && lane > k - s && cx == c) v -= lik * u;
            }
        }
    }
    const double tol = fabs(__shfl(v, d - s)) * 1e-6;   // the drop rule of ilu0_row
    if (mine) lu[s + lane] = (s + lane != d && !(fabs(v) > tol)) ? 0.0 : v;
}

// One workgroup refactorises one system of a group, level by level (a workgroup barrier
// between levels instead of a launch per level), then makes the inverse pivots (ILU)
// and gathers the sweep values: the same per-row arithmetic as the per-level launches.
struct FactorJob {
    int32_t kind, nlevels;
    const int64_t* lvl_ptr;   // positions of each lower level in rows
    const int32_t* rows;
    const int32_t* rp;
    const int32_t* col;
    const int32_t* diag;
    const int32_t* tpos;
    const double* a;
    double* fac;
    double* inv_diag;
    int32_t* flags;
    const int32_t* lo_src;
    double* lo_val;
    const int32_t* up_src;
    double* up_val;
    const int32_t* lo_bsrc;   // blocked sweep layouts: factor slot per slot, -1 padding
    double* lo_bval;
    const int32_t* up_bsrc;
    double* up_bval;
    const int32_t* lo_brow;   // blocked layouts: row of each position, and the scale there
    double* lo_spos;
    const int32_t* up_brow;
    double* up_spos;
    int64_t n, nnz, lo_nnz, up_nnz, lo_bne, up_bne, lo_bn, up_bn;
};

constexpr int kFactorBlock = 512;

__global__ void __launch_bounds__(kFactorBlock) k_factor_team(const FactorJob* __restrict__ jobs) {
    const FactorJob J = jobs[blockIdx.x];
    if (J.kind == kIC) {
        for (int64_t k = threadIdx.x; k < J.nnz; k += kFactorBlock) {   // symmetry (linalg/mod.rs:73-77)
            const int32_t t = J.tpos[k];
            if (t >= 0 && !(fabs(J.a[k] - J.a[t]) < 1e-14)) atomicAdd(&J.flags[2], 1);
        }
    }
    for (int l = 0; l < J.nlevels; ++l) {
        const int64_t b = J.lvl_ptr[l], e = J.lvl_ptr[l + 1];
        for (int64_t p = b + threadIdx.x; p < e; p += kFactorBlock) {
            const int i = J.rows[p];
            if (i < 0) continue;
            if (J.kind == kIC)
                ic0_row(i, J.rp, J.col, J.diag, J.a, J.fac, J.inv_diag, J.flags);
            else
                ilu0_row(i, J.rp, J.col, J.diag, J.a, J.fac, J.flags);
        }
        __syncthreads();
    }
    if (J.kind == kILU) {
        for (int64_t i = threadIdx.x; i < J.n; i += kFactorBlock) J.inv_diag[i] = 1.0 / J.fac[J.diag[i]];
    }
    for (int64_t k = threadIdx.x; k < J.lo_nnz; k += kFactorBlock) J.lo_val[k] = J.fac[J.lo_src[k]];
    for (int64_t k = threadIdx.x; k < J.up_nnz; k += kFactorBlock) J.up_val[k] = J.fac[J.up_src[k]];
    for (int64_t k = threadIdx.x; k < J.lo_bne; k += kFactorBlock) {
        const int32_t e = J.lo_bsrc[k];
        J.lo_bval[k] = e >= 0 ? J.fac[e] : 0.0;
    }
    for (int64_t k = threadIdx.x; k < J.up_bne; k += kFactorBlock) {
        const int32_t e = J.up_bsrc[k];
        J.up_bval[k] = e >= 0 ? J.fac[e] : 0.0;
    }
    __syncthreads();   // inv_diag complete
    for (int64_t p = threadIdx.x; p < J.lo_bn; p += kFactorBlock) J.lo_spos[p] = J.inv_diag[J.lo_brow[p]];
    for (int64_t p = threadIdx.x; p < J.up_bn; p += kFactorBlock) J.up_spos[p] = J.inv_diag[J.up_brow[p]];
}

// blocked sweep values: slot e takes factor slot src[e], or 0 for padding
__global__ void k_gather_pad(int64_t ne, const int32_t* __restrict__ src, const double* __restrict__ fac,
                             double* __restrict__ out) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < ne) {
        const int32_t k = src[e];
        out[e] = k >= 0 ? fac[k] : 0.0;
    }
}

__global__ void k_inv_diag(int64_t n, const int32_t* __restrict__ diag, const double* __restrict__ v,
                           double* __restrict__ inv) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) inv[i] = 1.0 / v[diag[i]];
}

__global__ void k_gather(int64_t nnz, const int32_t* __restrict__ src, const double* __restrict__ fac,
                         double* __restrict__ dst) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nnz) dst[k] = fac[src[k]];
}

__global__ void k_sweep_level(int64_t begin, int64_t end, const int32_t* __restrict__ rows,
                              const int32_t* __restrict__ ptr, const int32_t* __restrict__ col,
                              const double* __restrict__ val, const double* __restrict__ scale,
                              const double* __restrict__ b, double* __restrict__ x,
                              const int32_t* __restrict__ skip) {
    if (skip && *skip) return;
    const int64_t p = begin + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= end) return;
    const int i = rows[p];
    if (i < 0) return;
    double s = b[i];
    const int e = ptr[p + 1];
    for (int k = ptr[p]; k < e; ++k) s -= val[k] * x[col[k]];
    x[i] = scale ? scale[i] * s : s;
}

// ---- single-launch sweep -----------------------------------------------------------
// Every row of the sweep is one lane; a wave64 owns one chunk of 64 positions of one
// level, and chunks are numbered in level order, so every row a lane waits for lives
// in a chunk with a smaller number.  The kernel is persistent: a cooperative launch
// makes all its W waves co-resident, and wave w takes chunks w, w+W, w+2W, ... in
// order.  The smallest unfinished chunk therefore always belongs to a running wave
// whose earlier chunks are done and whose dependencies are all finished: the sweep
// always drains.  W bounds how many waves poll at once (and how far ahead of the
// dependency front they run).
//
// The output vector doubles as the completion flag: it is filled with a signalling
// NaN bit pattern (pending) that no arithmetic result can carry, and each row's value
// is published with one agent-scope store; dependants read that word with
// agent-scope loads, which are coherent across XCDs.  Every wait is bounded: after
// kSpinLimit polls (or once another lane has timed out) the wait gives up, flags the
// error and lets the sweep finish, so a kernel never hangs.
template <class T>
struct Bits;
template <>
struct Bits<double> {
    using U = uint64_t;
    static constexpr U pending = kPendingBits;
    __device__ static double value(U u) { return __longlong_as_double((long long)u); }
    __device__ static U bits(double v) { return (U)__double_as_longlong(v); }
};
template <>
struct Bits<float> {
    using U = uint32_t;
    static constexpr U pending = 0x7FB0DEADu;
    __device__ static float value(U u) { return __uint_as_float(u); }
    __device__ static U bits(float v) { return __float_as_uint(v); }
};

constexpr int kSpinLimit = 1 << 18;   // polls before a wait gives up (FVG_TUNE_SPIN_LIMIT overrides)

static int spin_limit(const fvg_ctx* c) { return c->tune_spin_limit > 0 ? (int)c->tune_spin_limit : kSpinLimit; }
constexpr int kWavesPerGroup = 4;
constexpr int kMaxRowDeps = 6;     // dependencies loaded together before any wait

template <class T>
__global__ void k_fill_pending(int64_t n, T* __restrict__ x, const int32_t* __restrict__ skip) {
    if (skip && *skip) return;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) reinterpret_cast<typename Bits<T>::U*>(x)[i] = Bits<T>::pending;
}

template <class T>
__device__ __forceinline__ typename Bits<T>::U load_word(const T* x, int j) {
    return __hip_atomic_load(reinterpret_cast<const typename Bits<T>::U*>(x) + j, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
}

template <class T>
__device__ __forceinline__ T wait_value(const T* x, int j, int32_t* err, int spin_limit) {
    using U = typename Bits<T>::U;
    U v = load_word(x, j);
    int spins = 0;
    while (v == Bits<T>::pending) {
        if (++spins > spin_limit ||
            ((spins & 255) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
            __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
        }
        __builtin_amdgcn_s_sleep(1);
        v = load_word(x, j);
    }
    return Bits<T>::value(v);
}

// POLL: 0 re-poll at once; 1 re-poll after a short sleep; 2 keep two polls in flight.
// SCALE: the row result is multiplied by scale[i].
//
// A row of at most kMaxRowDeps dependencies issues every load it will ever need before
// its first wait: the column, the coefficient and the dependency word of each entry, and
// b[i] and scale[i].  Entry indices are clamped into the row (or onto the padding entry
// that follows the last row, when the row has no entries), so no load is conditional
// and all of them are in flight together.  Once the last dependency lands, the row is a
// few multiply-subtracts and one store away from publishing its own value: the chain
// from one level to the next is one store and one poll, with no load of static data in
// between.  A re-poll reloads every dependency word; a word that was ready reads the
// same again, since each row is published once.
template <class T, int POLL, bool SCALE>
__global__ void __launch_bounds__(kChunk * kWavesPerGroup)
    k_sweep_persistent(int64_t nchunks, int64_t nwaves, const int32_t* __restrict__ rows,
                       const int32_t* __restrict__ ptr, const int32_t* __restrict__ col,
                       const T* __restrict__ val, const T* __restrict__ scale, const T* __restrict__ b, T* x,
                       int32_t* __restrict__ err, const int32_t* __restrict__ skip, int spin_limit,
                       int withhold) {
    using U = typename Bits<T>::U;
    if (skip && *skip) return;
    const int64_t w0 = (int64_t)blockIdx.x * kWavesPerGroup + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    for (int64_t c = w0; c < nchunks; c += nwaves) {
        const int64_t p = c * kChunk + lane;
        const int i = rows[p];
        if (i < 0) continue;
        const int kb = ptr[p], ke = ptr[p + 1], len = ke - kb;
        T s = b[i];
        const T sc = SCALE ? scale[i] : T(1);
        if (len <= kMaxRowDeps) {
            const int last = len > 0 ? ke - 1 : kb;
            int cj[kMaxRowDeps];
            T a[kMaxRowDeps];
            U w[kMaxRowDeps];
#pragma unroll
            for (int q = 0; q < kMaxRowDeps; ++q) {
                const int k = min(kb + q, last);
                cj[q] = col[k];
                a[q] = val[k];
            }
            // only the row's own dependencies are polled (padding slots read as ready)
#pragma unroll
            for (int q = 0; q < kMaxRowDeps; ++q) w[q] = q < len ? load_word(x, cj[q]) : U(0);
            auto pending = [&](const U* v) {
                bool pend = false;
#pragma unroll
                for (int q = 0; q < kMaxRowDeps; ++q) pend |= (q < len) && v[q] == Bits<T>::pending;
                return pend;
            };
            int spins = 0;
            auto give_up = [&]() {
                if (++spins > spin_limit ||
                    ((spins & 255) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
                    __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    return true;
                }
                return false;
            };
            if (POLL == 2) {
                // two polls in flight, issued about half a round trip apart: each check
                // waits only for the older poll, so a dependency that lands is seen
                // sooner than with one poll per round trip
                if (pending(w)) {
                    U v[kMaxRowDeps];
#pragma unroll
                    for (int q = 0; q < kMaxRowDeps; ++q) v[q] = w[q];
#pragma unroll
                    for (int q = 0; q < kMaxRowDeps; ++q)
                        if (q < len && w[q] == Bits<T>::pending) w[q] = load_word(x, cj[q]);
                    __builtin_amdgcn_s_sleep(6);
#pragma unroll
                    for (int q = 0; q < kMaxRowDeps; ++q)
                        if (q < len && v[q] == Bits<T>::pending) v[q] = load_word(x, cj[q]);
                    for (;;) {
                        if (!pending(w) || give_up()) break;
                        // merge what the younger poll already saw, then re-poll the rest
#pragma unroll
                        for (int q = 0; q < kMaxRowDeps; ++q)
                            if (w[q] == Bits<T>::pending) w[q] = v[q];
                        if (!pending(w)) break;
#pragma unroll
                        for (int q = 0; q < kMaxRowDeps; ++q)
                            if (q < len && w[q] == Bits<T>::pending) w[q] = load_word(x, cj[q]);
                        if (!pending(v)) {
#pragma unroll
                            for (int q = 0; q < kMaxRowDeps; ++q) w[q] = v[q];
                            break;
                        }
#pragma unroll
                        for (int q = 0; q < kMaxRowDeps; ++q)
                            if (q < len && v[q] == Bits<T>::pending) v[q] = load_word(x, cj[q]);
                    }
                }
            } else {
                for (;;) {
                    if (!pending(w) || give_up()) break;
                    if (POLL == 1) __builtin_amdgcn_s_sleep(1);
                    // re-poll only the words still pending: lanes whose dependencies have
                    // landed issue no requests
#pragma unroll
                    for (int q = 0; q < kMaxRowDeps; ++q)
                        if (q < len && w[q] == Bits<T>::pending) w[q] = load_word(x, cj[q]);
                }
            }
            // subtract in ascending column order, as the sequential sweep does
#pragma unroll
            for (int q = 0; q < kMaxRowDeps; ++q)
                if (q < len) s -= a[q] * Bits<T>::value(w[q]);
        } else {
            for (int k = kb; k < ke; ++k) s -= val[k] * wait_value(x, col[k], err, spin_limit);
        }
        const T r = SCALE ? sc * s : s;
        if (!withhold)
            __hip_atomic_store(reinterpret_cast<U*>(x) + i, Bits<T>::bits(r), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
    }
}

// ---- the blocked sweep ----------------------------------------------------------------
// One workgroup per block of the blocked layout, all co-resident (cooperative launch).
// A block walks its levels in order with one barrier per level.  The layout guarantees
// that a dependency between blocks spans at least two levels (blocks of whole gap-1
// components, or delayed levels; by default three), so a row's dependencies are either
// values its own block made one or two levels before (kept in an LDS ring of kBlockRing
// levels) or values from other blocks (or older ones of its own) at least two levels
// before.  The last wave of the workgroup is the loader: during level l it awaits, with
// device-scope loads, the outside values level l + 1 will read, whose polls it issued
// during level l - 1, puts them in LDS and issues the polls of level l + 2.  The other
// waves only read LDS for their operands; their own loads are the static data of each
// row (packed LDS slots, coefficients, right-hand side, scale: all in position order),
// fetched kBlockAhead levels before use, so no load of theirs waits on a poll.  Solved
// values are published in position order (xpos), so a level's stores are contiguous
// and polls address positions; k_sweep_finish scatters them to their rows afterwards.
// (Scattered per-row stores inside this kernel measured 40 us slower per sweep: they
// load the memory system that the polls' latency depends on.)  Each row subtracts its
// products in ascending column order from b, as the sequential sweep does: the same
// bits.  Cross-block waits only ever target lower levels, so the block at the lowest
// level always progresses; every wait is bounded and raises the error flag.
constexpr int kBlockAhead = 4;
constexpr int kFetchBatch = 4;   // polls in flight per loader lane

// before each blocked sweep: the right-hand side in position order (threads [0, n)) and
// the pending fill of the positions that some block fetches from outside (threads
// [n, n + m)); no other position is ever polled
template <class T>
__global__ void k_sweep_prepare(int64_t n, const int32_t* __restrict__ brow, const T* __restrict__ b,
                                T* __restrict__ bpos, int64_t m, const int32_t* __restrict__ rows, T* __restrict__ x,
                                const int32_t* __restrict__ skip) {
    if (skip && *skip) return;
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n)
        bpos[k] = b[brow[k]];
    else if (k - n < m)
        reinterpret_cast<typename Bits<T>::U*>(x)[rows[k - n]] = Bits<T>::pending;
}

// after each blocked sweep: its values, in position order, to their rows
template <class T>
__global__ void k_sweep_finish(int64_t n, const int32_t* __restrict__ brow, const T* __restrict__ xpos,
                               T* __restrict__ x, const int32_t* __restrict__ skip) {
    if (skip && *skip) return;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) x[brow[p]] = xpos[p];
}

// scale in position order (after each refactorisation)
__global__ void k_gather_rows(int64_t n, const int32_t* __restrict__ rows, const double* __restrict__ v,
                              double* __restrict__ out) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) out[p] = v[rows[p]];
}

template <class T, int K, int FUSE = 0>
struct BlockRow {
    int p;
    bool ok;
    uint64_t dw[K / 4];   // K 16-bit LDS slots, unpacked when the row is solved
    double2 a[K / 2];     // K coefficients
    T bv, sc;
    // fused CG work (kFuseXR: ap, x and p of this position; kFuseRZ: r of its lower
    // position)
    T f[FUSE == 1 ? 3 : FUSE == 2 ? 1 : 0];
};

// SPIN: the wait limit is the spin_limit argument (a tuning override) instead of the
// compile-time kSpinLimit, and `withhold` != 0 keeps every solved value out of xpos (the
// block still uses its own values from LDS), so each fetch from outside waits until its
// limit: the forced time-out of the tests.  The default launch keeps the constant limit
// and always publishes (a runtime limit in the polling loop measured 12 us slower per
// sweep).
// One block's partial of a fused sweep (t, valid in thread 0; called by every thread of
// the block).  With a counter, the partial is published with agent-scope stores and the
// block counts itself in; the last block to arrive sums all partials (one per lane of its
// first wave, agent-scope loads: coherent across XCDs without a cache flush) and runs
// the solver's scalar step.
__device__ inline void sweep_fuse_finish(const SweepFuse& F, int blk, DD t, int* s_last) {
    if (!F.counter) {
        if (threadIdx.x == 0) put_dd(F.partials, 0, 0, blk, t);
        return;
    }
    const int nb = (int)gridDim.x;
    if (threadIdx.x == 0) {
        __hip_atomic_store(F.partials + 2 * blk, t.hi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(F.partials + 2 * blk + 1, t.lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __builtin_amdgcn_s_waitcnt(0);   // both stores have completed before the count
        *s_last = __hip_atomic_fetch_add(F.counter, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == nb - 1;
    }
    __syncthreads();
    if (!*s_last || threadIdx.x >= 64) return;
    DD a{0.0, 0.0};
    for (int q = (int)threadIdx.x; q < nb; q += 64)
        a = dd_add(a, DD{__hip_atomic_load(F.partials + 2 * q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                         __hip_atomic_load(F.partials + 2 * q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)});
    a = wave_sum(a);
    if (threadIdx.x == 0) {
        scalar_step(F.step, F.st, F.hist, a.hi, 0.0, F.active, F.all_done);
        __hip_atomic_store(F.counter, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

template <class T, bool SCALE, int K, bool SPIN, int FUSE>
__global__ void __launch_bounds__(1024)
    k_sweep_blocked(int nlevels, int64_t n, int W, int X, const int32_t* __restrict__ brange,
                    const int32_t* __restrict__ blvl, const int32_t* __restrict__ bxptr,
                    const int32_t* __restrict__ bxidx, const uint16_t* __restrict__ bdep,
                    const T* __restrict__ bval, const T* __restrict__ bpos,
                    const T* __restrict__ spos, T* xpos, int32_t* __restrict__ err,
                    const int32_t* __restrict__ skip, int spin_limit, int withhold, int brev, SweepFuse F) {
    using U = typename Bits<T>::U;
    extern __shared__ __align__(16) unsigned char smem[];
    if (skip && *skip) return;
    const int blk = blockIdx.x;
    const int Lb = brange[2 * blk], Le = brange[2 * blk + 1];
    if (Lb >= Le) {
        if constexpr (FUSE != kFuseNone) {
            __shared__ int s_last0;
            sweep_fuse_finish(F, blk, DD{0.0, 0.0}, &s_last0);
        }
        return;
    }
    T* ring = reinterpret_cast<T*>(smem);
    T* stage = ring + kBlockRing * W;
    int* s_lv = reinterpret_cast<int*>(stage + 2 * X);
    const int32_t* lv = blvl + (int64_t)blk * (nlevels + 1);
    const int32_t* xp = bxptr + (int64_t)blk * (nlevels + 1);
    for (int t = threadIdx.x; t <= nlevels; t += blockDim.x) s_lv[t] = lv[t];
    __syncthreads();
    // The loader is the last wave.  The test is made wave-uniform so that the loader and
    // the compute waves run separate loops (each with one barrier per level), and the
    // compiler's tracking of outstanding loads in the compute loop is not merged with the
    // loader's polling loop.
    const int nc = (int)blockDim.x - 64;
    const bool loader = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6) == (nc >> 6);
    const int ct = threadIdx.x, lt = (int)threadIdx.x - nc;
    // both loops run to a multiple of kBlockAhead levels (the steps past Le only meet the
    // barrier), so the unrolled compute loop has no exit but its back edge
    const int Lend = Lb + (Le - Lb + kBlockAhead - 1) / kBlockAhead * kBlockAhead;
    DD acc{0.0, 0.0};   // fused dot product of this thread's rows

    if (loader) {
        // The outside values of level l go into staging slot l & 1, by a pipeline of loads
        // that keeps the loader's only LDS accesses to those staging writes.  At the start of
        // step s (after the barrier that ends level s - 1) the loader issues the polls of
        // level s + 2, from the positions it loaded during step s - 1, loads the positions
        // polled for level s + 3 (from the offsets it loaded during step s - 1) and the
        // offsets of level s + 4; then it awaits the polls of level s + 1, issued a whole
        // step before, re-polling only the pending ones, and writes them to LDS.  Where the
        // layout puts three or more levels between blocks, each value was published before
        // its poll was issued.  A level's polls past the first 64 * kFetchBatch are issued
        // and awaited when the level is awaited.
        // A level's loads pass through one of four register slots over four steps (offsets,
        // positions, polls, staging); the loop is unrolled by four so that every slot keeps
        // registers of its own and no value still in flight is ever copied.
        struct Slot {
            int a, b;                // bxidx offsets of the slot's level
            int jj[kFetchBatch];     // positions it polls (-1: none)
            U w[kFetchBatch];        // the polls
        };
        auto offsets = [&](int l, Slot& S) {   // empty past the block's last level
            const int lc = l < Le ? l : Le - 1;
            S.a = xp[lc];
            S.b = xp[lc + 1];
            if (l >= Le) S.b = S.a;
        };
        // Every load of the pipeline is unconditional (indices clamped into range, unused
        // results replaced afterwards), so each step issues the same number of them and the
        // waits before use count only what is older.
        auto positions_at = [&](int a, int b, int (&jj)[kFetchBatch]) {
#pragma unroll
            for (int q = 0; q < kFetchBatch; ++q) {
                const int k = a + q * 64 + lt;
                const int v = bxidx[k < b ? k : (b > 0 ? b - 1 : 0)];
                jj[q] = k < b ? v : -1;
            }
        };
        auto poll = [&](const int (&jj)[kFetchBatch], U (&w)[kFetchBatch]) {
#pragma unroll
            for (int q = 0; q < kFetchBatch; ++q) w[q] = load_word(xpos, jj[q] >= 0 ? jj[q] : 0);
        };
        auto settle = [&](const int (&jj)[kFetchBatch], U (&w)[kFetchBatch]) {
#pragma unroll
            for (int q = 0; q < kFetchBatch; ++q)
                if (jj[q] < 0) w[q] = U(0);
            for (int spins = 0;; ++spins) {
                bool pend = false;
#pragma unroll
                for (int q = 0; q < kFetchBatch; ++q) pend |= w[q] == Bits<T>::pending;
                if (!pend) break;
                if (spins > (SPIN ? spin_limit : kSpinLimit) ||
                    ((spins & 255) == 255 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
                    __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
#pragma unroll
                for (int q = 0; q < kFetchBatch; ++q)
                    if (w[q] == Bits<T>::pending) w[q] = load_word(xpos, jj[q]);
            }
        };
        // level l's values into its staging slot: the polls in flight, then any further ones
        auto stage_level = [&](int l, Slot& S) {
            T* st = stage + (l & 1) * X;
            settle(S.jj, S.w);
#pragma unroll
            for (int q = 0; q < kFetchBatch; ++q)
                if (S.jj[q] >= 0) st[q * 64 + lt] = Bits<T>::value(S.w[q]);
            for (int kb = S.a + 64 * kFetchBatch; kb < S.b; kb += 64 * kFetchBatch) {
                int jr[kFetchBatch];
                U wr[kFetchBatch];
                positions_at(kb, S.b, jr);
                poll(jr, wr);
                settle(jr, wr);
#pragma unroll
                for (int q = 0; q < kFetchBatch; ++q)
                    if (jr[q] >= 0) st[kb - S.a + q * 64 + lt] = Bits<T>::value(wr[q]);
            }
        };
        // the step of level l: offsets of l + 4 (into the slot level l left), positions of
        // l + 3, polls of l + 2, and level l + 1 staged
        auto step = [&](int l, Slot& S4, Slot& S3, Slot& S2, Slot& S1) {
            offsets(l + 4, S4);
            positions_at(S3.a, S3.b, S3.jj);
            poll(S2.jj, S2.w);
            stage_level(l + 1, S1);
            __syncthreads();
        };
        Slot S0, S1, S2, S3;
        // level Lb at once; the pipeline primed with levels Lb + 1 .. Lb + 3
        offsets(Lb, S0);
        offsets(Lb + 1, S1);
        offsets(Lb + 2, S2);
        offsets(Lb + 3, S3);
        positions_at(S0.a, S0.b, S0.jj);
        positions_at(S1.a, S1.b, S1.jj);
        positions_at(S2.a, S2.b, S2.jj);
        poll(S0.jj, S0.w);
        poll(S1.jj, S1.w);
        stage_level(Lb, S0);
        __syncthreads();
        static_assert(kBlockAhead % 4 == 0, "the loader's loop runs whole groups of four steps to Lend");
        for (int L0 = Lb; L0 < Lend; L0 += 4) {   // slot of level v: S[(v - Lb) % 4]
            step(L0, S0, S3, S2, S1);
            step(L0 + 1, S1, S0, S3, S2);
            step(L0 + 2, S2, S1, S0, S3);
            step(L0 + 3, S3, S2, S1, S0);
        }
        if constexpr (FUSE == kFuseNone) return;
    } else {

    // compute threads: the data of this thread's row of level l, all in position order.
    // The loads are unconditional (indices clamped into range) so that each level issues
    // the same number of them and the waits before use count only what is older.  (A
    // gather of b by row index inside this kernel, two or three levels ahead, measured
    // 20 % slower than the separate gather into position order.)
    const T alpha = FUSE == kFuseXR ? *F.alpha : T(0);
    auto prefetch = [&](int l, BlockRow<T, K, FUSE>& r) {
        const int lc = l < Le ? l : Le - 1;
        const int p0 = s_lv[lc], p1 = s_lv[lc + 1];
        const bool ok = l < Le && p0 + ct < p1;
        int64_t p = ok ? p0 + ct : p0;
        p = p < n ? p : n - 1;
        r.p = (int)p;
        const uint64_t* dw = reinterpret_cast<const uint64_t*>(bdep) + p * (K / 4);
        const double2* a2 = reinterpret_cast<const double2*>(bval) + p * (K / 2);
#pragma unroll
        for (int g = 0; g < K / 4; ++g) r.dw[g] = dw[g];
#pragma unroll
        for (int h = 0; h < K / 2; ++h) r.a[h] = a2[h];
        // brev: the right-hand side in reverse order; kFuseXR reads r through F (it writes it)
        r.bv = FUSE == kFuseXR ? F.r[p] : bpos[brev ? n - 1 - p : p];
        r.sc = SCALE ? spos[p] : T(1);
        r.ok = ok;
        if constexpr (FUSE == kFuseXR) {
            r.f[0] = F.ap[p];
            r.f[1] = F.x[p];
            r.f[2] = F.pp[p];
        } else if constexpr (FUSE == kFuseRZ) {
            r.f[0] = F.r[n - 1 - p];
        }
    };
    auto solve = [&](int l, const BlockRow<T, K, FUSE>& r) {
        if (!r.ok) return;
        // (reading only the row's used slots measured faster than issuing all K reads at
        // once with the unused ones pointed at slot 0)
        T s = r.bv;
        if constexpr (FUSE == kFuseXR) {   // conjugate_gradient.rs:77-80, as k_cg_xr
            F.x[r.p] = r.f[1] + r.f[2] * alpha;
            const T ri = r.bv - r.f[0] * alpha;
            F.r[r.p] = ri;
            acc = dd_add(acc, ri * ri);
            s = ri;
        }
#pragma unroll
        for (int q = 0; q < K; ++q) {
            const unsigned d = (unsigned)(r.dw[q / 4] >> (16 * (q % 4))) & 0xFFFFu;
            const T aq = (q & 1) ? r.a[q / 2].y : r.a[q / 2].x;
            if (d != kNoSlot) s -= aq * ring[d];
        }
        const T res = SCALE ? r.sc * s : s;
        if constexpr (FUSE == kFuseRZ) acc = dd_add(acc, r.f[0] * res);   // r.z, as k_cg_rz_pos
        ring[(l % kBlockRing) * W + ct] = res;
        if (!SPIN || !withhold)
            __hip_atomic_store(reinterpret_cast<U*>(xpos) + r.p, Bits<T>::bits(res), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
    };
    BlockRow<T, K, FUSE> r0, r1, r2, r3;
    prefetch(Lb, r0);
    prefetch(Lb + 1, r1);
    prefetch(Lb + 2, r2);
    prefetch(Lb + 3, r3);
    __syncthreads();
    // step l: solve level l, then fetch the data of level l + 4 into the same registers
#define FVG_BLOCK_STEP(R, OFF)                        \
    {                                                 \
        const int l = L0 + (OFF);                     \
        solve(l, R);                                  \
        prefetch(l + kBlockAhead, R);                 \
        __syncthreads();                              \
    }
    for (int L0 = Lb; L0 < Lend; L0 += kBlockAhead) {
        FVG_BLOCK_STEP(r0, 0)
        FVG_BLOCK_STEP(r1, 1)
        FVG_BLOCK_STEP(r2, 2)
        FVG_BLOCK_STEP(r3, 3)
    }
#undef FVG_BLOCK_STEP
    }   // compute waves
    if constexpr (FUSE != kFuseNone) {
        // the block's partial of the fused dot product (the loader wave adds zero)
        __shared__ double s_red[2 * 16];
        __shared__ int s_last;
        const DD v = wave_sum(acc);
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
        if (lane == 0) {
            s_red[2 * w] = v.hi;
            s_red[2 * w + 1] = v.lo;
        }
        __syncthreads();
        DD t{0.0, 0.0};
        if (threadIdx.x == 0)
            for (int q = 0; q < (int)(blockDim.x >> 6); ++q) t = dd_add(t, DD{s_red[2 * q], s_red[2 * q + 1]});
        sweep_fuse_finish(F, blk, t, &s_last);
    }
}

// ---- FP32 inner preconditioner: Jacobi-iterated triangular solves ----------------------
// The mixed-precision inner solver may apply the incomplete factor inexactly: the FP64
// refinement around it corrects whatever the inner iterations leave.  A triangular system
// T y = v, T = D + N with N strictly triangular, is approximated by m Jacobi sweeps from
// y0 = D^-1 v, y_{k+1} = D^-1 (v - N y_k): the first m + 1 terms of the Neumann series
// sum_k (-D^-1 N)^k D^-1 v.  N is nilpotent, so the series is exact after as many terms as
// the triangular sweep has levels, and a few terms come close when the factor is
// diagonally dominant.  Each sweep is one fully parallel pass over the factor entries, with
// no row waiting for another, so the m passes are bound by HBM bandwidth where the exact
// sweep is bound by the latency of its level-to-level chain.  With the same m for the lower
// and the upper sweep of IC(0) (whose upper factor is the lower one transposed), the
// applied operator is P^T P with P the truncated lower series: symmetric positive
// definite, as CG requires.  (Iterative triangular solves for preconditioning: Chow and
// Patel, SIAM J. Sci. Comput. 37 (2015); Anzt, Chow and Dongarra, Euro-Par 2015.)
//
// Each factor is held as its own CSR in natural row order (strict lower part, or strict
// upper part with columns inside the square block), with FP32 values, so a sweep is an
// SpMV with an update: y_new = D^-1 (v - N y_old).  The kernel has the SpMV's shape: a
// workgroup per tile of kJacRows consecutive rows (tiles in XCD-contiguous order), the
// tile's (column, value) range streamed with coalesced loads, every thread issuing all its
// entry loads and then all its gathers before using any, products staged in LDS, and each
// row summing its own products.  FIRST folds the start y0 = D^-1 v into the first sweep.
constexpr int kJacRows = 256;
constexpr int kJacStage = 2048;
constexpr int kJacUnroll = 4;
// DELTA: columns read as 16-bit offsets from the tile's first row (dcol) instead of col
template <bool FIRST, bool SCALE, bool DELTA>
__global__ void __launch_bounds__(kJacRows)
    k_jacobi32(int n, const int32_t* __restrict__ ptr, const int32_t* __restrict__ col,
               const int16_t* __restrict__ dcol, const float* __restrict__ val,
               const float* __restrict__ scale, const float* __restrict__ v, const float* __restrict__ yin,
               float* __restrict__ yout, const int32_t* __restrict__ skip, const float* __restrict__ dotw,
               double* __restrict__ dpart) {
    __shared__ float prod[kJacStage];
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int r0 = tile * kJacRows;
    const int r1 = min(r0 + kJacRows, n);
    const int base = ptr[r0];
    const int cnt = ptr[r1] - base;
    const int r = r0 + (int)threadIdx.x;
    const int rc = min(r, r1 - 1);
    const int rb = ptr[rc], re = ptr[rc + 1];
    const float vr = v[rc];
    const float sr = SCALE ? scale[rc] : 1.f;
    if (skip && *skip) return;
    const float* __restrict__ src = FIRST ? v : yin;
    float sum = 0.f;
    if (cnt <= kJacStage) {
        for (int k0 = threadIdx.x; k0 < cnt; k0 += kJacRows * kJacUnroll) {
            int c[kJacUnroll];
            float a[kJacUnroll], yv[kJacUnroll], sv[kJacUnroll];
#pragma unroll
            for (int u = 0; u < kJacUnroll; ++u) {
                const int k = min(k0 + u * kJacRows, cnt - 1);
                c[u] = DELTA ? r0 + (int)__builtin_nontemporal_load(dcol + base + k)
                             : __builtin_nontemporal_load(col + base + k);
                a[u] = __builtin_nontemporal_load(val + base + k);
            }
#pragma unroll
            for (int u = 0; u < kJacUnroll; ++u) {
                yv[u] = src[c[u]];
                sv[u] = (FIRST && SCALE) ? scale[c[u]] : 1.f;
            }
#pragma unroll
            for (int u = 0; u < kJacUnroll; ++u)
                if (k0 + u * kJacRows < cnt) prod[k0 + u * kJacRows] = a[u] * ((FIRST && SCALE) ? sv[u] * yv[u] : yv[u]);
        }
        __syncthreads();
        if (r < r1)
            for (int k = rb - base; k < re - base; ++k) sum += prod[k];
    } else if (r < r1) {
        for (int k = rb; k < re; ++k) {
            const int j = DELTA ? r0 + (int)dcol[k] : col[k];
            sum += val[k] * ((FIRST && SCALE) ? scale[j] * src[j] : src[j]);
        }
    }
    const float y = SCALE ? sr * (vr - sum) : vr - sum;
    if (r < r1) yout[r] = y;
    if (dotw) {   // dotw . yout over the tile's rows, at the tile's slot (as a row-ordered dot kernel)
        __shared__ float red[kJacRows / 64];
        const float q = block_sum<kJacRows>(r < r1 ? dotw[r] * y : 0.f, red);
        if (threadIdx.x == 0) dpart[tile] = (double)q;
    }
}

__global__ void k_gather_f32(int64_t nnz, const int32_t* __restrict__ src, const double* __restrict__ fac,
                             float* __restrict__ out) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nnz) out[k] = (float)fac[src[k]];
}

__global__ void k_to_f32(int64_t n, const double* __restrict__ s, float* __restrict__ d) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = (float)s[i];
}

__global__ void k_scale32(int64_t n, const float* __restrict__ d, const float* __restrict__ r, float* __restrict__ z,
                          const int32_t* __restrict__ skip) {
    if (skip && *skip) return;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) z[i] = d ? d[i] * r[i] : r[i];
}

// Jacobi (d = 1/a_ii, DiagonalPreconditioner mod.rs:60-66) or identity copy (d == nullptr, mod.rs:25-31)
__global__ void k_scale(int64_t n, const double* __restrict__ d, const double* __restrict__ r,
                        double* __restrict__ z, const int32_t* __restrict__ skip) {
    if (skip && *skip) return;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) z[i] = d ? d[i] * r[i] : r[i];
}

// ---- host analysis ----------------------------------------------------------------------

static void free_sweep(fvg_ctx* c, Sweep& s) {
    dev_free_n(c, s.d_lvl_ptr, s.nlevels + 1);
    dev_free_n(c, s.val32, s.nnz + 1);
    dev_free_n(c, s.rows, s.npos);
    dev_free_n(c, s.ptr, s.npos + 1);
    dev_free_n(c, s.col, s.nnz + 1);
    dev_free_n(c, s.src, s.nnz);
    dev_free_n(c, s.val, s.nnz + 1);
    const int64_t L1 = s.bL + 1, K = s.bK;
    dev_free_n(c, s.brange, 2 * (int64_t)s.nblk);
    dev_free_n(c, s.blvl, s.nblk * L1);
    dev_free_n(c, s.bxptr, s.nblk * L1);
    dev_free_n(c, s.bxidx, s.nbx);
    dev_free_n(c, s.brow, s.n);
    dev_free_n(c, s.bdep, K * s.n);
    dev_free_n(c, s.bsrc, K * s.n);
    dev_free_n(c, s.bval, K * s.n);
    dev_free_n(c, s.spos, s.n);
    dev_free_n(c, s.bpos, s.n);
    dev_free_n(c, s.bfill, s.nfill);
    dev_free_n(c, s.xpos, s.n);
    dev_free_n(c, s.bfrom, s.n);
    dev_free_n(c, s.bposof, s.n);
    dev_free_n(c, s.bpolled, s.n);
    s.h_brow.clear();
    s.h_assign = BlockAssign{};
    s.mirrored = false;
    s.nblk = 0;
    s.nfill = 0;
}

constexpr int kBlockMaxDeps = 8;
constexpr int32_t kBlockGrowCap = 128;   // rows of one level a grown block may reach
constexpr int kBlockMaxRows = 960;   // 15 compute waves + the loader wave
// Rows packed by component into at most `target` blocks (see build_blocked): the block
// of each component; returns the number of blocks used.  A block's rows of one level are
// processed together and the fullest block-level sets the pace of its level, while every
// value a block reads from another block costs a device-scope poll.  So the components,
// in order of their first row, go greedily to the block that they share the most
// dependencies with among those that stay within `cap` rows on every level (cap: the
// widest component's rows of one level, rounded up to a whole wave), else to the block
// whose fullest level grows least.
static int32_t pack_components(int64_t n, const std::vector<int32_t>& rp, const std::vector<int32_t>& col,
                               bool upper, const std::vector<int32_t>& comp_of, size_t C,
                               const std::vector<int32_t>& lev, int nlevels, int target,
                               std::vector<int32_t>& blk_of_comp) {
    auto take = [&](int64_t i, int32_t j) { return upper ? (j > i && j < n) : (j < i); };
    const int32_t nb = (int32_t)std::min<size_t>((size_t)target, C);
    // rows per (component, level), and the components joined to each by a dependency
    // two or more levels apart
    std::vector<int64_t> cl_key;
    cl_key.reserve((size_t)n);
    for (int64_t i = 0; i < n; ++i)
        cl_key.push_back((int64_t)comp_of[(size_t)i] * nlevels + lev[(size_t)i]);
    std::sort(cl_key.begin(), cl_key.end());
    std::vector<int32_t> cl_ptr(C + 1, 0), cl_lev, cl_cnt;
    for (size_t k = 0; k < cl_key.size();) {
        size_t e = k;
        while (e < cl_key.size() && cl_key[e] == cl_key[k]) ++e;
        const int32_t q = (int32_t)(cl_key[k] / nlevels);
        cl_lev.push_back((int32_t)(cl_key[k] % nlevels));
        cl_cnt.push_back((int32_t)(e - k));
        cl_ptr[(size_t)q + 1]++;
        k = e;
    }
    int32_t widest = 0;
    for (int32_t v : cl_cnt) widest = std::max(widest, v);
    for (size_t q = 0; q < C; ++q) cl_ptr[q + 1] += cl_ptr[q];
    // component adjacency (each dependency between two components, both ways, with
    // multiplicity), by a two-pass counting sort
    std::vector<int64_t> nb_ptr(C + 1, 0);
    auto each_link = [&](auto&& f) {
        for (int64_t i = 0; i < n; ++i)
            for (int32_t k = rp[(size_t)i]; k < rp[(size_t)i + 1]; ++k) {
                const int32_t j = col[(size_t)k];
                if (!take(i, j)) continue;
                const int32_t a = comp_of[(size_t)i], b = comp_of[(size_t)j];
                if (a != b) f(a, b);
            }
    };
    each_link([&](int32_t a, int32_t b) {
        nb_ptr[(size_t)a + 1]++;
        nb_ptr[(size_t)b + 1]++;
    });
    for (size_t q = 0; q < C; ++q) nb_ptr[q + 1] += nb_ptr[q];
    std::vector<int32_t> nb_of((size_t)nb_ptr[C]);
    {
        std::vector<int64_t> next(nb_ptr.begin(), nb_ptr.end() - 1);
        each_link([&](int32_t a, int32_t b) {
            nb_of[(size_t)next[(size_t)a]++] = b;
            nb_of[(size_t)next[(size_t)b]++] = a;
        });
    }
    const int32_t cap = (widest + 63) / 64 * 64;
    std::vector<int32_t> load((size_t)nb * nlevels, 0), aff((size_t)nb, 0), peak((size_t)nb);
    std::fill(blk_of_comp.begin(), blk_of_comp.end(), -1);
    int32_t used = 0;
    for (size_t q = 0; q < C; ++q) {
        for (int64_t k = nb_ptr[q]; k < nb_ptr[q + 1]; ++k) {
            const int32_t o = blk_of_comp[(size_t)nb_of[(size_t)k]];
            if (o >= 0) aff[(size_t)o]++;
        }
        int32_t best = -1;
        for (int32_t b = 0; b < nb; ++b) {
            int32_t m = 0;
            for (int32_t k = cl_ptr[q]; k < cl_ptr[q + 1]; ++k)
                m = std::max(m, load[(size_t)b * nlevels + cl_lev[(size_t)k]] + cl_cnt[(size_t)k]);
            peak[(size_t)b] = m;
            if (best < 0) {
                best = b;
                continue;
            }
            const bool fb = peak[(size_t)best] <= cap, fx = m <= cap;
            if (fx != fb ? fx
                         : (fx ? (aff[(size_t)b] > aff[(size_t)best] ||
                                  (aff[(size_t)b] == aff[(size_t)best] && m < peak[(size_t)best]))
                               : m < peak[(size_t)best]))
                best = b;
        }
        blk_of_comp[q] = best;
        used = std::max(used, best + 1);
        for (int32_t k = cl_ptr[q]; k < cl_ptr[q + 1]; ++k)
            load[(size_t)best * nlevels + cl_lev[(size_t)k]] += cl_cnt[(size_t)k];
        for (int64_t k = nb_ptr[q]; k < nb_ptr[q + 1]; ++k) {
            const int32_t o = blk_of_comp[(size_t)nb_of[(size_t)k]];
            if (o >= 0) aff[(size_t)o] = 0;
        }
    }
    // blocks are filled in increasing order, so the used ones are 0 .. used - 1
    return used;
}

// Blocks of components for a layout that puts three or more levels between blocks (see
// assign_blocked); returns the number of blocks.  Delaying every dependency between blocks
// lengthens any dependency chain that crosses between blocks often, so blocks are made
// compact: first each component is paired with the one it shares the most dependencies
// with (on extruded-like orderings those two span the same levels side by side, and a
// chain can step back and forth between them), then the pairs are grown into blocks of
// about n / target rows.  A block starts from the lowest unassigned pair and adds, among
// the unassigned pairs that share a dependency with it, the one that raises its fullest
// level least (the most shared dependencies on a tie), while that stays within `cap`
// rows.  (On 1,572,864 Kuhn tetrahedra: 88 blocks and 269 levels at a gap of three,
// against 319 for the blocks of pack_components.)
static int32_t grow_blocks(int64_t n, const std::vector<int32_t>& rp, const std::vector<int32_t>& col, bool upper,
                           const std::vector<int32_t>& comp_of, size_t C, const std::vector<int32_t>& lev, int target,
                           int32_t cap, std::vector<int32_t>& blk_of_comp) {
    auto take = [&](int64_t i, int32_t j) { return upper ? (j > i && j < n) : (j < i); };
    // dependencies between components, as (weight, a, b) with a < b
    std::vector<uint64_t> keys;
    for (int64_t i = 0; i < n; ++i)
        for (int32_t k = rp[(size_t)i]; k < rp[(size_t)i + 1]; ++k) {
            const int32_t j = col[(size_t)k];
            if (!take(i, j)) continue;
            uint32_t a = (uint32_t)comp_of[(size_t)i], b = (uint32_t)comp_of[(size_t)j];
            if (a == b) continue;
            if (a > b) std::swap(a, b);
            keys.push_back((uint64_t)a << 32 | b);
        }
    std::sort(keys.begin(), keys.end());
    struct Link {
        int32_t w, a, b;
    };
    std::vector<Link> links;
    for (size_t k = 0; k < keys.size();) {
        size_t e = k;
        while (e < keys.size() && keys[e] == keys[k]) ++e;
        links.push_back({(int32_t)(e - k), (int32_t)(keys[k] >> 32), (int32_t)(keys[k] & 0xFFFFFFFFu)});
        k = e;
    }
    keys.clear();
    keys.shrink_to_fit();
    // pairs along the heaviest links (ties: lower components first)
    std::vector<Link> order(links);
    std::stable_sort(order.begin(), order.end(), [](const Link& x, const Link& y) { return x.w > y.w; });
    std::vector<int32_t> grp(C, -1);
    int32_t G = 0;
    for (const Link& l : order)
        if (grp[(size_t)l.a] < 0 && grp[(size_t)l.b] < 0) grp[(size_t)l.a] = grp[(size_t)l.b] = G++;
    for (size_t q = 0; q < C; ++q)
        if (grp[q] < 0) grp[q] = G++;
    // pairs in order of their lowest component
    {
        std::vector<int32_t> first((size_t)G, INT32_MAX);
        for (size_t q = 0; q < C; ++q) first[(size_t)grp[q]] = std::min(first[(size_t)grp[q]], (int32_t)q);
        std::vector<int32_t> rank((size_t)G);
        std::iota(rank.begin(), rank.end(), 0);
        std::sort(rank.begin(), rank.end(), [&](int32_t x, int32_t y) { return first[(size_t)x] < first[(size_t)y]; });
        std::vector<int32_t> renum((size_t)G);
        for (int32_t r = 0; r < G; ++r) renum[(size_t)rank[(size_t)r]] = r;
        for (size_t q = 0; q < C; ++q) grp[q] = renum[(size_t)grp[q]];
    }
    // rows of each pair per level (sorted (level, count) runs) and the pairs' links
    std::vector<int64_t> gl;
    gl.reserve((size_t)n);
    int nlev = 0;
    for (int64_t i = 0; i < n; ++i) {
        gl.push_back((int64_t)grp[(size_t)comp_of[(size_t)i]] << 32 | (uint32_t)lev[(size_t)i]);
        nlev = std::max(nlev, lev[(size_t)i] + 1);
    }
    std::sort(gl.begin(), gl.end());
    std::vector<int64_t> gp((size_t)G + 1, 0), gsize((size_t)G, 0);
    std::vector<int32_t> glev, gcnt;
    for (size_t k = 0; k < gl.size();) {
        size_t e = k;
        while (e < gl.size() && gl[e] == gl[k]) ++e;
        const int32_t g = (int32_t)(gl[k] >> 32);
        glev.push_back((int32_t)(gl[k] & 0xFFFFFFFF));
        gcnt.push_back((int32_t)(e - k));
        gp[(size_t)g + 1]++;
        gsize[(size_t)g] += (int64_t)(e - k);
        k = e;
    }
    for (int32_t g = 0; g < G; ++g) gp[(size_t)g + 1] += gp[(size_t)g];
    std::vector<int64_t> ap((size_t)G + 1, 0);
    for (const Link& l : links)
        if (grp[(size_t)l.a] != grp[(size_t)l.b]) {
            ap[(size_t)grp[(size_t)l.a] + 1]++;
            ap[(size_t)grp[(size_t)l.b] + 1]++;
        }
    for (int32_t g = 0; g < G; ++g) ap[(size_t)g + 1] += ap[(size_t)g];
    std::vector<std::pair<int32_t, int32_t>> adj((size_t)ap[(size_t)G]);   // (pair, weight)
    {
        std::vector<int64_t> nx(ap.begin(), ap.end() - 1);
        for (const Link& l : links) {
            const int32_t x = grp[(size_t)l.a], y = grp[(size_t)l.b];
            if (x == y) continue;
            adj[(size_t)nx[(size_t)x]++] = {y, l.w};
            adj[(size_t)nx[(size_t)y]++] = {x, l.w};
        }
    }
    // growth
    const int64_t goal = (n + target - 1) / target;
    std::vector<int32_t> blk_of_grp((size_t)G, -1), load((size_t)nlev, 0), touched;
    std::vector<int32_t> front;
    int32_t nb = 0;
    for (int32_t seed = 0; seed < G; ++seed) {
        if (blk_of_grp[(size_t)seed] >= 0) continue;
        const int32_t b = nb++;
        int64_t rows = 0;
        for (int32_t l : touched) load[(size_t)l] = 0;
        touched.clear();
        front.clear();
        auto add = [&](int32_t g) {
            blk_of_grp[(size_t)g] = b;
            for (int64_t k = gp[(size_t)g]; k < gp[(size_t)g + 1]; ++k) {
                if (load[(size_t)glev[(size_t)k]] == 0) touched.push_back(glev[(size_t)k]);
                load[(size_t)glev[(size_t)k]] += gcnt[(size_t)k];
            }
            rows += gsize[(size_t)g];
            for (int64_t k = ap[(size_t)g]; k < ap[(size_t)g + 1]; ++k)
                if (blk_of_grp[(size_t)adj[(size_t)k].first] < 0) front.push_back(adj[(size_t)k].first);
        };
        add(seed);
        while (rows < goal) {
            int32_t best = -1, best_peak = INT32_MAX, best_w = 0;
            for (int32_t g : front) {
                if (blk_of_grp[(size_t)g] >= 0) continue;
                int32_t peak = 0, wsum = 0;
                for (int64_t k = gp[(size_t)g]; k < gp[(size_t)g + 1]; ++k)
                    peak = std::max(peak, load[(size_t)glev[(size_t)k]] + gcnt[(size_t)k]);
                for (int64_t k = ap[(size_t)g]; k < ap[(size_t)g + 1]; ++k)
                    if (blk_of_grp[(size_t)adj[(size_t)k].first] == b) wsum += adj[(size_t)k].second;
                if (peak < best_peak || (peak == best_peak && (wsum > best_w || (wsum == best_w && g < best)))) {
                    best = g;
                    best_peak = peak;
                    best_w = wsum;
                }
            }
            if (best < 0 || best_peak > cap) break;
            add(best);
            front.erase(std::remove_if(front.begin(), front.end(),
                                       [&](int32_t g) { return blk_of_grp[(size_t)g] >= 0; }),
                        front.end());
        }
    }
    for (size_t q = 0; q < C; ++q) blk_of_comp[q] = blk_of_grp[(size_t)grp[q]];
    return nb;
}

// Delayed levels of a sweep whose rows are grouped by comp_of: a row's level is one past
// its deepest dependency in its own group and `gap` past one in another group; returns
// the number of levels.
static int delayed_levels(int64_t n, const std::vector<int32_t>& rp, const std::vector<int32_t>& col, bool upper,
                          const std::vector<int32_t>& comp_of, int gap, std::vector<int32_t>& dlev) {
    auto take = [&](int64_t i, int32_t j) { return upper ? (j > i && j < n) : (j < i); };
    dlev.assign((size_t)n, 0);
    int maxl = 0;
    for (int64_t t = 0; t < n; ++t) {
        const int64_t i = upper ? n - 1 - t : t;
        int l = 0;
        for (int32_t k = rp[(size_t)i]; k < rp[(size_t)i + 1]; ++k) {
            const int32_t j = col[(size_t)k];
            if (take(i, j))
                l = std::max(l, dlev[(size_t)j] + (comp_of[(size_t)j] == comp_of[(size_t)i] ? 1 : gap));
        }
        dlev[(size_t)i] = l;
        maxl = std::max(maxl, l);
    }
    return maxl + 1;
}

// Boxes of a coordinate grid: the bounding box of the rows' coordinates cut into equal
// slabs along each axis, `count` boxes (a power of two) in the most nearly cubic grid for
// the extents, boxes ordered z-major.  A row's slab along each axis is then raised to the
// highest slab among its dependencies (lower sweep, rows ascending; the lowest, rows
// descending, for an upper sweep), so that along every chain of dependencies the slab
// indices never go back and a chain crosses each cut at most once: a chain then meets at
// most (s0 - 1) + (s1 - 1) + (s2 - 1) delays, however long it is.  Returns the box count
// (1, no boxes, where a coordinate is not finite).
static int32_t coordinate_boxes(const std::vector<int32_t>& rp, const std::vector<int32_t>& col, int64_t n,
                                bool upper, const RowCoords& X, int count, std::vector<int32_t>& blk) {
    auto take = [&](int64_t i, int32_t j) { return upper ? (j > i && j < n) : (j < i); };
    const int dim = std::min(X.dim, 3);
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    for (int a = 0; a < dim; ++a) {
        lo[a] = hi[a] = X.v[a];
        for (int64_t i = 0; i < n; ++i) {
            const double v = X.v[(size_t)i * X.dim + a];
            if (!std::isfinite(v)) return 1;
            lo[a] = std::min(lo[a], v);
            hi[a] = std::max(hi[a], v);
        }
    }
    // the grid: s[a] = 2^e[a], sum e = log2(count); least box aspect ratio, ties to more
    // cuts along the later axes
    int k = 0;
    while ((2 << k) <= count) ++k;
    int s[3] = {1, 1, 1};
    double best = 1e300;
    for (int e0 = 0; e0 <= k; ++e0)
        for (int e1 = 0; e0 + e1 <= k; ++e1) {
            const int e[3] = {e0, e1, k - e0 - e1};
            double mx = 0.0, mn = 1e300;
            bool ok = true;
            for (int a = 0; a < 3; ++a) {
                const double ext = a < dim ? hi[a] - lo[a] : 0.0;
                if (!(ext > 0.0)) {
                    ok = ok && e[a] == 0;
                    continue;
                }
                const double side = ext / (double)(1 << e[a]);
                mx = std::max(mx, side);
                mn = std::min(mn, side);
            }
            if (!ok || mn > mx) continue;
            const double aspect = mx / mn;
            if (aspect < best) {
                best = aspect;
                for (int a = 0; a < 3; ++a) s[a] = 1 << e[a];
            }
        }
    std::vector<int32_t> slab[3];
    for (int a = 0; a < 3; ++a) {
        slab[a].assign((size_t)n, 0);
        if (s[a] == 1) continue;
        const double scale = (double)s[a] / (hi[a] - lo[a]);
        for (int64_t i = 0; i < n; ++i) {
            const int v = (int)(((double)X.v[(size_t)i * X.dim + a] - lo[a]) * scale);
            slab[a][(size_t)i] = std::min(std::max(v, 0), s[a] - 1);
        }
        for (int64_t t = 0; t < n; ++t) {
            const int64_t i = upper ? n - 1 - t : t;
            int32_t v = slab[a][(size_t)i];
            for (int32_t q = rp[(size_t)i]; q < rp[(size_t)i + 1]; ++q)
                if (take(i, col[(size_t)q]))
                    v = upper ? std::min(v, slab[a][(size_t)col[(size_t)q]]) : std::max(v, slab[a][(size_t)col[(size_t)q]]);
            slab[a][(size_t)i] = v;
        }
    }
    blk.resize((size_t)n);
    for (int64_t i = 0; i < n; ++i)
        blk[(size_t)i] = (slab[2][(size_t)i] * s[1] + slab[1][(size_t)i]) * s[0] + slab[0][(size_t)i];
    return s[0] * s[1] * s[2];
}

// The widest block-level (rows) of an assignment and how many block-levels hold rows.
static void block_level_stats(int64_t n, const std::vector<int32_t>& blk, int32_t nb, const std::vector<int32_t>& lev,
                              int nlevels, int32_t& widest, int64_t& busy) {
    std::vector<int32_t> cnt((size_t)nb * (size_t)nlevels, 0);
    widest = 0;
    busy = 0;
    for (int64_t i = 0; i < n; ++i) {
        const int32_t v = ++cnt[(size_t)blk[(size_t)i] * (size_t)nlevels + (size_t)lev[(size_t)i]];
        widest = std::max(widest, v);
        busy += v == 1;
    }
}

// The block and level assignment of a blocked sweep layout (see k_sweep_blocked).  Rows
// joined by a dependency one level apart are kept in one block (union-find over those
// edges) and the components are packed into at most `target` blocks (pack_components), or,
// with `gap` > 2, grown into compact blocks of about n / target rows (grow_blocks);
// when one component holds more than an eighth of the rows, the blocks are instead
// target_consecutive ranges of consecutive rows, or boxes of the rows' coordinates
// (coordinate_boxes) where X has them and they cost less.  Levels are then delayed
// (delayed_levels) so that every dependency between blocks spans at least `gap` levels
// (gap >= 2): with three or more, a block's loader wave has a whole level to fetch each
// value from outside (the components alone give two).  False (no blocked layout) when
// neither the consecutive blocks nor the boxes keep within half again the levels, 8
// blocks busy per level and kBlockMaxRows rows per block-level, or when a row has more
// than kBlockMaxDeps dependencies: the persistent sweep then serves.
static bool assign_blocked(const std::vector<int32_t>& rp, const std::vector<int32_t>& col, int64_t n, bool upper,
                           const std::vector<int32_t>& lev_in, int nlevels_in, int target, int target_consecutive,
                           int gap, const CoordSource& coords, const std::vector<int>& box_counts, BlockAssign& A) {
    auto take = [&](int64_t i, int32_t j) { return upper ? (j > i && j < n) : (j < i); };
    if (n < 4096 || target < 2) return false;
    std::vector<int32_t> par((size_t)n);
    for (int64_t i = 0; i < n; ++i) par[(size_t)i] = (int32_t)i;
    auto find = [&](int32_t a) {
        int32_t r = a;
        while (par[(size_t)r] != r) r = par[(size_t)r];
        while (par[(size_t)a] != r) {
            const int32_t nx = par[(size_t)a];
            par[(size_t)a] = r;
            a = nx;
        }
        return r;
    };
    int kmax = 0;
    for (int64_t i = 0; i < n; ++i) {
        int cnt = 0;
        for (int32_t k = rp[(size_t)i]; k < rp[(size_t)i + 1]; ++k) {
            const int32_t j = col[(size_t)k];
            if (!take(i, j)) continue;
            ++cnt;
            if (lev_in[(size_t)i] - lev_in[(size_t)j] == 1) {
                const int32_t a = find((int32_t)i), b = find(j);
                if (a != b) par[(size_t)a] = b;
            }
        }
        kmax = std::max(kmax, cnt);
    }
    if (kmax > kBlockMaxDeps) return false;
    kmax = kmax <= 4 ? 4 : kBlockMaxDeps;   // the kernel reads exactly 4 or 8 slots per row
    // components, ordered by their first row, packed into blocks
    std::vector<int32_t> comp_of((size_t)n), comp_size;
    std::vector<int32_t> id_of_root((size_t)n, -1);
    for (int64_t i = 0; i < n; ++i) {
        const int32_t r = find((int32_t)i);
        if (id_of_root[(size_t)r] < 0) {
            id_of_root[(size_t)r] = (int32_t)comp_size.size();
            comp_size.push_back(0);
        }
        comp_of[(size_t)i] = id_of_root[(size_t)r];
        comp_size[(size_t)comp_of[(size_t)i]]++;
    }
    bool contiguous = false;
    for (int32_t sz : comp_size)
        if ((int64_t)sz * 8 > n) contiguous = true;
    // One component holding much of the matrix (every dependency one level deep, as on
    // extruded and 2D meshes): blocks of consecutive rows instead, with delayed levels.
    // A row's level is then one past its deepest dependency in its own block and `gap`
    // past one in another block (two, the least the kernel requires, where `gap` would
    // cost more than half again the levels).  Built only when that costs at most half
    // again the levels and keeps at least 8 blocks busy per level on average (on a 2D mesh
    // the consecutive blocks would run one after another).
    std::vector<int32_t> dlev;
    int nlevels = nlevels_in;
    if (contiguous) {
        const int32_t nbc = (int32_t)std::min<int64_t>(target_consecutive, n / 64);
        comp_size.assign((size_t)nbc, 0);
        for (int64_t i = 0; i < n; ++i) {
            comp_of[(size_t)i] = (int32_t)(i * nbc / n);
            comp_size[(size_t)comp_of[(size_t)i]]++;
        }
        const int gap_asked = gap;
        nlevels = delayed_levels(n, rp, col, upper, comp_of, gap, dlev);
        if (gap > 2 && (int64_t)nlevels * 2 > (int64_t)nlevels_in * 3)   // the least delay instead
            nlevels = delayed_levels(n, rp, col, upper, comp_of, gap = 2, dlev);
        // A layout is usable when it costs at most half again the levels, keeps 8 blocks
        // busy per level on average and no block-level exceeds kBlockMaxRows rows; a level
        // costs about 19 + (waves of the widest block-level) units (on 2,097,152 prisms an
        // ILU(0) apply took 603.7 us with 512 blocks of consecutive rows, 459 levels of at
        // most 128 rows; 500.8 us with 64 boxes, 337 levels of 528 rows; 439.4 us with 128
        // boxes, 345 levels of 272 rows).
        auto cost = [&](const std::vector<int32_t>& blk, int32_t nbk, const std::vector<int32_t>& lv, int L) -> int64_t {
            if ((int64_t)L * 2 > (int64_t)nlevels_in * 3) return INT64_MAX;
            int32_t w = 0;
            int64_t busy = 0;
            block_level_stats(n, blk, nbk, lv, L, w, busy);
            if (busy < 8 * (int64_t)L || w > kBlockMaxRows) return INT64_MAX;
            return (int64_t)L * (19 + (w + 63) / 64);
        };
        int64_t best = cost(comp_of, nbc, dlev, nlevels);
        // boxes of the rows' coordinates, where the matrix has them: kept when they cost less
        const RowCoords X = coords && !box_counts.empty() ? coords() : RowCoords();
        if (X.v && X.dim > 0) {
            std::vector<int32_t> bblk, blev;
            for (int count : box_counts) {
                const int32_t nbx = coordinate_boxes(rp, col, n, upper, X, count, bblk);
                if (nbx < 2) continue;
                const int L = delayed_levels(n, rp, col, upper, bblk, gap_asked, blev);
                const int64_t k = cost(bblk, nbx, blev, L);
                if (k >= best) continue;
                best = k;
                nlevels = L;
                comp_of.swap(bblk);
                dlev.swap(blev);
                comp_size.assign((size_t)nbx, 0);
                for (int64_t i = 0; i < n; ++i) comp_size[(size_t)comp_of[(size_t)i]]++;
            }
        }
        if (best == INT64_MAX) return false;
    }
    // blocks: the consecutive row ranges as they are, or the components packed
    const size_t C = comp_size.size();
    int32_t nb = 0;
    std::vector<int32_t> blk_of_comp(C);
    if (contiguous) {
        for (size_t q = 0; q < C; ++q) blk_of_comp[q] = (int32_t)q;
        nb = (int32_t)C;
    } else if (gap > 2) {
        nb = grow_blocks(n, rp, col, upper, comp_of, C, lev_in, target, kBlockGrowCap, blk_of_comp);
    } else {
        nb = pack_components(n, rp, col, upper, comp_of, C, lev_in, nlevels_in, target, blk_of_comp);
    }
    A.nb = nb;
    A.kmax = kmax;
    A.blk.resize((size_t)n);
    for (int64_t i = 0; i < n; ++i) A.blk[(size_t)i] = blk_of_comp[(size_t)comp_of[(size_t)i]];
    if (contiguous) {
        A.nlevels = nlevels;
        A.lev = std::move(dlev);
    } else if (gap > 2) {   // the components give two levels between blocks; delay for more
        A.nlevels = delayed_levels(n, rp, col, upper, A.blk, gap, A.lev);
    } else {
        A.nlevels = nlevels_in;
        A.lev = lev_in;
    }
    return true;
}

// The blocked layout of an assignment (block and level of every row).  Positions run block
// by block and level by level; within a block-level rows ascend, or descend with
// `descending` (a mirrored layout).  Returns FVG_OK with sw.nblk = 0 when the layout does not
// fit the kernel (a block-level beyond kBlockMaxRows rows, LDS slots beyond 16 bits).
static int materialize_blocked(fvg_ctx* c, const std::vector<int32_t>& rp, const std::vector<int32_t>& col, int64_t n,
                               bool upper, const std::vector<int32_t>* src_map, const BlockAssign& A, bool descending,
                               Sweep& sw) {
    auto take = [&](int64_t i, int32_t j) { return upper ? (j > i && j < n) : (j < i); };
    const int32_t nb = A.nb;
    const int nlevels = A.nlevels, kmax = A.kmax;
    const std::vector<int32_t>& lev = A.lev;
    const std::vector<int32_t>& blk = A.blk;
    sw.bW = 0;
    sw.bF = 0;
    // positions: block by block, level by level, rows ascending (or descending)
    std::vector<int32_t> order((size_t)n);
    for (int64_t i = 0; i < n; ++i) order[(size_t)i] = (int32_t)i;
    std::sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
        const int32_t ba = blk[(size_t)a], bb = blk[(size_t)b];
        if (ba != bb) return ba < bb;
        if (lev[(size_t)a] != lev[(size_t)b]) return lev[(size_t)a] < lev[(size_t)b];
        return descending ? a > b : a < b;
    });
    const int64_t L1 = nlevels + 1;
    std::vector<int32_t> blvl((size_t)(nb * L1), 0), brange((size_t)(2 * nb), 0), pos_of((size_t)n);
    for (int64_t p = 0; p < n; ++p) pos_of[(size_t)order[(size_t)p]] = (int32_t)p;
    {
        int64_t p = 0;
        for (int32_t b = 0; b < nb; ++b) {
            int first = -1, last = -1;
            for (int l = 0; l < nlevels; ++l) {
                blvl[(size_t)(b * L1 + l)] = (int32_t)p;
                const int64_t p0 = p;
                while (p < n && blk[(size_t)order[(size_t)p]] == b && lev[(size_t)order[(size_t)p]] == l)
                    ++p;
                if (p > p0) {
                    if (first < 0) first = l;
                    last = l;
                }
                if (p - p0 > kBlockMaxRows) return FVG_OK;
                sw.bW = std::max<int32_t>(sw.bW, (int32_t)(p - p0));
            }
            blvl[(size_t)(b * L1 + nlevels)] = (int32_t)p;
            brange[(size_t)(2 * b)] = first < 0 ? 0 : first;
            brange[(size_t)(2 * b + 1)] = first < 0 ? 0 : last + 1;
        }
    }
    const int32_t W = sw.bW;
    auto blk_of_row = [&](int32_t i) { return blk[(size_t)i]; };
    // dependencies: a value of the same block less than kBlockRing levels back is read from
    // the block's LDS ring; any other is fetched into the level's staging slots
    std::vector<int32_t> bdep((size_t)(kmax * n), -1 - (1 << 30)), bsrc((size_t)(kmax * n), -1);
    std::vector<int32_t> bxptr((size_t)(nb * L1), 0), bxidx;
    std::vector<std::vector<int32_t>> need;   // per level of the block being built: fetched rows
    int32_t X = 0;
    for (int32_t b = 0; b < nb; ++b) {
        // pass 1: fetched rows per level, sorted and unique
        std::vector<std::vector<int32_t>> lists((size_t)nlevels);
        for (int l = brange[(size_t)(2 * b)]; l < brange[(size_t)(2 * b + 1)]; ++l)
            for (int32_t p = blvl[(size_t)(b * L1 + l)]; p < blvl[(size_t)(b * L1 + l + 1)]; ++p) {
                const int32_t i = order[(size_t)p];
                for (int32_t k = rp[(size_t)i]; k < rp[(size_t)i + 1]; ++k) {
                    const int32_t j = col[(size_t)k];
                    if (!take(i, j)) continue;
                    const bool ring = blk_of_row(j) == b && lev[(size_t)i] - lev[(size_t)j] < kBlockRing;
                    if (!ring) lists[(size_t)l].push_back(j);
                }
            }
        for (int l = 0; l < nlevels; ++l) {
            auto& v = lists[(size_t)l];
            std::sort(v.begin(), v.end());
            v.erase(std::unique(v.begin(), v.end()), v.end());
            bxptr[(size_t)(b * L1 + l)] = (int32_t)bxidx.size();
            for (int32_t j : v) bxidx.push_back(pos_of[(size_t)j]);   // polled by position
            X = std::max<int32_t>(X, (int32_t)v.size());
        }
        bxptr[(size_t)(b * L1 + nlevels)] = (int32_t)bxidx.size();
        sw.bF = std::max<int32_t>(sw.bF, bxptr[(size_t)(b * L1 + nlevels)] - bxptr[(size_t)(b * L1)]);
        need.swap(lists);
        // pass 2: descriptors (slots are resolved once X is known: stored as (kind, index))
        for (int l = brange[(size_t)(2 * b)]; l < brange[(size_t)(2 * b + 1)]; ++l)
            for (int32_t p = blvl[(size_t)(b * L1 + l)]; p < blvl[(size_t)(b * L1 + l + 1)]; ++p) {
                const int32_t i = order[(size_t)p];
                int q = 0;
                for (int32_t k = rp[(size_t)i]; k < rp[(size_t)i + 1]; ++k) {
                    const int32_t j = col[(size_t)k];
                    if (!take(i, j)) continue;
                    const size_t e = (size_t)((int64_t)p * kmax + q);
                    bsrc[e] = src_map ? (*src_map)[(size_t)k] : k;
                    if (blk_of_row(j) == b && lev[(size_t)i] - lev[(size_t)j] < kBlockRing) {
                        const int lj = lev[(size_t)j];
                        bdep[e] = (lj % kBlockRing) * W + (pos_of[(size_t)j] - blvl[(size_t)(b * L1 + lj)]);
                    } else {
                        const auto& v = need[(size_t)l];
                        const int32_t idx = (int32_t)(std::lower_bound(v.begin(), v.end(), j) - v.begin());
                        bdep[e] = -1 - ((l & 1) * (1 << 20) + idx);   // resolved below
                    }
                    ++q;
                }
            }
    }
    // LDS slots must fit 16 bits (kNoSlot marks the unused ones)
    if (X >= (1 << 20) || (int64_t)kBlockRing * W + 2 * (int64_t)X >= kNoSlot) return FVG_OK;
    std::vector<uint16_t> bdep16(bdep.size());
    for (size_t e = 0; e < bdep.size(); ++e) {
        int32_t d = bdep[e];
        if (d == -1 - (1 << 30)) {
            bdep16[e] = kNoSlot;
            continue;
        }
        if (d < 0) {
            const int32_t v = -1 - d;
            d = kBlockRing * W + (v >> 20) * X + (v & ((1 << 20) - 1));
        }
        bdep16[e] = (uint16_t)d;
    }
    std::vector<int32_t> fill(bxidx);
    std::sort(fill.begin(), fill.end());
    fill.erase(std::unique(fill.begin(), fill.end()), fill.end());
    if (bxidx.empty()) bxidx.push_back(0);   // the loader's clamped loads read one entry
    sw.h_brow = order;
    sw.bL = nlevels;
    sw.nblk = nb;
    sw.bX = X;
    sw.bK = kmax;
    sw.nbx = (int64_t)bxidx.size();
    sw.nfill = (int64_t)fill.size();
    int s;
    if ((s = dev_alloc_n(c, 2 * nb, &sw.brange)) || (s = dev_alloc_n(c, nb * L1, &sw.blvl)) ||
        (s = dev_alloc_n(c, nb * L1, &sw.bxptr)) || (s = dev_alloc_n(c, sw.nbx, &sw.bxidx)) ||
        (s = dev_alloc_n(c, n, &sw.brow)) ||
        (s = dev_alloc_n(c, kmax * n, &sw.bdep)) || (s = dev_alloc_n(c, kmax * n, &sw.bsrc)) ||
        (s = dev_alloc_n(c, kmax * n, &sw.bval)) || (s = dev_alloc_n(c, n, &sw.spos)) ||
        (s = dev_alloc_n(c, n, &sw.xpos)) ||
        (s = dev_alloc_n(c, n, &sw.bpos)) ||
        (s = dev_alloc_n(c, sw.nfill, &sw.bfill)) || (s = upload(c, sw.bfill, fill.data(), sw.nfill)) ||
        (s = upload(c, sw.brange, brange.data(), 2 * nb)) ||
        (s = upload(c, sw.blvl, blvl.data(), nb * L1)) || (s = upload(c, sw.bxptr, bxptr.data(), nb * L1)) ||
        (s = upload(c, sw.bxidx, bxidx.data(), sw.nbx)) || (s = upload(c, sw.brow, order.data(), n)) ||
        (s = upload(c, sw.bdep, bdep16.data(), kmax * n)) ||
        (s = upload(c, sw.bsrc, bsrc.data(), kmax * n)))
        return s;
    return FVG_OK;
}

// The blocked layout of a sweep (see k_sweep_blocked): assign_blocked, then the layout.
// The assignment is kept (sw.h_assign) so that a mirrored layout can be made from it.
static int build_blocked(fvg_ctx* c, const std::vector<int32_t>& rp, const std::vector<int32_t>& col, int64_t n,
                         bool upper, const std::vector<int32_t>* src_map, const std::vector<int32_t>& lev_in,
                         int nlevels_in, int target, int target_consecutive, int gap, const CoordSource& coords,
                         const std::vector<int>& box_counts, Sweep& sw) {
    BlockAssign A;
    if (!assign_blocked(rp, col, n, upper, lev_in, nlevels_in, target, target_consecutive, gap, coords, box_counts, A))
        return FVG_OK;
    FVG_TRY(materialize_blocked(c, rp, col, n, upper, src_map, A, false, sw));
    if (sw.nblk > 0) sw.h_assign = std::move(A);
    return FVG_OK;
}

// The upper layout mirrored from the blocked lower layout `lo`: block nb-1-b, level
// L-1-l and rows descending where the lower has block b, level l and rows ascending, so
// upper position q holds the row of lower position n-1-q.  A valid upper schedule when the
// upper dependencies are the lower ones transposed (the caller checks): an upper
// dependency i -> j (j > i) is then the lower dependency j -> i, one or more levels deep
// (two or more between blocks) in the lower layout and so in its mirror.  Leaves
// up.nblk = 0 when the layout does not fit the kernel.
static int build_mirrored(fvg_ctx* c, const std::vector<int32_t>& rp, const std::vector<int32_t>& col, int64_t n,
                          const std::vector<int32_t>* src_map, const Sweep& lo, Sweep& up) {
    const BlockAssign& L = lo.h_assign;
    if (lo.nblk == 0 || (int64_t)L.blk.size() != n) return FVG_OK;
    int kmax = 0;
    for (int64_t i = 0; i < n; ++i) {
        int cnt = 0;
        for (int32_t k = rp[(size_t)i]; k < rp[(size_t)i + 1]; ++k) {
            const int32_t j = col[(size_t)k];
            if (j > i && j < n) ++cnt;
        }
        kmax = std::max(kmax, cnt);
    }
    if (kmax > kBlockMaxDeps) return FVG_OK;
    BlockAssign U;
    U.nb = L.nb;
    U.nlevels = L.nlevels;
    U.kmax = kmax <= 4 ? 4 : kBlockMaxDeps;
    U.blk.resize((size_t)n);
    U.lev.resize((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        U.blk[(size_t)i] = L.nb - 1 - L.blk[(size_t)i];
        U.lev[(size_t)i] = L.nlevels - 1 - L.lev[(size_t)i];
    }
    FVG_TRY(materialize_blocked(c, rp, col, n, true, src_map, U, true, up));
    if (up.nblk == 0) return FVG_OK;
    bool ok = up.h_brow.size() == (size_t)n && lo.h_brow.size() == (size_t)n;
    for (int64_t q = 0; ok && q < n; ++q) ok = up.h_brow[(size_t)q] == lo.h_brow[(size_t)(n - 1 - q)];
    if (!ok) return fail(FVG_E_INVALID, "mirrored sweep layout does not reverse the lower one");
    up.mirrored = true;
    return FVG_OK;
}

// upper=false: entries j < i; upper=true: entries i < j < n, each row's entries taken in
// their stored order.  src_map (optional) maps a pattern slot to the factor slot holding
// that entry's value.  An upper sweep with `mirror` (the factor's blocked lower layout,
// whose dependencies the caller has checked are these transposed) takes the mirrored
// layout where it can.
static int build_sweep(fvg_ctx* c, const std::vector<int32_t>& rp, const std::vector<int32_t>& col, int64_t n,
                       bool upper, const std::vector<int32_t>* src_map, Sweep& sw, const Sweep* mirror = nullptr,
                       const CoordSource& coords = CoordSource()) {
    std::vector<int32_t> lev((size_t)n, 0);
    int maxlev = 0;
    auto take = [&](int64_t i, int32_t j) { return upper ? (j > i && j < n) : (j < i); };
    if (!upper) {
        for (int64_t i = 0; i < n; ++i) {
            int l = 0;
            for (int32_t k = rp[(size_t)i]; k < rp[(size_t)i + 1]; ++k)
                if (take(i, col[(size_t)k])) l = std::max(l, lev[(size_t)col[(size_t)k]] + 1);
            lev[(size_t)i] = l;
            maxlev = std::max(maxlev, l);
        }
    } else {
        for (int64_t i = n - 1; i >= 0; --i) {
            int l = 0;
            for (int32_t k = rp[(size_t)i]; k < rp[(size_t)i + 1]; ++k)
                if (take(i, col[(size_t)k])) l = std::max(l, lev[(size_t)col[(size_t)k]] + 1);
            lev[(size_t)i] = l;
            maxlev = std::max(maxlev, l);
        }
    }
    sw.n = n;
    sw.nlevels = maxlev + 1;
    std::vector<int64_t> count((size_t)sw.nlevels, 0);
    for (int64_t i = 0; i < n; ++i) count[(size_t)lev[(size_t)i]]++;
    sw.lvl_ptr.assign((size_t)sw.nlevels + 1, 0);
    for (int l = 0; l < sw.nlevels; ++l)
        sw.lvl_ptr[(size_t)l + 1] = sw.lvl_ptr[(size_t)l] + (count[(size_t)l] + kChunk - 1) / kChunk * kChunk;
    sw.npos = sw.lvl_ptr[(size_t)sw.nlevels];
    std::vector<int32_t> rows((size_t)sw.npos, -1);
    {
        std::vector<int64_t> cur(sw.lvl_ptr.begin(), sw.lvl_ptr.end() - 1);
        for (int64_t i = 0; i < n; ++i) rows[(size_t)cur[(size_t)lev[(size_t)i]]++] = (int32_t)i;   // ascending rows
    }
    std::vector<int32_t> ptr((size_t)sw.npos + 1, 0), ecol, esrc;
    ecol.reserve(col.size() / 2 + 1);
    esrc.reserve(col.size() / 2 + 1);
    for (int64_t p = 0; p < sw.npos; ++p) {
        const int32_t i = rows[(size_t)p];
        if (i >= 0) {
            for (int32_t k = rp[(size_t)i]; k < rp[(size_t)i + 1]; ++k) {
                const int32_t j = col[(size_t)k];
                if (!take(i, j)) continue;
                ecol.push_back(j);
                esrc.push_back(src_map ? (*src_map)[(size_t)k] : k);
            }
        }
        ptr[(size_t)p + 1] = (int32_t)ecol.size();
    }
    sw.nnz = (int64_t)ecol.size();
    ecol.push_back(0);   // padding entry after the last row: column 0, coefficient 0
    const double zero = 0.0;
    int s;
    if ((s = dev_alloc_n(c, sw.nlevels + 1, &sw.d_lvl_ptr)) || (s = upload(c, sw.d_lvl_ptr, sw.lvl_ptr.data(),
                                                                           sw.nlevels + 1)))
        return s;
    if ((s = dev_alloc_n(c, sw.npos, &sw.rows)) || (s = dev_alloc_n(c, sw.npos + 1, &sw.ptr)) ||
        (s = dev_alloc_n(c, sw.nnz + 1, &sw.col)) || (s = dev_alloc_n(c, sw.nnz, &sw.src)) ||
        (s = dev_alloc_n(c, sw.nnz + 1, &sw.val)) || (s = upload(c, sw.rows, rows.data(), sw.npos)) ||
        (s = upload(c, sw.ptr, ptr.data(), sw.npos + 1)) || (s = upload(c, sw.col, ecol.data(), sw.nnz + 1)) ||
        (s = upload(c, sw.val + sw.nnz, &zero, 1)) || (s = upload(c, sw.src, esrc.data(), sw.nnz)))
        return s;
    // packed components: about one block per CU, or with a gap of three or more, grown
    // blocks of about n / (CUs / 4) rows (on 1,572,864 tetrahedra ~95 blocks whose widest
    // level reaches 128 rows or one pair's); consecutive rows: two per CU (measured on
    // 2,097,152 prisms: 597 us per ILU(0) apply with 512 blocks, 621 with 256, 608 with 1024)
    const int gap = c->tune_sweep_gap ? (int)c->tune_sweep_gap : kBlockGap;
    const int packed = c->tune_sweep_blocks ? (int)c->tune_sweep_blocks
                                            : std::max(2, gap > 2 ? c->num_cus / 4 : c->num_cus);
    const int consecutive = c->tune_sweep_blocks ? (int)c->tune_sweep_blocks : std::max(2, 2 * c->num_cus);
    if (upper && mirror && mirror->nblk > 0) {
        FVG_TRY(build_mirrored(c, rp, col, n, src_map, *mirror, sw));
        if (sw.nblk > 0) return FVG_OK;
    }
    // boxes of the rows' coordinates: a grid of about CUs/2 boxes (a power of two; on
    // 2,097,152 prisms 64 boxes gave 337 levels of up to 528 rows, 128 boxes 345 of 272,
    // 256 boxes 353 of 272)
    std::vector<int> box_counts;
    if (c->tune_sweep_boxes >= 2)
        box_counts.push_back((int)c->tune_sweep_boxes);
    else if (c->tune_sweep_boxes == 0)
        box_counts.push_back(std::max(2, c->num_cus / 2));
    return build_blocked(c, rp, col, n, upper, src_map, lev, sw.nlevels, packed, consecutive, gap, coords, box_counts,
                         sw);
}

// When both layouts of a factor are blocked, the upper set-up reads the lower sweep's
// values in the lower layout's position order: up.bfrom[q] = lower position of row
// up.brow[q].
static int pair_layouts(fvg_ctx* c, Sweep& lo, Sweep& up) {
    int s = FVG_OK;
    if (lo.nblk > 0 && up.nblk > 0 && lo.n == up.n) {
        std::vector<int32_t> pos_lo((size_t)lo.n), from((size_t)up.n);
        for (int64_t p = 0; p < lo.n; ++p) pos_lo[(size_t)lo.h_brow[(size_t)p]] = (int32_t)p;
        for (int64_t q = 0; q < up.n; ++q) from[(size_t)q] = pos_lo[(size_t)up.h_brow[(size_t)q]];
        if ((s = dev_alloc_n(c, up.n, &up.bfrom)) == FVG_OK) s = upload(c, up.bfrom, from.data(), up.n);
        if (s == FVG_OK && (s = dev_alloc_n(c, lo.n, &lo.bposof)) == FVG_OK)
            s = upload(c, lo.bposof, pos_lo.data(), lo.n);
        // a mirrored pair: which positions each sweep polls
        std::vector<uint8_t> pl, pu;
        if (s == FVG_OK && up.mirrored) {
            pl.assign((size_t)lo.n, 0);
            pu.assign((size_t)up.n, 0);
            for (Sweep* sw : {&lo, &up}) {
                std::vector<int32_t> fill((size_t)sw->nfill);
                if ((s = download(c, fill.data(), sw->bfill, sw->nfill)) != FVG_OK) break;
                std::vector<uint8_t>& f = sw == &lo ? pl : pu;
                for (int32_t q : fill) f[(size_t)q] = 1;
            }
            if (s == FVG_OK && (s = dev_alloc_n(c, lo.n, &lo.bpolled)) == FVG_OK &&
                (s = dev_alloc_n(c, up.n, &up.bpolled)) == FVG_OK && (s = upload(c, lo.bpolled, pl.data(), lo.n)) ==
                    FVG_OK)
                s = upload(c, up.bpolled, pu.data(), up.n);
        }
        if (s == FVG_OK) s = hipStreamSynchronize(c->stream) == hipSuccess ? FVG_OK : FVG_E_HIP;   // host arrays
    }
    for (Sweep* sw : {&lo, &up}) {
        sw->h_brow.clear();
        sw->h_brow.shrink_to_fit();
        sw->h_assign = BlockAssign{};
    }
    return s;
}

static void free_tri32(fvg_ctx* c, Tri32& t) {
    dev_free_n(c, t.ptr, t.n + 1);
    dev_free_n(c, t.col, t.nnz);
    if (t.dcol) dev_free_n(c, t.dcol, t.nnz);
    dev_free_n(c, t.src, t.nnz);
    dev_free_n(c, t.val, t.nnz);
    t = Tri32{};
}

// pa->val[k] = a->val[pa_src[k]]
__global__ void k_gather_values(int64_t nnz, const int32_t* __restrict__ src, const double* __restrict__ v,
                                double* __restrict__ out) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nnz) out[k] = v[src[k]];
}

static void free_pos_matrix(fvg_ctx* c, const fvg_prec* p) {
    if (!p->pa) return;
    dev_free_n(c, p->pa->row_ptr, p->n + 1);
    dev_free_n(c, p->pa->col, p->a_nnz);
    dev_free_n(c, p->pa->val, p->a_nnz);
    dev_free_n(c, p->pa_src, p->a_nnz);
    delete p->pa;
    p->pa = nullptr;
    p->pa_of = nullptr;
}

static void free_prec(fvg_ctx* c, fvg_prec* p) {
    team_plan_release(c, p);
    dev_free_n(c, p->d_rp, p->n + 1);
    dev_free_n(c, p->d_col, p->nnz);
    dev_free_n(c, p->d_diag, p->n);
    dev_free_n(c, p->d_tpos, p->nnz);
    dev_free_n(c, p->d_fac, p->nnz + 1);
    dev_free_n(c, p->d_amap, p->a_nnz);
    dev_free_n(c, p->d_afill, p->nnz);
    dev_free_n(c, p->d_inv_diag, p->n);
    dev_free_n(c, p->d_tmp, p->n);
    dev_free_n(c, p->d_flags, 4);
    dev_free_n(c, p->d_inv_diag32, p->n);
    dev_free_n(c, p->d_tmp32, p->n);
    dev_free_n(c, p->d_tmp32b, p->n);
    free_tri32(c, p->jlo);
    free_tri32(c, p->jup);
    free_sweep(c, p->lower);
    free_sweep(c, p->upper);
    free_pos_matrix(c, p);
}

// pattern, diagonal slots and device copies shared by every kind
static int setup_pattern(fvg_ctx* c, int64_t n, std::vector<int32_t> rp, std::vector<int32_t> col, fvg_prec* p,
                         bool need_transpose) {
    p->n = n;
    p->nnz = (int64_t)col.size();
    p->h_rp = std::move(rp);
    p->h_col = std::move(col);
    p->h_diag.assign((size_t)n, -1);
    for (int64_t i = 0; i < n; ++i) {
        auto b = p->h_col.begin() + p->h_rp[(size_t)i], e = p->h_col.begin() + p->h_rp[(size_t)i + 1];
        auto it = std::lower_bound(b, e, (int32_t)i);
        if (it == e || *it != (int32_t)i)
            return fail(FVG_E_MISSING_ENTRY, "Sparse matrix index (" + std::to_string(i) + ", " + std::to_string(i) +
                                                 ") not found");
        p->h_diag[(size_t)i] = (int32_t)(it - p->h_col.begin());
    }
    if (need_transpose) {
        p->h_tpos.assign((size_t)p->nnz, -1);
        for (int64_t i = 0; i < n; ++i) {
            for (int32_t k = p->h_rp[(size_t)i]; k < p->h_rp[(size_t)i + 1]; ++k) {
                const int32_t j = p->h_col[(size_t)k];
                if (j >= n) continue;
                auto b = p->h_col.begin() + p->h_rp[(size_t)j], e = p->h_col.begin() + p->h_rp[(size_t)j + 1];
                auto it = std::lower_bound(b, e, (int32_t)i);
                if (it == e || *it != (int32_t)i)
                    return fail(FVG_E_MISSING_ENTRY, "Sparse matrix index (" + std::to_string(j) + ", " +
                                                         std::to_string(i) + ") not found");
                p->h_tpos[(size_t)k] = (int32_t)(it - p->h_col.begin());
            }
        }
    }
    int s;
    if ((s = dev_alloc_n(c, n + 1, &p->d_rp)) || (s = dev_alloc_n(c, p->nnz, &p->d_col)) ||
        (s = dev_alloc_n(c, n, &p->d_diag)) || (s = dev_alloc_n(c, n, &p->d_inv_diag)) ||
        (s = dev_alloc_n(c, 4, &p->d_flags)) || (s = upload(c, p->d_rp, p->h_rp.data(), n + 1)) ||
        (s = upload(c, p->d_col, p->h_col.data(), p->nnz)) || (s = upload(c, p->d_diag, p->h_diag.data(), n)))
        return s;
    if (need_transpose) {
        if ((s = dev_alloc_n(c, p->nnz, &p->d_tpos)) || (s = upload(c, p->d_tpos, p->h_tpos.data(), p->nnz))) return s;
    }
    return FVG_OK;
}

static int read_flags(fvg_ctx* c, const fvg_prec* p, int32_t* host4) {
    FVG_HIP(hipMemcpyAsync(c->pinned, p->d_flags, 4 * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    FVG_HIP(hipStreamSynchronize(c->stream));
    std::memcpy(host4, c->pinned, 4 * sizeof(int32_t));
    return FVG_OK;
}

// numeric factorisation for the current values of a (pattern already analysed).
// check: read the symmetry / pivot flags and fail on them (synchronises); otherwise the
// whole factorisation is only enqueued and the flags are read by fvg_prec_status.
static const double* factor_input(fvg_ctx* c, const fvg_prec* p, const fvg_csr* a);

// entries of the factor pattern's longest row (counted on first use)
static int32_t longest_row(fvg_prec* p) {
    if (p->max_row < 0) {
        int32_t w = 0;
        for (int64_t i = 0; i < p->n; ++i) w = std::max(w, p->h_rp[(size_t)i + 1] - p->h_rp[(size_t)i]);
        p->max_row = w;
    }
    return p->max_row;
}

static int factorize(fvg_ctx* c, fvg_prec* p, const fvg_csr* a, bool check = true) {
    FVG_HIP(hipMemsetAsync(p->d_flags, 0, 4 * sizeof(int32_t), c->stream));
    const Sweep& lo = p->lower;
    const double* av = factor_input(c, p, a);   // values on the factor pattern
    if (p->kind == kIC) {
        hipLaunchKernelGGL(k_sym_check, dim3(grid_for(p->nnz)), dim3(kBlock), 0, c->stream, p->nnz, p->d_tpos, av,
                           p->d_flags);
        if (check) {
            int32_t fl[4];
            FVG_TRY(read_flags(c, p, fl));
            if (fl[2]) return fail(FVG_E_NOT_SYMMETRIC, "IncompleteCholesky requires a symmetric matrix");
        }
        const bool wave = longest_row(p) <= 64;   // a wave per row holds the row in its lanes
        for (int l = 0; l < lo.nlevels; ++l) {
            const int64_t b = lo.lvl_ptr[(size_t)l], e = lo.lvl_ptr[(size_t)l + 1];
            if (wave)
                hipLaunchKernelGGL(k_ic0_level_wave, dim3(grid_for(64 * (e - b))), dim3(kBlock), 0, c->stream, b, e,
                                   lo.rows, p->d_rp, p->d_col, p->d_diag, av, p->d_fac, p->d_inv_diag, p->d_flags);
            else
                hipLaunchKernelGGL(k_ic0_level, dim3(grid_for(e - b)), dim3(kBlock), 0, c->stream, b, e, lo.rows,
                                   p->d_rp, p->d_col, p->d_diag, av, p->d_fac, p->d_inv_diag, p->d_flags);
        }
    } else {
        const bool wave = longest_row(p) <= 64;
        for (int l = 0; l < lo.nlevels; ++l) {
            const int64_t b = lo.lvl_ptr[(size_t)l], e = lo.lvl_ptr[(size_t)l + 1];
            if (wave)
                hipLaunchKernelGGL(k_ilu0_level_wave, dim3(grid_for(64 * (e - b))), dim3(kBlock), 0, c->stream, b,
                                   e, lo.rows, p->d_rp, p->d_col, p->d_diag, av, p->d_fac, p->d_flags);
            else
                hipLaunchKernelGGL(k_ilu0_level, dim3(grid_for(e - b)), dim3(kBlock), 0, c->stream, b, e, lo.rows,
                                   p->d_rp, p->d_col, p->d_diag, av, p->d_fac, p->d_flags);
        }
        hipLaunchKernelGGL(k_inv_diag, dim3(grid_for(p->n)), dim3(kBlock), 0, c->stream, p->n, p->d_diag, p->d_fac,
                           p->d_inv_diag);
    }
    FVG_HIP(hipGetLastError());
    if (p->kind == kILU && check) {
        int32_t fl[4];
        FVG_TRY(read_flags(c, p, fl));
        if (fl[0]) return fail(FVG_E_ZERO_PIVOT, "error, lu[k, k] too small");
    }
    hipLaunchKernelGGL(k_gather, dim3(grid_for(p->lower.nnz)), dim3(kBlock), 0, c->stream, p->lower.nnz, p->lower.src,
                       p->d_fac, p->lower.val);
    hipLaunchKernelGGL(k_gather, dim3(grid_for(p->upper.nnz)), dim3(kBlock), 0, c->stream, p->upper.nnz, p->upper.src,
                       p->d_fac, p->upper.val);
    for (const Sweep* sw : {&p->lower, &p->upper}) {
        const int64_t ne = (int64_t)sw->bK * sw->n;
        if (sw->nblk > 0 && ne > 0) {
            hipLaunchKernelGGL(k_gather_pad, dim3(grid_for(ne)), dim3(kBlock), 0, c->stream, ne, sw->bsrc, p->d_fac,
                               sw->bval);
            hipLaunchKernelGGL(k_gather_rows, dim3(grid_for(sw->n)), dim3(kBlock), 0, c->stream, sw->n, sw->brow,
                               p->d_inv_diag, sw->spos);
        }
    }
    FVG_HIP(hipGetLastError());
    if (check) FVG_HIP(hipStreamSynchronize(c->stream));
    p->fac_version += 1;
    return FVG_OK;
}

// largest grid of the persistent sweep that is guaranteed co-resident
template <class T>
static int max_coresident_groups(fvg_ctx* c) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_sweep_persistent<T, 0, true>, kChunk * kWavesPerGroup,
                                                     0) !=
        hipSuccess)
        return 1;
    const int m = per_cu * c->num_cus;
    return m > 0 ? m : 1;
}

// waves of the persistent sweep: about two and a half levels' worth of chunks in flight
// (fewer expose each chunk's start, more add polling traffic that slows every hop)
static int64_t sweep_waves(const fvg_ctx* c, const Sweep& sw, int max_groups) {
    int64_t w = c->tune_sweep_waves;
    if (w <= 0) {
        const int64_t chunks = sw.npos / kChunk;
        w = 5 * ((chunks + sw.nlevels - 1) / std::max(1, sw.nlevels)) / 2;
    }
    w = std::max<int64_t>(w, kWavesPerGroup);
    w = (w + kWavesPerGroup - 1) / kWavesPerGroup * kWavesPerGroup;
    return std::min<int64_t>(w, (int64_t)max_groups * kWavesPerGroup);
}

template <class T>
static const void* sweep_kernel(int64_t poll, bool scaled) {
    if (scaled)
        return poll == 1   ? (const void*)k_sweep_persistent<T, 1, true>
               : poll == 2 ? (const void*)k_sweep_persistent<T, 2, true>
                           : (const void*)k_sweep_persistent<T, 0, true>;
    return poll == 1   ? (const void*)k_sweep_persistent<T, 1, false>
           : poll == 2 ? (const void*)k_sweep_persistent<T, 2, false>
                       : (const void*)k_sweep_persistent<T, 0, false>;
}

// the persistent sweep proper; its output must already hold the pending pattern
template <class T>
static int launch_sweep(fvg_ctx* c, const Sweep& sw, const T* val, const T* scale, const T* b, T* x, int32_t* err,
                        const int32_t* skip) {
    if (sw.n == 0) return FVG_OK;
    int64_t nchunks = sw.npos / kChunk;
    int64_t nwaves = sweep_waves(c, sw, max_coresident_groups<T>(c));
    nwaves = std::min(nwaves, (nchunks + kWavesPerGroup - 1) / kWavesPerGroup * kWavesPerGroup);
    const int64_t groups = nwaves / kWavesPerGroup;
    int spin = spin_limit(c), withhold = c->tune_withhold ? 1 : 0;
    void* args[] = {&nchunks, &nwaves, (void*)&sw.rows, (void*)&sw.ptr, (void*)&sw.col, (void*)&val, (void*)&scale,
                    (void*)&b, (void*)&x, (void*)&err, (void*)&skip, (void*)&spin, (void*)&withhold};
    const void* fn = sweep_kernel<T>(c->tune_sweep_poll, scale != nullptr);
    FVG_HIP(launch_coresident(c, fn, (unsigned)groups, (unsigned)(kChunk * kWavesPerGroup), args, 0));
    return FVG_OK;
}

template <bool SCALE, int FUSE>
static const void* blocked_kernel_f(int K, bool spin) {
    if constexpr (FUSE != kFuseNone) {   // fused CG work with 4 slots per row only (8 would spill)
        if (K > 4) return nullptr;
        return spin ? (const void*)k_sweep_blocked<double, SCALE, 4, true, FUSE>
                    : (const void*)k_sweep_blocked<double, SCALE, 4, false, FUSE>;
    } else {
        if (spin)
            return K <= 4 ? (const void*)k_sweep_blocked<double, SCALE, 4, true, FUSE>
                          : (const void*)k_sweep_blocked<double, SCALE, 8, true, FUSE>;
        return K <= 4 ? (const void*)k_sweep_blocked<double, SCALE, 4, false, FUSE>
                      : (const void*)k_sweep_blocked<double, SCALE, 8, false, FUSE>;
    }
}

template <bool SCALE>
static const void* blocked_kernel(int K, bool spin, int fuse = kFuseNone) {
    if (fuse == kFuseXR) return blocked_kernel_f<SCALE, kFuseXR>(K, spin);
    if (fuse == kFuseRZ) return blocked_kernel_f<SCALE, kFuseRZ>(K, spin);
    return blocked_kernel_f<SCALE, kFuseNone>(K, spin);
}

static size_t blocked_lds(const Sweep& sw) {
    return (size_t)(kBlockRing * sw.bW + 2 * sw.bX) * sizeof(double) + (size_t)(sw.bL + 1) * sizeof(int32_t);
}

// whether every block of the blocked layout can be resident at once (decided on first use)
// the kernels of one family (fuse: kFuseNone, or both fused CG variants), scaled and not,
// with and without the test's spin limit, all co-resident for this layout
static bool blocked_family_fits(fvg_ctx* c, const Sweep& sw, bool fused) {
    const int threads = ((sw.bW + 63) / 64 + 1) * 64;
    const size_t lds = blocked_lds(sw);
    bool ok = threads <= 1024 && lds <= 160 * 1024;
    const int nf = fused ? 2 : 1;
    for (int v = 0; ok && v < 4 * nf; ++v) {
        const bool scaled = v & 1, spin = (v >> 1) & 1;
        const int fuse = fused ? (v >> 2 ? kFuseRZ : kFuseXR) : kFuseNone;
        const void* fn =
            scaled ? blocked_kernel<true>(sw.bK, spin, fuse) : blocked_kernel<false>(sw.bK, spin, fuse);
        if (lds > 64 * 1024 &&
            hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
            ok = false;
        int per_cu = 0;
        if (ok && (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, threads, lds) != hipSuccess ||
                   (int64_t)per_cu * c->num_cus < sw.nblk))
            ok = false;
    }
    return ok;
}

static bool blocked_fits(fvg_ctx* c, const Sweep& sw) {
    if (sw.nblk == 0) return false;
    if (sw.bOK < 0) sw.bOK = blocked_family_fits(c, sw, false) ? 1 : 0;
    return sw.bOK == 1;
}

static bool blocked_fused_fits(fvg_ctx* c, const Sweep& sw) {
    if (!blocked_fits(c, sw) || sw.bK > 4) return false;
    if (sw.bOKfuse < 0) sw.bOKfuse = blocked_family_fits(c, sw, true) ? 1 : 0;
    return sw.bOKfuse == 1;
}

// One blocked triangular solve.  The right-hand side is b[brow[p]] (natural order), or,
// with `from`, b[from[p]] (b in another layout's position order).  `scatter` = false
// leaves the values in sw.xpos only (a paired upper sweep reads them there).
static int launch_blocked(fvg_ctx* c, const Sweep& sw, const double* scale, const double* b, double* x, int32_t* err,
                          const int32_t* skip, const int32_t* from = nullptr, bool scatter = true,
                          bool prepare = true, const double* bsrc = nullptr, int brev = 0, int fuse = kFuseNone,
                          const SweepFuse* fz = nullptr) {
    // the timed span holds one whole triangular solve: set-up, sweep and scatter kernels
    // (without `prepare`, the caller has written the pending fill, and the right-hand side
    // in position order is sw.bpos, or bsrc, read in reverse with brev)
    const int tok = timing_begin(c);
    if (prepare)
        hipLaunchKernelGGL(k_sweep_prepare<double>, dim3(grid_for(sw.n + sw.nfill)), dim3(kBlock), 0, c->stream, sw.n,
                           from ? from : sw.brow, b, sw.bpos, sw.nfill, sw.bfill, sw.xpos, skip);
    int nlev = sw.bL, W = sw.bW, X = sw.bX, spin = spin_limit(c), withhold = c->tune_withhold ? 1 : 0;
    int64_t n = sw.n;
    const double* bp = bsrc ? bsrc : sw.bpos;
    const bool custom = c->tune_spin_limit > 0 || withhold;
    const void* fn = scale ? blocked_kernel<true>(sw.bK, custom, fuse) : blocked_kernel<false>(sw.bK, custom, fuse);
    SweepFuse F = fz ? *fz : SweepFuse{};
    void* args[] = {&nlev, &n, &W, &X, (void*)&sw.brange, (void*)&sw.blvl, (void*)&sw.bxptr, (void*)&sw.bxidx,
                    (void*)&sw.bdep, (void*)&sw.bval, (void*)&bp,
                    (void*)&sw.spos, (void*)&sw.xpos, (void*)&err, (void*)&skip, (void*)&spin, (void*)&withhold,
                    (void*)&brev, (void*)&F};
    const int threads = ((sw.bW + 63) / 64 + 1) * 64;
    FVG_HIP(launch_coresident(c, fn, (unsigned)sw.nblk, (unsigned)threads, args, blocked_lds(sw)));
    if (scatter)
        hipLaunchKernelGGL(k_sweep_finish<double>, dim3(grid_for(sw.n)), dim3(kBlock), 0, c->stream, sw.n, sw.brow,
                           (const double*)sw.xpos, x, skip);
    timing_end(c, tok, kSpanSweep, (prepare ? 2 : 1) + (scatter ? 1 : 0));
    FVG_HIP(hipGetLastError());
    return FVG_OK;
}

static int sweep(fvg_ctx* c, const Sweep& sw, const double* scale, const double* b, double* x, int32_t* err,
                 const int32_t* skip) {
    if (sw.n == 0) return FVG_OK;
    if (c->tune_sweep_mode == 0 && blocked_fits(c, sw))
        return launch_blocked(c, sw, scale, b, x, err, skip);
    if (c->tune_sweep_mode != 1) {   // the persistent sweep (mode 2, or no blocked layout)
        // the timed span holds one whole triangular solve: the pending fill and the sweep
        const int tok = timing_begin(c);
        hipLaunchKernelGGL(k_fill_pending<double>, dim3(grid_for(sw.n)), dim3(kBlock), 0, c->stream, sw.n, x, skip);
        FVG_TRY(launch_sweep<double>(c, sw, sw.val, scale, b, x, err, skip));
        timing_end(c, tok, kSpanSweep, 2);
        FVG_HIP(hipGetLastError());
        return FVG_OK;
    }
    const int tok = timing_begin(c);
    for (int l = 0; l < sw.nlevels; ++l) {
        const int64_t lb = sw.lvl_ptr[(size_t)l], le = sw.lvl_ptr[(size_t)l + 1];
        hipLaunchKernelGGL(k_sweep_level, dim3(grid_for(le - lb)), dim3(kBlock), 0, c->stream, lb, le, sw.rows,
                           sw.ptr, sw.col, sw.val, scale, b, x, skip);
    }
    timing_end(c, tok, kSpanSweep, sw.nlevels);
    FVG_HIP(hipGetLastError());
    return FVG_OK;
}

static int sweep32(fvg_ctx* c, const Sweep& sw, const float* scale, const float* b, float* x, int32_t* err,
                   const int32_t* skip) {
    if (sw.n == 0) return FVG_OK;
    const int tok = timing_begin(c);   // one whole triangular solve: the pending fill and the sweep
    hipLaunchKernelGGL(k_fill_pending<float>, dim3(grid_for(sw.n)), dim3(kBlock), 0, c->stream, sw.n, x, skip);
    FVG_TRY(launch_sweep<float>(c, sw, sw.val32, scale, b, x, err, skip));
    timing_end(c, tok, kSpanSweep, 2);
    FVG_HIP(hipGetLastError());
    return FVG_OK;
}

// one strict triangular factor of p in natural row order (the inexact FP32 apply):
// lower = slots before the diagonal; upper = slots after it with columns < n (IC: the
// value of slot (i, j) is l_ji, read through the transpose slot map)
static int build_tri32(fvg_ctx* c, const fvg_prec* p, bool upper, Tri32& t) {
    const int64_t n = p->n;
    std::vector<int32_t> ptr((size_t)n + 1, 0), col, src;
    col.reserve((size_t)p->nnz / 2 + 1);
    src.reserve((size_t)p->nnz / 2 + 1);
    for (int64_t i = 0; i < n; ++i) {
        const int32_t d = p->h_diag[(size_t)i];
        const int32_t kb = upper ? d + 1 : p->h_rp[(size_t)i], ke = upper ? p->h_rp[(size_t)i + 1] : d;
        for (int32_t k = kb; k < ke; ++k) {
            const int32_t j = p->h_col[(size_t)k];
            if (j >= n) break;
            col.push_back(j);
            src.push_back(upper && p->kind == kIC ? p->h_tpos[(size_t)k] : k);
        }
        ptr[(size_t)i + 1] = (int32_t)col.size();
    }
    t.n = n;
    t.nnz = (int64_t)col.size();
    // 16-bit column offsets from each tile's first row, where they all fit: the sweeps then
    // stream 2 bytes per entry for the columns instead of 4
    std::vector<int16_t> dcol((size_t)t.nnz);
    bool fits = t.nnz > 0;
    for (int64_t i = 0; i < n && fits; ++i) {
        const int64_t r0 = i / kJacRows * kJacRows;
        for (int32_t k = ptr[(size_t)i]; k < ptr[(size_t)i + 1]; ++k) {
            const int64_t d = (int64_t)col[(size_t)k] - r0;
            if (d < INT16_MIN || d > INT16_MAX) {
                fits = false;
                break;
            }
            dcol[(size_t)k] = (int16_t)d;
        }
    }
    int s;
    if ((s = dev_alloc_n(c, n + 1, &t.ptr)) || (s = dev_alloc_n(c, t.nnz, &t.col)) ||
        (s = dev_alloc_n(c, t.nnz, &t.src)) || (s = dev_alloc_n(c, t.nnz, &t.val)) ||
        (s = upload(c, t.ptr, ptr.data(), n + 1)) || (s = upload(c, t.col, col.data(), t.nnz)) ||
        (s = upload(c, t.src, src.data(), t.nnz)))
        return s;
    if (fits && ((s = dev_alloc_n(c, t.nnz, &t.dcol)) || (s = upload(c, t.dcol, dcol.data(), t.nnz)))) return s;
    FVG_HIP(hipStreamSynchronize(c->stream));   // the host vectors go out of scope
    return FVG_OK;
}


int ensure_prec32(fvg_ctx* c, fvg_prec* p) {
    if (p->kind == kIdentity) return FVG_OK;
    if (p->fac32_version == p->fac_version && p->d_inv_diag32) return FVG_OK;
    if (!p->d_inv_diag32) FVG_TRY(dev_alloc_n(c, p->n, &p->d_inv_diag32));
    if (!p->d_tmp32) FVG_TRY(dev_alloc_n(c, p->n, &p->d_tmp32));
    if (!p->d_tmp32b) FVG_TRY(dev_alloc_n(c, p->n, &p->d_tmp32b));
    hipLaunchKernelGGL(k_to_f32, dim3(grid_for(p->n)), dim3(kBlock), 0, c->stream, p->n, p->d_inv_diag,
                       p->d_inv_diag32);
    if (p->kind == kIC || p->kind == kILU) {   // FP32 factors as natural-order CSRs
        for (int u = 0; u < 2; ++u) {
            Tri32& t = u ? p->jup : p->jlo;
            if (!t.ptr) FVG_TRY(build_tri32(c, p, u == 1, t));
            if (t.nnz > 0)
                hipLaunchKernelGGL(k_gather_f32, dim3(grid_for(t.nnz)), dim3(kBlock), 0, c->stream, t.nnz, t.src,
                                   p->d_fac, t.val);
        }
    }
    for (Sweep* sw : {&p->lower, &p->upper}) {
        if (sw->nnz == 0 && sw->n == 0) continue;
        if (!sw->val32) FVG_TRY(dev_alloc_n(c, sw->nnz + 1, &sw->val32));
        hipLaunchKernelGGL(k_to_f32, dim3(grid_for(sw->nnz + 1)), dim3(kBlock), 0, c->stream, sw->nnz + 1, sw->val,
                           sw->val32);
    }
    FVG_HIP(hipGetLastError());
    p->fac32_version = p->fac_version;
    return FVG_OK;
}

// m Jacobi sweeps of one triangular factor (natural row order); sweep k writes `odd` for
// odd k and `even` for even k, so the result lands in `odd` when m is odd and in `even`
// otherwise.  With dotw, the last sweep also leaves the partials of dotw . result, one per
// tile of kJacRows rows at the tile's index (the layout of a one-row-per-thread dot kernel
// with workgroups of kJacRows threads).
template <bool FIRST, bool SCALE>
static void launch_jacobi32(fvg_ctx* c, dim3 grid, dim3 block, int n, const Tri32& t, const float* scale,
                            const float* v, const float* yin, float* out, const int32_t* skip, const float* dw,
                            double* dp) {
    if (t.dcol)
        hipLaunchKernelGGL((k_jacobi32<FIRST, SCALE, true>), grid, block, 0, c->stream, n, t.ptr, t.col, t.dcol, t.val,
                           scale, v, yin, out, skip, dw, dp);
    else
        hipLaunchKernelGGL((k_jacobi32<FIRST, SCALE, false>), grid, block, 0, c->stream, n, t.ptr, t.col, t.dcol, t.val,
                           scale, v, yin, out, skip, dw, dp);
}

static void jacobi32(fvg_ctx* c, const Tri32& t, int m, const float* scale, const float* v, float* odd, float* even,
                     const int32_t* skip, const float* dotw = nullptr, double* dpart = nullptr) {
    if (t.n == 0) return;
    const dim3 grid((unsigned)((t.n + kJacRows - 1) / kJacRows)), block(kJacRows);
    const int n = (int)t.n;
    const float* prev = nullptr;
    for (int k = 1; k <= m; ++k) {
        float* out = (k & 1) ? odd : even;
        const float* dw = k == m ? dotw : nullptr;
        if (k == 1) {
            if (scale)
                launch_jacobi32<true, true>(c, grid, block, n, t, scale, v, nullptr, out, skip, dw, dpart);
            else
                launch_jacobi32<true, false>(c, grid, block, n, t, nullptr, v, nullptr, out, skip, dw, dpart);
        } else {
            if (scale)
                launch_jacobi32<false, true>(c, grid, block, n, t, scale, v, prev, out, skip, dw, dpart);
            else
                launch_jacobi32<false, false>(c, grid, block, n, t, nullptr, v, prev, out, skip, dw, dpart);
        }
        prev = out;
    }
}

// the FP32 inner solver's factor application: exact level-scheduled sweeps, or m Jacobi
// sweeps per triangular factor (the default)
static int apply_factor32(fvg_ctx* c, const fvg_prec* p, const float* lo_scale, const float* r, float* z,
                          const int32_t* skip, const DotFuse& dot) {
    if (c->tune_inner_exact) {
        FVG_TRY(sweep32(c, p->lower, lo_scale, r, p->d_tmp32, p->d_flags + 3, skip));
        return sweep32(c, p->upper, p->d_inv_diag32, p->d_tmp32, z, p->d_flags + 3, skip);
    }
    const int m = c->tune_inner_sweeps > 0 ? (int)c->tune_inner_sweeps : kInnerSweeps;
    // lower: result in tmp32 (m odd) or tmp32b (m even); upper: its last sweep writes z
    jacobi32(c, p->jlo, m, lo_scale, r, p->d_tmp32, p->d_tmp32b, skip);
    const float* y = (m & 1) ? p->d_tmp32 : p->d_tmp32b;
    float* other = (m & 1) ? p->d_tmp32b : p->d_tmp32;
    // the dot with z fused into the upper factor's last sweep where its partials land as the
    // caller's (one per kJacRows rows)
    const bool fuse = dot.w && dot.done && p->jup.n > 0 && (p->jup.n + kJacRows - 1) / kJacRows == dot.count;
    jacobi32(c, p->jup, m, p->d_inv_diag32, y, (m & 1) ? z : other, (m & 1) ? other : z, skip,
             fuse ? dot.w : nullptr, fuse ? dot.part : nullptr);
    if (fuse) *dot.done = true;
    FVG_HIP(hipGetLastError());
    return FVG_OK;
}

int apply_prec32(fvg_ctx* c, const fvg_prec* p, const float* r, float* z, const int32_t* skip, const DotFuse& dot) {
    if (dot.done) *dot.done = false;
    switch (p->kind) {
    case kIdentity:
        hipLaunchKernelGGL(k_scale32, dim3(grid_for(p->n)), dim3(kBlock), 0, c->stream, p->n, nullptr, r, z, skip);
        break;
    case kDiagonal:
        hipLaunchKernelGGL(k_scale32, dim3(grid_for(p->n)), dim3(kBlock), 0, c->stream, p->n, p->d_inv_diag32, r, z,
                           skip);
        break;
    case kIC:
        return apply_factor32(c, p, p->d_inv_diag32, r, z, skip, dot);
    case kILU:
        return apply_factor32(c, p, nullptr, r, z, skip, dot);
    default:
        return fail(FVG_E_INVALID, "unknown preconditioner");
    }
    FVG_HIP(hipGetLastError());
    return FVG_OK;
}

int check_sweep_status(fvg_ctx* c, const fvg_prec* p) {
    if (!p->d_flags) return FVG_OK;
    int32_t fl[4];
    FVG_TRY(read_flags(c, p, fl));
    if (fl[3]) {
        FVG_HIP(hipMemsetAsync(p->d_flags + 3, 0, sizeof(int32_t), c->stream));
        FVG_HIP(hipStreamSynchronize(c->stream));
        // a time-out that no tuning override forced: later co-resident launches of this
        // context are cooperative (some workgroup may have been kept from residency)
        if (c->tune_spin_limit == 0 && c->tune_withhold == 0) c->sweep_timed_out = true;
        return fail(FVG_E_TIMEOUT, "a triangular-sweep wait exceeded its spin limit");
    }
    return FVG_OK;
}

// z = U^-1 L^-1 r: the lower sweep into d_tmp, then the upper sweep.  When both layouts
// are blocked, the upper set-up reads the lower sweep's values in its position order
// and d_tmp is not written.
static int apply_factor(fvg_ctx* c, const fvg_prec* p, const double* lo_scale, const double* r, double* z,
                        const int32_t* skip) {
    const Sweep& lo = p->lower;
    const Sweep& up = p->upper;
    if (c->tune_sweep_mode == 0 && up.bfrom && blocked_fits(c, lo) && blocked_fits(c, up)) {
        FVG_TRY(launch_blocked(c, lo, lo_scale, r, p->d_tmp, p->d_flags + 3, skip, nullptr, false));
        return launch_blocked(c, up, p->d_inv_diag, lo.xpos, z, p->d_flags + 3, skip, up.bfrom, true);
    }
    FVG_TRY(sweep(c, p->lower, lo_scale, r, p->d_tmp, p->d_flags + 3, skip));
    return sweep(c, p->upper, p->d_inv_diag, p->d_tmp, z, p->d_flags + 3, skip);
}

// The matrix with rows and columns in the lower layout's position order, each row's entries
// in the original (ascending original column) order, so a product row sums the same terms
// in the same order as the original matrix's row.
static int build_pos_matrix(fvg_ctx* c, const fvg_prec* p) {
    const int64_t n = p->n, nnz = p->a_nnz;
    const auto& arp = p->level ? p->h_arp : p->h_rp;
    const auto& acol = p->level ? p->h_acol : p->h_col;
    std::vector<int32_t> row((size_t)n), pos((size_t)n);
    FVG_TRY(download(c, row.data(), p->lower.brow, n));
    for (int64_t q = 0; q < n; ++q) pos[(size_t)row[(size_t)q]] = (int32_t)q;
    std::vector<int32_t> prp((size_t)n + 1, 0), pcol((size_t)nnz), psrc((size_t)nnz);
    for (int64_t q = 0; q < n; ++q) {
        const int32_t i = row[(size_t)q];
        int32_t o = prp[(size_t)q];
        for (int32_t k = arp[(size_t)i]; k < arp[(size_t)i + 1]; ++k, ++o) {
            pcol[(size_t)o] = pos[(size_t)acol[(size_t)k]];
            psrc[(size_t)o] = k;
        }
        prp[(size_t)q + 1] = o;
    }
    auto* pa = new fvg_csr();
    pa->nrows = pa->ncols = n;
    pa->nnz = nnz;
    p->pa = pa;
    int s;
    if ((s = dev_alloc_n(c, n + 1, &pa->row_ptr)) || (s = dev_alloc_n(c, nnz, &pa->col)) ||
        (s = dev_alloc_n(c, nnz, &pa->val)) || (s = dev_alloc_n(c, nnz, &p->pa_src)) ||
        (s = upload(c, pa->row_ptr, prp.data(), n + 1)) || (s = upload(c, pa->col, pcol.data(), nnz)) ||
        (s = upload(c, p->pa_src, psrc.data(), nnz))) {
        (void)hipStreamSynchronize(c->stream);
        free_pos_matrix(c, p);
        return s;
    }
    FVG_HIP(hipStreamSynchronize(c->stream));   // the host arrays
    return FVG_OK;
}

int pos_solve(fvg_ctx* c, const fvg_prec* p, const fvg_csr* a, bool allow_alloc, PosSolve* out, bool* ok) {
    *ok = false;
    if (c->tune_solve_order != 0 || (p->kind != kIC && p->kind != kILU)) return FVG_OK;
    const Sweep& lo = p->lower;
    const Sweep& up = p->upper;
    if (!up.mirrored || !lo.bpolled || !up.bpolled || c->tune_sweep_mode != 0 || !blocked_fits(c, lo) ||
        !blocked_fits(c, up))
        return FVG_OK;
    if (a->nrows != p->n || a->ncols != p->n || a->nnz != p->a_nnz) return FVG_OK;
    if (p->pa_of != a) {   // a matrix met for the first time: the pattern must be p's
        if (a->h_row_ptr != (p->level ? p->h_arp : p->h_rp) || a->h_col != (p->level ? p->h_acol : p->h_col))
            return FVG_OK;
        if (!p->pa) {
            if (!allow_alloc) return FVG_OK;
            FVG_TRY(build_pos_matrix(c, p));
        }
        p->pa_of = a;
        p->pa_ver = ~0ull;
    }
    if (p->pa_ver != a->version) {
        hipLaunchKernelGGL(k_gather_values, dim3(grid_for(p->a_nnz)), dim3(kBlock), 0, c->stream, p->a_nnz,
                           p->pa_src, (const double*)a->val, p->pa->val);
        FVG_HIP(hipGetLastError());
        p->pa_ver = a->version;
    }
    out->n = p->n;
    out->row = lo.brow;
    out->lo_x = lo.xpos;
    out->up_x = up.xpos;
    out->lo_poll = lo.bpolled;
    out->up_poll = up.bpolled;
    out->pa = p->pa;
    *ok = true;
    return FVG_OK;
}

int apply_pos(fvg_ctx* c, const fvg_prec* p, const double* r_pos, const int32_t* skip) {
    const Sweep& lo = p->lower;
    const Sweep& up = p->upper;
    const double* lo_scale = p->kind == kIC ? p->d_inv_diag : nullptr;
    FVG_TRY(launch_blocked(c, lo, lo_scale, nullptr, nullptr, p->d_flags + 3, skip, nullptr, false, false, r_pos, 0));
    return launch_blocked(c, up, p->d_inv_diag, nullptr, nullptr, p->d_flags + 3, skip, nullptr, false, false, lo.xpos,
                          1);
}

int apply_pos_fused(fvg_ctx* c, const fvg_prec* p, double* r_pos, const int32_t* skip, const SweepFuse& lo_f,
                    const SweepFuse& up_f) {
    const Sweep& lo = p->lower;
    const Sweep& up = p->upper;
    const double* lo_scale = p->kind == kIC ? p->d_inv_diag : nullptr;
    FVG_TRY(launch_blocked(c, lo, lo_scale, nullptr, nullptr, p->d_flags + 3, skip, nullptr, false, false, r_pos, 0,
                           kFuseXR, &lo_f));
    return launch_blocked(c, up, p->d_inv_diag, nullptr, nullptr, p->d_flags + 3, skip, nullptr, false, false, lo.xpos,
                          1, kFuseRZ, &up_f);
}

bool pos_fusable(fvg_ctx* c, const fvg_prec* p) {
    return blocked_fused_fits(c, p->lower) && blocked_fused_fits(c, p->upper);
}

int apply_prec(fvg_ctx* c, const fvg_prec* p, const double* r, double* z, const int32_t* skip) {
    switch (p->kind) {
    case kIdentity:
        hipLaunchKernelGGL(k_scale, dim3(grid_for(p->n)), dim3(kBlock), 0, c->stream, p->n, nullptr, r, z, skip);
        FVG_HIP(hipGetLastError());
        return FVG_OK;
    case kDiagonal:
        hipLaunchKernelGGL(k_scale, dim3(grid_for(p->n)), dim3(kBlock), 0, c->stream, p->n, p->d_inv_diag, r, z, skip);
        FVG_HIP(hipGetLastError());
        return FVG_OK;
    case kIC:
        return apply_factor(c, p, p->d_inv_diag, r, z, skip);
    case kILU:
        return apply_factor(c, p, nullptr, r, z, skip);
    }
    return fail(FVG_E_INVALID, "unknown preconditioner");
}

// ---- incomplete factorisations with fill (level > 0) --------------------------------------
// Row i's fill follows the reference's pattern walks (incomplete_cholesky.rs:35-85,
// incomplete_lower_upper.rs:30-137): a depth-first walk from i over the matrix graph, a
// stack of vertices, each vertex's columns taken in their stored order, every vertex met
// at most once per row (marked on first meeting, even when not explored), and explored
// only while the walk's length to it is below the level.
struct Walk {
    std::vector<int32_t> mark, length, stack;   // mark[j] == i: met in the walk of row i
    explicit Walk(int64_t n) : mark((size_t)n, -1), length((size_t)n, 0) {}
    // kind 0 (IC): every vertex met from a vertex whose length is below the level joins;
    // 1 (ILU upper): vertices above i join, those below are explored while within the level;
    // 2 (ILU lower): vertices below i join, those above are explored while within the level
    template <int KIND>
    void run(const std::vector<int32_t>& rp, const std::vector<int32_t>& col, int64_t n, int32_t i, int level,
             std::vector<int32_t>& out) {
        out.clear();
        stack.assign(1, i);
        mark[(size_t)i] = i;
        length[(size_t)i] = 0;
        while (!stack.empty()) {
            const int32_t k = stack.back();
            stack.pop_back();
            const int lk = length[(size_t)k];
            for (int32_t kj = rp[(size_t)k]; kj < rp[(size_t)k + 1]; ++kj) {
                const int32_t j = col[(size_t)kj];
                if (j == k || j >= n) continue;   // the square block
                if (mark[(size_t)j] == i) continue;
                mark[(size_t)j] = i;
                if (KIND == 0) {
                    if (lk < level) {
                        stack.push_back(j);
                        length[(size_t)j] = lk + 1;
                        out.push_back(j);
                    }
                } else if (KIND == 1) {
                    if (j < i && lk < level) {
                        stack.push_back(j);
                        length[(size_t)j] = lk + 1;
                    } else if (j > i) {
                        out.push_back(j);
                    }
                } else {
                    if (j > i && lk < level) {
                        stack.push_back(j);
                        length[(size_t)j] = lk + 1;
                    } else if (j < i) {
                        out.push_back(j);
                    }
                }
            }
        }
        std::sort(out.begin(), out.end());
    }
};

// The factor pattern of IC(level) / ILU(level) of a square matrix pattern (rp, col):
// frp/fcol (rows sorted), amap (matrix slot -> factor slot, -1 for a column >= n), and for
// IC the upper sweep's entries in the order of the reference's transposed pattern sp_t
// (each row: the matrix's columns above the diagonal, then the fill's, each ascending) with
// the factor slot of each (-1: an sp_t entry without a factor entry, coefficient 0).
// FVG_E_MISSING_ENTRY where the reference's l_vt construction would fail (a factor entry
// whose transpose is not in sp_t).
struct FillPattern {
    std::vector<int32_t> frp, fcol, amap, urp, ucol, usrc;
    bool transposed = true;   // IC: the upper entries are exactly the lower ones transposed
};

static int fill_pattern(const std::vector<int32_t>& rp, const std::vector<int32_t>& col, int64_t n, int kind,
                        int level, FillPattern& F) {
    std::vector<std::vector<int32_t>> rows((size_t)n);
    std::vector<int32_t> s1, s2, arow;
    if (kind == kIC) {
        Walk w(n);
        std::vector<std::vector<int32_t>> lrows((size_t)n), trows((size_t)n);
        for (int64_t i = 0; i < n; ++i) {
            w.run<0>(rp, col, n, (int32_t)i, level, s1);
            arow.clear();
            for (int32_t k = rp[(size_t)i]; k < rp[(size_t)i + 1]; ++k)
                if (col[(size_t)k] < n) arow.push_back(col[(size_t)k]);
            auto& L = lrows[(size_t)i];
            auto& T = trows[(size_t)i];
            for (int32_t j : arow) {
                if (j <= i) L.push_back(j);
                if (j >= i) T.push_back(j);
            }
            for (int32_t j : s1) {
                if (std::binary_search(arow.begin(), arow.end(), j)) continue;   // arow is sorted
                if (j <= i) L.push_back(j);
                if (j >= i) T.push_back(j);
            }
            std::sort(L.begin(), L.end());   // from_sparsity_and_values sorts the factor's rows
        }
        // the factor pattern: L and its transpose, rows sorted
        for (int64_t i = 0; i < n; ++i)
            for (int32_t j : lrows[(size_t)i]) {
                rows[(size_t)i].push_back(j);
                if (j != i) rows[(size_t)j].push_back((int32_t)i);
            }
        for (auto& r : rows) std::sort(r.begin(), r.end());
        F.frp.assign((size_t)n + 1, 0);
        for (int64_t i = 0; i < n; ++i) F.frp[(size_t)i + 1] = F.frp[(size_t)i] + (int32_t)rows[(size_t)i].size();
        F.fcol.reserve((size_t)F.frp[(size_t)n]);
        for (auto& r : rows) F.fcol.insert(F.fcol.end(), r.begin(), r.end());
        auto slot = [&](int32_t i, int32_t j) -> int32_t {   // factor slot of (i, j), or -1
            auto b = F.fcol.begin() + F.frp[(size_t)i], e = F.fcol.begin() + F.frp[(size_t)i + 1];
            auto it = std::lower_bound(b, e, j);
            return (it != e && *it == j) ? (int32_t)(it - F.fcol.begin()) : -1;
        };
        // every factor entry (r, c), c < r, needs r in sp_t row c
        for (int64_t r = 0; r < n; ++r)
            for (int32_t c : lrows[(size_t)r]) {
                if (c == r) continue;
                const auto& T = trows[(size_t)c];
                if (std::find(T.begin(), T.end(), (int32_t)r) == T.end())
                    return fail(FVG_E_MISSING_ENTRY, "IncompleteCholesky: transpose entry (" + std::to_string(c) +
                                                         ", " + std::to_string(r) + ") missing from the fill pattern");
            }
        // the upper sweep: sp_t order, coefficient l_ji (the factor slot of (j, i)) or 0
        F.urp.assign((size_t)n + 1, 0);
        for (int64_t i = 0; i < n; ++i) {
            for (int32_t j : trows[(size_t)i]) {
                if (j <= i) continue;
                const int32_t sl = std::binary_search(lrows[(size_t)j].begin(), lrows[(size_t)j].end(), (int32_t)i)
                                       ? slot(j, (int32_t)i) : -1;
                if (sl < 0) F.transposed = false;
                F.ucol.push_back(j);
                F.usrc.push_back(sl);
            }
            F.urp[(size_t)i + 1] = (int32_t)F.ucol.size();
        }
        // every factor entry above the diagonal is an sp_t entry (checked above), so the
        // upper entries are the lower ones transposed unless sp_t has extra entries
        int64_t lower_strict = 0;
        for (const auto& L : lrows) lower_strict += (int64_t)L.size() - 1;
        if ((int64_t)F.ucol.size() != lower_strict) F.transposed = false;
    } else {
        Walk wu(n), wl(n);
        for (int64_t i = 0; i < n; ++i) {
            wl.run<2>(rp, col, n, (int32_t)i, level, s1);
            wu.run<1>(rp, col, n, (int32_t)i, level, s2);
            auto& r = rows[(size_t)i];
            r = s1;
            r.push_back((int32_t)i);
            r.insert(r.end(), s2.begin(), s2.end());
        }
        F.frp.assign((size_t)n + 1, 0);
        for (int64_t i = 0; i < n; ++i) F.frp[(size_t)i + 1] = F.frp[(size_t)i] + (int32_t)rows[(size_t)i].size();
        for (auto& r : rows) F.fcol.insert(F.fcol.end(), r.begin(), r.end());
    }
    // the matrix's slots in the factor pattern
    F.amap.assign(col.size(), -1);
    for (int64_t i = 0; i < n; ++i)
        for (int32_t k = rp[(size_t)i]; k < rp[(size_t)i + 1]; ++k) {
            const int32_t j = col[(size_t)k];
            if (j >= n) continue;
            auto b = F.fcol.begin() + F.frp[(size_t)i], e = F.fcol.begin() + F.frp[(size_t)i + 1];
            auto it = std::lower_bound(b, e, j);
            if (it != e && *it == j) F.amap[(size_t)k] = (int32_t)(it - F.fcol.begin());
        }
    return FVG_OK;
}

// whether every entry (i, j), j < n, of the pattern has (j, i) too
static bool structurally_symmetric(const std::vector<int32_t>& rp, const std::vector<int32_t>& col, int64_t n) {
    for (int64_t i = 0; i < n; ++i)
        for (int32_t k = rp[(size_t)i]; k < rp[(size_t)i + 1]; ++k) {
            const int32_t j = col[(size_t)k];
            if (j >= n) continue;
            auto b = col.begin() + rp[(size_t)j], e = col.begin() + rp[(size_t)j + 1];
            if (!std::binary_search(b, e, (int32_t)i)) return false;
        }
    return true;
}

// the factorisation's input: the matrix values, or (with fill) scattered into the factor
// pattern with zeros elsewhere
__global__ void k_scatter_values(int64_t nnz, const int32_t* __restrict__ amap, const double* __restrict__ a,
                                 double* __restrict__ out) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nnz) {
        const int32_t t = amap[k];
        if (t >= 0) out[t] = a[k];
    }
}

static const double* factor_input(fvg_ctx* c, const fvg_prec* p, const fvg_csr* a) {
    if (p->level == 0) return a->val;
    (void)launch_fill(c, p->nnz, p->d_afill, 0.0);
    hipLaunchKernelGGL(k_scatter_values, dim3(grid_for(p->a_nnz)), dim3(kBlock), 0, c->stream, p->a_nnz, p->d_amap,
                       (const double*)a->val, p->d_afill);
    return p->d_afill;
}

// the matrix a has the pattern p was analysed for
static bool same_pattern(const fvg_prec* p, const fvg_csr* a) {
    if (a->nrows != p->n || a->nnz != p->a_nnz) return false;
    if (a->pattern_id != 0 && a->pattern_id == p->pattern_id) return true;
    return a->h_col == (p->level == 0 ? p->h_col : p->h_acol);
}

static int create_factored(fvg_ctx* c, const fvg_csr* a, int32_t level, int kind, fvg_prec** out) {
    if (!c || !a || !out) return fail(FVG_E_INVALID, "null argument");
    *out = nullptr;
    if (level < 0) return fail(FVG_E_INVALID, "negative fill level");
    if (a->nrows <= 0) return fail(FVG_E_INVALID, "empty matrix");
    const int64_t n = a->nrows;
    auto* p = new fvg_prec();
    p->kind = kind;
    p->level = level;
    p->pattern_id = a->pattern_id;
    p->a_nnz = a->nnz;
    FillPattern F;
    int s = FVG_OK;
    if (level > 0) {
        s = fill_pattern(a->h_row_ptr, a->h_col, n, kind, level, F);
        p->h_arp = a->h_row_ptr;
        p->h_acol = a->h_col;
    }
    if (s == FVG_OK)
        s = level > 0 ? setup_pattern(c, n, F.frp, F.fcol, p, kind == kIC)
                      : setup_pattern(c, n, a->h_row_ptr, a->h_col, p, kind == kIC);
    if (s == FVG_OK && level > 0) {
        if ((s = dev_alloc_n(c, p->a_nnz, &p->d_amap)) == FVG_OK && (s = dev_alloc_n(c, p->nnz, &p->d_afill)) == FVG_OK)
            s = upload(c, p->d_amap, F.amap.data(), p->a_nnz);
    }
    // one slot past the factor values holds 0: the coefficient of sweep entries without a
    // factor entry
    if (s == FVG_OK) s = dev_alloc_n(c, p->nnz + 1, &p->d_fac);
    if (s == FVG_OK) s = dev_alloc_n(c, n, &p->d_tmp);
    // the rows' coordinates: the matrix's own, or its mesh's cell centres, read once and
    // only if a layout analysis prices boxes of them
    std::vector<float> centre;
    int centre_dim = 0;
    bool centre_read = false;
    const CoordSource X = [&]() -> RowCoords {
        if (a->coord_dim > 0 && (int64_t)a->h_coords.size() == n * a->coord_dim)
            return RowCoords{a->h_coords.data(), a->coord_dim};
        if (a->coord_mesh && !centre_read) {
            centre_read = true;
            if (mesh_cell_centres(c, a->coord_mesh, a->coord_mesh_id, centre, centre_dim) != FVG_OK ||
                (int64_t)centre.size() != n * centre_dim)
                centre.clear();
        }
        return centre.empty() ? RowCoords() : RowCoords{centre.data(), centre_dim};
    };
    if (s == FVG_OK) s = build_sweep(c, p->h_rp, p->h_col, n, false, nullptr, p->lower, nullptr, X);
    if (s == FVG_OK) {
        if (kind == kIC && level > 0) {
            for (int32_t& v : F.usrc)
                if (v < 0) v = (int32_t)p->nnz;
            s = build_sweep(c, F.urp, F.ucol, n, true, &F.usrc, p->upper, F.transposed ? &p->lower : nullptr, X);
        } else {
            // IC(0): the upper entries are the lower ones transposed; ILU: when the pattern is symmetric
            const bool tr = kind == kIC || structurally_symmetric(p->h_rp, p->h_col, n);
            s = build_sweep(c, p->h_rp, p->h_col, n, true, kind == kIC ? &p->h_tpos : nullptr, p->upper,
                            tr ? &p->lower : nullptr, X);
        }
    }
    if (s == FVG_OK) s = pair_layouts(c, p->lower, p->upper);
    if (s == FVG_OK) s = launch_fill(c, p->nnz + 1, p->d_fac, 0.0);
    if (s == FVG_OK) s = factorize(c, p, a);
    if (s != FVG_OK) {
        (void)hipStreamSynchronize(c->stream);
        free_prec(c, p);
        delete p;
        return s;
    }
    *out = p;
    return FVG_OK;
}

}  // namespace fvg

using namespace fvg;

extern "C" {

int fvg_identity_create(fvg_ctx* c, int64_t n, fvg_prec** out) {
    if (!c || !out || n < 0) return fail(FVG_E_INVALID, "bad argument");
    auto* p = new fvg_prec();
    p->kind = kIdentity;
    p->n = n;
    *out = p;
    return FVG_OK;
}

int fvg_diagonal_create(fvg_ctx* c, const fvg_csr* a, fvg_prec** out) {
    if (!c || !a || !out) return fail(FVG_E_INVALID, "null argument");
    *out = nullptr;
    auto* p = new fvg_prec();
    p->kind = kDiagonal;
    p->pattern_id = a->pattern_id;
    p->a_nnz = a->nnz;
    int s = setup_pattern(c, a->nrows, a->h_row_ptr, a->h_col, p, false);
    if (s == FVG_OK) {
        hipLaunchKernelGGL(k_inv_diag, dim3(grid_for(p->n)), dim3(kBlock), 0, c->stream, p->n, p->d_diag, a->val,
                           p->d_inv_diag);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess)
            s = fail(FVG_E_HIP, "diagonal preconditioner launch failed");
    }
    if (s != FVG_OK) {
        free_prec(c, p);
        delete p;
        return s;
    }
    *out = p;
    return FVG_OK;
}

int fvg_ic_create(fvg_ctx* c, const fvg_csr* a, int32_t level, fvg_prec** out) {
    return create_factored(c, a, level, kIC, out);
}

int fvg_ilu_create(fvg_ctx* c, const fvg_csr* a, int32_t level, fvg_prec** out) {
    return create_factored(c, a, level, kILU, out);
}

static int refactor(fvg_ctx* c, fvg_prec* p, const fvg_csr* a, bool check) {
    if (!c || !p || !a) return fail(FVG_E_INVALID, "null argument");
    if (p->kind == kIdentity) return a->nrows == p->n ? FVG_OK : fail(FVG_E_INVALID, "matrix size differs");
    // the same mesh pattern is recognised by its id; other matrices are compared entry by entry
    if (!same_pattern(p, a)) return fail(FVG_E_INVALID, "matrix pattern differs from the analysed one");
    if (p->kind == kDiagonal) {
        hipLaunchKernelGGL(k_inv_diag, dim3(grid_for(p->n)), dim3(kBlock), 0, c->stream, p->n, p->d_diag, a->val,
                           p->d_inv_diag);
        FVG_HIP(hipGetLastError());
        p->fac_version += 1;
        return FVG_OK;
    }
    return factorize(c, p, a, check);
}

int fvg_prec_sweep_layout(const fvg_prec* p, int32_t* info, int32_t count) {
    if (!p || (count > 0 && !info)) return fail(FVG_E_INVALID, "null argument");
    const int64_t v[] = {p->lower.nblk, p->upper.nblk, p->lower.nlevels, p->lower.bW, p->lower.bX,
                         p->lower.nbx,  p->upper.bW,   p->upper.bX,       p->lower.bK, p->upper.bK,
                         p->upper.mirrored ? 1 : 0,    p->lower.nblk ? p->lower.bL : 0};
    for (int k = 0; k < count && k < (int)(sizeof(v) / sizeof(v[0])); ++k)
        info[k] = (int32_t)std::min<int64_t>(v[k], INT32_MAX);
    return FVG_OK;
}

int fvg_prec_refactor(fvg_ctx* c, fvg_prec* p, const fvg_csr* a) { return refactor(c, p, a, true); }

int fvg_prec_refactor_async(fvg_ctx* c, fvg_prec* p, const fvg_csr* a) { return refactor(c, p, a, false); }

int fvg_prec_refactor_many(fvg_ctx* c, int32_t count, fvg_prec* const* ps, const fvg_csr* const* as) {
    if (!c || count < 0 || (count && (!ps || !as))) return fail(FVG_E_INVALID, "null argument");
    std::vector<FactorJob> jobs;
    for (int k = 0; k < count; ++k) {
        fvg_prec* p = ps[k];
        const fvg_csr* a = as[k];
        if (!p || !a) return fail(FVG_E_INVALID, "null preconditioner or matrix");
        if (p->kind != kIC && p->kind != kILU) {   // Jacobi / identity: nothing to batch
            FVG_TRY(refactor(c, p, a, false));
            continue;
        }
        if (!same_pattern(p, a)) return fail(FVG_E_INVALID, "matrix pattern differs from the analysed one");
        FactorJob J{};
        J.kind = p->kind;
        J.nlevels = p->lower.nlevels;
        J.lvl_ptr = p->lower.d_lvl_ptr;
        J.rows = p->lower.rows;
        J.rp = p->d_rp;
        J.col = p->d_col;
        J.diag = p->d_diag;
        J.tpos = p->d_tpos;
        J.a = factor_input(c, p, a);
        J.fac = p->d_fac;
        J.inv_diag = p->d_inv_diag;
        J.flags = p->d_flags;
        J.lo_src = p->lower.src;
        J.lo_val = p->lower.val;
        J.up_src = p->upper.src;
        J.up_val = p->upper.val;
        J.lo_bsrc = p->lower.bsrc;
        J.lo_bval = p->lower.bval;
        J.up_bsrc = p->upper.bsrc;
        J.up_bval = p->upper.bval;
        J.lo_bne = p->lower.nblk ? (int64_t)p->lower.bK * p->lower.n : 0;
        J.up_bne = p->upper.nblk ? (int64_t)p->upper.bK * p->upper.n : 0;
        J.lo_brow = p->lower.brow;
        J.lo_spos = p->lower.spos;
        J.up_brow = p->upper.brow;
        J.up_spos = p->upper.spos;
        J.lo_bn = p->lower.nblk ? p->lower.n : 0;
        J.up_bn = p->upper.nblk ? p->upper.n : 0;
        J.n = p->n;
        J.nnz = p->nnz;
        J.lo_nnz = p->lower.nnz;
        J.up_nnz = p->upper.nnz;
        jobs.push_back(J);
        FVG_HIP(hipMemsetAsync(p->d_flags, 0, 4 * sizeof(int32_t), c->stream));
        p->fac_version += 1;
    }
    if (jobs.empty()) return FVG_OK;
    // job table: a grow-only device buffer of the context, copied from a pinned host twin;
    // the twin is rewritten only after the previous copy out of it has run
    const size_t bytes = sizeof(FactorJob) * jobs.size();
    if (c->jobs_cap < bytes) {
        FVG_HIP(hipStreamSynchronize(c->stream));
        if (c->jobs) dev_free(c, c->jobs, c->jobs_cap);
        if (c->jobs_host) (void)hipHostFree(c->jobs_host);
        c->jobs = nullptr;
        c->jobs_host = nullptr;
        c->jobs_cap = 0;
        size_t cap = 4096;
        while (cap < bytes) cap <<= 1;
        FVG_TRY(dev_alloc(c, cap, &c->jobs));
        FVG_HIP(hipHostMalloc(&c->jobs_host, cap, hipHostMallocDefault));
        c->jobs_cap = cap;
    }
    if (!c->jobs_evt) FVG_HIP(hipEventCreateWithFlags(&c->jobs_evt, hipEventDisableTiming));
    else FVG_HIP(hipEventSynchronize(c->jobs_evt));
    std::memcpy(c->jobs_host, jobs.data(), bytes);
    FVG_HIP(hipMemcpyAsync(c->jobs, c->jobs_host, bytes, hipMemcpyHostToDevice, c->stream));
    FVG_HIP(hipEventRecord(c->jobs_evt, c->stream));
    hipLaunchKernelGGL(k_factor_team, dim3((unsigned)jobs.size()), dim3(kFactorBlock), 0, c->stream,
                       static_cast<const FactorJob*>(c->jobs));
    FVG_HIP(hipGetLastError());
    return FVG_OK;
}

int fvg_prec_status(fvg_ctx* c, const fvg_prec* p) {
    if (!c || !p) return fail(FVG_E_INVALID, "null argument");
    if (!p->d_flags || p->kind == kDiagonal || p->kind == kIdentity) return FVG_OK;
    int32_t fl[4];
    FVG_TRY(read_flags(c, p, fl));
    if (p->kind == kIC && fl[2]) return fail(FVG_E_NOT_SYMMETRIC, "IncompleteCholesky requires a symmetric matrix");
    if (p->kind == kILU && fl[0]) return fail(FVG_E_ZERO_PIVOT, "error, lu[k, k] too small");
    return FVG_OK;
}

int fvg_prec_destroy(fvg_ctx* c, fvg_prec* p) {
    if (!p) return FVG_OK;
    if (c) {
        (void)hipStreamSynchronize(c->stream);
        free_prec(c, p);
    }
    delete p;
    return FVG_OK;
}

int fvg_precondition(fvg_ctx* c, const fvg_prec* p, fvg_vec* solution, const fvg_vec* rhs) {
    if (!c || !p || !solution || !rhs) return fail(FVG_E_INVALID, "null argument");
    if (solution->n != p->n || rhs->n != p->n) return fail(FVG_E_INVALID, "vector length mismatch");
    if (solution->d == rhs->d) return fail(FVG_E_INVALID, "solution and rhs must not alias");
    FVG_TRY(apply_prec(c, p, rhs->d, solution->d, nullptr));
    FVG_HIP(hipStreamSynchronize(c->stream));
    return check_sweep_status(c, p);
}

int fvg_prec_factor_pattern(const fvg_prec* p, int64_t* n, int64_t* nnz, int32_t* row_ptr, int32_t* col) {
    if (!p) return fail(FVG_E_INVALID, "null argument");
    if (n) *n = p->n;
    if (nnz) *nnz = p->nnz;
    if (row_ptr && !p->h_rp.empty()) std::memcpy(row_ptr, p->h_rp.data(), sizeof(int32_t) * p->h_rp.size());
    if (col && !p->h_col.empty()) std::memcpy(col, p->h_col.data(), sizeof(int32_t) * p->h_col.size());
    return FVG_OK;
}

int fvg_prec_factor_download(fvg_ctx* c, const fvg_prec* p, double* values, double* inv_diag) {
    if (!c || !p) return fail(FVG_E_INVALID, "null argument");
    if (values) {
        if (!p->d_fac) return fail(FVG_E_INVALID, "preconditioner has no factor");
        FVG_TRY(download(c, values, p->d_fac, p->nnz));
    }
    if (inv_diag) {
        if (!p->d_inv_diag) return fail(FVG_E_INVALID, "preconditioner has no diagonal");
        FVG_TRY(download(c, inv_diag, p->d_inv_diag, p->n));
    }
    return FVG_OK;
}

int fvg_prec_levels(const fvg_prec* p, int32_t* lower_levels, int32_t* upper_levels) {
    if (!p) return fail(FVG_E_INVALID, "null argument");
    if (lower_levels) *lower_levels = p->lower.nlevels;
    if (upper_levels) *upper_levels = p->upper.nlevels;
    return FVG_OK;
}

int fvg_prec_nonpositive_pivots(fvg_ctx* c, const fvg_prec* p, int64_t* count) {
    if (!c || !p || !count) return fail(FVG_E_INVALID, "null argument");
    *count = 0;
    if (p->kind != kIC) return FVG_OK;
    int32_t fl[4];
    FVG_TRY(read_flags(c, p, fl));
    *count = fl[1];
    return FVG_OK;
}

}  // extern "C"
