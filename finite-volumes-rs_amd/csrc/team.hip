// team.hip -- batched solves: one workgroup runs one system's whole Krylov solve.
//
// A batch holds many independent systems A_k x_k = b_k (each with its own matrix and
// IC(0) / ILU(0) / Jacobi / identity preconditioner).  The solve of one system is a
// single persistent workgroup: SpMV, triangular sweeps, vector updates and dot products
// all run inside it, separated by workgroup barriers, so a system never waits on another
// workgroup and the whole batch is one kernel launch (replayed from a captured hipGraph).
// Each result is therefore independent of what else is in the batch: a system gives the
// same bits alone and in any mix, and batches on different streams overlap freely.
//
// Storage is permuted into "positions": rows sorted by the level of the lower triangular
// sweep (dependency depth), so a level is a contiguous range.  The upper sweep runs the
// same levels in reverse (valid for structurally symmetric patterns, which mesh matrices
// are).  Entries keep the original ascending column order, so per-row sums are the
// reference's (dmatrix.rs:136-146, incomplete_cholesky.rs:217-254,
// incomplete_lower_upper.rs:287-328).
//
// Triangular sweeps inside the workgroup: the rows of a sweep are cut into steps (a
// level, or a slice of a level) and the steps into groups that fit one LDS stage buffer.
// While a group is computed step by step (one barrier per step), the next group's static
// data (step bounds, row pointers, dependency positions, factor values, scales) and its
// right-hand side are already in flight into registers; they land in the second stage
// buffer at the end of the group.  Solved values go to an LDS ring indexed by position,
// from which later rows read their dependencies (when every dependency lies within the
// ring's span, which the host checks per sweep), and to HBM for the passes after the
// sweep.
//
// Solver variants (the algorithm steps are those of krylov.hip / mixed.hip):
//   CG64 / BI64   conjugate_gradient / bi_conjugate_gradient_stab in FP64, with the
//                 reference's stopping tests, iteration counting and history;
//   CG32 / BI32   FP32 inner iterations (FP32 values, factors, vectors and accumulation)
//                 inside FP64 iterative refinement on the true residual b - A x.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <type_traits>

#include "precond.hpp"
#include "reduce.hpp"

namespace fvg {

enum TeamVariant { kCG64 = 0, kBI64 = 1, kCG32 = 2, kBI32 = 3 };

constexpr int kRing64 = 16384;  // LDS ring of values by position (sweeps, SpMV window), FP64
constexpr int kRing32 = 32768;  // FP32
constexpr int kES = 4;          // factor entries per row held in registers (the rest read on use)
constexpr int kEA = 4;          // matrix entries per row held in registers (the rest read on use)
constexpr int kMaxGroup = 1024; // largest workgroup: the host's ring checks assume groups this large

struct TeamSweep {
    const int32_t* lvl;      // level of each processing index (non-decreasing)
    int32_t nnz1;            // index of the last entry (dep / val hold at least one element)
    int32_t pad1;
    const int32_t* ptr;      // entries of processing index q: [ptr[q], ptr[q+1])
    const int32_t* dep;      // position of each dependency
    const double* val;
    const float* val32;
    int32_t ring64;          // every dependency lies within the FP64 ring's span
    int32_t ring32;
    int32_t upper;           // processing index q is position n-1-q
    int32_t pad;
};

// device-side description of one system (uploaded once per batch, updated when b/x move)
struct TeamDesc {
    int32_t n, npad, kind, wspan;   // wspan: largest |position - column position| of A
    int32_t win64, win32, a_nnz1, pad1;   // the SpMV x window fits the FP64 / FP32 ring; last entry of A
    const int32_t* perm;     // natural row of each position
    const int32_t* a_ptr;    // A by position
    const int32_t* a_dep;
    const double* a_val;
    const float* a_val32;
    TeamSweep lo, up;
    const double* sc;        // per position: IC 1/l_ii, ILU 1/u_ii, Jacobi 1/a_ii
    const float* sc32;
    const double* b;         // caller vectors, natural order
    double* x;
    double* w64;             // workspace, vectors of npad elements
    float* w32;
    double* hist;
    int64_t hcap;
    fvg_solver_info* info;
    double atol, rtol, inner_tol;
    int64_t max_iter;
    int64_t* prof;           // optional: wall-clock ticks spent in SpMV, lower sweep, upper sweep, total
    int32_t* done;           // host-visible completion flag of this system (pinned host memory)
    const int32_t* pflags;   // the preconditioner's factorisation flags (IC / ILU), checked before solving
};

// ---- optional phase timing (thread 0's view of its workgroup) ----------------------------
__shared__ long long g_prof[4];
__shared__ int g_prof_on;
__shared__ int g_status;

struct ProfScope {
    int ph;
    long long t0;
    __device__ explicit ProfScope(int p) : ph(p), t0(0) {
        if (threadIdx.x == 0 && g_prof_on) t0 = wall_clock64();
    }
    __device__ ~ProfScope() {
        if (threadIdx.x == 0 && g_prof_on) g_prof[ph] += wall_clock64() - t0;
    }
};

// ---- typed memory access -------------------------------------------------------------------
// The descriptor lives in LDS, so every pointer read from it is generic, and generic
// (flat) accesses count against both the vector-memory and the LDS wait counters: a wait
// for an LDS value then also waits for every global load in flight, which undoes the
// prefetching.  The hot loops therefore access memory through pointers typed global (gm)
// or LDS (lm).
template <class T>
using gptr_t = __attribute__((address_space(1))) T*;
template <class T>
using lptr_t = __attribute__((address_space(3))) T*;
template <class T>
__device__ __forceinline__ gptr_t<T> gm(T* p) {
    return (gptr_t<T>)p;
}
template <class T>
__device__ __forceinline__ lptr_t<T> lm(T* p) {
    return (lptr_t<T>)p;
}

// ---- workgroup reductions (result in every thread, fixed order) -------------------------
template <int NT, class T>
struct Reducer {
    T* buf;   // 2 parities x 2 values x NT/64 waves
    int par;
    __device__ T sum(T v) {
        v = wave_sum(v);
        const lptr_t<T> r = lm(buf) + par * 2 * (NT / 64);
        par ^= 1;
        if ((threadIdx.x & 63) == 0) r[threadIdx.x >> 6] = v;
        __syncthreads();
        T s = T(0);
#pragma unroll
        for (int i = 0; i < NT / 64; ++i) s += r[i];
        return s;
    }
    __device__ void sum2(T& a, T& b) {
        a = wave_sum(a);
        b = wave_sum(b);
        const lptr_t<T> r = lm(buf) + par * 2 * (NT / 64);
        par ^= 1;
        if ((threadIdx.x & 63) == 0) {
            r[threadIdx.x >> 6] = a;
            r[NT / 64 + (threadIdx.x >> 6)] = b;
        }
        __syncthreads();
        T s = T(0), t = T(0);
#pragma unroll
        for (int i = 0; i < NT / 64; ++i) {
            s += r[i];
            t += r[NT / 64 + i];
        }
        a = s;
        b = t;
    }
};

template <class T>
__device__ __forceinline__ const T* pick(const double* d, const float* f);
template <>
__device__ __forceinline__ const double* pick<double>(const double* d, const float*) { return d; }
template <>
__device__ __forceinline__ const float* pick<float>(const double*, const float* f) { return f; }

// ---- triangular sweeps: one row per thread per group of NT rows ------------------------
// Rows are taken in processing order in groups of NT; thread t owns row t of each group.
// A group spans levels [La, Lb] of the sweep; the group is computed level by level with
// one barrier per level, each thread solving its row at its level.  A row's data (first
// kES entries, right-hand side, scale, level) is fetched into registers while the
// previous group is being computed, and its entry range two groups ahead.  Every fetch is
// unconditional (indices clamped into range), so each stage issues the same number of
// loads and the wait for a stage's data never includes the loads issued after it.
// Solved values go to the LDS ring (indexed by position) and to HBM.
template <class T>
struct SweepRow {
    int kb, cnt, lvl;
    int dep[kES];
    T val[kES];
    T b, sc;
};

template <int NT>
__device__ __forceinline__ void sweep_ptr(const TeamSweep& S, int n, int g, int& kb, int& ke) {
    const int q = g * NT + (int)threadIdx.x;
    const int qc = min(q, n - 1);
    kb = gm(S.ptr)[qc];
    ke = gm(S.ptr)[qc + 1];
    if (q >= n) ke = kb;
}

template <int NT, class T>
__device__ __forceinline__ void sweep_row(SweepRow<T>& R, const TeamSweep& S, const T* __restrict__ val,
                                          const T* __restrict__ sc, const T* __restrict__ b, int n, int g, int kb,
                                          int ke) {
    const int q = g * NT + (int)threadIdx.x;
    const int qc = min(q, n - 1);
    const int pos = S.upper ? n - 1 - qc : qc;
    R.kb = kb;
    R.cnt = ke - kb;
    const int lv = gm(S.lvl)[qc];
    R.lvl = q < n ? lv : -1;
    R.b = gm(b)[pos];
    R.sc = gm(sc)[pos];
    const int last = max(R.cnt - 1, 0);
#pragma unroll
    for (int e = 0; e < kES; ++e) {
        const int k = min(kb + min(e, last), S.nnz1);
        R.dep[e] = gm(S.dep)[k];
        R.val[e] = gm(val)[k];
    }
}

// levels spanned by group g
template <int NT>
__device__ __forceinline__ void group_levels(const TeamSweep& S, int n, int g, int& la, int& lb) {
    const int q0 = g * NT;
    const int q1 = min(q0 + NT, n);
    la = gm(S.lvl)[q0];
    lb = gm(S.lvl)[q1 - 1];
}

template <int NT, class T, int RING>
__device__ __forceinline__ void sweep_compute(const SweepRow<T>& R, const TeamSweep& S, const T* __restrict__ val,
                                              bool has_sc, bool ring_ok, int n, int g, int la, int lb, T* out,
                                              T* ring) {
    const int q = g * NT + (int)threadIdx.x;
    const int pos = S.upper ? n - 1 - q : q;
    for (int L = la; L <= lb; ++L) {
        if (R.lvl == L) {
            T s = R.b;
            if (ring_ok) {
                // every slot's read issued at once (a slot past the row's count holds the
                // row's last dependency), each product kept only for the row's own slots
                T v[kES];
#pragma unroll
                for (int e = 0; e < kES; ++e) v[e] = lm(ring)[R.dep[e] & (RING - 1)];
#pragma unroll
                for (int e = 0; e < kES; ++e) {
                    const T t = s - R.val[e] * v[e];
                    s = e < R.cnt ? t : s;
                }
                for (int k = R.kb + kES; k < R.kb + R.cnt; ++k) s -= gm(val)[k] * lm(ring)[gm(S.dep)[k] & (RING - 1)];
            } else {
#pragma unroll
                for (int e = 0; e < kES; ++e)
                    if (e < R.cnt) s -= R.val[e] * gm(out)[R.dep[e]];
                for (int k = R.kb + kES; k < R.kb + R.cnt; ++k) s -= gm(val)[k] * gm(out)[gm(S.dep)[k]];
            }
            const T r = has_sc ? R.sc * s : s;
            if (ring_ok) lm(ring)[pos & (RING - 1)] = r;
            gm(out)[pos] = r;
        }
        __syncthreads();
    }
}

// x = sweep(b); sc may be absent (unit diagonal), in which case scl is still a valid
// array of n values (its loads are discarded)
template <int NT, class T, int RING>
__device__ void team_sweep(const TeamSweep& Sd, const T* val, const T* sc, const T* scl, const T* b, T* out, int n,
                           T* ring) {
    const TeamSweep S = Sd;   // registers: no descriptor reads inside the pipelined loop
    ProfScope prof(S.upper ? 2 : 1);
    const bool ring_ok = ring != nullptr && (sizeof(T) == 8 ? S.ring64 != 0 : S.ring32 != 0);
    const bool has_sc = sc != nullptr;
    const int G = (n + NT - 1) / NT;
    SweepRow<T> A, B;
    int akb, ake, bkb, bke;
    int ala, alb, bla, blb;
    sweep_ptr<NT>(S, n, 0, akb, ake);
    sweep_row<NT, T>(A, S, val, scl, b, n, 0, akb, ake);
    group_levels<NT>(S, n, 0, ala, alb);
    sweep_ptr<NT>(S, n, min(1, G - 1), bkb, bke);
    for (int g = 0; g < G; g += 2) {
        const int g1 = min(g + 1, G - 1), g2 = min(g + 2, G - 1), g3 = min(g + 3, G - 1);
        sweep_row<NT, T>(B, S, val, scl, b, n, g1, bkb, bke);
        group_levels<NT>(S, n, g1, bla, blb);
        sweep_ptr<NT>(S, n, g2, akb, ake);
        sweep_compute<NT, T, RING>(A, S, val, has_sc, ring_ok, n, g, ala, alb, out, ring);
        if (g + 1 >= G) break;
        sweep_row<NT, T>(A, S, val, scl, b, n, g2, akb, ake);
        group_levels<NT>(S, n, g2, ala, alb);
        sweep_ptr<NT>(S, n, g3, bkb, bke);
        sweep_compute<NT, T, RING>(B, S, val, has_sc, ring_ok, n, g + 1, bla, blb, out, ring);
    }
}

// z = M^-1 r (ends with a barrier); tmp holds the intermediate vector
template <int NT, class T, int RING>
__device__ void team_prec(const TeamDesc& D, const T* r, T* z, T* tmp, T* ring) {
    const T* sc = pick<T>(D.sc, D.sc32);
    switch (D.kind) {
    case kIdentity:
        for (int p = threadIdx.x; p < D.n; p += NT) gm(z)[p] = gm(r)[p];
        __syncthreads();
        break;
    case kDiagonal:
        for (int p = threadIdx.x; p < D.n; p += NT) gm(z)[p] = gm(sc)[p] * gm(r)[p];
        __syncthreads();
        break;
    case kIC:
        team_sweep<NT, T, RING>(D.lo, pick<T>(D.lo.val, D.lo.val32), sc, sc, r, tmp, D.n, ring);
        team_sweep<NT, T, RING>(D.up, pick<T>(D.up.val, D.up.val32), sc, sc, tmp, z, D.n, ring);
        break;
    default:   // kILU: unit lower factor
        team_sweep<NT, T, RING>(D.lo, pick<T>(D.lo.val, D.lo.val32), (const T*)nullptr, sc, r, tmp, D.n, ring);
        team_sweep<NT, T, RING>(D.up, pick<T>(D.up.val, D.up.val32), sc, sc, tmp, z, D.n, ring);
        break;
    }
}

// ---- SpMV: y_p = sum_k a_k x_{dep_k} in stored (ascending column) order ------------------
// Thread t owns row t of each tile of NT rows.  Row data is fetched one tile ahead (entry
// ranges two tiles ahead), together with the row's value of one auxiliary vector that the
// epilogue needs.  When the system's column spread allows (|p - dep| <= w for every
// entry), x is read from an LDS window that slides with the tiles: the slice the next
// tile adds is fetched with contiguous loads while the current tile computes.  As in the
// sweeps every fetch is unconditional; rows past n (up to the padded length) compute
// zeros.  epi(p, y_p, aux_p) consumes each row.
template <class T>
struct SpRow {
    int kb, cnt;
    int dep[kEA];
    T val[kEA];
    T aux, xs;
};

struct SpA {   // the matrix arrays, held in registers during a product
    const int32_t* ptr;
    const int32_t* dep;
    int n, nnz1, npad, w;
};

template <int NT>
__device__ __forceinline__ void spmv_ptr(const SpA& M, int g, int& kb, int& ke) {
    const int q = g * NT + (int)threadIdx.x;
    const int qc = min(q, M.n - 1);
    kb = gm(M.ptr)[qc];
    ke = gm(M.ptr)[qc + 1];
    if (q >= M.n) ke = kb;
}

template <int NT, class T>
__device__ __forceinline__ void spmv_row(SpRow<T>& R, const SpA& M, const T* __restrict__ val,
                                         const T* __restrict__ aux, const T* __restrict__ x, int g, int kb, int ke) {
    const int q = g * NT + (int)threadIdx.x;
    R.kb = kb;
    R.cnt = ke - kb;
    const int last = max(R.cnt - 1, 0);
#pragma unroll
    for (int e = 0; e < kEA; ++e) {
        const int k = min(kb + min(e, last), M.nnz1);
        R.dep[e] = gm(M.dep)[k];
        R.val[e] = gm(val)[k];
    }
    R.aux = gm(aux)[min(q, M.npad - 1)];
    R.xs = gm(x)[min(q + NT + M.w, M.npad - 1)];   // window slice this tile adds for the next one
}

template <int NT, class T, int RING, class Epi>
__device__ __forceinline__ void spmv_compute(const SpRow<T>& R, const SpA& M, const T* __restrict__ val,
                                             const T* __restrict__ x, const T* ring, bool win, int g, Epi& epi) {
    const int q = g * NT + (int)threadIdx.x;
    T s = T(0);
    if (win) {
        // every slot's window read issued at once (a slot past the row's count holds its
        // last entry), each product kept only for the row's own slots
        T xv[kEA];
#pragma unroll
        for (int e = 0; e < kEA; ++e) xv[e] = lm(ring)[R.dep[e] & (RING - 1)];
#pragma unroll
        for (int e = 0; e < kEA; ++e) {
            const T t = s + R.val[e] * xv[e];
            s = e < R.cnt ? t : s;
        }
        for (int k = R.kb + kEA; k < R.kb + R.cnt; ++k) s += gm(val)[k] * lm(ring)[gm(M.dep)[k] & (RING - 1)];
    } else {
#pragma unroll
        for (int e = 0; e < kEA; ++e)
            if (e < R.cnt) s += R.val[e] * gm(x)[R.dep[e]];
        for (int k = R.kb + kEA; k < R.kb + R.cnt; ++k) s += gm(val)[k] * gm(x)[gm(M.dep)[k]];
    }
    epi(q, s, R.aux);
}

template <int NT, class T, int RING, class Epi>
__device__ void team_spmv(const TeamDesc& Dd, const T* __restrict__ val, const T* __restrict__ x,
                          const T* __restrict__ aux, T* ring, Epi epi) {
    ProfScope prof(0);
    const SpA M{Dd.a_ptr, Dd.a_dep, Dd.n, Dd.a_nnz1, Dd.npad, Dd.wspan};   // registers from here on
    const int n = M.n;
    const bool win = ring != nullptr && (sizeof(T) == 8 ? Dd.win64 != 0 : Dd.win32 != 0);
    const int G = (n + NT - 1) / NT;
    if (win) {   // window for tile 0: positions [0, NT + w)
        const int e0 = min(M.npad, NT + M.w);
        for (int p = threadIdx.x; p < e0; p += NT) lm(ring)[p & (RING - 1)] = gm(x)[p];
    }
    SpRow<T> A, B;
    int akb, ake, bkb, bke;
    spmv_ptr<NT>(M, 0, akb, ake);
    spmv_row<NT, T>(A, M, val, aux, x, 0, akb, ake);
    spmv_ptr<NT>(M, min(1, G - 1), bkb, bke);
    __syncthreads();
    for (int g = 0; g < G; g += 2) {
        const int g1 = min(g + 1, G - 1), g2 = min(g + 2, G - 1), g3 = min(g + 3, G - 1);
        spmv_row<NT, T>(B, M, val, aux, x, g1, bkb, bke);
        spmv_ptr<NT>(M, g2, akb, ake);
        spmv_compute<NT, T, RING>(A, M, val, x, ring, win, g, epi);
        if (g + 1 >= G) break;
        if (win) {   // tile g+1 needs positions up to (g+2) NT + w: the slice A fetched
            const int pa = (g + 1) * NT + M.w + (int)threadIdx.x;
            if (pa < M.npad) lm(ring)[pa & (RING - 1)] = A.xs;
            __syncthreads();
        }
        spmv_row<NT, T>(A, M, val, aux, x, g2, akb, ake);
        spmv_ptr<NT>(M, g3, bkb, bke);
        spmv_compute<NT, T, RING>(B, M, val, x, ring, win, g + 1, epi);
        if (win && g + 2 < G) {
            const int pb = (g + 2) * NT + M.w + (int)threadIdx.x;
            if (pb < M.npad) lm(ring)[pb & (RING - 1)] = B.xs;
            __syncthreads();
        }
    }
}

// ---- element-wise passes over the padded vectors, four elements per thread step ----------
template <class T>
struct alignas(4 * sizeof(T)) V4 {
    T e[4];
};
template <class T>
using ev4 = T __attribute__((ext_vector_type(4)));
template <class T>
__device__ __forceinline__ V4<T> ld4(const T* p, int v) {
    const ev4<T> a = *reinterpret_cast<gptr_t<const ev4<T>>>(gm(p + 4 * v));
    return V4<T>{{a.x, a.y, a.z, a.w}};
}
template <class T>
__device__ __forceinline__ void st4(T* p, int v, const V4<T>& a) {
    *reinterpret_cast<gptr_t<ev4<T>>>(gm(p + 4 * v)) = ev4<T>{a.e[0], a.e[1], a.e[2], a.e[3]};
}

// p = z; returns partial r.z
template <int NT, class T>
__device__ T pass_start(int nv, const T* __restrict__ r, const T* __restrict__ z, T* __restrict__ p) {
    T q = T(0);
    for (int v = threadIdx.x; v < nv; v += NT) {
        const V4<T> R = ld4(r, v), Z = ld4(z, v);
        st4(p, v, Z);
#pragma unroll
        for (int e = 0; e < 4; ++e) q += R.e[e] * Z.e[e];
    }
    return q;
}

template <int NT, class T>
__device__ T pass_dot(int nv, const T* __restrict__ a, const T* __restrict__ b) {
    T q = T(0);
    for (int v = threadIdx.x; v < nv; v += 2 * NT) {
        const int v2 = v + NT;
        const V4<T> A = ld4(a, v), B = ld4(b, v);
        V4<T> A2{}, B2{};
        if (v2 < nv) {
            A2 = ld4(a, v2);
            B2 = ld4(b, v2);
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) q += A.e[e] * B.e[e];
        if (v2 < nv) {
#pragma unroll
            for (int e = 0; e < 4; ++e) q += A2.e[e] * B2.e[e];
        }
    }
    return q;
}

// FP64 CG (conjugate_gradient.rs:77-80): x += p*alpha; r -= ap*alpha; returns partial r.r
template <int NT>
__device__ double pass_cg_xr(int nv, const double* __restrict__ p, const double* __restrict__ ap,
                             double* __restrict__ x, double* __restrict__ r, double alpha) {
    double q = 0.0;
    for (int v = threadIdx.x; v < nv; v += NT) {
        const V4<double> P = ld4(p, v), AP = ld4(ap, v);
        V4<double> X = ld4(x, v), R = ld4(r, v);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            X.e[e] += P.e[e] * alpha;
            R.e[e] = R.e[e] - AP.e[e] * alpha;
            q += R.e[e] * R.e[e];
        }
        st4(x, v, X);
        st4(r, v, R);
    }
    return q;
}

// FP64 CG (:98-100): p = z + p*beta
template <int NT>
__device__ void pass_cg_p(int nv, const double* __restrict__ z, double* __restrict__ p, double beta) {
    for (int v = threadIdx.x; v < nv; v += NT) {
        const V4<double> Z = ld4(z, v);
        V4<double> P = ld4(p, v);
#pragma unroll
        for (int e = 0; e < 4; ++e) P.e[e] = Z.e[e] + P.e[e] * beta;
        st4(p, v, P);
    }
}

// FP64 BiCGStab (:213-216): x = x + y*alpha; s = r - v*alpha; returns partial s.s
template <int NT>
__device__ double pass_bi_hs(int nv, const double* __restrict__ y, const double* __restrict__ vv,
                             const double* __restrict__ r, double* __restrict__ x, double* __restrict__ s,
                             double alpha) {
    double q = 0.0;
    for (int v = threadIdx.x; v < nv; v += NT) {
        const V4<double> Y = ld4(y, v), VV = ld4(vv, v), R = ld4(r, v);
        V4<double> X = ld4(x, v), S;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            X.e[e] = X.e[e] + Y.e[e] * alpha;
            S.e[e] = R.e[e] - VV.e[e] * alpha;
            q += S.e[e] * S.e[e];
        }
        st4(x, v, X);
        st4(s, v, S);
    }
    return q;
}

// FP64 BiCGStab (:246-249): x = x + z*omega; r = s - t*omega; partials rh.r and r.r
template <int NT>
__device__ void pass_bi_xr(int nv, const double* __restrict__ z, const double* __restrict__ t,
                           const double* __restrict__ s, const double* __restrict__ rh, double* __restrict__ x,
                           double* __restrict__ r, double omega, double& q0, double& q1) {
    for (int v = threadIdx.x; v < nv; v += NT) {
        const V4<double> Z = ld4(z, v), TT = ld4(t, v), S = ld4(s, v), RH = ld4(rh, v);
        V4<double> X = ld4(x, v), R;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            X.e[e] = X.e[e] + Z.e[e] * omega;
            R.e[e] = S.e[e] - TT.e[e] * omega;
            q0 += RH.e[e] * R.e[e];
            q1 += R.e[e] * R.e[e];
        }
        st4(x, v, X);
        st4(r, v, R);
    }
}

// FP64 BiCGStab (:268-282): p = r + (p - v*omega)*beta, or the restart rh = r, p = r
template <int NT>
__device__ void pass_bi_p(int nv, const double* __restrict__ r, const double* __restrict__ vv, double* __restrict__ p,
                          double* __restrict__ rh, double omega, double beta, int mode) {
    for (int v = threadIdx.x; v < nv; v += NT) {
        const V4<double> R = ld4(r, v);
        if (mode == 0) {
            const V4<double> VV = ld4(vv, v);
            V4<double> P = ld4(p, v);
#pragma unroll
            for (int e = 0; e < 4; ++e) P.e[e] = R.e[e] + (P.e[e] - VV.e[e] * omega) * beta;
            st4(p, v, P);
        } else {
            st4(rh, v, R);
            st4(p, v, R);
        }
    }
}

// ---- per-system bookkeeping ----------------------------------------------------------------
__device__ __forceinline__ void team_push(const TeamDesc& D, int64_t& len, double e) {
    if (threadIdx.x == 0 && len < D.hcap) gm(D.hist)[len] = e;
    len += 1;
}

// gather b and x into positions; returns the partial |b|^2
template <int NT>
__device__ double team_gather(const TeamDesc& D, double* __restrict__ b, double* __restrict__ x) {
    double q = 0.0;
    const int32_t* __restrict__ perm = D.perm;
    for (int p0 = 0; p0 < D.n; p0 += 2 * NT) {
        int row[2];
        double bb[2], xx[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int p = p0 + u * NT + (int)threadIdx.x;
            row[u] = p < D.n ? gm(perm)[p] : 0;
        }
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            bb[u] = gm(D.b)[row[u]];
            xx[u] = gm(D.x)[row[u]];
        }
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int p = p0 + u * NT + (int)threadIdx.x;
            if (p < D.n) {
                gm(b)[p] = bb[u];
                gm(x)[p] = xx[u];
                q += bb[u] * bb[u];
            }
        }
    }
    return q;
}

template <int NT>
__device__ void team_scatter(const TeamDesc& D, const double* __restrict__ x) {
    for (int p = threadIdx.x; p < D.n; p += NT) gm(D.x)[gm(D.perm)[p]] = gm(x)[p];
}

__device__ void team_info(const TeamDesc& D, int64_t iters, double init, double fin, int64_t hlen, int64_t restarts,
                          int64_t outer, double true_res) {
    if (threadIdx.x != 0) return;
    fvg_solver_info I;
    I.iterations = iters;
    I.initial_residual = init;
    I.final_residual = fin;
    I.history_len = hlen;
    I.restarts = restarts;
    I.outer_iterations = outer;
    I.true_residual = true_res;
    I.status = FVG_OK;
    I.flags = 0;
    *D.info = I;
}

// |b| below 1e-14: x = 0 and one history entry (conjugate_gradient.rs:29-36)
template <int NT>
__device__ bool team_zero_rhs(const TeamDesc& D, double bnorm, bool mixed) {
    if (!(bnorm < 1e-14)) return false;
    for (int p = threadIdx.x; p < D.n; p += NT) gm(D.x)[gm(D.perm)[p]] = 0.0;
    int64_t len = 0;
    team_push(D, len, bnorm);
    team_info(D, 0, bnorm, bnorm, len, 0, 0, mixed ? bnorm : 0.0);
    return true;
}

// ---- FP64 conjugate_gradient (conjugate_gradient.rs:12-112) ------------------------------
template <int NT>
__device__ void team_cg64(const TeamDesc& D, Reducer<NT, double>& R, double* ring) {
    const int np = D.npad, nv = D.npad / 4;
    double* b = D.w64;
    double* x = b + np;
    double* r = x + np;
    double* z = r + np;
    double* p = z + np;
    double* ap = p + np;   // also the intermediate vector of the two sweeps
    const double bnorm = sqrt(R.sum(team_gather<NT>(D, b, x)));
    if (team_zero_rhs<NT>(D, bnorm, false)) return;
    __syncthreads();
    // r = (A x - b) * -1 and |r| (:47-52)
    double q = 0.0;
    team_spmv<NT, double, kRing64>(D, D.a_val, x, b, ring, [&](int i, double s, double bi) {
        const double ri = (s - bi) * -1.0;
        gm(r)[i] = ri;
        q += ri * ri;
    });
    const double init_res = sqrt(R.sum(q)) / bnorm;
    double error = 1.0;
    int64_t iter = 0, hlen = 0;
    if ((error > D.atol) && !(iter > D.max_iter)) {
        team_prec<NT, double, kRing64>(D, r, z, ap, ring);
        double rz = R.sum(pass_start<NT, double>(nv, r, z, p));
        for (;;) {
            q = 0.0;   // ap = A p, p.ap (:72-74)
            team_spmv<NT, double, kRing64>(D, D.a_val, p, p, ring, [&](int i, double s, double pi) {
                gm(ap)[i] = s;
                q += pi * s;
            });
            const double alpha = rz / R.sum(q);
            error = sqrt(R.sum(pass_cg_xr<NT>(nv, p, ap, x, r, alpha))) / bnorm;   // (:77-89)
            if ((error < D.atol) || (error / init_res < D.rtol)) {
                team_push(D, hlen, error);
                iter += 1;
                break;
            }
            team_prec<NT, double, kRing64>(D, r, z, ap, ring);
            const double rz_new = R.sum(pass_dot<NT, double>(nv, r, z));
            const double beta = rz_new / rz;   // (:94-105)
            rz = rz_new;
            team_push(D, hlen, error);
            iter += 1;
            if (!(error > D.atol) || iter > D.max_iter) break;
            pass_cg_p<NT>(nv, z, p, beta);
            __syncthreads();
        }
    }
    __syncthreads();
    team_scatter<NT>(D, x);
    team_info(D, iter, init_res, error, hlen, 0, 0, 0.0);
}

// ---- FP64 bi_conjugate_gradient_stab (bi_conjugate_gradient_stab.rs:13-179) --------------
template <int NT>
__device__ void team_bi64(const TeamDesc& D, Reducer<NT, double>& R, double* ring) {
    const int n = D.n, np = D.npad, nv = D.npad / 4;
    double* b = D.w64;
    double* x = b + np;
    double* r = x + np;
    double* rh = r + np;
    double* p = rh + np;
    double* y = p + np;
    double* v = y + np;
    double* s = v + np;
    double* z = s + np;
    double* t = z + np;   // also the intermediate vector of the sweeps
    const double bnorm = sqrt(R.sum(team_gather<NT>(D, b, x)));
    if (team_zero_rhs<NT>(D, bnorm, false)) return;
    __syncthreads();
    double q = 0.0;   // r = b - A x, rh = p = r (:165-190)
    team_spmv<NT, double, kRing64>(D, D.a_val, x, b, ring, [&](int i, double sv, double bi) {
        const double ri = (sv - bi) * -1.0;
        gm(r)[i] = ri;
        gm(rh)[i] = ri;
        gm(p)[i] = ri;
        q += ri * ri;
    });
    const double rr0 = R.sum(q);
    const double init_res = sqrt(rr0) / bnorm;
    double rho = rr0, error = 1.0, alpha = 0.0, omega = 0.0;
    int64_t iter = 0, hlen = 0, restarts = 0;
    bool run = true;
    if (init_res < D.atol) {   // (:176-178)
        error = init_res;
        team_push(D, hlen, init_res);
        run = false;
    } else if (!(error > D.atol) || iter > D.max_iter) {
        run = false;
    }
    (void)n;
    while (run) {
        team_prec<NT, double, kRing64>(D, p, y, t, ring);   // y = M^-1 p
        q = 0.0;                                                 // v = A y, rh.v (:208-209)
        team_spmv<NT, double, kRing64>(D, D.a_val, y, rh, ring, [&](int i, double sv, double rhi) {
            gm(v)[i] = sv;
            q += rhi * sv;
        });
        alpha = rho / R.sum(q);
        const double ss = sqrt(R.sum(pass_bi_hs<NT>(nv, y, v, r, x, s, alpha))) / bnorm;   // (:213-231)
        if ((ss < D.atol) || (ss / init_res < D.rtol)) {
            error = ss;
            team_push(D, hlen, ss);
            iter += 1;
            break;
        }
        team_prec<NT, double, kRing64>(D, s, z, t, ring);   // z = M^-1 s
        double q0 = 0.0, q1 = 0.0;                               // t = A z; t.s, t.t (:239-242)
        team_spmv<NT, double, kRing64>(D, D.a_val, z, s, ring, [&](int i, double tv, double si) {
            gm(t)[i] = tv;
            q0 += si * tv;
            q1 += tv * tv;
        });
        R.sum2(q0, q1);
        omega = q0 / q1;
        q0 = 0.0;
        q1 = 0.0;
        pass_bi_xr<NT>(nv, z, t, s, rh, x, r, omega, q0, q1);   // (:246-249)
        R.sum2(q0, q1);
        const double rr = sqrt(q1) / bnorm;   // (:252-287)
        if ((rr < D.atol) || (rr / init_res < D.rtol)) {
            error = rr;
            team_push(D, hlen, rr);
            iter += 1;
            break;
        }
        const double beta = (q0 / rho) * (alpha / omega);
        rho = q0;
        int mode = 0;
        if (!(fabs(rho) > (1e-15 * bnorm))) {
            mode = 1;
            restarts += 1;
            rho = q1;
        }
        error = rr / bnorm;
        team_push(D, hlen, error);
        iter += 1;
        if (!(error > D.atol) || iter > D.max_iter) break;
        pass_bi_p<NT>(nv, r, v, p, rh, omega, beta, mode);
        __syncthreads();
    }
    __syncthreads();
    team_scatter<NT>(D, x);
    team_info(D, iter, init_res, error, hlen, restarts, 0, 0.0);
}

// ---- mixed precision: FP32 inner solves inside FP64 refinement ---------------------------
// Reduction asked of the inner solve of a refinement step: the remaining factor to the
// caller's tolerance, with a margin, not below the floor `lo` and not above 0.1.
__device__ __forceinline__ double inner_target(const TeamDesc& D, double e, double init_res) {
    const double goal = fmax(D.rtol * init_res, D.atol);
    double t = 0.3 * goal / e;
    t = fmin(t, 0.1);
    return fmax(t, D.inner_tol);
}

template <int NT>
__device__ void team_mixed(const TeamDesc& D, Reducer<NT, double>& R, Reducer<NT, float>& F, bool bicg,
                           float* ring) {
    const int tid = threadIdx.x;
    const int n = D.n, np = D.npad, nv = D.npad / 4;
    double* b = D.w64;
    double* x = b + np;
    double* xa = x + np;  // the last accepted solution
    float* r = D.w32;
    float* d = r + np;
    float* z = d + np;
    float* p = z + np;
    float* ap = p + np;   // CG: A p and the sweeps' intermediate
    float* rh = ap + np;  // BiCGStab only
    float* y = rh + np;
    float* v = y + np;
    float* s = v + np;
    float* t = s + np;    // BiCGStab: A z and the sweeps' intermediate
    const double bnorm = sqrt(R.sum(team_gather<NT>(D, b, x)));
    if (team_zero_rhs<NT>(D, bnorm, true)) return;
    __syncthreads();
    const int64_t max_inner = D.max_iter > 0 ? D.max_iter : 1;
    double init_res = 0.0, true_res = 0.0, e_acc = 0.0;
    int64_t total = 0, outer = 0, restarts = 0, hlen = 0;
    int64_t cap = 400;     // inner iterations per refinement step; shrinks after a rejected step
    bool check = false;    // the current x carries a correction not yet accepted
    int rejected = 0;      // corrections taken back in a row
    int weak = 0;          // accepted corrections in a row that did not halve the residual
    bool fallback = false;
    (void)n;
    for (int k = 0; k < 64; ++k) {
        // r64 = b - A x in FP64 (values, products and sum), r32 = (float) r64, d = 0
        double q = 0.0;
        team_spmv<NT, double, kRing64>(D, D.a_val, x, b, (double*)nullptr, [&](int i, double sv, double bi) {
            const double ri = (sv - bi) * -1.0;
            gm(r)[i] = (float)ri;
            gm(d)[i] = 0.f;
            q += ri * ri;
        });
        const double e = sqrt(R.sum(q)) / bnorm;
        if (outer == 0) init_res = e;
        outer += 1;
        // refinement keeps only corrections that lower the true residual: an FP32 inner
        // solve that stagnated or diverged is taken back and retried with a shorter budget
        if (check && !(e < e_acc)) {
            for (int i = tid; i < np; i += NT) gm(x)[i] = gm(xa)[i];
            __syncthreads();
            check = false;
            cap = cap / 4 > 8 ? cap / 4 : 8;
            if (++rejected >= 2) {   // FP32 iterations make no progress on this system: go on in FP64
                fallback = true;
                break;
            }
            if (total >= max_inner) break;
            continue;
        }
        if (check) rejected = 0;   // a correction was accepted (not just the restored x re-measured)
        if (check && e > 0.5 * e_acc) {   // accepted, but the refinement is barely progressing
            if (++weak >= 3) {
                e_acc = e;
                true_res = e;
                team_push(D, hlen, e);
                check = false;   // keep this x: it is the best so far
                fallback = true;
                break;
            }
        } else {
            weak = 0;
        }
        e_acc = e;
        true_res = e;
        check = false;
        team_push(D, hlen, e);
        if ((e < D.atol) || (e / init_res < D.rtol) || total >= max_inner) break;
        for (int i = tid; i < np; i += NT) gm(xa)[i] = gm(x)[i];
        const int64_t inner_end = total + cap;
        const double tol = inner_target(D, e, init_res);
        if (!bicg) {
            team_prec<NT, float, kRing32>(D, r, z, ap, ring);   // z = M r; p = z; |r|, r.z
            float q0 = 0.f, q1 = 0.f;
            for (int vi = threadIdx.x; vi < nv; vi += NT) {
                const V4<float> RR = ld4(r, vi), Z = ld4(z, vi);
                st4(p, vi, Z);
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    q0 += RR.e[j] * RR.e[j];
                    q1 += RR.e[j] * Z.e[j];
                }
            }
            F.sum2(q0, q1);
            const double rin0 = sqrt((double)q0);
            double rz = (double)q1;
            if (!(rin0 > 0.0)) break;   // the FP32 residual underflowed: refinement cannot progress
            for (;;) {
                float qa = 0.f;
                team_spmv<NT, float, kRing32>(D, D.a_val32, p, p, ring, [&](int i, float sv, float pi) {
                    gm(ap)[i] = sv;
                    qa += pi * sv;
                });
                const double alpha_d = rz / (double)F.sum(qa);
                if (!isfinite(alpha_d)) break;   // FP32 breakdown: refine from the correction so far
                const float alpha = (float)alpha_d;
                float qr = 0.f;
                for (int vi = threadIdx.x; vi < nv; vi += NT) {
                    const V4<float> P = ld4(p, vi), AP = ld4(ap, vi);
                    V4<float> DD = ld4(d, vi), RR = ld4(r, vi);
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        DD.e[j] += alpha * P.e[j];
                        RR.e[j] = RR.e[j] - alpha * AP.e[j];
                        qr += RR.e[j] * RR.e[j];
                    }
                    st4(d, vi, DD);
                    st4(r, vi, RR);
                }
                const double rin = sqrt((double)F.sum(qr));
                total += 1;
                if (rin <= tol * rin0 || total >= max_inner || total >= inner_end) break;
                if (!(rin < 1e2 * rin0)) break;   // diverging: let the refinement judge this correction
                team_prec<NT, float, kRing32>(D, r, z, ap, ring);
                const double rz_new = (double)F.sum(pass_dot<NT, float>(nv, r, z));
                const float beta = (float)(rz_new / rz);
                rz = rz_new;
                for (int vi = threadIdx.x; vi < nv; vi += NT) {
                    const V4<float> Z = ld4(z, vi);
                    V4<float> P = ld4(p, vi);
#pragma unroll
                    for (int j = 0; j < 4; ++j) P.e[j] = Z.e[j] + beta * P.e[j];
                    st4(p, vi, P);
                }
                __syncthreads();
            }
        } else {
            float q0 = 0.f;
            for (int vi = threadIdx.x; vi < nv; vi += NT) {
                const V4<float> RR = ld4(r, vi);
                st4(rh, vi, RR);
                st4(p, vi, RR);
#pragma unroll
                for (int j = 0; j < 4; ++j) q0 += RR.e[j] * RR.e[j];
            }
            q0 = F.sum(q0);
            const double rin0 = sqrt((double)q0);
            double rho = (double)q0, alpha = 0.0, omega = 0.0;
            if (!(rin0 > 0.0)) break;
            for (;;) {
                team_prec<NT, float, kRing32>(D, p, y, t, ring);
                float qv = 0.f;
                team_spmv<NT, float, kRing32>(D, D.a_val32, y, rh, ring, [&](int i, float sv, float rhi) {
                    gm(v)[i] = sv;
                    qv += rhi * sv;
                });
                alpha = rho / (double)F.sum(qv);
                if (!isfinite(alpha)) break;     // FP32 breakdown: refine from the correction so far
                const float af = (float)alpha;
                float qs = 0.f;
                for (int vi = threadIdx.x; vi < nv; vi += NT) {
                    const V4<float> Y = ld4(y, vi), VV = ld4(v, vi), RR = ld4(r, vi);
                    V4<float> DD = ld4(d, vi), S;
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        DD.e[j] += af * Y.e[j];
                        S.e[j] = RR.e[j] - af * VV.e[j];
                        qs += S.e[j] * S.e[j];
                    }
                    st4(d, vi, DD);
                    st4(s, vi, S);
                }
                const double sn = sqrt((double)F.sum(qs));
                if (sn <= tol * rin0) {
                    total += 1;
                    break;
                }
                team_prec<NT, float, kRing32>(D, s, z, t, ring);
                float qts = 0.f, qtt = 0.f;
                team_spmv<NT, float, kRing32>(D, D.a_val32, z, s, ring, [&](int i, float tv, float si) {
                    gm(t)[i] = tv;
                    qts += tv * si;
                    qtt += tv * tv;
                });
                F.sum2(qts, qtt);
                omega = (double)qts / (double)qtt;
                if (!isfinite(omega) || omega == 0.0) {   // breakdown after h = d + alpha y: keep h
                    total += 1;
                    break;
                }
                const float of = (float)omega;
                float qhr = 0.f, qrr = 0.f;
                for (int vi = threadIdx.x; vi < nv; vi += NT) {
                    const V4<float> Z = ld4(z, vi), TT = ld4(t, vi), S = ld4(s, vi), RH = ld4(rh, vi);
                    V4<float> DD = ld4(d, vi), RR;
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        DD.e[j] += of * Z.e[j];
                        RR.e[j] = S.e[j] - of * TT.e[j];
                        qhr += RH.e[j] * RR.e[j];
                        qrr += RR.e[j] * RR.e[j];
                    }
                    st4(d, vi, DD);
                    st4(r, vi, RR);
                }
                F.sum2(qhr, qrr);
                const double rin = sqrt((double)qrr);
                total += 1;
                if (rin <= tol * rin0 || total >= max_inner || total >= inner_end) break;
                if (!(rin < 1e2 * rin0)) break;   // diverging: let the refinement judge this correction
                const double rho_new = (double)qhr;
                const double beta = (rho_new / rho) * (alpha / omega);
                rho = rho_new;
                int mode = 0;
                if (!(fabs(rho) > 1e-30 * rin0 * rin0)) {
                    mode = 1;
                    restarts += 1;
                    rho = (double)qrr;
                }
                const float bf = (float)beta;
                for (int vi = threadIdx.x; vi < nv; vi += NT) {
                    const V4<float> RR = ld4(r, vi);
                    if (mode == 0) {
                        const V4<float> VV = ld4(v, vi);
                        V4<float> P = ld4(p, vi);
#pragma unroll
                        for (int j = 0; j < 4; ++j) P.e[j] = RR.e[j] + bf * (P.e[j] - of * VV.e[j]);
                        st4(p, vi, P);
                    } else {
                        st4(rh, vi, RR);
                        st4(p, vi, RR);
                    }
                }
                __syncthreads();
            }
        }
        for (int vi = threadIdx.x; vi < np / 2; vi += NT) {   // FP64 solution update x += d
            const int i = 2 * vi;
            gm(x)[i] += (double)gm(d)[i];
            gm(x)[i + 1] += (double)gm(d)[i + 1];
        }
        check = true;
        __syncthreads();
    }
    if (check) {   // the loop ended on an unjudged correction (outer-step bound): keep the accepted x
        for (int i = tid; i < np; i += NT) gm(x)[i] = gm(xa)[i];
        __syncthreads();
    }
    if (fallback) {
        // the reference FP64 iterations from the accepted solution (no LDS ring: this kernel's
        // ring holds FP32 values); its info replaces this one, with the refinement steps kept
        team_scatter<NT>(D, x);
        __syncthreads();
        if (bicg)
            team_bi64<NT>(D, R, (double*)nullptr);
        else
            team_cg64<NT>(D, R, (double*)nullptr);
        __syncthreads();
        if (threadIdx.x == 0) {
            D.info->iterations += total;
            D.info->outer_iterations = outer;
            D.info->true_residual = D.info->final_residual;
            D.info->flags |= FVG_INFO_FP64_FALLBACK;
        }
        return;
    }
    __syncthreads();
    team_scatter<NT>(D, x);
    team_info(D, total, init_res, true_res, hlen, restarts, outer, true_res);
}

template <int NT, int VAR>
__global__ void __launch_bounds__(NT) k_team(const TeamDesc* __restrict__ descs, const int32_t* __restrict__ count) {
    const int n_inst = *count;   // systems of this launch (the grid is the batch's capacity)
    using T = typename std::conditional<VAR == kCG32 || VAR == kBI32, float, double>::type;
    constexpr int RING = sizeof(T) == 8 ? kRing64 : kRing32;
    __shared__ T ring[RING];
    __shared__ double red64[4 * (NT / 64)];
    __shared__ float red32[4 * (NT / 64)];
    __shared__ TeamDesc sdesc;
    static_assert(sizeof(TeamDesc) % 8 == 0 && sizeof(TeamDesc) / 8 <= NT, "descriptor copy");
    Reducer<NT, double> R{red64, 0};
    Reducer<NT, float> F{red32, 0};
    for (int k = blockIdx.x; k < n_inst; k += gridDim.x) {
        // the descriptor lives in LDS for the whole solve: reading it never waits on the
        // vector-memory counter that the prefetched loads are counted by
        __syncthreads();
        if (threadIdx.x < sizeof(TeamDesc) / 8)
            reinterpret_cast<uint64_t*>(&sdesc)[threadIdx.x] = reinterpret_cast<const uint64_t*>(descs + k)[threadIdx.x];
        __syncthreads();
        const TeamDesc& D = sdesc;
        long long t0 = 0;
        if (threadIdx.x == 0) {
            g_prof_on = D.prof != nullptr;
            for (int j = 0; j < 4; ++j) g_prof[j] = 0;
            t0 = wall_clock64();
        }
        // a factorisation that failed its checks (made without waiting for them) is reported,
        // not used: IC of an unsymmetric matrix, ILU with a zero pivot
        if (threadIdx.x == 0) {
            int st = FVG_OK;
            if (D.pflags && D.kind == kIC && D.pflags[2]) st = FVG_E_NOT_SYMMETRIC;
            if (D.pflags && D.kind == kILU && D.pflags[0]) st = FVG_E_ZERO_PIVOT;
            g_status = st;
        }
        __syncthreads();
        if (g_status != FVG_OK) {
            if (threadIdx.x == 0) {
                fvg_solver_info I{};
                I.status = g_status;
                *D.info = I;
            }
        } else {
            if constexpr (VAR == kCG64) team_cg64<NT>(D, R, ring);
            if constexpr (VAR == kBI64) team_bi64<NT>(D, R, ring);
            if constexpr (VAR == kCG32) team_mixed<NT>(D, R, F, false, ring);
            if constexpr (VAR == kBI32) team_mixed<NT>(D, R, F, true, ring);
        }
        if (threadIdx.x == 0 && g_prof_on) {
            g_prof[3] = wall_clock64() - t0;
            for (int j = 0; j < 4; ++j) D.prof[j] = g_prof[j];
        }
        // every wave's writes (solution, info) reach memory before the completion flag
        __threadfence_system();
        __syncthreads();
        if (threadIdx.x == 0) __hip_atomic_store(D.done, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// out64[k] = in[src[k]], out32[k] = (float) in[src[k]]
__global__ void k_team_gather(int64_t n, const int32_t* __restrict__ src, const double* __restrict__ in,
                              double* __restrict__ out64, float* __restrict__ out32) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double v = in[src[k]];
    out64[k] = v;
    out32[k] = (float)v;
}

#include "cluster.hpp"

// ---- host side ---------------------------------------------------------------------------

template <int NT>
static const void* team_kernel(int var) {
    switch (var) {
    case kCG64: return (const void*)k_team<NT, kCG64>;
    case kBI64: return (const void*)k_team<NT, kBI64>;
    case kCG32: return (const void*)k_team<NT, kCG32>;
    default: return (const void*)k_team<NT, kBI32>;
    }
}

static const void* team_kernel_for(int nt, int var) {
    switch (nt) {
    case 256: return team_kernel<256>(var);
    case 512: return team_kernel<512>(var);
    default: return team_kernel<1024>(var);
    }
}

}  // namespace fvg

// one sweep's tables, as offsets into its plan's index array
struct TeamSweepLayout {
    int64_t lvl = 0, ptr = 0, dep = 0, src = 0;
    int64_t nnz = 0, val = 0;
    int32_t ring64 = 0, ring32 = 0;
};

// arrays of one system, as offsets into its plan's arrays
struct TeamLayout {
    int64_t n = 0, npad = 0, nnz_a = 0;
    int32_t wspan = 0, win64 = 0, win32 = 0;
    int64_t perm = 0, a_ptr = 0, a_dep = 0, a_src = 0;   // int arena
    TeamSweepLayout lo, up;
    int64_t a_val = 0, sc = 0;                           // value arrays
};

// Per-system plan, cached on the preconditioner: the analysis of the matrix pattern,
// the gathered values and the work vectors.  A batch only points at plans, so building
// a batch over already-planned systems costs no analysis and no large allocation.
struct TeamPlan {
    const fvg_csr* a = nullptr;   // matrix the values are gathered from
    TeamLayout L;                 // offsets into this plan's arrays
    int32_t* ints = nullptr;
    int64_t ni = 0;
    double* v64 = nullptr;
    float* v32 = nullptr;
    int64_t nv = 0;
    double* w64 = nullptr;        // work vectors: 13 of npad doubles (b, x, x accepted and the ten
                                  // inner vectors of a multi-workgroup FP64 fallback)
    float* w32 = nullptr;         // 10 of npad floats
    int64_t nw = 0, nw32 = 0;
    uint64_t a_ver = ~0ull, p_ver = ~0ull;
    bool busy = false;            // part of a launched batch
};

struct fvg_batch {
    int n_inst = 0;
    int capacity = 0;                 // systems the allocations and the captured launch hold
    int solver = 0;      // 0 CG, 1 BiCGStab
    int precision = 0;
    int variant = 0;
    int nt = 512;
    fvg_solver_options opts{};
    std::vector<const fvg_csr*> A;
    std::vector<fvg_prec*> P;
    std::vector<TeamPlan*> plan;
    std::vector<int> order;           // launch slot -> system (largest first)
    double* hist = nullptr;
    int64_t nhist = 0, hcap = 0;
    fvg_solver_info* h_info = nullptr;  // pinned host memory the kernel writes each system's info to
    int32_t* h_done = nullptr;          // pinned host completion flags
    int64_t* d_prof = nullptr;        // 4 counters per system when phase timing is on
    fvg::TeamDesc* d_desc = nullptr;  // capacity descriptors followed by the system count
    int32_t* d_count = nullptr;
    std::vector<fvg::TeamDesc> desc;  // by launch slot
    fvg::TeamDesc* h_desc = nullptr;  // pinned copy the uploads read from (capacity + 1: the count)
    bool desc_dirty = true;
    hipStream_t stream = nullptr;
    hipEvent_t ready = nullptr;       // recorded on the context stream before each launch
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int graph_nt = 0, graph_cap = 0;
    bool launched = false;
    fvg_ctx* flight_ctx = nullptr;    // the context whose stream a launch followed (counts it in flight)
    // multi-workgroup solves of mixed-precision batches
    int cl_rows = 4096;                    // rows per workgroup (0: one workgroup per system)
    // per size bucket: systems of at most cl_limit[i] rows (ascending) get cl_brows[i] rows
    // per workgroup; larger ones cl_rows (fvg_batch_set_cluster_buckets)
    std::vector<int64_t> cl_limit;
    std::vector<int32_t> cl_brows;
    int cl_workers = 0;                    // worker pool (0: as many as can be co-resident)
    int cl_resident = 1;                   // workgroups of the multi-workgroup kernel that fit the device
    std::vector<int32_t> cl_groups;        // workgroups of each launch slot
    std::vector<int32_t> cl_tickets;       // (slot, rank) pairs, largest systems first
    int32_t* h_cl = nullptr;               // pinned staging: groups, then tickets
    int32_t* d_cl_groups = nullptr;
    int2* d_cl_tickets = nullptr;
    fvg::ClusterHead* d_cl_head = nullptr;
    fvg::ClusterCtl* d_cl_ctl = nullptr;
    bool cl_dirty = true;
    int graph_kind = -1, graph_workers = 0, graph_sweeps = 0;
    bool cluster() const { return precision == FVG_PRECISION_MIXED && cl_rows > 0; }
};

namespace fvg {

static void batch_free(fvg_ctx* c, fvg_batch* B) {
    if (B->stream) (void)hipStreamSynchronize(B->stream);
    if (B->launched) {
        for (TeamPlan* t : B->plan) t->busy = false;
        B->launched = false;
        if (B->flight_ctx) B->flight_ctx->batches_in_flight -= 1;
        B->flight_ctx = nullptr;
    }
    if (B->exec) (void)hipGraphExecDestroy(B->exec);
    if (B->graph) (void)hipGraphDestroy(B->graph);
    dev_free_n(c, B->hist, B->nhist);
    if (B->h_info) (void)hipHostFree(B->h_info);
    if (B->h_done) (void)hipHostFree(B->h_done);
    if (B->h_desc) (void)hipHostFree(B->h_desc);
    B->h_info = nullptr;
    B->h_done = nullptr;
    B->h_desc = nullptr;
    dev_free_n(c, B->d_prof, 4 * B->capacity);
    dev_free_n(c, B->d_desc, B->capacity);
    dev_free_n(c, B->d_count, 1);
    if (B->h_cl) (void)hipHostFree(B->h_cl);
    B->h_cl = nullptr;
    dev_free_n(c, B->d_cl_groups, B->capacity);
    dev_free_n(c, B->d_cl_tickets, (int64_t)B->capacity * kClMaxGroups);
    dev_free_n(c, B->d_cl_ctl, B->capacity);
    dev_free_n(c, B->d_cl_head, 1);
    if (B->ready) (void)hipEventDestroy(B->ready);
    if (B->stream) (void)hipStreamDestroy(B->stream);
}

static void plan_free(fvg_ctx* c, TeamPlan* t) {
    dev_free_n(c, t->ints, t->ni);
    dev_free_n(c, t->v64, t->nv);
    dev_free_n(c, t->v32, t->nv);
    dev_free_n(c, t->w64, t->nw);
    dev_free_n(c, t->w32, t->nw32);
}

void team_plan_release(fvg_ctx* c, fvg_prec* p) {
    if (!p->team) return;
    auto* t = static_cast<TeamPlan*>(p->team);
    if (t->busy) (void)hipDeviceSynchronize();   // destroyed while a batch using it runs: let it finish
    plan_free(c, t);
    delete t;
    p->team = nullptr;
}

static int64_t put_ints(std::vector<int32_t>& ints, const std::vector<int32_t>& v) {
    const int64_t off = (int64_t)ints.size();
    ints.insert(ints.end(), v.begin(), v.end());
    return off;
}

// Ring check of one sweep: rows are solved in groups of up to kMaxGroup processing
// indices, and a row reads a dependency from the ring slot of its position, so no row
// written up to the end of the reader's group may share that slot: every dependency
// must lie less than the ring size before the end of its reader's group.
static void ring_check(int64_t n, const std::vector<int32_t>& ptr, const std::vector<int32_t>& qdep, int32_t& ring64,
                       int32_t& ring32) {
    int64_t span = 0;
    for (int64_t q = 0; q < n; ++q) {
        const int64_t qe = std::min<int64_t>((q / kMaxGroup + 1) * kMaxGroup, n);
        for (int32_t k = ptr[(size_t)q]; k < ptr[(size_t)q + 1]; ++k)
            span = std::max<int64_t>(span, qe - 1 - qdep[(size_t)k]);
    }
    ring64 = span < kRing64 ? 1 : 0;
    ring32 = span < kRing32 ? 1 : 0;
}

// Host analysis of one system: level order, permuted A, lower and upper factor
// entries with their source slots, sweep plans.  Appends index data to `ints`.
static int team_analyse(const fvg_csr* a, const fvg_prec* p, TeamLayout& L, std::vector<int32_t>& ints) {
    const int64_t n = a->nrows;
    const auto& rp = a->h_row_ptr;
    const auto& col = a->h_col;
    const bool fac = p->kind == kIC || p->kind == kILU;
    if (fac) {   // structural symmetry (the upper sweep runs the lower levels in reverse)
        for (int64_t i = 0; i < n; ++i)
            for (int32_t k = rp[(size_t)i]; k < rp[(size_t)i + 1]; ++k) {
                const int32_t j = col[(size_t)k];
                auto b = col.begin() + rp[(size_t)j], e = col.begin() + rp[(size_t)j + 1];
                if (!std::binary_search(b, e, (int32_t)i))
                    return fail(FVG_E_UNSUPPORTED, "batched solves need a structurally symmetric matrix pattern");
            }
    }
    std::vector<int32_t> lev((size_t)n, 0);
    int32_t maxlev = 0;
    if (fac) {
        for (int64_t i = 0; i < n; ++i) {
            int32_t l = 0;
            for (int32_t k = rp[(size_t)i]; k < rp[(size_t)i + 1]; ++k) {
                const int32_t j = col[(size_t)k];
                if (j < i) l = std::max(l, lev[(size_t)j] + 1);
            }
            lev[(size_t)i] = l;
            maxlev = std::max(maxlev, l);
        }
    }
    const int32_t nlev = maxlev + 1;
    std::vector<int32_t> lvl((size_t)nlev + 1, 0);
    for (int64_t i = 0; i < n; ++i) lvl[(size_t)lev[(size_t)i] + 1]++;
    for (int32_t l = 0; l < nlev; ++l) lvl[(size_t)l + 1] += lvl[(size_t)l];
    std::vector<int32_t> perm((size_t)n), pos((size_t)n);
    {
        std::vector<int32_t> cur(lvl.begin(), lvl.end() - 1);
        for (int64_t i = 0; i < n; ++i) {
            const int32_t q = cur[(size_t)lev[(size_t)i]]++;
            perm[(size_t)q] = (int32_t)i;
            pos[(size_t)i] = q;
        }
    }
    std::vector<int32_t> a_ptr((size_t)n + 1, 0), a_dep, a_src;
    std::vector<int32_t> l_ptr((size_t)n + 1, 0), l_dep, l_src, l_qdep;
    std::vector<int32_t> u_ptr((size_t)n + 1, 0), u_dep, u_src, u_qdep;
    a_dep.reserve(col.size());
    a_src.reserve(col.size());
    for (int64_t q = 0; q < n; ++q) {
        const int32_t i = perm[(size_t)q];
        for (int32_t k = rp[(size_t)i]; k < rp[(size_t)i + 1]; ++k) {
            const int32_t j = col[(size_t)k];
            a_dep.push_back(pos[(size_t)j]);
            a_src.push_back(k);
            if (fac && j < i) {
                l_dep.push_back(pos[(size_t)j]);
                l_qdep.push_back(pos[(size_t)j]);
                l_src.push_back(k);
            }
        }
        a_ptr[(size_t)q + 1] = (int32_t)a_dep.size();
        l_ptr[(size_t)q + 1] = (int32_t)l_dep.size();
    }
    if (fac) {
        for (int64_t q = 0; q < n; ++q) {   // processing index q -> position n-1-q
            const int32_t i = perm[(size_t)(n - 1 - q)];
            for (int32_t k = rp[(size_t)i]; k < rp[(size_t)i + 1]; ++k) {
                const int32_t j = col[(size_t)k];
                if (j > i) {
                    u_dep.push_back(pos[(size_t)j]);
                    u_qdep.push_back((int32_t)(n - 1 - pos[(size_t)j]));
                    u_src.push_back(p->kind == kIC ? p->h_tpos[(size_t)k] : k);
                }
            }
            u_ptr[(size_t)q + 1] = (int32_t)u_dep.size();
        }
    }
    L.n = n;
    L.npad = (n + kMaxGroup - 1) / kMaxGroup * kMaxGroup;   // whole tiles: rows past n compute zeros
    L.nnz_a = (int64_t)a_dep.size();
    L.perm = put_ints(ints, perm);
    L.a_ptr = put_ints(ints, a_ptr);
    L.a_dep = put_ints(ints, a_dep);
    L.a_src = put_ints(ints, a_src);
    // SpMV window: largest distance between a row's position and its columns' positions
    int64_t wspan = 0;
    for (int64_t q = 0; q < n; ++q)
        for (int32_t k = a_ptr[(size_t)q]; k < a_ptr[(size_t)q + 1]; ++k)
            wspan = std::max<int64_t>(wspan, std::abs((int64_t)a_dep[(size_t)k] - q));
    L.wspan = (int32_t)std::min<int64_t>(wspan, INT32_MAX / 4);
    L.win64 = 2 * kMaxGroup + 2 * wspan <= kRing64 ? 1 : 0;
    L.win32 = 2 * kMaxGroup + 2 * wspan <= kRing32 ? 1 : 0;
    if (fac) {
        std::vector<int32_t> llev((size_t)n), ulev((size_t)n);   // level of each processing index
        for (int64_t q = 0; q < n; ++q) {
            llev[(size_t)q] = lev[(size_t)perm[(size_t)q]];
            ulev[(size_t)q] = nlev - 1 - lev[(size_t)perm[(size_t)(n - 1 - q)]];
        }
        for (int side = 0; side < 2; ++side) {
            TeamSweepLayout& S = side == 0 ? L.lo : L.up;
            const auto& sptr = side == 0 ? l_ptr : u_ptr;
            ring_check(n, sptr, side == 0 ? l_qdep : u_qdep, S.ring64, S.ring32);
            S.lvl = put_ints(ints, side == 0 ? llev : ulev);
            S.ptr = put_ints(ints, sptr);
            std::vector<int32_t> dep = side == 0 ? l_dep : u_dep;
            if (dep.empty()) dep.push_back(0);   // the sweep reads at least one (unused) entry
            S.dep = put_ints(ints, dep);
            S.src = put_ints(ints, side == 0 ? l_src : u_src);
            S.nnz = side == 0 ? (int64_t)l_dep.size() : (int64_t)u_dep.size();
        }
    }
    return FVG_OK;
}

// copy matrix and factor values into the plan's permuted arrays (FP64 and FP32)
static int plan_refresh(fvg_ctx* c, TeamPlan* t, const fvg_prec* p, hipStream_t s) {
    const TeamLayout& L = t->L;
    const fvg_csr* a = t->a;
    auto gather = [&](int64_t cnt, int64_t src_off, const double* in, int64_t dst) {
        if (cnt <= 0 || !in) return;
        hipLaunchKernelGGL(k_team_gather, dim3(grid_for(cnt)), dim3(kBlock), 0, s, cnt, t->ints + src_off, in,
                           t->v64 + dst, t->v32 + dst);
    };
    gather(L.nnz_a, L.a_src, a->val, L.a_val);
    if (p->kind == kIC || p->kind == kILU) {
        gather(L.lo.nnz, L.lo.src, p->d_fac, L.lo.val);
        gather(L.up.nnz, L.up.src, p->d_fac, L.up.val);
    }
    if (p->kind != kIdentity) gather(L.n, L.perm, p->d_inv_diag, L.sc);
    FVG_HIP(hipGetLastError());
    t->a_ver = a->version;
    t->p_ver = p->fac_version;
    return FVG_OK;
}

// the plan of system (a, p): made on first use, remade when p is used with another matrix
static int get_plan(fvg_ctx* c, const fvg_csr* a, fvg_prec* p, TeamPlan** out) {
    if (p->team && static_cast<TeamPlan*>(p->team)->a == a) {
        *out = static_cast<TeamPlan*>(p->team);
        return FVG_OK;
    }
    if (p->team && static_cast<TeamPlan*>(p->team)->busy)
        return fail(FVG_E_INVALID, "preconditioner is part of a batch in flight");
    team_plan_release(c, p);
    auto* t = new TeamPlan();
    t->a = a;
    std::vector<int32_t> ints;
    int s = team_analyse(a, p, t->L, ints);
    TeamLayout& L = t->L;
    if (s == FVG_OK) {
        L.a_val = 0;
        L.lo.val = L.a_val + L.nnz_a;
        L.up.val = L.lo.val + std::max<int64_t>(L.lo.nnz, 1);
        L.sc = L.up.val + std::max<int64_t>(L.up.nnz, 1);
        t->nv = L.sc + L.n;
        t->nw = 13 * L.npad;
        t->nw32 = 10 * L.npad;
        t->ni = (int64_t)ints.size();
        if ((s = dev_alloc_n(c, t->ni, &t->ints)) == FVG_OK && (s = dev_alloc_n(c, t->nv, &t->v64)) == FVG_OK &&
            (s = dev_alloc_n(c, t->nv, &t->v32)) == FVG_OK && (s = dev_alloc_n(c, t->nw, &t->w64)) == FVG_OK &&
            (s = dev_alloc_n(c, t->nw32, &t->w32)) == FVG_OK)
            s = upload(c, t->ints, ints.data(), t->ni);
        // padding elements of the work vectors stay zero for the life of the plan
        if (s == FVG_OK && (hipMemsetAsync(t->w64, 0, sizeof(double) * t->nw, c->stream) != hipSuccess ||
                            hipMemsetAsync(t->w32, 0, sizeof(float) * t->nw32, c->stream) != hipSuccess))
            s = fail(FVG_E_HIP, "workspace clear failed");
    }
    if (s != FVG_OK) {
        (void)hipStreamSynchronize(c->stream);
        plan_free(c, t);
        delete t;
        return s;
    }
    p->team = t;
    *out = t;
    return FVG_OK;
}

// worker pool of the multi-workgroup kernel: everything that can be co-resident, or the
// caller's smaller share
static int cluster_workers(const fvg_batch* B) {
    int w = B->cl_resident;
    if (B->cl_workers > 0) w = std::min(w, B->cl_workers);
    return std::max(w, 1);
}

static int batch_capture(fvg_ctx* c, fvg_batch* B) {
    if (B->exec) {
        (void)hipGraphExecDestroy(B->exec);
        B->exec = nullptr;
    }
    if (B->graph) {
        (void)hipGraphDestroy(B->graph);
        B->graph = nullptr;
    }
    const bool cl = B->cluster();
    const int workers = cl ? cluster_workers(B) : 0;
    const int sweeps = c->tune_inner_sweeps > 0 ? (int)c->tune_inner_sweeps : kInnerSweeps;
    long long tmo = 0;
    if (cl) {
        int khz = 0;
        FVG_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c->device));
        tmo = (long long)std::max(khz, 1000) * 1000LL * 10LL;   // a barrier wait gives up after 10 s
    }
    const void* fn = cl ? (B->solver == FVG_SOLVER_CG ? (const void*)k_cluster<false> : (const void*)k_cluster<true>)
                        : team_kernel_for(B->nt, B->variant);
    int32_t sw = sweeps;
    void* targs[] = {(void*)&B->d_desc, (void*)&B->d_count};
    void* cargs[] = {(void*)&B->d_desc, (void*)&B->d_cl_groups, (void*)&B->d_cl_tickets, (void*)&B->d_cl_head,
                     (void*)&B->d_cl_ctl, (void*)&sw, (void*)&tmo};
    FVG_HIP(hipStreamBeginCapture(B->stream, hipStreamCaptureModeThreadLocal));
    hipError_t le;
    if (cl) {
        hipLaunchKernelGGL(k_cluster_reset, dim3((unsigned)((B->capacity + 255) / 256)), dim3(256), 0, B->stream,
                           B->d_cl_ctl, B->capacity, B->d_cl_head);
        le = hipGetLastError();
        if (le == hipSuccess)
            le = hipLaunchKernel(fn, dim3((unsigned)workers), dim3(kClThreads), cargs, 0, B->stream);
    } else {
        le = hipLaunchKernel(fn, dim3((unsigned)B->capacity), dim3((unsigned)B->nt), targs, 0, B->stream);
    }
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(B->stream, &g);
    if (le != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return fail(FVG_E_HIP, std::string("batch kernel launch: ") + hipGetErrorString(le));
    }
    FVG_HIP(ce);
    B->graph = g;
    FVG_HIP(hipGraphInstantiate(&B->exec, B->graph, nullptr, nullptr, 0));
    B->graph_nt = B->nt;
    B->graph_cap = B->capacity;
    B->graph_kind = cl ? 1 : 0;
    B->graph_workers = B->cl_workers;
    B->graph_sweeps = sweeps;
    return FVG_OK;
}

static void fill_sweep_desc(const TeamPlan* B, const TeamSweepLayout& S, bool upper, TeamSweep& T) {
    const int32_t* I = B->ints;
    T.lvl = I + S.lvl;
    T.nnz1 = (int32_t)std::max<int64_t>(S.nnz - 1, 0);
    T.ptr = I + S.ptr;
    T.dep = I + S.dep;
    T.val = B->v64 + S.val;
    T.val32 = B->v32 + S.val;
    T.ring64 = S.ring64;
    T.ring32 = S.ring32;
    T.upper = upper ? 1 : 0;
}

}  // namespace fvg

namespace fvg {

static int check_members(int32_t n, const fvg_csr* const* a, const fvg_prec* const* p) {
    if (!a || !p || n <= 0) return fail(FVG_E_INVALID, "bad argument");
    for (int k = 0; k < n; ++k) {
        if (!a[k] || !p[k]) return fail(FVG_E_INVALID, "null matrix or preconditioner");
        if (a[k]->nrows != a[k]->ncols || a[k]->nrows <= 0) return fail(FVG_E_INVALID, "batched systems must be square");
        if (a[k]->nrows > INT32_MAX / 4) return fail(FVG_E_INVALID, "system too large for a batch");
        if (p[k]->n != a[k]->nrows) return fail(FVG_E_INVALID, "preconditioner size does not match its matrix");
        if (p[k]->level > 0)
            return fail(FVG_E_UNSUPPORTED, "batched solves take level-0 incomplete factorisations");
        const bool same = a[k]->pattern_id != 0 && a[k]->pattern_id == p[k]->pattern_id && a[k]->nnz == p[k]->nnz;
        if ((p[k]->kind == kIC || p[k]->kind == kILU) && !same && p[k]->h_col != a[k]->h_col)
            return fail(FVG_E_INVALID, "preconditioner pattern differs from its matrix");
    }
    std::vector<const fvg_prec*> seen(p, p + n);
    std::sort(seen.begin(), seen.end());
    if (std::adjacent_find(seen.begin(), seen.end()) != seen.end())
        return fail(FVG_E_INVALID, "a preconditioner appears twice in one batch");
    return FVG_OK;
}

// per-system storage for `cap` systems (history, pinned status, descriptors)
static int batch_reserve(fvg_ctx* c, fvg_batch* B, int cap) {
    if (cap <= B->capacity) return FVG_OK;
    if (B->capacity > 0) {   // grow: release the smaller storage first
        dev_free_n(c, B->hist, B->nhist);
        (void)hipHostFree(B->h_info);
        (void)hipHostFree(B->h_done);
        (void)hipHostFree(B->h_desc);
        B->h_info = nullptr;
        B->h_done = nullptr;
        B->h_desc = nullptr;
        dev_free_n(c, B->d_prof, 4 * B->capacity);
        dev_free_n(c, B->d_desc, B->capacity);
        (void)hipHostFree(B->h_cl);
        B->h_cl = nullptr;
        dev_free_n(c, B->d_cl_groups, B->capacity);
        dev_free_n(c, B->d_cl_tickets, (int64_t)B->capacity * kClMaxGroups);
        dev_free_n(c, B->d_cl_ctl, B->capacity);
    }
    B->capacity = cap;
    B->nhist = B->hcap * cap;
    FVG_TRY(dev_alloc_n(c, B->nhist, &B->hist));
    if (hipHostMalloc((void**)&B->h_info, sizeof(fvg_solver_info) * cap, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&B->h_done, sizeof(int32_t) * cap, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&B->h_desc, sizeof(TeamDesc) * (cap + 1), hipHostMallocDefault) != hipSuccess)
        return fail(FVG_E_HIP, "pinned batch status allocation failed");
    std::memset(B->h_info, 0, sizeof(fvg_solver_info) * cap);
    std::memset(B->h_done, 0, sizeof(int32_t) * cap);
    FVG_TRY(dev_alloc_n(c, cap, &B->d_desc));
    if (!B->d_count) FVG_TRY(dev_alloc_n(c, 1, &B->d_count));
    if (hipHostMalloc((void**)&B->h_cl, sizeof(int32_t) * ((size_t)cap * (1 + 2 * kClMaxGroups) + 1),
                      hipHostMallocDefault) != hipSuccess)
        return fail(FVG_E_HIP, "pinned batch staging allocation failed");
    FVG_TRY(dev_alloc_n(c, cap, &B->d_cl_groups));
    FVG_TRY(dev_alloc_n(c, (int64_t)cap * kClMaxGroups, &B->d_cl_tickets));
    FVG_TRY(dev_alloc_n(c, cap, &B->d_cl_ctl));
    if (!B->d_cl_head) FVG_TRY(dev_alloc_n(c, 1, &B->d_cl_head));
    B->desc_dirty = true;
    B->cl_dirty = true;
    return FVG_OK;
}

// workgroups and tickets of the multi-workgroup path, by launch slot
static void cluster_plan(fvg_batch* B) {
    B->cl_groups.assign((size_t)B->n_inst, 1);
    B->cl_tickets.clear();
    if (!B->cluster()) return;
    for (int slot = 0; slot < B->n_inst; ++slot) {
        const int64_t n = B->plan[(size_t)B->order[(size_t)slot]]->L.n;
        int64_t rows = B->cl_rows;
        for (size_t q = 0; q < B->cl_limit.size(); ++q)
            if (n <= B->cl_limit[q]) {
                rows = B->cl_brows[q];
                break;
            }
        const int64_t k = (n + rows - 1) / rows;
        // a system's workgroups must all run at once: never more than the launch's pool
        const int kmax = std::min(kClMaxGroups, std::max(B->cl_workers > 0 ? std::min(B->cl_workers, B->cl_resident)
                                                                            : B->cl_resident,
                                                         1));
        const int K = (int)std::min<int64_t>(std::max<int64_t>(k, 1), kmax);
        B->cl_groups[(size_t)slot] = K;
        for (int r = 0; r < K; ++r) {
            B->cl_tickets.push_back(slot);
            B->cl_tickets.push_back(r);
        }
    }
    B->cl_dirty = true;
}

// the systems of the next launches: plans, launch order and descriptors (uploaded at launch)
static int batch_bind(fvg_ctx* c, fvg_batch* B, int32_t n, const fvg_csr* const* a, const fvg_prec* const* p) {
    FVG_TRY(check_members(n, a, p));
    FVG_TRY(batch_reserve(c, B, n));
    B->n_inst = n;
    B->A.assign(a, a + n);
    B->P.resize((size_t)n);
    B->plan.assign((size_t)n, nullptr);
    for (int k = 0; k < n; ++k) {
        B->P[(size_t)k] = const_cast<fvg_prec*>(p[k]);   // the plan is a cache owned by the preconditioner
        FVG_TRY(get_plan(c, a[k], B->P[(size_t)k], &B->plan[(size_t)k]));
    }
    // largest systems first, so they start first and spread over the XCDs
    B->order.resize((size_t)n);
    std::iota(B->order.begin(), B->order.end(), 0);
    std::stable_sort(B->order.begin(), B->order.end(),
                     [&](int x, int y) { return B->plan[(size_t)x]->L.n > B->plan[(size_t)y]->L.n; });
    const fvg_solver_options* o = &B->opts;
    B->desc.assign((size_t)n, TeamDesc{});
    for (int slot = 0; slot < n; ++slot) {
        const int k = B->order[(size_t)slot];
        const TeamPlan* t = B->plan[(size_t)k];
        const TeamLayout& L = t->L;
        TeamDesc& D = B->desc[(size_t)slot];
        std::memset(&D, 0, sizeof(D));
        D.n = (int32_t)L.n;
        D.npad = (int32_t)L.npad;
        D.wspan = L.wspan;
        D.a_nnz1 = (int32_t)(L.nnz_a - 1);
        D.win64 = L.win64;
        D.win32 = L.win32;
        D.kind = p[k]->kind;
        const int32_t* I = t->ints;
        D.perm = I + L.perm;
        D.a_ptr = I + L.a_ptr;
        D.a_dep = I + L.a_dep;
        D.a_val = t->v64 + L.a_val;
        D.a_val32 = t->v32 + L.a_val;
        fill_sweep_desc(t, L.lo, false, D.lo);
        fill_sweep_desc(t, L.up, true, D.up);
        D.sc = t->v64 + L.sc;
        D.sc32 = t->v32 + L.sc;
        D.w64 = t->w64;
        D.w32 = t->w32;
        D.hist = B->hist + B->hcap * k;
        D.hcap = B->hcap;
        D.info = B->h_info + k;
        D.done = B->h_done + k;
        D.pflags = (p[k]->kind == kIC || p[k]->kind == kILU) ? p[k]->d_flags : nullptr;
        D.prof = B->d_prof ? B->d_prof + 4 * k : nullptr;
        D.atol = o->absolute_tolerance;
        D.rtol = o->relative_tolerance;
        D.inner_tol = o->inner_tolerance > 0.0 ? o->inner_tolerance : 1e-4;
        D.max_iter = o->max_iterations;
    }
    B->desc_dirty = true;
    cluster_plan(B);
    return FVG_OK;
}

}  // namespace fvg

using namespace fvg;

extern "C" {

int fvg_batch_create(fvg_ctx* c, int32_t n, const fvg_csr* const* a, const fvg_prec* const* p, int32_t solver,
                     const fvg_solver_options* o, fvg_batch** out) {
    if (!c || !a || !p || !o || !out || n <= 0) return fail(FVG_E_INVALID, "bad argument");
    if (solver != FVG_SOLVER_CG && solver != FVG_SOLVER_BICGSTAB) return fail(FVG_E_INVALID, "unknown solver");
    if (o->max_iterations < 0) return fail(FVG_E_INVALID, "max_iterations must be non-negative");
    *out = nullptr;
    FVG_TRY(check_members(n, a, p));
    auto* B = new fvg_batch();
    B->solver = solver;
    B->precision = o->precision;
    B->opts = *o;
    B->variant = (solver == FVG_SOLVER_CG ? 0 : 1) + (o->precision == FVG_PRECISION_MIXED ? 2 : 0);
    B->hcap = o->precision == FVG_PRECISION_MIXED ? 64 : o->max_iterations + 2;
    {
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_cluster<true>, kClThreads, 0) != hipSuccess ||
            per_cu <= 0)
            per_cu = 1;
        B->cl_resident = per_cu * std::max(c->num_cus, 1);
    }
    int s = batch_bind(c, B, n, a, p);
    if (s == FVG_OK && hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking) != hipSuccess)
        s = fail(FVG_E_HIP, "batch stream creation failed");
    if (s == FVG_OK && hipEventCreateWithFlags(&B->ready, hipEventDisableTiming) != hipSuccess)
        s = fail(FVG_E_HIP, "batch event creation failed");
    if (s != FVG_OK) {
        (void)hipStreamSynchronize(c->stream);
        batch_free(c, B);
        delete B;
        return s;
    }
    *out = B;
    return FVG_OK;
}

int fvg_batch_reserve(fvg_ctx* c, fvg_batch* B, int32_t capacity) {
    if (!c || !B || capacity <= 0) return fail(FVG_E_INVALID, "bad argument");
    if (B->launched) return fail(FVG_E_INVALID, "batch in flight: finish it first");
    if (capacity <= B->capacity) return FVG_OK;
    FVG_TRY(batch_reserve(c, B, capacity));
    // descriptors point into the per-system storage just moved: rebuild them
    std::vector<const fvg_prec*> ps(B->P.begin(), B->P.end());
    return batch_bind(c, B, B->n_inst, B->A.data(), ps.data());
}

int fvg_batch_rebind(fvg_ctx* c, fvg_batch* B, int32_t n, const fvg_csr* const* a, const fvg_prec* const* p) {
    if (!c || !B) return fail(FVG_E_INVALID, "null argument");
    if (B->launched) return fail(FVG_E_INVALID, "batch in flight: finish it first");
    return batch_bind(c, B, n, a, p);
}

int fvg_batch_destroy(fvg_ctx* c, fvg_batch* B) {
    if (!c || !B) return fail(FVG_E_INVALID, "null argument");
    batch_free(c, B);
    delete B;
    return FVG_OK;
}

int fvg_batch_size(const fvg_batch* B, int32_t* n_systems, int64_t* total_rows) {
    if (!B) return fail(FVG_E_INVALID, "null argument");
    if (n_systems) *n_systems = B->n_inst;
    if (total_rows) {
        int64_t t = 0;
        for (const TeamPlan* pl : B->plan) t += pl->L.n;
        *total_rows = t;
    }
    return FVG_OK;
}

int fvg_batch_launch(fvg_ctx* c, fvg_batch* B, const fvg_vec* const* b, fvg_vec* const* x) {
    if (!c || !B || !b || !x) return fail(FVG_E_INVALID, "null argument");
    if (B->launched) return fail(FVG_E_INVALID, "batch already in flight: finish it first");
    for (const TeamPlan* t : B->plan)
        if (t->busy) return fail(FVG_E_INVALID, "a system of this batch is in another batch in flight");
    bool moved = false;
    for (int slot = 0; slot < B->n_inst; ++slot) {
        const int k = B->order[(size_t)slot];
        const fvg_vec* bk = b[k];
        fvg_vec* xk = x[k];
        if (!bk || !xk) return fail(FVG_E_INVALID, "null vector");
        if (bk->n != B->plan[(size_t)k]->L.n || xk->n != B->plan[(size_t)k]->L.n)
            return fail(FVG_E_INVALID, "vector size does not match its system");
        if (bk->d == xk->d) return fail(FVG_E_INVALID, "b and x must not alias");
        TeamDesc& D = B->desc[(size_t)slot];
        if (D.b != bk->d || D.x != xk->d) {
            D.b = bk->d;
            D.x = xk->d;
            moved = true;
        }
    }
    moved = moved || B->desc_dirty;
    // refresh values assembled or refactored since the last launch
    for (int k = 0; k < B->n_inst; ++k) {
        TeamPlan* t = B->plan[(size_t)k];
        if (t->a_ver != t->a->version || t->p_ver != B->P[(size_t)k]->fac_version)
            FVG_TRY(plan_refresh(c, t, B->P[(size_t)k], c->stream));
    }
    if (moved) {
        // through pinned memory, so the copies never wait for the host; the pinned copy is
        // rewritten only at a later launch, after this batch has finished
        std::memcpy(B->h_desc, B->desc.data(), sizeof(TeamDesc) * B->desc.size());
        int32_t* cnt = reinterpret_cast<int32_t*>(B->h_desc + B->capacity);
        *cnt = B->n_inst;
        FVG_HIP(hipMemcpyAsync(B->d_desc, B->h_desc, sizeof(TeamDesc) * B->desc.size(), hipMemcpyHostToDevice,
                               c->stream));
        FVG_HIP(hipMemcpyAsync(B->d_count, cnt, sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
        B->desc_dirty = false;
    }
    if (B->cluster() && B->cl_dirty) {
        const int nt = (int)(B->cl_tickets.size() / 2);
        int32_t* g = B->h_cl;
        int32_t* tk = B->h_cl + B->capacity;
        std::memcpy(g, B->cl_groups.data(), sizeof(int32_t) * B->cl_groups.size());
        std::memcpy(tk, B->cl_tickets.data(), sizeof(int32_t) * B->cl_tickets.size());
        int32_t* cnt = tk + 2 * (size_t)B->capacity * kClMaxGroups;   // last staging word: the ticket count
        *cnt = nt;
        FVG_HIP(hipMemcpyAsync(B->d_cl_groups, g, sizeof(int32_t) * B->cl_groups.size(), hipMemcpyHostToDevice,
                               c->stream));
        if (nt > 0)
            FVG_HIP(hipMemcpyAsync(B->d_cl_tickets, tk, sizeof(int32_t) * B->cl_tickets.size(), hipMemcpyHostToDevice,
                                   c->stream));
        FVG_HIP(hipMemcpyAsync(&B->d_cl_head->nticket, cnt, sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
        B->cl_dirty = false;
    }
    std::memset(B->h_done, 0, sizeof(int32_t) * (size_t)B->n_inst);
    // the batch stream starts after everything enqueued on the context stream so far
    FVG_HIP(hipEventRecord(B->ready, c->stream));
    FVG_HIP(hipStreamWaitEvent(B->stream, B->ready, 0));
    const int sweeps = c->tune_inner_sweeps > 0 ? (int)c->tune_inner_sweeps : kInnerSweeps;
    const bool stale = B->cluster() ? (B->graph_kind != 1 || B->graph_workers != B->cl_workers ||
                                       B->graph_sweeps != sweeps)
                                    : (B->graph_kind != 0 || B->graph_nt != B->nt);
    if (!B->exec || stale || B->graph_cap != B->capacity) FVG_TRY(batch_capture(c, B));
    FVG_HIP(hipGraphLaunch(B->exec, B->stream));
    B->launched = true;
    B->flight_ctx = c;
    c->batches_in_flight += 1;   // until finished: co-resident sweeps of c launch cooperatively
    for (TeamPlan* t : B->plan) t->busy = true;
    return FVG_OK;
}

int fvg_batch_finish(fvg_ctx* c, fvg_batch* B, fvg_solver_info* infos, double* const* histories, int64_t cap) {
    if (!c || !B) return fail(FVG_E_INVALID, "null argument");
    if (!B->launched) return fail(FVG_E_INVALID, "batch not launched");
    B->launched = false;
    if (B->flight_ctx) B->flight_ctx->batches_in_flight -= 1;
    B->flight_ctx = nullptr;
    for (TeamPlan* t : B->plan) t->busy = false;
    FVG_HIP(hipStreamSynchronize(B->stream));
    if (infos) std::memcpy(infos, B->h_info, sizeof(fvg_solver_info) * (size_t)B->n_inst);
    if (!infos) infos = B->h_info;
    if (histories && cap > 0) {
        for (int k = 0; k < B->n_inst; ++k) {
            if (!histories[k]) continue;
            const int64_t m = std::min(infos[k].history_len, std::min(cap, B->hcap));
            if (m > 0)
                FVG_HIP(hipMemcpyAsync(histories[k], B->hist + B->hcap * k, sizeof(double) * (size_t)m,
                                       hipMemcpyDeviceToHost, B->stream));
        }
        FVG_HIP(hipStreamSynchronize(B->stream));
    }
    return FVG_OK;
}

int fvg_batch_poll(fvg_ctx* c, fvg_batch* B, int32_t* done, fvg_solver_info* infos) {
    if (!c || !B || !done) return fail(FVG_E_INVALID, "null argument");
    for (int k = 0; k < B->n_inst; ++k) {
        done[k] = __atomic_load_n(B->h_done + k, __ATOMIC_ACQUIRE);
        if (infos && done[k]) infos[k] = B->h_info[k];
    }
    return FVG_OK;
}

int fvg_batch_query(fvg_ctx* c, fvg_batch* B, int32_t* done) {
    if (!c || !B || !done) return fail(FVG_E_INVALID, "null argument");
    if (!B->launched) {
        *done = 1;
        return FVG_OK;
    }
    const hipError_t e = hipStreamQuery(B->stream);
    if (e == hipSuccess) {
        *done = 1;
        return FVG_OK;
    }
    if (e == hipErrorNotReady) {
        *done = 0;
        return FVG_OK;
    }
    return fail(FVG_E_HIP, std::string("batch stream: ") + hipGetErrorString(e));
}

int fvg_batch_solve(fvg_ctx* c, fvg_batch* B, const fvg_vec* const* b, fvg_vec* const* x, fvg_solver_info* infos,
                    double* const* histories, int64_t cap) {
    FVG_TRY(fvg_batch_launch(c, B, b, x));
    return fvg_batch_finish(c, B, infos, histories, cap);
}

int fvg_batch_set_threads(fvg_ctx* c, fvg_batch* B, int32_t threads) {
    if (!c || !B) return fail(FVG_E_INVALID, "null argument");
    if (threads != 256 && threads != 512 && threads != 1024)
        return fail(FVG_E_INVALID, "workgroup size must be 256, 512 or 1024");
    if (B->launched) return fail(FVG_E_INVALID, "batch in flight");
    B->nt = threads;
    return FVG_OK;
}

int fvg_batch_set_phase_timing(fvg_ctx* c, fvg_batch* B, int32_t enable) {
    if (!c || !B) return fail(FVG_E_INVALID, "null argument");
    if (B->launched) return fail(FVG_E_INVALID, "batch in flight");
    if (enable && !B->d_prof) {
        FVG_TRY(dev_alloc_n(c, 4 * (int64_t)B->capacity, &B->d_prof));
        FVG_HIP(hipMemsetAsync(B->d_prof, 0, sizeof(int64_t) * 4 * B->capacity, c->stream));
    }
    for (int slot = 0; slot < B->n_inst; ++slot)
        B->desc[(size_t)slot].prof = enable ? B->d_prof + 4 * B->order[(size_t)slot] : nullptr;
    B->desc_dirty = true;
    return FVG_OK;
}

int fvg_batch_phase_times(fvg_ctx* c, fvg_batch* B, double* us) {
    if (!c || !B || !us) return fail(FVG_E_INVALID, "null argument");
    if (!B->d_prof) return fail(FVG_E_INVALID, "phase timing is not enabled");
    std::vector<int64_t> t((size_t)4 * B->n_inst);
    FVG_TRY(download(c, t.data(), B->d_prof, (int64_t)t.size()));
    int khz = 0;
    FVG_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c->device));
    if (khz <= 0) return fail(FVG_E_HIP, "wall clock rate unavailable");
    for (size_t i = 0; i < t.size(); ++i) us[i] = (double)t[i] * 1e3 / (double)khz;
    return FVG_OK;
}

int fvg_batch_set_cluster(fvg_ctx* c, fvg_batch* B, int32_t rows_per_group, int32_t max_workers) {
    if (!c || !B) return fail(FVG_E_INVALID, "null argument");
    if (rows_per_group < 0 || max_workers < 0) return fail(FVG_E_INVALID, "negative cluster setting");
    if (B->launched) return fail(FVG_E_INVALID, "batch in flight");
    B->cl_rows = rows_per_group;
    B->cl_workers = max_workers;
    cluster_plan(B);
    return FVG_OK;
}

int fvg_batch_set_cluster_buckets(fvg_ctx* c, fvg_batch* B, int32_t count, const int64_t* max_rows,
                                  const int32_t* rows_per_group) {
    if (!c || !B || count < 0 || (count > 0 && (!max_rows || !rows_per_group)))
        return fail(FVG_E_INVALID, "null argument or negative count");
    if (B->launched) return fail(FVG_E_INVALID, "batch in flight");
    for (int32_t q = 0; q < count; ++q) {
        if (rows_per_group[q] <= 0 || max_rows[q] <= 0) return fail(FVG_E_INVALID, "bucket sizes must be positive");
        if (q > 0 && max_rows[q] <= max_rows[q - 1]) return fail(FVG_E_INVALID, "bucket limits must ascend");
    }
    B->cl_limit.assign(max_rows, max_rows + count);
    B->cl_brows.assign(rows_per_group, rows_per_group + count);
    cluster_plan(B);
    return FVG_OK;
}

int fvg_batch_get_cluster(const fvg_batch* B, int32_t* rows_per_group, int32_t* workers) {
    if (!B) return fail(FVG_E_INVALID, "null argument");
    if (rows_per_group) *rows_per_group = B->cluster() ? B->cl_rows : 0;
    if (workers) *workers = B->cluster() ? cluster_workers(B) : 0;
    return FVG_OK;
}

int fvg_batch_get_threads(const fvg_batch* B, int32_t* threads) {
    if (!B || !threads) return fail(FVG_E_INVALID, "null argument");
    *threads = B->nt;
    return FVG_OK;
}

}  // extern "C"
