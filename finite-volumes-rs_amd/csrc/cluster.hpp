// cluster.hpp -- mixed-precision batched solves spread over several workgroups per system.
// Included by team.hip inside namespace fvg, after the system descriptor (TeamDesc).
//
// One workgroup moves at most a few tens of GB/s from HBM (its outstanding misses over the
// memory latency), so a system of 10^4..10^6 rows solved by one workgroup takes as long
// as its data takes to stream through one CU.  Here a system of n rows is solved by
// K = ceil(n / rows_per_group) workgroups (a "cluster", 1 <= K <= kClMaxGroups), each
// owning a contiguous range of positions; the cluster meets at agent-scope barriers, and
// every dot product is a fixed-order sum of the K workgroup partials, so each workgroup
// takes the same branch and a result depends only on the system and on K (not on the
// batch it runs in).
//
// The algorithm is the FP32-inner / FP64-refinement solve of the single-workgroup mixed
// path (refinement on the true residual b - A x in FP64, monotone acceptance of
// corrections), with two differences that the refinement makes legitimate:
//  * the incomplete factor is applied by Jacobi-iterated triangular solves (m sweeps per
//    factor, every sweep one parallel pass), since a level-by-level sweep would need one
//    cluster barrier per level;
//  * when the inner solves stop making progress the refinement applies the factor with
//    more sweeps, then runs the inner iterations in FP64 (flag FVG_INFO_FP64_FALLBACK).
//
// Scheduling: the grid is a pool of workers (at most what is co-resident).  A worker
// takes tickets in order from a launch-wide counter; ticket t names (system, rank), with
// the ranks of one system consecutive and systems largest first.  A ticket is only ever
// held by a running worker, and all tickets before the last taken one belong to systems
// whose ranks are all held, so the earliest unfinished system always has its whole
// cluster running: the batch drains.  Every barrier wait is bounded by a wall-clock limit;
// a wait that expires marks the system aborted (FVG_E_TIMEOUT) and every later barrier
// of that system returns at once, so a launch always ends.
constexpr int kClThreads = 512;
constexpr int kClMaxGroups = 64;
constexpr int kClWaves = kClThreads / 64;

struct alignas(128) ClusterCtl {   // per system; arrive and abort are zeroed before a launch
    int32_t arrive;
    int32_t abort;
    int32_t pad[30];
    double part[2][kClMaxGroups][4];   // workgroup partial sums, alternating by reduction
};

struct ClusterHead {   // per launch: the ticket counter (zeroed) and the ticket count
    int32_t next;
    int32_t nticket;
    int32_t pad[30];
};

struct Grp {
    ClusterCtl* C;
    long long tmo;     // barrier wait limit, wall-clock ticks
    int K, rank;
    int p0, p1;        // positions [p0, p1) belong to this workgroup
    int epoch;         // barriers passed
    int red;           // reductions done (selects the partial buffer)
    int par;           // LDS partial buffer parity
    double* lds;       // 2 x 4 x kClWaves partials
    double* bc;        // 4 broadcast slots
};

// Vectors that another workgroup of the cluster reads are written with device-scope stores
// (write-through to the device's coherence point) and read by the other workgroups with
// device-scope loads; a workgroup reads its own positions with ordinary loads (each
// position is only ever touched by one thread of its workgroup).  So a barrier needs no
// cache write-back or invalidation -- which on a multi-XCD device would flush the L2 of
// every workgroup's XCD at every barrier -- only that this workgroup's stores have
// completed before it arrives.
template <class T>
__device__ __forceinline__ T ld_sh(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ void st_sh(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// cluster barrier: every workgroup's shared-vector writes before it are visible to every
// workgroup's device-scope loads after it
__device__ void grp_sync(Grp& G) {
    G.epoch += 1;
    __builtin_amdgcn_s_waitcnt(0);   // this wave's stores have completed
    __syncthreads();
    if (G.K > 1 && threadIdx.x == 0) {
        __hip_atomic_fetch_add(&G.C->arrive, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const int target = G.epoch * G.K;
        const long long t0 = wall_clock64();
        while (__hip_atomic_load(&G.C->arrive, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (__hip_atomic_load(&G.C->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
            if (wall_clock64() - t0 > G.tmo) {
                __hip_atomic_store(&G.C->abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
    }
    __syncthreads();
}

// cluster-wide sums of N values (result in every thread of every workgroup, same bits);
// doubles as a cluster barrier
template <int N>
__device__ void grp_sum(Grp& G, double (&v)[N]) {
    static_assert(N >= 1 && N <= 4, "at most four sums at once");
    double* L = G.lds + G.par * 4 * kClWaves;
    G.par ^= 1;
#pragma unroll
    for (int j = 0; j < N; ++j) v[j] = wave_sum(v[j]);
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int j = 0; j < N; ++j) L[j * kClWaves + (threadIdx.x >> 6)] = v[j];
    __syncthreads();
    double s[N];
#pragma unroll
    for (int j = 0; j < N; ++j) {
        s[j] = 0.0;
#pragma unroll
        for (int w = 0; w < kClWaves; ++w) s[j] += L[j * kClWaves + w];
    }
    if (G.K == 1) {
#pragma unroll
        for (int j = 0; j < N; ++j) v[j] = s[j];
        return;
    }
    double* P = G.C->part[G.red & 1][G.rank];
    G.red += 1;
    if (threadIdx.x == 0)
#pragma unroll
        for (int j = 0; j < N; ++j) __hip_atomic_store(&P[j], s[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    grp_sync(G);
    if (threadIdx.x < 64) {
        const int g = threadIdx.x;
        const double* Q = G.C->part[(G.red - 1) & 1][g < G.K ? g : 0];
#pragma unroll
        for (int j = 0; j < N; ++j) {
            double t = g < G.K ? __hip_atomic_load(&Q[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0.0;
            t = wave_sum(t);   // fixed tree over the K partials
            if (g == 0) G.bc[j] = t;
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < N; ++j) v[j] = G.bc[j];
}

// Row passes (SpMV, factor sweeps) take kClRows rows per thread at a time and issue the
// loads of all of them before using any: row bounds, then the first kClEnt entries'
// columns and coefficients (indices clamped into the row, or onto the last entry, so no
// load is conditional), then the gathered vector values.  Longer rows finish their
// remaining entries one by one.  Sums run over the entries in stored order.
template <class T>
struct ClUnroll {
    static constexpr int rows = sizeof(T) == 4 ? 4 : 2;   // FP64 rows take twice the registers
};
constexpr int kClEnt = 4;

// y = A x over this workgroup's positions (entries in stored order, from zero)
template <class T, class Epi>
__device__ __forceinline__ void cl_spmv(const TeamDesc& D, const Grp& G, const T* __restrict__ val,
                                        const T* __restrict__ x, Epi epi) {
    constexpr int kClRows = ClUnroll<T>::rows;
    const int32_t* __restrict__ ap = D.a_ptr;
    const int32_t* __restrict__ ad = D.a_dep;
    const int nnz1 = D.a_nnz1;
    for (int base = G.p0 + (int)threadIdx.x; base < G.p1; base += kClThreads * kClRows) {
        int kb[kClRows], cnt[kClRows];
#pragma unroll
        for (int u = 0; u < kClRows; ++u) {
            const int p = base + u * kClThreads;
            const int pc = p < G.p1 ? p : G.p0;
            kb[u] = ap[pc];
            const int ke = ap[pc + 1];
            cnt[u] = p < G.p1 ? ke - kb[u] : 0;
        }
        int dj[kClRows][kClEnt];
        T a[kClRows][kClEnt], xv[kClRows][kClEnt];
#pragma unroll
        for (int u = 0; u < kClRows; ++u) {
            const int last = cnt[u] > 0 ? cnt[u] - 1 : 0;
#pragma unroll
            for (int e = 0; e < kClEnt; ++e) {
                const int k = min(kb[u] + min(e, last), nnz1);
                dj[u][e] = ad[k];
                a[u][e] = val[k];
            }
        }
#pragma unroll
        for (int u = 0; u < kClRows; ++u)
#pragma unroll
            for (int e = 0; e < kClEnt; ++e) xv[u][e] = ld_sh(x + dj[u][e]);
#pragma unroll
        for (int u = 0; u < kClRows; ++u) {
            T s = T(0);
#pragma unroll
            for (int e = 0; e < kClEnt; ++e)
                if (e < cnt[u]) s += a[u][e] * xv[u][e];
            for (int k = kb[u] + kClEnt; k < kb[u] + cnt[u]; ++k) s += val[k] * ld_sh(x + ad[k]);
            const int p = base + u * kClThreads;
            if (p < G.p1) epi(p, s);
        }
    }
}

// one Jacobi sweep of a triangular factor over this workgroup's positions:
// out = sc (v - N yin) (sc = 1 unless scaled), with yin = sc v on the first sweep
template <class T, bool FIRST>
__device__ __forceinline__ void cl_jacobi(const TeamSweep& S, const T* __restrict__ val, const T* __restrict__ sc,
                                          bool scaled, const T* __restrict__ v, const T* __restrict__ yin,
                                          T* __restrict__ out, int n, const Grp& G) {
    const int32_t* __restrict__ ptr = S.ptr;
    const int32_t* __restrict__ dep = S.dep;
    constexpr int kClRows = ClUnroll<T>::rows;
    const int nnz1 = S.nnz1;
    const T* __restrict__ src = FIRST ? v : yin;
    for (int base = G.p0 + (int)threadIdx.x; base < G.p1; base += kClThreads * kClRows) {
        int kb[kClRows], cnt[kClRows];
        T vp[kClRows], sp[kClRows];
#pragma unroll
        for (int u = 0; u < kClRows; ++u) {
            const int p = base + u * kClThreads;
            const int pc = p < G.p1 ? p : G.p0;
            const int q = S.upper ? n - 1 - pc : pc;
            kb[u] = ptr[q];
            const int ke = ptr[q + 1];
            cnt[u] = p < G.p1 ? ke - kb[u] : 0;
            vp[u] = v[pc];
            sp[u] = scaled ? sc[pc] : T(1);
        }
        int dj[kClRows][kClEnt];
        T a[kClRows][kClEnt], yv[kClRows][kClEnt], sj[kClRows][kClEnt];
#pragma unroll
        for (int u = 0; u < kClRows; ++u) {
            const int last = cnt[u] > 0 ? cnt[u] - 1 : 0;
#pragma unroll
            for (int e = 0; e < kClEnt; ++e) {
                const int k = min(kb[u] + min(e, last), nnz1);
                dj[u][e] = dep[k];
                a[u][e] = val[k];
            }
        }
#pragma unroll
        for (int u = 0; u < kClRows; ++u)
#pragma unroll
            for (int e = 0; e < kClEnt; ++e) {
                yv[u][e] = ld_sh(src + dj[u][e]);
                sj[u][e] = (FIRST && scaled) ? sc[dj[u][e]] : T(1);
            }
#pragma unroll
        for (int u = 0; u < kClRows; ++u) {
            T s = vp[u];
#pragma unroll
            for (int e = 0; e < kClEnt; ++e)
                if (e < cnt[u]) s -= a[u][e] * ((FIRST && scaled) ? sj[u][e] * yv[u][e] : yv[u][e]);
            for (int k = kb[u] + kClEnt; k < kb[u] + cnt[u]; ++k) {
                const int j = dep[k];
                const T yj = ld_sh(src + j);
                s -= val[k] * ((FIRST && scaled) ? sc[j] * yj : yj);
            }
            const int p = base + u * kClThreads;
            if (p < G.p1) st_sh(out + p, scaled ? sp[u] * s : s);
        }
    }
}

// m Jacobi sweeps of one factor; sweep k writes `odd` (k odd) or `even`; a cluster
// barrier follows every sweep but the last when last_sync is false
template <class T>
__device__ void cl_factor(const TeamSweep& S, const T* val, const T* sc, bool scaled, int m, const T* v, T* odd,
                          T* even, int n, Grp& G, bool last_sync) {
    const T* prev = nullptr;
    for (int k = 1; k <= m; ++k) {
        T* out = (k & 1) ? odd : even;
        if (k == 1)
            cl_jacobi<T, true>(S, val, sc, scaled, v, nullptr, out, n, G);
        else
            cl_jacobi<T, false>(S, val, sc, scaled, v, prev, out, n, G);
        prev = out;
        if (k < m || last_sync) grp_sync(G);
    }
}

// z = M^-1 r (approximate for IC / ILU: m Jacobi sweeps per factor); t1, t2 are
// temporaries distinct from r and z.  Ends without a barrier: the caller's next
// reduction (or barrier) publishes z.
template <class T>
__device__ void cl_prec(const TeamDesc& D, Grp& G, int m, const T* r, T* z, T* t1, T* t2) {
    const T* sc = pick<T>(D.sc, D.sc32);
    const int n = D.n;
    if (D.kind == kIdentity || D.kind == kDiagonal) {
        const bool dg = D.kind == kDiagonal;
        for (int p = G.p0 + (int)threadIdx.x; p < G.p1; p += kClThreads) st_sh(z + p, dg ? sc[p] * r[p] : r[p]);
        return;
    }
    // lower factor: IC scaled by 1/l_ii, ILU unit diagonal; the result lands in t1 (m odd) or t2
    cl_factor<T>(D.lo, pick<T>(D.lo.val, D.lo.val32), sc, D.kind == kIC, m, r, t1, t2, n, G, true);
    const T* y = (m & 1) ? t1 : t2;
    T* o = (m & 1) ? t2 : t1;
    // upper factor, scaled by 1/l_ii (IC) or 1/u_ii (ILU); its last sweep writes z
    cl_factor<T>(D.up, pick<T>(D.up.val, D.up.val32), sc, true, m, y, (m & 1) ? z : o, (m & 1) ? o : z, n, G, false);
}

__device__ __forceinline__ double cl_inner_target(const TeamDesc& D, double e, double init_res) {
    const double goal = fmax(D.rtol * init_res, D.atol);
    double t = 0.3 * goal / e;
    t = fmin(t, 0.1);
    return fmax(t, D.inner_tol);
}

// inner vectors of one precision: FP32 in the plan's FP32 workspace, FP64 after b, x, xa
template <class T>
struct ClVecs {
    T *r, *d, *z, *p, *w, *rh, *y, *v, *s, *t;
};

template <class T>
__device__ ClVecs<T> cl_vecs(const TeamDesc& D);
template <>
__device__ ClVecs<float> cl_vecs<float>(const TeamDesc& D) {
    float* b = D.w32;
    const int np = D.npad;
    return ClVecs<float>{b, b + np, b + 2 * np, b + 3 * np, b + 4 * np, b + 5 * np, b + 6 * np, b + 7 * np,
                         b + 8 * np, b + 9 * np};
}
template <>
__device__ ClVecs<double> cl_vecs<double>(const TeamDesc& D) {
    double* b = D.w64 + 3 * (int64_t)D.npad;
    const int np = D.npad;
    return ClVecs<double>{b, b + np, b + 2 * np, b + 3 * np, b + 4 * np, b + 5 * np, b + 6 * np, b + 7 * np,
                          b + 8 * np, b + 9 * np};
}

// r = b - A x in FP64 (stored as T), d = 0; returns |r|^2 of the cluster (a barrier)
template <class T>
__device__ double cl_residual(const TeamDesc& D, Grp& G, const double* b, const double* x) {
    ClVecs<T> V = cl_vecs<T>(D);
    double q[1] = {0.0};
    cl_spmv<double>(D, G, D.a_val, x, [&](int p, double sv) {
        const double ri = (sv - b[p]) * -1.0;
        st_sh(V.r + p, (T)ri);
        V.d[p] = T(0);
        q[0] += ri * ri;
    });
    grp_sum<1>(G, q);
    return q[0];
}

// Inner iterations in T arithmetic on A d = r (r from cl_residual, d = 0), for at most
// `inner_end` iterations in total or until |r| <= tol |r0|.  Thread partials accumulate in
// T; the few workgroup partials of a dot product are combined in FP64.
template <class T>
__device__ void cl_cg(const TeamDesc& D, Grp& G, int m, double tol, int64_t& total, int64_t inner_end,
                      int64_t max_inner) {
    ClVecs<T> V = cl_vecs<T>(D);
    const T* aval = pick<T>(D.a_val, D.a_val32);
    cl_prec<T>(D, G, m, V.r, V.z, V.w, V.t);   // z = M r; p = z; |r|^2, r.z
    T a0 = T(0), a1 = T(0);
    for (int p = G.p0 + (int)threadIdx.x; p < G.p1; p += kClThreads) {
        const T rr = V.r[p], zz = V.z[p];
        st_sh(V.p + p, zz);
        a0 += rr * rr;
        a1 += rr * zz;
    }
    double q[2] = {(double)a0, (double)a1};
    grp_sum<2>(G, q);
    const double rin0 = sqrt(q[0]);
    double rz = q[1];
    if (!(rin0 > 0.0)) return;   // the T residual underflowed: refinement cannot progress
    for (;;) {
        T qa = T(0);
        cl_spmv<T>(D, G, aval, V.p, [&](int p, T sv) {
            st_sh(V.w + p, sv);   // w doubles as a factor-sweep temporary (read across the cluster)
            qa += V.p[p] * sv;
        });
        double s1[1] = {(double)qa};
        grp_sum<1>(G, s1);
        const double alpha_d = rz / s1[0];
        if (!isfinite(alpha_d)) break;   // breakdown: refine from the correction so far
        const T alpha = (T)alpha_d;
        T qr = T(0);
        for (int p = G.p0 + (int)threadIdx.x; p < G.p1; p += kClThreads) {
            V.d[p] += alpha * V.p[p];
            const T rn = V.r[p] - alpha * V.w[p];
            st_sh(V.r + p, rn);
            qr += rn * rn;
        }
        double s2[1] = {(double)qr};
        grp_sum<1>(G, s2);
        const double rin = sqrt(s2[0]);
        total += 1;
        if (rin <= tol * rin0 || total >= max_inner || total >= inner_end) break;
        if (!(rin < 1e2 * rin0)) break;   // diverging: let the refinement judge this correction
        cl_prec<T>(D, G, m, V.r, V.z, V.w, V.t);
        T qz = T(0);
        for (int p = G.p0 + (int)threadIdx.x; p < G.p1; p += kClThreads) qz += V.r[p] * V.z[p];
        double s3[1] = {(double)qz};
        grp_sum<1>(G, s3);
        const T beta = (T)(s3[0] / rz);
        rz = s3[0];
        for (int p = G.p0 + (int)threadIdx.x; p < G.p1; p += kClThreads) st_sh(V.p + p, V.z[p] + beta * V.p[p]);
        grp_sync(G);
    }
}

template <class T>
__device__ void cl_bicg(const TeamDesc& D, Grp& G, int m, double tol, int64_t& total, int64_t inner_end,
                        int64_t max_inner, int64_t& restarts) {
    ClVecs<T> V = cl_vecs<T>(D);
    const T* aval = pick<T>(D.a_val, D.a_val32);
    T a0 = T(0);
    for (int p = G.p0 + (int)threadIdx.x; p < G.p1; p += kClThreads) {
        const T rr = V.r[p];
        V.rh[p] = rr;
        st_sh(V.p + p, rr);
        a0 += rr * rr;
    }
    double q[1] = {(double)a0};
    grp_sum<1>(G, q);
    const double rin0 = sqrt(q[0]);
    double rho = q[0], alpha = 0.0, omega = 0.0;
    if (!(rin0 > 0.0)) return;
    for (;;) {
        cl_prec<T>(D, G, m, V.p, V.y, V.z, V.t);   // y = M p
        grp_sync(G);
        T qv = T(0);
        cl_spmv<T>(D, G, aval, V.y, [&](int p, T sv) {
            V.v[p] = sv;
            qv += V.rh[p] * sv;
        });
        double s1[1] = {(double)qv};
        grp_sum<1>(G, s1);
        alpha = rho / s1[0];
        if (!isfinite(alpha)) break;   // breakdown: refine from the correction so far
        const T af = (T)alpha;
        T qs = T(0);
        for (int p = G.p0 + (int)threadIdx.x; p < G.p1; p += kClThreads) {
            V.d[p] += af * V.y[p];
            const T sv = V.r[p] - af * V.v[p];
            st_sh(V.s + p, sv);
            qs += sv * sv;
        }
        double s2[1] = {(double)qs};
        grp_sum<1>(G, s2);
        if (sqrt(s2[0]) <= tol * rin0) {
            total += 1;
            break;
        }
        cl_prec<T>(D, G, m, V.s, V.z, V.t, V.y);   // z = M s
        grp_sync(G);
        T qts = T(0), qtt = T(0);
        cl_spmv<T>(D, G, aval, V.z, [&](int p, T tv) {
            st_sh(V.t + p, tv);   // t doubles as a factor-sweep temporary
            qts += tv * V.s[p];
            qtt += tv * tv;
        });
        double s3[2] = {(double)qts, (double)qtt};
        grp_sum<2>(G, s3);
        omega = s3[0] / s3[1];
        if (!isfinite(omega) || omega == 0.0) {   // breakdown after h = d + alpha y: keep h
            total += 1;
            break;
        }
        const T of = (T)omega;
        T qhr = T(0), qrr = T(0);
        for (int p = G.p0 + (int)threadIdx.x; p < G.p1; p += kClThreads) {
            V.d[p] += of * V.z[p];
            const T rn = V.s[p] - of * V.t[p];
            st_sh(V.r + p, rn);
            qhr += V.rh[p] * rn;
            qrr += rn * rn;
        }
        double s4[2] = {(double)qhr, (double)qrr};
        grp_sum<2>(G, s4);
        const double rin = sqrt(s4[1]);
        total += 1;
        if (rin <= tol * rin0 || total >= max_inner || total >= inner_end) break;
        if (!(rin < 1e2 * rin0)) break;
        const double rho_new = s4[0];
        const double beta = (rho_new / rho) * (alpha / omega);
        rho = rho_new;
        bool restart = false;
        if (!(fabs(rho) > 1e-30 * rin0 * rin0)) {   // rho lost: restart with r^ = r
            restart = true;
            restarts += 1;
            rho = s4[1];
        }
        const T bf = (T)beta;
        for (int p = G.p0 + (int)threadIdx.x; p < G.p1; p += kClThreads) {
            const T rr = V.r[p];
            if (!restart) {
                st_sh(V.p + p, rr + bf * (V.p[p] - of * V.v[p]));
            } else {
                V.rh[p] = rr;
                st_sh(V.p + p, rr);
            }
        }
        grp_sync(G);
    }
}

// x_T += d over this workgroup's positions
template <class T>
__device__ __forceinline__ void cl_update(const TeamDesc& D, const Grp& G, double* x) {
    const T* d = cl_vecs<T>(D).d;
    for (int p = G.p0 + (int)threadIdx.x; p < G.p1; p += kClThreads) st_sh(x + p, x[p] + (double)d[p]);
}

__device__ __forceinline__ void cl_push(const TeamDesc& D, const Grp& G, int64_t& len, double e) {
    if (G.rank == 0 && threadIdx.x == 0 && len < D.hcap) D.hist[len] = e;
    len += 1;
}

// the refinement: FP64 residual and solution, inner iterations in FP32 (FP64 after a
// fallback); the same monotone acceptance as the single-workgroup mixed solver
__device__ void cl_solve(const TeamDesc& D, Grp& G, bool bicg, int m) {
    const int np = D.npad;
    double* b = D.w64;
    double* x = b + np;
    double* xa = x + np;   // the last accepted solution
    double q[1] = {0.0};
    for (int p = G.p0 + (int)threadIdx.x; p < G.p1; p += kClThreads) {
        const int row = D.perm[p];
        const double bb = D.b[row];
        b[p] = bb;
        st_sh(x + p, D.x[row]);
        q[0] += bb * bb;
    }
    grp_sum<1>(G, q);
    const double bnorm = sqrt(q[0]);
    int64_t hlen = 0;
    fvg_solver_info I{};
    if (bnorm < 1e-14) {   // x = 0 and one history entry (conjugate_gradient.rs:29-36)
        for (int p = G.p0 + (int)threadIdx.x; p < G.p1; p += kClThreads) D.x[D.perm[p]] = 0.0;
        grp_sync(G);
        cl_push(D, G, hlen, bnorm);
        I.initial_residual = I.final_residual = I.true_residual = bnorm;
        I.history_len = hlen;
    } else {
        const int64_t max_inner = D.max_iter > 0 ? D.max_iter : 1;
        double init_res = 0.0, true_res = 0.0, e_acc = 0.0;
        int64_t total = 0, outer = 0, restarts = 0;
        int64_t cap = 400;     // inner iterations per refinement step; shrinks after a rejected step
        bool check = false;    // x carries a correction not yet accepted
        // When the inner solves stop making progress the refinement escalates: FP32 with m
        // Jacobi sweeps per factor, FP32 with 4m, FP64 with 4m, FP64 with 16m; past the
        // last level it ends with the best x so far.
        int level = 0;
        bool used64 = false;
        int rejected = 0, weak = 0;
        for (int k = 0; k < 64; ++k) {
            const bool use64 = level >= 2;
            const double rr = use64 ? cl_residual<double>(D, G, b, x) : cl_residual<float>(D, G, b, x);
            const double e = sqrt(rr) / bnorm;
            if (outer == 0) init_res = e;
            outer += 1;
            if (check && !(e < e_acc)) {   // the correction raised the true residual: take it back
                for (int p = G.p0 + (int)threadIdx.x; p < G.p1; p += kClThreads) st_sh(x + p, xa[p]);
                grp_sync(G);
                check = false;
                cap = cap / 4 > 8 ? cap / 4 : 8;
                if (++rejected >= 2) {
                    if (level >= 3) break;
                    level += 1;
                    rejected = 0;
                    weak = 0;
                    cap = 400;
                }
                if (total >= max_inner) break;
                continue;
            }
            if (check) rejected = 0;
            bool remeasure = false;
            if (check && e > 0.5 * e_acc) {   // accepted, but barely progressing
                if (++weak >= 3) {
                    if (level >= 3) {
                        e_acc = e;
                        true_res = e;
                        cl_push(D, G, hlen, e);
                        check = false;
                        break;
                    }
                    level += 1;
                    weak = 0;
                    cap = 400;
                    remeasure = level == 2;   // the residual is wanted in the FP64 vectors now
                }
            } else {
                weak = 0;
            }
            e_acc = e;
            true_res = e;
            check = false;
            cl_push(D, G, hlen, e);
            if ((e < D.atol) || (e / init_res < D.rtol) || total >= max_inner) break;
            if (remeasure) cl_residual<double>(D, G, b, x);   // the same residual, into the FP64 vectors
            for (int p = G.p0 + (int)threadIdx.x; p < G.p1; p += kClThreads) xa[p] = x[p];
            const int64_t inner_end = total + cap;
            const double tol = cl_inner_target(D, e, init_res);
            const int ms = level == 0 ? m : level <= 2 ? 4 * m : 16 * m;
            if (level >= 2) {
                used64 = true;
                if (bicg)
                    cl_bicg<double>(D, G, ms, tol, total, inner_end, max_inner, restarts);
                else
                    cl_cg<double>(D, G, ms, tol, total, inner_end, max_inner);
                cl_update<double>(D, G, x);
            } else {
                if (bicg)
                    cl_bicg<float>(D, G, ms, tol, total, inner_end, max_inner, restarts);
                else
                    cl_cg<float>(D, G, ms, tol, total, inner_end, max_inner);
                cl_update<float>(D, G, x);
            }
            check = true;
            grp_sync(G);
        }
        const bool keep = check;   // the loop ended on an unjudged correction: keep the accepted x
        for (int p = G.p0 + (int)threadIdx.x; p < G.p1; p += kClThreads) D.x[D.perm[p]] = keep ? xa[p] : x[p];
        grp_sync(G);
        I.iterations = total;
        I.initial_residual = init_res;
        I.final_residual = true_res;
        I.true_residual = true_res;
        I.history_len = hlen;
        I.restarts = restarts;
        I.outer_iterations = outer;
        I.flags = used64 ? FVG_INFO_FP64_FALLBACK : 0;
    }
    if (G.rank == 0 && threadIdx.x == 0) {
        const bool aborted = G.K > 1 && __hip_atomic_load(&G.C->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        I.status = aborted ? FVG_E_TIMEOUT : FVG_OK;
        *D.info = I;
    }
}

// zero the barrier counters of `n` systems and the ticket counter
__global__ void k_cluster_reset(ClusterCtl* __restrict__ ctl, int n, ClusterHead* __restrict__ head) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        ctl[i].arrive = 0;
        ctl[i].abort = 0;
    }
    if (i == 0) head->next = 0;
}

template <bool BICG>
__global__ void __launch_bounds__(kClThreads)
    k_cluster(const TeamDesc* __restrict__ descs, const int32_t* __restrict__ groups, const int2* __restrict__ tickets,
              ClusterHead* __restrict__ head, ClusterCtl* __restrict__ ctl, int32_t sweeps, long long tmo) {
    __shared__ TeamDesc sdesc;
    __shared__ double lds[2 * 4 * kClWaves];
    __shared__ double bc[4];
    __shared__ int s_t, s_status;
    static_assert(sizeof(TeamDesc) % 8 == 0 && sizeof(TeamDesc) / 8 <= kClThreads, "descriptor copy");
    const int nticket = head->nticket;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) s_t = __hip_atomic_fetch_add(&head->next, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        const int t = s_t;
        if (t >= nticket) return;
        const int2 tk = tickets[t];
        const int s = tk.x;
        if (threadIdx.x < sizeof(TeamDesc) / 8)
            reinterpret_cast<uint64_t*>(&sdesc)[threadIdx.x] = reinterpret_cast<const uint64_t*>(descs + s)[threadIdx.x];
        if (threadIdx.x == 0) {
            // a factorisation that failed its checks is reported, not used
            const TeamDesc& E = descs[s];
            int st = FVG_OK;
            if (E.pflags && E.kind == kIC && E.pflags[2]) st = FVG_E_NOT_SYMMETRIC;
            if (E.pflags && E.kind == kILU && E.pflags[0]) st = FVG_E_ZERO_PIVOT;
            s_status = st;
        }
        __syncthreads();
        const TeamDesc& D = sdesc;
        Grp G;
        G.C = ctl + s;
        G.tmo = tmo;
        G.K = groups[s];
        G.rank = tk.y;
        G.p0 = (int)((int64_t)D.n * G.rank / G.K);
        G.p1 = (int)((int64_t)D.n * (G.rank + 1) / G.K);
        G.epoch = 0;
        G.red = 0;
        G.par = 0;
        G.lds = lds;
        G.bc = bc;
        if (s_status != FVG_OK) {
            if (G.rank == 0 && threadIdx.x == 0) {
                fvg_solver_info I{};
                I.status = s_status;
                *D.info = I;
            }
        } else {
            cl_solve(D, G, BICG, sweeps);
        }
        if (G.rank == 0) {
            // the cluster's writes (solution, info) reach memory before the completion flag
            __threadfence_system();
            __syncthreads();
            if (threadIdx.x == 0) __hip_atomic_store(D.done, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
}
