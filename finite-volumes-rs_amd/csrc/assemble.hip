// assemble.hip -- fvm::assembly::assemble (src/fvm/assembly.rs:11-79).
//
// Assembly is row-parallel: one thread per cell builds its whole CSR row and its
// right-hand side.  The thread starts from the cell terms, then applies every face the
// cell owns or neighbours in ascending face id, from the per-cell lists the mesh build
// made.  Each slot therefore receives its updates in the order of the reference's
// sequential face loop, so the assembled values are bit-identical to its.  An interior
// face's terms are evaluated once for each side (the same operations, the same bits).
// The equation arrives as a postfix program of reference terms evaluated per face /
// per cell in the reference's operation order.
#include <cstring>

#include "reduce.hpp"
#include "schemes.hpp"

namespace fvg {

constexpr int kMaxFields = 8;
constexpr int kMaxStack = 8;

struct Triple {
    double li, lj, r;
};
struct Pair {
    double l, r;
};
__device__ __forceinline__ Triple pick(bool c, const Triple& a, const Triple& b) {
    return Triple{c ? a.li : b.li, c ? a.lj : b.lj, c ? a.r : b.r};
}
__device__ __forceinline__ Pair pick(bool c, const Pair& a, const Pair& b) {
    return Pair{c ? a.l : b.l, c ? a.r : b.r};
}

// The evaluation stack of a term program.  Every thread runs the same program, so the
// stack depth is the same in every lane.  Each slot is read and written only as a
// member-wise select on the depth (no branch per slot, no computed address), so the
// stack stays in registers; a runtime-indexed private array would live in scratch.
template <class T, int S>
struct RegStack {
    T s[S];
    int sp = 0;
    __device__ __forceinline__ T at(int i) const {
        T v = s[0];
#pragma unroll
        for (int k = 1; k < S; ++k) v = pick(i == k, s[k], v);
        return v;
    }
    __device__ __forceinline__ void put(int i, const T& v) {
#pragma unroll
        for (int k = 0; k < S; ++k) s[k] = pick(i == k, v, s[k]);
    }
    __device__ __forceinline__ void push(const T& v) { put(sp++, v); }
    __device__ __forceinline__ T pop() { return at(--sp); }
};

// Backward differencing coefficients (schemes/time/backward.rs:24-29), powi(2) as x * x
struct BackwardCoef {
    double a_next, a_last, a_last_2;
};
__host__ __device__ inline BackwardCoef backward_coef(double dt, double dt_2) {
    const double wn = dt / dt_2;
    return BackwardCoef{(1.0 + 2.0 * wn) / (1.0 + wn), ((1.0 + wn) * (1.0 + wn)) / (1.0 + wn), (wn * wn) / (1.0 + wn)};
}

struct TermProgram {
    int n_ops;
    fvg_term_op ops[FVG_MAX_TERM_OPS];
    const double* face_fields[kMaxFields];
    const double* cell_fields[kMaxFields];
    MeshView geo;   // geometry for the Corrected, Linear and LimitedLinear schemes
};

// Term::face_terms of the whole expression for E faces f[j] (original ids) at once, so
// that each term's loads for all E faces are in flight together.  laplacian.rs:35-45 +
// orthogonal.rs:26-44 or corrected.rs:34-67, convection.rs:36-44 + upwind.rs:26-33,
// linear.rs:22-44 or limitedlinear.rs:24-72, divergence.rs:36-45, algebra operations.rs.
template <int D, int S, int E, bool G = true>
__device__ __forceinline__ void face_terms(const TermProgram& P, const int (&f)[E], const double (&tinv)[E],
                                           Triple (&out)[E]) {
    RegStack<Triple, S> st[E];
    for (int k = 0; k < P.n_ops; ++k) {
        const fvg_term_op o = P.ops[k];
        switch (o.op) {
        case FVG_OP_PARAM:   // a parameter of the word before
            break;
        case FVG_OP_USER_TERM:
#pragma unroll
            for (int j = 0; j < E; ++j) {
                if (o.arg0 >= 0)
                    st[j].push(Triple{P.face_fields[o.arg0][f[j]], P.face_fields[o.arg0 + 1][f[j]],
                                      P.face_fields[o.arg0 + 2][f[j]]});
                else
                    st[j].push(Triple{0.0, 0.0, 0.0});
            }
            break;
        case FVG_OP_LAPLACIAN_ORTHOGONAL:
#pragma unroll
            for (int j = 0; j < E; ++j) {
                const double t = tinv[j];
                const double l0 = 1.0 * (-t), l1 = 1.0 * t, r = 0.0;
                const double diff = P.face_fields[o.arg0][f[j]];
                st[j].push(Triple{diff * l0, diff * l1, -(diff * r)});
            }
            break;
        case FVG_OP_CONVECTION_UPWIND:
#pragma unroll
            for (int j = 0; j < E; ++j) {
                const double w = P.face_fields[o.arg1][f[j]];
                const double l0 = w > 0.0 ? 1.0 : 0.0, l1 = w > 0.0 ? 0.0 : 1.0, r = 0.0;
                const double flux = P.face_fields[o.arg0][f[j]];
                st[j].push(Triple{l0 * flux, l1 * flux, (-r) * flux});
            }
            break;
        case FVG_OP_DIVERGENCE:
#pragma unroll
            for (int j = 0; j < E; ++j) st[j].push(Triple{0.0, 0.0, -P.face_fields[o.arg0][f[j]]});
            break;
        case FVG_OP_LAPLACIAN_CORRECTED:
            if constexpr (!G) break;   // G = false: the program uses no scheme that reads geometry
#pragma unroll
            for (int j = 0; j < E; ++j) {
                double l0, l1, r;
                corrected_terms<D, 1>(P.geo, f[j], P.geo.owner[f[j]], P.geo.neigh[f[j]], P.cell_fields[o.arg1],
                                      o.scalar, l0, l1, &r);
                const double diff = P.face_fields[o.arg0][f[j]];
                st[j].push(Triple{diff * l0, diff * l1, -(diff * r)});
            }
            break;
        case FVG_OP_CONVECTION_LINEAR:
            if constexpr (!G) break;
#pragma unroll
            for (int j = 0; j < E; ++j) {
                const double w = linear_weight<D>(P.geo, f[j], P.geo.owner[f[j]], P.geo.neigh[f[j]]);
                const double l0 = 1.0 * w, l1 = 1.0 * (1.0 - w), r = 0.0;
                const double flux = P.face_fields[o.arg0][f[j]];
                st[j].push(Triple{l0 * flux, l1 * flux, (-r) * flux});
            }
            break;
        case FVG_OP_CONVECTION_LIMITED_LINEAR:
            if constexpr (!G) break;
#pragma unroll
            for (int j = 0; j < E; ++j) {
                const int nb = P.geo.neigh[f[j]];
                const double sflux = P.face_fields[o.arg1][f[j]];
                const double lim = P.face_fields[o.arg2][f[j]];
                const double w = linear_weight<D>(P.geo, f[j], P.geo.owner[f[j]], nb);
                // t[u] = wu lim + (1 - lim), t[d] = (1 - wu) lim with u the upwind side
                const bool up0 = sflux >= 0.0;
                const double wu = up0 ? w : 1.0 - w;
                const double tu = wu * lim + (1.0 - lim), td = (1.0 - wu) * lim;
                double l0, l1;
                if (nb >= 0) {
                    l0 = 1.0 * (up0 ? tu : td);
                    l1 = 1.0 * (up0 ? td : tu);
                } else if (sflux > 0.0) {   // boundary: outlet / inlet
                    l0 = 1.0;
                    l1 = 0.0;
                } else {
                    l0 = 0.0;
                    l1 = 1.0;
                }
                const double r = 0.0, flux = P.face_fields[o.arg0][f[j]];
                st[j].push(Triple{l0 * flux, l1 * flux, (-r) * flux});
            }
            break;
        case FVG_OP_ADD:
#pragma unroll
            for (int j = 0; j < E; ++j) {
                const Triple b = st[j].pop(), a = st[j].pop();
                st[j].push(Triple{a.li + b.li, a.lj + b.lj, a.r + b.r});
            }
            break;
        case FVG_OP_SUB:
#pragma unroll
            for (int j = 0; j < E; ++j) {
                const Triple b = st[j].pop(), a = st[j].pop();
                st[j].push(Triple{a.li - b.li, a.lj - b.lj, a.r - b.r});
            }
            break;
        case FVG_OP_NEG:
#pragma unroll
            for (int j = 0; j < E; ++j) {
                const Triple a = st[j].pop();
                st[j].push(Triple{-a.li, -a.lj, -a.r});
            }
            break;
        default:   // cell-only terms contribute the Term::face_terms default
#pragma unroll
            for (int j = 0; j < E; ++j) st[j].push(Triple{0.0, 0.0, 0.0});
            break;
        }
    }
#pragma unroll
    for (int j = 0; j < E; ++j) out[j] = st[j].at(0);
}

template <int D, int S>
__device__ __forceinline__ Triple face_terms(const TermProgram& P, int f, double tinv) {
    const int fs[1] = {f};
    const double ts[1] = {tinv};
    Triple t[1];
    face_terms<D, S, 1>(P, fs, ts, t);
    return t[0];
}

// Term::cell_terms of the whole expression for cell c.  source.rs:34-39, 72-77,
// terms/time.rs:31-34 + schemes/time/euler.rs:27-37, schemes/time/backward.rs:21-62.

template <int S>
__device__ __forceinline__ void cell_terms(const TermProgram& P, int c, double& l_out, double& r_out) {
    RegStack<Pair, S> st;
    for (int k = 0; k < P.n_ops; ++k) {
        const fvg_term_op o = P.ops[k];
        switch (o.op) {
        case FVG_OP_PARAM:   // a parameter of the word before
            break;
        case FVG_OP_SOURCE:
            st.push(Pair{0.0, -P.cell_fields[o.arg0][c]});
            break;
        case FVG_OP_LINEAR_SOURCE:
            st.push(Pair{P.cell_fields[o.arg0][c], 0.0});
            break;
        case FVG_OP_TIME_EULER: {
            double r0 = 1.0, r1 = 1.0;
            if (o.arg1 >= 0) {
                r0 = P.cell_fields[o.arg1][c];
                r1 = P.cell_fields[o.arg2][c];
            }
            const double dt = o.scalar;
            st.push(Pair{1.0 * (1.0 / dt) * r0, P.cell_fields[o.arg0][c] / dt * r1});
            break;
        }
        case FVG_OP_TIME_BACKWARD: {   // schemes/time/backward.rs:45-61
            const double dt = o.scalar;
            const BackwardCoef a = backward_coef(dt, P.ops[k + 1].scalar);
            double r0 = 1.0, r1 = 1.0, r2 = 1.0;
            if (o.arg2 >= 0) {
                r0 = P.cell_fields[o.arg2][c];
                r1 = P.cell_fields[o.arg2 + 1][c];
                r2 = P.cell_fields[o.arg2 + 2][c];
            }
            st.push(Pair{1.0 * (a.a_next / dt) * r0,
                         (P.cell_fields[o.arg0][c] * a.a_last * r1 - P.cell_fields[o.arg1][c] * a.a_last_2 * r2) / dt});
            break;
        }
        case FVG_OP_USER_TERM:
            if (o.arg1 >= 0)
                st.push(Pair{P.cell_fields[o.arg1][c], P.cell_fields[o.arg1 + 1][c]});
            else
                st.push(Pair{0.0, 0.0});
            break;
        case FVG_OP_ADD: {
            const Pair b = st.pop(), a = st.pop();
            st.push(Pair{a.l + b.l, a.r + b.r});
            break;
        }
        case FVG_OP_SUB: {
            const Pair b = st.pop(), a = st.pop();
            st.push(Pair{a.l - b.l, a.r - b.r});
            break;
        }
        case FVG_OP_NEG: {
            const Pair a = st.pop();
            st.push(Pair{-a.l, -a.r});
            break;
        }
        default:   // face-only terms contribute the Term::cell_terms default
            st.push(Pair{0.0, 0.0});
            break;
        }
    }
    const Pair t = st.at(0);
    l_out = t.l;
    r_out = t.r;
}

// Row c of the cell loop (assembly.rs:26-32: A_cc = l*vol, b_c = r*vol, every other
// slot of the row from zero) followed by the face loop (assembly.rs:35-76) restricted
// to the faces that update row c, in face order.
template <int D, int S>
__global__ void __launch_bounds__(kBlock) k_assemble_rows(TermProgram P, int n, const int32_t* __restrict__ row_ptr,
                                                          const int32_t* __restrict__ diagpos,
                                                          const double* __restrict__ vol,
                                                          const int32_t* __restrict__ ent_ptr,
                                                          const int2* __restrict__ ent,
                                                          const double* __restrict__ area,
                                                          const double* __restrict__ tinv,
                                                          const double* __restrict__ bc_i,
                                                          const double* __restrict__ bc_e, double* __restrict__ val,
                                                          double* __restrict__ rhs) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    double l, r;
    cell_terms<S>(P, c, l, r);
    const double v = vol[c];
    const int d = diagpos[c];
    for (int k = row_ptr[c]; k < row_ptr[c + 1]; ++k)
        if (k != d) val[k] = 0.0;
    double a_cc = l * v, b_c = r * v;
    for (int e = ent_ptr[c]; e < ent_ptr[c + 1]; ++e) {
        const int2 en = ent[e];
        const int f = en.x & kFaceMask, role = en.x >> kRoleShift;
        const Triple t = face_terms<D, S>(P, f, tinv[f]);
        const double A = area[f];
        const double li = t.li * A, lj = t.lj * A, rf = t.r * A;
        if (role == kRoleOwner) {          // lhs[i,i] += li; lhs[i,j] += lj; rhs[i] += r
            a_cc += li;
            val[en.y] += lj;
            b_c += rf;
        } else if (role == kRoleNeighbour) {   // lhs[j,i] -= li; lhs[j,j] -= lj; rhs[j] -= r
            val[en.y] -= li;
            a_cc -= lj;
            b_c -= rf;
        } else {                           // boundary: value_j = bndi * value_i + bnde
            a_cc += li;
            a_cc += lj * bc_i[en.y];
            b_c += rf - lj * bc_e[en.y];
        }
    }
    val[d] = a_cc;
    rhs[c] = b_c;
}

// The same assembly for meshes whose off-diagonal slots each belong to one face, with
// lanes over cell-face entries instead of rows, so that the entries' gathers are in
// flight at once.  A block owns `cpb` consecutive cells and walks their entries (which
// are consecutive in `ent`) kAsmEntriesPerLane block-widths at a time: each lane
// evaluates the face terms of its entries, sets their off-diagonal slots (0 + lj or
// 0 - li, the reference's update of a slot that starts at zero), and leaves their
// diagonal and right-hand-side contributions in LDS; then lane c of the block adds its
// cell's contributions to the cell terms, in ascending face id, as the row kernel does.
// x - y is x + (-y) in IEEE arithmetic, so the neighbour side's subtractions are stored
// negated.
constexpr int kAsmBlock = 256;
constexpr int kAsmSlots = 1024;   // CSR values one block stages in LDS (8 KiB)

template <int D, int S, int E, bool G>
__global__ void __launch_bounds__(kAsmBlock) k_assemble_entries(TermProgram P, int n, int cpb, int stage_cap,
                                                                const int32_t* __restrict__ row_ptr,
                                                                const int32_t* __restrict__ diagpos,
                                                                const double* __restrict__ vol,
                                                                const int32_t* __restrict__ ent_ptr,
                                                                const int2* __restrict__ ent,
                                                                const double2* __restrict__ tarea,
                                                                const double* __restrict__ bc_i,
                                                                const double* __restrict__ bc_e,
                                                                double* __restrict__ val, double* __restrict__ rhs) {
    constexpr int CH = E * kAsmBlock;
    __shared__ double s_a[CH], s_a2[CH], s_b[CH];
    __shared__ double s_val[kAsmSlots];
    __shared__ unsigned char s_bnd[CH];
    const int tid = threadIdx.x;
    // consecutive cell ranges on one XCD, so that neighbouring cells' face lines meet in
    // one L2
    const int c0 = xcd_tile(blockIdx.x, gridDim.x) * cpb;
    const int c1 = min(c0 + cpb, n);
    const int e0 = ent_ptr[c0], e1 = ent_ptr[c1];
    // the block's rows are one range of CSR values: built in LDS and written out in full
    // lines when it fits, else stored slot by slot
    const int r0 = row_ptr[c0], nslot = row_ptr[c1] - r0;
    const bool staged = nslot <= stage_cap;   // the same in every lane of the block
    auto put = [&](int slot, double v) {
        if (staged) s_val[slot - r0] = v;
        else val[slot] = v;
    };
    const int c = c0 + tid;
    const bool mine = c < c1;   // this lane sums cell c
    double a = 0.0, b = 0.0;
    int ecur = 0, eend = 0;
    if (mine) {
        double l, r;
        cell_terms<S>(P, c, l, r);
        const double v = vol[c];
        a = l * v;
        b = r * v;
        ecur = ent_ptr[c];
        eend = ent_ptr[c + 1];
    }
    for (int base = e0; base < e1; base += CH) {
        // E entries per lane, strided by the block width; lanes past the end evaluate
        // entry e0 again and discard it
        int2 en[E];
        int f[E];
        double tv[E], A[E];
#pragma unroll
        for (int j = 0; j < E; ++j) {
            const int e = base + tid + j * kAsmBlock;
            en[j] = ent[e < e1 ? e : e0];
            f[j] = en[j].x & kFaceMask;
        }
#pragma unroll
        for (int j = 0; j < E; ++j) {
            const double2 ta = tarea[f[j]];
            tv[j] = ta.x;
            A[j] = ta.y;
        }
        Triple t[E];
        face_terms<D, S, E, G>(P, f, tv, t);
#pragma unroll
        for (int j = 0; j < E; ++j) {
            if (base + tid + j * kAsmBlock >= e1) continue;
            const int k = tid + j * kAsmBlock, role = en[j].x >> kRoleShift;
            const double li = t[j].li * A[j], lj = t[j].lj * A[j], rf = t[j].r * A[j];
            if (role == kRoleOwner) {          // lhs[i,i] += li; lhs[i,j] += lj; rhs[i] += r
                put(en[j].y, 0.0 + lj);
                s_a[k] = li;
                s_b[k] = rf;
                s_bnd[k] = 0;
            } else if (role == kRoleNeighbour) {   // lhs[j,i] -= li; lhs[j,j] -= lj; rhs[j] -= r
                put(en[j].y, 0.0 - li);
                s_a[k] = -lj;
                s_b[k] = -rf;
                s_bnd[k] = 0;
            } else {                           // boundary: value_j = bndi * value_i + bnde
                s_a[k] = li;
                s_a2[k] = lj * bc_i[en[j].y];
                s_b[k] = rf - lj * bc_e[en[j].y];
                s_bnd[k] = 1;
            }
        }
        __syncthreads();
        if (mine) {
            const int stop = min(eend, base + CH);
            for (; ecur < stop; ++ecur) {
                const int k = ecur - base;
                a += s_a[k];
                if (s_bnd[k]) a += s_a2[k];
                b += s_b[k];
            }
        }
        __syncthreads();
    }
    if (mine) {
        put(diagpos[c], a);
        rhs[c] = b;
    }
    if (staged) {
        __syncthreads();
        for (int k = tid; k < nslot; k += kAsmBlock) val[r0 + k] = s_val[k];
    }
}

// cells per block of the entry kernel: on average one pass of entries per block (E per
// lane), at most one cell per lane
inline int assemble_cells_per_block(int64_t n_cells, int64_t n_ent, int E) {
    const int64_t per = n_ent > 0 ? (E * kAsmBlock * n_cells) / n_ent : kAsmBlock;
    return (int)std::max<int64_t>(1, std::min<int64_t>(kAsmBlock, per));
}

// stack-depth check so the device evaluator never over- or under-runs
static int validate_program(const fvg_term_op* prog, int n, int nff, int ncf, int* max_depth) {
    if (n <= 0 || n > FVG_MAX_TERM_OPS) return fail(FVG_E_INVALID, "term program length out of range");
    int depth = 0;
    *max_depth = 0;
    for (int k = 0; k < n; ++k) {
        const fvg_term_op& o = prog[k];
        switch (o.op) {
        case FVG_OP_LAPLACIAN_ORTHOGONAL:
        case FVG_OP_DIVERGENCE:
        case FVG_OP_CONVECTION_LINEAR:
            if (o.arg0 < 0 || o.arg0 >= nff) return fail(FVG_E_INVALID, "face field index out of range");
            ++depth;
            break;
        case FVG_OP_LAPLACIAN_CORRECTED:
            if (o.arg0 < 0 || o.arg0 >= nff) return fail(FVG_E_INVALID, "face field index out of range");
            if (o.arg1 < 0 || o.arg1 >= ncf) return fail(FVG_E_INVALID, "cell field index out of range");
            ++depth;
            break;
        case FVG_OP_CONVECTION_LIMITED_LINEAR:
            if (o.arg0 < 0 || o.arg0 >= nff || o.arg1 < 0 || o.arg1 >= nff || o.arg2 < 0 || o.arg2 >= nff)
                return fail(FVG_E_INVALID, "face field index out of range");
            ++depth;
            break;
        case FVG_OP_CONVECTION_UPWIND:
            if (o.arg0 < 0 || o.arg0 >= nff || o.arg1 < 0 || o.arg1 >= nff)
                return fail(FVG_E_INVALID, "face field index out of range");
            ++depth;
            break;
        case FVG_OP_SOURCE:
        case FVG_OP_LINEAR_SOURCE:
            if (o.arg0 < 0 || o.arg0 >= ncf) return fail(FVG_E_INVALID, "cell field index out of range");
            ++depth;
            break;
        case FVG_OP_TIME_EULER:
            if (o.arg0 < 0 || o.arg0 >= ncf) return fail(FVG_E_INVALID, "cell field index out of range");
            if (o.arg1 >= 0 && (o.arg1 >= ncf || o.arg2 < 0 || o.arg2 >= ncf))
                return fail(FVG_E_INVALID, "density field index out of range");
            ++depth;
            break;
        case FVG_OP_TIME_BACKWARD:
            if (o.arg0 < 0 || o.arg0 >= ncf || o.arg1 < 0 || o.arg1 >= ncf)
                return fail(FVG_E_INVALID, "cell field index out of range");
            if (o.arg2 != -1 && (o.arg2 < 0 || o.arg2 + 2 >= ncf))
                return fail(FVG_E_INVALID, "density field index out of range");
            if (k + 1 >= n || prog[k + 1].op != FVG_OP_PARAM)
                return fail(FVG_E_INVALID, "time(Backward) needs an FVG_OP_PARAM word with dt_2 after it");
            ++depth;
            ++k;   // its parameter word
            break;
        case FVG_OP_USER_TERM:
            if (o.arg0 != -1 && (o.arg0 < 0 || o.arg0 + 2 >= nff))
                return fail(FVG_E_INVALID, "user term face field index out of range");
            if (o.arg1 != -1 && (o.arg1 < 0 || o.arg1 + 1 >= ncf))
                return fail(FVG_E_INVALID, "user term cell field index out of range");
            ++depth;
            break;
        case FVG_OP_ADD:
        case FVG_OP_SUB:
            if (depth < 2) return fail(FVG_E_INVALID, "binary term operator needs two operands");
            --depth;
            break;
        case FVG_OP_NEG:
            if (depth < 1) return fail(FVG_E_INVALID, "negation needs an operand");
            break;
        default:
            return fail(FVG_E_INVALID, "unknown term opcode " + std::to_string(o.op));
        }
        if (depth > kMaxStack) return fail(FVG_E_INVALID, "term program too deep");
        *max_depth = std::max(*max_depth, depth);
    }
    if (depth != 1) return fail(FVG_E_INVALID, "term program must leave exactly one term");
    return FVG_OK;
}

struct LaunchArgs {
    const TermProgram& P;
    int n;
    const fvg_mesh* m;
    const double *bi, *be;
    double *val, *rhs;
    bool entries;
    int stage_cap;   // CSR values a block may stage in LDS (0: store slot by slot)
    hipStream_t stream;
};

// the kernels for a program whose evaluation stack needs at most S slots; G: it uses a
// scheme that reads face or cell geometry.  The entry kernel takes two entries per lane
// where the stack is shallow enough for the registers.
template <int S, bool G>
void launch_assemble(const LaunchArgs& L) {
    const fvg_mesh* m = L.m;
    if (L.entries) {
        constexpr int E = S <= 4 ? 2 : 1;
        const int cpb = assemble_cells_per_block(m->n_cells, m->n_ent, E);
        const dim3 grid((unsigned)((L.n + cpb - 1) / cpb));
        auto k = m->dim == 2 ? k_assemble_entries<2, S, E, G> : k_assemble_entries<3, S, E, G>;
        hipLaunchKernelGGL(k, grid, dim3(kAsmBlock), 0, L.stream, L.P, L.n, cpb, L.stage_cap, m->d_row_ptr,
                           m->d_diagpos, m->d_cell_vol, m->d_ent_ptr, m->d_ent, m->d_tarea, L.bi, L.be, L.val,
                           L.rhs);
    } else {
        auto k = m->dim == 2 ? k_assemble_rows<2, S> : k_assemble_rows<3, S>;
        hipLaunchKernelGGL(k, dim3(grid_for(L.n)), dim3(kBlock), 0, L.stream, L.P, L.n, m->d_row_ptr, m->d_diagpos,
                           m->d_cell_vol, m->d_ent_ptr, m->d_ent, m->d_area, m->d_tinv, L.bi, L.be, L.val, L.rhs);
    }
}

template <int S>
void launch_assemble(const LaunchArgs& L, bool geometry) {
    if (geometry) launch_assemble<S, true>(L);
    else launch_assemble<S, false>(L);
}

}  // namespace fvg

using namespace fvg;

extern "C" {

int fvg_assemble(fvg_ctx* c, const fvg_mesh* m, const fvg_term_op* program, int32_t n_ops,
                 const fvg_vec* const* face_fields, int32_t n_face_fields, const fvg_vec* const* cell_fields,
                 int32_t n_cell_fields, const fvg_vec* bc_i, const fvg_vec* bc_e, fvg_csr* lhs, fvg_vec* rhs) {
    if (!c || !m || !program || !lhs || !rhs) return fail(FVG_E_INVALID, "null argument");
    if (n_face_fields < 0 || n_face_fields > kMaxFields || n_cell_fields < 0 || n_cell_fields > kMaxFields)
        return fail(FVG_E_INVALID, "too many fields");
    int depth = 0;
    FVG_TRY(validate_program(program, n_ops, n_face_fields, n_cell_fields, &depth));
    if (lhs->pattern_id != m->pattern_id) return fail(FVG_E_INVALID, "lhs was not created from this mesh");
    if (rhs->n != m->n_cells) return fail(FVG_E_INVALID, "rhs length must equal the number of cells");
    if (m->n_bnd > 0 && (!bc_i || !bc_e || bc_i->n != m->n_bnd || bc_e->n != m->n_bnd))
        return fail(FVG_E_INVALID, "boundary closure arrays must have one entry per boundary face");
    TermProgram P{};
    P.n_ops = n_ops;
    for (int k = 0; k < n_ops; ++k) P.ops[k] = program[k];
    for (int k = 0; k < n_face_fields; ++k) {
        if (!face_fields[k] || face_fields[k]->n != m->n_faces)
            return fail(FVG_E_INVALID, "face field length must equal the number of faces");
        P.face_fields[k] = face_fields[k]->d;
    }
    // a cell field read as a Corrected gradient holds n_cells x dim entries, any other n_cells
    bool is_grad[kMaxFields] = {}, is_value[kMaxFields] = {};
    for (int k = 0; k < n_ops; ++k) {
        const fvg_term_op& o = program[k];
        if (o.op == FVG_OP_LAPLACIAN_CORRECTED) is_grad[o.arg1] = true;
        if (o.op == FVG_OP_SOURCE || o.op == FVG_OP_LINEAR_SOURCE) is_value[o.arg0] = true;
        if (o.op == FVG_OP_TIME_EULER) {
            is_value[o.arg0] = true;
            if (o.arg1 >= 0) is_value[o.arg1] = is_value[o.arg2] = true;
        }
        if (o.op == FVG_OP_TIME_BACKWARD) {
            is_value[o.arg0] = is_value[o.arg1] = true;
            if (o.arg2 >= 0) is_value[o.arg2] = is_value[o.arg2 + 1] = is_value[o.arg2 + 2] = true;
            ++k;
        }
        if (o.op == FVG_OP_USER_TERM && o.arg1 >= 0) is_value[o.arg1] = is_value[o.arg1 + 1] = true;
    }
    for (int k = 0; k < n_cell_fields; ++k) {
        if (is_grad[k] && is_value[k]) return fail(FVG_E_INVALID, "a cell field cannot be a value and a gradient");
        const int64_t want = is_grad[k] ? m->n_cells * m->dim : m->n_cells;
        if (!cell_fields[k] || cell_fields[k]->n != want)
            return fail(FVG_E_INVALID, is_grad[k] ? "a gradient field must hold n_cells x dim entries"
                                                  : "cell field length must equal the number of cells");
        P.cell_fields[k] = cell_fields[k]->d;
    }
    P.geo = mesh_view(m);
    const int n = (int)m->n_cells;
    const double* bi = m->n_bnd ? bc_i->d : nullptr;
    const double* be = m->n_bnd ? bc_e->d : nullptr;
    const LaunchArgs L{P, n, m, bi, be, lhs->val, rhs->d, m->unique_slots && c->tune_assemble != 1,
                       c->tune_assemble == 2 ? 0 : kAsmSlots, c->stream};
    bool geometry = false;
    for (int k = 0; k < n_ops; ++k)
        geometry |= program[k].op == FVG_OP_LAPLACIAN_CORRECTED || program[k].op == FVG_OP_CONVECTION_LINEAR ||
                    program[k].op == FVG_OP_CONVECTION_LIMITED_LINEAR;
    if (depth <= 2) launch_assemble<2>(L, geometry);
    else if (depth <= 4) launch_assemble<4>(L, geometry);
    else launch_assemble<kMaxStack>(L, geometry);
    FVG_HIP(hipGetLastError());
    lhs->version += 1;
    return FVG_OK;
}

}  // extern "C"
