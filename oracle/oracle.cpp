// oracle.cpp -- CPU FP64 restatement of the finite-volumes assemble-and-solve path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker the GPU library is compared
// against; it is loaded by tests/, by __graft_entry__.smoke() and by the CPU-baseline
// leg of bench.py, and by nothing else.  The product (libfvg.so) never links it.
//
// Every function restates the reference crate's Rust code in the same operation order
// (sequential sums, reciprocal-multiply, no fused multiply-add: build with
// -ffp-contract=off), citing the file:line it follows.  File paths are relative to the
// root of the finite-volumes-rs crate.
//
// Exported C functions (prefix orc_) are bound from Python through ctypes.

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <set>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <type_traits>
#include <vector>

namespace orc {

// ---------------------------------------------------------------------------------
// Sparsity<T>  (src/core/sparsity.rs:8-161)
// ---------------------------------------------------------------------------------
struct Sparsity {
    std::vector<size_t> minors;
    std::vector<size_t> major_starts{0};
    size_t max_minor = 0;

    void push_to_major(size_t m) {                       // sparsity.rs:25-28
        minors.push_back(m);
        max_minor = std::max(max_minor, m);
    }
    void close_major() { major_starts.push_back(minors.size()); }   // :31-33
    size_t major_start(size_t i) const { return major_starts[i]; }  // :41-43
    size_t major_end(size_t i) const { return major_starts[i + 1]; }
    size_t major_len() const { return major_starts.size() - 1; }
    size_t minor_len() const { return minors.size(); }
    // sparsity.rs:102-111 (linear search)
    long find_flat_index(size_t major, size_t minor) const {
        for (size_t k = major_start(major); k < major_end(major); ++k)
            if (minors[k] == minor) return (long)k;
        return -1;
    }
    // sparsity.rs:123-138: per-row sort
    void sort() {
        for (size_t i = 0; i < major_len(); ++i)
            std::sort(minors.begin() + major_start(i), minors.begin() + major_end(i));
    }
    // sparsity.rs:141-157: per-row sort carrying values (stable, like Rust's sort_by)
    template <class D>
    void sort_with(std::vector<D>& values) {
        for (size_t i = 0; i < major_len(); ++i) {
            size_t rs = major_start(i), re = major_end(i);
            std::vector<std::pair<size_t, D>> merged;
            merged.reserve(re - rs);
            for (size_t k = rs; k < re; ++k) merged.emplace_back(minors[k], values[k]);
            std::stable_sort(merged.begin(), merged.end(),
                             [](const auto& a, const auto& b) { return a.first < b.first; });
            for (size_t k = rs; k < re; ++k) {
                minors[k] = merged[k - rs].first;
                values[k] = merged[k - rs].second;
            }
        }
    }
};

// ---------------------------------------------------------------------------------
// Small vector  (src/core/vector.rs)
// ---------------------------------------------------------------------------------
template <int N>
struct Vec {
    double d[N];
    Vec() { for (int i = 0; i < N; ++i) d[i] = 0.0; }
    double& operator[](int i) { return d[i]; }
    double operator[](int i) const { return d[i]; }
};
template <int N> Vec<N> operator+(Vec<N> a, const Vec<N>& b) { for (int i = 0; i < N; ++i) a[i] += b[i]; return a; }   // vector.rs:195-203
template <int N> Vec<N> operator-(Vec<N> a, const Vec<N>& b) { for (int i = 0; i < N; ++i) a[i] -= b[i]; return a; }   // :206-214
template <int N> Vec<N> operator*(Vec<N> a, double s) { for (int i = 0; i < N; ++i) a[i] *= s; return a; }             // :216-224
template <int N> Vec<N> operator/(Vec<N> a, double s) { for (int i = 0; i < N; ++i) a[i] /= s; return a; }             // :227-235
template <int N> double dot(const Vec<N>& a, const Vec<N>& b) {   // vector.rs:37-43
    double x = 0.0;
    for (int i = 0; i < N; ++i) x += a[i] * b[i];
    return x;
}
template <int N> double norm(const Vec<N>& a) { return std::sqrt(dot(a, a)); }  // vector.rs:45-47

// Normal::normal  (vector.rs:301-329)
template <int N> Vec<N> normal_of(const Vec<N>& s, const Vec<N>& r) {
    Vec<N> n;
    if (N == 1) {
        n[0] = 1.0;
    } else if (N == 2) {
        n[0] = -s[1];
        n[1] = s[0];
    } else {
        n[0] = s[1 % N] * r[2 % N] - s[2 % N] * r[1 % N];
        n[1 % N] = s[2 % N] * r[0] - s[0] * r[2 % N];
        n[2 % N] = s[0] * r[1 % N] - s[1 % N] * r[0];
        for (int i = 0; i < N; ++i) n[i] *= 0.5;
    }
    return n;
}

// ---------------------------------------------------------------------------------
// Mesh<DIM>  (src/core/mesh/mod.rs, src/core/mesh/compute.rs)
// ---------------------------------------------------------------------------------
// Matrix<N, N>  (src/core/matrix.rs): rows of Vector<N>
// ---------------------------------------------------------------------------------
template <int N>
struct Mat {
    Vec<N> r[N];
    Mat() { for (int i = 0; i < N; ++i) for (int j = 0; j < N; ++j) r[i][j] = 0.0; }
};

// Vector::outer (matrix.rs:105-115)
template <int N>
Mat<N> outer(const Vec<N>& a, const Vec<N>& b) {
    Mat<N> m;
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) m.r[i][j] = a[i] * b[j];
    return m;
}

// Matrix += Matrix (matrix.rs:160-168)
template <int N>
void add_to(Mat<N>& a, const Mat<N>& b) {
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) a.r[i][j] += b.r[i][j];
}

// Matrix * Vector (matrix.rs:221-230): one dot product per row
template <int N>
Vec<N> matvec(const Mat<N>& a, const Vec<N>& v) {
    Vec<N> out;
    for (int i = 0; i < N; ++i) out[i] = dot(a.r[i], v);
    return out;
}

// Matrix::inv (matrix.rs:320-383): Gauss-Jordan without pivoting unless a pivot is
// below 1e-14, then a swap with row `maxi`, where maxi is the largest entry of ROW k
// right of the diagonal (the reference's search, kept as written).  False when the
// reference returns Err(SingularMatrix).
template <int N>
bool inverse(Mat<N> a, Mat<N>& out) {
    out = Mat<N>();
    for (int i = 0; i < N; ++i) out.r[i][i] = 1.0;
    for (int k = 0; k < N - 1; ++k) {
        double piv = a.r[k][k];
        if (std::fabs(piv) < 1e-14) {
            int maxi = k;
            double maxp = std::fabs(piv);
            for (int j = k + 1; j < N; ++j) {
                double pj = std::fabs(a.r[k][j]);
                if (pj > maxp) { maxi = j; maxp = pj; }
            }
            if (maxi == k) return false;
            std::swap(a.r[k], a.r[maxi]);
            std::swap(out.r[k], out.r[maxi]);
        }
        piv = a.r[k][k];
        for (int i = k + 1; i < N; ++i) {
            double f = -a.r[i][k] / piv;
            a.r[i] = a.r[i] + a.r[k] * f;
            out.r[i] = out.r[i] + out.r[k] * f;
        }
    }
    for (int k = 0; k < N; ++k) {
        double piv = a.r[k][k];
        if (std::fabs(piv) < 1e-14) return false;
        a.r[k] = a.r[k] / piv;
        out.r[k] = out.r[k] / piv;
    }
    for (int k = N - 1; k >= 1; --k) {
        double piv = a.r[k][k];
        for (int i = 0; i < k; ++i) {
            double f = -a.r[i][k] / piv;
            a.r[i] = a.r[i] + a.r[k] * f;
            out.r[i] = out.r[i] + out.r[k] * f;
        }
    }
    return true;
}

// ---------------------------------------------------------------------------------
constexpr size_t NONE = (size_t)-1;

template <int DIM>
struct FaceData {
    double area = 0.0;
    Vec<DIM> center, normal;
    size_t owner = NONE;
    int kind = 0;          // 0 none, 1 cell, 2 boundary  (InternalFaceNeighbor, mod.rs:117-122)
    size_t neighbor = 0;   // cell index or patch id
};
template <int DIM>
struct CellData {
    double volume = 0.0;
    Vec<DIM> center;
};

template <int DIM>
struct Mesh {
    std::vector<Vec<DIM>> nodes;
    Sparsity face_nodes, cell_faces, cell_to_cell;
    std::vector<FaceData<DIM>> face_data;
    std::vector<CellData<DIM>> cell_data;
    std::vector<Vec<DIM>> cell_face_grad;   // per cell-face listing (compute.rs:156)
    std::vector<Vec<DIM>> cell_diag_grad;   // per cell (compute.rs:157)

    void add_node(const Vec<DIM>& p) { nodes.push_back(p); }   // mod.rs:444-449

    // mod.rs:451-476: boundary faces are tagged now, interior faces get their
    // neighbour from the second add_cell that names them.
    void add_face(const int32_t* nd, size_t len, int32_t patch) {
        for (size_t k = 0; k < len; ++k) face_nodes.push_to_major((size_t)nd[k]);
        face_nodes.close_major();
        FaceData<DIM> fd;
        if (patch >= 0) { fd.kind = 2; fd.neighbor = (size_t)patch; }
        face_data.push_back(fd);
    }

    // mod.rs:486-527
    void add_cell(const int32_t* fc, size_t len) {
        for (size_t k = 0; k < len; ++k) cell_faces.push_to_major((size_t)fc[k]);
        cell_faces.close_major();
        cell_data.push_back(CellData<DIM>());
        size_t me = cell_data.size() - 1;
        for (size_t k = 0; k < len; ++k) {
            FaceData<DIM>& f = face_data[(size_t)fc[k]];
            if (f.owner == NONE) {
                f.owner = me;
            } else if (f.kind == 0) {
                f.kind = 1;
                f.neighbor = me;
            }
        }
    }

    size_t n_cells() const { return cell_faces.major_len(); }
    size_t n_faces() const { return face_nodes.major_len(); }

    // compute.rs:267-321
    void calc_face_data(size_t f, const Vec<DIM>& node_ave) {
        FaceData<DIM>& fd = face_data[f];
        size_t s = face_nodes.major_start(f), len = face_nodes.major_end(f) - s;
        if (DIM == 2) {
            if (len != 2) throw std::runtime_error("2D face must have two nodes");
            Vec<DIM> n0 = nodes[face_nodes.minors[s]];
            Vec<DIM> n1 = nodes[face_nodes.minors[s + 1]];
            Vec<DIM> dn = n1 - n0;
            Vec<DIM> dsn = normal_of(dn, dn);
            fd.area = norm(dsn);
            fd.normal = dsn / fd.area;
            fd.center = (n0 + n1) * 0.5;
        } else {
            Vec<DIM> center, normal;
            double area = 0.0;
            for (size_t i = 0; i < len; ++i) {
                Vec<DIM> n0 = nodes[face_nodes.minors[s + i]];
                Vec<DIM> n1 = nodes[face_nodes.minors[s + (i == len - 1 ? 0 : i + 1)]];
                Vec<DIM> dn0 = n0 - node_ave;
                Vec<DIM> dn1 = n1 - node_ave;
                Vec<DIM> dsn = normal_of(dn0, dn1);
                double area_i = norm(dsn);
                area += area_i;
                normal = dsn;                                  // compute.rs:305 (assignment)
                center = center + ((n0 + n1 + node_ave) * 1.0 / 3.0) * area_i;
            }
            if (!(area > 2.220446049250313e-16)) throw std::runtime_error("degenerate face");
            normal = normal / area;
            center = center / area;
            fd.area = area;
            fd.center = center;
            fd.normal = normal;
        }
    }

    // compute.rs:325-375
    void calc_cell_data(size_t c, const Vec<DIM>& facec_ave) {
        Vec<DIM> center;
        double volume = 0.0;
        for (size_t k = cell_faces.major_start(c); k < cell_faces.major_end(c); ++k) {
            const FaceData<DIM>& face = face_data[cell_faces.minors[k]];
            Vec<DIM> dfc = face.center - facec_ave;
            double height = std::fabs(dot(dfc, face.normal));
            double volume_i, w;
            Vec<DIM> center_i;
            if (DIM == 2) {
                volume_i = height * face.area / 2.0;
                center_i = (face.center * 2.0 + facec_ave) / 3.0;
            } else {
                volume_i = height * face.area / 3.0;
                center_i = (face.center * 3.0 + facec_ave) / 4.0;
            }
            (void)w;
            center = center + center_i * volume_i;
            volume += volume_i;
        }
        cell_data[c].volume = volume;
        cell_data[c].center = center / volume;
    }

    // compute.rs:12-152 (serial: every cell and face owned)
    void compute() {
        for (size_t f = 0; f < n_faces(); ++f) {                       // :15-18
            if (face_data[f].owner == NONE || face_data[f].owner >= n_cells())
                throw std::runtime_error("face without owner cell");
            if (face_data[f].kind == 0)                                 // :21-31
                throw std::runtime_error("face without neighbour in a serial mesh");
        }
        for (size_t f = 0; f < n_faces(); ++f) {                       // :35-45
            size_t s = face_nodes.major_start(f), size = face_nodes.major_end(f) - s;
            Vec<DIM> c;
            for (size_t k = 0; k < size; ++k) c = c + nodes[face_nodes.minors[s + k]];
            c = c / (double)size;
            calc_face_data(f, c);
        }
        for (size_t c = 0; c < n_cells(); ++c) {                       // :52-69
            size_t s = cell_faces.major_start(c), size = cell_faces.major_end(c) - s;
            Vec<DIM> center;
            for (size_t k = 0; k < size; ++k) center = center + face_data[cell_faces.minors[s + k]].center;
            center = center / (double)size;
            calc_cell_data(c, center);
        }
        for (size_t f = 0; f < n_faces(); ++f) {                       // :97-131
            Vec<DIM> inside = cell_data[face_data[f].owner].center;
            if (dot(face_data[f].normal, face_data[f].center - inside) < 0.0)
                face_data[f].normal = face_data[f].normal * -1.0;
        }
        for (size_t c = 0; c < n_cells(); ++c) {                       // :135-152
            std::set<size_t> others;   // HashSet in the reference; order is removed by the later sort
            for (size_t k = cell_faces.major_start(c); k < cell_faces.major_end(c); ++k) {
                const FaceData<DIM>& fd = face_data[cell_faces.minors[k]];
                if (fd.kind == 1) others.insert(fd.neighbor == c ? fd.owner : fd.neighbor);  // FaceRef::other_cell mod.rs:326-338
            }
            cell_to_cell.push_to_major(c);
            for (size_t o : others) cell_to_cell.push_to_major(o);
            cell_to_cell.close_major();
        }
        // compute.rs:154-200: least-squares gradient coefficients.  For every listing k
        // of cell c, dgc is the vector to the other cell's centre (or to the face centre
        // on a boundary); g = sum dgc dgc^T, coef_k = g^-1 dgc, diag_c = -sum coef_k.
        cell_face_grad.assign(cell_faces.minor_len(), Vec<DIM>());
        cell_diag_grad.assign(n_cells(), Vec<DIM>());
        for (size_t c = 0; c < n_cells(); ++c) {
            Mat<DIM> g;
            const Vec<DIM> xc = cell_data[c].center;
            auto dgc_of = [&](size_t k) {
                const FaceData<DIM>& fd = face_data[cell_faces.minors[k]];
                if (fd.kind == 2) return fd.center - xc;
                size_t co = fd.neighbor == c ? fd.owner : fd.neighbor;
                return cell_data[co].center - xc;
            };
            for (size_t k = cell_faces.major_start(c); k < cell_faces.major_end(c); ++k) {
                Vec<DIM> dgc = dgc_of(k);
                add_to(g, outer(dgc, dgc));
            }
            Mat<DIM> ginv;
            if (!inverse(g, ginv)) throw std::runtime_error("singular gradient matrix");
            for (size_t k = cell_faces.major_start(c); k < cell_faces.major_end(c); ++k) {
                Vec<DIM> gi = matvec(ginv, dgc_of(k));
                cell_face_grad[k] = gi;
                cell_diag_grad[c] = cell_diag_grad[c] - gi;
            }
        }
    }
};

// ---------------------------------------------------------------------------------
// DistributedMatrix<f64>  (src/linalg/dmatrix.rs)
// ---------------------------------------------------------------------------------
struct Matrix {
    Sparsity sp;
    std::vector<double> values;
    size_t nrows = 0, ncols = 0;

    // dmatrix.rs:38-50
    static Matrix from_cut_sparsity(const Sparsity& s, size_t nrows) {
        Matrix m;
        for (size_t i = 0; i < nrows; ++i) {
            for (size_t k = s.major_start(i); k < s.major_end(i); ++k) m.sp.push_to_major(s.minors[k]);
            m.sp.close_major();
        }
        m.ncols = m.sp.max_minor + 1;
        m.values.assign(m.sp.minor_len(), 0.0);
        m.sp.sort();
        m.nrows = nrows;
        return m;
    }
    // dmatrix.rs:53-58
    static Matrix from_sparsity_and_values(Sparsity s, std::vector<double> v) {
        Matrix m;
        m.nrows = s.major_len();
        m.ncols = s.max_minor + 1;
        s.sort_with(v);
        m.sp = std::move(s);
        m.values = std::move(v);
        return m;
    }
    // binary search inside a sorted row: Index<[usize;2]> dmatrix.rs:262-286
    long pos(size_t i, size_t j) const {
        if (i >= sp.major_len()) return -1;
        auto b = sp.minors.begin() + sp.major_start(i), e = sp.minors.begin() + sp.major_end(i);
        auto it = std::lower_bound(b, e, j);
        if (it == e || *it != j) return -1;
        return (long)(it - sp.minors.begin());
    }
    double& at(size_t i, size_t j) {
        long p = pos(i, j);
        if (p < 0) throw std::runtime_error("Sparse matrix index not found");
        return values[(size_t)p];
    }
    double index(size_t i, size_t j) const {   // Index<[usize; 2]> (dmatrix.rs:262-276): panics if absent
        long p = pos(i, j);
        if (p < 0) throw std::runtime_error("Sparse matrix index not found");
        return values[(size_t)p];
    }
    double get(size_t i, size_t j) const {   // dmatrix.rs:105-117
        long p = pos(i, j);
        return p < 0 ? 0.0 : values[(size_t)p];
    }
    // dmatrix.rs:136-146
    void imul(std::vector<double>& result, const std::vector<double>& rhs) const {
        for (size_t i = 0; i < nrows; ++i) {
            result[i] = 0.0;
            for (size_t k = sp.major_start(i); k < sp.major_end(i); ++k)
                result[i] += values[k] * rhs[sp.minors[k]];
        }
    }
    // dmatrix.rs:214-229
    Matrix square_block() const {
        Sparsity s;
        std::vector<double> v;
        for (size_t i = 0; i < nrows; ++i) {
            for (size_t k = sp.major_start(i); k < sp.major_end(i); ++k) {
                if (sp.minors[k] < nrows) { s.push_to_major(sp.minors[k]); v.push_back(values[k]); }
            }
            s.close_major();
        }
        return from_sparsity_and_values(std::move(s), std::move(v));
    }
    // dmatrix.rs:76-88 with ApproximateCmp for f64 (linalg/mod.rs:73-77)
    bool is_symmetric() const {
        for (size_t i = 0; i < nrows; ++i)
            for (size_t k = sp.major_start(i); k < sp.major_end(i); ++k) {
                size_t j = sp.minors[k];
                if (j < nrows) {
                    long p = pos(j, i);
                    if (p < 0) throw std::runtime_error("Sparse matrix index not found");
                    if (!(std::fabs(values[k] - values[(size_t)p]) < 1e-14)) return false;
                }
            }
        return true;
    }
};

// ---------------------------------------------------------------------------------
// Terms and assembly (src/fvm/terms/*, src/fvm/schemes/*, src/fvm/assembly.rs)
// The expression tree arrives as a postfix program; each opcode is one reference term.
// ---------------------------------------------------------------------------------
enum {
    OP_LAPLACIAN_ORTHOGONAL = 1,
    OP_CONVECTION_UPWIND = 2,
    OP_SOURCE = 3,
    OP_LINEAR_SOURCE = 4,
    OP_TIME_EULER = 5,
    OP_DIVERGENCE = 6,
    OP_LAPLACIAN_CORRECTED = 7,
    OP_CONVECTION_LINEAR = 8,
    OP_CONVECTION_LIMITED_LINEAR = 9,
    OP_TIME_BACKWARD = 10,   // arg0 previous, arg1 previous_2, arg2 first density or -1, scalar dt
    OP_USER_TERM = 11,       // another impl Term: arg0 = face fields li, lj, r (or -1), arg1 = cell fields l, r (or -1)
    OP_PARAM = 15,           // dt_2 of the OP_TIME_BACKWARD before it
    OP_ADD = 16,
    OP_SUB = 17,
    OP_NEG = 18,
};
struct TermOp { int32_t op, arg0, arg1, arg2; double scalar; };

struct Triple { double li, lj, r; };
struct Pair { double l, r; };

// FaceNormalGradientScheme::terms for a value with nc components: (gi, gj, grhs[nc]).
// Orthogonal (facengrad/orthogonal.rs:26-44) when grads is null, else
// Corrected(grads, reduction) (facengrad/corrected.rs:34-67) with grads holding nc x DIM
// entries per cell; for a vector value the gradient is a Matrix<nc, DIM> and every
// operation acts row by row.
template <int DIM>
void fngrad_terms(const Mesh<DIM>& mesh, size_t f, int nc, const double* grads, double reduction, double& gi,
                  double& gj, double* grhs) {
    const FaceData<DIM>& fd = mesh.face_data[f];
    const Vec<DIM> ci = mesh.cell_data[fd.owner].center;
    if (!grads) {
        double dx = fd.kind == 1 ? norm(mesh.cell_data[fd.neighbor].center - ci) : norm(fd.center - ci);
        double t = 1.0 / dx;
        gi = 1.0 * (-t);
        gj = 1.0 * t;
        for (int c = 0; c < nc; ++c) grhs[c] = 0.0;
        return;
    }
    auto grad_row = [&](size_t cell, int c) {
        Vec<DIM> g;
        for (int d = 0; d < DIM; ++d) g[d] = grads[(cell * nc + c) * DIM + d];
        return g;
    };
    const Vec<DIM> n = fd.normal;
    if (fd.kind == 1) {
        Vec<DIM> d = mesh.cell_data[fd.neighbor].center - ci;
        double c_corr = 1.0 / dot(n, d);
        double w = dot(fd.center - ci, n) / dot(d, n);
        for (int c = 0; c < nc; ++c) {
            Vec<DIM> gf = grad_row(fd.owner, c) * (1.0 - w) + grad_row(fd.neighbor, c) * w;
            double expl = dot(gf, n - d * c_corr);
            grhs[c] = expl * reduction;
        }
        gi = 1.0 * (-c_corr);
        gj = 1.0 * c_corr;
    } else {
        Vec<DIM> d = fd.center - ci;
        double c_corr = 1.0 / dot(n, d);
        for (int c = 0; c < nc; ++c) {
            double expl = dot(grad_row(fd.owner, c), n - d * c_corr);
            grhs[c] = expl * reduction;
        }
        gi = 1.0 * (-c_corr);
        gj = 1.0 * c_corr;
    }
}

// Linear interpolation weight (faceinterp/linear.rs:22-44, limitedlinear.rs:46-55)
template <int DIM>
double linear_weight(const Mesh<DIM>& mesh, size_t f) {
    const FaceData<DIM>& fd = mesh.face_data[f];
    const Vec<DIM> fc = fd.center, cc = mesh.cell_data[fd.owner].center;
    const Vec<DIM> nc = fd.kind == 1 ? mesh.cell_data[fd.neighbor].center : fd.center;
    return dot(fd.normal, nc - fc) / std::fmax(dot(fd.normal, nc - cc), 1e-14);
}

template <int DIM>
struct Assembler {
    const Mesh<DIM>& mesh;
    const TermOp* prog;
    int nops;
    const double* const* face_fields;
    const double* const* cell_fields;

    // Term::face_terms of the whole expression for face f
    Triple face_terms(size_t f) const {
        Triple st[32];
        int sp = 0;
        const FaceData<DIM>& fd = mesh.face_data[f];
        for (int k = 0; k < nops; ++k) {
            const TermOp& o = prog[k];
            switch (o.op) {
            case OP_LAPLACIAN_ORTHOGONAL: {
                // orthogonal.rs:26-44 then laplacian.rs:35-45
                double dx;
                if (fd.kind == 1) dx = norm(mesh.cell_data[fd.neighbor].center - mesh.cell_data[fd.owner].center);
                else dx = norm(fd.center - mesh.cell_data[fd.owner].center);
                double t = 1.0 / dx;
                double l0 = 1.0 * (-t), l1 = 1.0 * t, r = 0.0;
                double diff = face_fields[o.arg0][f];
                st[sp++] = {diff * l0, diff * l1, -(diff * r)};
                break;
            }
            case OP_CONVECTION_UPWIND: {
                // upwind.rs:26-33 then convection.rs:36-44; arg0 = convected flux, arg1 = upwind flux
                double uflux = face_fields[o.arg1][f];
                double l0, l1, r = 0.0;
                if (uflux > 0.0) { l0 = 1.0; l1 = 0.0; } else { l0 = 0.0; l1 = 1.0; }
                double flux = face_fields[o.arg0][f];
                st[sp++] = {l0 * flux, l1 * flux, (-r) * flux};
                break;
            }
            case OP_DIVERGENCE: {   // divergence.rs:36-45
                double flux = face_fields[o.arg0][f];
                st[sp++] = {0.0, 0.0, -flux};
                break;
            }
            case OP_LAPLACIAN_CORRECTED: {
                // corrected.rs:34-67 then laplacian.rs:35-45; arg1 = cell gradient field
                // (n x DIM), scalar = corrector_reduction
                double l0, l1, r;
                fngrad_terms(mesh, f, 1, cell_fields[o.arg1], o.scalar, l0, l1, &r);
                double diff = face_fields[o.arg0][f];
                st[sp++] = {diff * l0, diff * l1, -(diff * r)};
                break;
            }
            case OP_CONVECTION_LINEAR: {   // linear.rs:22-44 then convection.rs:36-44
                double w = linear_weight(mesh, f);
                double l0 = 1.0 * w, l1 = 1.0 * (1.0 - w), r = 0.0;
                double flux = face_fields[o.arg0][f];
                st[sp++] = {l0 * flux, l1 * flux, (-r) * flux};
                break;
            }
            case OP_CONVECTION_LIMITED_LINEAR: {
                // limitedlinear.rs:24-72 then convection.rs:36-44; arg0 = convected flux,
                // arg1 = the scheme's flux (upwind side), arg2 = face limiters
                double sflux = face_fields[o.arg1][f];
                int u = sflux >= 0.0 ? 0 : 1, dn = u == 0 ? 1 : 0;
                double t[2] = {0.0, 0.0};
                double lim = face_fields[o.arg2][f];
                double w = linear_weight(mesh, f);
                double wu = u == 0 ? w : 1.0 - w;
                t[u] = wu * lim + (1.0 - lim);
                t[dn] = (1.0 - wu) * lim;
                double l0, l1, r = 0.0;
                if (fd.kind == 1) { l0 = 1.0 * t[0]; l1 = 1.0 * t[1]; }
                else if (sflux > 0.0) { l0 = 1.0; l1 = 0.0; }
                else { l0 = 0.0; l1 = 1.0; }
                double flux = face_fields[o.arg0][f];
                st[sp++] = {l0 * flux, l1 * flux, (-r) * flux};
                break;
            }
            case OP_USER_TERM:   // any other impl Term (terms/mod.rs:24-40): its face_terms as given
                if (o.arg0 >= 0) st[sp++] = {face_fields[o.arg0][f], face_fields[o.arg0 + 1][f], face_fields[o.arg0 + 2][f]};
                else st[sp++] = {0.0, 0.0, 0.0};
                break;
            case OP_PARAM:
                break;
            case OP_SOURCE: case OP_LINEAR_SOURCE: case OP_TIME_EULER: case OP_TIME_BACKWARD:
                st[sp++] = {0.0, 0.0, 0.0};   // Term::face_terms default (terms/mod.rs:32-34)
                break;
            case OP_ADD: { Triple b = st[--sp], a = st[--sp]; st[sp++] = {a.li + b.li, a.lj + b.lj, a.r + b.r}; break; }   // operations.rs:34-41, 56-62
            case OP_SUB: { Triple b = st[--sp], a = st[--sp]; st[sp++] = {a.li - b.li, a.lj - b.lj, a.r - b.r}; break; }   // operations.rs:71-77
            case OP_NEG: { Triple a = st[--sp]; st[sp++] = {-a.li, -a.lj, -a.r}; break; }                                  // operations.rs:170-176
            default: throw std::runtime_error("bad term opcode");
            }
        }
        return st[0];
    }

    // Term::cell_terms of the whole expression for cell c
    Pair cell_terms(size_t c) const {
        Pair st[32];
        int sp = 0;
        for (int k = 0; k < nops; ++k) {
            const TermOp& o = prog[k];
            switch (o.op) {
            case OP_SOURCE:          // source.rs:34-39
                st[sp++] = {0.0, -cell_fields[o.arg0][c]};
                break;
            case OP_LINEAR_SOURCE:   // source.rs:72-77
                st[sp++] = {cell_fields[o.arg0][c], 0.0};
                break;
            case OP_TIME_EULER: {    // terms/time.rs:31-34 -> schemes/time/euler.rs:27-37
                double r0 = 1.0, r1 = 1.0;
                if (o.arg1 >= 0) { r0 = cell_fields[o.arg1][c]; r1 = cell_fields[o.arg2][c]; }
                double dt = o.scalar;
                st[sp++] = {1.0 * (1.0 / dt) * r0, cell_fields[o.arg0][c] / dt * r1};
                break;
            }
            case OP_TIME_BACKWARD: {   // terms/time.rs:31-34 -> schemes/time/backward.rs:21-62
                // Backward::new: wn = dt / dt_2, powi(2) is one rounded product
                double dt = o.scalar, dt_2 = prog[k + 1].scalar;
                double wn = dt / dt_2;
                double a_next = (1.0 + 2.0 * wn) / (1.0 + wn);
                double a_last = ((1.0 + wn) * (1.0 + wn)) / (1.0 + wn);
                double a_last_2 = (wn * wn) / (1.0 + wn);
                double r0 = 1.0, r1 = 1.0, r2 = 1.0;
                if (o.arg2 >= 0) { r0 = cell_fields[o.arg2][c]; r1 = cell_fields[o.arg2 + 1][c]; r2 = cell_fields[o.arg2 + 2][c]; }
                double prev = cell_fields[o.arg0][c], prev2 = cell_fields[o.arg1][c];
                st[sp++] = {1.0 * (a_next / dt) * r0, (prev * a_last * r1 - prev2 * a_last_2 * r2) / dt};
                break;
            }
            case OP_USER_TERM:   // any other impl Term: its cell_terms as given
                if (o.arg1 >= 0) st[sp++] = {cell_fields[o.arg1][c], cell_fields[o.arg1 + 1][c]};
                else st[sp++] = {0.0, 0.0};
                break;
            case OP_PARAM:
                break;
            case OP_LAPLACIAN_ORTHOGONAL: case OP_CONVECTION_UPWIND: case OP_DIVERGENCE:
            case OP_LAPLACIAN_CORRECTED: case OP_CONVECTION_LINEAR: case OP_CONVECTION_LIMITED_LINEAR:
                st[sp++] = {0.0, 0.0};   // Term::cell_terms default (terms/mod.rs:28-30)
                break;
            case OP_ADD: { Pair b = st[--sp], a = st[--sp]; st[sp++] = {a.l + b.l, a.r + b.r}; break; }
            case OP_SUB: { Pair b = st[--sp], a = st[--sp]; st[sp++] = {a.l - b.l, a.r - b.r}; break; }
            case OP_NEG: { Pair a = st[--sp]; st[sp++] = {-a.l, -a.r}; break; }
            default: throw std::runtime_error("bad term opcode");
            }
        }
        return st[0];
    }

    // assembly.rs:11-79.  bc_i/bc_e are the boundary closure values, in boundary-face order.
    void assemble(Matrix& lhs, std::vector<double>& rhs, const double* bc_i, const double* bc_e) const {
        size_t n = mesh.n_cells();
        lhs = Matrix::from_cut_sparsity(mesh.cell_to_cell, n);
        rhs.assign(n, 0.0);
        for (size_t c = 0; c < n; ++c) {
            Pair p = cell_terms(c);
            lhs.at(c, c) = p.l * mesh.cell_data[c].volume;
            rhs[c] = p.r * mesh.cell_data[c].volume;
        }
        size_t b = 0;
        for (size_t f = 0; f < mesh.n_faces(); ++f) {
            Triple t = face_terms(f);
            const FaceData<DIM>& fd = mesh.face_data[f];
            double li = t.li * fd.area, lj = t.lj * fd.area, r = t.r * fd.area;
            size_t i = fd.owner;
            if (fd.kind == 1) {
                size_t j = fd.neighbor;
                lhs.at(i, i) += li;
                lhs.at(i, j) += lj;
                rhs[i] += r;
                lhs.at(j, i) -= li;
                lhs.at(j, j) -= lj;
                rhs[j] -= r;
            } else {
                double bndi = bc_i[b], bnde = bc_e[b];
                ++b;
                lhs.at(i, i) += li;
                lhs.at(i, i) += lj * bndi;
                rhs[i] += r - lj * bnde;
            }
        }
    }
};

// ---------------------------------------------------------------------------------
// fvm tools: compute_gradients (tools/gradients.rs:41-79) and compute_limiters
// (tools/limiters.rs:47-126) for values with nc components (nc = 1: f64 fields;
// nc > 1: Vector<nc> fields with Matrix<nc, DIM> gradients).  Boundary closures are
// (bci per boundary face, bce per boundary face and component), in boundary-face order.
// ---------------------------------------------------------------------------------
template <int DIM>
void compute_gradients(const Mesh<DIM>& mesh, const std::vector<size_t>& bnd_of_face, int nc, const double* values,
                       const double* bci, const double* bce, double* out) {
    for (size_t c = 0; c < mesh.n_cells(); ++c) {
        for (int q = 0; q < nc; ++q) {
            const double vc = values[c * nc + q];
            Vec<DIM> grad = mesh.cell_diag_grad[c] * vc;            // mul_to_gradient(own_grad)
            for (size_t k = mesh.cell_faces.major_start(c); k < mesh.cell_faces.major_end(c); ++k) {
                const size_t f = mesh.cell_faces.minors[k];
                const FaceData<DIM>& fd = mesh.face_data[f];
                const Vec<DIM>& g = mesh.cell_face_grad[k];
                if (fd.kind == 1) {
                    size_t o = fd.neighbor == c ? fd.owner : fd.neighbor;   // FaceRef::other_cell
                    grad = grad + g * values[o * nc + q];
                } else {
                    size_t b = bnd_of_face[f];
                    double fval = bci[b] * vc + bce[b * nc + q];
                    grad = grad + g * fval;
                }
            }
            for (int d = 0; d < DIM; ++d) out[(c * nc + q) * DIM + d] = grad[d];
        }
    }
}

// limiters/mod.rs:17-82
inline double limiter_value(int kind, double k, double r) {
    switch (kind) {
    case 0: return std::fmin(2.0 * r * k, 1.0);                            // LimitedLinear(k)
    case 1: return std::fmin(1.0 * r, 1.0);                                // MinMod
    case 2: return 2.0 * r / (1.0 + r);                                    // VanLeer
    case 3: return (r * r + r) / (r * r + 1.0);                            // VanAlbada
    case 4: case 5: {                                                      // Sweby(k), SuperBee = Sweby(2)
        double beta = kind == 5 ? 2.0 : k;
        double a = std::fmin(r, beta), b = std::fmin(beta * r, 1.0);
        return std::fmax(std::fmax(a, b), 0.0);
    }
    }
    throw std::runtime_error("bad limiter kind");
}

// interp: 0 Linear (linear.rs), 1 Upwind(flux) (upwind.rs:26-33).  fgrads: null for
// Orthogonal, else the Corrected scheme's gradient field with `reduction`.
template <int DIM>
void compute_limiters(const Mesh<DIM>& mesh, const std::vector<size_t>& bnd_of_face, int nc, const double* values,
                      const double* grads, int limiter, double k, int interp, const double* flux,
                      const double* fgrads, double reduction, const double* bci, const double* bce, double* out) {
    std::vector<double> grhs((size_t)nc), fng((size_t)nc), gdd((size_t)nc);
    for (size_t f = 0; f < mesh.n_faces(); ++f) {
        const FaceData<DIM>& fd = mesh.face_data[f];
        const size_t i = fd.owner;
        const Vec<DIM> ci = mesh.cell_data[i].center;
        double gi, gj;
        fngrad_terms(mesh, f, nc, fgrads, reduction, gi, gj, grhs.data());
        Vec<DIM> delta;
        if (fd.kind == 1) {
            const size_t j = fd.neighbor;
            delta = mesh.cell_data[j].center - ci;
            double ii, ij;
            if (interp == 0) { double w = linear_weight(mesh, f); ii = 1.0 * w; ij = 1.0 * (1.0 - w); }
            else if (flux[f] > 0.0) { ii = 1.0; ij = 0.0; }
            else { ii = 0.0; ij = 1.0; }
            for (int q = 0; q < nc; ++q) {
                double v = 0.0;
                v += gi * values[i * nc + q];
                v += gj * values[j * nc + q];
                v += grhs[q];
                fng[q] = v;
                Vec<DIM> gface;
                for (int d = 0; d < DIM; ++d) {
                    gface[d] += grads[(i * nc + q) * DIM + d] * ii;
                    gface[d] += grads[(j * nc + q) * DIM + d] * ij;
                }
                gdd[q] = dot(gface, delta);
            }
        } else {
            const size_t b = bnd_of_face[f];
            delta = fd.center - ci;
            for (int q = 0; q < nc; ++q) {
                double fval = bci[b] * values[i * nc + q] + bce[b * nc + q];
                double v = 0.0;
                v += gi * values[i * nc + q];
                v += gj * fval;
                v += grhs[q];
                fng[q] = v;
                Vec<DIM> gface;
                for (int d = 0; d < DIM; ++d) gface[d] = grads[(i * nc + q) * DIM + d];
                gdd[q] = dot(gface, delta);
            }
        }
        double r;   // LimiterFrom::to_limiter_r (limiters.rs:17-44)
        if (nc == 1) {
            r = std::fabs(fng[0]) > 1e-12 ? std::fmax(2.0 * gdd[0] / (norm(delta) * fng[0]) - 1.0, 0.0) : 1.0;
        } else {
            double ff = 0.0, fg = 0.0;
            for (int q = 0; q < nc; ++q) ff += fng[q] * fng[q];
            for (int q = 0; q < nc; ++q) fg += fng[q] * gdd[q];
            r = std::sqrt(ff) > 1e-12 ? std::fmax(2.0 * fg / (norm(delta) * ff) - 1.0, 0.0) : 1.0;
        }
        out[f] = limiter_value(limiter, k, r);
    }
}

// ---------------------------------------------------------------------------------
// Preconditioners  (src/linalg/preconditioners/*)
// ---------------------------------------------------------------------------------
struct Prec {
    virtual ~Prec() {}
    virtual void precondition(std::vector<double>& sol, const std::vector<double>& rhs) const = 0;
};

struct IdentityPrec : Prec {   // preconditioners/mod.rs:22-33
    void precondition(std::vector<double>& sol, const std::vector<double>& rhs) const override {
        for (size_t i = 0; i < rhs.size(); ++i) sol[i] = rhs[i];
    }
};

struct DiagonalPrec : Prec {   // preconditioners/mod.rs:38-68
    std::vector<double> diagonal;
    explicit DiagonalPrec(const Matrix& m) {
        diagonal.resize(m.nrows);
        for (size_t i = 0; i < m.nrows; ++i) {
            long p = m.pos(i, i);
            if (p < 0) throw std::runtime_error("Sparse matrix index not found");
            diagonal[i] = 1.0 / m.values[(size_t)p];
        }
    }
    void precondition(std::vector<double>& sol, const std::vector<double>& rhs) const override {
        for (size_t i = 0; i < diagonal.size(); ++i) sol[i] = diagonal[i] * rhs[i];
    }
};

struct ICPrec : Prec {   // incomplete_cholesky.rs:19-255
    Matrix l;
    Sparsity sp_t;
    std::vector<double> l_vt, inv_diag;

    // incomplete_cholesky.rs:35-85 (stack-ordered walk, HashMap visited/length)
    static void row_i_sparsity_edit(std::set<size_t>& s, size_t i, const Matrix& m, size_t level) {
        s.clear();
        std::vector<size_t> q{i};
        std::unordered_map<size_t, size_t> visited{{i, i}}, length{{i, 0}};
        while (!q.empty()) {
            size_t k = q.back();
            q.pop_back();
            size_t lk = length.at(k);
            for (size_t kj = m.sp.major_start(k); kj < m.sp.major_end(k); ++kj) {
                size_t j = m.sp.minors[kj];
                if (j == k) continue;
                auto it = visited.find(j);
                if (it != visited.end() && it->second == i) continue;
                visited[j] = i;
                if (lk < level) {
                    q.push_back(j);
                    length[j] = lk + 1;
                    s.insert(j);
                }
            }
        }
    }

    // incomplete_cholesky.rs:148-202
    ICPrec(const Matrix& matrix_in, size_t level) {
        Matrix matrix = matrix_in.square_block();
        if (!matrix.is_symmetric()) throw std::invalid_argument("not symmetric");
        Sparsity sp;
        std::vector<double> values;
        std::set<size_t> row_sp;
        for (size_t i = 0; i < matrix.nrows; ++i) {
            row_i_sparsity_edit(row_sp, i, matrix, level);
            for (size_t k = matrix.sp.major_start(i); k < matrix.sp.major_end(i); ++k) {
                size_t j = matrix.sp.minors[k];
                if (j <= i) { sp.push_to_major(j); values.push_back(matrix.values[k]); }
                if (j >= i) sp_t.push_to_major(j);
                row_sp.erase(j);
            }
            for (size_t j : row_sp) {
                if (j <= i) { sp.push_to_major(j); values.push_back(0.0); }
                if (j >= i) sp_t.push_to_major(j);
            }
            sp.close_major();
            sp_t.close_major();
        }
        inv_diag.assign(matrix.nrows, 0.0);
        l_vt.assign(values.size(), 0.0);
        l = Matrix::from_sparsity_and_values(std::move(sp), std::move(values));
        compute();
    }

    // incomplete_cholesky.rs:87-145
    void compute() {
        for (size_t k = 0; k < l.nrows; ++k) {
            double lkk = l.at(k, k);
            lkk = std::sqrt(lkk);
            double lkkinv = 1.0 / lkk;
            inv_diag[k] = lkkinv;
            l.at(k, k) = lkk;
            for (size_t k_i = sp_t.major_start(k); k_i < sp_t.major_end(k); ++k_i) {
                size_t i = sp_t.minors[k_i];
                if (i <= k) continue;
                long p = l.pos(i, k);
                if (p >= 0) l.values[(size_t)p] = lkkinv * l.values[(size_t)p];
            }
            for (size_t k_j = sp_t.major_start(k); k_j < sp_t.major_end(k); ++k_j) {
                for (size_t k_i = sp_t.major_start(k); k_i < sp_t.major_end(k); ++k_i) {
                    size_t i = sp_t.minors[k_i];
                    size_t j = sp_t.minors[k_j];
                    if (i < j) continue;
                    if (j <= k) continue;
                    double lik = l.get(i, k);
                    double ljk = l.get(j, k);
                    long p = l.pos(i, j);
                    if (p >= 0) l.values[(size_t)p] -= lik * ljk;
                }
            }
        }
        for (size_t row = 0; row < l.nrows; ++row)
            for (size_t k = l.sp.major_start(row); k < l.sp.major_end(row); ++k) {
                long kk = sp_t.find_flat_index(l.sp.minors[k], row);
                if (kk < 0) throw std::runtime_error("transpose entry missing");
                l_vt[(size_t)kk] = l.values[k];
            }
    }

    // incomplete_cholesky.rs:217-254
    void precondition(std::vector<double>& solution, const std::vector<double>& rhs) const override {
        for (size_t i = 0; i < l.nrows; ++i) {
            double si = rhs[i];
            for (size_t k = l.sp.major_start(i); k < l.sp.major_end(i); ++k) {
                size_t j = l.sp.minors[k];
                if (i != j) si -= l.values[k] * solution[j];
            }
            solution[i] = inv_diag[i] * si;
        }
        for (size_t ii = l.nrows; ii-- > 0;) {
            size_t i = ii;
            double si = solution[i];
            for (size_t k = sp_t.major_start(i); k < sp_t.major_end(i); ++k) {
                size_t j = sp_t.minors[k];
                if (i != j) si -= l_vt[k] * solution[j];
            }
            solution[i] = inv_diag[i] * si;
        }
    }
};

struct ILUPrec : Prec {   // incomplete_lower_upper.rs:17-328
    Matrix lu;

    // incomplete_lower_upper.rs:30-81
    static std::set<size_t> row_i_u_sparsity(size_t i, const Matrix& m, size_t level) {
        std::set<size_t> s;
        std::vector<size_t> q{i};
        std::map<size_t, size_t> visited{{i, i}}, length{{i, 0}};
        while (!q.empty()) {
            size_t k = q.back();
            q.pop_back();
            size_t lk = length.at(k);
            for (size_t kj = m.sp.major_start(k); kj < m.sp.major_end(k); ++kj) {
                size_t j = m.sp.minors[kj];
                if (j == k) continue;
                auto it = visited.find(j);
                if (it != visited.end() && it->second == i) continue;
                visited[j] = i;
                if (j < i && lk < level) { q.push_back(j); length[j] = lk + 1; }
                else if (j > i) s.insert(j);
            }
        }
        return s;
    }
    // incomplete_lower_upper.rs:84-137
    static std::set<size_t> row_i_l_sparsity(size_t i, const Matrix& m, size_t level) {
        std::set<size_t> s;
        std::vector<size_t> q{i};
        std::map<size_t, size_t> visited{{i, i}}, length{{i, 0}};
        while (!q.empty()) {
            size_t k = q.back();
            q.pop_back();
            size_t lk = length.at(k);
            for (size_t kj = m.sp.major_start(k); kj < m.sp.major_end(k); ++kj) {
                size_t j = m.sp.minors[kj];
                if (j == k) continue;
                auto it = visited.find(j);
                if (it != visited.end() && it->second == i) continue;
                visited[j] = i;
                if (j > i && lk < level) { q.push_back(j); length[j] = lk + 1; }
                else if (j < i) s.insert(j);
            }
        }
        return s;
    }

    // incomplete_lower_upper.rs:200-263.  A zero pivot throws (the reference panics).
    ILUPrec(const Matrix& matrix_in, size_t level) {
        Matrix matrix = matrix_in.square_block();
        size_t n = matrix.nrows;
        lu.nrows = 0;
        for (size_t i = 0; i < n; ++i) {
            std::set<size_t> l_row_sp = row_i_l_sparsity(i, matrix, level);
            std::set<size_t> u_row_sp = row_i_u_sparsity(i, matrix, level);
            std::unordered_map<size_t, double> lu_row;   // SparseArray (sparse_array.rs:8-84)
            for (size_t k : l_row_sp) lu_row[k] = matrix.get(i, k);
            lu_row[i] = matrix.get(i, i);
            for (size_t k : u_row_sp) lu_row[k] = matrix.get(i, k);
            for (size_t k : l_row_sp) {
                if (k >= i) continue;
                double lukk = lu.at(k, k);
                if (std::fabs(lukk) < 1e-16) throw std::domain_error("zero pivot");
                double lukkinv = 1.0 / lukk;
                lu_row.at(k) = lukkinv * lu_row.at(k);
                double lik = lu_row.at(k);
                for (size_t q = lu.sp.major_start(k); q < lu.sp.major_end(k); ++q) {
                    size_t j = lu.sp.minors[q];
                    if (j <= k) continue;
                    double term = lik * lu.values[q];
                    auto it = lu_row.find(j);
                    if (it == lu_row.end()) continue;
                    it->second -= term;
                }
            }
            // sparsify_btreeset: keep the diagonal, drop |v| <= |u_ii|*1e-6 (sparse_array.rs:48-50)
            double tol = std::fabs(lu_row.at(i)) * 1e-6;
            std::vector<std::pair<size_t, double>> row;
            for (auto& kv : lu_row)
                if (kv.first == i || std::fabs(kv.second) > tol) row.push_back(kv);
            std::sort(row.begin(), row.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
            for (auto& kv : row) { lu.sp.push_to_major(kv.first); lu.values.push_back(kv.second); }   // push_row dmatrix.rs:191-200
            lu.sp.close_major();
            lu.nrows += 1;
        }
        lu.ncols = lu.sp.max_minor + 1;
    }

    // incomplete_lower_upper.rs:291-328
    void precondition(std::vector<double>& solution, const std::vector<double>& rhs) const override {
        for (size_t i = 0; i < lu.nrows; ++i) {
            double si = rhs[i];
            for (size_t k = lu.sp.major_start(i); k < lu.sp.major_end(i); ++k) {
                size_t j = lu.sp.minors[k];
                if (j < i) si -= lu.values[k] * solution[j];
            }
            solution[i] = si;
        }
        for (size_t ii = lu.nrows; ii-- > 0;) {
            size_t i = ii;
            double si = solution[i];
            for (size_t k = lu.sp.major_start(i); k < lu.sp.major_end(i); ++k) {
                size_t j = lu.sp.minors[k];
                if (j > i) si -= lu.values[k] * solution[j];
            }
            double uii = lu.values[(size_t)lu.pos(i, i)];
            solution[i] = (1.0 / uii) * si;
        }
    }
};

// ---------------------------------------------------------------------------------
// Solvers  (src/linalg/solvers/*)
// ---------------------------------------------------------------------------------
static double vdot(const std::vector<double>& a, const std::vector<double>& b) {   // dvector.rs:63-71
    double out = 0.0;
    for (size_t i = 0; i < a.size(); ++i) out += a[i] * b[i];
    return out;
}

struct Info {
    int64_t iterations = 0;
    double initial_residual = 0, final_residual = 0;
    int64_t restarts = 0;
    std::vector<double> history;
};

// conjugate_gradient.rs:12-112
static Info conjugate_gradient(std::vector<double>& solution, const Matrix& lhs, const std::vector<double>& rhs,
                               const Prec& prec, double rtol, double atol, int64_t max_iter) {
    Info info;
    double tolerance = atol, relative_tolerance = rtol;
    double bnorm = std::sqrt(vdot(rhs, rhs));
    if (bnorm < 1e-14) {
        for (auto& v : solution) v = 0.0;
        info.initial_residual = info.final_residual = bnorm;
        info.history = {bnorm};
        return info;
    }
    size_t n = solution.size();
    std::vector<double> p(n, 0.0), r(rhs.size(), 0.0), z(n, 0.0), ap(rhs.size(), 0.0);
    lhs.imul(r, solution);
    for (size_t i = 0; i < r.size(); ++i) r[i] -= rhs[i];
    for (size_t i = 0; i < r.size(); ++i) r[i] *= -1.0;
    double initial_residual = std::sqrt(vdot(r, r)) / bnorm;
    prec.precondition(z, r);
    p = z;
    double error = 1.0;
    int64_t iter = 0;
    while (error > tolerance) {
        if (iter > max_iter) break;
        lhs.imul(ap, p);
        double r_dot_z = vdot(r, z);
        double p_dot_ap = vdot(p, ap);
        double alpha = r_dot_z / p_dot_ap;
        for (size_t i = 0; i < ap.size(); ++i) {
            solution[i] += p[i] * alpha;
            r[i] -= ap[i] * alpha;
        }
        double r_dot_r = vdot(r, r);
        error = std::sqrt(r_dot_r) / bnorm;
        if ((error < tolerance) || (error / initial_residual < relative_tolerance)) {
            info.history.push_back(error);
            iter += 1;
            break;
        }
        prec.precondition(z, r);
        double r_dot_z_new = vdot(r, z);
        double beta = r_dot_z_new / r_dot_z;
        for (size_t i = 0; i < ap.size(); ++i) p[i] = z[i] + p[i] * beta;
        info.history.push_back(error);
        iter += 1;
    }
    info.iterations = iter;
    info.initial_residual = initial_residual;
    info.final_residual = error;
    return info;
}

// bi_conjugate_gradient_stab.rs:13-179 (right preconditioned; Q4 double division kept)
static Info bicgstab(std::vector<double>& solution, const Matrix& lhs, const std::vector<double>& rhs,
                     const Prec& prec, double rtol, double atol, int64_t max_iter) {
    Info info;
    double tolerance = atol, relative_tolerance = rtol;
    double bnorm = std::sqrt(vdot(rhs, rhs));
    if (bnorm < 1e-14) {
        for (auto& v : solution) v = 0.0;
        info.initial_residual = info.final_residual = bnorm;
        info.history = {bnorm};
        return info;
    }
    size_t n = solution.size(), m = rhs.size();
    std::vector<double> p(n, 0.0), r(m, 0.0), rh(m, 0.0), z(n, 0.0), y(n, 0.0), v(m, 0.0), h(m, 0.0),
        s(m, 0.0), t(m, 0.0);
    lhs.imul(r, solution);
    for (size_t i = 0; i < m; ++i) r[i] -= rhs[i];
    for (size_t i = 0; i < m; ++i) r[i] *= -1.0;
    for (size_t i = 0; i < m; ++i) rh[i] = r[i];
    double initial_residual = std::sqrt(vdot(r, r)) / bnorm;
    if (initial_residual < tolerance) {
        info.initial_residual = info.final_residual = initial_residual;
        info.history = {initial_residual};
        return info;
    }
    for (size_t i = 0; i < std::min(n, m); ++i) p[i] = r[i];
    double rho = vdot(rh, r);
    double error = 1.0;
    int64_t iter = 0;
    while (error > tolerance) {
        if (iter > max_iter) break;
        prec.precondition(y, p);
        lhs.imul(v, y);
        double rh_dot_v = vdot(rh, v);
        double alpha = rho / rh_dot_v;
        for (size_t i = 0; i < m; ++i) {
            h[i] = solution[i] + y[i] * alpha;
            s[i] = r[i] - v[i] * alpha;
        }
        double s_dot_s_sqrt = std::sqrt(vdot(s, s)) / bnorm;
        if ((s_dot_s_sqrt < tolerance) || (s_dot_s_sqrt / initial_residual < relative_tolerance)) {
            error = s_dot_s_sqrt;
            info.history.push_back(error);
            for (size_t i = 0; i < m; ++i) solution[i] = h[i];
            iter += 1;
            break;
        }
        prec.precondition(z, s);
        lhs.imul(t, z);
        double t_dot_s = vdot(t, s);
        double t_dot_t = vdot(t, t);
        double omega = t_dot_s / t_dot_t;
        for (size_t i = 0; i < m; ++i) {
            solution[i] = h[i] + z[i] * omega;
            r[i] = s[i] - t[i] * omega;
        }
        double r_dot_r_sqrt = std::sqrt(vdot(r, r)) / bnorm;
        if ((r_dot_r_sqrt < tolerance) || (r_dot_r_sqrt / initial_residual < relative_tolerance)) {
            error = r_dot_r_sqrt;
            info.history.push_back(error);
            iter += 1;
            break;
        }
        double rho_new = vdot(rh, r);
        double beta = (rho_new / rho) * (alpha / omega);
        rho = rho_new;
        if (std::fabs(rho) > (1e-15 * bnorm)) {
            for (size_t i = 0; i < m; ++i) p[i] = r[i] + (p[i] - v[i] * omega) * beta;
        } else {
            info.restarts += 1;
            for (size_t i = 0; i < m; ++i) { rh[i] = r[i]; p[i] = r[i]; }
            rho = vdot(rh, r);
        }
        error = r_dot_r_sqrt / bnorm;
        info.history.push_back(error);
        iter += 1;
    }
    info.iterations = iter;
    info.initial_residual = initial_residual;
    info.final_residual = error;
    return info;
}

}  // namespace orc

// =================================================================================
// C interface (ctypes)
// =================================================================================
using namespace orc;

namespace {
thread_local std::string g_err;

// ---------------------------------------------------------------------------------
// SIMPLE helpers (tools/hbya.rs) and the p-Poisson wall distance (tools/wall_distance.rs)
// Vector<DIM> cell fields are n x DIM, face closures as in assembly (bci per boundary
// face, bce per boundary face and component), in boundary-face order.
// ---------------------------------------------------------------------------------
// Face interpolation scheme terms (t0, t1) with a zero right-hand side: 0 Linear
// (linear.rs:25-43), 1 Upwind(flux) (upwind.rs:26-33), 2 LimitedLinear(flux, limiters)
// (limitedlinear.rs:30-76).
template <int DIM>
void interp_terms(const Mesh<DIM>& mesh, size_t f, int kind, const double* flux, const double* lim, double& t0,
                  double& t1) {
    const FaceData<DIM>& fd = mesh.face_data[f];
    if (kind == 0) {
        double w = linear_weight(mesh, f);
        t0 = 1.0 * w;
        t1 = 1.0 * (1.0 - w);
    } else if (kind == 1) {
        if (flux[f] > 0.0) { t0 = 1.0; t1 = 0.0; }
        else { t0 = 0.0; t1 = 1.0; }
    } else if (kind == 2) {
        double fl = flux[f];
        int u = fl >= 0.0 ? 0 : 1, dn = u == 0 ? 1 : 0;
        double t[2] = {0.0, 0.0};
        double l = lim[f];
        double w = linear_weight(mesh, f);
        double wu = u == 0 ? w : 1.0 - w;
        t[u] = wu * l + (1.0 - l);
        t[dn] = (1.0 - wu) * l;
        if (fd.kind == 1) { t0 = 1.0 * t[0]; t1 = 1.0 * t[1]; }
        else if (fl > 0.0) { t0 = 1.0; t1 = 0.0; }
        else { t0 = 0.0; t1 = 1.0; }
    } else {
        throw std::runtime_error("bad interpolation scheme");
    }
}

// compute_hbya_ainv (hbya.rs:15-50): ainv_i = vol_i / a_ii; hbya_i = (mrhs_i -
// sum_{j != i} u_j a_ij) / a_ii over row i in column order.
template <int DIM>
void compute_hbya_ainv(const Mesh<DIM>& mesh, const Matrix& a, const double* mrhs, const double* vel, double* hbya,
                       double* ainv) {
    for (size_t i = 0; i < a.nrows; ++i) ainv[i] = mesh.cell_data[i].volume / a.index(i, i);
    for (size_t i = 0; i < a.nrows; ++i) {
        Vec<DIM> hi;
        for (int d = 0; d < DIM; ++d) hi[d] = mrhs[i * DIM + d];
        for (size_t k = a.sp.major_start(i); k < a.sp.major_end(i); ++k) {
            const size_t j = a.sp.minors[k];
            if (i != j) {
                Vec<DIM> uj;
                for (int d = 0; d < DIM; ++d) uj[d] = vel[j * DIM + d];
                hi = hi - uj * a.values[k];
            }
        }
        hi = hi / a.index(i, i);
        for (int d = 0; d < DIM; ++d) hbya[i * DIM + d] = hi[d];
    }
}

// intepolate_hbya_ainv_faces (hbya.rs:56-102)
template <int DIM>
void interpolate_hbya_ainv_faces(const Mesh<DIM>& mesh, const std::vector<size_t>& bnd_of_face, const double* hbya,
                                 const double* ainv, int hkind, const double* hflux, const double* hlim, int akind,
                                 const double* aflux, const double* alim, const double* ubci, const double* ubce,
                                 const double* pbci, const double* pbce, double* hbyan_face, double* ainv_face) {
    auto cellv = [&](const double* v, size_t c) {
        Vec<DIM> x;
        for (int d = 0; d < DIM; ++d) x[d] = v[c * DIM + d];
        return x;
    };
    for (size_t f = 0; f < mesh.n_faces(); ++f) {
        const FaceData<DIM>& fd = mesh.face_data[f];
        const size_t i = fd.owner;
        double hi, hj, ai, aj;
        interp_terms(mesh, f, hkind, hflux, hlim, hi, hj);
        interp_terms(mesh, f, akind, aflux, alim, ai, aj);
        if (fd.kind == 1) {
            const size_t j = fd.neighbor;
            Vec<DIM> zero;
            Vec<DIM> hf = cellv(hbya, i) * hi + cellv(hbya, j) * hj + zero;   // the schemes' zero rhs
            hbyan_face[f] = dot(hf, fd.normal);
            ainv_face[f] = ai * ainv[i] + aj * ainv[j] + 0.0;
        } else {
            const size_t b = bnd_of_face[f];
            const double blhsu = ubci[b], blhsp = pbci[b], bvp = pbce[b];
            Vec<DIM> bvu;
            for (int d = 0; d < DIM; ++d) bvu[d] = ubce[b * DIM + d];
            const bool p_wall = (std::fabs(blhsp - 1.0) < 1e-10) && (std::fabs(bvp) < 1e-10);
            const bool u_wall = (std::fabs(blhsu) < 1e-10) && (std::fabs(dot(bvu, fd.normal)) < 1e-10);
            hbyan_face[f] = p_wall && u_wall ? 0.0 : dot(cellv(hbya, i), fd.normal);
            ainv_face[f] = ainv[i];
        }
    }
}

// correct_phi (hbya.rs:108-157): fgrads null for Orthogonal, else Corrected(fgrads, reduction)
template <int DIM>
void correct_phi(const Mesh<DIM>& mesh, const std::vector<size_t>& bnd_of_face, const double* hbyan_face,
                 const double* ainv_face, const double* pressure, const double* fgrads, double reduction,
                 const double* pbci, const double* pbce, double* phi) {
    for (size_t f = 0; f < mesh.n_faces(); ++f) {
        const FaceData<DIM>& fd = mesh.face_data[f];
        const size_t i = fd.owner;
        const double ainv = ainv_face[f], hbyan = hbyan_face[f];
        double gi, gj, grhs;
        fngrad_terms(mesh, f, 1, fgrads, reduction, gi, gj, &grhs);
        double pgrad_n;
        if (fd.kind == 1) {
            pgrad_n = gi * pressure[i] + gj * pressure[fd.neighbor] + grhs;
        } else {
            const size_t b = bnd_of_face[f];
            const double p_face = pbci[b] * pressure[i] + pbce[b];
            pgrad_n = gi * pressure[i] + gj * p_face + grhs;
        }
        phi[f] = hbyan - ainv * pgrad_n;
    }
}

// correct_velocity (hbya.rs:161-179)
template <int DIM>
void correct_velocity(const Mesh<DIM>& mesh, const double* hbya, const double* ainv, const double* pgrad,
                      double* vel) {
    for (size_t c = 0; c < mesh.n_cells(); ++c)
        for (int d = 0; d < DIM; ++d) vel[c * DIM + d] = hbya[c * DIM + d] - pgrad[c * DIM + d] * ainv[c];
}

// f64::powi for the small positive exponents used here, as binary exponentiation
// (the expansion the Rust compiler emits; the products commute, so the order of the
// two factors of each product does not change a bit)
inline double powi_pos(double x, int n) {
    double r = 1.0;
    bool first = true;
    while (n) {
        if (n & 1) { r = first ? x : r * x; first = false; }
        n >>= 1;
        if (n) x = x * x;
    }
    return r;
}

// wall_distance.rs:92-103: face diffusivity |grad_f|^(p-2), grad_f the Linear
// interpolation of the cell gradients (the owner's on a boundary face)
template <int DIM>
void wall_face_viscosity(const Mesh<DIM>& mesh, const double* grads, int p, double* visc) {
    for (size_t f = 0; f < mesh.n_faces(); ++f) {
        const FaceData<DIM>& fd = mesh.face_data[f];
        Vec<DIM> g0, gf;
        for (int d = 0; d < DIM; ++d) g0[d] = grads[fd.owner * DIM + d];
        if (fd.kind == 1) {
            Vec<DIM> g1;
            for (int d = 0; d < DIM; ++d) g1[d] = grads[fd.neighbor * DIM + d];
            double t0, t1;
            interp_terms(mesh, f, 0, nullptr, nullptr, t0, t1);
            gf = g0 * t0 + g1 * t1;
        } else {
            gf = g0;
        }
        visc[f] = powi_pos(norm(gf), p - 2);
    }
}

// wall_distance.rs:107-116
template <int DIM>
void wall_normalize(const Mesh<DIM>& mesh, const double* grads, int p, double* dist) {
    for (size_t c = 0; c < mesh.n_cells(); ++c) {
        Vec<DIM> g;
        for (int d = 0; d < DIM; ++d) g[d] = grads[c * DIM + d];
        const double gnorm = norm(g), u = dist[c], pf = (double)p;
        dist[c] = -powi_pos(gnorm, p - 1) + std::pow((pf / (pf - 1.)) * u + powi_pos(gnorm, p), (pf - 1.) / pf);
    }
}

// compute_wall_distance (wall_distance.rs:8-119) with p = 3: wall[b] marks boundary face
// b (boundary order) as a wall (closure (0, 0)), the others are (1, 0).  Returns the
// number of solves; info receives the last solve's record.
template <int DIM>
int compute_wall_distance(const Mesh<DIM>& mesh, const std::vector<size_t>& bnd_of_face, size_t n_bnd,
                          const uint8_t* wall, int max_iterations, double tolerance, double rtol, double atol,
                          int64_t max_iter, double* dist, Info* last) {
    const int p = 3;
    const size_t n = mesh.n_cells(), F = mesh.n_faces();
    std::vector<double> grad(n * DIM, 0.0), visc(F, 1.0), src(n, -1.0), bci(n_bnd), bce(n_bnd, 0.0);
    for (size_t b = 0; b < n_bnd; ++b) bci[b] = wall[b] ? 0.0 : 1.0;
    // - laplacian(Corrected(&grad, 1.0), &visc) + source(&src)
    const TermOp prog[4] = {{OP_LAPLACIAN_CORRECTED, 0, 1, 0, 1.0}, {OP_NEG, 0, 0, 0, 0.0},
                            {OP_SOURCE, 0, 0, 0, 0.0}, {OP_ADD, 0, 0, 0, 0.0}};
    int solves = 0;
    for (int iter = 1; iter <= max_iterations; ++iter) {
        const double* ff[1] = {visc.data()};
        const double* cf[2] = {src.data(), grad.data()};
        Matrix lhs;
        std::vector<double> rhs;
        Assembler<DIM>{mesh, prog, 4, ff, cf}.assemble(lhs, rhs, bci.data(), bce.data());
        std::vector<double> sol(dist, dist + n);
        ICPrec prec(lhs, 1);
        Info info = conjugate_gradient(sol, lhs, rhs, prec, rtol, atol, max_iter);
        ++solves;
        for (size_t c = 0; c < n; ++c) dist[c] = sol[c];
        if (last) *last = info;
        if (info.initial_residual < tolerance) break;
        compute_gradients(mesh, bnd_of_face, 1, dist, bci.data(), bce.data(), grad.data());
        wall_face_viscosity(mesh, grad.data(), p, visc.data());
    }
    wall_normalize(mesh, grad.data(), p, dist);
    return solves;
}

struct MeshHandle {
    int dim;
    Mesh<2>* m2 = nullptr;
    Mesh<3>* m3 = nullptr;
    std::vector<size_t> bnd_of_face;   // boundary ordinal per face, NONE for interior
    size_t n_bnd = 0;
    ~MeshHandle() { delete m2; delete m3; }
};

template <int DIM>
Mesh<DIM>* build_mesh(int64_t nn, const double* nodes, int64_t nf, const int64_t* fptr, const int32_t* fnodes,
                      const int32_t* fpatch, int64_t nc, const int64_t* cptr, const int32_t* cfaces) {
    auto* m = new Mesh<DIM>();
    for (int64_t i = 0; i < nn; ++i) {
        Vec<DIM> p;
        for (int d = 0; d < DIM; ++d) p[d] = nodes[i * DIM + d];
        m->add_node(p);
    }
    for (int64_t f = 0; f < nf; ++f) m->add_face(fnodes + fptr[f], (size_t)(fptr[f + 1] - fptr[f]), fpatch[f]);
    for (int64_t c = 0; c < nc; ++c) m->add_cell(cfaces + cptr[c], (size_t)(cptr[c + 1] - cptr[c]));
    m->compute();
    return m;
}

struct MatrixHandle { Matrix m; };
struct PrecHandle { Prec* p = nullptr; ~PrecHandle() { delete p; } };

Matrix csr_to_matrix(int64_t n, const int32_t* row_ptr, const int32_t* col, const double* val) {
    Matrix m;
    for (int64_t i = 0; i < n; ++i) {
        for (int64_t k = row_ptr[i]; k < row_ptr[i + 1]; ++k) {
            m.sp.push_to_major((size_t)col[k]);
            m.values.push_back(val[k]);
        }
        m.sp.close_major();
    }
    m.nrows = (size_t)n;
    m.ncols = m.sp.max_minor + 1;
    return m;
}
}  // namespace

template <int DIM>
static void geometry_out(const Mesh<DIM>& m, double* farea, double* fcen, double* fnorm, int32_t* owner, int32_t* neigh,
                         double* cvol, double* ccen) {
    for (size_t f = 0; f < m.n_faces(); ++f) {
        const auto& fd = m.face_data[f];
        if (farea) farea[f] = fd.area;
        for (int d = 0; d < DIM; ++d) {
            if (fcen) fcen[f * DIM + d] = fd.center[d];
            if (fnorm) fnorm[f * DIM + d] = fd.normal[d];
        }
        if (owner) owner[f] = (int32_t)fd.owner;
        if (neigh) neigh[f] = fd.kind == 1 ? (int32_t)fd.neighbor : -1;
    }
    for (size_t c = 0; c < m.n_cells(); ++c) {
        if (cvol) cvol[c] = m.cell_data[c].volume;
        for (int d = 0; d < DIM; ++d)
            if (ccen) ccen[c * DIM + d] = m.cell_data[c].center[d];
    }
}

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

// Mesh::add_node/add_face/add_cell + Mesh::compute.  Returns NULL on error.
void* orc_mesh_new(int dim, int64_t nn, const double* nodes, int64_t nf, const int64_t* fptr, const int32_t* fnodes,
                   const int32_t* fpatch, int64_t nc, const int64_t* cptr, const int32_t* cfaces) {
    try {
        auto* h = new MeshHandle();
        h->dim = dim;
        if (dim == 2) h->m2 = build_mesh<2>(nn, nodes, nf, fptr, fnodes, fpatch, nc, cptr, cfaces);
        else if (dim == 3) h->m3 = build_mesh<3>(nn, nodes, nf, fptr, fnodes, fpatch, nc, cptr, cfaces);
        else { delete h; g_err = "invalid dimension"; return nullptr; }
        h->bnd_of_face.assign((size_t)nf, NONE);
        for (int64_t f = 0; f < nf; ++f)
            if (fpatch[f] >= 0) h->bnd_of_face[(size_t)f] = h->n_bnd++;
        return h;
    } catch (std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

void orc_mesh_free(void* h) { delete (MeshHandle*)h; }

int64_t orc_mesh_nnz(void* hv) {
    auto* h = (MeshHandle*)hv;
    return h->dim == 2 ? (int64_t)h->m2->cell_to_cell.minor_len() : (int64_t)h->m3->cell_to_cell.minor_len();
}

void orc_mesh_geometry(void* hv, double* farea, double* fcen, double* fnorm, int32_t* owner, int32_t* neigh,
                       double* cvol, double* ccen) {
    auto* h = (MeshHandle*)hv;
    if (h->dim == 2) geometry_out(*h->m2, farea, fcen, fnorm, owner, neigh, cvol, ccen);
    else geometry_out(*h->m3, farea, fcen, fnorm, owner, neigh, cvol, ccen);
}

// assemble(): outputs the sorted CSR pattern (int32), values and rhs.
int orc_assemble(void* hv, const TermOp* prog, int nops, const double* const* face_fields,
                 const double* const* cell_fields, const double* bc_i, const double* bc_e, int32_t* row_ptr,
                 int32_t* col_idx, double* values, double* rhs) {
    try {
        auto* h = (MeshHandle*)hv;
        Matrix A;
        std::vector<double> b;
        if (h->dim == 2) Assembler<2>{*h->m2, prog, nops, face_fields, cell_fields}.assemble(A, b, bc_i, bc_e);
        else Assembler<3>{*h->m3, prog, nops, face_fields, cell_fields}.assemble(A, b, bc_i, bc_e);
        for (size_t i = 0; i <= A.nrows; ++i) row_ptr[i] = (int32_t)A.sp.major_starts[i];
        for (size_t k = 0; k < A.sp.minor_len(); ++k) { col_idx[k] = (int32_t)A.sp.minors[k]; values[k] = A.values[k]; }
        for (size_t i = 0; i < b.size(); ++i) rhs[i] = b[i];
        return 0;
    } catch (std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

// Gradient coefficients of Mesh::compute: per cell-face listing (DIM each) and per cell.
void orc_mesh_grad_coefficients(void* hv, double* face_coef, double* diag_coef) {
    auto* h = (MeshHandle*)hv;
    auto out = [&](auto& m, int dim) {
        for (size_t k = 0; k < m.cell_face_grad.size(); ++k)
            for (int d = 0; d < dim; ++d) face_coef[k * dim + d] = m.cell_face_grad[k][d];
        for (size_t c = 0; c < m.cell_diag_grad.size(); ++c)
            for (int d = 0; d < dim; ++d) diag_coef[c * dim + d] = m.cell_diag_grad[c][d];
    };
    if (h->dim == 2) out(*h->m2, 2);
    else out(*h->m3, 3);
}

int orc_gradients(void* hv, int nc, const double* values, const double* bci, const double* bce, double* out) {
    auto* h = (MeshHandle*)hv;
    if (h->dim == 2) compute_gradients(*h->m2, h->bnd_of_face, nc, values, bci, bce, out);
    else compute_gradients(*h->m3, h->bnd_of_face, nc, values, bci, bce, out);
    return 0;
}

int orc_limiters(void* hv, int nc, const double* values, const double* grads, int limiter, double k, int interp,
                 const double* flux, const double* fgrads, double reduction, const double* bci, const double* bce,
                 double* out) {
    try {
        auto* h = (MeshHandle*)hv;
        if (h->dim == 2)
            compute_limiters(*h->m2, h->bnd_of_face, nc, values, grads, limiter, k, interp, flux, fgrads, reduction,
                             bci, bce, out);
        else
            compute_limiters(*h->m3, h->bnd_of_face, nc, values, grads, limiter, k, interp, flux, fgrads, reduction,
                             bci, bce, out);
        return 0;
    } catch (std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

// Mat<N>::inverse on one row-major N x N matrix (N = 2 or 3); 0 on success, 1 singular.
int orc_matrix_inverse(int n, const double* a, double* out) {
    auto run = [&](auto tag) {
        constexpr int N = decltype(tag)::value;
        Mat<N> m, r;
        for (int i = 0; i < N; ++i)
            for (int j = 0; j < N; ++j) m.r[i][j] = a[i * N + j];
        if (!inverse(m, r)) return 1;
        for (int i = 0; i < N; ++i)
            for (int j = 0; j < N; ++j) out[i * N + j] = r.r[i][j];
        return 0;
    };
    if (n == 2) return run(std::integral_constant<int, 2>());
    if (n == 3) return run(std::integral_constant<int, 3>());
    return 1;
}

void orc_spmv(int64_t n, const int32_t* row_ptr, const int32_t* col, const double* val, const double* x, double* y) {
    Matrix m = csr_to_matrix(n, row_ptr, col, val);
    std::vector<double> xx(x, x + m.ncols), yy((size_t)n);
    m.imul(yy, xx);
    for (int64_t i = 0; i < n; ++i) y[i] = yy[(size_t)i];
}

// kind: 0 identity, 1 diagonal, 2 IC(level), 3 ILU(level).
// status: 0 ok, 4 zero pivot, 5 not symmetric, 1 other.
void* orc_prec_new(int kind, int64_t n, const int32_t* row_ptr, const int32_t* col, const double* val, int level,
                   int* status) {
    *status = 0;
    try {
        Matrix m = csr_to_matrix(n, row_ptr, col, val);
        auto* h = new PrecHandle();
        if (kind == 0) h->p = new IdentityPrec();
        else if (kind == 1) h->p = new DiagonalPrec(m);
        else if (kind == 2) h->p = new ICPrec(m, (size_t)level);
        else if (kind == 3) h->p = new ILUPrec(m, (size_t)level);
        else { delete h; *status = 1; g_err = "bad preconditioner kind"; return nullptr; }
        return h;
    } catch (std::invalid_argument& e) {
        g_err = e.what(); *status = 5; return nullptr;
    } catch (std::domain_error& e) {
        g_err = e.what(); *status = 4; return nullptr;
    } catch (std::exception& e) {
        g_err = e.what(); *status = 1; return nullptr;
    }
}

void orc_prec_free(void* h) { delete (PrecHandle*)h; }

// Factor access: IC -> L (lower, with diagonal) and inv_diag; ILU -> LU rows.
int64_t orc_prec_factor_nnz(void* hv) {
    auto* h = (PrecHandle*)hv;
    if (auto* ic = dynamic_cast<ICPrec*>(h->p)) return (int64_t)ic->l.sp.minor_len();
    if (auto* ilu = dynamic_cast<ILUPrec*>(h->p)) return (int64_t)ilu->lu.sp.minor_len();
    return 0;
}
int orc_prec_factor(void* hv, int32_t* row_ptr, int32_t* col, double* val, double* inv_diag) {
    auto* h = (PrecHandle*)hv;
    const Matrix* m = nullptr;
    if (auto* ic = dynamic_cast<ICPrec*>(h->p)) {
        m = &ic->l;
        if (inv_diag) for (size_t i = 0; i < ic->inv_diag.size(); ++i) inv_diag[i] = ic->inv_diag[i];
    } else if (auto* ilu = dynamic_cast<ILUPrec*>(h->p)) {
        m = &ilu->lu;
    } else {
        return 1;
    }
    for (size_t i = 0; i <= m->nrows; ++i) row_ptr[i] = (int32_t)m->sp.major_starts[i];
    for (size_t k = 0; k < m->sp.minor_len(); ++k) { col[k] = (int32_t)m->sp.minors[k]; val[k] = m->values[k]; }
    return 0;
}

void orc_precondition(void* hv, int64_t n, const double* rhs, double* sol) {
    auto* h = (PrecHandle*)hv;
    std::vector<double> r(rhs, rhs + n), s(sol, sol + n);
    h->p->precondition(s, r);
    for (int64_t i = 0; i < n; ++i) sol[i] = s[(size_t)i];
}

// solver: 0 CG, 1 BiCGStab.  out_info: [iterations, initial_residual, final_residual, restarts, history_len]
int orc_solve(int solver, int64_t n, const int32_t* row_ptr, const int32_t* col, const double* val, void* prec,
              const double* b, double* x, double rtol, double atol, int64_t max_iter, double* out_info,
              double* history, int64_t history_cap) {
    try {
        Matrix m = csr_to_matrix(n, row_ptr, col, val);
        std::vector<double> bb(b, b + n), xx(x, x + n);
        const Prec& p = *((PrecHandle*)prec)->p;
        Info info = solver == 0 ? conjugate_gradient(xx, m, bb, p, rtol, atol, max_iter)
                                : bicgstab(xx, m, bb, p, rtol, atol, max_iter);
        for (int64_t i = 0; i < n; ++i) x[i] = xx[(size_t)i];
        out_info[0] = (double)info.iterations;
        out_info[1] = info.initial_residual;
        out_info[2] = info.final_residual;
        out_info[3] = (double)info.restarts;
        out_info[4] = (double)info.history.size();
        if (history)
            for (size_t k = 0; k < info.history.size() && (int64_t)k < history_cap; ++k) history[k] = info.history[k];
        return 0;
    } catch (std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

// SIMPLE helpers (tools/hbya.rs).  Vector fields are n x DIM (DIM = the mesh's).
int orc_hbya_ainv(void* hv, int64_t n, const int32_t* row_ptr, const int32_t* col, const double* val,
                  const double* mrhs, const double* vel, double* hbya, double* ainv) {
    try {
        auto* h = (MeshHandle*)hv;
        Matrix a = csr_to_matrix(n, row_ptr, col, val);
        if (h->dim == 2) compute_hbya_ainv(*h->m2, a, mrhs, vel, hbya, ainv);
        else compute_hbya_ainv(*h->m3, a, mrhs, vel, hbya, ainv);
        return 0;
    } catch (std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

int orc_interpolate_hbya_ainv_faces(void* hv, const double* hbya, const double* ainv, int hkind, const double* hflux,
                                    const double* hlim, int akind, const double* aflux, const double* alim,
                                    const double* ubci, const double* ubce, const double* pbci, const double* pbce,
                                    double* hbyan_face, double* ainv_face) {
    try {
        auto* h = (MeshHandle*)hv;
        if (h->dim == 2)
            interpolate_hbya_ainv_faces(*h->m2, h->bnd_of_face, hbya, ainv, hkind, hflux, hlim, akind, aflux, alim,
                                        ubci, ubce, pbci, pbce, hbyan_face, ainv_face);
        else
            interpolate_hbya_ainv_faces(*h->m3, h->bnd_of_face, hbya, ainv, hkind, hflux, hlim, akind, aflux, alim,
                                        ubci, ubce, pbci, pbce, hbyan_face, ainv_face);
        return 0;
    } catch (std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

int orc_correct_phi(void* hv, const double* hbyan_face, const double* ainv_face, const double* pressure,
                    const double* fgrads, double reduction, const double* pbci, const double* pbce, double* phi) {
    auto* h = (MeshHandle*)hv;
    if (h->dim == 2) correct_phi(*h->m2, h->bnd_of_face, hbyan_face, ainv_face, pressure, fgrads, reduction, pbci, pbce, phi);
    else correct_phi(*h->m3, h->bnd_of_face, hbyan_face, ainv_face, pressure, fgrads, reduction, pbci, pbce, phi);
    return 0;
}

int orc_correct_velocity(void* hv, const double* hbya, const double* ainv, const double* pgrad, double* vel) {
    auto* h = (MeshHandle*)hv;
    if (h->dim == 2) correct_velocity(*h->m2, hbya, ainv, pgrad, vel);
    else correct_velocity(*h->m3, hbya, ainv, pgrad, vel);
    return 0;
}

// wall distance pieces: the face diffusivity update and the final normalisation (p = 3)
int orc_wall_face_viscosity(void* hv, const double* grads, double* visc) {
    auto* h = (MeshHandle*)hv;
    if (h->dim == 2) wall_face_viscosity(*h->m2, grads, 3, visc);
    else wall_face_viscosity(*h->m3, grads, 3, visc);
    return 0;
}

int orc_wall_normalize(void* hv, const double* grads, double* dist) {
    auto* h = (MeshHandle*)hv;
    if (h->dim == 2) wall_normalize(*h->m2, grads, 3, dist);
    else wall_normalize(*h->m3, grads, 3, dist);
    return 0;
}

// compute_wall_distance; out_info: [solves, iterations, initial_residual, final_residual]
// of the last solve
int orc_wall_distance(void* hv, const uint8_t* wall, int max_iterations, double tolerance, double rtol, double atol,
                      int64_t max_iter, double* dist, double* out_info) {
    try {
        auto* h = (MeshHandle*)hv;
        Info last;
        int solves = h->dim == 2 ? compute_wall_distance(*h->m2, h->bnd_of_face, h->n_bnd, wall, max_iterations,
                                                         tolerance, rtol, atol, max_iter, dist, &last)
                                 : compute_wall_distance(*h->m3, h->bnd_of_face, h->n_bnd, wall, max_iterations,
                                                         tolerance, rtol, atol, max_iter, dist, &last);
        out_info[0] = solves;
        out_info[1] = (double)last.iterations;
        out_info[2] = last.initial_residual;
        out_info[3] = last.final_residual;
        return 0;
    } catch (std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

}  // extern "C"
