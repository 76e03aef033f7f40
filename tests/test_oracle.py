"""Pins the CPU FP64 oracle before it is trusted as the parity reference.

The reference crate's own tests hold no golden vectors for assembly or the solvers;
its single mesh test (src/core/mesh/io.rs:569-609) runs on square.su2.  Everything
else is pinned by what the mathematics fixes: closed forms (including the crate's
3D face-normal and volume scaling, compute.rs:305 and :359-361), defining properties
of the factorisations, exact solutions of TPFA on orthogonal grids, and an
independent direct solver (scipy).
"""
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

import meshgen
import oracle as O
import problems

HERE = os.path.dirname(os.path.abspath(__file__))
EPS = np.finfo(np.float64).eps


def _su2():
    import sys
    sys.path.insert(0, os.path.join(HERE, "golden"))
    import su2
    return su2.read(os.path.join(HERE, "golden", "square.su2"))


def test_square_su2_reference_invariants():
    """io.rs:569-589: 121 nodes, 220 faces, 100 cells, sum vol = 1 and sum_bnd n*A = 0 within 10 eps."""
    m = _su2()
    assert m["nodes"].shape[0] == 121
    om = O.OracleMesh(m)
    assert om.n_faces == 220 and om.n_cells == 100
    g = om.geometry()
    assert abs(g["cell_volume"].sum() - 1.0) < EPS * 10.0
    b = m["face_patch"] >= 0
    assert np.linalg.norm((g["face_normal"][b] * g["face_area"][b, None]).sum(axis=0)) < EPS * 10.0


def _single_cell_mesh(nodes, faces):
    fptr = np.zeros(len(faces) + 1, np.int64)
    fptr[1:] = np.cumsum([len(f) for f in faces])
    return dict(dim=nodes.shape[1], nodes=np.asarray(nodes, np.float64), face_ptr=fptr,
                face_nodes=np.array([n for f in faces for n in f], np.int32),
                face_patch=np.zeros(len(faces), np.int32), cell_ptr=np.array([0, len(faces)], np.int64),
                cell_faces=np.arange(len(faces), dtype=np.int32))


def test_closed_form_triangle():
    m = _single_cell_mesh(np.array([[0, 0], [1, 0], [0, 1]], float), [(0, 1), (1, 2), (2, 0)])
    g = O.OracleMesh(m).geometry()
    assert np.allclose(g["cell_volume"], [0.5], rtol=0, atol=1e-15)
    assert np.allclose(g["cell_center"][0], [1 / 3, 1 / 3], atol=1e-15)
    assert np.allclose(np.linalg.norm(g["face_normal"], axis=1), 1.0, atol=1e-15)   # 2D normals are unit
    assert np.allclose(sorted(g["face_area"]), [1, 1, np.sqrt(2)], atol=1e-15)
    # outward normals
    d = g["face_center"] - g["cell_center"][0]
    assert np.all((d * g["face_normal"]).sum(1) > 0)


def test_closed_form_tet_normal_and_volume_scaling():
    """3D normals have |n| = 1/3 on triangles (compute.rs:305); volume = true/3 (compute.rs:359-361)."""
    nodes = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
    m = _single_cell_mesh(nodes, [(0, 2, 1), (0, 1, 3), (1, 2, 3), (0, 3, 2)])
    g = O.OracleMesh(m).geometry()
    assert np.allclose(np.linalg.norm(g["face_normal"], axis=1), 1.0 / 3.0, atol=1e-15)
    assert np.allclose(sorted(g["face_area"]), [0.5, 0.5, 0.5, np.sqrt(3) / 2], atol=1e-15)
    assert abs(g["cell_volume"][0] - (1.0 / 6.0) / 3.0) < 1e-16
    assert np.allclose(g["cell_center"][0], [0.25, 0.25, 0.25], atol=1e-15)


def test_closed_form_prism_and_hex_scaling():
    """Quad faces carry |n| = 1/4; a right prism's volume is 5/18 of true, a cube's 1/4."""
    pr = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [0, 1, 1]], float)
    m = _single_cell_mesh(pr, [(0, 1, 2), (3, 4, 5), (0, 1, 4, 3), (1, 2, 5, 4), (2, 0, 3, 5)])
    g = O.OracleMesh(m).geometry()
    nn = np.linalg.norm(g["face_normal"], axis=1)
    assert np.allclose(nn[:2], 1 / 3, atol=1e-15) and np.allclose(nn[2:], 1 / 4, atol=1e-15)
    assert abs(g["cell_volume"][0] - 0.5 * 5.0 / 18.0) < 1e-15
    hx = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
    m = _single_cell_mesh(hx, [(0, 3, 2, 1), (4, 5, 6, 7), (0, 1, 5, 4), (1, 2, 6, 5), (2, 3, 7, 6), (3, 0, 4, 7)])
    g = O.OracleMesh(m).geometry()
    assert abs(g["cell_volume"][0] - 0.25) < 1e-15


@pytest.mark.parametrize("gen", [lambda: meshgen.tri2d(12, seed=1), lambda: meshgen.tet3d(3, seed=2),
                                 lambda: meshgen.prism3d(5, layers=3, seed=3)])
def test_generated_meshes_volume_and_closure(gen):
    m = gen()
    g = O.OracleMesh(m).geometry()
    scale = {2: 1.0}.get(m["dim"], None)
    vol = g["cell_volume"].sum()
    if m["dim"] == 2:
        assert abs(vol - 1.0) < 1e-13
    elif "tet" in m["name"]:
        assert abs(vol - 1.0 / 3.0) < 1e-13
    else:
        assert abs(vol - 5.0 / 18.0) < 1e-13
    # every cell closed: sum of outward |n|-scaled area vectors is zero up to the per-face scaling
    b = m["face_patch"] >= 0
    assert scale is None or np.linalg.norm((g["face_normal"][b] * g["face_area"][b, None]).sum(0)) < 1e-13


def test_pattern_brute_force_and_counts():
    """Row i = sorted({i} U face neighbours); nnz = n + 2 F_int (cfg1 size from the workload table)."""
    m = meshgen.tri2d(64, seed=1)
    s = meshgen.mesh_sizes(m)
    assert (s["n"], s["f_int"], s["f_bnd"], s["nnz"]) == (8192, 12160, 256, 32512)
    om = O.OracleMesh(m)
    g = om.geometry()
    prob = problems.diffusion(m, g)
    rp, ci, va, b = om.assemble(prob["prog"], prob["face_fields"], prob["cell_fields"], prob["bc_i"], prob["bc_e"])
    nbrs = [{i} for i in range(s["n"])]
    for o, nb in zip(g["owner"], g["neighbor"]):
        if nb >= 0:
            nbrs[o].add(nb)
            nbrs[nb].add(o)
    for i in range(s["n"]):
        assert list(ci[rp[i]:rp[i + 1]]) == sorted(nbrs[i])


def test_tpfa_five_point_stencil_on_square_su2():
    """Uniform 10x10 quads, k = 1: A_ij = -1 for face neighbours, A_ii = #faces + #boundary faces."""
    m = _su2()
    om = O.OracleMesh(m)
    g = om.geometry()
    n = om.n_cells
    prob = problems.diffusion(m, g, kc=np.ones(n), f=np.zeros(n))
    rp, ci, va, b = om.assemble(prob["prog"], prob["face_fields"], prob["cell_fields"], prob["bc_i"], prob["bc_e"])
    A = sp.csr_matrix((va, ci, rp), shape=(n, n)).toarray()
    nb = np.bincount(g["owner"][m["face_patch"] >= 0], minlength=n)
    off = A - np.diag(np.diag(A))
    assert np.allclose(off[off != 0], -1.0, atol=1e-11)
    assert np.allclose(np.diag(A), 4.0 + nb, atol=1e-11)
    assert np.abs(A - A.T).max() == 0.0


def test_tpfa_exact_for_linear_solution():
    """u = x with Dirichlet u = x_f on an orthogonal grid: TPFA reproduces u_c = x_c."""
    m = _su2()
    om = O.OracleMesh(m)
    g = om.geometry()
    n = om.n_cells
    prob = problems.diffusion(m, g, kc=np.ones(n), f=np.zeros(n), bc="x")
    rp, ci, va, b = om.assemble(prob["prog"], prob["face_fields"], prob["cell_fields"], prob["bc_i"], prob["bc_e"])
    P = O.OraclePrec("ic", rp, ci, va, 0)
    x, info = O.solve("cg", rp, ci, va, P, b, rtol=1e-12, atol=0.0)
    assert np.abs(x - g["cell_center"][:, 0]).max() < 1e-10


def test_laplacian_properties_on_jittered_mesh():
    m = meshgen.tet3d(5, seed=4)
    om = O.OracleMesh(m)
    g = om.geometry()
    n = om.n_cells
    kc = meshgen.lognormal_cells(n, 1.5, seed=3)
    prob = problems.diffusion(m, g, kc=kc, f=np.zeros(n))
    rp, ci, va, b = om.assemble(prob["prog"], prob["face_fields"], prob["cell_fields"], prob["bc_i"], prob["bc_e"])
    A = sp.csr_matrix((va, ci, rp), shape=(n, n))
    assert abs(A - A.T).max() == 0.0
    d = A.diagonal()
    off = A - sp.diags(d)
    assert (off.data <= 0).all() and (d > 0).all()       # M-matrix sign pattern
    interior = np.ones(n, bool)
    interior[g["owner"][m["face_patch"] >= 0]] = False
    rs = np.asarray(A.sum(axis=1)).ravel()
    assert np.abs(rs[interior]).max() < 1e-12 * d.max()   # zero interior row sums


def test_upwind_convection_column_sums():
    """Upwind convection is conservative: every interior face adds +F and -F to the same column, so
    column c sums to the outflow through c's boundary faces (inflow closed by bndi = bnde = 0)."""
    m = meshgen.tri2d(8, seed=2)
    om = O.OracleMesh(m)
    g = om.geometry()
    F = om.n_faces
    flux = np.random.default_rng(1).normal(size=F)
    nb = om.n_bnd
    prog = [(problems.CONV, 0, 0, 0, 0.0)]
    rp, ci, va, b = om.assemble(prog, [flux], [], np.zeros(nb), np.zeros(nb))
    n = om.n_cells
    A = sp.csr_matrix((va, ci, rp), shape=(n, n))
    fa = flux * g["face_area"]
    bo = g["neighbor"] < 0
    net = np.zeros(n)
    np.add.at(net, g["owner"][bo], np.where(fa[bo] > 0, fa[bo], 0.0))
    cs = np.asarray(A.sum(axis=0)).ravel()
    assert np.allclose(cs, net, atol=1e-13)
    assert np.abs(net).max() > 0.1


def test_spmv_matches_dense_product():
    m = meshgen.tri2d(10, seed=2)
    om = O.OracleMesh(m)
    g = om.geometry()
    prob = problems.convection_diffusion(m, g, seed=1)
    rp, ci, va, b = om.assemble(prob["prog"], prob["face_fields"], prob["cell_fields"], prob["bc_i"], prob["bc_e"])
    n = om.n_cells
    x = np.random.default_rng(3).normal(size=n)
    y = O.spmv(rp, ci, va, x)
    ref = sp.csr_matrix((va, ci, rp), shape=(n, n)).toarray() @ x
    assert np.allclose(y, ref, rtol=1e-14, atol=1e-14 * np.abs(ref).max())


def _tridiag(n, seed):
    rng = np.random.default_rng(seed)
    off = -rng.uniform(0.5, 1.0, n - 1)
    d = np.abs(np.concatenate([[0], off])) + np.abs(np.concatenate([off, [0]])) + rng.uniform(0.1, 1.0, n)
    A = np.diag(d) + np.diag(off, 1) + np.diag(off, -1)
    S = sp.csr_matrix(A)
    S.sort_indices()
    return A, S.indptr.astype(np.int32), S.indices.astype(np.int32), S.data.copy()


def test_ic0_equals_cholesky_on_a_chain():
    """No fill on a tridiagonal matrix: IC(0) is the exact Cholesky factor; PCG converges in one step."""
    A, rp, ci, va = _tridiag(40, 1)
    P = O.OraclePrec("ic", rp, ci, va, 0)
    lrp, lci, lva, inv = P.factor()
    L = sp.csr_matrix((lva, lci, lrp), shape=A.shape).toarray()
    Lref = np.linalg.cholesky(A)
    assert np.allclose(L, Lref, rtol=1e-14, atol=1e-14)
    assert np.allclose(inv, 1.0 / np.diag(Lref), rtol=1e-14)
    b = np.random.default_rng(2).normal(size=40)
    x, info = O.solve("cg", rp, ci, va, P, b, rtol=1e-10, atol=0.0)
    assert info["iterations"] == 1
    assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-12)


def test_ic0_defining_property_on_mesh_matrix():
    """(L L^T)_ij = A_ij on the pattern of A."""
    m = meshgen.tri2d(10, seed=5)
    om = O.OracleMesh(m)
    g = om.geometry()
    prob = problems.diffusion(m, g, kc=meshgen.lognormal_cells(om.n_cells, 1.0, 1))
    rp, ci, va, b = om.assemble(prob["prog"], prob["face_fields"], prob["cell_fields"], prob["bc_i"], prob["bc_e"])
    n = om.n_cells
    P = O.OraclePrec("ic", rp, ci, va, 0)
    lrp, lci, lva, _ = P.factor()
    L = sp.csr_matrix((lva, lci, lrp), shape=(n, n)).toarray()
    LLt = L @ L.T
    A = sp.csr_matrix((va, ci, rp), shape=(n, n)).toarray()
    mask = A != 0
    assert np.allclose(LLt[mask], A[mask], rtol=1e-12, atol=1e-12 * np.abs(A).max())
    assert not np.allclose(LLt, A)   # but not a complete factorisation
    # preconditioner apply = (L L^T)^-1 r
    r = np.random.default_rng(4).normal(size=n)
    assert np.allclose(P.apply(r), np.linalg.solve(LLt, r), rtol=1e-10)


def test_ilu0_equals_lu_on_a_chain_and_defining_property():
    A, rp, ci, va = _tridiag(30, 3)
    A = A + np.diag(0.3 * np.ones(29), 1)   # make it nonsymmetric
    S = sp.csr_matrix(A)
    S.sort_indices()
    rp, ci, va = S.indptr.astype(np.int32), S.indices.astype(np.int32), S.data.copy()
    P = O.OraclePrec("ilu", rp, ci, va, 0)
    lrp, lci, lva, _ = P.factor()
    LU = sp.csr_matrix((lva, lci, lrp), shape=A.shape).toarray()
    L = np.tril(LU, -1) + np.eye(30)
    U = np.triu(LU)
    assert np.allclose(L @ U, A, rtol=1e-13, atol=1e-13)   # exact LU (no fill on a chain)
    r = np.random.default_rng(5).normal(size=30)
    assert np.allclose(P.apply(r), np.linalg.solve(A, r), rtol=1e-11)


def test_ilu0_drop_rule():
    """Entries with |v| <= 1e-6 |u_ii| are removed after the row is eliminated (incomplete_lower_upper.rs:255-258)."""
    A = np.array([[4.0, 1e-7, 1.0], [1e-7 * 4, 4.0, 1.0], [1.0, 1.0, 4.0]])
    S = sp.csr_matrix(A)
    S.sort_indices()
    P = O.OraclePrec("ilu", S.indptr.astype(np.int32), S.indices.astype(np.int32), S.data.copy(), 0)
    lrp, lci, lva, _ = P.factor()
    assert list(lci[lrp[0]:lrp[1]]) == [0, 2]       # 1e-7 <= 1e-6 * 4 dropped from row 0
    assert list(lci[lrp[1]:lrp[2]]) == [1, 2]       # l_10 = 1e-7 dropped from row 1
    assert list(lci[lrp[2]:lrp[3]]) == [0, 1, 2]


@pytest.mark.parametrize("solver,prec,kind", [("cg", "ic", "sym"), ("cg", "diagonal", "sym"),
                                              ("bicgstab", "ilu", "nonsym"), ("bicgstab", "identity", "nonsym")])
def test_solvers_against_direct_solve(solver, prec, kind):
    m = meshgen.tri2d(24, seed=6)
    om = O.OracleMesh(m)
    g = om.geometry()
    if kind == "sym":
        prob = problems.diffusion(m, g, kc=meshgen.lognormal_cells(om.n_cells, 1.0, 2), seed=5)
    else:
        prob = problems.convection_diffusion(m, g, pe=(1, 20), seed=2)
    rp, ci, va, b = om.assemble(prob["prog"], prob["face_fields"], prob["cell_fields"], prob["bc_i"], prob["bc_e"])
    n = om.n_cells
    P = O.OraclePrec(prec, rp, ci, va, 0)
    x, info = O.solve(solver, rp, ci, va, P, b, rtol=1e-10, atol=0.0, max_iter=5000)
    A = sp.csr_matrix((va, ci, rp), shape=(n, n))
    xs = spl.spsolve(A.tocsc(), b)
    assert np.linalg.norm(x - xs) <= 1e-6 * np.linalg.norm(xs)
    assert np.linalg.norm(b - A @ x) <= 1e-8 * np.linalg.norm(b)
    assert info["history"].shape[0] == info["iterations"]


def test_bicgstab_divides_by_bnorm_twice_in_loop_test():
    """Q4: the while test sees |r|/|b|^2 (bi_conjugate_gradient_stab.rs:139, 171); with |b| = 1e4 and
    absolute_tolerance 1e-6 the loop leaves while |r|/|b| is still ~1e-2."""
    m = meshgen.tri2d(24, seed=6)
    om = O.OracleMesh(m)
    g = om.geometry()
    prob = problems.convection_diffusion(m, g, pe=(1, 20), seed=2)
    rp, ci, va, b = om.assemble(prob["prog"], prob["face_fields"], prob["cell_fields"], prob["bc_i"], prob["bc_e"])
    b = b / np.linalg.norm(b) * 1e4
    P = O.OraclePrec("identity", rp, ci, va, 0)
    x, info = O.solve("bicgstab", rp, ci, va, P, b, rtol=0.0, atol=1e-6, max_iter=5000)
    n = om.n_cells
    A = sp.csr_matrix((va, ci, rp), shape=(n, n))
    rel = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
    assert info["final_residual"] <= 1e-6
    assert rel > 1e-5   # a single division would have iterated down to 1e-6
    # the recorded history is |r|/|b|^2
    assert info["history"][-1] <= 1e-6 and info["history"][-1] * 1e4 == pytest.approx(rel, rel=1e-6)


def test_cg_runs_max_iterations_plus_one():
    A, rp, ci, va = _tridiag(200, 7)
    b = np.random.default_rng(1).normal(size=200)
    P = O.OraclePrec("identity", rp, ci, va, 0)
    x, info = O.solve("cg", rp, ci, va, P, b, rtol=1e-30, atol=0.0, max_iter=5)
    assert info["iterations"] == 6


def test_ic_rejects_nonsymmetric():
    A = np.array([[4.0, 1.0], [1.5, 4.0]])
    S = sp.csr_matrix(A)
    with pytest.raises(O.NotSymmetric):
        O.OraclePrec("ic", S.indptr.astype(np.int32), S.indices.astype(np.int32), S.data.copy(), 0)


# ---- fvm tools: gradient coefficients, compute_gradients, compute_limiters and the
#      Corrected / Linear / LimitedLinear schemes.  The reference holds no tests for
#      them; what pins them is the least-squares identity, exactness on linear fields,
#      the closed form of the limiter ratio and the schemes' limiting cases. ----------

TOOL_MESHES = {
    "tri": lambda: meshgen.tri2d(8, seed=3),
    "tet": lambda: meshgen.tet3d(4, seed=5),
    "prism": lambda: meshgen.prism3d(4, layers=3, seed=2),
}


def _linear_setup(raw, ncomp=1):
    om = O.OracleMesh(raw)
    g = om.geometry()
    D = om.dim
    rng = np.random.default_rng(ncomp + 7)
    a = rng.normal(size=(ncomp, D))
    b0 = rng.normal(size=ncomp)
    u = g["cell_center"] @ a.T + b0                      # (n, ncomp)
    bf = np.nonzero(raw["face_patch"] >= 0)[0]
    bci = np.zeros(len(bf))
    bce = g["face_center"][bf] @ a.T + b0
    return om, g, a, u, bci, bce, bf


def test_matrix_inverse_against_numpy_and_its_pivot_rule():
    rng = np.random.default_rng(3)
    for n in (2, 3):
        for _ in range(20):
            a = rng.normal(size=(n, n)) + n * np.eye(n)
            assert np.allclose(O.matrix_inverse(a) @ a, np.eye(n), atol=1e-12)
    # zero leading pivot: rows swap with the one the largest entry of row 0 names
    p = np.array([[0.0, 1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0]])
    assert np.array_equal(O.matrix_inverse(p), p)
    # singular matrices return Err(SingularMatrix)
    assert O.matrix_inverse(np.array([[1.0, 2.0], [2.0, 4.0]])) is None
    assert O.matrix_inverse(np.zeros((3, 3))) is None


@pytest.mark.parametrize("name", sorted(TOOL_MESHES))
def test_gradient_coefficients_reconstruct_the_identity(name):
    """coef_k = G^-1 dgc_k with G = sum dgc dgc^T, so sum_k coef_k dgc_k^T = I and
    diag = -sum_k coef_k (compute.rs:154-200)."""
    raw = TOOL_MESHES[name]()
    om = O.OracleMesh(raw)
    g = om.geometry()
    fc, dc = om.grad_coefficients()
    cp, cf = raw["cell_ptr"], raw["cell_faces"]
    D = om.dim
    for c in range(om.n_cells):
        S = np.zeros((D, D))
        tot = np.zeros(D)
        for k in range(cp[c], cp[c + 1]):
            f = cf[k]
            if g["neighbor"][f] < 0:
                dgc = g["face_center"][f] - g["cell_center"][c]
            else:
                o = g["owner"][f] if g["neighbor"][f] == c else g["neighbor"][f]
                dgc = g["cell_center"][o] - g["cell_center"][c]
            S += np.outer(fc[k], dgc)
            tot += fc[k]
        assert np.allclose(S, np.eye(D), atol=1e-10), c
        assert np.allclose(dc[c], -tot, rtol=1e-12, atol=1e-12 * np.abs(tot).max())


@pytest.mark.parametrize("name", sorted(TOOL_MESHES))
def test_gradients_exact_for_linear_fields_and_componentwise(name):
    om, g, a, u, bci, bce, bf = _linear_setup(TOOL_MESHES[name](), ncomp=2)
    gr = om.gradients(u, bci, bce)
    assert gr.shape == (om.n_cells, 2, om.dim)
    assert np.abs(gr - a[None]).max() <= 1e-10 * np.abs(a).max()
    # a Vector<2> field is two f64 fields, bit for bit
    for q in range(2):
        assert np.array_equal(om.gradients(u[:, q].copy(), bci, bce[:, q].copy()), gr[:, q, :])


@pytest.mark.parametrize("name", sorted(TOOL_MESHES))
def test_limiters_of_linear_fields(name):
    """Orthogonal face gradient: r = 1 on every face of a linear field, so each limiter
    takes its value at 1.  Corrected face gradient with the exact gradient: fngrad = a.n,
    so r = max(2 (a.d) / (|d| a.n) - 1, 0) in closed form (limiters.rs:17-44)."""
    om, g, a, u, bci, bce, bf = _linear_setup(TOOL_MESHES[name](), ncomp=1)
    u, bce, a1 = u[:, 0].copy(), bce[:, 0].copy(), a[0]
    gr = om.gradients(u, bci, bce)
    for lim, k, want, tol in (("limited_linear", 1.0, 1.0, 0.0), ("minmod", 1.0, 1.0, 1e-9),
                              ("vanleer", 1.0, 1.0, 1e-9), ("vanalbada", 1.0, 1.0, 1e-9),
                              ("sweby", 1.5, 1.0, 1e-9), ("superbee", 1.0, 1.0, 1e-9)):
        L = om.limiters(u, gr, lim, k, "linear", None, None, 1.0, bci, bce)
        assert np.abs(L - want).max() <= tol, lim
    exact = np.tile(a1, (om.n_cells, 1))
    rng = np.random.default_rng(2)
    flux = rng.normal(size=om.n_faces)
    L = om.limiters(u, exact, "limited_linear", 0.5, "upwind", flux, exact, 1.0, bci, bce)
    inside = g["neighbor"] >= 0
    xj = np.where(inside[:, None], g["cell_center"][np.maximum(g["neighbor"], 0)], g["face_center"])
    d = xj - g["cell_center"][g["owner"]]
    an = g["face_normal"] @ a1
    r = np.maximum(2.0 * (d @ a1) / (np.linalg.norm(d, axis=1) * an) - 1.0, 0.0)
    want = np.minimum(r, 1.0)
    ok = np.abs(an) > 1e-6
    assert ok.sum() > 0.9 * om.n_faces
    assert np.allclose(L[ok], want[ok], rtol=1e-9, atol=1e-9)


def test_vector_limiter_ratio_closed_form():
    om, g, a, u, bci, bce, bf = _linear_setup(TOOL_MESHES["tet"](), ncomp=2)
    exact = np.tile(a, (om.n_cells, 1, 1))
    L = om.limiters(u, exact, "limited_linear", 0.5, "linear", None, exact, 1.0, bci, bce)
    inside = g["neighbor"] >= 0
    xj = np.where(inside[:, None], g["cell_center"][np.maximum(g["neighbor"], 0)], g["face_center"])
    d = xj - g["cell_center"][g["owner"]]
    fng = g["face_normal"] @ a.T                         # (F, 2)
    gdd = d @ a.T
    ff = (fng * fng).sum(1)
    r = np.maximum(2.0 * (fng * gdd).sum(1) / (np.linalg.norm(d, axis=1) * ff) - 1.0, 0.0)
    ok = np.sqrt(ff) > 1e-6
    assert np.allclose(L[ok], np.minimum(r, 1.0)[ok], rtol=1e-9, atol=1e-9)


@pytest.mark.parametrize("name", sorted(TOOL_MESHES))
def test_corrected_laplacian_is_exact_for_linear_fields(name):
    """With the exact gradient the Corrected face-normal gradient is a.n on every face,
    so -laplacian(Corrected) with constant diffusivity and exact Dirichlet values leaves
    A u - b = -k sum_f A_f a.n_f = 0 in every cell; the Orthogonal scheme does not on
    these non-orthogonal meshes."""
    om, g, a, u, bci, bce, bf = _linear_setup(TOOL_MESHES[name](), ncomp=1)
    u, bce, a1 = u[:, 0].copy(), bce[:, 0].copy(), a[0]
    kf = np.full(om.n_faces, 1.7)
    grads = np.tile(a1, (om.n_cells, 1)).ravel()
    res = {}
    for scheme, prog in (("corrected", [(O.OP["LAPLACIAN_CORRECTED"], 0, 0, 0, 1.0), (O.OP["NEG"], 0, 0, 0, 0.0)]),
                         ("orthogonal", [(O.OP["LAPLACIAN_ORTHOGONAL"], 0, 0, 0, 0.0), (O.OP["NEG"], 0, 0, 0, 0.0)])):
        rp, ci, va, b = om.assemble(prog, [kf], [grads], bci, bce)
        S = sp.csr_matrix((va, ci, rp), shape=(om.n_cells, om.n_cells))
        res[scheme] = np.abs(S @ u - b).max() / (np.abs(va).max() * np.abs(u).max())
    assert res["corrected"] < 1e-12
    assert res["orthogonal"] > 1e-6


@pytest.mark.parametrize("name", sorted(TOOL_MESHES))
def test_limited_linear_limits_are_upwind_and_linear(name):
    """LimitedLinear with every limiter 0 is Upwind, bit for bit; with every limiter 1 it
    is Linear interpolation (limitedlinear.rs:24-72)."""
    raw = TOOL_MESHES[name]()
    om = O.OracleMesh(raw)
    rng = np.random.default_rng(4)
    flux = rng.normal(size=om.n_faces)
    nb = om.n_bnd
    bci, bce = rng.uniform(0, 1, nb), rng.uniform(0, 1, nb)
    ll = lambda lim: om.assemble([(O.OP["CONVECTION_LIMITED_LINEAR"], 0, 0, 1, 0.0)], [flux, lim], [], bci, bce)
    up = om.assemble([(O.OP["CONVECTION_UPWIND"], 0, 0, 0, 0.0)], [flux], [], bci, bce)
    z = ll(np.zeros(om.n_faces))
    assert np.array_equal(z[2], up[2]) and np.array_equal(z[3], up[3])
    one = ll(np.ones(om.n_faces))
    lin = om.assemble([(O.OP["CONVECTION_LINEAR"], 0, 0, 0, 0.0)], [flux], [], bci, bce)
    inside_only = np.allclose(one[2], lin[2], rtol=1e-12, atol=1e-14)
    # boundary faces keep the upwind closure in LimitedLinear, Linear interpolates there
    g = om.geometry()
    if nb == 0:
        assert inside_only
    S1 = sp.csr_matrix((one[2], one[1], one[0]), shape=(om.n_cells, om.n_cells))
    SL = sp.csr_matrix((lin[2], lin[1], lin[0]), shape=(om.n_cells, om.n_cells))
    interior_cells = np.setdiff1d(np.arange(om.n_cells), g["owner"][g["neighbor"] < 0])
    D = (S1 - SL).tocsr()[interior_cells]
    assert np.abs(D.data).max(initial=0.0) <= 1e-12 * np.abs(lin[2]).max()


def test_linear_interpolation_weight_closed_form():
    """Linear (linear.rs:22-44) puts weight w = n.(x_j - x_f) / n.(x_j - x_i) on the owner:
    with the flux of a single interior face set, its owner row holds w * flux on the
    diagonal and (1 - w) * flux at the neighbour."""
    raw = TOOL_MESHES["tet"]()
    om = O.OracleMesh(raw)
    g = om.geometry()
    inner = np.nonzero(g["neighbor"] >= 0)[0][:12]
    for f in inner:
        flux = np.zeros(om.n_faces)
        flux[f] = 2.5
        rp, ci, va, b = om.assemble([(O.OP["CONVECTION_LINEAR"], 0, 0, 0, 0.0)], [flux], [], np.zeros(om.n_bnd),
                                    np.zeros(om.n_bnd))
        S = sp.csr_matrix((va, ci, rp), shape=(om.n_cells, om.n_cells))
        i, j = g["owner"][f], g["neighbor"][f]
        n, xf = g["face_normal"][f], g["face_center"][f]
        xi, xj = g["cell_center"][i], g["cell_center"][j]
        w = n @ (xj - xf) / (n @ (xj - xi))
        A = g["face_area"][f]
        assert np.isclose(S[i, i], w * 2.5 * A, rtol=1e-12)
        assert np.isclose(S[i, j], (1.0 - w) * 2.5 * A, rtol=1e-12)
        assert np.isclose(S[j, i], -w * 2.5 * A, rtol=1e-12)


# ---- time(Euler), linear_source, divergence: terms the assembly tests above do not reach ----

def _assemble_terms(m, prog, face_fields, cell_fields, bnde=None):
    om = O.OracleMesh(m)
    g = om.geometry()
    nb = om.n_bnd
    be = np.zeros(nb) if bnde is None else bnde
    rp, ci, va, b = om.assemble(prog, face_fields, cell_fields, np.ones(nb), be)
    n = om.n_cells
    return om, g, sp.csr_matrix((va, ci, rp), shape=(n, n)), b


@pytest.mark.parametrize("mesh,total_volume", [("tri", 1.0), ("tet", 1.0 / 3.0)])
def test_time_euler_is_mass_over_dt(mesh, total_volume):
    """time(Euler) (schemes/time/euler.rs:27-37) as a cell term assembled as l*vol, r*vol
    (assembly.rs:26-32): a diagonal matrix rho*vol/dt and rhs rho_prev*u_prev*vol/dt; no
    face terms, so a boundary value changes nothing.  Summed over the cells the diagonal
    times dt is the mesh volume: 1 on the unit square, 1/3 for tets (computed tet volume is
    a third of the true one, compute.rs:359-361)."""
    m = meshgen.tri2d(12, seed=3) if mesh == "tri" else meshgen.tet3d(4, seed=3)
    n = m["cell_ptr"].shape[0] - 1
    rng = np.random.default_rng(4)
    prev, rho, rho0 = rng.normal(size=n), rng.uniform(0.5, 2.0, n), rng.uniform(0.5, 2.0, n)
    dt = 0.0125
    TE = O.OP["TIME_EULER"]
    nb = int((m["face_patch"] >= 0).sum())
    om, g, A, b = _assemble_terms(m, [(TE, 0, -1, -1, dt)], [], [prev], bnde=rng.normal(size=nb))
    vol = g["cell_volume"]
    assert A.nnz == n + 2 * int((g["neighbor"] >= 0).sum())
    off = A - sp.diags(A.diagonal())
    assert abs(off).max() == 0.0
    assert np.allclose(A.diagonal(), vol / dt, rtol=4 * EPS, atol=0)
    assert np.allclose(b, prev * vol / dt, rtol=4 * EPS, atol=0)
    assert abs(A.diagonal().sum() * dt - total_volume) < 1e-12
    # with density: rho on the diagonal, rho_prev on the right-hand side
    om, g, A, b = _assemble_terms(m, [(TE, 0, 1, 2, dt)], [], [prev, rho, rho0])
    assert np.allclose(A.diagonal(), rho * vol / dt, rtol=4 * EPS, atol=0)
    assert np.allclose(b, prev * rho0 * vol / dt, rtol=4 * EPS, atol=0)
    assert not np.allclose(b, prev * rho * vol / dt)   # the two densities are not swapped


def test_linear_source_adds_c_vol_to_the_diagonal():
    """linear_source(c) (source.rs:69-78) contributes (c, 0) per cell: diagonal c*vol, no
    right-hand side.  Added to a Laplacian it changes only the diagonal, by exactly that
    amount (term algebra, operations.rs:34-41); with c = 1 the diagonal sums to the area 1."""
    m = meshgen.tri2d(10, seed=5)
    n = m["cell_ptr"].shape[0] - 1
    F = m["face_ptr"].shape[0] - 1
    c = np.random.default_rng(6).uniform(0.1, 3.0, n)
    LS, LAP, ADD = O.OP["LINEAR_SOURCE"], O.OP["LAPLACIAN_ORTHOGONAL"], O.OP["ADD"]
    om, g, A, b = _assemble_terms(m, [(LS, 0, 0, 0, 0.0)], [], [c])
    vol = g["cell_volume"]
    assert abs(A - sp.diags(A.diagonal())).max() == 0.0 and not b.any()
    assert np.allclose(A.diagonal(), c * vol, rtol=4 * EPS, atol=0)
    om, g, A1, _ = _assemble_terms(m, [(LS, 0, 0, 0, 0.0)], [], [np.ones(n)])
    assert abs(A1.diagonal().sum() - 1.0) < 1e-12
    k = np.random.default_rng(7).uniform(0.5, 2.0, F)
    _, _, L, bl = _assemble_terms(m, [(LAP, 0, 0, 0, 0.0)], [k], [])
    _, _, LC, blc = _assemble_terms(m, [(LAP, 0, 0, 0, 0.0), (LS, 0, 0, 0, 0.0), (ADD, 0, 0, 0, 0.0)], [k], [c])
    d = (LC - L).toarray()
    assert np.allclose(np.diag(d), c * vol, rtol=1e-12, atol=1e-15)
    assert np.abs(d - np.diag(np.diag(d))).max() == 0.0
    assert np.array_equal(blc, bl)


@pytest.mark.parametrize("mesh", ["tri", "tet"])
def test_divergence_rhs_is_minus_the_net_outflow(mesh):
    """divergence(flux) (divergence.rs:36-45) gives face terms (0, 0, -flux), times the face
    area in assembly: no matrix entries, and rhs_c = -(net outflow of c) = -sum over c's
    faces of flux_f*area_f, + for faces c owns and - for faces it neighbours.  For a
    constant velocity u, flux_f = u . n_f, the outflow of every closed cell is zero
    (divergence theorem; the 3D normals of triangles are all scaled by 1/3, which keeps
    the sum zero on tets)."""
    m = meshgen.tri2d(9, seed=8) if mesh == "tri" else meshgen.tet3d(4, seed=8)
    F = m["face_ptr"].shape[0] - 1
    DIV = O.OP["DIVERGENCE"]
    flux = np.random.default_rng(9).normal(size=F)
    om, g, A, b = _assemble_terms(m, [(DIV, 0, 0, 0, 0.0)], [flux], [])
    n = om.n_cells
    assert not A.data.any()
    fa = flux * g["face_area"]
    net = np.bincount(g["owner"], weights=fa, minlength=n)
    inner = g["neighbor"] >= 0
    net -= np.bincount(g["neighbor"][inner], weights=fa[inner], minlength=n)
    assert np.allclose(b, -net, rtol=0, atol=1e-13)
    assert abs(b.sum() + fa[~inner].sum()) < 1e-12   # interior faces cancel pairwise
    u = np.array([0.7, -1.3, 0.4])[: om.dim]
    flux_u = g["face_normal"] @ u
    _, _, _, b0 = _assemble_terms(m, [(DIV, 0, 0, 0, 0.0)], [flux_u], [])
    assert np.abs(b0).max() < 1e-13
    assert np.abs(flux_u * g["face_area"]).max() > 1e-3


def test_bicgstab_restart_when_rho_collapses():
    """The restart branch (bi_conjugate_gradient_stab.rs:155-169).  A = [[2,1,-1],[1,4,1],
    [1,0,2]], b = e1, x0 = 0, no preconditioner: alpha = 1/2 and s = (0, -1/2, -1/2), so
    t = A s has t_0 = 0 and r_1 = s - omega t keeps r_1[0] = 0: rho_1 = r^.r_1 = 0 exactly,
    every number so far exact in binary.  The solver restarts (rh = p = r_1) once and then
    converges to A^-1 b.  Without the restart rho would stay 0 and the iteration stall."""
    A = np.array([[2.0, 1.0, -1.0], [1.0, 4.0, 1.0], [1.0, 0.0, 2.0]])
    S = sp.csr_matrix(A)
    S.sort_indices()
    rp, ci, va = S.indptr.astype(np.int32), S.indices.astype(np.int32), S.data.astype(np.float64)
    b = np.array([1.0, 0.0, 0.0])
    P = O.OraclePrec("identity", rp, ci, va, 0)
    x, info = O.solve("bicgstab", rp, ci, va, P, b, rtol=1e-12, atol=0.0, max_iter=50)
    assert info["restarts"] == 1
    assert np.abs(x - np.linalg.solve(A, b)).max() < 1e-12
    assert 2 <= info["iterations"] <= 5
    # the same system with b = e2 never meets rho = 0: no restart
    x2, info2 = O.solve("bicgstab", rp, ci, va, P, np.array([0.0, 1.0, 0.0]), rtol=1e-12, atol=0.0, max_iter=50)
    assert info2["restarts"] == 0
    assert np.abs(x2 - np.linalg.solve(A, [0.0, 1.0, 0.0])).max() < 1e-12


# ---- IC(level) / ILU(level) with fill (incomplete_cholesky.rs:35-85, incomplete_lower_upper.rs:30-137) ----

def _mesh_system(kind, nx=8, seed=5):
    m = meshgen.tri2d(nx, seed=seed)
    om = O.OracleMesh(m)
    g = om.geometry()
    if kind == "sym":
        prob = problems.diffusion(m, g, kc=meshgen.lognormal_cells(om.n_cells, 1.0, 1))
    else:
        prob = problems.convection_diffusion(m, g, pe=(1, 20), seed=3)
    rp, ci, va, b = om.assemble(prob["prog"], prob["face_fields"], prob["cell_fields"], prob["bc_i"], prob["bc_e"])
    return rp, ci, va, om.n_cells


def _neighbours(rp, ci, n):
    return [set(int(j) for j in ci[rp[i]:rp[i + 1]] if j != i and j < n) for i in range(n)]


def test_ic_level_one_walk_adds_no_fill():
    """The reference's IC walk inserts a vertex only when met from a vertex whose walk length is
    below the level (incomplete_cholesky.rs:72-76): at level 1 only row i's own neighbours,
    which are already in A's pattern, so IC(1) is IC(0) bit for bit."""
    rp, ci, va, n = _mesh_system("sym")
    f0 = O.OraclePrec("ic", rp, ci, va, 0).factor()
    f1 = O.OraclePrec("ic", rp, ci, va, 1).factor()
    for a, b in zip(f0, f1):
        assert np.array_equal(a, b)


def test_ic_level_two_defining_property_and_fill():
    """IC(2) fills distance-two couplings; on whatever pattern it fills, the incomplete factor
    satisfies (L L^T)_ij = A_ij on that pattern (A_ij = 0 at fill positions), closer to A than
    IC(0) off the pattern, and every fill entry couples two vertices at graph distance 2."""
    rp, ci, va, n = _mesh_system("sym")
    A = sp.csr_matrix((va, ci, rp), shape=(n, n)).toarray()
    nb = _neighbours(rp, ci, n)
    lrp, lci, lva, _ = O.OraclePrec("ic", rp, ci, va, 2).factor()
    L = sp.csr_matrix((lva, lci, lrp), shape=(n, n)).toarray()
    mask = np.zeros((n, n), bool)
    for i in range(n):
        for j in lci[lrp[i]:lrp[i + 1]]:
            mask[i, j] = mask[j, i] = True
    fill = mask & (A == 0)
    assert fill.sum() > 0
    for i, j in zip(*np.nonzero(np.tril(fill))):
        assert j not in nb[i] and nb[i] & nb[j], (i, j)    # distance 2
    LLt = L @ L.T
    assert np.allclose(LLt[mask], A[mask], rtol=1e-12, atol=1e-12 * np.abs(A).max())
    l0rp, l0ci, l0va, _ = O.OraclePrec("ic", rp, ci, va, 0).factor()
    L0 = sp.csr_matrix((l0va, l0ci, l0rp), shape=(n, n)).toarray()
    assert np.linalg.norm(LLt - A) < np.linalg.norm(L0 @ L0.T - A)


def test_ic_high_level_is_the_exact_cholesky_factor():
    """A level beyond the graph's diameter fills every row completely (every vertex met is
    explored and inserted): IC(level) is the exact Cholesky factor and PCG converges at once."""
    rp, ci, va, n = _mesh_system("sym", nx=4)
    A = sp.csr_matrix((va, ci, rp), shape=(n, n)).toarray()
    P = O.OraclePrec("ic", rp, ci, va, n)
    lrp, lci, lva, inv = P.factor()
    L = sp.csr_matrix((lva, lci, lrp), shape=(n, n)).toarray()
    assert np.allclose(L, np.linalg.cholesky(A), rtol=1e-12, atol=1e-12)
    b = np.random.default_rng(1).normal(size=n)
    x, info = O.solve("cg", rp, ci, va, P, b, rtol=1e-10, atol=0.0)
    assert info["iterations"] == 1 and np.allclose(x, np.linalg.solve(A, b), rtol=1e-10)


def test_ilu_level_one_pattern_and_defining_property():
    """ILU(1) (incomplete_lower_upper.rs:30-137): row i's upper part holds the j > i adjacent to
    i or to a lower neighbour of i, its lower part the j < i adjacent to i or to a higher
    neighbour of i (brute force below).  On that pattern, without drops, (L U)_ij = A_ij."""
    rp, ci, va, n = _mesh_system("nonsym")
    nb = _neighbours(rp, ci, n)
    A = sp.csr_matrix((va, ci, rp), shape=(n, n)).toarray()
    P = O.OraclePrec("ilu", rp, ci, va, 1)
    frp, fci, fva, _ = P.factor()
    for i in range(n):
        up = {j for j in nb[i] if j > i} | {j for k in nb[i] if k < i for j in nb[k] if j > i}
        lo = {j for j in nb[i] if j < i} | {j for k in nb[i] if k > i for j in nb[k] if j < i}
        got = set(int(j) for j in fci[frp[i]:frp[i + 1]])
        # entries may be dropped (|v| <= 1e-6 |u_ii|), never added
        assert got <= up | lo | {i} and i in got
    LU = sp.csr_matrix((fva, fci, frp), shape=(n, n)).toarray()
    L = np.tril(LU, -1) + np.eye(n)
    U = np.triu(LU)
    mask = np.zeros((n, n), bool)
    for i in range(n):
        mask[i, list(fci[frp[i]:frp[i + 1]])] = True
    prod = L @ U
    assert np.allclose(prod[mask], A[mask], rtol=1e-10, atol=1e-10 * np.abs(A).max())
    # the fill makes it a closer factorisation than ILU(0)
    f0 = O.OraclePrec("ilu", rp, ci, va, 0).factor()
    LU0 = sp.csr_matrix((f0[2], f0[1], f0[0]), shape=(n, n)).toarray()
    assert np.linalg.norm(prod - A) < np.linalg.norm((np.tril(LU0, -1) + np.eye(n)) @ np.triu(LU0) - A)


# ---- time(Backward) and user terms ----

@pytest.mark.parametrize("omega", [1.0, 0.4, 2.5])
def test_time_backward_is_exact_for_quadratics_in_time(omega):
    """time(Backward) (schemes/time/backward.rs:21-62) is variable-step BDF2: with
    w = dt/dt_2 the cell terms are ((1+2w)/(1+w)/dt, ((1+w) u_n - w^2/(1+w) u_{n-1})/dt),
    times the cell volume (assembly.rs:26-32).  BDF2 differentiates quadratics in time
    exactly, so time(Backward) - source(u'(t_{n+1})) returns u(t_{n+1}) for
    u(t) = a + b t + c t^2 sampled at t_{n-1} = 0, t_n = dt_2; a swapped level, a wrong
    sign or a coefficient for the other step ratio misses it.  For w = 1 the
    coefficients are the textbook 3/2, 2, 1/2."""
    m = meshgen.tet3d(3, seed=6)
    n = m["cell_ptr"].shape[0] - 1
    rng = np.random.default_rng(12)
    a, bb, c = rng.normal(size=n), rng.normal(size=n), rng.normal(size=n)
    dt_2 = 0.03
    dt = omega * dt_2
    u = lambda t: a + bb * t + c * t * t
    t1 = dt_2 + dt
    TB, PAR, SRC, SUB = O.OP["TIME_BACKWARD"], O.OP["PARAM"], O.OP["SOURCE"], O.OP["SUB"]
    prog = [(TB, 0, 1, -1, dt), (PAR, 0, 0, 0, dt_2), (SRC, 2, 0, 0, 0.0), (SUB, 0, 0, 0, 0.0)]
    om, g, A, b = _assemble_terms(m, prog, [], [u(dt_2), u(0.0), bb + 2.0 * c * t1])
    vol = g["cell_volume"]
    assert abs(A - sp.diags(A.diagonal())).max() == 0.0
    assert np.allclose(A.diagonal(), (1 + 2 * omega) / (1 + omega) / dt * vol, rtol=8 * EPS, atol=0)
    x = b / A.diagonal()
    assert np.allclose(x, u(t1), rtol=1e-11, atol=1e-12)
    if omega == 1.0:
        _, _, A1, b1 = _assemble_terms(m, prog[:2], [], [u(dt_2), u(0.0)])
        assert np.allclose(A1.diagonal(), 1.5 / dt * vol, rtol=4 * EPS, atol=0)
        assert np.allclose(b1, (2.0 * u(dt_2) - 0.5 * u(0.0)) / dt * vol, rtol=1e-12, atol=1e-13)


def test_time_backward_densities():
    """Backward::new_with_density: density scales the unknown's coefficient, density_last
    the previous level and density_last2 the one before (backward.rs:45-60); all three at
    one give the bits of Backward::new."""
    m = meshgen.tri2d(10, seed=2)
    n = m["cell_ptr"].shape[0] - 1
    rng = np.random.default_rng(13)
    p1, p2 = rng.normal(size=n), rng.normal(size=n)
    r0, r1, r2 = (rng.uniform(0.5, 2.0, n) for _ in range(3))
    dt, dt_2 = 0.02, 0.025
    TB, PAR = O.OP["TIME_BACKWARD"], O.OP["PARAM"]
    om, g, A, b = _assemble_terms(m, [(TB, 0, 1, 2, dt), (PAR, 0, 0, 0, dt_2)], [], [p1, p2, r0, r1, r2])
    vol = g["cell_volume"]
    w = dt / dt_2
    assert np.allclose(A.diagonal(), (1 + 2 * w) / (1 + w) / dt * r0 * vol, rtol=8 * EPS, atol=0)
    want = ((1 + w) * p1 * r1 - w * w / (1 + w) * p2 * r2) / dt * vol
    assert np.allclose(b, want, rtol=1e-12, atol=1e-12)
    ones = np.ones(n)
    _, _, A1, b1 = _assemble_terms(m, [(TB, 0, 1, 2, dt), (PAR, 0, 0, 0, dt_2)], [], [p1, p2, ones, ones, ones])
    _, _, A0, b0 = _assemble_terms(m, [(TB, 0, 1, -1, dt), (PAR, 0, 0, 0, dt_2)], [], [p1, p2])
    assert np.array_equal(A1.data, A0.data) and np.array_equal(b1, b0)


@pytest.mark.parametrize("mesh", ["tri", "tet"])
def test_user_term_reproduces_builtin_terms(mesh):
    """A user `impl Term` (terms/mod.rs:24-40) supplying the same face and cell terms as a
    built-in one assembles to the same bits: convection(Upwind) has face terms
    (flux if w > 0 else 0*flux, 0*flux if w > 0 else flux, -0*flux) (upwind.rs:26-33,
    convection.rs:36-44), source(f) has cell terms (0, -f) (source.rs:34-39).  Swapping
    li and lj, or l and r, breaks the match."""
    m = meshgen.tri2d(8, seed=4) if mesh == "tri" else meshgen.tet3d(3, seed=4)
    om = O.OracleMesh(m)
    n, F, nb = om.n_cells, om.n_faces, om.n_bnd
    rng = np.random.default_rng(14)
    flux, w, f = rng.normal(size=F), rng.normal(size=F), rng.normal(size=n)
    bci, bce = rng.uniform(0, 1, nb), rng.normal(size=nb)
    OP = O.OP
    ref = om.assemble([(OP["CONVECTION_UPWIND"], 0, 1, 0, 0.0), (OP["SOURCE"], 0, 0, 0, 0.0), (OP["SUB"], 0, 0, 0, 0.0)],
                      [flux, w], [f], bci, bce)
    li = np.where(w > 0.0, 1.0, 0.0) * flux
    lj = np.where(w > 0.0, 0.0, 1.0) * flux
    r = (-0.0) * flux
    prog = [(OP["USER_TERM"], 0, -1, 0, 0.0), (OP["USER_TERM"], -1, 0, 0, 0.0), (OP["SUB"], 0, 0, 0, 0.0)]
    got = om.assemble(prog, [li, lj, r], [np.zeros(n), -f], bci, bce)
    for k in range(4):
        assert np.array_equal(got[k], ref[k])
    swapped = om.assemble(prog, [lj, li, r], [np.zeros(n), -f], bci, bce)
    assert not np.array_equal(swapped[2], ref[2])
    swapped = om.assemble(prog, [li, lj, r], [-f, np.zeros(n)], bci, bce)
    assert not np.array_equal(swapped[3], ref[3])


# ---- SIMPLE helpers (tools/hbya.rs) and the wall distance (tools/wall_distance.rs) ----

def _momentum_like(m, seed):
    """A diagonally dominant matrix on the mesh pattern: time(Euler) + convection(Upwind)
    - laplacian(Orthogonal), assembled by the oracle."""
    om = O.OracleMesh(m)
    rng = np.random.default_rng(seed)
    n, F, nb = om.n_cells, om.n_faces, om.n_bnd
    OP = O.OP
    prog = [(OP["TIME_EULER"], 0, -1, -1, 0.01), (OP["CONVECTION_UPWIND"], 0, 0, 0, 0.0), (OP["ADD"], 0, 0, 0, 0.0),
            (OP["LAPLACIAN_ORTHOGONAL"], 1, 0, 0, 0.0), (OP["SUB"], 0, 0, 0, 0.0)]
    rp, ci, va, b = om.assemble(prog, [rng.normal(size=F), rng.uniform(0.5, 2.0, F)], [rng.normal(size=n)],
                                rng.uniform(0, 1, nb), rng.normal(size=nb))
    return om, rp, ci, va


@pytest.mark.parametrize("mesh", ["tri", "tet"])
def test_hbya_reproduces_the_velocity_that_solves_the_momentum_rows(mesh):
    """compute_hbya_ainv (hbya.rs:15-50): hbya_i = (mrhs_i - sum_{j != i} a_ij u_j) / a_ii.
    With mrhs = A u (component by component) that is u itself, up to rounding; ainv is
    vol_i / a_ii.  Dropping the diagonal exclusion or a sign breaks the first."""
    m = meshgen.tri2d(8, seed=2) if mesh == "tri" else meshgen.tet3d(3, seed=2)
    om, rp, ci, va = _momentum_like(m, 31)
    n, D = om.n_cells, om.dim
    A = sp.csr_matrix((va, ci, rp), shape=(n, n))
    u = np.random.default_rng(32).normal(size=(n, D))
    hb, ai = om.hbya_ainv(rp, ci, va, A @ u, u)
    assert np.allclose(hb, u, rtol=1e-11, atol=1e-12)
    assert np.array_equal(ai, om.geometry()["cell_volume"] / A.diagonal())
    with pytest.raises(RuntimeError):     # a row without its diagonal: the reference's index panics
        keep = ci != np.repeat(np.arange(n), np.diff(rp))
        rp2 = np.concatenate([[0], np.cumsum([keep[rp[i]:rp[i + 1]].sum() for i in range(n)])]).astype(np.int32)
        om.hbya_ainv(rp2, ci[keep], va[keep], A @ u, u)


@pytest.mark.parametrize("mesh", ["tri", "tet"])
def test_hbya_face_interpolation_closed_forms(mesh):
    """intepolate_hbya_ainv_faces (hbya.rs:56-102): a constant hbya U interpolates to U.n on
    every face with Linear and LimitedLinear weights (they sum to one), a constant ainv to
    itself; Upwind takes the upwind cell's value exactly; on a boundary face the owner's
    U.n unless both closures describe a wall ((0, u.n = 0) and (1, 0)), where it is 0."""
    m = meshgen.tri2d(8, seed=3) if mesh == "tri" else meshgen.tet3d(3, seed=3)
    om = O.OracleMesh(m)
    g = om.geometry()
    n, F, nb, D = om.n_cells, om.n_faces, om.n_bnd, om.dim
    rng = np.random.default_rng(33)
    U = rng.normal(size=D)
    hb = np.tile(U, (n, 1))
    ai = np.full(n, 0.37)
    walls = rng.uniform(size=nb) < 0.5
    ubi, ube = np.where(walls, 0.0, 1.0), np.zeros((nb, D))
    pbi, pbe = np.ones(nb), np.zeros(nb)
    pbi[~walls] = 0.0                       # outflow-like faces: fixed pressure, not a wall
    bf = np.nonzero(m["face_patch"] >= 0)[0]
    un = g["face_normal"] @ U
    flux, lim = rng.normal(size=F), rng.uniform(0, 1, F)
    for scheme in ("linear", ("limited_linear", flux, lim)):
        hn, af = om.interpolate_hbya_ainv_faces(hb, ai, scheme, scheme, ubi, ube, pbi, pbe)
        want = un.copy()
        want[bf[walls]] = 0.0
        assert np.allclose(hn, want, rtol=1e-12, atol=1e-13)
        assert np.allclose(af, 0.37, rtol=1e-14, atol=0)
        assert np.all(hn[bf[walls]] == 0.0) and np.array_equal(af[bf], np.full(nb, 0.37))
    hbv = rng.normal(size=(n, D))
    hn, af = om.interpolate_hbya_ainv_faces(hbv, rng.normal(size=n), ("upwind", flux), "linear", ubi, ube, pbi, pbe)
    inner = g["neighbor"] >= 0
    up = np.where(flux > 0, g["owner"], np.maximum(g["neighbor"], 0))
    want = (hbv[up] * g["face_normal"]).sum(1)
    assert np.allclose(hn[inner], want[inner], rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("mesh", ["tri", "tet"])
def test_correct_phi_with_the_exact_pressure_gradient(mesh):
    """correct_phi (hbya.rs:108-157) with Corrected and the exact gradient of a linear
    pressure p = a.x (exact Dirichlet values on the boundary) takes the face-normal
    gradient a.n on every face, so phi = hbyan - ainv a.n; Orthogonal does not on these
    non-orthogonal meshes."""
    m = meshgen.tri2d(8, seed=4) if mesh == "tri" else meshgen.tet3d(3, seed=4)
    om = O.OracleMesh(m)
    g = om.geometry()
    n, F, nb, D = om.n_cells, om.n_faces, om.n_bnd, om.dim
    rng = np.random.default_rng(34)
    a = rng.normal(size=D)
    p = g["cell_center"] @ a
    bf = np.nonzero(m["face_patch"] >= 0)[0]
    pbi, pbe = np.zeros(nb), g["face_center"][bf] @ a
    hn, af = rng.normal(size=F), rng.uniform(0.5, 2.0, F)
    grads = np.tile(a, (n, 1))
    phi = om.correct_phi(hn, af, p, pbi, pbe, corrected_grads=grads, reduction=1.0)
    want = hn - af * (g["face_normal"] @ a)
    assert np.allclose(phi, want, rtol=1e-10, atol=1e-11)
    phi_o = om.correct_phi(hn, af, p, pbi, pbe)
    assert not np.allclose(phi_o, want, rtol=1e-6, atol=1e-8)


def test_correct_velocity_undoes_the_pressure_correction():
    """correct_velocity (hbya.rs:161-179): velocity = hbya - ainv * grad p, so hbya = u +
    ainv * g gives back u."""
    om = O.OracleMesh(meshgen.tet3d(3, seed=5))
    n, D = om.n_cells, om.dim
    rng = np.random.default_rng(35)
    u, gp, ai = rng.normal(size=(n, D)), rng.normal(size=(n, D)), rng.uniform(0.1, 2.0, n)
    v = om.correct_velocity(u + ai[:, None] * gp, ai, gp)
    assert np.allclose(v, u, rtol=1e-14, atol=1e-14)


def test_wall_face_viscosity_and_normalisation_closed_forms():
    """wall_distance.rs:92-116 with p = 3.  The face diffusivity is |Linear interpolation
    of the gradients|, so a constant gradient a gives |a| on every face.  The final
    normalisation -|g|^2 + (3/2 u + |g|^3)^(2/3) is exact for the 1D p-Poisson solution
    with a wall at y = 0 and zero gradient at y = h: u = (2/3)(h^1.5 - (h - y)^1.5),
    |u'| = (h - y)^0.5, which it maps to y."""
    m = meshgen.tri2d(10, seed=6)
    om = O.OracleMesh(m)
    g = om.geometry()
    n = om.n_cells
    a = np.array([0.3, -1.2])
    visc = om.wall_face_viscosity(np.tile(a, (n, 1)))
    assert np.allclose(visc, np.linalg.norm(a), rtol=1e-14, atol=0)
    h = 1.25
    y = g["cell_center"][:, 1]
    u = (2.0 / 3.0) * (h ** 1.5 - (h - y) ** 1.5)
    grads = np.stack([np.zeros(n), np.sqrt(h - y)], axis=1)
    d = om.wall_normalize(grads, u)
    assert np.allclose(d, y, rtol=1e-12, atol=1e-13)


def test_wall_distance_loop_against_independent_solves():
    """compute_wall_distance (wall_distance.rs:59-105): the first pass solves
    -laplacian(Corrected(0, 1), 1) + source(-1) with walls at y = 0 and y = 1 (closure
    (0, 0)) and zero gradient elsewhere ((1, 0)); the second pass the same with the new
    gradients and face diffusivity |grad_f|.  Each pass is checked against scipy's direct
    solve of the oracle-assembled system, and the result against the normalisation of the
    last pass's solution with the gradients taken after it.  A tolerance above the first
    initial residual stops after one solve, with zero gradients in the normalisation."""
    m = meshgen.tri2d(12, seed=7)
    om = O.OracleMesh(m)
    n, F, nb = om.n_cells, om.n_faces, om.n_bnd
    patch = m["face_patch"][m["face_patch"] >= 0]
    wall = (patch == 0) | (patch == 2)
    bci, bce = np.where(wall, 0.0, 1.0), np.zeros(nb)
    OP = O.OP
    prog = [(OP["LAPLACIAN_CORRECTED"], 0, 1, 0, 1.0), (OP["NEG"], 0, 0, 0, 0.0), (OP["SOURCE"], 0, 0, 0, 0.0),
            (OP["ADD"], 0, 0, 0, 0.0)]
    visc, grads, src = np.ones(F), np.zeros(n * 2), np.full(n, -1.0)
    for rounds in (1, 2):
        rp, ci, va, b = om.assemble(prog, [visc], [src, grads], bci, bce)
        u = spl.spsolve(sp.csr_matrix((va, ci, rp), shape=(n, n)).tocsc(), b)
        grads = om.gradients(u, bci, bce).ravel()
        visc = om.wall_face_viscosity(grads.reshape(n, 2))
        d, info = om.wall_distance(wall, rounds, 1e-30, 1e-12, 0.0, 2000)
        assert info["solves"] == rounds
        want = om.wall_normalize(grads.reshape(n, 2), u)
        assert np.allclose(d, want, rtol=1e-8, atol=1e-10), rounds
    d1, info = om.wall_distance(wall, 5, 10.0, 1e-12, 0.0, 2000)
    rp, ci, va, b = om.assemble(prog, [np.ones(F)], [src, np.zeros(n * 2)], bci, bce)
    u1 = spl.spsolve(sp.csr_matrix((va, ci, rp), shape=(n, n)).tocsc(), b)
    assert info["solves"] == 1
    assert np.allclose(d1, (1.5 * u1) ** (2.0 / 3.0), rtol=1e-8, atol=1e-10)
