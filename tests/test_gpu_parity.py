"""GPU parity: libfvg.so against the CPU FP64 oracle on the same seeded inputs.

Bars (from the workload definition): CSR row_ptr/col_idx bit-exact, assembled values
within 1e-12 relative, SpMV / factors / preconditioner sweeps bit-exact (same
operation order), solutions within 1e-6 relative at a 1e-8 relative residual and
iteration counts within +-2.
"""
import numpy as np
import pytest

import fvg
import meshgen
import oracle as O
import problems

pytestmark = pytest.mark.gpu

MESHES = {
    "tri16": lambda: meshgen.tri2d(16, seed=3),
    "tri64": lambda: meshgen.tri2d(64, seed=1),
    "tet4": lambda: meshgen.tet3d(4, seed=5),
    "tet10": lambda: meshgen.tet3d(10, seed=7),
    "prism8": lambda: meshgen.prism3d(8, layers=4, seed=2),
}


def _mk(name):
    raw = MESHES[name]()
    om = O.OracleMesh(raw)
    return raw, om, om.geometry()


def _assemble_both(ctx, raw, om, geom, prob):
    mesh = fvg.Mesh(ctx, raw)
    orc = om.assemble(prob["prog"], prob["face_fields"], prob["cell_fields"], prob["bc_i"], prob["bc_e"])
    lhs, rhs = fvg.assemble_program(mesh, prob["prog"], prob["face_fields"], prob["cell_fields"], prob["bc_i"],
                                    prob["bc_e"])
    return mesh, orc, lhs, rhs


def _values_close(gpu, ref, rel=1e-12):
    scale = np.abs(ref).max()
    return np.all(np.abs(gpu - ref) <= rel * np.abs(ref) + 1e-15 * scale)


@pytest.mark.parametrize("name", ["tri16", "tet4", "prism8"])
def test_geometry_bit_exact(ctx, name):
    raw, om, geom = _mk(name)
    mesh = fvg.Mesh(ctx, raw)
    g = mesh.geometry()
    for k in ("face_area", "face_center", "face_normal", "owner", "neighbor", "cell_volume", "cell_center"):
        assert np.array_equal(g[k], geom[k]), k


@pytest.mark.parametrize("name", ["tri64", "tet10", "prism8"])
def test_assembly_diffusion(ctx, name):
    raw, om, geom = _mk(name)
    n = om.n_cells
    kc = meshgen.lognormal_cells(n, 1.0, seed=11)
    prob = problems.diffusion(raw, geom, kc=kc, seed=4, bc="x")
    mesh, (rp, ci, va, rhs_o), lhs, rhs = _assemble_both(ctx, raw, om, geom, prob)
    grp, gci, gva = lhs.csr()
    assert np.array_equal(grp, rp) and np.array_equal(gci, ci)
    assert _values_close(gva, va)
    assert _values_close(rhs.numpy(), rhs_o)


@pytest.mark.parametrize("name", ["tri64", "prism8"])
def test_assembly_convection_diffusion(ctx, name):
    raw, om, geom = _mk(name)
    prob = problems.convection_diffusion(raw, geom, seed=9)
    mesh, (rp, ci, va, rhs_o), lhs, rhs = _assemble_both(ctx, raw, om, geom, prob)
    grp, gci, gva = lhs.csr()
    assert np.array_equal(grp, rp) and np.array_equal(gci, ci)
    assert _values_close(gva, va)
    assert _values_close(rhs.numpy(), rhs_o)


def test_assembly_term_algebra(ctx):
    """time(Euler) + linear_source + divergence - 2-level nested algebra, via the term API."""
    raw, om, geom = _mk("tri16")
    n, F = om.n_cells, om.n_faces
    rng = np.random.default_rng(5)
    prev, lin, dens = rng.uniform(0, 1, n), rng.uniform(0, 2, n), rng.uniform(1, 2, n)
    flux, kf = rng.normal(size=F), rng.uniform(0.5, 1.5, F)
    eq = (fvg.time(fvg.Euler(prev, 0.1, dens, dens)) - fvg.linear_source(lin)) + \
        -(fvg.laplacian(fvg.Orthogonal(), kf) - fvg.divergence(flux))
    ops, ff, cf = eq.compile()
    nb = om.n_bnd
    bci, bce = rng.uniform(0, 1, nb), rng.uniform(0, 1, nb)
    mesh = fvg.Mesh(ctx, raw)
    lhs, rhs = fvg.assemble(eq, (bci, bce), mesh)
    rp, ci, va, rhs_o = om.assemble(ops, ff, cf, bci, bce)
    assert _values_close(lhs.csr()[2], va)
    assert _values_close(rhs.numpy(), rhs_o)


@pytest.mark.parametrize("name", ["tri64", "tet10"])
def test_spmv_bit_exact(ctx, name):
    raw, om, geom = _mk(name)
    prob = problems.diffusion(raw, geom, kc=meshgen.lognormal_cells(om.n_cells, 1.0, 2), seed=1)
    rp, ci, va, _ = om.assemble(prob["prog"], prob["face_fields"], prob["cell_fields"], prob["bc_i"], prob["bc_e"])
    A = fvg.DistributedMatrix.from_csr(ctx, rp, ci, va)
    x = np.random.default_rng(0).normal(size=om.n_cells)
    X = fvg.DistributedVector.from_data(ctx, x)
    Y = fvg.DistributedVector(ctx, om.n_cells)
    A.imul(Y, X)
    assert np.array_equal(Y.numpy(), O.spmv(rp, ci, va, x))


def _oracle_system(name, kind):
    raw, om, geom = _mk(name)
    if kind == "sym":
        prob = problems.diffusion(raw, geom, kc=meshgen.lognormal_cells(om.n_cells, 1.0, 8), seed=2, bc="x")
    else:
        prob = problems.convection_diffusion(raw, geom, seed=3)
    rp, ci, va, b = om.assemble(prob["prog"], prob["face_fields"], prob["cell_fields"], prob["bc_i"], prob["bc_e"])
    return rp, ci, va, b


@pytest.mark.parametrize("name", ["tri64", "tet10"])
def test_ic0_factor_and_sweeps_bit_exact(ctx, name):
    rp, ci, va, b = _oracle_system(name, "sym")
    A = fvg.DistributedMatrix.from_csr(ctx, rp, ci, va)
    ic = fvg.IncompleteCholesky.from_matrix(A, 0)
    P = O.OraclePrec("ic", rp, ci, va, 0)
    lrp, lci, lva, linv = P.factor()
    gvals, ginv = ic.factor_values()
    # gather the GPU factor on the oracle's L pattern (lower part of A's pattern)
    n = rp.shape[0] - 1
    low = np.concatenate([np.arange(rp[i], rp[i + 1])[ci[rp[i]:rp[i + 1]] <= i] for i in range(n)])
    assert np.array_equal(ci[low], lci)
    assert np.array_equal(gvals[low], lva)
    assert np.array_equal(ginv, linv)
    r = np.random.default_rng(1).normal(size=n)
    z = fvg.DistributedVector(ctx, n)
    ic.precondition(z, fvg.DistributedVector.from_data(ctx, r))
    assert np.array_equal(z.numpy(), P.apply(r))


@pytest.mark.parametrize("name", ["tri64", "prism8"])
def test_ilu0_factor_and_sweeps_bit_exact(ctx, name):
    rp, ci, va, b = _oracle_system(name, "nonsym")
    A = fvg.DistributedMatrix.from_csr(ctx, rp, ci, va)
    ilu = fvg.IncompleteLowerUpper.from_matrix(A, 0)
    P = O.OraclePrec("ilu", rp, ci, va, 0)
    lrp, lci, lva, _ = P.factor()
    gvals, _ = ilu.factor_values()
    n = rp.shape[0] - 1
    for i in range(n):
        cols = ci[rp[i]:rp[i + 1]]
        g = gvals[rp[i]:rp[i + 1]]
        ocols = lci[lrp[i]:lrp[i + 1]]
        ovals = lva[lrp[i]:lrp[i + 1]]
        keep = np.isin(cols, ocols)
        assert np.array_equal(cols[keep], ocols)
        assert np.array_equal(g[keep], ovals)
        assert np.all(g[~keep] == 0.0)   # dropped entries are explicit zeros
    r = np.random.default_rng(2).normal(size=n)
    z = fvg.DistributedVector(ctx, n)
    ilu.precondition(z, fvg.DistributedVector.from_data(ctx, r))
    assert np.array_equal(z.numpy(), P.apply(r))


def _solve_pair(ctx, rp, ci, va, b, solver, prec, rtol=1e-8):
    n = rp.shape[0] - 1
    P = O.OraclePrec(prec, rp, ci, va, 0)
    xo, io = O.solve(solver, rp, ci, va, P, b, rtol=rtol, atol=0.0, max_iter=2000)
    A = fvg.DistributedMatrix.from_csr(ctx, rp, ci, va)
    M = (fvg.IncompleteCholesky if prec == "ic" else fvg.IncompleteLowerUpper).from_matrix(A, 0)
    x = fvg.DistributedVector(ctx, n)
    B = fvg.DistributedVector.from_data(ctx, b)
    opts = fvg.LinearSolverOptions(relative_tolerance=rtol, absolute_tolerance=0.0, max_iterations=2000)
    fn = fvg.conjugate_gradient if solver == "cg" else fvg.bi_conjugate_gradient_stab
    info = fn(x, A, B, M, opts)
    return xo, io, x.numpy(), info


@pytest.mark.parametrize("name", ["tri64", "tet10"])
def test_cg_ic0_parity(ctx, name):
    rp, ci, va, b = _oracle_system(name, "sym")
    xo, io, xg, info = _solve_pair(ctx, rp, ci, va, b, "cg", "ic")
    assert abs(info.iterations - io["iterations"]) <= 2, (info.iterations, io["iterations"])
    assert np.linalg.norm(xg - xo) <= 1e-6 * np.linalg.norm(xo)
    assert info.final_residual / info.initial_residual < 1e-8
    assert abs(info.initial_residual - io["initial_residual"]) <= 1e-12 * io["initial_residual"]
    assert len(info.history) == info.iterations


@pytest.mark.parametrize("name", ["tri64", "prism8"])
def test_bicgstab_ilu0_parity(ctx, name):
    rp, ci, va, b = _oracle_system(name, "nonsym")
    xo, io, xg, info = _solve_pair(ctx, rp, ci, va, b, "bicgstab", "ilu")
    assert abs(info.iterations - io["iterations"]) <= 2, (info.iterations, io["iterations"])
    assert np.linalg.norm(xg - xo) <= 1e-6 * np.linalg.norm(xo)


def test_end_to_end_cfg1_reference_api(ctx):
    """The 64x64 2D workload through the reference-shaped API: assemble -> IC(0) -> CG."""
    raw = meshgen.tri2d(64, jitter=0.25, seed=1)
    om = O.OracleMesh(raw)
    geom = om.geometry()
    n = om.n_cells
    f = np.random.default_rng(2).uniform(0.0, 1.0, n)
    kf = np.ones(om.n_faces)
    mesh = fvg.Mesh.from_arrays(ctx, raw)
    eq = -fvg.laplacian(fvg.Orthogonal(), kf) - fvg.source(f)
    lhs, rhs = fvg.assemble(eq, fvg.dirichlet(np.zeros(om.n_bnd)), mesh)
    ic = fvg.IncompleteCholesky.from_matrix(lhs, 0)
    x = fvg.DistributedVector.from_size(ctx, n)
    info = fvg.conjugate_gradient(x, lhs, rhs, ic, fvg.LinearSolverOptions(1e-8, 0.0, 1000))
    ops, ff, cf = eq.compile()
    rp, ci, va, b = om.assemble(ops, ff, cf, np.zeros(om.n_bnd), np.zeros(om.n_bnd))
    P = O.OraclePrec("ic", rp, ci, va, 0)
    xo, io = O.solve("cg", rp, ci, va, P, b, rtol=1e-8, atol=0.0)
    assert abs(info.iterations - io["iterations"]) <= 2
    xg = x.numpy()
    assert np.linalg.norm(xg - xo) <= 1e-6 * np.linalg.norm(xo)
    assert "ConjugateGradient" in str(info)


def test_zero_rhs_and_max_iterations(ctx):
    rp, ci, va, b = _oracle_system("tri16", "sym")
    n = rp.shape[0] - 1
    A = fvg.DistributedMatrix.from_csr(ctx, rp, ci, va)
    M = fvg.IncompleteCholesky.from_matrix(A, 0)
    x = fvg.DistributedVector.from_data(ctx, np.ones(n))
    info = fvg.conjugate_gradient(x, A, fvg.DistributedVector(ctx, n), M, fvg.LinearSolverOptions(1e-8, 0.0, 100))
    assert info.iterations == 0 and np.all(x.numpy() == 0.0) and len(info.history) == 1
    # max_iterations = 3 runs 4 iterations and still returns (conjugate_gradient.rs:67)
    x = fvg.DistributedVector(ctx, n)
    info = fvg.conjugate_gradient(x, A, fvg.DistributedVector.from_data(ctx, b), M,
                                  fvg.LinearSolverOptions(1e-30, 0.0, 3))
    P = O.OraclePrec("ic", rp, ci, va, 0)
    _, io = O.solve("cg", rp, ci, va, P, b, rtol=1e-30, atol=0.0, max_iter=3)
    assert info.iterations == io["iterations"] == 4


def test_error_paths(ctx):
    # IC on a nonsymmetric matrix -> the reference's assert (incomplete_cholesky.rs:153)
    rp, ci, va, b = _oracle_system("tri64", "nonsym")
    A = fvg.DistributedMatrix.from_csr(ctx, rp, ci, va)
    with pytest.raises(fvg.NotSymmetricError):
        fvg.IncompleteCholesky.from_matrix(A, 0)
    # ILU zero pivot -> the reference's panic (incomplete_lower_upper.rs:234-236)
    rp2 = np.array([0, 2, 4], np.int32)
    ci2 = np.array([0, 1, 0, 1], np.int32)
    va2 = np.array([0.0, 1.0, 1.0, 1.0])
    with pytest.raises(fvg.ZeroPivotError):
        fvg.IncompleteLowerUpper.from_matrix(fvg.DistributedMatrix.from_csr(ctx, rp2, ci2, va2), 0)
    with pytest.raises(O.ZeroPivot):
        O.OraclePrec("ilu", rp2, ci2, va2, 0)
    # HBM budget
    small = fvg.Context(0, hbm_budget=1 << 20)
    with pytest.raises(fvg.BudgetError):
        fvg.DistributedVector(small, 1 << 20)
    small.close()
