"""GPU parity: libfvg.so against the CPU FP64 oracle on the same seeded inputs.

Bars (from the workload definition): CSR row_ptr/col_idx bit-exact, assembled values
within 1e-12 relative, SpMV / factors / preconditioner sweeps bit-exact (same
operation order), solutions within 1e-6 relative at a 1e-8 relative residual and
iteration counts within +-2.  Geometry and assembled values are held to the tighter
bar of bit-exactness: the device applies every face in the reference's order.
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
    "tet16": lambda: meshgen.tet3d(16, seed=9),
    "prism8": lambda: meshgen.prism3d(8, layers=4, seed=2),
    "prism24": lambda: meshgen.prism3d(24, layers=12, seed=2),
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
    assert np.array_equal(gva, va)
    assert np.array_equal(rhs.numpy(), rhs_o)


@pytest.mark.parametrize("name", ["tri64", "prism8"])
def test_assembly_convection_diffusion(ctx, name):
    raw, om, geom = _mk(name)
    prob = problems.convection_diffusion(raw, geom, seed=9)
    mesh, (rp, ci, va, rhs_o), lhs, rhs = _assemble_both(ctx, raw, om, geom, prob)
    grp, gci, gva = lhs.csr()
    assert np.array_equal(grp, rp) and np.array_equal(gci, ci)
    assert np.array_equal(gva, va)
    assert np.array_equal(rhs.numpy(), rhs_o)


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
    assert np.array_equal(lhs.csr()[2], va)
    assert np.array_equal(rhs.numpy(), rhs_o)


def _cell_lists(raw):
    cp, cf = raw["cell_ptr"], raw["cell_faces"]
    return [list(cf[cp[c]:cp[c + 1]]) for c in range(len(cp) - 1)]


def _with_cells(raw, lists, **over):
    r = dict(raw, **over)
    r["cell_ptr"] = np.concatenate([[0], np.cumsum([len(x) for x in lists])]).astype(np.int64)
    r["cell_faces"] = np.concatenate([np.asarray(x, np.int32) for x in lists])
    return r


def test_mesh_listing_edge_cases(ctx):
    """Face lists in shuffled order, a boundary face listed twice by its cell and an
    interior face listed by a third cell (add_cell ignores it, FaceRef::other_cell still
    adds the neighbour to that cell's pattern row): owners, neighbours, geometry, the
    pattern and the assembled values against the oracle, bit for bit."""
    raw = meshgen.tri2d(6, seed=4)
    g0 = O.OracleMesh(raw).geometry()
    fp = raw["face_patch"]
    lists = _cell_lists(raw)
    rng = np.random.default_rng(8)
    for x in lists:
        rng.shuffle(x)
    fb = int(np.nonzero(fp >= 0)[0][3])
    lists[int(g0["owner"][fb])].append(fb)
    fi = int(np.nonzero(fp < 0)[0][5])
    i, j = int(g0["owner"][fi]), int(g0["neighbor"][fi])
    k = max(c for c in range(len(lists)) if c not in (i, j))
    assert k > j
    lists[k].append(fi)
    edited = _with_cells(raw, lists)
    om = O.OracleMesh(edited)
    geom = om.geometry()
    mesh = fvg.Mesh(ctx, edited)
    g = mesh.geometry()
    for key in ("face_area", "face_center", "face_normal", "owner", "neighbor", "cell_volume", "cell_center"):
        assert np.array_equal(g[key], geom[key]), key
    prob = problems.diffusion(edited, geom, kc=meshgen.lognormal_cells(om.n_cells, 1.0, seed=2), seed=4, bc="x")
    _, (rp, ci, va, rhs_o), lhs, rhs = _assemble_both(ctx, edited, om, geom, prob)
    grp, gci, gva = lhs.csr()
    assert j in gci[grp[k]:grp[k + 1]]
    assert np.array_equal(grp, rp) and np.array_equal(gci, ci)
    assert np.array_equal(gva, va)
    assert np.array_equal(rhs.numpy(), rhs_o)


def test_mesh_errors(ctx):
    """Every invalid mesh is refused with FVG_E_MESH naming the face or cell a sequential
    build meets first; the oracle (the reference's asserts and Err returns) refuses the
    ones it checks too."""
    raw = meshgen.tri2d(3, seed=2)
    fp, fptr, fn = raw["face_patch"], raw["face_ptr"], raw["face_nodes"]
    F, lists = len(fp), _cell_lists(raw)

    def refused(r, text, oracle_too):
        with pytest.raises(fvg.MeshError, match=text):
            fvg.Mesh(ctx, r)
        if oracle_too:
            with pytest.raises(RuntimeError):
                O.OracleMesh(r)

    # a face that no cell lists (compute.rs:15-18)
    refused(dict(raw, face_ptr=np.append(fptr, fptr[-1] + 2), face_nodes=np.append(fn, [0, 1]).astype(np.int32),
                 face_patch=np.append(fp, 0).astype(np.int32)), f"face {F} has no owner", True)
    # an interior face with a single cell (compute.rs:21-31)
    fb = int(np.nonzero(fp >= 0)[0][2])
    p2 = fp.copy()
    p2[fb] = -1
    refused(dict(raw, face_patch=p2), f"interior face {fb} has no neighbour", True)
    # the same face listed twice by one cell as an interior face: owner == neighbour
    l2 = [list(x) for x in lists]
    owner_b = next(c for c, x in enumerate(lists) if fb in x)
    l2[owner_b].append(fb)
    refused(_with_cells(raw, l2, face_patch=p2), f"interior face {fb} has one cell as both", False)
    # a 2D face with three nodes (compute.rs:271-276)
    fptr3 = fptr.copy()
    fptr3[1:] += 1
    fn3 = np.insert(fn, 2, fn[0]).astype(np.int32)
    refused(dict(raw, face_ptr=fptr3, face_nodes=fn3), "degenerate face 0", True)
    # a node index out of range
    fn4 = fn.copy()
    fn4[7] = raw["nodes"].shape[0]
    refused(dict(raw, face_nodes=fn4), "face 3 names a node out of range", False)
    # a cell naming a face out of range
    l5 = [list(x) for x in lists]
    l5[2][0] = F
    refused(_with_cells(raw, l5), "cell 2 names a face out of range", False)
    # offsets that decrease are refused before anything reaches the device
    cp = raw["cell_ptr"].copy()
    cp[1] = cp[2] + 1
    with pytest.raises(fvg.FvgError, match="FVG_E_INVALID"):
        fvg.Mesh(ctx, dict(raw, cell_ptr=cp))


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


def test_bicgstab_restart_branch(ctx):
    """rho collapses to 0 after the first iteration of A x = e1 with A = [[2,1,-1],[1,4,1],
    [1,0,2]] and no preconditioner (exact in binary; the oracle's pin of the restart branch,
    bi_conjugate_gradient_stab.rs:155-169): device-wide and batched FP64 BiCGStab restart
    once, as the oracle does, and reach its solution."""
    A = np.array([[2.0, 1.0, -1.0], [1.0, 4.0, 1.0], [1.0, 0.0, 2.0]])
    rp = np.array([0, 3, 6, 8], np.int32)
    ci = np.array([0, 1, 2, 0, 1, 2, 0, 2], np.int32)
    va = np.array([2.0, 1.0, -1.0, 1.0, 4.0, 1.0, 1.0, 2.0])
    b = np.array([1.0, 0.0, 0.0])
    P = O.OraclePrec("identity", rp, ci, va, 0)
    xo, io = O.solve("bicgstab", rp, ci, va, P, b, rtol=1e-12, atol=0.0, max_iter=50)
    assert io["restarts"] == 1
    M = fvg.DistributedMatrix.from_csr(ctx, rp, ci, va)
    I = fvg.IdentityPreconditioner.new(ctx, 3)
    B = fvg.DistributedVector.from_data(ctx, b)
    opts = fvg.LinearSolverOptions(1e-12, 0.0, 50)
    x = fvg.DistributedVector(ctx, 3)
    info = fvg.bi_conjugate_gradient_stab(x, M, B, I, opts)
    assert info.restarts == io["restarts"]
    assert abs(info.iterations - io["iterations"]) <= 2
    assert np.abs(x.numpy() - np.linalg.solve(A, b)).max() < 1e-12
    batch = fvg.Batch(ctx, [M], [I], "bicgstab", opts)
    try:
        xb = fvg.DistributedVector(ctx, 3)
        (bi,) = batch.solve([B], [xb])
        assert bi.restarts == io["restarts"]
        assert np.abs(xb.numpy() - np.linalg.solve(A, b)).max() < 1e-12
    finally:
        batch.close()


@pytest.mark.parametrize("name,kind,solver,prec", [("tet10", "sym", "cg", "ic"), ("prism24", "nonsym", "bicgstab", "ilu")])
def test_position_order_solve_matches_natural_order(ctx, name, kind, solver, prec):
    """With both sweep layouts blocked and the upper one the lower one mirrored (symmetric
    patterns), FP64 solves iterate with every vector in the lower sweep's position order
    (FVG_TUNE_SOLVE_ORDER 0, the default) and launch one kernel per triangular solve; in
    natural order (1) each solve also gathers its right-hand side (and the last scatters).
    Element by element the iteration arithmetic is the same and the dot products are
    accumulated in double-double, so both give the same iterations and x to 1e-12, and
    both meet the oracle's bars."""
    rp, ci, va, b = _oracle_system(name, kind)
    n = rp.shape[0] - 1
    P = O.OraclePrec(prec, rp, ci, va, 0)
    A = fvg.DistributedMatrix.from_csr(ctx, rp, ci, va)
    M = (fvg.IncompleteCholesky if prec == "ic" else fvg.IncompleteLowerUpper).from_matrix(A, 0)
    lay = M.sweep_layout()
    assert lay["lower_blocks"] > 0 and lay["upper_blocks"] > 0 and lay["mirrored"] == 1
    B = fvg.DistributedVector.from_data(ctx, b)
    opts = fvg.LinearSolverOptions(1e-8, 0.0, 2000)
    fn = fvg.conjugate_gradient if solver == "cg" else fvg.bi_conjugate_gradient_stab
    out = {}
    try:
        for order in (1, 0):
            ctx.set_tuning(fvg.TUNE_SOLVE_ORDER, order)
            x = fvg.DistributedVector.from_data(ctx, np.full(n, 0.5))
            ctx.enable_timing(True)
            ctx.kernel_stats(reset=True)
            info = fn(x, A, B, M, opts)
            st = ctx.kernel_stats(reset=True)
            ctx.enable_timing(False)
            out[order] = (x.numpy(), info, st)
    finally:
        ctx.set_tuning(fvg.TUNE_SOLVE_ORDER, 0)
        ctx.enable_timing(False)
    (xn, inat, snat), (xp, ipos, spos) = out[1], out[0]
    xo, io = O.solve(solver, rp, ci, va, P, b, x0=np.full(n, 0.5), rtol=1e-8, atol=0.0, max_iter=2000)
    for xg, ig in ((xn, inat), (xp, ipos)):
        assert abs(ig.iterations - io["iterations"]) <= 2, (ig.iterations, io["iterations"])
        assert np.linalg.norm(xg - xo) <= 1e-6 * np.linalg.norm(xo)
        assert ig.final_residual <= 1e-8 * ig.initial_residual
    assert ipos.iterations == inat.iterations and ipos.restarts == inat.restarts
    assert np.linalg.norm(xp - xn) <= 1e-12 * np.linalg.norm(xn)
    assert spos["sweep_launches"] == spos["sweep_calls"]          # one kernel per triangular solve
    assert snat["sweep_launches"] > snat["sweep_calls"]


@pytest.mark.parametrize("name", ["tet10", "tet16"])
def test_fused_cg_iteration_matches_separate_kernels(ctx, name):
    """In position order, FP64 CG does its residual update and r.r inside the lower sweep
    and r.z inside the upper sweep (FVG_TUNE_CG_FUSION 0, the default) instead of in
    separate kernels (1).  The arithmetic per element is the same and the dot products are
    double-double sums, so x, the iteration count and the residual history are the same
    bits, and both meet the oracle's bars."""
    rp, ci, va, b = _oracle_system(name, "sym")
    n = rp.shape[0] - 1
    A = fvg.DistributedMatrix.from_csr(ctx, rp, ci, va)
    M = fvg.IncompleteCholesky.from_matrix(A, 0)
    lay = M.sweep_layout()
    assert lay["mirrored"] == 1
    B = fvg.DistributedVector.from_data(ctx, b)
    opts = fvg.LinearSolverOptions(1e-10, 0.0, 2000)
    out = {}
    try:
        for mode in (1, 0):
            ctx.set_tuning(fvg.TUNE_CG_FUSION, mode)
            x = fvg.DistributedVector.from_data(ctx, np.full(n, 0.25))
            info = fvg.conjugate_gradient(x, A, B, M, opts)
            out[mode] = (x.numpy(), info)
    finally:
        ctx.set_tuning(fvg.TUNE_CG_FUSION, 0)
    (xs, isep), (xf, ifus) = out[1], out[0]
    assert ifus.iterations == isep.iterations
    assert np.array_equal(ifus.history, isep.history)
    assert np.array_equal(xf, xs)
    P = O.OraclePrec("ic", rp, ci, va, 0)
    xo, io = O.solve("cg", rp, ci, va, P, b, x0=np.full(n, 0.25), rtol=1e-10, atol=0.0, max_iter=2000)
    assert abs(ifus.iterations - io["iterations"]) <= 2
    assert np.linalg.norm(xf - xo) <= 1e-6 * np.linalg.norm(xo)
