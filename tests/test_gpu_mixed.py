"""Mixed precision (FP32 inner iterations, FP64 refinement) and the per-level sweep mode.

The mixed solver is not the reference algorithm, so its bar is the FP64 oracle's
solution: true relative residual <= 1e-8 and x within 1e-6 relative of the oracle's x
(stated for every case below), with the inner factor applied both ways: exact
level-scheduled sweeps and Jacobi-iterated triangular solves."""
import numpy as np
import pytest
import scipy.sparse as sp

import fvg
import oracle as O
from test_gpu_parity import _oracle_system

pytestmark = pytest.mark.gpu


# inner factor application: (exact level-scheduled sweeps, Jacobi sweeps per factor; 0 = default)
INNER = {"jacobi": (0, 0), "jacobi1": (0, 1), "exact": (1, 0)}


@pytest.mark.parametrize("inner", sorted(INNER))
@pytest.mark.parametrize("name,kind,solver,prec", [("tri64", "sym", "cg", "ic"), ("tet10", "sym", "cg", "ic"),
                                                   ("tri64", "nonsym", "bicgstab", "ilu"),
                                                   ("prism8", "nonsym", "bicgstab", "ilu")])
def test_mixed_precision_reaches_fp64_accuracy(ctx, name, kind, solver, prec, inner):
    exact, sweeps = INNER[inner]
    ctx.set_tuning(fvg.TUNE_INNER_EXACT, exact)
    ctx.set_tuning(fvg.TUNE_INNER_SWEEPS, sweeps)
    try:
        _check_mixed(ctx, name, kind, solver, prec)
    finally:
        ctx.set_tuning(fvg.TUNE_INNER_EXACT, 0)
        ctx.set_tuning(fvg.TUNE_INNER_SWEEPS, 0)


def _check_mixed(ctx, name, kind, solver, prec):
    rp, ci, va, b = _oracle_system(name, kind)
    n = rp.shape[0] - 1
    P = O.OraclePrec(prec, rp, ci, va, 0)
    xo, io = O.solve(solver, rp, ci, va, P, b, rtol=1e-8, atol=0.0, max_iter=3000)
    A = fvg.DistributedMatrix.from_csr(ctx, rp, ci, va)
    M = (fvg.IncompleteCholesky if prec == "ic" else fvg.IncompleteLowerUpper).from_matrix(A, 0)
    x = fvg.DistributedVector(ctx, n)
    B = fvg.DistributedVector.from_data(ctx, b)
    opts = fvg.LinearSolverOptions(1e-8, 0.0, 3000, precision=fvg.PRECISION_MIXED)
    fn = fvg.conjugate_gradient if solver == "cg" else fvg.bi_conjugate_gradient_stab
    info = fn(x, A, B, M, opts)
    xg = x.numpy()
    S = sp.csr_matrix((va, ci, rp), shape=(n, n))
    true_rel = np.linalg.norm(b - S @ xg) / np.linalg.norm(b)
    assert true_rel <= 1e-8, true_rel
    assert abs(info.true_residual - true_rel) <= 1e-3 * true_rel + 1e-15
    assert np.linalg.norm(xg - xo) <= 1e-6 * np.linalg.norm(xo)
    assert info.outer_iterations >= 2
    # the FP32 inner solver stays within a small multiple of the FP64 iteration count
    # (one Jacobi sweep per factor is the weakest inner preconditioner tested)
    assert info.iterations <= 4 * io["iterations"] + 20


@pytest.mark.parametrize("solver,prec,kind", [("cg", "ic", "sym"), ("bicgstab", "ilu", "nonsym")])
def test_mixed_graph_replay_and_direct_launches_agree(ctx, solver, prec, kind):
    """The inner iterations replayed from a captured graph and launched one by one run the
    same kernels in the same order: identical solutions, bit for bit."""
    rp, ci, va, b = _oracle_system("tet10", kind)
    n = rp.shape[0] - 1
    A = fvg.DistributedMatrix.from_csr(ctx, rp, ci, va)
    M = (fvg.IncompleteCholesky if prec == "ic" else fvg.IncompleteLowerUpper).from_matrix(A, 0)
    B = fvg.DistributedVector.from_data(ctx, b)
    opts = fvg.LinearSolverOptions(1e-8, 0.0, 3000, precision=fvg.PRECISION_MIXED)
    fn = fvg.conjugate_gradient if solver == "cg" else fvg.bi_conjugate_gradient_stab
    out = []
    try:
        for direct in (0, 1, 0):
            ctx.set_tuning(fvg.TUNE_GRAPHS, direct)
            x = fvg.DistributedVector(ctx, n)
            info = fn(x, A, B, M, opts)
            out.append((x.numpy(), info.iterations, info.outer_iterations))
    finally:
        ctx.set_tuning(fvg.TUNE_GRAPHS, 0)
    for xg, it, outer in out[1:]:
        assert np.array_equal(xg, out[0][0]) and it == out[0][1] and outer == out[0][2]


@pytest.mark.parametrize("mode,waves,poll", [(1, 0, 0), (2, 4, 0), (2, 64, 0), (2, 0, 0), (2, 0, 1), (2, 0, 2),
                                             (0, 0, 0)])
def test_sweep_launch_shapes_give_identical_bits(ctx, mode, waves, poll):
    """One launch per level, the persistent sweep at several wave counts and poll modes,
    and the blocked sweep all give the oracle's bits (the schedule never changes the
    arithmetic)."""
    rp, ci, va, b = _oracle_system("tet10", "sym")
    n = rp.shape[0] - 1
    A = fvg.DistributedMatrix.from_csr(ctx, rp, ci, va)
    ic = fvg.IncompleteCholesky.from_matrix(A, 0)
    ref = O.OraclePrec("ic", rp, ci, va, 0).apply(np.random.default_rng(1).normal(size=n))
    try:
        ctx.set_tuning(fvg.TUNE_SWEEP_MODE, mode)
        ctx.set_tuning(fvg.TUNE_SWEEP_WAVES, waves)
        ctx.set_tuning(fvg.TUNE_SWEEP_POLL, poll)
        if mode == 0:
            assert ic.sweep_layout()["lower_blocks"] > 0
        z = fvg.DistributedVector(ctx, n)
        ic.precondition(z, fvg.DistributedVector.from_data(ctx, np.random.default_rng(1).normal(size=n)))
        assert np.array_equal(z.numpy(), ref)
    finally:
        ctx.set_tuning(fvg.TUNE_SWEEP_MODE, 0)
        ctx.set_tuning(fvg.TUNE_SWEEP_WAVES, 0)
        ctx.set_tuning(fvg.TUNE_SWEEP_POLL, 0)


@pytest.mark.parametrize("name,kind,prec", [("tet10", "sym", "ic"), ("tet10", "nonsym", "ilu"),
                                           ("prism24", "nonsym", "ilu"), ("prism24", "sym", "ic")])
def test_blocked_sweep_gaps_give_identical_bits(ctx, name, kind, prec):
    """Layouts that put 2, 3 (the default), 4 or 6 levels between a value of one workgroup
    and its use in another (components packed, or paired and grown into compact blocks, or
    consecutive rows) all give the oracle's apply bits, and FP64 solves with them the same
    iterations and solution bits: the layout changes the schedule, never the arithmetic.
    A delayed layout has at least the dependency levels."""
    rp, ci, va, b = _oracle_system(name, kind)
    n = rp.shape[0] - 1
    A = fvg.DistributedMatrix.from_csr(ctx, rp, ci, va)
    B = fvg.DistributedVector.from_data(ctx, b)
    r = np.random.default_rng(5).normal(size=n)
    ref = O.OraclePrec(prec, rp, ci, va, 0).apply(r)
    fn = fvg.conjugate_gradient if prec == "ic" else fvg.bi_conjugate_gradient_stab
    opts = fvg.LinearSolverOptions(1e-8, 0.0, 2000)
    for bad in (1, 9, -1):
        with pytest.raises(fvg.FvgError):
            ctx.set_tuning(fvg.TUNE_SWEEP_GAP, bad)
    solves = []
    try:
        for gap in (2, 3, 4, 6):
            ctx.set_tuning(fvg.TUNE_SWEEP_GAP, gap)
            P = (fvg.IncompleteCholesky if prec == "ic" else fvg.IncompleteLowerUpper).from_matrix(A, 0)
            lay = P.sweep_layout()
            assert lay["lower_blocks"] > 0 and lay["block_levels"] >= lay["levels"], (gap, lay)
            z = fvg.DistributedVector(ctx, n)
            P.precondition(z, fvg.DistributedVector.from_data(ctx, r))
            assert np.array_equal(z.numpy(), ref), gap
            x = fvg.DistributedVector(ctx, n)
            info = fn(x, A, B, P, opts)
            solves.append((gap, info.iterations, x.numpy()))
            P.close()
    finally:
        ctx.set_tuning(fvg.TUNE_SWEEP_GAP, 0)
    for gap, it, xg in solves[1:]:
        assert it == solves[0][1] and np.array_equal(xg, solves[0][2]), gap


@pytest.mark.parametrize("name,kind,prec", [("prism24", "nonsym", "ilu"), ("prism24", "sym", "ic"),
                                           ("tri64", "sym", "ic")])
def test_coordinate_boxes_give_identical_bits(ctx, name, kind, prec):
    """Rows that carry their cell centres may take a blocked layout of boxes of the
    coordinates (one dependency component: extruded and 2D meshes).  Without boxes, with
    the default choice and with grids of 16 and 64 boxes the preconditioner applies give the
    oracle's bits and the FP64 solves the same iterations and solution bits.  On the prism
    mesh a forced grid of 64 boxes is built as 64 blocks, and boxes off, coordinates
    dropped or a coordinate that is not finite give the layout of a matrix without
    coordinates."""
    from test_gpu_parity import _mk
    rp, ci, va, b = _oracle_system(name, kind)
    n = rp.shape[0] - 1
    centre = _mk(name)[2]["cell_center"]
    A = fvg.DistributedMatrix.from_csr(ctx, rp, ci, va)
    B = fvg.DistributedVector.from_data(ctx, b)
    r = np.random.default_rng(7).normal(size=n)
    ref = O.OraclePrec(prec, rp, ci, va, 0).apply(r)
    cls = fvg.IncompleteCholesky if prec == "ic" else fvg.IncompleteLowerUpper
    fn = fvg.conjugate_gradient if prec == "ic" else fvg.bi_conjugate_gradient_stab
    opts = fvg.LinearSolverOptions(1e-8, 0.0, 2000)
    for bad in (-1, 1025):
        with pytest.raises(fvg.FvgError):
            ctx.set_tuning(fvg.TUNE_SWEEP_BOXES, bad)
    with pytest.raises(ValueError):
        A.set_row_coordinates(centre[:-1])
    plain = cls.from_matrix(A, 0).sweep_layout()
    A.set_row_coordinates(centre)
    solves, layouts = [], {}
    try:
        for boxes in (1, 0, 16, 64):
            ctx.set_tuning(fvg.TUNE_SWEEP_BOXES, boxes)
            P = cls.from_matrix(A, 0)
            layouts[boxes] = P.sweep_layout()
            z = fvg.DistributedVector(ctx, n)
            P.precondition(z, fvg.DistributedVector.from_data(ctx, r))
            assert np.array_equal(z.numpy(), ref), boxes
            x = fvg.DistributedVector(ctx, n)
            info = fn(x, A, B, P, opts)
            solves.append((boxes, info.iterations, x.numpy()))
            P.close()
    finally:
        ctx.set_tuning(fvg.TUNE_SWEEP_BOXES, 0)
    for boxes, it, xg in solves[1:]:
        assert it == solves[0][1] and np.array_equal(xg, solves[0][2]), boxes
    assert layouts[1] == plain
    if name == "prism24":
        assert layouts[64]["lower_blocks"] == 64 and layouts[64]["upper_blocks"] == 64, layouts[64]
        assert layouts[64]["block_levels"] < plain["block_levels"], (layouts[64], plain)
    A.set_row_coordinates(None)
    assert cls.from_matrix(A, 0).sweep_layout() == plain
    # a coordinate that is not finite: no boxes
    bad = np.array(centre, dtype=np.float64)
    bad[n // 2, 0] = np.nan
    A.set_row_coordinates(bad)
    assert cls.from_matrix(A, 0).sweep_layout() == plain


def test_mesh_matrices_carry_cell_centres(ctx):
    """A matrix made from a device mesh uses the mesh's cell centres: on the 13,824-prism
    mesh its ILU(0) takes a layout of 64 coordinate boxes when asked for that grid (fewer
    levels than the same matrix made from CSR arrays) and applies with the oracle's bits.
    Once the mesh is destroyed the matrix has no coordinates: the layout of a matrix made
    from CSR arrays."""
    from test_gpu_parity import _mk
    import problems
    raw, om, geom = _mk("prism24")
    prob = problems.convection_diffusion(raw, geom, seed=3)
    mesh = fvg.Mesh(ctx, raw)
    lhs, rhs = fvg.assemble_program(mesh, prob["prog"], prob["face_fields"], prob["cell_fields"], prob["bc_i"],
                                    prob["bc_e"])
    rp, ci, va = lhs.csr()
    n = rp.shape[0] - 1
    try:
        ctx.set_tuning(fvg.TUNE_SWEEP_BOXES, 64)
        boxed = fvg.IncompleteLowerUpper.from_matrix(lhs, 0)
        plain = fvg.IncompleteLowerUpper.from_matrix(fvg.DistributedMatrix.from_csr(ctx, rp, ci, va), 0)
    finally:
        ctx.set_tuning(fvg.TUNE_SWEEP_BOXES, 0)
    lb, lp = boxed.sweep_layout(), plain.sweep_layout()
    assert lb["lower_blocks"] == 64 and lb["block_levels"] < lp["block_levels"], (lb, lp)
    r = np.random.default_rng(8).normal(size=n)
    z = fvg.DistributedVector(ctx, n)
    boxed.precondition(z, fvg.DistributedVector.from_data(ctx, r))
    assert np.array_equal(z.numpy(), O.OraclePrec("ilu", rp, ci, va, 0).apply(r))
    mesh.close()
    try:
        ctx.set_tuning(fvg.TUNE_SWEEP_BOXES, 64)
        after = fvg.IncompleteLowerUpper.from_matrix(lhs, 0)
    finally:
        ctx.set_tuning(fvg.TUNE_SWEEP_BOXES, 0)
    assert after.sweep_layout() == lp
    after.precondition(z, fvg.DistributedVector.from_data(ctx, r))
    assert np.array_equal(z.numpy(), O.OraclePrec("ilu", rp, ci, va, 0).apply(r))


@pytest.mark.parametrize("name,kind,prec,blocked", [("tet10", "sym", "ic", True), ("tet10", "nonsym", "ilu", True),
                                                    ("prism24", "nonsym", "ilu", True), ("prism24", "sym", "ic", True),
                                                    ("tri64", "sym", "ic", False)])
def test_blocked_sweep_is_built_and_exact_after_refactor(ctx, name, kind, prec, blocked):
    """The default sweep mode runs the blocked sweep on both factors of the tet mesh
    (blocks of dependency-one-level-apart components) and of the prism mesh (one
    component: blocks of consecutive rows with delayed levels).  On the triangle mesh
    consecutive blocks would run one after another, so no layout is built and the
    persistent sweep serves.  Either way the results are the oracle's bits, also after a
    refactorisation with new values (the blocked layout's coefficients are gathered
    again by the factor kernel)."""
    rp, ci, va, b = _oracle_system(name, kind)
    n = rp.shape[0] - 1
    A = fvg.DistributedMatrix.from_csr(ctx, rp, ci, va)
    P = (fvg.IncompleteCholesky if prec == "ic" else fvg.IncompleteLowerUpper).from_matrix(A, 0)
    lay = P.sweep_layout()
    assert (lay["lower_blocks"] > 0 and lay["upper_blocks"] > 0) == blocked
    assert blocked or (lay["lower_blocks"] == 0 and lay["upper_blocks"] == 0)
    r = np.random.default_rng(4).normal(size=n)
    z = fvg.DistributedVector(ctx, n)
    P.precondition(z, fvg.DistributedVector.from_data(ctx, r))
    assert np.array_equal(z.numpy(), O.OraclePrec(prec, rp, ci, va, 0).apply(r))
    # new values on the same pattern: a diagonal shift and a rescaled off-diagonal part
    diag = np.repeat(np.arange(n), np.diff(rp)) == ci
    va2 = np.where(diag, va * 1.25 + 0.5, va * 0.75)
    A2 = fvg.DistributedMatrix.from_csr(ctx, rp, ci, va2)
    P.refactor(A2)
    P.precondition(z, fvg.DistributedVector.from_data(ctx, r))
    assert np.array_equal(z.numpy(), O.OraclePrec(prec, rp, ci, va2, 0).apply(r))


@pytest.mark.parametrize("prec,kind", [("ic", "sym"), ("ilu", "nonsym")])
def test_blocked_sweep_with_eight_dependency_slots(ctx, prec, kind):
    """Rows with 5 to 8 lower (and upper) dependencies take the blocked kernel's 8-slot
    form: the tet mesh's matrix plus long-range couplings of every 7th row to the rows
    1500, 2500, 3500 and 4500 before and after it (many levels apart, so the layout's
    components stay small), the diagonal raised to keep the factorisation stable.
    Oracle bits."""
    rp, ci, va, b = _oracle_system("tet10", kind)
    n = rp.shape[0] - 1
    A = sp.csr_matrix((va, ci, rp), shape=(n, n)).tolil()
    rng = np.random.default_rng(6)
    for i in range(0, n, 7):
        for off in (-4500, -3500, -2500, -1500, 1500, 2500, 3500, 4500):
            j = i + off
            if 0 <= j < n:
                A[i, j] = -0.05
                A[j, i] = -0.05 if kind == "sym" else -0.05 * rng.uniform(0.5, 1.5)
    A = A.tocsr()
    A.setdiag(A.diagonal() + 0.5)
    A.sort_indices()
    rp2, ci2, va2 = A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.astype(np.float64)
    M = fvg.DistributedMatrix.from_csr(ctx, rp2, ci2, va2)
    P = (fvg.IncompleteCholesky if prec == "ic" else fvg.IncompleteLowerUpper).from_matrix(M, 0)
    lay = P.sweep_layout()
    assert lay["lower_blocks"] > 0 and lay["upper_blocks"] > 0
    assert lay["lower_slots"] == 8 and lay["upper_slots"] == 8
    r = np.random.default_rng(8).normal(size=n)
    z = fvg.DistributedVector(ctx, n)
    P.precondition(z, fvg.DistributedVector.from_data(ctx, r))
    assert np.array_equal(z.numpy(), O.OraclePrec(prec, rp2, ci2, va2, 0).apply(r))


@pytest.mark.parametrize("sweep_mode", [0, 2])
def test_sweep_launch_modes_give_the_same_bits(ctx, sweep_mode):
    """The triangular sweeps whose workgroups wait on one another run from an ordinary
    launch (1), a cooperative one (2) or the automatic choice (0: ordinary while one
    context is alive on the device, which a second context here turns to cooperative).
    Each gives the oracle's bits, for the blocked (sweep mode 0) and the persistent
    (sweep mode 2) sweep, and the mode each call used is read back; a launch mode
    outside 0-2 is rejected."""
    rp, ci, va, b = _oracle_system("tet10", "sym")
    n = rp.shape[0] - 1
    want = O.OraclePrec("ic", rp, ci, va, 0).apply(b)
    with pytest.raises(fvg.FvgError):
        ctx.set_tuning(fvg.TUNE_SWEEP_LAUNCH, 3)
    ctx.set_tuning(fvg.TUNE_SWEEP_MODE, sweep_mode)
    other = None
    try:
        A = fvg.DistributedMatrix.from_csr(ctx, rp, ci, va)
        P = fvg.IncompleteCholesky.from_matrix(A, 0)
        assert P.sweep_layout()["lower_blocks"] > 0
        rhs = fvg.DistributedVector.from_data(ctx, b)
        z = fvg.DistributedVector(ctx, n)
        for launch, used in ((1, 1), (2, 2), (0, 1)):
            ctx.set_tuning(fvg.TUNE_SWEEP_LAUNCH, launch)
            assert ctx.get_tuning(fvg.TUNE_SWEEP_LAUNCH) == launch
            z.fill(0.0)
            P.precondition(z, rhs)
            assert np.array_equal(z.numpy(), want), launch
            assert ctx.sweep_launch_state() == (used, 0), launch
        other = fvg.Context(0)
        z.fill(0.0)
        P.precondition(z, rhs)
        assert np.array_equal(z.numpy(), want)
        assert ctx.sweep_launch_state() == (2, 0)
    finally:
        if other is not None:
            other.close()
        ctx.set_tuning(fvg.TUNE_SWEEP_LAUNCH, 0)
        ctx.set_tuning(fvg.TUNE_SWEEP_MODE, 0)


@pytest.mark.parametrize("mode", [0, 2])
def test_sweep_wait_time_out_is_reported_and_recovers(ctx, mode):
    """Every wait of a triangular sweep is bounded.  With the sweeps withholding their
    values from the workgroups and waves that wait on them (a test-only knob) and a small
    spin limit, every such wait runs to its limit, sets the error flag that ends all
    other waits, and the kernel exits normally: each call returns FVG_E_TIMEOUT.  With the
    knobs cleared, the next call gives the oracle's bits again (blocked sweep: mode 0;
    persistent sweep: mode 2)."""
    rp, ci, va, b = _oracle_system("tet10", "sym")
    n = rp.shape[0] - 1
    A = fvg.DistributedMatrix.from_csr(ctx, rp, ci, va)
    ic = fvg.IncompleteCholesky.from_matrix(A, 0)
    lay = ic.sweep_layout()
    if mode == 0:
        # values cross blocks, so the blocked sweep has waits that can be starved
        assert lay["lower_blocks"] > 1 and lay["lower_fetch_total"] > 0
    assert lay["levels"] > 1
    r = fvg.DistributedVector.from_data(ctx, np.random.default_rng(2).normal(size=n))
    ref = O.OraclePrec("ic", rp, ci, va, 0).apply(r.numpy())
    z = fvg.DistributedVector(ctx, n)
    statuses = []
    try:
        ctx.set_tuning(fvg.TUNE_SWEEP_MODE, mode)
        ctx.set_tuning(fvg.TUNE_SPIN_LIMIT, 16)
        ctx.set_tuning(fvg.TUNE_SWEEP_WITHHOLD, 1)
        for _ in range(3):
            try:
                ic.precondition(z, r)
                statuses.append("FVG_OK")
            except fvg.FvgError as e:
                statuses.append(fvg.STATUS[e.status])
    finally:
        ctx.set_tuning(fvg.TUNE_SWEEP_WITHHOLD, 0)
        ctx.set_tuning(fvg.TUNE_SPIN_LIMIT, 0)
        ctx.set_tuning(fvg.TUNE_SWEEP_MODE, 0)
    assert statuses == ["FVG_E_TIMEOUT"] * 3, statuses
    ic.precondition(z, r)
    assert np.array_equal(z.numpy(), ref)
    # a forced time-out does not make the context's later launches cooperative
    assert ctx.sweep_launch_state()[0] in (1, 2)


def test_sweep_while_a_batch_is_in_flight(ctx):
    """A batch of this context runs on a stream of its own (persistent kernels holding
    compute units).  While it is in flight, the co-resident IC(0) sweep launches
    cooperatively, and gives the oracle's bits; once the batch is finished, the
    automatic choice is the ordinary launch again."""
    import meshgen
    rp, ci, va, b = _oracle_system("tet10", "sym")
    n = rp.shape[0] - 1
    want = O.OraclePrec("ic", rp, ci, va, 0).apply(b)
    A = fvg.DistributedMatrix.from_csr(ctx, rp, ci, va)
    P = fvg.IncompleteCholesky.from_matrix(A, 0)
    rhs = fvg.DistributedVector.from_data(ctx, b)
    z = fvg.DistributedVector(ctx, n)
    # a batch of a few large systems, so it is still running when the sweep is enqueued
    mats, precs, bs, xs = [], [], [], []
    for j in range(4):
        raw = meshgen.tri2d(120, seed=50 + j)
        om = O.OracleMesh(raw)
        brp, bci, bva, bb = om.assemble(*_diffusion_program(om))
        Aj = fvg.DistributedMatrix.from_csr(ctx, brp, bci, bva)
        mats.append(Aj)
        precs.append(fvg.IncompleteCholesky.from_matrix(Aj, 0))
        bs.append(fvg.DistributedVector.from_data(ctx, bb))
        xs.append(fvg.DistributedVector(ctx, brp.shape[0] - 1))
    batch = fvg.Batch(ctx, mats, precs, "cg", fvg.LinearSolverOptions(1e-8, 0.0, 3000))
    try:
        batch.launch(bs, xs)
        assert ctx.sweep_launch_state()[1] == 1
        P.precondition(z, rhs)
        assert ctx.sweep_launch_state() == (2, 1)
        assert np.array_equal(z.numpy(), want)
        infos = batch.finish()
        assert all(i.final_residual <= 1e-8 * i.initial_residual for i in infos)
        assert ctx.sweep_launch_state()[1] == 0
        z.fill(0.0)
        P.precondition(z, rhs)
        assert ctx.sweep_launch_state() == (1, 0)
        assert np.array_equal(z.numpy(), want)
    finally:
        batch.close()


def _diffusion_program(om):
    """-laplacian(Orthogonal, k = 1) - source(1) with u = 0 on the boundary."""
    eq = -fvg.laplacian(fvg.Orthogonal(), np.ones(om.n_faces)) - fvg.source(np.ones(om.n_cells))
    ops, ff, cf = eq.compile()
    return ops, ff, cf, np.zeros(om.n_bnd), np.zeros(om.n_bnd)


def test_blocked_sweeps_on_an_unsymmetric_pattern(ctx):
    """ILU(0) on a pattern that is not symmetric (the tet mesh's matrix plus couplings of
    every 5th row to the rows 1500 and 3000 before it, below the diagonal only): the lower
    layout has 8 slots per row and the upper 4, their blocks differ, and the upper
    sweep's set-up reads the lower sweep's values through the pairing index.  Oracle
    bits."""
    rp, ci, va, b = _oracle_system("tet10", "nonsym")
    n = rp.shape[0] - 1
    A = sp.csr_matrix((va, ci, rp), shape=(n, n)).tolil()
    for i in range(0, n, 5):
        for off in (1500, 3000):
            if i - off >= 0:
                A[i, i - off] = -0.04
    A = A.tocsr()
    A.setdiag(A.diagonal() + 0.3)
    A.sort_indices()
    rp2, ci2, va2 = A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.astype(np.float64)
    M = fvg.DistributedMatrix.from_csr(ctx, rp2, ci2, va2)
    P = fvg.IncompleteLowerUpper.from_matrix(M, 0)
    lay = P.sweep_layout()
    assert lay["lower_blocks"] > 0 and lay["upper_blocks"] > 0
    assert lay["lower_slots"] == 8 and lay["upper_slots"] == 4
    r = np.random.default_rng(9).normal(size=n)
    z = fvg.DistributedVector(ctx, n)
    P.precondition(z, fvg.DistributedVector.from_data(ctx, r))
    assert np.array_equal(z.numpy(), O.OraclePrec("ilu", rp2, ci2, va2, 0).apply(r))
