"""IC(level) and ILU(level) with fill (level > 0) on the device against the oracle.

The factor pattern is the reference's (its pattern walks, incomplete_cholesky.rs:35-85 and
incomplete_lower_upper.rs:30-137), the factorisation runs the level-0 kernels on it, and
IC's upper sweep takes its terms in the order of the reference's transposed pattern.  Bars:
factor pattern equal, factor values and inverse pivots bit-exact, one preconditioner apply
bit-exact, and the solvers within the usual bars (iterations +-2, x within 1e-6)."""
import numpy as np
import pytest

import fvg
import oracle as O
from test_gpu_parity import _oracle_system

pytestmark = pytest.mark.gpu


def _check_factor(ctx, kind, name, level):
    sym = "sym" if kind == "ic" else "nonsym"
    rp, ci, va, b = _oracle_system(name, sym)
    n = rp.shape[0] - 1
    A = fvg.DistributedMatrix.from_csr(ctx, rp, ci, va)
    Prec = fvg.IncompleteCholesky if kind == "ic" else fvg.IncompleteLowerUpper
    M = Prec.from_matrix(A, level)
    P = O.OraclePrec(kind, rp, ci, va, level)
    orp, oci, ova, oinv = P.factor()
    frp, fci = M.factor_pattern()
    gva, ginv = M.factor_values()
    for i in range(n):
        cols = fci[frp[i]:frp[i + 1]]
        vals = gva[frp[i]:frp[i + 1]]
        ocols = oci[orp[i]:orp[i + 1]]
        ovals = ova[orp[i]:orp[i + 1]]
        if kind == "ic":   # the oracle's L is the lower part of the device's symmetric pattern
            keep = cols <= i
            assert np.array_equal(cols[keep], ocols), i
            assert np.array_equal(vals[keep], ovals), i
        else:              # ILU: dropped entries are explicit zeros on the device
            keep = np.isin(cols, ocols)
            assert np.array_equal(cols[keep], ocols), i
            assert np.array_equal(vals[keep], ovals), i
            assert np.all(vals[~keep] == 0.0)
    if kind == "ic":
        assert np.array_equal(ginv, oinv)
    r = np.random.default_rng(level).normal(size=n)
    z = fvg.DistributedVector(ctx, n)
    M.precondition(z, fvg.DistributedVector.from_data(ctx, r))
    assert np.array_equal(z.numpy(), P.apply(r))
    return rp, ci, va, b, A, M, P, frp


@pytest.mark.parametrize("name", ["tri64", "tet10"])
@pytest.mark.parametrize("level", [1, 2])
def test_ic_fill_factor_and_apply_bit_exact(ctx, name, level):
    rp, ci, va, b, A, M, P, frp = _check_factor(ctx, "ic", name, level)
    if level == 1:   # the reference's walk adds no fill at level 1
        assert frp[-1] == rp[-1]
    else:
        assert frp[-1] > rp[-1]


@pytest.mark.parametrize("name", ["tri64", "tet10"])
def test_ic_level_three_is_refused_like_the_reference(ctx, name):
    """At level 3 the reference's walk (each vertex met once per row, explored only within
    the level) leaves factor entries whose transpose is missing from its transposed pattern,
    and its l_vt lookup panics (incomplete_cholesky.rs:138-143); the oracle raises and the
    device refuses with FVG_E_MISSING_ENTRY."""
    rp, ci, va, b = _oracle_system(name, "sym")
    with pytest.raises(RuntimeError):
        O.OraclePrec("ic", rp, ci, va, 3)
    A = fvg.DistributedMatrix.from_csr(ctx, rp, ci, va)
    with pytest.raises(fvg.FvgError) as e:
        fvg.IncompleteCholesky.from_matrix(A, 3)
    assert fvg.STATUS[e.value.status] == "FVG_E_MISSING_ENTRY"


@pytest.mark.parametrize("name", ["tri64", "prism8"])
@pytest.mark.parametrize("level", [1, 2])
def test_ilu_fill_factor_and_apply_bit_exact(ctx, name, level):
    rp, ci, va, b, A, M, P, frp = _check_factor(ctx, "ilu", name, level)
    assert frp[-1] > rp[-1]


@pytest.mark.parametrize("name,kind,solver,prec,level", [("tet10", "sym", "cg", "ic", 2),
                                                         ("tri64", "sym", "cg", "ic", 1),
                                                         ("prism24", "nonsym", "bicgstab", "ilu", 1),
                                                         ("tri64", "nonsym", "bicgstab", "ilu", 2)])
def test_solvers_with_fill_match_the_oracle(ctx, name, kind, solver, prec, level):
    """Solves with fill: iterations +-2 and x within 1e-6 of the oracle, then a refactorisation
    of new values gives the oracle's apply bits.  (IC(2)'s transposed pattern can hold entries
    beyond its factor's transpose, and ILU(level) patterns are not symmetric, so these solves
    may iterate in natural order: the sweep layout then is not mirrored.)"""
    rp, ci, va, b = _oracle_system(name, kind)
    n = rp.shape[0] - 1
    P = O.OraclePrec(prec, rp, ci, va, level)
    xo, io = O.solve(solver, rp, ci, va, P, b, rtol=1e-8, atol=0.0, max_iter=2000)
    A = fvg.DistributedMatrix.from_csr(ctx, rp, ci, va)
    M = (fvg.IncompleteCholesky if prec == "ic" else fvg.IncompleteLowerUpper).from_matrix(A, level)
    x = fvg.DistributedVector(ctx, n)
    fn = fvg.conjugate_gradient if solver == "cg" else fvg.bi_conjugate_gradient_stab
    info = fn(x, A, fvg.DistributedVector.from_data(ctx, b), M, fvg.LinearSolverOptions(1e-8, 0.0, 2000))
    assert abs(info.iterations - io["iterations"]) <= 2, (info.iterations, io["iterations"])
    assert np.linalg.norm(x.numpy() - xo) <= 1e-6 * np.linalg.norm(xo)
    # a refactorisation after new values gives the oracle's factor of those values
    va2 = va * 1.5
    A.set_values(va2)
    M.refactor(A)
    r = np.random.default_rng(5).normal(size=n)
    z = fvg.DistributedVector(ctx, n)
    M.precondition(z, fvg.DistributedVector.from_data(ctx, r))
    assert np.array_equal(z.numpy(), O.OraclePrec(prec, rp, ci, va2, level).apply(r))


def test_batches_refuse_fill_levels(ctx):
    rp, ci, va, b = _oracle_system("tri16", "sym")
    A = fvg.DistributedMatrix.from_csr(ctx, rp, ci, va)
    M = fvg.IncompleteCholesky.from_matrix(A, 2)
    with pytest.raises(fvg.FvgError) as e:
        fvg.Batch(ctx, [A], [M], "cg", fvg.LinearSolverOptions(1e-8, 0.0, 100))
    assert fvg.STATUS[e.value.status] == "FVG_E_UNSUPPORTED"


@pytest.mark.parametrize("kind", ["ilu", "ic"])
def test_fill_layouts_with_coordinate_boxes(ctx, kind):
    """IC(1) and ILU(1) of the prism mesh whose rows carry their cell centres: with the
    coordinate boxes off, at the default and forced to grids of 16 and 64, every apply
    gives the oracle's bits.  ILU(1)'s fill pattern is not symmetric, so its upper layout
    is assigned on its own, with the slab indices closed along the upper dependencies."""
    from test_gpu_parity import _mk
    centre = _mk("prism24")[2]["cell_center"]
    rp, ci, va, b = _oracle_system("prism24", "nonsym" if kind == "ilu" else "sym")
    n = rp.shape[0] - 1
    A = fvg.DistributedMatrix.from_csr(ctx, rp, ci, va)
    A.set_row_coordinates(centre)
    r = np.random.default_rng(11).normal(size=n)
    want = O.OraclePrec(kind, rp, ci, va, 1).apply(r)
    Prec = fvg.IncompleteCholesky if kind == "ic" else fvg.IncompleteLowerUpper
    layouts = {}
    try:
        for boxes in (1, 0, 16, 64):
            ctx.set_tuning(fvg.TUNE_SWEEP_BOXES, boxes)
            M = Prec.from_matrix(A, 1)
            layouts[boxes] = M.sweep_layout()
            z = fvg.DistributedVector(ctx, n)
            M.precondition(z, fvg.DistributedVector.from_data(ctx, r))
            assert np.array_equal(z.numpy(), want), (kind, boxes, layouts[boxes])
            M.close()
    finally:
        ctx.set_tuning(fvg.TUNE_SWEEP_BOXES, 0)
    # a grid of 64 boxes is built for both factors: mirrored for IC(1), on its own for ILU(1)
    assert layouts[64]["lower_blocks"] == 64 and layouts[64]["upper_blocks"] == 64, layouts[64]
    assert layouts[64]["mirrored"] == (1 if kind == "ic" else 0), layouts[64]
    assert layouts[64]["block_levels"] < layouts[1]["block_levels"], (layouts[64], layouts[1])


def _banded(n, half, arrow, seed):
    """Nonsymmetric, diagonally dominant CSR: a band of half-width `half`, plus a dense
    first row and column when `arrow` (a row of n entries)."""
    rng = np.random.default_rng(seed)
    rp, ci, va = [0], [], []
    for i in range(n):
        cols = sorted(set(range(max(0, i - half), min(n, i + half + 1))) | ({0} if arrow else set())
                      | (set(range(n)) if arrow and i == 0 else set()))
        vals = rng.uniform(-1.0, 1.0, size=len(cols))
        vals[cols.index(i)] = 2.0 * len(cols) + (n if arrow and i == 0 else 0)
        ci += cols
        va += list(vals)
        rp.append(len(ci))
    return np.array(rp, np.int32), np.array(ci, np.int32), np.array(va, np.float64)


@pytest.mark.parametrize("n,half,arrow", [(300, 30, False), (200, 2, True)])
def test_ilu0_rows_near_and_past_a_wave(ctx, n, half, arrow):
    """The ILU refactorisation holds a row in the lanes of one wave when every row has at
    most 64 entries (61 here) and falls back to a thread per row otherwise (an arrow matrix
    whose first row holds all 200 columns): both give the oracle's factor and apply bits."""
    rp, ci, va = _banded(n, half, arrow, seed=n)
    A = fvg.DistributedMatrix.from_csr(ctx, rp, ci, va)
    M = fvg.IncompleteLowerUpper.from_matrix(A, 0)
    P = O.OraclePrec("ilu", rp, ci, va, 0)
    orp, oci, ova, oinv = P.factor()
    frp, fci = M.factor_pattern()
    gva, ginv = M.factor_values()
    assert np.array_equal(frp, rp) and np.array_equal(fci, ci)
    for i in range(n):   # the oracle leaves out what the drop rule zeroes
        cols, vals = fci[frp[i]:frp[i + 1]], gva[frp[i]:frp[i + 1]]
        keep = np.isin(cols, oci[orp[i]:orp[i + 1]])
        assert np.array_equal(cols[keep], oci[orp[i]:orp[i + 1]]), i
        assert np.array_equal(vals[keep], ova[orp[i]:orp[i + 1]]), i
        assert np.all(vals[~keep] == 0.0), i
    r = np.random.default_rng(1).normal(size=n)
    z = fvg.DistributedVector(ctx, n)
    M.precondition(z, fvg.DistributedVector.from_data(ctx, r))
    assert np.array_equal(z.numpy(), P.apply(r))
    M.refactor(A)   # a second factorisation over the first gives the same bits
    M.precondition(z, fvg.DistributedVector.from_data(ctx, r))
    assert np.array_equal(z.numpy(), P.apply(r))
