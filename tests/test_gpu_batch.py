"""Batched solves: many independent systems in one launch.

FP64 batches solve each system with one workgroup; mixed batches spread a system over
ceil(n / rows_per_group) workgroups (rows_per_group 0: one workgroup per system).

Bars:
- FP64 batches run the reference algorithm per system. Each system's solution is within
  1e-6 relative of the oracle's at a 1e-8 relative residual, and its iteration count is
  within +-2 of the oracle's.
- Mixed batches (FP32 inner, FP64 refinement) reach a true relative residual <= 1e-8, and x
  is within 1e-6 relative of the FP64 oracle solution, for one and for many workgroups
  per system.
- Every system gives the same bits alone and in any mix.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import fvg
import meshgen
import oracle as O
import problems

pytestmark = pytest.mark.gpu

OPTS = dict(relative_tolerance=1e-8, absolute_tolerance=0.0, max_iterations=3000)


def _systems(kind, meshes, seed0):
    out = []
    for j, raw in enumerate(meshes):
        om = O.OracleMesh(raw)
        geom = om.geometry()
        n = om.n_cells
        if kind == "sym":
            f = np.random.default_rng(seed0 + j).normal(size=n)
            prob = problems.diffusion(raw, geom, kc=meshgen.lognormal_cells(n, 1.0, seed0 + j), f=f, bc="zero")
        else:
            prob = problems.convection_diffusion(raw, geom, pe=(1.0, 50.0), seed=seed0 + j)
        out.append(om.assemble(prob["prog"], prob["face_fields"], prob["cell_fields"], prob["bc_i"], prob["bc_e"]))
    return out


def _sym_meshes():
    return [meshgen.tri2d(12, seed=1), meshgen.tri2d(40, seed=2), meshgen.tet3d(5, seed=3), meshgen.tri2d(7, seed=4),
            meshgen.tri2d(25, seed=5), meshgen.tri2d(3, seed=6)]


def _nonsym_meshes():
    return [meshgen.tri2d(20, seed=7), meshgen.prism3d(6, layers=3, seed=8), meshgen.tri2d(9, seed=9)]


class _Batch:
    """Device objects for a list of systems plus a batch over them."""

    def __init__(self, ctx, systems, prec, solver, precision=fvg.PRECISION_FP64, order=None, rows=None):
        self.systems = systems
        order = list(range(len(systems))) if order is None else order
        self.order = order
        self.A, self.M, self.b, self.x = [], [], [], []
        for k in order:
            rp, ci, va, b = systems[k]
            A = fvg.DistributedMatrix.from_csr(ctx, rp, ci, va)
            kinds = prec if isinstance(prec, list) else [prec] * len(systems)
            pk = kinds[k]
            if pk == "ic":
                M = fvg.IncompleteCholesky.from_matrix(A, 0)
            elif pk == "ilu":
                M = fvg.IncompleteLowerUpper.from_matrix(A, 0)
            elif pk == "diagonal":
                M = fvg.DiagonalPreconditioner.from_matrix(A)
            else:
                M = fvg.IdentityPreconditioner.new(ctx, rp.shape[0] - 1)
            self.A.append(A)
            self.M.append(M)
            self.b.append(fvg.DistributedVector.from_data(ctx, b))
            self.x.append(fvg.DistributedVector(ctx, rp.shape[0] - 1))
        self.opts = fvg.LinearSolverOptions(**OPTS, precision=precision)
        self.batch = fvg.Batch(ctx, self.A, self.M, solver, self.opts)
        if rows is not None:
            self.batch.set_cluster(rows)

    def solve(self):
        for x in self.x:
            x.fill(0.0)
        infos = self.batch.solve(self.b, self.x)
        return {k: (self.x[j].numpy(), infos[j]) for j, k in enumerate(self.order)}


def _check_against_oracle(systems, results, prec, solver):
    kinds = prec if isinstance(prec, list) else [prec] * len(systems)
    for k, (rp, ci, va, b) in enumerate(systems):
        P = O.OraclePrec(kinds[k], rp, ci, va, 0)
        xo, io = O.solve(solver, rp, ci, va, P, b, rtol=1e-8, atol=0.0, max_iter=3000)
        xg, ig = results[k]
        assert abs(ig.iterations - io["iterations"]) <= 2, (k, ig.iterations, io["iterations"])
        assert np.linalg.norm(xg - xo) <= 1e-6 * np.linalg.norm(xo), k
        assert ig.final_residual < 1e-8 * ig.initial_residual
        assert abs(ig.initial_residual - io["initial_residual"]) <= 1e-12 * io["initial_residual"]
        assert ig.history.shape[0] == ig.iterations
        assert ig.history[-1] == ig.final_residual


def test_batch_cg_ic0_matches_oracle_per_system(ctx):
    systems = _systems("sym", _sym_meshes(), 10)
    B = _Batch(ctx, systems, "ic", "cg")
    _check_against_oracle(systems, B.solve(), "ic", "cg")


def test_batch_bicgstab_ilu0_matches_oracle_per_system(ctx):
    systems = _systems("nonsym", _nonsym_meshes(), 20)
    B = _Batch(ctx, systems, "ilu", "bicgstab")
    _check_against_oracle(systems, B.solve(), "ilu", "bicgstab")


def test_batch_mixes_preconditioner_kinds(ctx):
    systems = _systems("sym", [meshgen.tri2d(10, seed=31), meshgen.tri2d(14, seed=32), meshgen.tri2d(6, seed=33)], 30)
    kinds = ["ic", "diagonal", "identity"]
    B = _Batch(ctx, systems, kinds, "cg")
    _check_against_oracle(systems, B.solve(), kinds, "cg")


@pytest.mark.parametrize("solver,kind,prec,precision,rows", [("cg", "sym", "ic", fvg.PRECISION_FP64, None),
                                                             ("cg", "sym", "ic", fvg.PRECISION_MIXED, 0),
                                                             ("cg", "sym", "ic", fvg.PRECISION_MIXED, 256),
                                                             ("bicgstab", "nonsym", "ilu", fvg.PRECISION_FP64, None),
                                                             ("bicgstab", "nonsym", "ilu", fvg.PRECISION_MIXED, 0),
                                                             ("bicgstab", "nonsym", "ilu", fvg.PRECISION_MIXED, 256)])
def test_batch_result_is_the_same_alone_and_in_the_mix(ctx, solver, kind, prec, precision, rows):
    meshes = _sym_meshes() if kind == "sym" else _nonsym_meshes()
    systems = _systems(kind, meshes, 40)
    full = _Batch(ctx, systems, prec, solver, precision, rows=rows).solve()
    rev = _Batch(ctx, systems, prec, solver, precision, order=list(reversed(range(len(systems)))), rows=rows).solve()
    for k in range(len(systems)):
        alone = _Batch(ctx, [systems[k]], prec, solver, precision, rows=rows).solve()[0]
        for other in (full[k], rev[k]):
            assert np.array_equal(alone[0], other[0]), k
            assert alone[1].iterations == other[1].iterations
            assert alone[1].final_residual == other[1].final_residual
            assert np.array_equal(alone[1].history, other[1].history)


def test_batch_size_buckets(ctx):
    """Rows per workgroup per size bucket: each system of a bucketed mixed batch gives the
    bits it gives alone with its bucket's rows per workgroup (a result depends on its system
    and its workgroup count only), every answer meets the accuracy bar, the bucket autotuner
    sets one choice per bucket, and malformed buckets are refused."""
    systems = _systems("sym", _sym_meshes(), 70)
    limits, rows, rest = [500, 1500], [64, 256], 4096

    def rows_of(n):
        return next((r for lim, r in zip(limits, rows) if n <= lim), rest)

    B = _Batch(ctx, systems, "ic", "cg", fvg.PRECISION_MIXED, rows=rest)
    B.batch.set_cluster_buckets(limits, rows)
    res = B.solve()
    for k, (rp, ci, va, b) in enumerate(systems):
        n = rp.shape[0] - 1
        alone = _Batch(ctx, [systems[k]], "ic", "cg", fvg.PRECISION_MIXED, rows=rows_of(n)).solve()[0]
        assert np.array_equal(alone[0], res[k][0]), (k, n)
        S = sp.csr_matrix((va, ci, rp), shape=(n, n))
        assert np.linalg.norm(b - S @ res[k][0]) / np.linalg.norm(b) <= 1e-8, k
    with pytest.raises(fvg.FvgError):
        B.batch.set_cluster_buckets([1500, 500], [64, 256])
    with pytest.raises(fvg.FvgError):
        B.batch.set_cluster_buckets([500], [0])
    times = B.batch.tune_buckets(B.b, B.x, limits=(500, 1500), candidates=(64, 256, 1024))
    assert set(times) == {500, 1500, None} and all(len(t) == 3 for t in times.values())
    assert B.batch.buckets[0] == [500, 1500] and len(B.batch.buckets[1]) == 2
    assert B.batch.rows_per_group in (64, 256, 1024)
    res = B.solve()
    for k, (rp, ci, va, b) in enumerate(systems):
        n = rp.shape[0] - 1
        S = sp.csr_matrix((va, ci, rp), shape=(n, n))
        assert np.linalg.norm(b - S @ res[k][0]) / np.linalg.norm(b) <= 1e-8, k


@pytest.mark.parametrize("rows", [0, 64, 256, 4096])
@pytest.mark.parametrize("solver,kind,prec", [("cg", "sym", "ic"), ("bicgstab", "nonsym", "ilu")])
def test_batch_mixed_precision_reaches_fp64_accuracy(ctx, solver, kind, prec, rows):
    """rows 64 puts the largest systems on 50 workgroups; 4096 gives every system one."""
    meshes = _sym_meshes() if kind == "sym" else _nonsym_meshes()
    systems = _systems(kind, meshes, 50)
    res = _Batch(ctx, systems, prec, solver, fvg.PRECISION_MIXED, rows=rows).solve()
    for k, (rp, ci, va, b) in enumerate(systems):
        n = rp.shape[0] - 1
        xg, ig = res[k]
        S = sp.csr_matrix((va, ci, rp), shape=(n, n))
        true_rel = np.linalg.norm(b - S @ xg) / np.linalg.norm(b)
        assert true_rel <= 1e-8, (k, true_rel)
        assert abs(ig.true_residual - true_rel) <= 1e-3 * true_rel + 1e-15
        xo, _ = O.solve(solver, rp, ci, va, O.OraclePrec(prec, rp, ci, va, 0), b, rtol=1e-8, atol=0.0, max_iter=3000)
        assert np.linalg.norm(xg - xo) <= 1e-6 * np.linalg.norm(xo), k


def test_batch_graph_replays_are_identical(ctx):
    """The batch launch is a captured graph; replays with eager work in between give the same bits."""
    systems = _systems("sym", _sym_meshes()[:4], 60)
    B = _Batch(ctx, systems, "ic", "cg")
    first = B.solve()
    scratch = fvg.DistributedVector(ctx, systems[0][0].shape[0] - 1)
    for _ in range(3):
        B.A[0].imul(scratch, B.b[0])   # eager work on the context stream between replays
        ctx.sync()
        again = B.solve()
        for k in first:
            assert np.array_equal(first[k][0], again[k][0])
            assert first[k][1].iterations == again[k][1].iterations


def test_batch_picks_up_new_values_after_refactor(ctx):
    systems = _systems("sym", _sym_meshes()[:3], 70)
    B = _Batch(ctx, systems, "ic", "cg")
    B.solve()
    scaled = []
    for j, (rp, ci, va, b) in enumerate(systems):
        va2 = va * (1.5 + j)
        B.A[j].set_values(va2)
        B.M[j].refactor(B.A[j])
        scaled.append((rp, ci, va2, b))
    _check_against_oracle(scaled, B.solve(), "ic", "cg")


@pytest.mark.parametrize("threads", [256, 512])
def test_batch_workgroup_sizes(ctx, threads):
    systems = _systems("sym", _sym_meshes(), 80)
    B = _Batch(ctx, systems, "ic", "cg")
    B.batch.threads = threads
    assert B.batch.threads == threads
    _check_against_oracle(systems, B.solve(), "ic", "cg")


def test_batch_zero_rhs_and_errors(ctx):
    systems = _systems("sym", [meshgen.tri2d(8, seed=90), meshgen.tri2d(5, seed=91)], 90)
    rp, ci, va, b = systems[1]
    systems[1] = (rp, ci, va, np.zeros_like(b))
    B = _Batch(ctx, systems, "ic", "cg")
    for x in B.x:
        x.fill(3.0)
    infos = B.batch.solve(B.b, B.x)
    assert np.all(B.x[1].numpy() == 0.0)
    assert infos[1].iterations == 0 and infos[1].final_residual == 0.0
    assert infos[0].iterations > 0
    # a structurally unsymmetric pattern is refused for factorised preconditioners
    rp3 = np.array([0, 2, 3, 4], np.int32)
    ci3 = np.array([0, 1, 1, 2], np.int32)
    va3 = np.array([2.0, -1.0, 2.0, 2.0])
    A3 = fvg.DistributedMatrix.from_csr(ctx, rp3, ci3, va3)
    M3 = fvg.IncompleteLowerUpper.from_matrix(A3, 0)
    with pytest.raises(fvg.FvgError):
        fvg.Batch(ctx, [A3], [M3], "bicgstab", fvg.LinearSolverOptions(**OPTS))


@pytest.mark.parametrize("rows", [0, 32])
def test_batch_mixed_zero_rhs_and_preconditioner_kinds(ctx, rows):
    """Mixed batches: a zero right-hand side gives x = 0 after no iterations, and Jacobi /
    identity preconditioners solve like IC(0), with one or several workgroups per system."""
    systems = _systems("sym", [meshgen.tri2d(10, seed=61), meshgen.tri2d(8, seed=62), meshgen.tri2d(9, seed=63),
                               meshgen.tri2d(7, seed=64)], 60)
    rp, ci, va, b = systems[1]
    systems[1] = (rp, ci, va, np.zeros_like(b))
    kinds = ["ic", "ic", "diagonal", "identity"]
    B = _Batch(ctx, systems, kinds, "cg", fvg.PRECISION_MIXED, rows=rows)
    for x in B.x:
        x.fill(3.0)
    infos = B.batch.solve(B.b, B.x)
    assert np.all(B.x[1].numpy() == 0.0)
    assert infos[1].iterations == 0 and infos[1].status == 0
    for k in (0, 2, 3):
        rp, ci, va, b = systems[k]
        n = rp.shape[0] - 1
        S = sp.csr_matrix((va, ci, rp), shape=(n, n))
        xg = B.x[k].numpy()
        assert infos[k].status == 0 and infos[k].iterations > 0
        # the tolerance is relative to the initial residual (x0 = 3)
        r0 = np.linalg.norm(b - S @ np.full(n, 3.0))
        assert np.linalg.norm(b - S @ xg) <= 1e-8 * r0, k


@pytest.mark.parametrize("prec", ["ic", "ilu"])
def test_refactor_many_gives_the_per_level_factors(ctx, prec):
    """One launch refactorising several systems gives the bits of the per-level launches
    (and so the oracle's), and defers the pivot / symmetry checks to status()."""
    systems = _systems("sym" if prec == "ic" else "nonsym",
                       _sym_meshes()[:3] if prec == "ic" else _nonsym_meshes(), 100)
    cls = fvg.IncompleteCholesky if prec == "ic" else fvg.IncompleteLowerUpper
    A, M, scaled = [], [], []
    for j, (rp, ci, va, b) in enumerate(systems):
        a = fvg.DistributedMatrix.from_csr(ctx, rp, ci, va)
        A.append(a)
        M.append(cls.from_matrix(a, 0))
        va2 = va * (2.0 + j)
        a.set_values(va2)
        scaled.append((rp, ci, va2))
    fvg.refactor_many(ctx, M, A)
    ctx.sync()
    for m, (rp, ci, va2) in zip(M, scaled):
        m.status()
        got = m.factor_values()
        lrp, lci, lva, linv = O.OraclePrec(prec, rp, ci, va2, 0).factor()
        ref = cls.from_matrix(fvg.DistributedMatrix.from_csr(ctx, rp, ci, va2), 0).factor_values()
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
    if prec == "ic":   # an unsymmetric matrix is reported by status(), not by the launch
        rp, ci, va, b = systems[0]
        bad = va.copy()
        off = np.nonzero(ci != np.repeat(np.arange(rp.shape[0] - 1), np.diff(rp)))[0][0]
        bad[off] *= 1.5
        A[0].set_values(bad)
        fvg.refactor_many(ctx, [M[0]], [A[0]])
        with pytest.raises(fvg.NotSymmetricError):
            M[0].status()


@pytest.mark.parametrize("precision,rows", [(fvg.PRECISION_FP64, None), (fvg.PRECISION_MIXED, 256)])
def test_batch_reports_a_failed_deferred_factorisation(ctx, precision, rows):
    """A system whose deferred refactorisation failed its check is reported in its info
    (status) and left unsolved; the other systems of the batch are solved as usual."""
    systems = _systems("sym", _sym_meshes()[:3], 110)
    B = _Batch(ctx, systems, "ic", "cg", precision, rows=rows)
    rp, ci, va, b = systems[1]
    bad = va.copy()
    off = np.nonzero(ci != np.repeat(np.arange(rp.shape[0] - 1), np.diff(rp)))[0][0]
    bad[off] *= 1.5                      # no longer symmetric
    B.A[1].set_values(bad)
    B.M[1].refactor_async(B.A[1])
    for x in B.x:
        x.fill(0.0)
    infos = B.batch.solve(B.b, B.x)
    assert infos[1].status == 5          # FVG_E_NOT_SYMMETRIC
    assert np.all(B.x[1].numpy() == 0.0)
    assert infos[0].status == 0 and infos[2].status == 0
    if precision == fvg.PRECISION_FP64:
        assert infos[0].final_residual < 1e-8 * infos[0].initial_residual
    else:   # mixed reports residuals relative to |b|
        assert infos[0].true_residual <= 1e-8
