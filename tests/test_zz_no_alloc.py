"""The FP64 solvers that allocate nothing (fvg_cg_noalloc / fvg_bicgstab_noalloc), called
through the C ABI, against the allocating entry points (fvg_cg / fvg_bicgstab) on the
same systems: the same kernels on the same data, so the same bits (solution, iteration
count, residuals, history).  Two calls in a row, the second on a smaller system, since
every call shares the one scratch array and the first call's contents are left in it."""
import numpy as np
import pytest

import fvg
import meshgen
import oracle as O
import problems

pytestmark = pytest.mark.gpu


def _system(kind, raw, seed):
    om = O.OracleMesh(raw)
    g = om.geometry()
    if kind == "sym":
        pr = problems.diffusion(raw, g, kc=meshgen.lognormal_cells(om.n_cells, 1.0, seed), seed=seed)
    else:
        pr = problems.convection_diffusion(raw, g, pe=(1.0, 50.0), seed=seed)
    return om.assemble(pr["prog"], pr["face_fields"], pr["cell_fields"], pr["bc_i"], pr["bc_e"])


def _both(ctx, rp, ci, va, b, solver):
    n = rp.shape[0] - 1
    A = fvg.DistributedMatrix.from_csr(ctx, rp, ci, va)
    M = (fvg.IncompleteCholesky if solver == "cg" else fvg.IncompleteLowerUpper).from_matrix(A, 0)
    B = fvg.DistributedVector.from_data(ctx, b)
    opts = fvg.LinearSolverOptions(1e-8, 0.0, 3000)
    x_ref = fvg.DistributedVector.from_data(ctx, np.full(n, 0.25))
    x_new = fvg.DistributedVector.from_data(ctx, np.full(n, 0.25))
    if solver == "cg":
        ref = fvg.conjugate_gradient(x_ref, A, B, M, opts)
        used = ctx.bytes_in_use()
        new = fvg.conjugate_gradient_noalloc(x_new, A, B, M, opts)
    else:
        ref = fvg.bi_conjugate_gradient_stab(x_ref, A, B, M, opts)
        used = ctx.bytes_in_use()
        new = fvg.bi_conjugate_gradient_stab_noalloc(x_new, A, B, M, opts)
    assert ctx.bytes_in_use() == used        # nothing allocated through the context
    assert np.array_equal(x_new.numpy(), x_ref.numpy())
    for k in ("iterations", "initial_residual", "final_residual", "restarts"):
        assert getattr(new, k) == getattr(ref, k), k
    assert np.array_equal(new.history, ref.history)
    assert ref.final_residual <= 1e-8 * ref.initial_residual
    return ref


@pytest.mark.parametrize("solver,kind,meshes", [
    ("cg", "sym", (lambda: meshgen.tet3d(12, seed=21), lambda: meshgen.tri2d(30, seed=22))),
    ("bicgstab", "nonsym", (lambda: meshgen.prism3d(16, layers=8, seed=23), lambda: meshgen.tri2d(20, seed=24))),
])
def test_noalloc_solvers_match_the_allocating_ones(ctx, solver, kind, meshes):
    sizes = []
    for j, mk in enumerate(meshes):
        rp, ci, va, b = _system(kind, mk(), 40 + j)
        print(f"no-alloc {solver} call {j}: n={rp.shape[0] - 1}", flush=True)
        _both(ctx, rp, ci, va, b, solver)
        sizes.append(rp.shape[0] - 1)
    assert sizes[1] < sizes[0]


def test_noalloc_refuses_what_it_cannot_hold(ctx):
    rp, ci, va, b = _system("sym", meshgen.tri2d(6, seed=25), 45)
    n = rp.shape[0] - 1
    A = fvg.DistributedMatrix.from_csr(ctx, rp, ci, va)
    M = fvg.IncompleteCholesky.from_matrix(A, 0)
    B = fvg.DistributedVector.from_data(ctx, b)
    x = fvg.DistributedVector(ctx, n)
    cap = fvg.noalloc_scratch_bytes()
    assert cap == 160 << 20
    with pytest.raises(fvg.FvgError) as e:   # the history of 30M iterations does not fit
        fvg.conjugate_gradient_noalloc(x, A, B, M, fvg.LinearSolverOptions(1e-8, 0.0, cap // 8 - 100))
    assert fvg.STATUS[e.value.status] == "FVG_E_INVALID"
    with pytest.raises(fvg.FvgError) as e:
        fvg.conjugate_gradient_noalloc(x, A, B, M, fvg.LinearSolverOptions(1e-8, 0.0, 100,
                                                                            precision=fvg.PRECISION_MIXED))
    assert fvg.STATUS[e.value.status] == "FVG_E_UNSUPPORTED"
    # after the refusals the entry point still works
    info = fvg.conjugate_gradient_noalloc(x, A, B, M, fvg.LinearSolverOptions(1e-8, 0.0, 100))
    assert info.final_residual <= 1e-8 * info.initial_residual
