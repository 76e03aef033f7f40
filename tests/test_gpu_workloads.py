"""The workload generators through the device's own front half and its FP64 solvers.

Each case runs what a crate user runs: fvg.Mesh (Mesh::compute on the device) ->
fields from the device geometry -> Assembler into DistributedMatrix::from_mesh ->
IncompleteCholesky / IncompleteLowerUpper::from_matrix(&A, 0) -> conjugate_gradient /
bi_conjugate_gradient_stab from x = 0 to a 1e-8 relative residual, and compares with the
FP64 oracle on the same inputs.

Bars: geometry, CSR row_ptr / col_idx, assembled values and right-hand side bit-exact
(the device applies faces in the reference's order; the workload bar for values is
1e-12 relative); one preconditioner apply bit-exact; iterations within +-2 of the
oracle's; x within 1e-6 relative of the oracle's x; true relative residual <= 1e-8.

Reduced sizes compare with the oracle run here.  At the full workload sizes the oracle's
numbers come from tests/golden/workloads_full.json (written by
tests/golden/make_workload_fixture.py): SHA-256 of the CSR arrays, values and
right-hand side, iteration counts, the solution's norm and its values at 512 seeded
indices (the sampled relative error stands in for the full one).
"""
import hashlib
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp

import fvg
import oracle as O
import workloads

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GEOM_KEYS = ("face_area", "face_center", "face_normal", "owner", "neighbor", "cell_volume", "cell_center")


def _device_solve(ctx, raw, fields):
    mesh = fvg.Mesh(ctx, raw)
    geom = mesh.geometry()
    pr = fields(geom)
    asm = fvg.Assembler(mesh, pr["prog"], pr["face_fields"], pr["cell_fields"], pr["bc_i"], pr["bc_e"])
    lhs = fvg.DistributedMatrix.from_mesh(ctx, mesh)
    rhs = fvg.DistributedVector.from_size(ctx, mesh.n_cells)
    asm.run(lhs, rhs)
    Prec = fvg.IncompleteCholesky if pr["prec"] == "ic" else fvg.IncompleteLowerUpper
    M = Prec.from_matrix(lhs, 0)
    x = fvg.DistributedVector.from_size(ctx, mesh.n_cells)
    fn = fvg.conjugate_gradient if pr["solver"] == "cg" else fvg.bi_conjugate_gradient_stab
    info = fn(x, lhs, rhs, M, fvg.LinearSolverOptions(1e-8, 0.0, 5000))
    return dict(mesh=mesh, geom=geom, pr=pr, lhs=lhs, rhs=rhs, M=M, x=x, info=info)


def _true_residual(ctx, d):
    y = fvg.DistributedVector.from_size(ctx, d["mesh"].n_cells)
    d["lhs"].imul(y, d["x"])
    b = d["rhs"].numpy()
    return float(np.linalg.norm(b - y.numpy()) / np.linalg.norm(b))


REDUCED = [("diffusion_2d", {}),                       # the full 8,192-triangle workload
           ("diffusion_3d", {"nx": 12}), ("diffusion_3d", {"nx": 16}),
           ("convdiff_prism", {"nx": 24, "layers": 12})]


@pytest.mark.parametrize("name,kw", REDUCED, ids=[f"{n}-{'-'.join(map(str, k.values())) or 'full'}"
                                                  for n, k in REDUCED])
def test_workload_against_the_oracle(ctx, name, kw):
    raw, fields = workloads.WORKLOADS[name](**kw)
    print(f"workload {name} {kw}", flush=True)
    d = _device_solve(ctx, raw, fields)
    om = O.OracleMesh(raw)
    og = om.geometry()
    for k in GEOM_KEYS:
        assert np.array_equal(d["geom"][k], og[k]), k
    pr = fields(og)
    rp, ci, va, b = om.assemble(pr["prog"], pr["face_fields"], pr["cell_fields"], pr["bc_i"], pr["bc_e"])
    grp, gci, gva = d["lhs"].csr()
    assert np.array_equal(grp, rp) and np.array_equal(gci, ci)
    assert np.array_equal(gva, va)
    assert np.array_equal(d["rhs"].numpy(), b)
    P = O.OraclePrec(pr["prec"], rp, ci, va, 0)
    n = om.n_cells
    r = np.random.default_rng(3).normal(size=n)
    z = fvg.DistributedVector(ctx, n)
    d["M"].precondition(z, fvg.DistributedVector.from_data(ctx, r))
    assert np.array_equal(z.numpy(), P.apply(r))
    xo, io = O.solve(pr["solver"], rp, ci, va, P, b, rtol=1e-8, atol=0.0, max_iter=5000)
    info, xg = d["info"], d["x"].numpy()
    assert abs(info.iterations - io["iterations"]) <= 2, (info.iterations, io["iterations"])
    assert np.linalg.norm(xg - xo) <= 1e-6 * np.linalg.norm(xo)
    assert abs(info.initial_residual - io["initial_residual"]) <= 1e-12 * io["initial_residual"]
    assert info.final_residual <= 1e-8 * info.initial_residual
    S = sp.csr_matrix((va, ci, rp), shape=(n, n))
    assert np.linalg.norm(b - S @ xg) <= 1e-8 * np.linalg.norm(b) * 1.5
    assert _true_residual(ctx, d) <= 1e-8 * 1.5


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.mark.parametrize("name", ["diffusion_2d", "diffusion_3d", "convdiff_prism"])
def test_full_size_workload_against_the_oracle_fixture(ctx, name):
    """1,572,864 tets (CG + IC(0)) and 2,097,152 prisms (BiCGStab + ILU(0)) at the
    benchmark's sizes, and the 8,192-triangle workload, against the oracle's recorded run."""
    with open(os.path.join(HERE, "golden", "workloads_full.json")) as f:
        ref = json.load(f)[name]
    raw, fields = workloads.WORKLOADS[name]()
    print(f"workload {name} full size", flush=True)
    d = _device_solve(ctx, raw, fields)
    assert d["mesh"].n_cells == ref["n_cells"] and d["lhs"].nnz == ref["nnz"]
    grp, gci, gva = d["lhs"].csr()
    assert _sha(grp) == ref["sha_row_ptr"] and _sha(gci) == ref["sha_col"]
    assert _sha(gva) == ref["sha_values"]
    assert _sha(d["rhs"].numpy()) == ref["sha_rhs"]
    info, xg = d["info"], d["x"].numpy()
    assert abs(info.iterations - ref["iterations"]) <= 2, (info.iterations, ref["iterations"])
    assert abs(info.initial_residual - ref["initial_residual"]) <= 1e-12 * ref["initial_residual"]
    assert info.final_residual <= 1e-8 * info.initial_residual
    assert abs(np.linalg.norm(xg) - ref["x_norm"]) <= 1e-6 * ref["x_norm"]
    idx = np.array(ref["sample_idx"])
    xo = np.array(ref["x_samples"])
    assert np.linalg.norm(xg[idx] - xo) <= 1e-6 * np.linalg.norm(xo)
    assert _true_residual(ctx, d) <= 1e-8 * 1.5
