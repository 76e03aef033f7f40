"""GPU parity of time(Backward) and user-defined terms against the oracle, bit for bit.

time(Backward) follows schemes/time/backward.rs:21-62 (coefficients recomputed on the
device in the same operation order); a user term is any other `impl Term`
(terms/mod.rs:24-40) whose per-face (li, lj, r) and per-cell (l, r) values the caller
supplies.  Also the program validation the device path applies before a launch.
"""
import numpy as np
import pytest

import fvg
import meshgen
import oracle as O

pytestmark = pytest.mark.gpu

MESHES = {
    "tri16": lambda: meshgen.tri2d(16, seed=3),
    "tet4": lambda: meshgen.tet3d(4, seed=5),
    "prism8": lambda: meshgen.prism3d(8, layers=4, seed=2),
}


def _fields(raw, seed):
    om = O.OracleMesh(raw)
    rng = np.random.default_rng(seed)
    n, F, nb = om.n_cells, om.n_faces, om.n_bnd
    return om, rng, n, F, rng.uniform(0.0, 1.0, nb), rng.normal(size=nb)


def _check(lhs, rhs, ref):
    rp, ci, va, b = ref
    grp, gci, gva = lhs.csr()
    assert np.array_equal(grp, rp) and np.array_equal(gci, ci)
    assert np.array_equal(gva, va)
    assert np.array_equal(rhs.numpy(), b)


@pytest.mark.parametrize("name", sorted(MESHES))
@pytest.mark.parametrize("with_density", [False, True])
def test_backward_assemble_bit_exact(ctx, name, with_density):
    """time(Backward) + convection(Upwind) - laplacian(Orthogonal) - source, the shape of
    a second-order transient step."""
    raw = MESHES[name]()
    om, rng, n, F, bci, bce = _fields(raw, 21)
    mesh = fvg.Mesh(ctx, raw)
    p1, p2, src = rng.normal(size=n), rng.normal(size=n), rng.normal(size=n)
    flux, kf = rng.normal(size=F), rng.uniform(0.5, 2.0, F)
    dens = [rng.uniform(0.5, 2.0, n) for _ in range(3)] if with_density else [None] * 3
    dt, dt_2 = 0.013, 0.017
    eq = (fvg.time(fvg.Backward(p1, p2, dt, dt_2, *dens)) + fvg.convection(fvg.Upwind(flux), flux)
          - fvg.laplacian(fvg.Orthogonal(), kf) - fvg.source(src))
    lhs, rhs = fvg.assemble(eq, fvg.FaceConstraints(bci, bce), mesh)
    OP = O.OP
    cf = [p1, p2] + (dens if with_density else []) + [src]
    s = len(cf) - 1
    prog = [(OP["TIME_BACKWARD"], 0, 1, 2 if with_density else -1, dt), (OP["PARAM"], 0, 0, 0, dt_2),
            (OP["CONVECTION_UPWIND"], 0, 0, 0, 0.0), (OP["ADD"], 0, 0, 0, 0.0),
            (OP["LAPLACIAN_ORTHOGONAL"], 1, 0, 0, 0.0), (OP["SUB"], 0, 0, 0, 0.0),
            (OP["SOURCE"], s, 0, 0, 0.0), (OP["SUB"], 0, 0, 0, 0.0)]
    _check(lhs, rhs, om.assemble(prog, [flux, kf], cf, bci, bce))


@pytest.mark.parametrize("name", sorted(MESHES))
def test_user_term_assemble_bit_exact(ctx, name):
    """A user term with face and cell parts, one with only face terms and one with only
    cell terms, combined with built-in terms; the arrays live on the device."""
    raw = MESHES[name]()
    om, rng, n, F, bci, bce = _fields(raw, 22)
    mesh = fvg.Mesh(ctx, raw)
    fa = [rng.normal(size=F) for _ in range(3)]
    ca = [rng.uniform(0.5, 1.5, n), rng.normal(size=n)]
    fb = [rng.normal(size=F) for _ in range(3)]
    cc = [rng.normal(size=n), rng.normal(size=n)]
    kf = rng.uniform(0.5, 2.0, F)
    dev = lambda a: fvg.DistributedVector.from_data(ctx, a)
    eq = (fvg.user_term([dev(a) for a in fa], [dev(a) for a in ca]) - fvg.laplacian(fvg.Orthogonal(), kf)
          + fvg.user_term(face_terms=fb) - (-fvg.user_term(cell_terms=cc)))
    lhs, rhs = fvg.assemble(eq, fvg.FaceConstraints(bci, bce), mesh)
    OP = O.OP
    prog = [(OP["USER_TERM"], 0, 0, 0, 0.0), (OP["LAPLACIAN_ORTHOGONAL"], 3, 0, 0, 0.0), (OP["SUB"], 0, 0, 0, 0.0),
            (OP["USER_TERM"], 4, -1, 0, 0.0), (OP["ADD"], 0, 0, 0, 0.0),
            (OP["USER_TERM"], -1, 2, 0, 0.0), (OP["NEG"], 0, 0, 0, 0.0), (OP["SUB"], 0, 0, 0, 0.0)]
    _check(lhs, rhs, om.assemble(prog, fa + [kf] + fb, ca + cc, bci, bce))


def test_deep_program_bit_exact(ctx):
    """A program that fills the device's eight-slot evaluation stack: eight terms pushed
    before the first operator (a right-leaning sum), against the oracle."""
    raw = MESHES["tet4"]()
    om, rng, n, F, bci, bce = _fields(raw, 23)
    mesh = fvg.Mesh(ctx, raw)
    flux = [rng.normal(size=F) for _ in range(4)]
    srcs = [rng.normal(size=n) for _ in range(4)]
    OP = O.OP
    prog = []
    for k in range(4):
        prog.append((OP["CONVECTION_UPWIND"], k, k, 0, 0.0))
        prog.append((OP["SOURCE"], k, 0, 0, 0.0))
    for k in range(7):
        prog.append((OP["SUB"] if k % 2 else OP["ADD"], 0, 0, 0, 0.0))
    lhs, rhs = fvg.assemble_program(mesh, prog, flux, srcs, bci, bce)
    _check(lhs, rhs, om.assemble(prog, flux, srcs, bci, bce))


def test_program_validation(ctx):
    raw = MESHES["tri16"]()
    om, rng, n, F, bci, bce = _fields(raw, 24)
    mesh = fvg.Mesh(ctx, raw)
    p = rng.normal(size=n)
    f = rng.normal(size=F)
    OP = O.OP
    bad = [
        [(OP["TIME_BACKWARD"], 0, 1, -1, 0.1)],                                  # no PARAM word
        [(OP["TIME_BACKWARD"], 0, 1, -1, 0.1), (OP["ADD"], 0, 0, 0, 0.0)],       # wrong word after it
        [(OP["PARAM"], 0, 0, 0, 0.1)],                                           # PARAM on its own
        [(OP["TIME_BACKWARD"], 0, 1, 0, 0.1), (OP["PARAM"], 0, 0, 0, 0.1)],      # densities 0..2 of 2 fields
        [(OP["USER_TERM"], 0, -1, 0, 0.0)],                                      # faces 0..2 of 1 field
        [(OP["USER_TERM"], -1, 1, 0, 0.0)],                                      # cells 1..2 of 2 fields
        [(OP["USER_TERM"], -2, -1, 0, 0.0)],
        [(OP["SOURCE"], 0, 0, 0, 0.0)] * 9 + [(OP["ADD"], 0, 0, 0, 0.0)] * 8,   # deeper than 8
    ]
    for prog in bad:
        with pytest.raises(fvg.FvgError) as e:
            fvg.assemble_program(mesh, prog, [f], [p, p], bci, bce)
        assert e.value.status == 1, prog
    # a gradient-length array as a Backward level is refused (cell values are n_cells long)
    with pytest.raises(fvg.FvgError):
        fvg.assemble_program(mesh, [(OP["TIME_BACKWARD"], 0, 1, -1, 0.1), (OP["PARAM"], 0, 0, 0, 0.1)], [],
                             [p, np.zeros(2 * n)], bci, bce)


PROGRAMS = {
    # -laplacian(Orthogonal) + convection(Upwind) - source: the steady workloads' shape
    "orthogonal": lambda OP, s: ([(OP["LAPLACIAN_ORTHOGONAL"], 0, 0, 0, 0.0), (OP["NEG"], 0, 0, 0, 0.0),
                                  (OP["CONVECTION_UPWIND"], 1, 1, 0, 0.0), (OP["ADD"], 0, 0, 0, 0.0),
                                  (OP["SOURCE"], 0, 0, 0, 0.0), (OP["SUB"], 0, 0, 0, 0.0)], [s["k"], s["flux"]],
                                 [s["src"]]),
    # time(Euler) + convection(Linear) - laplacian(Corrected): gradients gathered per face
    "corrected": lambda OP, s: ([(OP["TIME_EULER"], 0, -1, -1, 0.02), (OP["CONVECTION_LINEAR"], 1, 0, 0, 0.0),
                                 (OP["ADD"], 0, 0, 0, 0.0), (OP["LAPLACIAN_CORRECTED"], 0, 1, 0, 0.7),
                                 (OP["SUB"], 0, 0, 0, 0.0)], [s["k"], s["flux"]], [s["src"], s["grads"]]),
    # a depth-4 program: a - (b - (c + d))
    "deep4": lambda OP, s: ([(OP["LAPLACIAN_ORTHOGONAL"], 0, 0, 0, 0.0), (OP["CONVECTION_UPWIND"], 1, 1, 0, 0.0),
                             (OP["SOURCE"], 0, 0, 0, 0.0), (OP["LINEAR_SOURCE"], 1, 0, 0, 0.0),
                             (OP["ADD"], 0, 0, 0, 0.0), (OP["SUB"], 0, 0, 0, 0.0), (OP["SUB"], 0, 0, 0, 0.0)],
                            [s["k"], s["flux"]], [s["src"], s["c"]]),
}


def _program_fields(om, seed):
    rng = np.random.default_rng(seed)
    n, F = om.n_cells, om.n_faces
    return dict(k=rng.uniform(0.5, 2.0, F), flux=rng.normal(size=F), src=rng.normal(size=n),
                grads=rng.normal(size=n * om.dim), c=rng.uniform(0.1, 1.0, n))


@pytest.mark.parametrize("name", sorted(MESHES))
@pytest.mark.parametrize("prog_name", sorted(PROGRAMS))
def test_entry_kernel_matches_row_kernel(ctx, name, prog_name):
    """One lane per cell-face entry (the default where every off-diagonal slot belongs to
    one face; its rows staged in LDS, or stored slot by slot) and one lane per row give
    the same bits, and all the oracle's."""
    raw = MESHES[name]()
    om, rng, n, F, bci, bce = _fields(raw, 25)
    mesh = fvg.Mesh(ctx, raw)
    prog, ff, cf = PROGRAMS[prog_name](O.OP, _program_fields(om, 26))
    ref = om.assemble(prog, ff, cf, bci, bce)
    try:
        for mode in (0, 1, 2):
            ctx.set_tuning(fvg.TUNE_ASSEMBLE, mode)
            lhs, rhs = fvg.assemble_program(mesh, prog, ff, cf, bci, bce)
            _check(lhs, rhs, ref)
    finally:
        ctx.set_tuning(fvg.TUNE_ASSEMBLE, 0)


@pytest.mark.parametrize("prog_name", ["orthogonal", "corrected"])
def test_faces_sharing_a_slot(ctx, prog_name):
    """Cells joined by two faces (interior faces split in two): both faces add into the
    same off-diagonal slot, in face order, so assembly keeps one lane per row."""
    raw0 = meshgen.tri2d(12, seed=6)
    inner = np.nonzero(raw0["face_patch"] < 0)[0]
    raw = meshgen.split_faces_2d(raw0, inner[::5])
    om, rng, n, F, bci, bce = _fields(raw, 27)
    assert om.nnz == O.OracleMesh(raw0).nnz and F > raw0["face_patch"].shape[0]
    mesh = fvg.Mesh(ctx, raw)
    prog, ff, cf = PROGRAMS[prog_name](O.OP, _program_fields(om, 28))
    lhs, rhs = fvg.assemble_program(mesh, prog, ff, cf, bci, bce)
    _check(lhs, rhs, om.assemble(prog, ff, cf, bci, bce))
