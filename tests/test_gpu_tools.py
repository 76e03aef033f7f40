"""GPU parity of the fvm tools and the Corrected / Linear / LimitedLinear schemes.

compute_gradients, compute_limiters and assembly with the new schemes are held to
bit-exactness against the oracle (same operation order); the transient step of the
crate's ex4 (gradients -> limiters -> assembly -> BiCGStab + ILU(0)) to the solver bar:
x within 1e-6 relative of the oracle's at a 1e-8 relative residual.
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
LIMITERS = [("limited_linear", fvg.limiters.LimitedLinear(1.0)), ("minmod", fvg.limiters.MinMod()),
            ("vanleer", fvg.limiters.VanLeer()), ("vanalbada", fvg.limiters.VanAlbada()),
            ("sweby", fvg.limiters.Sweby(1.5)), ("superbee", fvg.limiters.SuperBee())]


def _setup(name, ncomp, seed=0):
    raw = MESHES[name]()
    om = O.OracleMesh(raw)
    rng = np.random.default_rng(seed)
    n, F, nb = om.n_cells, om.n_faces, om.n_bnd
    shape = (n,) if ncomp == 1 else (n, ncomp)
    values = rng.normal(size=shape)
    bci = rng.uniform(0.0, 1.0, nb)
    bce = rng.normal(size=(nb,) if ncomp == 1 else (nb, ncomp))
    flux = rng.normal(size=F)
    return raw, om, values, bci, bce, flux


@pytest.mark.parametrize("name", sorted(MESHES))
@pytest.mark.parametrize("ncomp", [1, 3])
def test_gradients_bit_exact(ctx, name, ncomp):
    raw, om, values, bci, bce, flux = _setup(name, ncomp, seed=ncomp)
    mesh = fvg.Mesh(ctx, raw)
    v = fvg.DistributedVector.from_data(ctx, values.ravel())
    g = fvg.DistributedVector(ctx, om.n_cells * ncomp * om.dim)
    fvg.compute_gradients(g, v, fvg.FaceConstraints(bci, bce), mesh, ncomp=ncomp)
    ref = om.gradients(values, bci, bce)
    assert np.array_equal(g.numpy(), ref.ravel())


@pytest.mark.parametrize("name", sorted(MESHES))
@pytest.mark.parametrize("lname,limiter", LIMITERS)
def test_limiters_bit_exact(ctx, name, lname, limiter):
    raw, om, values, bci, bce, flux = _setup(name, 1, seed=5)
    mesh = fvg.Mesh(ctx, raw)
    bc = fvg.FaceConstraints(bci, bce)
    v = fvg.DistributedVector.from_data(ctx, values)
    g = fvg.DistributedVector(ctx, om.n_cells * om.dim)
    fvg.compute_gradients(g, v, bc, mesh)
    grads = om.gradients(values, bci, bce)
    lim = fvg.DistributedVector(ctx, om.n_faces)
    for interp, fng in (("upwind", "corrected"), ("linear", "orthogonal"), ("linear", "corrected"),
                        ("upwind", "orthogonal")):
        isch = fvg.Upwind(flux) if interp == "upwind" else fvg.Linear()
        fsch = fvg.Corrected(g, 0.75) if fng == "corrected" else fvg.Orthogonal()
        fvg.compute_limiters(lim, v, g, limiter, isch, fsch, bc, mesh)
        ref = om.limiters(values, grads, lname, limiter.k, interp, flux if interp == "upwind" else None,
                          grads if fng == "corrected" else None, 0.75, bci, bce)
        assert np.array_equal(lim.numpy(), ref), (interp, fng)


@pytest.mark.parametrize("name", ["tet4", "prism8"])
@pytest.mark.parametrize("ncomp", [2, 3])
def test_vector_limiters_bit_exact(ctx, name, ncomp):
    """Field<Vector<ncomp>> with Matrix<ncomp, dim> gradients (ex6's velocity limiters)."""
    raw, om, values, bci, bce, flux = _setup(name, ncomp, seed=7)
    mesh = fvg.Mesh(ctx, raw)
    bc = fvg.FaceConstraints(bci, bce)
    v = fvg.DistributedVector.from_data(ctx, values.ravel())
    g = fvg.DistributedVector(ctx, om.n_cells * ncomp * om.dim)
    fvg.compute_gradients(g, v, bc, mesh, ncomp=ncomp)
    grads = om.gradients(values, bci, bce)
    lim = fvg.DistributedVector(ctx, om.n_faces)
    fvg.compute_limiters(lim, v, g, fvg.limiters.LimitedLinear(1.0), fvg.Upwind(flux), fvg.Corrected(g, 1.0), bc,
                         mesh, ncomp=ncomp)
    ref = om.limiters(values, grads, "limited_linear", 1.0, "upwind", flux, grads, 1.0, bci, bce)
    assert np.array_equal(lim.numpy(), ref)


@pytest.mark.parametrize("name", sorted(MESHES))
def test_new_schemes_assemble_bit_exact(ctx, name):
    """time(Euler) + convection(LimitedLinear) + convection(Linear) - laplacian(Corrected)
    - source, with the gradient and limiter fields resident on the device."""
    raw, om, values, bci, bce, flux = _setup(name, 1, seed=9)
    mesh = fvg.Mesh(ctx, raw)
    rng = np.random.default_rng(11)
    n, F = om.n_cells, om.n_faces
    kf = rng.uniform(0.5, 2.0, F)
    src = rng.normal(size=n)
    w = rng.normal(size=F)
    bc = fvg.FaceConstraints(bci, bce)
    v = fvg.DistributedVector.from_data(ctx, values)
    g = fvg.DistributedVector(ctx, n * om.dim)
    fvg.compute_gradients(g, v, bc, mesh)
    lim = fvg.DistributedVector(ctx, F)
    fvg.compute_limiters(lim, v, g, fvg.limiters.LimitedLinear(1.0), fvg.Upwind(flux), fvg.Corrected(g, 1.0), bc,
                         mesh)
    eq = (fvg.time(fvg.Euler(values, 0.05)) + fvg.convection(fvg.LimitedLinear(w, lim), flux)
          + fvg.convection(fvg.Linear(), 0.5 * flux) - fvg.laplacian(fvg.Corrected(g, 0.8), kf)
          - fvg.source(src))
    lhs, rhs = fvg.assemble(eq, bc, mesh)
    grads = om.gradients(values, bci, bce)
    lims = lim.numpy()
    OPc = O.OP
    prog = [(OPc["TIME_EULER"], 0, -1, -1, 0.05), (OPc["CONVECTION_LIMITED_LINEAR"], 0, 1, 2, 0.0), (OPc["ADD"], 0, 0, 0, 0.0),
            (OPc["CONVECTION_LINEAR"], 3, 0, 0, 0.0), (OPc["ADD"], 0, 0, 0, 0.0),
            (OPc["LAPLACIAN_CORRECTED"], 4, 1, 0, 0.8), (OPc["SUB"], 0, 0, 0, 0.0),
            (OPc["SOURCE"], 2, 0, 0, 0.0), (OPc["SUB"], 0, 0, 0, 0.0)]
    rp, ci, va, b = om.assemble(prog, [flux, w, lims, 0.5 * flux, kf], [values, grads.ravel(), src], bci, bce)
    grp, gci, gva = lhs.csr()
    assert np.array_equal(grp, rp) and np.array_equal(gci, ci)
    assert np.array_equal(gva, va)
    assert np.array_equal(rhs.numpy(), b)


def test_transient_step_like_ex4(ctx):
    """One time step of the crate's ex4 on the device against the oracle: gradients,
    limiters (LimitedLinear(1), Upwind, Corrected), assembly of time + convection
    (LimitedLinear) - laplacian(Corrected) - source, then BiCGStab + ILU(1), as
    examples/ex4/ex4.rs:137-155 does."""
    raw = meshgen.prism3d(12, layers=6, seed=4)
    om = O.OracleMesh(raw)
    geom = om.geometry()
    n, F, nb = om.n_cells, om.n_faces, om.n_bnd
    rng = np.random.default_rng(3)
    u0 = np.exp(-8.0 * ((geom["cell_center"] - 0.5) ** 2).sum(1))
    vel = np.array([1.0, 0.5, 0.25])
    flux = geom["face_normal"] @ vel
    kf = np.full(F, 1e-2)
    src = rng.uniform(0.0, 0.1, n)
    bf = np.nonzero(raw["face_patch"] >= 0)[0]
    inflow = flux[bf] < 0
    bci = np.where(inflow, 0.0, 1.0)          # inflow: fixed value 0; outflow: zero gradient
    bce = np.zeros(nb)
    mesh = fvg.Mesh(ctx, raw)
    bc = fvg.FaceConstraints(bci, bce)
    u = fvg.DistributedVector.from_data(ctx, u0)
    g = fvg.DistributedVector(ctx, n * 3)
    lim = fvg.DistributedVector(ctx, F)
    fvg.compute_gradients(g, u, bc, mesh)
    fvg.compute_limiters(lim, u, g, fvg.limiters.LimitedLinear(1.0), fvg.Upwind(flux), fvg.Corrected(g, 1.0), bc,
                         mesh)
    eq = (fvg.time(fvg.Euler(u0, 0.01)) + fvg.convection(fvg.LimitedLinear(flux, lim), flux)
          - fvg.laplacian(fvg.Corrected(g, 1.0), kf) - fvg.source(src))
    lhs, rhs = fvg.assemble(eq, bc, mesh)
    ilu = fvg.IncompleteLowerUpper.from_matrix(lhs, 1)
    x = fvg.DistributedVector(ctx, n)
    info = fvg.bi_conjugate_gradient_stab(x, lhs, rhs, ilu, fvg.LinearSolverOptions(1e-8, 0.0, 1000))
    # the same step on the oracle
    grads = om.gradients(u0, bci, bce)
    lims = om.limiters(u0, grads, "limited_linear", 1.0, "upwind", flux, grads, 1.0, bci, bce)
    assert np.array_equal(g.numpy(), grads.ravel()) and np.array_equal(lim.numpy(), lims)
    OPc = O.OP
    prog = [(OPc["TIME_EULER"], 0, -1, -1, 0.01), (OPc["CONVECTION_LIMITED_LINEAR"], 0, 0, 1, 0.0),
            (OPc["ADD"], 0, 0, 0, 0.0), (OPc["LAPLACIAN_CORRECTED"], 2, 1, 0, 1.0), (OPc["SUB"], 0, 0, 0, 0.0),
            (OPc["SOURCE"], 2, 0, 0, 0.0), (OPc["SUB"], 0, 0, 0, 0.0)]
    rp, ci, va, b = om.assemble(prog, [flux, lims, kf], [u0, grads.ravel(), src], bci, bce)
    assert np.array_equal(lhs.csr()[2], va) and np.array_equal(rhs.numpy(), b)
    P = O.OraclePrec("ilu", rp, ci, va, 1)
    xo, io = O.solve("bicgstab", rp, ci, va, P, b, rtol=1e-8, atol=0.0, max_iter=1000)
    assert abs(info.iterations - io["iterations"]) <= 2
    assert np.linalg.norm(x.numpy() - xo) <= 1e-6 * np.linalg.norm(xo)
