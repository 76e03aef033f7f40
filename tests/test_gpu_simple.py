"""GPU parity of the SIMPLE helpers (tools/hbya.rs) and the wall distance
(tools/wall_distance.rs) against the oracle.

compute_hbya_ainv, intepolate_hbya_ainv_faces, correct_phi and correct_velocity repeat the
reference's operation order and are held to bit-exactness.  compute_wall_distance runs CG
solves, so its rounds are held to the solver bar (1e-6 relative of the oracle's result);
with no rounds it is the final normalisation alone, whose pow() may differ from the
C library's by an ulp (held to 4 ulp).
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


def _momentum(om, rng):
    n, F, nb = om.n_cells, om.n_faces, om.n_bnd
    OP = O.OP
    prog = [(OP["TIME_EULER"], 0, -1, -1, 0.01), (OP["CONVECTION_UPWIND"], 0, 0, 0, 0.0), (OP["ADD"], 0, 0, 0, 0.0),
            (OP["LAPLACIAN_ORTHOGONAL"], 1, 0, 0, 0.0), (OP["SUB"], 0, 0, 0, 0.0)]
    ff = [rng.normal(size=F), rng.uniform(0.5, 2.0, F)]
    cf = [rng.normal(size=n)]
    bci, bce = rng.uniform(0, 1, nb), rng.normal(size=nb)
    return prog, ff, cf, bci, bce


@pytest.mark.parametrize("name", sorted(MESHES))
def test_hbya_ainv_bit_exact(ctx, name):
    raw = MESHES[name]()
    om = O.OracleMesh(raw)
    rng = np.random.default_rng(41)
    n, D = om.n_cells, om.dim
    mesh = fvg.Mesh(ctx, raw)
    prog, ff, cf, bci, bce = _momentum(om, rng)
    lhs, _ = fvg.assemble_program(mesh, prog, ff, cf, bci, bce)
    rp, ci, va = lhs.csr()
    mrhs, vel = rng.normal(size=(n, D)), rng.normal(size=(n, D))
    hb = fvg.DistributedVector(ctx, n * D)
    ai = fvg.DistributedVector(ctx, n)
    fvg.compute_hbya_ainv(hb, ai, vel.ravel(), lhs, mrhs.ravel(), mesh)
    ref_h, ref_a = om.hbya_ainv(rp, ci, va, mrhs, vel)
    assert np.array_equal(hb.numpy(), ref_h.ravel()) and np.array_equal(ai.numpy(), ref_a)


def test_hbya_ainv_missing_diagonal(ctx):
    raw = MESHES["tri16"]()
    om = O.OracleMesh(raw)
    n, D = om.n_cells, om.dim
    mesh = fvg.Mesh(ctx, raw)
    rp = np.arange(n + 1, dtype=np.int32)
    col = np.roll(np.arange(n, dtype=np.int32), 1)      # one off-diagonal entry per row
    a = fvg.DistributedMatrix.from_csr(ctx, rp, col, np.ones(n), n)
    hb = fvg.DistributedVector(ctx, n * D)
    ai = fvg.DistributedVector(ctx, n)
    with pytest.raises(fvg.FvgError) as e:
        fvg.compute_hbya_ainv(hb, ai, np.zeros(n * D), a, np.zeros(n * D), mesh)
    assert e.value.status == 6
    with pytest.raises(RuntimeError):
        om.hbya_ainv(rp, col, np.ones(n), np.zeros((n, D)), np.zeros((n, D)))


def _closures(om, rng, walls_frac=0.5):
    nb, D = om.n_bnd, om.dim
    walls = rng.uniform(size=nb) < walls_frac
    ubi, ube = np.where(walls, 0.0, 1.0), np.where(walls[:, None], 0.0, rng.normal(size=(nb, D)))
    pbi = np.where(walls, 1.0, 0.0)
    pbe = np.where(walls, 0.0, rng.normal(size=nb))
    return ubi, ube, pbi, pbe


@pytest.mark.parametrize("name", sorted(MESHES))
def test_interpolate_hbya_ainv_faces_bit_exact(ctx, name):
    raw = MESHES[name]()
    om = O.OracleMesh(raw)
    rng = np.random.default_rng(42)
    n, F, D = om.n_cells, om.n_faces, om.dim
    mesh = fvg.Mesh(ctx, raw)
    hbya, ainv = rng.normal(size=(n, D)), rng.uniform(0.1, 1.0, n)
    flux, lim = rng.normal(size=F), rng.uniform(0, 1, F)
    ubi, ube, pbi, pbe = _closures(om, rng)
    ubc, pbc = fvg.FaceConstraints(ubi, ube), fvg.FaceConstraints(pbi, pbe)
    hn = fvg.DistributedVector(ctx, F)
    af = fvg.DistributedVector(ctx, F)
    schemes = [(fvg.Linear(), "linear"), (fvg.Upwind(flux), ("upwind", flux)),
               (fvg.LimitedLinear(flux, lim), ("limited_linear", flux, lim))]
    for hs, hr in schemes:
        for as_, ar in schemes:
            fvg.intepolate_hbya_ainv_faces(hn, af, hbya.ravel(), ainv, ubc, pbc, hs, as_, mesh)
            ref_hn, ref_af = om.interpolate_hbya_ainv_faces(hbya, ainv, hr, ar, ubi, ube, pbi, pbe)
            assert np.array_equal(hn.numpy(), ref_hn), (hr, ar)
            assert np.array_equal(af.numpy(), ref_af), (hr, ar)


@pytest.mark.parametrize("name", sorted(MESHES))
def test_correct_phi_and_velocity_bit_exact(ctx, name):
    raw = MESHES[name]()
    om = O.OracleMesh(raw)
    rng = np.random.default_rng(43)
    n, F, D = om.n_cells, om.n_faces, om.dim
    mesh = fvg.Mesh(ctx, raw)
    hn, af, p = rng.normal(size=F), rng.uniform(0.5, 2.0, F), rng.normal(size=n)
    grads = rng.normal(size=(n, D))
    _, _, pbi, pbe = _closures(om, rng)
    pbc = fvg.FaceConstraints(pbi, pbe)
    phi = fvg.DistributedVector(ctx, F)
    fvg.correct_phi(phi, hn, af, p, fvg.Corrected(grads.ravel(), 0.7), pbc, mesh)
    assert np.array_equal(phi.numpy(), om.correct_phi(hn, af, p, pbi, pbe, corrected_grads=grads, reduction=0.7))
    fvg.correct_phi(phi, hn, af, p, fvg.Orthogonal(), pbc, mesh)
    assert np.array_equal(phi.numpy(), om.correct_phi(hn, af, p, pbi, pbe))
    hbya, ainv, pg = rng.normal(size=(n, D)), rng.uniform(0.1, 1.0, n), rng.normal(size=(n, D))
    vel = fvg.DistributedVector(ctx, n * D)
    fvg.correct_velocity(vel, hbya.ravel(), ainv, pg.ravel(), mesh)
    assert np.array_equal(vel.numpy(), om.correct_velocity(hbya, ainv, pg).ravel())


def _channel_walls(raw):
    patch = raw["face_patch"][raw["face_patch"] >= 0]
    return [0, 2], (patch == 0) | (patch == 2)       # the y = 0 and y = 1 sides


def test_wall_distance_normalisation_only(ctx):
    """No rounds: the distance is the normalisation of the initial guess with zero
    gradients, (3/2 u)^(2/3)."""
    raw = MESHES["tri16"]()
    om = O.OracleMesh(raw)
    mesh = fvg.Mesh(ctx, raw)
    u0 = np.random.default_rng(44).uniform(0.0, 0.4, om.n_cells)
    d = fvg.DistributedVector.from_data(ctx, u0)
    solves, _ = fvg.compute_wall_distance(d, mesh, [0, 2], 0, 1e-3)
    ref, info = om.wall_distance(_channel_walls(raw)[1], 0, 1e-3, 1e-6, 1e-6, 1000, initial=u0)
    assert solves == 0 and info["solves"] == 0
    ulp = np.abs(d.numpy().view(np.int64) - ref.view(np.int64))
    assert ulp.max() <= 4


@pytest.mark.parametrize("name,rounds", [("tri16", 1), ("tri16", 3), ("tet4", 2)])
def test_wall_distance_rounds_against_the_oracle(ctx, name, rounds):
    """compute_wall_distance with walls on two opposite sides: each round's CG solve and the
    normalisation against the oracle's, at the solver bar."""
    raw = MESHES[name]()
    om = O.OracleMesh(raw)
    mesh = fvg.Mesh(ctx, raw)
    patch = raw["face_patch"][raw["face_patch"] >= 0]
    walls = [2, 3] if om.dim == 3 else [0, 2]            # y = 0 / y = 1 sides
    wall = np.isin(patch, walls)
    d = fvg.DistributedVector(ctx, om.n_cells)
    opts = fvg.LinearSolverOptions(1e-10, 0.0, 2000)
    solves, info = fvg.compute_wall_distance(d, mesh, walls, rounds, 1e-30, opts)
    ref, rinfo = om.wall_distance(wall, rounds, 1e-30, 1e-10, 0.0, 2000)
    assert solves == rounds == rinfo["solves"]
    assert abs(info.iterations - rinfo["iterations"]) <= 2
    assert np.linalg.norm(d.numpy() - ref) <= 1e-6 * np.linalg.norm(ref)


def test_wall_distance_default_patches_by_name(ctx):
    """wall_patches = None takes every patch whose name contains "wall" and not "slip"
    (wall_distance.rs:17-27)."""
    raw = dict(MESHES["tri16"]())
    raw["patch_names"] = ["bottom_wall", "outlet", "top_wall", "slip_wall"]
    mesh = fvg.Mesh(ctx, raw)
    a = fvg.DistributedVector(ctx, mesh.n_cells)
    b = fvg.DistributedVector(ctx, mesh.n_cells)
    fvg.compute_wall_distance(a, mesh, None, 2, 1e-30)
    fvg.compute_wall_distance(b, mesh, [0, 2], 2, 1e-30)
    assert np.array_equal(a.numpy(), b.numpy())
    raw.pop("patch_names")
    with pytest.raises(fvg.FvgError):
        fvg.compute_wall_distance(a, fvg.Mesh(ctx, raw), None, 1, 1e-3)
