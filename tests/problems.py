"""Workload problems shared by the tests: a mesh, its fields, the equation and the
boundary closure, evaluated from Mesh::compute geometry the way a crate user would."""
import numpy as np

import meshgen

LAP, CONV, SRC, LIN, TIME, DIV, ADD, SUB, NEG = 1, 2, 3, 4, 5, 6, 16, 17, 18


def diffusion(mesh, geom, kc=None, f=None, seed=0, bc="zero"):
    """-laplacian(Orthogonal, k_f) - source(f); Dirichlet u=0 (bc='zero') or u=x (bc='x')."""
    n = mesh["cell_ptr"].shape[0] - 1
    rng = np.random.default_rng(seed)
    if kc is None:
        kc = np.ones(n)
    if f is None:
        f = rng.uniform(0.0, 1.0, n)
    kf = meshgen.harmonic_face_values(kc, geom["owner"], geom["neighbor"])
    bnd = np.nonzero(mesh["face_patch"] >= 0)[0]
    bci = np.zeros(bnd.size)
    bce = np.zeros(bnd.size) if bc == "zero" else geom["face_center"][bnd, 0].copy()
    prog = [(LAP, 0, 0, 0, 0.0), (NEG, 0, 0, 0, 0.0), (SRC, 0, 0, 0, 0.0), (SUB, 0, 0, 0, 0.0)]
    return dict(prog=prog, face_fields=[kf], cell_fields=[f], bc_i=bci, bc_e=bce, symmetric=True)


def convection_diffusion(mesh, geom, kdiff=1e-2, velocity=(1.0, 0.5, 0.25), pe=(1.0, 100.0), seed=0):
    """convection(Upwind(&F), &F) - laplacian(Orthogonal, k); inflow u=1, outflow zero gradient.

    Velocity is scaled per cell so that |u| h / k ~ U[pe]; the face flux is the mean of
    the two cells' velocity dotted with face.normal() (ex4.rs:59-65 style).
    """
    n = mesh["cell_ptr"].shape[0] - 1
    dim = mesh["dim"]
    rng = np.random.default_rng(seed)
    u = np.asarray(velocity[:dim], dtype=np.float64)
    u = u / np.linalg.norm(u)
    speed = rng.uniform(pe[0], pe[1], n) * kdiff / mesh["h"]
    uc = speed[:, None] * u[None, :]
    own, nb = geom["owner"], geom["neighbor"]
    uf = uc[own].copy()
    it = nb >= 0
    uf[it] = 0.5 * (uc[own[it]] + uc[nb[it]])
    flux = (uf * geom["face_normal"]).sum(axis=1)
    kf = np.full(geom["owner"].shape[0], kdiff)
    bnd = np.nonzero(mesh["face_patch"] >= 0)[0]
    inflow = flux[bnd] < 0.0
    bci = np.where(inflow, 0.0, 1.0)
    bce = np.where(inflow, 1.0, 0.0)
    prog = [(CONV, 0, 0, 0, 0.0), (LAP, 1, 0, 0, 0.0), (SUB, 0, 0, 0, 0.0)]
    return dict(prog=prog, face_fields=[flux, kf], cell_fields=[], bc_i=bci, bc_e=bce, symmetric=False)
