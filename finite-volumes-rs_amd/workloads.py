"""The synthetic assemble-and-solve workloads, built the way a crate user builds them.

Each builder returns the raw mesh (meshgen form) and a ``fields(geom)`` function that
evaluates the equation's fields and boundary closure from Mesh::compute geometry --
exactly as user code written against the crate would (e.g. a flux ``u . face.normal()``
as in examples/ex4/ex4.rs:59-65, a Dirichlet value from ``face.center()``).

Equation programs are postfix lists of (opcode, arg0, arg1, arg2, scalar); see
include/fvg.h for the opcodes.

Workloads (sizes as in the benchmark definition):
  diffusion_2d     64x64 quads -> 8,192 triangles, jitter 0.25h, k = 1, f ~ U[0,1],
                   u = 0 on the boundary, CG + IC(0)
  diffusion_3d     64^3 hexes -> 1,572,864 Kuhn tets, jitter 0.2h, ln k ~ N(0, 1.5^2)
                   box-filtered over 3x3x3 hexes, u = x on the boundary, f ~ N(0,1), CG + IC(0)
  convdiff_prism   128x128x2 triangles extruded through 64 layers -> 2,097,152 prisms,
                   velocity (1,0.5,0.25) scaled per cell to Peclet U[1,100], k = 1e-2,
                   upwind, inflow u = 1 / outflow zero gradient, BiCGStab + ILU(0)
  batched_2d       256 independent jittered triangle meshes, cells log-uniform in
                   [1k, 64k], ln k ~ N(0, 1) per cell, f ~ N(0,1), u = 0, CG + IC(0)
  serving          Poisson stream over a pool of 2D triangle (1k-64k cells, 70%) and 3D
                   tet (50k-500k cells, 30%) meshes; diffusion (CG + IC(0)) or
                   convection-diffusion with cell Peclet U[1,50] (BiCGStab + ILU(0))
"""
from __future__ import annotations

import numpy as np

import meshgen

LAP, CONV, SRC, LIN, TIME, DIV, ADD, SUB, NEG = 1, 2, 3, 4, 5, 6, 16, 17, 18

# -laplacian(Orthogonal, &k) - source(&f)      (face field 0 = k, cell field 0 = f)
DIFFUSION_PROGRAM = [(LAP, 0, 0, 0, 0.0), (NEG, 0, 0, 0, 0.0), (SRC, 0, 0, 0, 0.0), (SUB, 0, 0, 0, 0.0)]
# convection(Upwind(&F), &F) - laplacian(Orthogonal, &k)   (face fields 0 = F, 1 = k)
CONVDIFF_PROGRAM = [(CONV, 0, 0, 0, 0.0), (LAP, 1, 0, 0, 0.0), (SUB, 0, 0, 0, 0.0)]


def diffusion_fields(raw, geom, kc, f, bc="zero"):
    """Face k = harmonic mean of the cells (owner value on the boundary); Dirichlet
    u = 0 (bc='zero') or u = x_f (bc='x') as (bndi, bnde) = (0, g)."""
    kf = meshgen.harmonic_face_values(kc, geom["owner"], geom["neighbor"])
    bnd = np.nonzero(raw["face_patch"] >= 0)[0]
    bci = np.zeros(bnd.size)
    bce = np.zeros(bnd.size) if bc == "zero" else geom["face_center"][bnd, 0].copy()
    return dict(prog=DIFFUSION_PROGRAM, face_fields=[kf], cell_fields=[np.asarray(f, np.float64)], bc_i=bci,
                bc_e=bce, solver="cg", prec="ic")


def convdiff_fields(raw, geom, kdiff=1e-2, velocity=(1.0, 0.5, 0.25), pe=(1.0, 100.0), seed=0):
    """Per-cell velocity scaled to |u| h / k ~ U[pe]; face flux = mean cell velocity . face.normal();
    inflow faces (flux < 0) Dirichlet u = 1 -> (0, 1), others zero gradient -> (1, 0) (ex4.rs:76-86)."""
    n = raw["cell_ptr"].shape[0] - 1
    dim = raw["dim"]
    rng = np.random.default_rng(seed)
    u = np.asarray(velocity[:dim], dtype=np.float64)
    u = u / np.linalg.norm(u)
    speed = rng.uniform(pe[0], pe[1], n) * kdiff / raw["h"]
    uc = speed[:, None] * u[None, :]
    own, nb = geom["owner"], geom["neighbor"]
    uf = uc[own].copy()
    it = nb >= 0
    uf[it] = 0.5 * (uc[own[it]] + uc[nb[it]])
    flux = (uf * geom["face_normal"]).sum(axis=1)
    kf = np.full(own.shape[0], kdiff)
    bnd = np.nonzero(raw["face_patch"] >= 0)[0]
    inflow = flux[bnd] < 0.0
    return dict(prog=CONVDIFF_PROGRAM, face_fields=[flux, kf], cell_fields=[], bc_i=np.where(inflow, 0.0, 1.0),
                bc_e=np.where(inflow, 1.0, 0.0), solver="bicgstab", prec="ilu")


def diffusion_2d(nx=64, seed=1):
    raw = meshgen.tri2d(nx, jitter=0.25, seed=seed)
    n = raw["cell_ptr"].shape[0] - 1
    f = np.random.default_rng(seed + 1000).uniform(0.0, 1.0, n)

    def fields(geom):
        return diffusion_fields(raw, geom, np.ones(n), f, bc="zero")

    return raw, fields


def diffusion_3d(nx=64, seed=1, sigma=1.5):
    raw = meshgen.tet3d(nx, jitter=0.2, seed=seed)
    n = raw["cell_ptr"].shape[0] - 1
    kc = meshgen.lognormal_hex_boxfiltered(nx, nx, nx, sigma, seed + 2000)
    f = np.random.default_rng(seed + 1000).standard_normal(n)

    def fields(geom):
        return diffusion_fields(raw, geom, kc, f, bc="x")

    return raw, fields


def convdiff_prism(nx=128, layers=64, seed=1, pe=(1.0, 100.0)):
    raw = meshgen.prism3d(nx, layers=layers, jitter=0.25, seed=seed)

    def fields(geom):
        return convdiff_fields(raw, geom, pe=pe, seed=seed + 3000)

    return raw, fields


def batched_2d(n_inst=256, seed=1, lo=1000, hi=64000, sigma=1.0):
    """Independent 2D diffusion instances with cell counts log-uniform in [lo, hi]:
    jittered triangle meshes (2 nx^2 cells with nx = round(sqrt(n/2))), per-cell log-normal
    k with the given sigma, f ~ N(0,1) per cell (each instance its own right-hand side),
    u = 0 on the boundary, CG + IC(0).  Returns a list of (raw mesh, fields) pairs."""
    rng = np.random.default_rng(seed)
    targets = np.exp(rng.uniform(np.log(lo), np.log(hi), n_inst))
    out = []
    for k, t in enumerate(targets):
        nx = max(2, int(round(np.sqrt(t / 2.0))))
        s = seed * 100003 + k
        raw = meshgen.tri2d(nx, jitter=0.25, seed=s)
        n = raw["cell_ptr"].shape[0] - 1
        kc = meshgen.lognormal_cells(n, sigma, s + 1)
        f = np.random.default_rng(s + 2).standard_normal(n)

        def fields(geom, raw=raw, kc=kc, f=f):
            return diffusion_fields(raw, geom, kc, f, bc="zero")

        out.append((raw, fields))
    return out


def serving_pool(n2d=12, n3d=4, seed=1, tri=(1000, 64000), tet=(50000, 500000)):
    """Distinct meshes for the serving stream: n2d jittered triangle meshes and n3d Kuhn
    tetrahedral meshes with cell counts log-uniform in the given ranges.  Returns
    {key: raw mesh}; keys start with 'tri' or 'tet'."""
    rng = np.random.default_rng(seed)
    pool = {}
    for k, t in enumerate(np.exp(rng.uniform(np.log(tri[0]), np.log(tri[1]), n2d))):
        nx = max(2, int(round(np.sqrt(t / 2.0))))
        pool[f"tri{k}_{2 * nx * nx}"] = meshgen.tri2d(nx, jitter=0.25, seed=seed * 7919 + k)
    for k, t in enumerate(np.exp(rng.uniform(np.log(tet[0]), np.log(tet[1]), n3d))):
        nx = max(2, int(round((t / 6.0) ** (1.0 / 3.0))))
        pool[f"tet{k}_{6 * nx ** 3}"] = meshgen.tet3d(nx, jitter=0.2, seed=seed * 104729 + k)
    return pool


def serving_requests(pool, geometry_of, n, qps, frac_3d=0.3, frac_conv=0.3, seed=2, pe=(1.0, 50.0)):
    """n requests with Poisson arrivals at rate qps over the pool's meshes: a 3D mesh with
    probability frac_3d (else 2D), convection-diffusion (BiCGStab + ILU(0), Peclet U[pe])
    with probability frac_conv, else diffusion (CG + IC(0)) with ln k ~ N(0, 1) per cell
    and f ~ N(0,1).  geometry_of(key) returns the mesh's Mesh::compute geometry.
    Returns a list of serving.Request."""
    from serving import Request, poisson_arrivals
    rng = np.random.default_rng(seed)
    keys2 = [k for k in pool if k.startswith("tri")]
    keys3 = [k for k in pool if k.startswith("tet")]
    arrivals = poisson_arrivals(n, qps, seed + 1)
    out = []
    for i in range(n):
        keys = keys3 if (keys3 and rng.uniform() < frac_3d) else keys2
        key = keys[int(rng.integers(len(keys)))]
        raw = pool[key]
        geom = geometry_of(key)
        nc = raw["cell_ptr"].shape[0] - 1
        s = int(rng.integers(1 << 30))
        if rng.uniform() < frac_conv:
            f = convdiff_fields(raw, geom, pe=pe, seed=s)
            kind = "bicgstab"
        else:
            kc = meshgen.lognormal_cells(nc, 1.0, s)
            f = diffusion_fields(raw, geom, kc, np.random.default_rng(s + 1).standard_normal(nc), bc="zero")
            kind = "cg"
        out.append(Request(i, key, kind, f["prog"], f["face_fields"], f["cell_fields"], f["bc_i"], f["bc_e"],
                           arrival=float(arrivals[i])))
    return out


def serving_fresh_requests(geometry_of_raw, n, qps, frac_3d=0.3, frac_conv=0.3, seed=3, pe=(1.0, 50.0),
                           tri=(1000, 64000), tet=(50000, 500000)):
    """n requests with Poisson arrivals at rate qps, each on a mesh of its own: a jittered
    Kuhn tetrahedral mesh with probability frac_3d, else a jittered triangle mesh, cell
    counts log-uniform in the given ranges; fields as serving_requests.  The mesh travels
    with the request (Request.raw), so the server builds it, and analyses the factor's
    pattern, inside the request's latency.  geometry_of_raw(raw) returns the
    Mesh::compute geometry the client uses to make the fields."""
    from serving import Request, poisson_arrivals
    rng = np.random.default_rng(seed)
    arrivals = poisson_arrivals(n, qps, seed + 1)
    out = []
    for i in range(n):
        s = int(rng.integers(1 << 30))
        if rng.uniform() < frac_3d:
            t = np.exp(rng.uniform(np.log(tet[0]), np.log(tet[1])))
            raw = meshgen.tet3d(max(2, int(round((t / 6.0) ** (1.0 / 3.0)))), jitter=0.2, seed=s)
        else:
            t = np.exp(rng.uniform(np.log(tri[0]), np.log(tri[1])))
            raw = meshgen.tri2d(max(2, int(round(np.sqrt(t / 2.0)))), jitter=0.25, seed=s)
        geom = geometry_of_raw(raw)
        nc = raw["cell_ptr"].shape[0] - 1
        if rng.uniform() < frac_conv:
            f = convdiff_fields(raw, geom, pe=pe, seed=s + 2)
            kind = "bicgstab"
        else:
            kc = meshgen.lognormal_cells(nc, 1.0, s + 3)
            f = diffusion_fields(raw, geom, kc, np.random.default_rng(s + 4).standard_normal(nc), bc="zero")
            kind = "cg"
        out.append(Request(i, f"fresh{seed}_{i}", kind, f["prog"], f["face_fields"], f["cell_fields"], f["bc_i"],
                           f["bc_e"], arrival=float(arrivals[i]), raw=raw))
    return out


WORKLOADS = {"diffusion_2d": diffusion_2d, "diffusion_3d": diffusion_3d, "convdiff_prism": convdiff_prism}
