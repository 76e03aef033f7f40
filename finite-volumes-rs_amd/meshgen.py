"""Seeded synthetic meshes and fields for the assemble-and-solve workloads.

A mesh is returned in the raw form the crate's builder consumes
(``Mesh::add_node / add_face / add_cell``, src/core/mesh/mod.rs:444-527):

* ``nodes``      float64 (n_nodes, dim)
* ``face_ptr``   int64 (n_faces + 1), ``face_nodes`` int32: node list of each face
* ``face_patch`` int32 (n_faces): boundary patch id, -1 for an interior face
* ``cell_ptr``   int64 (n_cells + 1), ``cell_faces`` int32: face list of each cell

Faces are numbered as the SU2 reader numbers them (src/core/mesh/io.rs:364-486):
interior faces first in order of first appearance, then boundary faces grouped
by patch, each in order of first appearance.  A face keeps the node order of the
cell that first names it.  Cells are numbered naturally: ``2*(i + nx*j) + t``
for triangles, ``6*hex + perm`` for Kuhn tetrahedra and ``tri + n_tri*layer``
for prisms.
"""
from __future__ import annotations

import numpy as np

# Kuhn split of a hex: one tet per axis permutation, all sharing the (0,0,0)-(1,1,1) diagonal.
_PERMS = ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0))
# local faces of a tet with vertices (P0, P1, P2, P3)
_TET_FACES = ((0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3))


def _faces_from_cells(local, n_local_per_cell, patch_of_key):
    """Deduplicate per-cell local faces.

    ``local``: int64 (n_cells * n_local_per_cell, K) node lists, K-padded with -1.
    ``patch_of_key``: callable(nodes (m, K)) -> patch ids for boundary faces.
    """
    m, K = local.shape
    key = np.sort(np.where(local < 0, np.iinfo(np.int64).max, local), axis=1)
    # unique rows; first occurrence index gives creation order
    kv = np.ascontiguousarray(key).view(np.dtype((np.void, key.dtype.itemsize * K))).ravel()
    _, first, inverse, counts = np.unique(kv, return_index=True, return_inverse=True, return_counts=True)
    if counts.max() > 2:
        raise ValueError("non-manifold mesh: a face is shared by more than two cells")
    interior = counts == 2
    patch = np.full(first.shape[0], -1, dtype=np.int32)
    bidx = np.nonzero(~interior)[0]
    patch[bidx] = patch_of_key(local[first[bidx]])
    # final order: interior by first appearance, then boundary by (patch, first appearance)
    int_ids = np.nonzero(interior)[0]
    order_int = int_ids[np.argsort(first[int_ids], kind="stable")]
    bnd_sorted = bidx[np.lexsort((first[bidx], patch[bidx]))]
    order = np.concatenate([order_int, bnd_sorted])
    final_of_unique = np.empty_like(order)
    final_of_unique[order] = np.arange(order.shape[0])
    face_local_nodes = local[first[order]]
    nnodes = (face_local_nodes >= 0).sum(axis=1)
    face_ptr = np.zeros(order.shape[0] + 1, dtype=np.int64)
    face_ptr[1:] = np.cumsum(nnodes)
    face_nodes = face_local_nodes[face_local_nodes >= 0].astype(np.int32)
    cell_faces = final_of_unique[inverse].astype(np.int32)
    n_cells = m // n_local_per_cell
    cell_ptr = np.arange(n_cells + 1, dtype=np.int64) * n_local_per_cell
    return dict(face_ptr=face_ptr, face_nodes=face_nodes, face_patch=patch[order].astype(np.int32),
                cell_ptr=cell_ptr, cell_faces=cell_faces)


def _jitter(rng, shape, amount):
    return rng.uniform(-amount, amount, size=shape)


def tri2d(nx, ny=None, jitter=0.25, seed=0):
    """Unit square, nx*ny quads each split along its (i,j)-(i+1,j+1) diagonal.

    Interior nodes are moved by U[-jitter*h, jitter*h] in each coordinate.
    Boundary patches: 0 bottom, 1 right, 2 top, 3 left.
    """
    ny = nx if ny is None else ny
    rng = np.random.default_rng(seed)
    hx, hy = 1.0 / nx, 1.0 / ny
    I, J = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), indexing="xy")
    I, J = I.ravel(), J.ravel()
    nodes = np.stack([I * hx, J * hy], axis=1).astype(np.float64)
    inner = (I > 0) & (I < nx) & (J > 0) & (J < ny)
    nodes[inner, 0] += _jitter(rng, inner.sum(), jitter * hx)
    nodes[inner, 1] += _jitter(rng, inner.sum(), jitter * hy)

    def nid(i, j):
        return i + (nx + 1) * j

    qi, qj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")
    qi, qj = qi.ravel(), qj.ravel()
    n00, n10, n11, n01 = nid(qi, qj), nid(qi + 1, qj), nid(qi + 1, qj + 1), nid(qi, qj + 1)
    tris = np.empty((2 * qi.size, 3), dtype=np.int64)
    tris[0::2] = np.stack([n00, n10, n11], axis=1)
    tris[1::2] = np.stack([n00, n11, n01], axis=1)
    local = np.stack([tris[:, [0, 1]], tris[:, [1, 2]], tris[:, [2, 0]]], axis=1).reshape(-1, 2)

    def patch_of(fn):
        i, j = fn % (nx + 1), fn // (nx + 1)
        p = np.full(fn.shape[0], -1, dtype=np.int32)
        p[(j == 0).all(axis=1)] = 0
        p[(i == nx).all(axis=1)] = 1
        p[(j == ny).all(axis=1)] = 2
        p[(i == 0).all(axis=1)] = 3
        return p

    mesh = _faces_from_cells(local, 3, patch_of)
    mesh.update(dim=2, nodes=nodes, h=min(hx, hy), name=f"tri2d_{nx}x{ny}")
    return mesh


def split_faces_2d(mesh, faces):
    """A copy of a 2D mesh with each listed interior face split at its midpoint into two
    faces between the same two cells (those cells gain a vertex on a straight edge), so
    that two faces update the same off-diagonal matrix slot."""
    nodes = mesh["nodes"]
    fp, fn = mesh["face_ptr"], mesh["face_nodes"]
    cp, cf = mesh["cell_ptr"], mesh["cell_faces"]
    face_list = [list(fn[fp[f]:fp[f + 1]]) for f in range(len(fp) - 1)]
    patches = list(mesh["face_patch"])
    cells = [list(cf[cp[c]:cp[c + 1]]) for c in range(len(cp) - 1)]
    extra = []
    for f in faces:
        if patches[f] >= 0 or len(face_list[f]) != 2:
            raise ValueError("only interior 2D faces can be split")
        a, b = face_list[f]
        mid = len(nodes) + len(extra)
        extra.append(0.5 * (nodes[a] + nodes[b]))
        face_list[f] = [a, mid]
        g = len(face_list)
        face_list.append([mid, b])
        patches.append(-1)
        for cl in cells:
            if f in cl:
                cl.append(g)
    out = dict(mesh)
    out.update(nodes=np.vstack([nodes] + [e[None] for e in extra]),
               face_ptr=np.arange(0, 2 * len(face_list) + 1, 2, dtype=np.int64),
               face_nodes=np.array(sum(face_list, []), dtype=np.int32),
               face_patch=np.array(patches, dtype=np.int32),
               cell_ptr=np.cumsum([0] + [len(c) for c in cells]).astype(np.int64),
               cell_faces=np.array(sum(cells, []), dtype=np.int32), name=mesh["name"] + "_split")
    return out


def tet3d(nx, ny=None, nz=None, jitter=0.2, seed=0):
    """Unit cube, nx*ny*nz hexes each split into 6 Kuhn tetrahedra.

    Interior nodes are moved by U[-jitter*h, jitter*h] per coordinate.
    Boundary patches: 0 x=0, 1 x=1, 2 y=0, 3 y=1, 4 z=0, 5 z=1.
    """
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    rng = np.random.default_rng(seed)
    h = np.array([1.0 / nx, 1.0 / ny, 1.0 / nz])
    K, J, I = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    I, J, K = I.ravel(), J.ravel(), K.ravel()
    nodes = np.stack([I * h[0], J * h[1], K * h[2]], axis=1).astype(np.float64)
    inner = (I > 0) & (I < nx) & (J > 0) & (J < ny) & (K > 0) & (K < nz)
    for d in range(3):
        nodes[inner, d] += _jitter(rng, inner.sum(), jitter * h[d])

    def nid(i, j, k):
        return i + (nx + 1) * (j + (ny + 1) * k)

    hk, hj, hi = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    hi, hj, hk = hi.ravel(), hj.ravel(), hk.ravel()
    nh = hi.size
    tets = np.empty((nh, 6, 4), dtype=np.int64)
    for t, perm in enumerate(_PERMS):
        c = np.zeros((3, nh), dtype=np.int64)
        verts = [nid(hi, hj, hk)]
        for ax in perm[:2]:
            c[ax] += 1
            verts.append(nid(hi + c[0], hj + c[1], hk + c[2]))
        verts.append(nid(hi + 1, hj + 1, hk + 1))
        tets[:, t, :] = np.stack(verts, axis=1)
    tets = tets.reshape(-1, 4)
    local = np.stack([tets[:, list(f)] for f in _TET_FACES], axis=1).reshape(-1, 3)

    def patch_of(fn):
        i = fn % (nx + 1)
        j = (fn // (nx + 1)) % (ny + 1)
        k = fn // ((nx + 1) * (ny + 1))
        p = np.full(fn.shape[0], -1, dtype=np.int32)
        for pid, cond in enumerate([i == 0, i == nx, j == 0, j == ny, k == 0, k == nz]):
            p[cond.all(axis=1)] = pid
        return p

    mesh = _faces_from_cells(local, 4, patch_of)
    mesh.update(dim=3, nodes=nodes, h=float(h.min()), name=f"tet3d_{nx}x{ny}x{nz}")
    return mesh


def prism3d(nx, ny=None, layers=None, jitter=0.25, seed=0):
    """Jittered tri2d(nx, ny) base extruded through ``layers`` uniform layers in z.

    Cells are numbered layer-major (tri + n_tri*layer).  Faces per prism: bottom
    and top triangles, then three side quads.  Boundary patches: 0 bottom(y=0),
    1 right(x=1), 2 top(y=1), 3 left(x=0), 4 z=0, 5 z=1.
    """
    ny = nx if ny is None else ny
    layers = nx if layers is None else layers
    base = tri2d(nx, ny, jitter=jitter, seed=seed)
    nb = base["nodes"].shape[0]
    dz = 1.0 / layers
    nodes = np.concatenate([np.concatenate([base["nodes"], np.full((nb, 1), l * dz)], axis=1)
                            for l in range(layers + 1)], axis=0)
    # recover base triangles from the tri2d numbering
    qi, qj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")
    qi, qj = qi.ravel(), qj.ravel()
    nid = lambda i, j: i + (nx + 1) * j
    n00, n10, n11, n01 = nid(qi, qj), nid(qi + 1, qj), nid(qi + 1, qj + 1), nid(qi, qj + 1)
    tris = np.empty((2 * qi.size, 3), dtype=np.int64)
    tris[0::2] = np.stack([n00, n10, n11], axis=1)
    tris[1::2] = np.stack([n00, n11, n01], axis=1)
    ntri = tris.shape[0]
    L = np.repeat(np.arange(layers), ntri)
    T = np.tile(tris, (layers, 1))
    a0, b0, c0 = T[:, 0] + nb * L, T[:, 1] + nb * L, T[:, 2] + nb * L
    a1, b1, c1 = a0 + nb, b0 + nb, c0 + nb
    pad = np.full_like(a0, -1)
    local = np.stack([
        np.stack([a0, b0, c0, pad], axis=1),
        np.stack([a1, b1, c1, pad], axis=1),
        np.stack([a0, b0, b1, a1], axis=1),
        np.stack([b0, c0, c1, b1], axis=1),
        np.stack([c0, a0, a1, c1], axis=1),
    ], axis=1).reshape(-1, 4)

    def patch_of(fn):
        valid = fn >= 0
        bn = np.where(valid, fn % nb, 0)
        lz = np.where(valid, fn // nb, 0)
        i, j = bn % (nx + 1), bn // (nx + 1)
        p = np.full(fn.shape[0], -1, dtype=np.int32)
        allv = lambda cond: (cond | ~valid).all(axis=1)
        p[allv(j == 0)] = 0
        p[allv(i == nx)] = 1
        p[allv(j == ny)] = 2
        p[allv(i == 0)] = 3
        p[allv(lz == 0)] = 4
        p[allv(lz == layers)] = 5
        return p

    mesh = _faces_from_cells(local, 5, patch_of)
    mesh.update(dim=3, nodes=nodes, h=min(1.0 / nx, 1.0 / ny, dz), name=f"prism3d_{nx}x{ny}x{layers}")
    return mesh


def mesh_sizes(mesh):
    fp = mesh["face_patch"]
    n_cells = mesh["cell_ptr"].shape[0] - 1
    f_int = int((fp < 0).sum())
    return dict(n=n_cells, f_int=f_int, f_bnd=int(fp.shape[0] - f_int), nnz=n_cells + 2 * f_int)


# ----------------------------------------------------------------------------------
# Fields for the workloads.  ``geom`` holds the mesh geometry as computed by
# Mesh::compute (face_area, face_center, face_normal, owner, neighbor, cell_volume,
# cell_center), so that user fields built from it match what a crate user gets.
# ----------------------------------------------------------------------------------

def harmonic_face_values(kc, owner, neighbor):
    """Face conductivity: harmonic mean of the two cells, owner value on boundary faces."""
    kf = kc[owner].copy()
    it = neighbor >= 0
    a, b = kc[owner[it]], kc[neighbor[it]]
    kf[it] = 2.0 * a * b / (a + b)
    return kf


def lognormal_cells(n, sigma, seed):
    rng = np.random.default_rng(seed)
    return np.exp(sigma * rng.standard_normal(n))


def lognormal_hex_boxfiltered(nx, ny, nz, sigma, seed):
    """ln k ~ N(0, sigma^2) per hex, correlated by a 3x3x3 box filter, 6 tets per hex."""
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((nz, ny, nx))
    pad = np.pad(g, 1, mode="constant")
    cnt = np.pad(np.ones_like(g), 1, mode="constant")
    s = np.zeros_like(g)
    c = np.zeros_like(g)
    for dk in range(3):
        for dj in range(3):
            for di in range(3):
                s += pad[dk:dk + nz, dj:dj + ny, di:di + nx]
                c += cnt[dk:dk + nz, dj:dj + ny, di:di + nx]
    sm = s / c
    sm = (sm - sm.mean()) / sm.std()
    return np.repeat(np.exp(sigma * sm.ravel()), 6)
