"""Minimal SU2 reader producing the raw mesh form of meshgen (test fixture loader).

Faces are numbered as the crate's reader numbers them (src/core/mesh/io.rs:364-486):
interior faces in order of first appearance, then boundary faces grouped by marker
(markers numbered 0.. in file order, io.rs:300-360), each in first-appearance order.
Element face lists follow the VTK conventions the crate uses (src/core/mesh/vtk.rs).
"""
import numpy as np

_FACES = {
    3: ((0,), (1,)),                                   # line
    5: ((0, 1), (1, 2), (2, 0)),                       # triangle
    9: ((0, 1), (1, 2), (2, 3), (3, 0)),               # quad
    10: ((0, 2, 1), (0, 1, 3), (1, 2, 3), (0, 3, 2)),  # tetrahedron
    12: ((0, 3, 2, 1), (4, 5, 6, 7), (0, 1, 5, 4), (1, 2, 6, 5), (2, 3, 7, 6), (3, 0, 4, 7)),   # hexahedron
}
_NODES = {3: 2, 5: 3, 9: 4, 10: 4, 12: 8}


def read(path):
    lines = [l.strip() for l in open(path) if l.strip()]
    i = 0
    dim = None
    elems, nodes, markers = [], None, []
    while i < len(lines):
        l = lines[i]
        if l.startswith("NDIME="):
            dim = int(l.split("=")[1])
        elif l.startswith("NELEM="):
            ne = int(l.split("=")[1])
            for k in range(ne):
                v = [int(t) for t in lines[i + 1 + k].split()]
                elems.append((v[0], v[1:1 + _NODES[v[0]]]))
            i += ne
        elif l.startswith("NPOIN="):
            npt = int(l.split("=")[1].split()[0])
            nodes = np.array([[float(t) for t in lines[i + 1 + k].split()[:dim]] for k in range(npt)])
            i += npt
        elif l.startswith("MARKER_ELEMS="):
            nm = int(l.split("=")[1])
            mk = len(markers)
            keys = set()
            for k in range(nm):
                v = [int(t) for t in lines[i + 1 + k].split()]
                keys.add(tuple(sorted(v[1:])))
            markers.append(keys)
            i += nm
        i += 1
    face_nodes, face_key, cell_faces = [], {}, []
    for kind, nd in elems:
        cf = []
        for loc in _FACES[kind]:
            fn = tuple(nd[j] for j in loc)
            key = tuple(sorted(fn))
            if key not in face_key:
                face_key[key] = len(face_nodes)
                face_nodes.append(fn)
            cf.append(face_key[key])
        cell_faces.append(cf)
    patch_of = {}
    for mid, keys in enumerate(markers):
        for key in keys:
            patch_of[key] = mid
    first = list(range(len(face_nodes)))
    patches = [patch_of.get(tuple(sorted(fn)), -1) for fn in face_nodes]
    interior = [f for f in first if patches[f] < 0]
    boundary = sorted([f for f in first if patches[f] >= 0], key=lambda f: (patches[f], f))
    order = interior + boundary
    final = {f: q for q, f in enumerate(order)}
    fptr = np.zeros(len(order) + 1, np.int64)
    fptr[1:] = np.cumsum([len(face_nodes[f]) for f in order])
    fnodes = np.array([n for f in order for n in face_nodes[f]], np.int32)
    fpatch = np.array([patches[f] for f in order], np.int32)
    cptr = np.zeros(len(cell_faces) + 1, np.int64)
    cptr[1:] = np.cumsum([len(c) for c in cell_faces])
    cfaces = np.array([final[f] for c in cell_faces for f in c], np.int32)
    return dict(dim=dim, nodes=nodes, face_ptr=fptr, face_nodes=fnodes, face_patch=fpatch, cell_ptr=cptr,
                cell_faces=cfaces, h=0.1, name="square.su2")
