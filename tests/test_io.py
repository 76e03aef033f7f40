"""Mesh files and results: the SU2 and native-format readers (src/core/mesh/io.rs) and the
PVTU writer (src/post/write_pvtu.rs), on the host.

The reference's own test (io.rs:567-609) reads square.su2, checks 121 nodes, 220 faces,
100 cells, a total volume of 1 and a closed boundary (both within 10 eps), writes the
mesh in its own format, reads it back and checks again; the same round trip runs here.
The SU2 reader is also held, array for array, to the fixture reader in tests/golden
(an independent restatement of read_su2 and compute_vtk_faces), on square.su2 and on
generated tetrahedral and hexahedral files.  The PVTU files are parsed back and checked
against a restatement of collect_vtu_mesh_data / write_polygons / write_polyhedra.
"""
import os
import sys
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import fvg
import meshgen
import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import su2  # noqa: E402

EPS = np.finfo(np.float64).eps
SQUARE = os.path.join(HERE, "golden", "square.su2")
KEYS = ("face_ptr", "face_nodes", "face_patch", "cell_ptr", "cell_faces")


def _same_topology(a, b):
    assert a["dim"] == b["dim"]
    assert np.array_equal(np.asarray(a["nodes"], np.float64).reshape(-1, a["dim"]),
                          np.asarray(b["nodes"], np.float64).reshape(-1, b["dim"]))
    for k in KEYS:
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k


def _square_invariants(raw):
    om = O.OracleMesh(raw)
    g = om.geometry()
    assert om.n_cells == 100 and om.n_faces == 220 and raw["nodes"].shape[0] == 121
    assert abs(g["cell_volume"].sum() - 1.0) < EPS * 10.0
    b = raw["face_patch"] >= 0
    sf = (g["face_normal"][b] * g["face_area"][b, None]).sum(0)
    assert np.linalg.norm(sf) < EPS * 10.0


def _rust_f64(v):
    """Rust's `{}` for f64: shortest round-trip digits, never in exponent form."""
    r = repr(float(v))
    if "e" not in r and "E" not in r:
        return r[:-2] if r.endswith(".0") else r
    from decimal import Decimal
    s = format(Decimal(r), "f")
    return s.rstrip("0").rstrip(".") if "." in s else s


def _write_native(raw, path, names):
    """Mesh::write (io.rs:10-64) of a serial mesh: every face and cell owned ("-")."""
    d = raw["dim"]
    fp = raw["face_patch"]
    lines = [f"NDIME={d}", f"NPATCH={len(names)}"]
    for pid, name in enumerate(names):
        idx = np.nonzero(fp == pid)[0]
        lines.append(f"{name} {pid} {int(idx[0]) if len(idx) else 0} {len(idx)}")
    lines.append(f"NNODES={raw['nodes'].shape[0]}")
    lines += [" ".join(_rust_f64(x) for x in p) for p in raw["nodes"]]
    nf = raw["face_ptr"].shape[0] - 1
    lines.append(f"NFACES={nf}")
    for f in range(nf):
        nd = raw["face_nodes"][raw["face_ptr"][f]:raw["face_ptr"][f + 1]]
        tag = "-" if fp[f] < 0 else str(fp[f])
        lines.append(" ".join([str(len(nd))] + [str(x) for x in nd] + [tag, str(f), "-"]))
    nc = raw["cell_ptr"].shape[0] - 1
    lines.append(f"NCELLS={nc}")
    for c in range(nc):
        cf = raw["cell_faces"][raw["cell_ptr"][c]:raw["cell_ptr"][c + 1]]
        lines.append(" ".join([str(len(cf))] + [str(x) for x in cf] + [str(c), "-"]))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def test_square_su2_reference_round_trip(tmp_path):
    t = fvg.Topology.read_su2(SQUARE)
    raw = t.raw()
    assert (t.dim, t.n_nodes, t.n_faces, t.n_cells) == (2, 121, 220, 100)
    _same_topology(raw, su2.read(SQUARE))
    _square_invariants(raw)
    names = [t.patch_name(p) for p in range(t.n_patches)]
    assert names[0] == "PhysicalLine0"
    path = str(tmp_path / "square.mesh")
    _write_native(raw, path, names)
    back = fvg.Topology.read(path)
    _same_topology(back.raw(), raw)
    _square_invariants(back.raw())
    assert [back.patch_name(p) for p in range(back.n_patches)] == names


def _su2_text(dim, nodes, elems, markers):
    out = [f"NDIME= {dim}", f"NELEM= {len(elems)}"]
    out += [" ".join(str(x) for x in [k] + list(nd) + [i]) for i, (k, nd) in enumerate(elems)]
    out.append(f"NPOIN= {len(nodes)}")
    out += [" ".join(_rust_f64(x) for x in p) + f" {i}" for i, p in enumerate(nodes)]
    out.append(f"NMARK= {len(markers)}")
    for name, (kind, faces) in markers:
        out.append(f"MARKER_TAG= {name}")
        out.append(f"MARKER_ELEMS= {len(faces)}")
        out += [" ".join(str(x) for x in [kind] + list(f)) + " " for f in faces]
    return "\n".join(out) + "\n"


def _hex_su2(n, seed):
    rng = np.random.default_rng(seed)
    h = 1.0 / n
    K, J, I = np.meshgrid(np.arange(n + 1), np.arange(n + 1), np.arange(n + 1), indexing="ij")
    I, J, K = I.ravel(), J.ravel(), K.ravel()
    nodes = np.stack([I * h, J * h, K * h], axis=1)
    inner = (I > 0) & (I < n) & (J > 0) & (J < n) & (K > 0) & (K < n)
    nodes[inner] += rng.uniform(-0.15 * h, 0.15 * h, size=(inner.sum(), 3))
    nid = lambda i, j, k: i + (n + 1) * (j + (n + 1) * k)
    elems = []
    for k in range(n):
        for j in range(n):
            for i in range(n):
                elems.append((12, [nid(i, j, k), nid(i + 1, j, k), nid(i + 1, j + 1, k), nid(i, j + 1, k),
                                   nid(i, j, k + 1), nid(i + 1, j, k + 1), nid(i + 1, j + 1, k + 1),
                                   nid(i, j + 1, k + 1)]))
    sides = {"xmin": [], "xmax": [], "ymin": [], "ymax": [], "zmin": [], "zmax": []}
    for a in range(n):
        for b in range(n):
            sides["xmin"].append([nid(0, a, b), nid(0, a + 1, b), nid(0, a + 1, b + 1), nid(0, a, b + 1)])
            sides["xmax"].append([nid(n, a, b), nid(n, a + 1, b), nid(n, a + 1, b + 1), nid(n, a, b + 1)])
            sides["ymin"].append([nid(a, 0, b), nid(a + 1, 0, b), nid(a + 1, 0, b + 1), nid(a, 0, b + 1)])
            sides["ymax"].append([nid(a, n, b), nid(a + 1, n, b), nid(a + 1, n, b + 1), nid(a, n, b + 1)])
            sides["zmin"].append([nid(a, b, 0), nid(a + 1, b, 0), nid(a + 1, b + 1, 0), nid(a, b + 1, 0)])
            sides["zmax"].append([nid(a, b, n), nid(a + 1, b, n), nid(a + 1, b + 1, n), nid(a, b + 1, n)])
    return _su2_text(3, nodes, elems, [(k, (9, v)) for k, v in sides.items()])


def _tet_su2(raw):
    """The tetrahedra of a meshgen mesh as SU2 (nodes from the cells' faces), with the
    boundary triangles in one marker per patch."""
    fp, fn, fptr = raw["face_patch"], raw["face_nodes"], raw["face_ptr"]
    elems = []
    for c in range(raw["cell_ptr"].shape[0] - 1):
        nd = []
        for f in raw["cell_faces"][raw["cell_ptr"][c]:raw["cell_ptr"][c + 1]]:
            for x in fn[fptr[f]:fptr[f + 1]]:
                if x not in nd:
                    nd.append(int(x))
        elems.append((10, nd))
    markers = []
    for pid in range(fp.max() + 1):
        faces = [list(fn[fptr[f]:fptr[f + 1]]) for f in np.nonzero(fp == pid)[0]]
        markers.append((f"side{pid}", (5, faces)))
    return _su2_text(3, raw["nodes"], elems, markers)


@pytest.mark.parametrize("kind", ["tet", "hex"])
def test_su2_reader_3d_against_fixture_reader(tmp_path, kind):
    """3D files: the same arrays as the fixture reader.  (The crate scales 3D face normals
    and volumes, compute.rs:305 and :359-361, so the 2D volume and closure identities do
    not carry over; the tetrahedral file is held to the cell volumes of the mesh it was
    written from instead.)"""
    src = meshgen.tet3d(3, seed=2)
    text = _tet_su2(src) if kind == "tet" else _hex_su2(3, seed=5)
    path = str(tmp_path / f"{kind}.su2")
    with open(path, "w") as f:
        f.write(text)
    t = fvg.Topology.read_su2(path)
    raw = t.raw()
    _same_topology(raw, su2.read(path))
    assert t.n_patches == 6
    g = O.OracleMesh(raw).geometry()
    if kind == "tet":
        assert t.n_cells == 6 * 27 and t.n_faces == src["face_ptr"].shape[0] - 1
        ref = O.OracleMesh(src).geometry()["cell_volume"]
        assert np.allclose(g["cell_volume"], ref, rtol=1e-12, atol=0.0)
    else:
        assert t.n_cells == 27 and t.n_faces == 3 * 4 * 9
        assert (raw["face_patch"] >= 0).sum() == 6 * 9 and np.all(g["cell_volume"] > 0.0)


def test_reader_errors(tmp_path):
    p = tmp_path / "bad.su2"
    p.write_text("NDIME= 2\nNELEM= 1\n7 0 1 2 0\nNPOIN= 3\n0 0 0\n1 0 1\n0 1 2\n")
    with pytest.raises(fvg.FvgError, match="unsupported kind"):
        fvg.Topology.read_su2(str(p))
    q = tmp_path / "remote.mesh"
    q.write_text("NDIME=2\nNNODES=2\n0 0\n1 0\nNFACES=1\n2 0 1 - 0 3\nNCELLS=0\n")
    with pytest.raises(fvg.FvgError, match="FVG_E_UNSUPPORTED"):
        fvg.Topology.read(str(q))
    with pytest.raises(fvg.FvgError, match="cannot open"):
        fvg.Topology.read_su2(str(tmp_path / "missing.su2"))


def _vtu_restatement(raw):
    """collect_vtu_mesh_data (write_pvtu.rs:161-204) in Python: per cell its faces'
    local node ids, nodes numbered in order of first appearance."""
    local, nodes, cells = {}, [], []
    for c in range(raw["cell_ptr"].shape[0] - 1):
        faces = []
        for f in raw["cell_faces"][raw["cell_ptr"][c]:raw["cell_ptr"][c + 1]]:
            ids = []
            for x in raw["face_nodes"][raw["face_ptr"][f]:raw["face_ptr"][f + 1]]:
                if int(x) not in local:
                    local[int(x)] = len(nodes)
                    nodes.append(raw["nodes"][x])
                ids.append(local[int(x)])
            faces.append(ids)
        cells.append(faces)
    return np.array(nodes), cells


def _arrays(piece_xml, tag):
    out = {}
    for da in piece_xml.iter("DataArray"):
        out[da.get("Name")] = (da.get("type"), da.get("NumberOfComponents"), da.text.split())
    return out


@pytest.mark.parametrize("which", ["square", "tet"])
def test_pvtu_writer(tmp_path, which):
    raw = su2.read(SQUARE) if which == "square" else meshgen.tet3d(2, seed=4)
    t = fvg.Topology.from_raw(raw)
    n = t.n_cells
    rng = np.random.default_rng(1)
    u = rng.normal(size=n)
    v = rng.normal(size=(n, raw["dim"]))
    w = fvg.PvtuWriter(t).with_("u", u).with_("vel", v)
    os.makedirs(tmp_path / "out", exist_ok=True)
    w.write(str(tmp_path / "out" / "res.pvtu"))
    parent = (tmp_path / "out" / "res.pvtu").read_text()
    want = ('<VTKFile type="PUnstructuredGrid">\n  <PUnstructuredGrid GhostLevel="0">\n    <PCellData>\n'
            '      <PDataArray Name="u" type="Float32"/>\n'
            f'      <PDataArray Name="vel" type="Float32" NumberOfComponents="{raw["dim"]}"/>\n'
            '    </PCellData>\n    <PPoints>\n      <PDataArray type="Float32" NumberOfComponents="3"/>\n'
            '    </PPoints>\n    <Piece Source="res_0.vtu"/>\n  </PUnstructuredGrid>\n</VTKFile>\n')
    assert parent == want
    root = ET.parse(str(tmp_path / "out" / "res_0.vtu")).getroot()
    piece = root.find("UnstructuredGrid/Piece")
    nodes, cells = _vtu_restatement(raw)
    assert int(piece.get("NumberOfPoints")) == len(nodes) and int(piece.get("NumberOfCells")) == n
    A = _arrays(piece, "DataArray")
    pts = np.array([float(x) for x in A["Points"][2]]).reshape(-1, 3)
    assert np.array_equal(pts[:, :raw["dim"]], nodes)
    assert np.all(pts[:, raw["dim"]:] == 0.0)
    conn = [int(x) for x in A["connectivity"][2]]
    offs = [int(x) for x in A["offsets"][2]]
    types = set(A["types"][2])
    if raw["dim"] == 2:
        assert types == {"7"}
        # write_polygons: the node each face shares with the next face is skipped
        want_conn = []
        for faces in cells:
            for fi in range(len(faces)):
                a, b = faces[fi], faces[(fi + 1) % len(faces)]
                want_conn.append(a[1] if a[0] in b else a[0])
        assert conn == want_conn
        assert offs == list(np.cumsum([len(f) for f in cells]))
    else:
        assert types == {"42"}
        stream = [int(x) for x in A["faces"][2]]
        want_stream, fo = [], []
        for faces in cells:
            want_stream.append(len(faces))
            for f in faces:
                want_stream += [len(f)] + f
            fo.append(len(want_stream))
        assert stream == want_stream
        assert [int(x) for x in A["faceoffsets"][2]] == fo
        start = 0
        for c, faces in enumerate(cells):
            got = conn[start:offs[c]]
            assert sorted(got) == sorted({x for f in faces for x in f})
            start = offs[c]
    assert np.array_equal(np.array([float(x) for x in A["u"][2]]), u)
    assert A["vel"][1] == str(raw["dim"])
    assert np.array_equal(np.array([float(x) for x in A["vel"][2]]).reshape(n, -1), v)
    assert all(x == _rust_f64(float(x)) for x in A["u"][2][:50])


def test_pvtu_file_naming_quirks(tmp_path):
    """file_without_extension / file_without_directory scan from the end and never look at
    the first character (write_pvtu.rs:14-24, 416-424)."""
    t = fvg.Topology.from_raw(su2.read(SQUARE))
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        fvg.PvtuWriter(t).write("a.b.pvtu")
        assert os.path.exists("a.b.pvtu") and os.path.exists("a.b_0.vtu")
        assert 'Source="a.b_0.vtu"' in open("a.b.pvtu").read()
    finally:
        os.chdir(cwd)
