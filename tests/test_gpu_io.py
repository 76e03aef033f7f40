"""GPU side of the mesh files and results: a device mesh built from a topology read from
SU2, and the asynchronous PVTU write (device fields copied out on the context stream,
files written by a host thread) giving the bytes of the synchronous write."""
import os
import sys

import numpy as np
import pytest

import fvg
import oracle as O

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
SQUARE = os.path.join(HERE, "golden", "square.su2")


def test_mesh_from_su2_topology_bit_exact(ctx):
    t = fvg.Topology.read_su2(SQUARE)
    raw = t.raw()
    mesh = fvg.Mesh.from_topology(ctx, t)
    g = mesh.geometry()
    ref = O.OracleMesh(raw).geometry()
    for k in ("face_area", "face_center", "face_normal", "owner", "neighbor", "cell_volume", "cell_center"):
        assert np.array_equal(g[k], ref[k]), k


def test_async_pvtu_write_matches_sync(ctx, tmp_path):
    t = fvg.Topology.read_su2(SQUARE)
    n = t.n_cells
    rng = np.random.default_rng(2)
    u, v = rng.normal(size=n), rng.normal(size=(n, 2))
    du = fvg.DistributedVector.from_data(ctx, u)
    dv = fvg.DistributedVector.from_data(ctx, v.ravel())
    fvg.PvtuWriter(t).with_("u", u).with_("vel", v).write(str(tmp_path / "sync.pvtu"))
    w = fvg.PvtuWriter(t).with_("u", du).with_("vel", dv)
    w.write_async(ctx, str(tmp_path / "async.pvtu"))
    # the GPU carries on while the host thread writes: overwrite the fields right away
    du.fill(0.0)
    dv.fill(0.0)
    w.wait()
    a = (tmp_path / "async_0.vtu").read_bytes()
    b = (tmp_path / "sync_0.vtu").read_bytes()
    assert a == b
    assert (tmp_path / "async.pvtu").read_text().replace("async_0", "sync_0") == (tmp_path / "sync.pvtu").read_text()
    # a second write through the same staging, then destroy without an explicit wait
    du.upload(u)
    w.write_async(ctx, str(tmp_path / "again.pvtu"))
    w.close()
    vtu = (tmp_path / "again_0.vtu").read_text()
    assert "NumberOfCells=\"100\"" in vtu
