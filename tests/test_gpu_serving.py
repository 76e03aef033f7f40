"""Serving loop on the GPU: continuous batching over a mixed stream.

Bars: every answer reaches a true relative residual <= 1e-8 (mixed precision, FP64
refinement) and is within 1e-6 relative of the FP64 oracle's solution; each request gives
the same bits alone and in the mix; the HBM budget is never exceeded.
"""
import copy

import numpy as np
import pytest
import scipy.sparse as sp

import meshgen
import oracle as O
import workloads
from serving import Server, latency_summary

pytestmark = pytest.mark.gpu


def _pool():
    return {"tri0_800": meshgen.tri2d(20, seed=11), "tri1_2450": meshgen.tri2d(35, seed=12),
            "tri2_288": meshgen.tri2d(12, seed=13), "tet0_1296": meshgen.tet3d(6, seed=14)}


def _oracle_system(raw, r):
    om = O.OracleMesh(raw)
    return om.assemble(r.prog, r.face_fields, r.cell_fields, r.bc_i, r.bc_e)


def _server(budget, lanes, max_batch):
    pool = _pool()
    srv = Server(budget_bytes=budget, lanes=lanes, max_batch=max_batch)
    for k, raw in pool.items():
        srv.register_mesh(k, raw)
    return srv, pool


def _requests(srv, pool, n, seed):
    return workloads.serving_requests(pool, lambda k: srv.meshes[k].mesh.geometry(), n, qps=3000.0, frac_3d=0.25,
                                      frac_conv=0.4, seed=seed)


def _serve(budget, lanes, max_batch, n, seed):
    srv, pool = _server(budget, lanes, max_batch)
    reqs = _requests(srv, pool, n, seed)
    wall = srv.run(reqs)
    return srv, pool, reqs, wall


def test_serving_answers_are_accurate_and_the_same_alone_and_in_the_mix():
    srv, pool, reqs, wall = _serve(2 * 10**9, lanes=2, max_batch=6, n=18, seed=5)
    try:
        assert all(r.done >= r.arrival for r in reqs)
        assert max(r.batch_size for r in reqs) > 1          # requests really were batched together
        for r in reqs:
            rp, ci, va, b = _oracle_system(pool[r.mesh], r)
            n = rp.shape[0] - 1
            S = sp.csr_matrix((va, ci, rp), shape=(n, n))
            rel = np.linalg.norm(b - S @ r.x) / np.linalg.norm(b)
            assert rel <= 1e-8, (r.rid, rel)
        for r in reqs[:6]:
            rp, ci, va, b = _oracle_system(pool[r.mesh], r)
            P = O.OraclePrec("ic" if r.kind == "cg" else "ilu", rp, ci, va, 0)
            xo, _ = O.solve(r.kind, rp, ci, va, P, b, rtol=1e-8, atol=0.0, max_iter=5000)
            assert np.linalg.norm(r.x - xo) <= 1e-6 * np.linalg.norm(xo)
        # the same requests one at a time
        for r in reqs:
            alone = copy.copy(r)
            alone.arrival, alone.done, alone.x = 0.0, -1.0, None
            srv.run([alone])
            assert alone.batch_size == 1
            assert np.array_equal(alone.x, r.x), r.rid
        s = latency_summary(reqs, wall)
        assert s["count"] == len(reqs) and s["p99_ms"] >= s["p50_ms"] > 0.0
    finally:
        srv.close()


def test_serving_stays_within_the_hbm_budget():
    srv, pool = _server(2 * 10**9, lanes=2, max_batch=4)
    try:
        # learn every slot's footprint, then leave room for about two slots only:
        # slots are evicted and rebuilt, and nothing exceeds the budget
        srv.run(_requests(srv, pool, 24, seed=3))
        biggest = max(e.slot_bytes for e in srv.meshes.values() if e.slot_bytes)
        srv.drop_idle_slots()
        srv.budget = srv.ctx.bytes_in_use() + int(2.2 * biggest)
        srv.peak_bytes = 0
        reqs = _requests(srv, pool, 14, seed=9)
        srv.run(reqs)
        assert all(r.done >= 0 for r in reqs)
        assert srv.peak_bytes <= srv.budget
        assert srv.ctx.bytes_in_use() <= srv.budget
        assert srv.evictions > 0
        for r in reqs:
            rp, ci, va, b = _oracle_system(pool[r.mesh], r)
            n = rp.shape[0] - 1
            S = sp.csr_matrix((va, ci, rp), shape=(n, n))
            assert np.linalg.norm(b - S @ r.x) / np.linalg.norm(b) <= 1e-8
    finally:
        srv.close()


def test_mixed_bicgstab_recovers_from_an_fp32_breakdown():
    """A convection-diffusion request of the serving stream whose FP32 inner BiCGStab breaks
    down (non-finite step) is still answered to a 1e-8 true residual: the inner solve stops
    before the bad step and the FP64 refinement restarts it, on the batched path and on the
    single-system path."""
    import fvg
    pool = workloads.serving_pool()
    ctx = fvg.Context(0)
    meshes = {}

    def geo(k):
        if k not in meshes:
            meshes[k] = fvg.Mesh(ctx, pool[k])
        return meshes[k].geometry()

    reqs = workloads.serving_requests(pool, geo, 600, 80.0, seed=10)
    r = [q for q in reqs if q.mesh.startswith("tri1_") and q.kind == "bicgstab"][1]
    rp, ci, va, b = _oracle_system(pool[r.mesh], r)
    n = rp.shape[0] - 1
    S = sp.csr_matrix((va, ci, rp), shape=(n, n))
    mesh = meshes[r.mesh]
    lhs, rhs = fvg.assemble_program(mesh, r.prog, r.face_fields, r.cell_fields, r.bc_i, r.bc_e)
    M = fvg.IncompleteLowerUpper.from_matrix(lhs, 0)
    opts = fvg.LinearSolverOptions(1e-8, 0.0, 5000, precision=fvg.PRECISION_MIXED)
    x = fvg.DistributedVector(ctx, n)
    batch = fvg.Batch(ctx, [lhs], [M], "bicgstab", opts)
    info = batch.solve([rhs], [x])[0]
    xb = x.numpy()
    assert np.all(np.isfinite(xb)) and info.iterations < 5000
    assert np.linalg.norm(b - S @ xb) / np.linalg.norm(b) <= 1e-8
    x.fill(0.0)
    info = fvg.bi_conjugate_gradient_stab(x, lhs, rhs, M, opts)
    xs = x.numpy()
    assert np.all(np.isfinite(xs))
    assert np.linalg.norm(b - S @ xs) / np.linalg.norm(b) <= 1e-8
    batch.close()


def test_requests_on_their_own_meshes():
    """Requests that carry their own mesh: the server builds each mesh when it places the
    request (inside its latency), answers to the accuracy bar, drops the mesh and its slot
    once the request is answered (nothing of theirs stays resident), and gives the same bits
    as the same request on a registered mesh."""
    srv = Server(budget_bytes=2 * 10**9, lanes=2, max_batch=4)
    try:
        base = srv.ctx.bytes_in_use()
        reqs = workloads.serving_fresh_requests(lambda raw: O.OracleMesh(raw).geometry(), 10, qps=2000.0,
                                                frac_3d=0.3, frac_conv=0.4, seed=21, tri=(200, 2000),
                                                tet=(1296, 6000))
        srv.run(reqs)
        assert srv.stats["meshes_built"] == len(reqs)
        assert not any(e.transient for e in srv.meshes.values())
        assert not srv.slots and srv.ctx.bytes_in_use() <= base + (1 << 20)
        for r in reqs:
            rp, ci, va, b = _oracle_system(r.raw, r)
            n = rp.shape[0] - 1
            S = sp.csr_matrix((va, ci, rp), shape=(n, n))
            assert np.linalg.norm(b - S @ r.x) / np.linalg.norm(b) <= 1e-8, r.rid
        # the same request on a registered mesh: the same bits
        r = reqs[0]
        srv.register_mesh("kept", r.raw)
        again = copy.copy(r)
        again.mesh, again.raw, again.arrival, again.done, again.x = "kept", None, 0.0, -1.0, None
        srv.run([again])
        assert np.array_equal(again.x, r.x)
    finally:
        srv.close()


def test_slots_skip_the_device_wide_sweep_layout():
    """A slot is only ever solved in batches, so its preconditioner is made without the
    blocked layout of device-wide sweeps (most of its set-up time), the context's tuning
    is left as it was, and the answers meet the accuracy bar."""
    import fvg
    srv, pool = _server(2 * 10**9, lanes=2, max_batch=4)
    try:
        srv.ctx.set_tuning(fvg.TUNE_SWEEP_BLOCKS, 0)
        reqs = _requests(srv, pool, 8, seed=4)
        srv.run(reqs)
        assert srv.ctx.get_tuning(fvg.TUNE_SWEEP_BLOCKS) == 0
        slots = list(srv.slots)
        assert slots
        for s in slots:
            lay = s.M.sweep_layout()
            assert lay["lower_blocks"] == 0 and lay["upper_blocks"] == 0, lay
        for r in reqs:
            rp, ci, va, b = _oracle_system(pool[r.mesh], r)
            n = rp.shape[0] - 1
            S = sp.csr_matrix((va, ci, rp), shape=(n, n))
            assert np.linalg.norm(b - S @ r.x) / np.linalg.norm(b) <= 1e-8, r.rid
    finally:
        srv.close()
