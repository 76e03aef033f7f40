"""Host logic of the serving loop (no GPU): arrivals, statistics and batch formation."""
import numpy as np

import fvg
import meshgen
import oracle as O
from serving import Request, form_group, latency_summary, poisson_arrivals, raw_cells, raw_nnz_bound


def _req(i, kind="cg", mesh="m", arrival=0.0):
    return Request(i, mesh, kind, [], [], [], np.zeros(0), np.zeros(0), arrival=arrival)


def test_poisson_arrivals_rate_and_order():
    t = poisson_arrivals(20000, 250.0, seed=3)
    assert np.all(np.diff(t) > 0)
    assert abs(t.size / t[-1] - 250.0) / 250.0 < 0.03   # mean rate
    gaps = np.diff(t)
    assert abs(gaps.std() / gaps.mean() - 1.0) < 0.05    # exponential gaps: cv = 1


def test_latency_summary_percentiles():
    reqs = []
    for i in range(100):
        r = _req(i, arrival=0.0)
        r.done = (i + 1) * 1e-3
        reqs.append(r)
    s = latency_summary(reqs, wall=2.0)
    assert s["count"] == 100
    assert abs(s["p50_ms"] - np.percentile(np.arange(1, 101), 50)) < 1e-9
    assert abs(s["p99_ms"] - np.percentile(np.arange(1, 101), 99)) < 1e-9
    assert s["throughput"] == 50.0
    r = _req(100)
    assert latency_summary([r], 1.0)["count"] == 0   # unfinished requests are not counted


def test_form_group_keeps_kind_order_capacity_and_placement():
    pending = [_req(0, "cg"), _req(1, "bicgstab"), _req(2, "cg", "big"), _req(3, "cg"), _req(4, "cg"), _req(5, "cg")]
    free = {0, 2, 3, 4, 5}
    g = form_group(pending, lambda r: r.rid in free, max_batch=3)
    assert [r.rid for r in g] == [0, 2, 3]               # first kind wins, arrival order, capacity 3
    g = form_group(pending, lambda r: r.rid in free, max_batch=10, allow_large=False, large_rows=100,
                   rows_of=lambda r: 1000 if r.mesh == "big" else 10)
    assert [r.rid for r in g] == [0, 3, 4, 5]            # large meshes left waiting
    g = form_group(pending, lambda r: r.rid == 1, max_batch=10)
    assert [r.rid for r in g] == [1]                      # an unplaceable head does not block the rest
    assert form_group(pending, lambda r: False, max_batch=4) == []


def test_raw_mesh_size_estimates_bound_the_pattern():
    """A request carrying its own mesh is admitted on estimates from its topology alone:
    the cell count is exact and the CSR entry bound is at least the oracle's pattern."""
    for raw in (meshgen.tri2d(9, seed=1), meshgen.tet3d(4, seed=2), meshgen.prism3d(5, layers=3, seed=3)):
        om = O.OracleMesh(raw)
        assert raw_cells(raw) == om.n_cells
        eq = -fvg.laplacian(fvg.Orthogonal(), np.ones(om.n_faces)) - fvg.source(np.ones(om.n_cells))
        ops, ff, cf = eq.compile()
        rp = om.assemble(ops, ff, cf, np.zeros(om.n_bnd), np.zeros(om.n_bnd))[0]
        assert raw_nnz_bound(raw) >= int(rp[-1])
