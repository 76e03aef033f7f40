"""Benchmark: assemble-and-solve throughput on one MI355X.

A step is one complete solve of the 3D diffusion workload (64^3 hexes split into
1,572,864 jittered Kuhn tetrahedra, log-normal k with sigma 1.5 box-filtered over
3x3x3 hexes, Dirichlet u = x, f ~ N(0,1)): device assembly into the CSR pattern,
IC(0) factorisation, and PCG from x = 0 to a 1e-8 relative residual.  Mesh geometry,
the CSR pattern, the face colouring and the triangular-solve level analysis are
per-mesh setup done before timing; fields and boundary data are resident in HBM.

Prints one JSON line.  `roofline` is the dominant kernel's algorithmic HBM bytes per
call divided by its average duration, measured with HIP events recorded on the
library's stream around every call in the timed region.  `cpu_baseline` times the
CPU FP64 oracle (the reference's algorithm restated in C++, one core) on one full
solve of the same inputs.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "finite-volumes-rs_amd"))

HBM_PEAK_GBS = 8000.0
TIMING_EVERY = 16
METRIC = "PCG/BiCGStab solves/sec + SpMV HBM GB/s at 1e-8 rel. residual; p50/p99 latency"


def spmv_bytes(n, nnz, w=8):
    return (w + 4) * nnz + 4 * (n + 1) + 2 * w * n


def sweep_bytes(n, nnz_strict, w=8):
    return (w + 4) * nnz_strict + 4 * (n + 1) + 3 * w * n


# Bytes per row a fused CG sweep streams besides its triangular solve, averaged over the
# pair: the lower sweep forms its right-hand side r - ap * alpha itself and updates x, so
# it also reads ap, x and p and writes x and r (5 * 8 = 40); the upper one reads r for r.z
# (8).
FUSED_CG_BYTES_PER_ROW = (40 + 8) // 2


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nx", type=int, default=64, help="hexes per direction (64 = the benchmark workload)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--direct-launches", action="store_true",
                    help="launch the mixed-precision inner iterations kernel by kernel instead of replaying a "
                         "captured graph (same kernels and work; for kernel tracers)")
    ap.add_argument("--legs", default="batched,serving,convdiff,transient",
                    help="extra legs beside the headline solve: batched, serving, convdiff, transient "
                         "(comma list, or none)")
    ap.add_argument("--batch-systems", type=int, default=256)
    ap.add_argument("--serve-requests", type=int, default=400)
    ap.add_argument("--serve-qps", type=float, default=80.0)
    ap.add_argument("--serve-fresh", type=int, default=60, help="requests on meshes of their own (serving leg)")
    ap.add_argument("--serve-fresh-qps", type=float, default=10.0)
    ap.add_argument("--serve-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run a plain device copy of known size (scales rocprof FETCH_SIZE / WRITE_SIZE "
                         "counters of the same run)")
    return ap.parse_args()


def layout_summary(prec):
    """The blocked sweep layout of a preconditioner: workgroups, levels (the dependency
    levels, and those of the blocked layout) and the widest level of one workgroup."""
    lay = prec.sweep_layout()
    return {"workgroups": lay["lower_blocks"], "dependency_levels": lay["levels"],
            "block_levels": lay["block_levels"], "widest_rows": lay["lower_rows"]}


def leg_batched(args, fvg, workloads):
    """256 independent 2D diffusion systems (cells log-uniform in [1k, 64k]) solved by one
    graph-launched batch: FP64 (one workgroup per system, workgroup size autotuned) and
    mixed precision (each system on ceil(n / rows) workgroups, rows per workgroup autotuned)."""
    inst = workloads.batched_2d(args.batch_systems)
    ctx = fvg.Context(0)
    A, M, B, X, ns = [], [], [], [], []
    for raw, fields in inst:
        mesh = fvg.Mesh(ctx, raw)
        pr = fields(mesh.geometry())
        lhs, rhs = fvg.assemble_program(mesh, pr["prog"], pr["face_fields"], pr["cell_fields"], pr["bc_i"],
                                        pr["bc_e"])
        A.append(lhs)
        B.append(rhs)
        M.append(fvg.IncompleteCholesky.from_matrix(lhs, 0))
        X.append(fvg.DistributedVector(ctx, mesh.n_cells))
        ns.append(mesh.n_cells)
    ctx.sync()
    res = {}
    xs = {}
    # fp64 and mixed_one_workgroup share the schedule (one workgroup per system, exact
    # level-scheduled sweeps), so their ratio is the precision alone; mixed spreads each
    # system over a cluster of workgroups with Jacobi-iterated factor applies
    for name, prec in (("fp64", fvg.PRECISION_FP64), ("mixed_one_workgroup", fvg.PRECISION_MIXED),
                       ("mixed", fvg.PRECISION_MIXED)):
        opts = fvg.LinearSolverOptions(1e-8, 0.0, 5000, precision=prec)
        bt = fvg.Batch(ctx, A, M, "cg", opts)
        if name == "mixed_one_workgroup":
            bt.set_cluster(0)
        tune = bt.tune(B, X)
        bucket = None
        if name == "mixed":
            # rows per workgroup per size bucket, each bucket's systems timed on their own;
            # kept when the whole batch runs faster than with the best single choice
            single = bt.rows_per_group
            btimes = bt.tune_buckets(B, X)
            t_b = min(bt._time_solve(B, X, 1) for _ in range(2))
            kept = t_b < min(tune.values())
            bucket = {"limits": bt.buckets[0], "rows": bt.buckets[1] + [bt.rows_per_group],
                      "ms_per_batch": 1e3 * t_b, "kept": kept,
                      "bucket_autotune_ms": {str(k): {str(r): 1e3 * v for r, v in t.items()}
                                             for k, t in btimes.items()}}
            if not kept:
                bt.set_cluster_buckets([], [])
                bt.set_cluster(single, bt.max_workers)
        times = []
        for _ in range(max(1, args.steps)):
            for x in X:
                x.fill(0.0)
            ctx.sync()
            t = time.perf_counter()
            infos = bt.solve(B, X, history=False)
            times.append(time.perf_counter() - t)
        xs[name] = [x.numpy() for x in X]
        worst = 0.0
        for k in range(len(A)):   # true relative residual of every system, FP64 SpMV on the device
            y = fvg.DistributedVector(ctx, ns[k])
            A[k].imul(y, X[k])
            bh = B[k].numpy()
            worst = max(worst, float(np.linalg.norm(bh - y.numpy()) / np.linalg.norm(bh)))
        it = np.array([i.iterations for i in infos])
        best = min(times)
        shape = ({"rows_per_workgroup": bt.rows_per_group, "resident_workgroups": bt.workers}
                 if bt.rows_per_group > 0 else {"threads_per_system": bt.threads})
        if bucket is not None:
            shape["size_buckets"] = bucket
        res[name] = {"solves_per_s": len(A) / best, "ms_per_batch": 1e3 * best, "launch_shape": shape,
                     "autotune_ms": {str(k): 1e3 * v for k, v in tune.items()},
                     "iterations_max": int(it.max()), "iterations_mean": float(it.mean()),
                     "worst_true_relative_residual": worst}
        if bt.rows_per_group == 0:
            res[name]["phases"] = batch_phase_rates(bt, A, ns, B, X, best, 8 if prec == fvg.PRECISION_FP64 else 4)
        bt.close()
    for name in ("mixed_one_workgroup", "mixed"):
        err = max(float(np.linalg.norm(a - b) / np.linalg.norm(b)) for a, b in zip(xs[name], xs["fp64"]))
        res[name]["speedup_vs_fp64"] = res[name]["solves_per_s"] / res["fp64"]["solves_per_s"]
        res[name]["max_rel_error_vs_fp64_solution"] = err
    res["systems"] = len(A)
    res["total_cells"] = int(sum(ns))
    res["largest_system"] = int(max(ns))
    res["workload"] = ("batched_2d: 256 jittered triangle meshes, cells log-uniform in [1k, 64k], ln k ~ N(0,1) "
                       "per cell, f ~ N(0,1), u = 0; PCG + IC(0) to 1e-8 rel. residual, x0 = 0; one graph "
                       "launch per batch")
    ctx.close()
    return res


def batch_phase_rates(bt, A, ns, B, X, batch_s, w):
    """SpMV and triangular-sweep rates of a one-workgroup-per-system batch from its phase
    timers (one more solve with timing on): per system, (iterations + 1) SpMVs and lower
    and upper IC(0) sweeps of the algorithmic bytes of spmv_bytes / sweep_bytes at value
    width w.  per_workgroup: a phase's bytes over its time summed over the systems (the
    rate one CU reaches); device: all systems' bytes of the phase over the batch's wall
    time (what the device moved for that phase while the batch ran)."""
    bt.set_phase_timing(True)
    for x in X:
        x.fill(0.0)
    infos = bt.solve(B, X, history=False)
    us = bt.phase_times()
    bt.set_phase_timing(False)
    out = {}
    for j, name in ((0, "spmv"), (1, "lower_sweep"), (2, "upper_sweep")):
        tot_b, tot_us = 0.0, 0.0
        for k, (a, n) in enumerate(zip(A, ns)):
            calls = infos[k].iterations + 1
            nnz_s = (a.nnz - n) // 2
            tot_b += calls * (spmv_bytes(n, a.nnz, w) if j == 0 else sweep_bytes(n, nnz_s, w))
            tot_us += us[k, j]
        per_wg = tot_b / (tot_us * 1e-6) / 1e9 if tot_us > 0 else 0.0
        dev = tot_b / batch_s / 1e9
        out[name] = {"algorithmic_bytes": tot_b, "busy_ms_summed": 1e-3 * tot_us, "per_workgroup_GB/s": per_wg,
                     "device_GB/s": dev, "frac": dev / HBM_PEAK_GBS}
    out["solve_ms_largest_system"] = 1e-3 * float(us[int(np.argmax(ns)), 3])
    return out


def leg_serving(args, fvg, workloads):
    """A Poisson stream of mixed-size solves through the continuous-batching server, in a
    process of its own (its streams then get hardware queues of their own)."""
    import subprocess
    cmd = [sys.executable, os.path.abspath(__file__), "--serve-child", "--serve-requests", str(args.serve_requests),
           "--serve-qps", str(args.serve_qps), "--serve-fresh", str(args.serve_fresh),
           "--serve-fresh-qps", str(args.serve_fresh_qps)]
    # the server keeps several streams busy at once (per lane: preparation and batches;
    # allocation, result copies): give each a hardware queue of its own (read when the HIP
    # runtime starts)
    env = dict(os.environ)
    env["GPU_MAX_HW_QUEUES"] = "32"
    res = subprocess.run(cmd, capture_output=True, text=True, timeout=600, env=env)
    if res.returncode != 0:
        raise RuntimeError("serving leg failed: " + res.stderr[-2000:])
    return json.loads(res.stdout.strip().splitlines()[-1])


def serve_child(args):
    import fvg
    import workloads
    out = _serve(args, fvg, workloads)
    print(json.dumps(out), flush=True)


def _serve(args, fvg, workloads):
    import serving
    pool = workloads.serving_pool()
    srv = serving.Server(budget_bytes=200 * 10**9, lanes=6, max_batch=48)
    for k, raw in pool.items():
        srv.register_mesh(k, raw)
    geo = {k: srv.meshes[k].mesh.geometry() for k in pool}
    # resident slots for every mesh and solver kind (setup, not serving), then a warm-up
    # stream that loads every kernel variant and sizes the staging buffers
    tmpl = {}
    for r in workloads.serving_requests(pool, geo.get, 400, args.serve_qps, seed=10):
        tmpl.setdefault((r.mesh, r.kind), r)
    for (key, kind), r in tmpl.items():
        srv.reserve(r, 6 if key.startswith("tet") else 4)
    srv.run(workloads.serving_requests(pool, geo.get, 96, args.serve_qps, seed=11))
    reqs = workloads.serving_requests(pool, geo.get, args.serve_requests, args.serve_qps, seed=12)
    wall = srv.run(reqs)
    s = serving.latency_summary(reqs, wall)
    # the same mix offered at three times the rate: latency near capacity
    hi = workloads.serving_requests(pool, geo.get, args.serve_requests, 3 * args.serve_qps, seed=13)
    hwall = srv.run(hi)
    hs = serving.latency_summary(hi, hwall)
    hs["offered_qps"] = 3 * args.serve_qps
    hs["failed"] = int(sum(1 for r in hi if r.error is not None or not (r.info.true_residual <= 1e-8)))
    # requests on meshes of their own: the server builds each mesh, assembles its pattern and
    # analyses its factor inside the request's latency (the client made the fields with the
    # geometry from a context of its own, before the run)
    client = fvg.Context(0)

    def geometry_of_raw(raw):
        m = fvg.Mesh(client, raw)
        g = m.geometry()
        m.close()
        return g

    fresh = workloads.serving_fresh_requests(geometry_of_raw, args.serve_fresh, args.serve_fresh_qps, seed=14)
    client.close()
    st0 = dict(srv.stats)
    fwall = srv.run(fresh)
    fs = serving.latency_summary(fresh, fwall)
    built = srv.stats.get("meshes_built", 0) - st0.get("meshes_built", 0)
    fs.update({"offered_qps": args.serve_fresh_qps, "meshes_built": built,
               "mesh_build_ms_mean": 1e3 * (srv.stats.get("mesh_build_s", 0.0) - st0.get("mesh_build_s", 0.0)) / max(1, built),
               "slot_setup_ms_mean": 1e3 * (srv.stats["slot_create_s"] - st0["slot_create_s"])
               / max(1, srv.stats["slots_created"] - st0["slots_created"]),
               "failed": int(sum(1 for r in fresh if r.error is not None or not (r.info.true_residual <= 1e-8))),
               "workload": "the same mix, each request on a new mesh of its own (sizes log-uniform as above); "
                           "latency includes the device mesh build, CSR pattern, assembler set-up and the "
                           "factor's layout analysis"})
    by = {}
    for name, sel in (("tri", lambda r: r.mesh.startswith("tri")), ("tet", lambda r: r.mesh.startswith("tet"))):
        sub = [r for r in reqs if sel(r)]
        if sub:
            lat = np.array([r.latency for r in sub]) * 1e3
            by[name] = {"count": len(sub), "p50_ms": float(np.percentile(lat, 50)),
                        "p99_ms": float(np.percentile(lat, 99))}
    wait = np.array([r.launched - r.arrival for r in reqs]) * 1e3
    run = np.array([r.done - r.launched for r in reqs]) * 1e3
    s["queue_ms"] = {"p50": float(np.percentile(wait, 50)), "p99": float(np.percentile(wait, 99))}
    s["solve_ms"] = {"p50": float(np.percentile(run, 50)), "p99": float(np.percentile(run, 99))}
    s.update({"offered_qps": args.serve_qps, "requests": len(reqs), "by_mesh": by, "lanes": len(srv.lanes),
              "max_batch": srv.max_batch, "resident_slots": len(srv.slots),
              "mean_batch": float(np.mean([r.batch_size for r in reqs])),
              "peak_hbm_bytes": int(srv.peak_bytes), "budget_bytes": srv.budget,
              "failed": int(sum(1 for r in reqs if r.error is not None or not (r.info.true_residual <= 1e-8))),
              "worst_true_relative_residual": float(max(r.info.true_residual for r in reqs)),
              "fp64_fallbacks": int(sum(1 for r in reqs if r.info.flags & 1)),
              "high_load": hs,
              "fresh_meshes": fs,
              "workload": "serving: Poisson arrivals; 70% 2D triangle meshes (1k-64k cells), 30% 3D Kuhn tet meshes "
                          "(50k-500k cells) from a pool of 16; 70% diffusion (ln k ~ N(0,1), CG + IC(0)), 30% "
                          "convection-diffusion (cell Peclet U[1,50], BiCGStab + ILU(0)); mixed precision to 1e-8; "
                          "latency = scheduled arrival to solution on the host"})
    srv.close()
    return s


def leg_convdiff(args, fvg, workloads):
    """Nonsymmetric convection-diffusion on 128x128x2 triangles x 64 layers of prisms
    (2,097,152 cells): assemble + ILU(0) + BiCGStab to 1e-8, FP64 and mixed."""
    raw, fields_of = workloads.convdiff_prism()
    ctx = new_context(fvg, args)
    mesh = fvg.Mesh(ctx, raw)
    pr = fields_of(mesh.geometry())
    asm = fvg.Assembler(mesh, pr["prog"], pr["face_fields"], pr["cell_fields"], pr["bc_i"], pr["bc_e"])
    lhs = fvg.DistributedMatrix.from_mesh(ctx, mesh)
    rhs = fvg.DistributedVector.from_size(ctx, mesh.n_cells)
    x = fvg.DistributedVector.from_size(ctx, mesh.n_cells)
    asm.run(lhs, rhs)
    ilu = fvg.IncompleteLowerUpper.from_matrix(lhs, 0)
    out = {"workload": "convdiff_prism: 128x128x2 jittered triangles x 64 layers = 2,097,152 prisms, velocity "
                       "(1,0.5,0.25) scaled to cell Peclet U[1,100], k = 1e-2, upwind, inflow u = 1 / outflow zero "
                       "gradient; assemble + ILU(0) + BiCGStab to 1e-8 rel. residual, x0 = 0",
           "n_cells": mesh.n_cells, "nnz": mesh.nnz, "mesh_build_ms": time_mesh_build(fvg, ctx, raw)}
    xs = {}
    n, nnz = mesh.n_cells, mesh.nnz
    nnz_s = (nnz - n) // 2
    # mixed: FP32 inner iterations applying ILU(0) by Jacobi-iterated sweeps (the default);
    # mixed_exact_sweeps: the same with exact level-scheduled FP32 sweeps
    for name, prec, exact in (("fp64", fvg.PRECISION_FP64, 0), ("mixed", fvg.PRECISION_MIXED, 0),
                              ("mixed_exact_sweeps", fvg.PRECISION_MIXED, 1)):
        opts = fvg.LinearSolverOptions(1e-8, 0.0, 4000, precision=prec)
        ctx.set_tuning(fvg.TUNE_INNER_EXACT, exact)

        def step():
            asm.run(lhs, rhs, sync=False)
            ilu.refactor(lhs)
            x.fill(0.0)
            return fvg.bi_conjugate_gradient_stab(x, lhs, rhs, ilu, opts)

        step()
        ctx.sync()
        if name == "fp64":
            ctx.kernel_stats(reset=True)
            ctx.enable_timing(True, every=TIMING_EVERY)
        t = time.perf_counter()
        steps = max(1, min(args.steps, 3))
        for _ in range(steps):
            info = step()
        ctx.sync()
        el = (time.perf_counter() - t) / steps
        y = fvg.DistributedVector.from_size(ctx, mesh.n_cells)
        lhs.imul(y, x)
        bh = rhs.numpy()
        xs[name] = x.numpy()
        out[name] = {"solves_per_s": 1.0 / el, "ms_per_solve": 1e3 * el, "iterations": info.iterations,
                     "true_relative_residual": float(np.linalg.norm(bh - y.numpy()) / np.linalg.norm(bh))}
        if name == "fp64":
            ctx.enable_timing(False)
            ks = ctx.kernel_stats(reset=True)
            sp_ms = ks["spmv_ms"] / max(1, ks["spmv_calls"])
            sw_ms = ks["sweep_ms"] / max(1, ks["sweep_calls"])
            # ILU(0) sweeps alternate lower (unit diagonal) and upper: their mean bytes
            sw_b = 0.5 * ((8 + 4) * nnz_s + 4 * (n + 1) + 2 * 8 * n) + 0.5 * sweep_bytes(n, nnz_s)
            out["fp64"]["kernels"] = {
                "spmv": {"avg_ms": sp_ms, "algorithmic_bytes": spmv_bytes(n, nnz),
                         "GB/s": spmv_bytes(n, nnz) / (sp_ms * 1e-3) / 1e9,
                         "frac": spmv_bytes(n, nnz) / (sp_ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
                "ilu0_sweep": {"span": "one triangular solve (lower or upper), HIP events around one launch group "
                                       f"in {TIMING_EVERY}", "avg_ms": sw_ms, "algorithmic_bytes": sw_b,
                               "GB/s": sw_b / (sw_ms * 1e-3) / 1e9, "frac": sw_b / (sw_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                               "levels": list(ilu.levels()), "layout": layout_summary(ilu)}}
    ctx.set_tuning(fvg.TUNE_INNER_EXACT, 0)
    for name in ("mixed", "mixed_exact_sweeps"):
        out[name]["speedup_vs_fp64"] = out[name]["solves_per_s"] / out["fp64"]["solves_per_s"]
        out[name]["rel_error_vs_fp64_solution"] = float(np.linalg.norm(xs[name] - xs["fp64"]) /
                                                        np.linalg.norm(xs["fp64"]))
    ctx.close()
    return out


def leg_transient(args, fvg, workloads):
    """One time step of the crate's ex4 on the 2,097,152-prism mesh of the convection-
    diffusion leg, everything on the device: compute_gradients, compute_limiters
    (LimitedLinear(1), Upwind on the flux, Corrected), assembly of time(Euler, dt = 0.01) +
    convection(LimitedLinear) - laplacian(Corrected) - source (a unit source in a ball of
    radius 0.1), then ILU(1) + BiCGStab to 1e-8, as examples/ex4/ex4.rs:137-155 does.
    The oracle's gradients + limiters + assembly of the same step run beside it on one
    host core."""
    raw, fields_of = workloads.convdiff_prism()
    ctx = new_context(fvg, args)
    mesh = fvg.Mesh(ctx, raw)
    geom = mesh.geometry()
    pr = fields_of(geom)
    flux, kf = pr["face_fields"]
    bci, bce = pr["bc_i"], pr["bc_e"]
    bc = fvg.FaceConstraints(bci, bce)
    n, F, nb, D = mesh.n_cells, mesh.n_faces, mesh.n_boundary_faces, 3
    u0 = np.exp(-8.0 * ((geom["cell_center"] - 0.5) ** 2).sum(1))
    src = (np.linalg.norm(geom["cell_center"] - np.array([0.25, 0.5, 0.5]), axis=1) < 0.1).astype(np.float64)
    dt = 0.01
    u = fvg.DistributedVector.from_data(ctx, u0)
    g = fvg.DistributedVector(ctx, n * D)
    lim = fvg.DistributedVector(ctx, F)
    fl = fvg.DistributedVector.from_data(ctx, flux)
    up, corr = fvg.Upwind(fl), fvg.Corrected(g, 1.0)
    llim = fvg.limiters.LimitedLinear(1.0)
    eq = (fvg.time(fvg.Euler(u, dt)) + fvg.convection(fvg.LimitedLinear(fl, lim), fl)
          - fvg.laplacian(fvg.Corrected(g, 1.0), kf) - fvg.source(src))
    ops, ff, cf = eq.compile()
    asm = fvg.Assembler(mesh, ops, ff, cf, bci, bce)
    lhs = fvg.DistributedMatrix.from_mesh(ctx, mesh)
    rhs = fvg.DistributedVector.from_size(ctx, n)
    x = fvg.DistributedVector.from_size(ctx, n)
    fvg.compute_gradients(g, u, bc, mesh)
    fvg.compute_limiters(lim, u, g, llim, up, corr, bc, mesh)
    asm.run(lhs, rhs)
    ilu = fvg.IncompleteLowerUpper.from_matrix(lhs, 1)
    opts = fvg.LinearSolverOptions(1e-8, 0.0, 1000)

    def timed(fn, reps=20):
        fn()
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        ctx.sync()
        return 1e3 * (time.perf_counter() - t0) / reps

    grad_ms = timed(lambda: fvg.compute_gradients(g, u, bc, mesh))
    lim_ms = timed(lambda: fvg.compute_limiters(lim, u, g, llim, up, corr, bc, mesh))
    asm_ms = timed(lambda: asm.run(lhs, rhs, sync=False))

    def step():
        fvg.compute_gradients(g, u, bc, mesh)
        fvg.compute_limiters(lim, u, g, llim, up, corr, bc, mesh)
        asm.run(lhs, rhs, sync=False)
        ilu.refactor(lhs)
        x.fill(0.0)
        return fvg.bi_conjugate_gradient_stab(x, lhs, rhs, ilu, opts)

    step()
    ctx.sync()
    steps = max(1, min(args.steps, 3))
    t0 = time.perf_counter()
    for _ in range(steps):
        info = step()
    ctx.sync()
    step_ms = 1e3 * (time.perf_counter() - t0) / steps
    L = int(raw["cell_ptr"][-1])
    # algorithmic bytes: every input read once, every output written once
    grad_bytes = n * (4 + 8 * D + 8 + 8 * D) + L * (4 + 8 * D) + nb * 16
    lim_bytes = F * (4 + 4 + 8 * D + 8 * D + 8 + 8) + n * (8 * D + 8 + 8 * D) + nb * (4 + 16)
    out = {"workload": "ex4 step on convdiff_prism: 2,097,152 prisms, LimitedLinear(1) limiters (Upwind, "
                       "Corrected), time(Euler, dt=0.01) + convection(LimitedLinear) - laplacian(Corrected) "
                       "- source (unit source in a ball of radius 0.1), ILU(1) + BiCGStab to 1e-8",
           "ilu_level": 1, "factor_nnz": int(ilu.factor_pattern()[0][-1]), "layout": layout_summary(ilu),
           "n_cells": n, "n_faces": F, "step_ms": step_ms, "steps_per_s": 1e3 / step_ms,
           "iterations": info.iterations,
           "gradients": {"avg_ms": grad_ms, "algorithmic_bytes": grad_bytes,
                         "GB/s": grad_bytes / (grad_ms * 1e-3) / 1e9,
                         "frac": grad_bytes / (grad_ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
           "limiters": {"avg_ms": lim_ms, "algorithmic_bytes": lim_bytes,
                        "GB/s": lim_bytes / (lim_ms * 1e-3) / 1e9,
                        "frac": lim_bytes / (lim_ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
           "assembly_ms": asm_ms}
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        om = O.OracleMesh(raw)
        t0 = time.perf_counter()
        go = om.gradients(u0, bci, bce)
        lo = om.limiters(u0, go, "limited_linear", 1.0, "upwind", flux, go, 1.0, bci, bce)
        prog = [(O.OP["TIME_EULER"], 0, -1, -1, dt), (O.OP["CONVECTION_LIMITED_LINEAR"], 0, 0, 1, 0.0),
                (O.OP["ADD"], 0, 0, 0, 0.0), (O.OP["LAPLACIAN_CORRECTED"], 2, 1, 0, 1.0), (O.OP["SUB"], 0, 0, 0, 0.0),
                (O.OP["SOURCE"], 2, 0, 0, 0.0), (O.OP["SUB"], 0, 0, 0, 0.0)]
        om.assemble(prog, [flux, lo, kf], [u0, go.ravel(), src], bci, bce)
        t_cpu = time.perf_counter() - t0
        out["cpu_oracle_tools_ms"] = 1e3 * t_cpu
        out["tools_speedup_vs_cpu"] = 1e3 * t_cpu / (grad_ms + lim_ms + asm_ms)
    ctx.close()
    return out


def measured_traffic(kernel, args):
    """HBM bytes per launch of `kernel` on this workload, from the rocprof FETCH_SIZE /
    WRITE_SIZE passes summarised in profiles/ (scaled by the copy calibration), or None."""
    path = os.path.join(ROOT, "profiles", "traffic_diffusion_3d.json")
    if args.nx != 64 or not os.path.exists(path):
        return None
    with open(path) as f:
        k = json.load(f)["kernels"].get(kernel)
    return None if k is None else k["traffic_bytes"]


def new_context(fvg, args):
    ctx = fvg.Context(0)
    if args.direct_launches:
        ctx.set_tuning(fvg.TUNE_GRAPHS, 1)
    return ctx


def assembly_bytes(mesh, pr):
    """Algorithmic HBM bytes of one assembly: per face its area, Orthogonal factor and
    face fields (read once), per cell its volume, row and entry offsets, diagonal slot,
    cell fields and right-hand side, the per-cell entry lists (8 bytes per face side),
    the boundary closure pair per boundary face, and every CSR value written once."""
    n, F, nb, nnz = mesh.n_cells, mesh.n_faces, mesh.n_boundary_faces, mesh.nnz
    sides = 2 * F - nb
    return (F * (16 + 8 * len(pr["face_fields"])) + n * (8 + 4 + 4 + 4 + 8 + 8 * len(pr["cell_fields"])) +
            8 * sides + 16 * nb + 8 * nnz)


def time_mesh_build(fvg, ctx, raw, reps=3):
    """Wall time (median of `reps`) of Mesh::compute plus the cell-to-cell CSR pattern
    on the device, the topology arrays starting in host memory."""
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        m = fvg.Mesh(ctx, raw)
        a = fvg.DistributedMatrix.from_mesh(ctx, m)
        ctx.sync()
        ts.append(time.perf_counter() - t0)
        a.close()
        m.close()
    return 1e3 * float(np.median(ts))


def time_assembly(ctx, asm, lhs, rhs, reps=20):
    """Device time of one assembly (a single row-gather launch), averaged over `reps`
    back-to-back launches."""
    asm.run(lhs, rhs)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        asm.run(lhs, rhs, sync=False)
    ctx.sync()
    return 1e3 * (time.perf_counter() - t0) / reps


CPU_RUNS = 5


def cpu_baseline(raw, fields_of):
    """Full solves of the same inputs by the oracle on one host core: the median of CPU_RUNS
    timed solves (assemble + IC(0) + CG) after an untimed assembly and factorisation."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    t0 = time.perf_counter()
    om = O.OracleMesh(raw)
    t_mesh = time.perf_counter() - t0
    pr = fields_of(om.geometry())
    rp, ci, va, b = om.assemble(pr["prog"], pr["face_fields"], pr["cell_fields"], pr["bc_i"], pr["bc_e"])
    O.OraclePrec("ic", rp, ci, va, 0)   # warm-up
    ts = []
    for _ in range(CPU_RUNS):
        t0 = time.perf_counter()
        rp, ci, va, b = om.assemble(pr["prog"], pr["face_fields"], pr["cell_fields"], pr["bc_i"], pr["bc_e"])
        P = O.OraclePrec("ic", rp, ci, va, 0)
        x, info = O.solve("cg", rp, ci, va, P, b, rtol=1e-8, atol=0.0, max_iter=2000)
        ts.append(time.perf_counter() - t0)
    t = float(np.median(ts))
    return dict(value=1.0 / t, unit="solves/s", cores=1, kind="port",
                sample=f"median of {CPU_RUNS} full solves of the same workload (assemble + IC(0) + CG to 1e-8, "
                       f"{info['iterations']} iterations; {min(ts):.2f}-{max(ts):.2f} s each) after an untimed "
                       f"assembly and factorisation; mesh geometry ({t_mesh:.2f} s) excluded",
                runs_s=ts, mesh_build_ms=1e3 * t_mesh)


def calibrate_copy():
    """Five launches of an elementwise kernel that reads and writes 512 MiB each: the
    HBM bytes a hardware counter should report per launch are known exactly."""
    import torch
    n = 1 << 27
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(5):
        torch.mul(a, 1.0, out=b)
    torch.cuda.synchronize()
    print("bench: calibration copy " + json.dumps({"kernel": "torch elementwise mul (copy)", "bytes_read": 4 * n,
                                                   "bytes_written": 4 * n, "launches": 5}), file=sys.stderr, flush=True)
    del a, b
    torch.cuda.empty_cache()


def main():
    args = parse()
    if args.serve_child:
        serve_child(args)
        return
    if args.calibrate_copy:
        calibrate_copy()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU path is not built.",
              file=sys.stderr)
        sys.exit(2)
    legs = [l for l in args.legs.split(",") if l and l != "none"]
    serving_out = None
    if "serving" in legs:
        # first, while this process holds no GPU queues, so that the server's streams get
        # hardware queues of their own
        print("bench: serving leg", file=sys.stderr, flush=True)
        serving_out = leg_serving(args, None, None)
    import fvg
    import workloads

    raw, fields_of = workloads.diffusion_3d(args.nx)
    ctx = new_context(fvg, args)
    mesh = fvg.Mesh(ctx, raw)
    pr = fields_of(mesh.geometry())
    asm = fvg.Assembler(mesh, pr["prog"], pr["face_fields"], pr["cell_fields"], pr["bc_i"], pr["bc_e"])
    lhs = fvg.DistributedMatrix.from_mesh(ctx, mesh)
    rhs = fvg.DistributedVector.from_size(ctx, mesh.n_cells)
    x = fvg.DistributedVector.from_size(ctx, mesh.n_cells)
    asm.run(lhs, rhs)
    mesh_build_ms = time_mesh_build(fvg, ctx, raw)
    asm_ms = time_assembly(ctx, asm, lhs, rhs)
    ic = fvg.IncompleteCholesky.from_matrix(lhs, 0)   # pattern analysis + first factorisation
    opts = fvg.LinearSolverOptions(relative_tolerance=1e-8, absolute_tolerance=0.0, max_iterations=2000)
    n, nnz = mesh.n_cells, mesh.nnz
    nnz_s = (nnz - n) // 2

    opts_mixed = fvg.LinearSolverOptions(relative_tolerance=1e-8, absolute_tolerance=0.0, max_iterations=4000,
                                         precision=fvg.PRECISION_MIXED)

    def step(mixed=False):
        asm.run(lhs, rhs, sync=False)
        ic.refactor(lhs)
        x.fill(0.0)
        return fvg.conjugate_gradient(x, lhs, rhs, ic, opts_mixed if mixed else opts)

    def true_residual():
        y = fvg.DistributedVector.from_size(ctx, n)
        lhs.imul(y, x)
        bh = rhs.numpy()
        return float(np.linalg.norm(bh - y.numpy()) / np.linalg.norm(bh))

    for _ in range(args.warmup):
        step()
    ctx.sync()
    ctx.kernel_stats(reset=True)
    # HIP events around one SpMV and one triangular solve in every TIMING_EVERY (the
    # markers of every call would slow the stream they measure)
    ctx.enable_timing(True, every=TIMING_EVERY)
    lat = []
    infos = []
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        infos.append(step())
        lat.append(time.perf_counter() - ts)
    ctx.sync()
    elapsed = time.perf_counter() - t0
    ctx.enable_timing(False)
    ks = ctx.kernel_stats(reset=True)

    true_rel = true_residual()
    x_fp64 = x.numpy()
    # the SpMV alone: back-to-back launches between two events (no per-call event overhead)
    y_spmv = fvg.DistributedVector.from_size(ctx, n)
    spmv_stream_ms = lhs.time_spmv(y_spmv, x, reps=50)

    # mixed precision: FP32 inner PCG + FP64 refinement, same inputs and tolerance
    for _ in range(args.warmup):
        step(mixed=True)
    ctx.sync()
    t0 = time.perf_counter()
    minfos = [step(mixed=True) for _ in range(args.steps)]
    ctx.sync()
    m_elapsed = time.perf_counter() - t0
    m_true = true_residual()
    m_err = float(np.linalg.norm(x.numpy() - x_fp64) / np.linalg.norm(x_fp64))
    # the same with exact level-scheduled FP32 sweeps in the inner iterations
    ctx.set_tuning(fvg.TUNE_INNER_EXACT, 1)
    step(mixed=True)
    ctx.sync()
    t0 = time.perf_counter()
    einfos = [step(mixed=True) for _ in range(args.steps)]
    ctx.sync()
    e_elapsed = time.perf_counter() - t0
    e_true = true_residual()
    e_err = float(np.linalg.norm(x.numpy() - x_fp64) / np.linalg.norm(x_fp64))
    ctx.set_tuning(fvg.TUNE_INNER_EXACT, 0)

    spmv_avg_ms = ks["spmv_ms"] / max(1, ks["spmv_calls"])
    sweep_avg_ms = ks["sweep_ms"] / max(1, ks["sweep_calls"])
    spmv_gbs = spmv_bytes(n, nnz) / (spmv_avg_ms * 1e-3) / 1e9
    fused = bool(infos[-1].flags & fvg.INFO_FUSED_SWEEPS)
    sweep_alg = sweep_bytes(n, nnz_s) + (FUSED_CG_BYTES_PER_ROW * n if fused else 0)
    sweep_gbs = sweep_alg / (sweep_avg_ms * 1e-3) / 1e9
    if ks["sweep_ms"] >= ks["spmv_ms"]:
        dom = dict(kernel="ic0_sweep", bytes=sweep_alg, gbs=sweep_gbs, ms=sweep_avg_ms)
    else:
        dom = dict(kernel="spmv", bytes=spmv_bytes(n, nnz), gbs=spmv_gbs, ms=spmv_avg_ms)
    lo_lv, up_lv = ic.levels()
    value = args.steps / elapsed
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"diffusion_3d: {args.nx}^3 hexes -> {n} jittered Kuhn tets, ln k ~ N(0,1.5^2) "
                               "3x3x3 box-filtered, u=x Dirichlet, f~N(0,1); assemble + IC(0) + PCG to 1e-8 "
                               "rel. residual, x0=0",
                   "n_cells": n, "nnz": nnz, "precision": "fp64", "solver": "cg+ic0"},
        "iterations": infos[-1].iterations,
        "final_residual": infos[-1].final_residual,
        "true_relative_residual": true_rel,
        "latency_ms": {"p50": 1e3 * float(np.percentile(lat, 50)), "p99": 1e3 * float(np.percentile(lat, 99))},
        "roofline": {"bound": "hbm", "kernel": dom["kernel"], "achieved": dom["gbs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": dom["gbs"] / HBM_PEAK_GBS, "traffic": measured_traffic(dom["kernel"], args),
                     "algorithmic_bytes": dom["bytes"], "avg_ms": dom["ms"]},
        "mixed": {"solves_per_s": args.steps / m_elapsed, "ms_per_step": 1e3 * m_elapsed / args.steps,
                  "speedup_vs_fp64": (args.steps / m_elapsed) / value, "inner_iterations": minfos[-1].iterations,
                  "outer_iterations": minfos[-1].outer_iterations, "true_relative_residual": m_true,
                  "rel_error_vs_fp64_solution": m_err, "inner_apply": "IC(0) by Jacobi-iterated FP32 sweeps",
                  "exact_sweeps": {"solves_per_s": args.steps / e_elapsed, "ms_per_step": 1e3 * e_elapsed / args.steps,
                                   "speedup_vs_fp64": (args.steps / e_elapsed) / value,
                                   "inner_iterations": einfos[-1].iterations,
                                   "outer_iterations": einfos[-1].outer_iterations, "true_relative_residual": e_true,
                                   "rel_error_vs_fp64_solution": e_err,
                                   "inner_apply": "IC(0) by exact level-scheduled FP32 sweeps"}},
        "mesh": {"build_ms": mesh_build_ms, "assembly_ms": asm_ms,
                 "assembly_algorithmic_bytes": assembly_bytes(mesh, pr),
                 "assembly_GB/s": assembly_bytes(mesh, pr) / (asm_ms * 1e-3) / 1e9},
        "kernels": {
            "spmv": {"avg_ms": spmv_avg_ms, "algorithmic_bytes": spmv_bytes(n, nnz), "GB/s": spmv_gbs,
                     "frac": spmv_gbs / HBM_PEAK_GBS, "calls": ks["spmv_calls"], "total_ms": ks["spmv_ms"],
                     "back_to_back": {"launches": 50, "avg_ms": spmv_stream_ms,
                                      "GB/s": spmv_bytes(n, nnz) / (spmv_stream_ms * 1e-3) / 1e9,
                                      "frac": spmv_bytes(n, nnz) / (spmv_stream_ms * 1e-3) / 1e9 / HBM_PEAK_GBS}},
            "ic0_sweep": {"span": "one triangular solve: one k_sweep_blocked launch (the solve iterates in the "
                                  "lower sweep's position order with the upper layout mirrored, so no set-up or "
                                  f"scatter kernel); HIP events around one launch group in {TIMING_EVERY}",
                          "fused_cg_work": fused,
                          "algorithmic_bytes_note": "triangular solve (values, columns, row offsets, b, x, scale) "
                                                    + (f"plus the fused CG vectors, {FUSED_CG_BYTES_PER_ROW} B per "
                                                       "row on average over the lower and upper sweep" if fused
                                                       else ""),
                          "avg_ms": sweep_avg_ms, "algorithmic_bytes": sweep_alg, "GB/s": sweep_gbs,
                          "frac": sweep_gbs / HBM_PEAK_GBS, "calls": ks["sweep_calls"], "launches": ks["sweep_launches"],
                          "total_ms": ks["sweep_ms"], "levels": [lo_lv, up_lv], "layout": layout_summary(ic)},
        },
    }
    # each leg below runs as the only live context on the device, as the headline did (the
    # sweep launch mode follows the number of live contexts)
    for obj in (ic, y_spmv, x, rhs, lhs, mesh):
        obj.close()
    ctx.close()
    for name, fn in (("batched", leg_batched), ("convdiff", leg_convdiff), ("transient", leg_transient)):
        if name in legs:
            print(f"bench: {name} leg", file=sys.stderr, flush=True)
            out[name] = fn(args, fvg, workloads)
    if serving_out is not None:
        out["serving"] = serving_out
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(raw, fields_of)
        out["mesh"]["cpu_oracle_build_ms"] = out["cpu_baseline"].pop("mesh_build_ms")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
