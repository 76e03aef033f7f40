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
METRIC = "PCG/BiCGStab solves/sec + SpMV HBM GB/s at 1e-8 rel. residual; p50/p99 latency"


def spmv_bytes(n, nnz, w=8):
    return (w + 4) * nnz + 4 * (n + 1) + 2 * w * n


def sweep_bytes(n, nnz_strict, w=8):
    return (w + 4) * nnz_strict + 4 * (n + 1) + 3 * w * n


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nx", type=int, default=64, help="hexes per direction (64 = the benchmark workload)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    return ap.parse_args()


def cpu_baseline(raw, fields_of):
    """One full solve of the same inputs by the oracle on one host core."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    t0 = time.perf_counter()
    om = O.OracleMesh(raw)
    t_mesh = time.perf_counter() - t0
    pr = fields_of(om.geometry())
    t0 = time.perf_counter()
    rp, ci, va, b = om.assemble(pr["prog"], pr["face_fields"], pr["cell_fields"], pr["bc_i"], pr["bc_e"])
    P = O.OraclePrec("ic", rp, ci, va, 0)
    x, info = O.solve("cg", rp, ci, va, P, b, rtol=1e-8, atol=0.0, max_iter=2000)
    t = time.perf_counter() - t0
    return dict(value=1.0 / t, unit="solves/s", cores=1, kind="port",
                sample="one full solve of the same workload (assemble + IC(0) + CG to 1e-8), "
                       f"{info['iterations']} iterations in {t:.2f} s; mesh geometry ({t_mesh:.2f} s) excluded")


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU path is not built.",
              file=sys.stderr)
        sys.exit(2)
    import fvg
    import workloads

    raw, fields_of = workloads.diffusion_3d(args.nx)
    ctx = fvg.Context(0)
    mesh = fvg.Mesh(ctx, raw)
    pr = fields_of(mesh.geometry())
    asm = fvg.Assembler(mesh, pr["prog"], pr["face_fields"], pr["cell_fields"], pr["bc_i"], pr["bc_e"])
    lhs = fvg.DistributedMatrix.from_mesh(ctx, mesh)
    rhs = fvg.DistributedVector.from_size(ctx, mesh.n_cells)
    x = fvg.DistributedVector.from_size(ctx, mesh.n_cells)
    asm.run(lhs, rhs)
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
    ctx.enable_timing(True)
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

    spmv_avg_ms = ks["spmv_ms"] / max(1, ks["spmv_calls"])
    sweep_avg_ms = ks["sweep_ms"] / max(1, ks["sweep_calls"])
    spmv_gbs = spmv_bytes(n, nnz) / (spmv_avg_ms * 1e-3) / 1e9
    sweep_gbs = sweep_bytes(n, nnz_s) / (sweep_avg_ms * 1e-3) / 1e9
    if ks["sweep_ms"] >= ks["spmv_ms"]:
        dom = dict(kernel="ic0_sweep", bytes=sweep_bytes(n, nnz_s), gbs=sweep_gbs, ms=sweep_avg_ms)
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
                     "unit": "GB/s", "frac": dom["gbs"] / HBM_PEAK_GBS, "traffic": None,
                     "algorithmic_bytes": dom["bytes"], "avg_ms": dom["ms"]},
        "mixed": {"solves_per_s": args.steps / m_elapsed, "ms_per_step": 1e3 * m_elapsed / args.steps,
                  "speedup_vs_fp64": (args.steps / m_elapsed) / value, "inner_iterations": minfos[-1].iterations,
                  "outer_iterations": minfos[-1].outer_iterations, "true_relative_residual": m_true,
                  "rel_error_vs_fp64_solution": m_err},
        "kernels": {
            "spmv": {"avg_ms": spmv_avg_ms, "algorithmic_bytes": spmv_bytes(n, nnz), "GB/s": spmv_gbs,
                     "frac": spmv_gbs / HBM_PEAK_GBS, "calls": ks["spmv_calls"], "total_ms": ks["spmv_ms"]},
            "ic0_sweep": {"avg_ms": sweep_avg_ms, "algorithmic_bytes": sweep_bytes(n, nnz_s), "GB/s": sweep_gbs,
                          "frac": sweep_gbs / HBM_PEAK_GBS, "calls": ks["sweep_calls"], "launches": ks["sweep_launches"],
                          "total_ms": ks["sweep_ms"], "levels": [lo_lv, up_lv]},
        },
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(raw, fields_of)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
