"""Writes workloads_full.json: the FP64 oracle's results on the full-size workloads.

Run from the repository root (about a minute on one core):
    python tests/golden/make_workload_fixture.py

For each workload the oracle (the reference's algorithm restated in C++, FP64) builds
the mesh with Mesh::compute, assembles the equation, factors IC(0) / ILU(0) and solves
from x = 0 to a 1e-8 relative residual.  Recorded: sizes, SHA-256 of the CSR row
pointers, columns, values and right-hand side (the device must reproduce them bit for
bit), the iteration and restart counts, the initial and final residuals, the solution's
2-norm and its values at 512 seeded sample indices.  The GPU tests compare against
these numbers without running the oracle on the GPU machine.
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "finite-volumes-rs_amd"), os.path.join(ROOT, "oracle")]

import oracle as O  # noqa: E402
import workloads  # noqa: E402

SAMPLES = 512


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def sample_indices(n):
    return np.sort(np.random.default_rng(12345).choice(n, size=min(SAMPLES, n), replace=False))


def record(name, builder):
    raw, fields = builder()
    om = O.OracleMesh(raw)
    pr = fields(om.geometry())
    rp, ci, va, b = om.assemble(pr["prog"], pr["face_fields"], pr["cell_fields"], pr["bc_i"], pr["bc_e"])
    P = O.OraclePrec(pr["prec"], rp, ci, va, 0)
    x, info = O.solve(pr["solver"], rp, ci, va, P, b, rtol=1e-8, atol=0.0, max_iter=5000)
    idx = sample_indices(om.n_cells)
    print(name, om.n_cells, info["iterations"], info["restarts"], flush=True)
    return {"n_cells": om.n_cells, "nnz": int(ci.shape[0]), "solver": pr["solver"], "prec": pr["prec"],
            "sha_row_ptr": sha(rp), "sha_col": sha(ci), "sha_values": sha(va), "sha_rhs": sha(b),
            "iterations": info["iterations"], "restarts": info["restarts"],
            "initial_residual": info["initial_residual"], "final_residual": info["final_residual"],
            "x_norm": float(np.linalg.norm(x)), "sample_idx": idx.tolist(), "x_samples": x[idx].tolist()}


def main():
    out = {"diffusion_2d": record("diffusion_2d", workloads.diffusion_2d),
           "diffusion_3d": record("diffusion_3d", workloads.diffusion_3d),
           "convdiff_prism": record("convdiff_prism", workloads.convdiff_prism)}
    with open(os.path.join(HERE, "workloads_full.json"), "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
