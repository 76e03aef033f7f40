"""Python host mirror of the finite-volumes assemble-and-solve API over libfvg.so.

Names and argument meanings follow the crate:

    mesh = Mesh.from_arrays(ctx, raw)                  # Mesh::new + add_* + compute
    eq = -laplacian(Orthogonal(), k_f) - source(f)     # fvm::terms algebra
    lhs, rhs = assemble(eq, bc, mesh)                  # fvm::assembly::assemble
    ic = IncompleteCholesky.from_matrix(lhs, 0)        # preconditioners
    info = conjugate_gradient(x, lhs, rhs, ic, LinearSolverOptions(...))

Everything numeric runs in libfvg.so on the GPU; this module only marshals host
arrays.  Importing it on a machine without the library or without a GPU raises:
there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfvg.so")

FVG_OK = 0
STATUS = {1: "FVG_E_INVALID", 2: "FVG_E_HIP", 3: "FVG_E_BUDGET", 4: "FVG_E_ZERO_PIVOT", 5: "FVG_E_NOT_SYMMETRIC",
          6: "FVG_E_MISSING_ENTRY", 7: "FVG_E_MESH", 8: "FVG_E_TIMEOUT", 9: "FVG_E_UNSUPPORTED"}

OP_LAPLACIAN_ORTHOGONAL, OP_CONVECTION_UPWIND, OP_SOURCE, OP_LINEAR_SOURCE, OP_TIME_EULER, OP_DIVERGENCE = 1, 2, 3, 4, 5, 6
OP_LAPLACIAN_CORRECTED, OP_CONVECTION_LINEAR, OP_CONVECTION_LIMITED_LINEAR = 7, 8, 9
OP_TIME_BACKWARD, OP_USER_TERM, OP_PARAM = 10, 11, 15
LIMITER_LIMITED_LINEAR, LIMITER_MINMOD, LIMITER_VANLEER, LIMITER_VANALBADA, LIMITER_SWEBY, LIMITER_SUPERBEE = range(6)
INTERP_LINEAR, INTERP_UPWIND, INTERP_LIMITED_LINEAR = 0, 1, 2
FNGRAD_ORTHOGONAL, FNGRAD_CORRECTED = 0, 1
OP_ADD, OP_SUB, OP_NEG = 16, 17, 18
PRECISION_FP64, PRECISION_MIXED = 0, 1
# fvg_solver_info.flags
INFO_FP64_FALLBACK, INFO_POSITION_ORDER, INFO_FUSED_SWEEPS = 1, 2, 4
(TUNE_SWEEP_MODE, TUNE_SWEEP_WAVES, TUNE_SWEEP_POLL, TUNE_INNER_EXACT, TUNE_INNER_SWEEPS, TUNE_GRAPHS,
 TUNE_SWEEP_BLOCKS, TUNE_SPIN_LIMIT, TUNE_SWEEP_LAUNCH, TUNE_SWEEP_WITHHOLD, TUNE_SOLVE_ORDER, TUNE_ASSEMBLE,
 TUNE_CG_FUSION, TUNE_SWEEP_GAP, TUNE_SWEEP_BOXES) = (1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15)


class FvgError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


class ZeroPivotError(FvgError):
    """IncompleteLowerUpper hit |u_kk| < 1e-16 (the reference panics)."""


class NotSymmetricError(FvgError):
    """IncompleteCholesky on a nonsymmetric matrix (the reference asserts)."""


class BudgetError(FvgError):
    """An allocation would exceed the context's HBM budget."""


class MeshError(FvgError):
    """Mesh::compute found an invalid topology or a degenerate face (the reference
    asserts or returns Err, compute.rs:15-31, 308)."""


_ERRORS = {3: BudgetError, 4: ZeroPivotError, 5: NotSymmetricError, 7: MeshError}


class TermOp(C.Structure):
    _fields_ = [("op", C.c_int32), ("arg0", C.c_int32), ("arg1", C.c_int32), ("arg2", C.c_int32),
                ("scalar", C.c_double)]


class LimiterSpec(C.Structure):
    _fields_ = [("limiter", C.c_int32), ("interp", C.c_int32), ("fngrad", C.c_int32), ("reserved", C.c_int32),
                ("k", C.c_double), ("reduction", C.c_double), ("flux", C.c_void_p), ("fn_grads", C.c_void_p)]


class InterpSpec(C.Structure):
    _fields_ = [("interp", C.c_int32), ("reserved", C.c_int32), ("flux", C.c_void_p), ("limiters", C.c_void_p)]


class SolverOptions(C.Structure):
    _fields_ = [("relative_tolerance", C.c_double), ("absolute_tolerance", C.c_double),
                ("max_iterations", C.c_int64), ("precision", C.c_int32), ("reserved", C.c_int32),
                ("inner_tolerance", C.c_double)]


class SolverInfo(C.Structure):
    _fields_ = [("iterations", C.c_int64), ("initial_residual", C.c_double), ("final_residual", C.c_double),
                ("history_len", C.c_int64), ("restarts", C.c_int64), ("outer_iterations", C.c_int64),
                ("true_residual", C.c_double), ("status", C.c_int32), ("flags", C.c_int32)]


class KernelStats(C.Structure):
    _fields_ = [("spmv_ms", C.c_double), ("sweep_ms", C.c_double), ("spmv_launches", C.c_int64),
                ("sweep_launches", C.c_int64), ("spmv_calls", C.c_int64), ("sweep_calls", C.c_int64)]


_lib = None
P = C.c_void_p


def lib():
    """Load libfvg.so; raises when it is missing (build it with __graft_entry__.build())."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FvgError(1, f"{LIB_PATH} is not built")
    L = C.CDLL(LIB_PATH)
    L.fvg_last_error.restype = C.c_char_p
    L.fvg_status_string.restype = C.c_char_p
    sig = {
        "fvg_ctx_create": [C.c_int, C.c_uint64, C.POINTER(P)],
        "fvg_ctx_destroy": [P],
        "fvg_ctx_sync": [P],
        "fvg_ctx_bytes_in_use": [P, C.POINTER(C.c_uint64)],
        "fvg_ctx_stream": [P, C.POINTER(P)],
        "fvg_ctx_enable_timing": [P, C.c_int],
        "fvg_ctx_set_tuning": [P, C.c_int32, C.c_int64],
        "fvg_ctx_get_tuning": [P, C.c_int32, C.POINTER(C.c_int64)],
        "fvg_noalloc_scratch_bytes": [C.POINTER(C.c_uint64)],
        "fvg_ctx_sweep_launch_state": [P, C.POINTER(C.c_int32), C.POINTER(C.c_int32)],
        "fvg_sweep_launch_rule": [C.c_int64, C.c_int32, C.c_int32, C.c_int32],
        "fvg_ctx_kernel_stats": [P, C.POINTER(KernelStats), C.c_int],
        "fvg_vec_create": [P, C.c_int64, C.POINTER(P)],
        "fvg_vec_wrap": [P, C.c_int64, P, C.POINTER(P)],
        "fvg_vec_destroy": [P, P],
        "fvg_vec_len": [P, C.POINTER(C.c_int64)],
        "fvg_vec_device_ptr": [P, C.POINTER(P)],
        "fvg_vec_upload": [P, P, P],
        "fvg_vec_download": [P, P, P],
        "fvg_vec_dot": [P, P, P, C.POINTER(C.c_double)],
        "fvg_vec_fill": [P, P, C.c_double],
        "fvg_mesh_create": [P, C.c_int, C.c_int64, P, C.c_int64, P, P, P, C.c_int64, P, P, C.POINTER(P)],
        "fvg_mesh_destroy": [P, P],
        "fvg_mesh_sizes": [P, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                           C.POINTER(C.c_int64), C.POINTER(C.c_int32)],
        "fvg_mesh_geometry": [P, P, P, P, P, P, P, P],
        "fvg_topology_read_su2": [C.c_char_p, C.POINTER(P)],
        "fvg_topology_read_native": [C.c_char_p, C.POINTER(P)],
        "fvg_topology_create": [C.c_int, C.c_int64, P, C.c_int64, P, P, P, C.c_int64, P, P, C.POINTER(P)],
        "fvg_topology_destroy": [P],
        "fvg_topology_sizes": [P, C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                               C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                               C.POINTER(C.c_int32)],
        "fvg_topology_arrays": [P, C.POINTER(P), C.POINTER(P), C.POINTER(P), C.POINTER(P), C.POINTER(P),
                                C.POINTER(P)],
        "fvg_topology_patch_name": [P, C.c_int32, C.POINTER(C.c_char_p)],
        "fvg_mesh_create_from_topology": [P, P, C.POINTER(P)],
        "fvg_pvtu_writer_create": [P, C.POINTER(P)],
        "fvg_pvtu_write": [P, C.c_char_p, C.c_int32, P, P, P],
        "fvg_pvtu_write_async": [P, P, C.c_char_p, C.c_int32, P, P, P],
        "fvg_pvtu_wait": [P],
        "fvg_pvtu_writer_destroy": [P],
        "fvg_prec_sweep_layout": [P, P, C.c_int32],
        "fvg_prec_factor_pattern": [P, C.POINTER(C.c_int64), C.POINTER(C.c_int64), P, P],
        "fvg_csr_from_mesh": [P, P, C.POINTER(P)],
        "fvg_csr_create": [P, C.c_int64, C.c_int64, C.c_int64, P, P, P, C.POINTER(P)],
        "fvg_csr_destroy": [P, P],
        "fvg_csr_sizes": [P, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64)],
        "fvg_csr_download": [P, P, P, P, P],
        "fvg_csr_upload_values": [P, P, P],
        "fvg_csr_set_row_coords": [P, C.c_int32, P],
        "fvg_spmv": [P, P, P, P],
        "fvg_spmv_time": [P, P, P, P, C.c_int32, C.POINTER(C.c_double)],
        "fvg_assemble": [P, P, P, C.c_int32, P, C.c_int32, P, C.c_int32, P, P, P, P],
        "fvg_compute_gradients": [P, P, C.c_int32, P, P, P, P],
        "fvg_compute_limiters": [P, P, C.c_int32, P, P, P, P, P, P],
        "fvg_compute_hbya_ainv": [P, P, P, P, P, P, P],
        "fvg_interpolate_hbya_ainv_faces": [P, P, P, P, P, P, P, P, P, P, P, P],
        "fvg_correct_phi": [P, P, P, P, P, C.c_int32, P, C.c_double, P, P, P],
        "fvg_correct_velocity": [P, P, P, P, P, P],
        "fvg_compute_wall_distance": [P, P, P, C.c_int32, C.c_double, P, P, P, P],
        "fvg_identity_create": [P, C.c_int64, C.POINTER(P)],
        "fvg_diagonal_create": [P, P, C.POINTER(P)],
        "fvg_ic_create": [P, P, C.c_int32, C.POINTER(P)],
        "fvg_ilu_create": [P, P, C.c_int32, C.POINTER(P)],
        "fvg_prec_refactor": [P, P, P],
        "fvg_prec_refactor_async": [P, P, P],
        "fvg_prec_status": [P, P],
        "fvg_vec_upload_many": [P, C.c_int32, P, P],
        "fvg_prec_refactor_many": [P, C.c_int32, P, P],
        "fvg_prec_destroy": [P, P],
        "fvg_precondition": [P, P, P, P],
        "fvg_prec_factor_download": [P, P, P, P],
        "fvg_prec_levels": [P, C.POINTER(C.c_int32), C.POINTER(C.c_int32)],
        "fvg_prec_nonpositive_pivots": [P, P, C.POINTER(C.c_int64)],
        "fvg_cg": [P, P, P, P, P, C.POINTER(SolverOptions), C.POINTER(SolverInfo), P, C.c_int64],
        "fvg_bicgstab": [P, P, P, P, P, C.POINTER(SolverOptions), C.POINTER(SolverInfo), P, C.c_int64],
        "fvg_cg_noalloc": [P, P, P, P, P, C.POINTER(SolverOptions), C.POINTER(SolverInfo), P, C.c_int64],
        "fvg_bicgstab_noalloc": [P, P, P, P, P, C.POINTER(SolverOptions), C.POINTER(SolverInfo), P, C.c_int64],
        "fvg_batch_create": [P, C.c_int32, P, P, C.c_int32, C.POINTER(SolverOptions), C.POINTER(P)],
        "fvg_batch_destroy": [P, P],
        "fvg_batch_rebind": [P, P, C.c_int32, P, P],
        "fvg_batch_reserve": [P, P, C.c_int32],
        "fvg_batch_size": [P, C.POINTER(C.c_int32), C.POINTER(C.c_int64)],
        "fvg_batch_launch": [P, P, P, P],
        "fvg_batch_query": [P, P, C.POINTER(C.c_int32)],
        "fvg_batch_poll": [P, P, P, P],
        "fvg_batch_finish": [P, P, P, P, C.c_int64],
        "fvg_batch_solve": [P, P, P, P, P, P, C.c_int64],
        "fvg_batch_set_threads": [P, P, C.c_int32],
        "fvg_batch_get_threads": [P, C.POINTER(C.c_int32)],
        "fvg_batch_set_cluster": [P, P, C.c_int32, C.c_int32],
        "fvg_batch_get_cluster": [P, C.POINTER(C.c_int32), C.POINTER(C.c_int32)],
        "fvg_batch_set_cluster_buckets": [P, P, C.c_int32, P, P],
        "fvg_batch_set_phase_timing": [P, P, C.c_int32],
        "fvg_batch_phase_times": [P, P, P],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    _lib = L
    return L


def check(status):
    if status != FVG_OK:
        msg = lib().fvg_last_error().decode()
        raise _ERRORS.get(status, FvgError)(status, msg)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(P)


def sweep_launch_rule(tune, live_contexts, batches_in_flight, timed_out_before):
    """Launch mode of the co-resident triangular sweeps (1 ordinary, 2 cooperative; 0 for
    arguments out of range): the rule the library applies, evaluated on the host."""
    return lib().fvg_sweep_launch_rule(int(tune), int(live_contexts), int(batches_in_flight),
                                       int(bool(timed_out_before)))


# ---------------------------------------------------------------------------------
# context, vectors
# ---------------------------------------------------------------------------------

class Context:
    """One device, one HIP stream, an optional HBM budget in bytes (0 = unlimited)."""

    def __init__(self, device=0, hbm_budget=0):
        h = P()
        check(lib().fvg_ctx_create(device, int(hbm_budget), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().fvg_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def sync(self):
        check(lib().fvg_ctx_sync(self.h))

    def bytes_in_use(self):
        v = C.c_uint64()
        check(lib().fvg_ctx_bytes_in_use(self.h, C.byref(v)))
        return v.value

    def set_tuning(self, knob, value):
        check(lib().fvg_ctx_set_tuning(self.h, int(knob), int(value)))

    def get_tuning(self, knob):
        v = C.c_int64()
        check(lib().fvg_ctx_get_tuning(self.h, int(knob), C.byref(v)))
        return v.value

    def sweep_launch_state(self):
        """(last co-resident sweep launch: 1 ordinary, 2 cooperative, 0 none yet; batches of
        this context in flight)"""
        mode, busy = C.c_int32(), C.c_int32()
        check(lib().fvg_ctx_sweep_launch_state(self.h, C.byref(mode), C.byref(busy)))
        return mode.value, busy.value

    def enable_timing(self, on=True, every=1):
        """HIP-event timing of SpMV / sweep launch groups; `every`: time one group in this many."""
        check(lib().fvg_ctx_enable_timing(self.h, int(every) if on else 0))

    def kernel_stats(self, reset=True):
        s = KernelStats()
        check(lib().fvg_ctx_kernel_stats(self.h, C.byref(s), 1 if reset else 0))
        return dict(spmv_ms=s.spmv_ms, sweep_ms=s.sweep_ms, spmv_launches=s.spmv_launches,
                    sweep_launches=s.sweep_launches, spmv_calls=s.spmv_calls, sweep_calls=s.sweep_calls)


def refactor_many(ctx, preconditioners, matrices):
    """Refactorise several preconditioners in one launch (checks deferred to status())."""
    n = len(preconditioners)
    if n == 0:
        return
    pp = (P * n)(*[p.h for p in preconditioners])
    mp = (P * n)(*[m.h for m in matrices])
    check(lib().fvg_prec_refactor_many(ctx.h, n, pp, mp))


def upload_many(ctx, vectors, arrays):
    """Upload arrays into vectors through one staging buffer, without waiting for the copies."""
    n = len(vectors)
    if n == 0:
        return
    keep = [np.ascontiguousarray(a, dtype=np.float64) for a in arrays]
    for v, a in zip(vectors, keep):
        if a.shape != (v.n,):
            raise ValueError("array length does not match its vector")
    vp = (P * n)(*[v.h for v in vectors])
    hp = (P * n)(*[_ptr(a) for a in keep])
    check(lib().fvg_vec_upload_many(ctx.h, n, vp, hp))


class DistributedVector:
    """DistributedVector<f64> (src/linalg/dvector.rs:10-179), device resident."""

    def __init__(self, ctx, n):
        self.ctx = ctx
        self.n = int(n)
        h = P()
        check(lib().fvg_vec_create(ctx.h, self.n, C.byref(h)))
        self.h = h

    @classmethod
    def from_size(cls, ctx, n):
        return cls(ctx, n)

    @classmethod
    def from_data(cls, ctx, data):
        data = np.ascontiguousarray(data, dtype=np.float64)
        v = cls(ctx, data.shape[0])
        v.upload(data)
        return v

    def upload(self, data):
        data = np.ascontiguousarray(data, dtype=np.float64)
        assert data.shape == (self.n,)
        check(lib().fvg_vec_upload(self.ctx.h, self.h, _ptr(data)))

    def numpy(self):
        out = np.empty(self.n)
        check(lib().fvg_vec_download(self.ctx.h, self.h, _ptr(out)))
        return out

    def fill(self, value, ctx=None):
        check(lib().fvg_vec_fill((ctx or self.ctx).h, self.h, float(value)))

    def dot(self, other):
        r = C.c_double()
        check(lib().fvg_vec_dot(self.ctx.h, self.h, other.h, C.byref(r)))
        return r.value

    def __len__(self):
        return self.n

    def close(self):
        if getattr(self, "h", None) and self.ctx.h:
            lib().fvg_vec_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        self.close()


# ---------------------------------------------------------------------------------
# mesh
# ---------------------------------------------------------------------------------

class Mesh:
    """Mesh<DIM> (src/core/mesh/mod.rs:200-234) built from raw topology and computed."""

    def __init__(self, ctx, raw, _topology=None):
        self.ctx = ctx
        h = P()
        if _topology is not None:
            self.dim = _topology.dim
            check(lib().fvg_mesh_create_from_topology(ctx.h, _topology.h, C.byref(h)))
            fp = _topology.raw()["face_patch"]
        else:
            self.dim = int(raw["dim"])
            nodes = np.ascontiguousarray(raw["nodes"], dtype=np.float64)
            fptr = np.ascontiguousarray(raw["face_ptr"], dtype=np.int64)
            fn = np.ascontiguousarray(raw["face_nodes"], dtype=np.int32)
            fp = np.ascontiguousarray(raw["face_patch"], dtype=np.int32)
            cptr = np.ascontiguousarray(raw["cell_ptr"], dtype=np.int64)
            cf = np.ascontiguousarray(raw["cell_faces"], dtype=np.int32)
            check(lib().fvg_mesh_create(ctx.h, self.dim, nodes.shape[0], _ptr(nodes), fptr.shape[0] - 1,
                                        _ptr(fptr), _ptr(fn), _ptr(fp), cptr.shape[0] - 1, _ptr(cptr), _ptr(cf),
                                        C.byref(h)))
        self.h = h
        nc, nf, nb, nnz, ncol = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64(), C.c_int32()
        check(lib().fvg_mesh_sizes(h, C.byref(nc), C.byref(nf), C.byref(nb), C.byref(nnz), C.byref(ncol)))
        self.n_cells, self.n_faces, self.n_boundary_faces, self.nnz, self.max_cell_faces = (
            nc.value, nf.value, nb.value, nnz.value, ncol.value)
        self.face_patch = fp
        self.topology = _topology
        self.names = None if raw is None else raw.get("patch_names")
        self._geom = None

    def patch_name(self, patch):
        """The patch's name from the mesh file (its topology) or the raw arrays'
        "patch_names"; None when the mesh carries no names."""
        if self.topology is not None:
            return self.topology.patch_name(patch)
        if self.names is not None:
            return self.names[patch] if isinstance(self.names, (list, tuple)) else self.names.get(patch)
        return None

    @classmethod
    def from_arrays(cls, ctx, raw):
        return cls(ctx, raw)

    @classmethod
    def from_topology(cls, ctx, topology):
        return cls(ctx, None, _topology=topology)

    def geometry(self):
        """face_area, face_center, face_normal, owner, neighbor, cell_volume, cell_center."""
        if self._geom is None:
            F, N, D = self.n_faces, self.n_cells, self.dim
            g = dict(face_area=np.empty(F), face_center=np.empty((F, D)), face_normal=np.empty((F, D)),
                     owner=np.empty(F, np.int32), neighbor=np.empty(F, np.int32), cell_volume=np.empty(N),
                     cell_center=np.empty((N, D)))
            check(lib().fvg_mesh_geometry(self.h, *[_ptr(g[k]) for k in (
                "face_area", "face_center", "face_normal", "owner", "neighbor", "cell_volume", "cell_center")]))
            self._geom = g
        return self._geom

    def boundary_faces(self):
        return np.nonzero(self.face_patch >= 0)[0]

    def close(self):
        if getattr(self, "h", None) and self.ctx.h:
            lib().fvg_mesh_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        self.close()


# ---------------------------------------------------------------------------------
# matrix
# ---------------------------------------------------------------------------------

# ---------------------------------------------------------------------------------
# mesh files and results (src/core/mesh/io.rs, src/post/write_pvtu.rs)
# ---------------------------------------------------------------------------------

class Topology:
    """A host mesh topology: what Mesh::add_node / add_face / add_cell receive."""

    def __init__(self, handle):
        self.h = handle
        d, nn, nf, nfn, nc, ncf, npat = (C.c_int32(), C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64(),
                                         C.c_int64(), C.c_int32())
        check(lib().fvg_topology_sizes(handle, C.byref(d), C.byref(nn), C.byref(nf), C.byref(nfn), C.byref(nc),
                                       C.byref(ncf), C.byref(npat)))
        self.dim, self.n_nodes, self.n_faces, self.n_cells, self.n_patches = d.value, nn.value, nf.value, \
            nc.value, npat.value
        self._sizes = (nfn.value, ncf.value)

    @classmethod
    def read_su2(cls, path):
        """Mesh::read_su2 (io.rs:489-533)."""
        h = P()
        check(lib().fvg_topology_read_su2(os.fsencode(path), C.byref(h)))
        return cls(h)

    @classmethod
    def read(cls, path):
        """Mesh::read (io.rs:68-211): the crate's own mesh format."""
        h = P()
        check(lib().fvg_topology_read_native(os.fsencode(path), C.byref(h)))
        return cls(h)

    @classmethod
    def from_raw(cls, raw):
        nodes = np.ascontiguousarray(raw["nodes"], dtype=np.float64)
        fptr = np.ascontiguousarray(raw["face_ptr"], dtype=np.int64)
        fn = np.ascontiguousarray(raw["face_nodes"], dtype=np.int32)
        fp = np.ascontiguousarray(raw["face_patch"], dtype=np.int32)
        cptr = np.ascontiguousarray(raw["cell_ptr"], dtype=np.int64)
        cf = np.ascontiguousarray(raw["cell_faces"], dtype=np.int32)
        h = P()
        check(lib().fvg_topology_create(int(raw["dim"]), nodes.shape[0], _ptr(nodes), fptr.shape[0] - 1, _ptr(fptr),
                                        _ptr(fn), _ptr(fp), cptr.shape[0] - 1, _ptr(cptr), _ptr(cf), C.byref(h)))
        return cls(h)

    def raw(self):
        """Copies of the arrays, in the raw-mesh layout fvg.Mesh takes."""
        ptrs = [P() for _ in range(6)]
        check(lib().fvg_topology_arrays(self.h, *[C.byref(q) for q in ptrs]))
        nfn, ncf = self._sizes

        def arr(p, ctype, n, dtype):
            if n == 0:
                return np.zeros(0, dtype)
            return np.ctypeslib.as_array(C.cast(p, C.POINTER(ctype)), shape=(n,)).astype(dtype, copy=True)

        return dict(dim=self.dim,
                    nodes=arr(ptrs[0], C.c_double, self.n_nodes * self.dim, np.float64).reshape(-1, self.dim),
                    face_ptr=arr(ptrs[1], C.c_int64, self.n_faces + 1, np.int64),
                    face_nodes=arr(ptrs[2], C.c_int32, nfn, np.int32),
                    face_patch=arr(ptrs[3], C.c_int32, self.n_faces, np.int32),
                    cell_ptr=arr(ptrs[4], C.c_int64, self.n_cells + 1, np.int64),
                    cell_faces=arr(ptrs[5], C.c_int32, ncf, np.int32))

    def patch_name(self, patch):
        v = C.c_char_p()
        check(lib().fvg_topology_patch_name(self.h, int(patch), C.byref(v)))
        return v.value.decode()

    def close(self):
        if getattr(self, "h", None):
            lib().fvg_topology_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class PvtuWriter:
    """PvtuWriter::new(&mesh).with(name, &field)...write(path) (src/post/write_pvtu.rs:30-58).
    Fields are host arrays (n_cells,) or (n_cells, ncomp) for write(), device vectors
    (with their component count) for write_async()."""

    def __init__(self, topology):
        h = P()
        check(lib().fvg_pvtu_writer_create(topology.h, C.byref(h)))
        self.h = h
        self.n_cells = topology.n_cells
        self.fields = []

    def with_(self, name, values):
        v = values if isinstance(values, DistributedVector) else np.ascontiguousarray(values, dtype=np.float64)
        nc = 1 if isinstance(v, DistributedVector) and v.n == self.n_cells else \
            (v.n // self.n_cells if isinstance(v, DistributedVector) else (1 if v.ndim == 1 else v.shape[1]))
        self.fields.append((name, nc, v))
        return self

    def _args(self):
        names = (C.c_char_p * max(1, len(self.fields)))(*[n.encode() for n, _, _ in self.fields])
        nc = np.array([c for _, c, _ in self.fields] or [1], np.int32)
        return names, nc

    def write(self, path):
        names, nc = self._args()
        vals = [np.ascontiguousarray(v.numpy() if isinstance(v, DistributedVector) else v, dtype=np.float64).ravel()
                for _, _, v in self.fields]
        vp = (P * max(1, len(vals)))(*[a.ctypes.data for a in vals])
        check(lib().fvg_pvtu_write(self.h, os.fsencode(path), len(self.fields), C.cast(names, P), _ptr(nc),
                                   C.cast(vp, P)))

    def write_async(self, ctx, path):
        """Copies the device fields out on ctx's stream and writes the files from a host
        thread; returns at once (wait() joins it)."""
        names, nc = self._args()
        for _, _, v in self.fields:
            if not isinstance(v, DistributedVector):
                raise FvgError(1, "write_async takes device vectors")
        self._keep = (names, nc)
        vp = (P * max(1, len(self.fields)))(*[v.h for _, _, v in self.fields])
        self._keep_vp = vp
        check(lib().fvg_pvtu_write_async(self.h, ctx.h, os.fsencode(path), len(self.fields), C.cast(names, P),
                                         _ptr(nc), C.cast(vp, P)))

    def wait(self):
        check(lib().fvg_pvtu_wait(self.h))

    def close(self):
        if getattr(self, "h", None):
            lib().fvg_pvtu_writer_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class DistributedMatrix:
    """DistributedMatrix<f64> in CSR (src/linalg/dmatrix.rs:14-301), device resident."""

    def __init__(self, ctx, handle):
        self.ctx = ctx
        self.h = handle
        nr, nc, nnz = C.c_int64(), C.c_int64(), C.c_int64()
        check(lib().fvg_csr_sizes(handle, C.byref(nr), C.byref(nc), C.byref(nnz)))
        self.nrows, self.ncols, self.nnz = nr.value, nc.value, nnz.value

    @classmethod
    def from_mesh(cls, ctx, mesh):
        """from_cut_sparsity(mesh.cell_to_cell_sparsity(), n_cells) (dmatrix.rs:38-50)."""
        h = P()
        check(lib().fvg_csr_from_mesh(ctx.h, mesh.h, C.byref(h)))
        return cls(ctx, h)

    @classmethod
    def from_csr(cls, ctx, row_ptr, col, values=None, ncols=None):
        rp = np.ascontiguousarray(row_ptr, dtype=np.int32)
        ci = np.ascontiguousarray(col, dtype=np.int32)
        va = None if values is None else np.ascontiguousarray(values, dtype=np.float64)
        n = rp.shape[0] - 1
        ncols = n if ncols is None else ncols
        h = P()
        check(lib().fvg_csr_create(ctx.h, n, ncols, ci.shape[0], _ptr(rp), _ptr(ci), _ptr(va), C.byref(h)))
        return cls(ctx, h)

    def csr(self):
        rp = np.empty(self.nrows + 1, np.int32)
        ci = np.empty(self.nnz, np.int32)
        va = np.empty(self.nnz)
        check(lib().fvg_csr_download(self.ctx.h, self.h, _ptr(rp), _ptr(ci), _ptr(va)))
        return rp, ci, va

    def set_values(self, values):
        va = np.ascontiguousarray(values, dtype=np.float64)
        check(lib().fvg_csr_upload_values(self.ctx.h, self.h, _ptr(va)))

    def set_row_coordinates(self, coords):
        """Coordinates of the rows ((nrows, dim) with dim 1 to 3, e.g. cell centres; None
        drops them) for the layout analysis of preconditioners created afterwards.  A
        matrix made by from_mesh uses its mesh's cell centres already, while the mesh is
        alive."""
        if coords is None:
            check(lib().fvg_csr_set_row_coords(self.h, 0, None))
            return
        xy = np.ascontiguousarray(coords, dtype=np.float64)
        if xy.ndim == 1:
            xy = xy[:, None]
        if xy.shape[0] != self.nrows:
            raise ValueError("one coordinate row per matrix row")
        check(lib().fvg_csr_set_row_coords(self.h, xy.shape[1], _ptr(xy)))

    def imul(self, result, rhs):
        """result = A rhs (dmatrix.rs:136-146)."""
        check(lib().fvg_spmv(self.ctx.h, self.h, rhs.h, result.h))

    def time_spmv(self, result, rhs, reps=50):
        """Average milliseconds of `reps` back-to-back result = A rhs launches (HIP events)."""
        ms = C.c_double()
        check(lib().fvg_spmv_time(self.ctx.h, self.h, rhs.h, result.h, int(reps), C.byref(ms)))
        return ms.value

    def close(self):
        if getattr(self, "h", None) and self.ctx.h:
            lib().fvg_csr_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        self.close()


# ---------------------------------------------------------------------------------
# terms (src/fvm/terms) -> postfix program
# ---------------------------------------------------------------------------------

def _field(a):
    """A field argument: a DistributedVector stays on the device, anything else becomes a
    contiguous float64 host array (uploaded when the equation is assembled)."""
    if isinstance(a, DistributedVector):
        return a
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a if a.ndim == 1 else a.ravel()


class Orthogonal:
    """facengrad::Orthogonal (src/fvm/schemes/facengrad/orthogonal.rs)."""


class Corrected:
    """facengrad::Corrected(&gradients, corrector_reduction)
    (src/fvm/schemes/facengrad/corrected.rs:10-30); gradients is n_cells x dim."""

    def __init__(self, gradients, corrector_reduction):
        self.gradients = _field(gradients)
        self.reduction = float(corrector_reduction)


class Upwind:
    """faceinterp::Upwind(&flux) (src/fvm/schemes/faceinterp/upwind.rs)."""

    def __init__(self, flux):
        self.flux = _field(flux)


class Linear:
    """faceinterp::Linear (src/fvm/schemes/faceinterp/linear.rs)."""


class LimitedLinear:
    """faceinterp::LimitedLinear(&flux, &limiters) (src/fvm/schemes/faceinterp/limitedlinear.rs)."""

    def __init__(self, flux, limiters):
        self.flux = _field(flux)
        self.limiters = _field(limiters)


class _Limiter:
    def __init__(self, kind, k=1.0):
        self.kind, self.k = kind, float(k)


class limiters:
    """schemes::limiters (src/fvm/schemes/limiters/mod.rs)."""

    @staticmethod
    def LimitedLinear(k):
        return _Limiter(LIMITER_LIMITED_LINEAR, k)

    MinMod = staticmethod(lambda: _Limiter(LIMITER_MINMOD))
    VanLeer = staticmethod(lambda: _Limiter(LIMITER_VANLEER))
    VanAlbada = staticmethod(lambda: _Limiter(LIMITER_VANALBADA))

    @staticmethod
    def Sweby(beta):
        return _Limiter(LIMITER_SWEBY, beta)

    SuperBee = staticmethod(lambda: _Limiter(LIMITER_SUPERBEE))


class Euler:
    """time::Euler (src/fvm/schemes/time/euler.rs)."""

    def __init__(self, previous, dt, density=None, density_previous=None):
        self.previous = _field(previous)
        self.dt = float(dt)
        self.density = density
        self.density_previous = density_previous


class Backward:
    """time::Backward (src/fvm/schemes/time/backward.rs:21-35): second-order backward
    differencing from the two previous levels, dt the step now and dt_2 the one before.
    With densities it is Backward::new_with_density (density, density_last,
    density_last2 per cell)."""

    def __init__(self, previous, previous_2, dt, dt_2, density=None, density_last=None, density_last2=None):
        self.previous = _field(previous)
        self.previous_2 = _field(previous_2)
        self.dt, self.dt_2 = float(dt), float(dt_2)
        dens = (density, density_last, density_last2)
        if any(d is not None for d in dens) and any(d is None for d in dens):
            raise FvgError(1, "Backward needs all three densities or none")
        self.densities = None if density is None else tuple(_field(d) for d in dens)


class Term:
    """An expression tree of reference terms; +, - and unary - build TermsBinaryOp / OpNeg."""

    def __init__(self, kind, **kw):
        self.kind = kind
        self.kw = kw

    def __add__(self, other):
        return Term("add", a=self, b=other)

    def __sub__(self, other):
        return Term("sub", a=self, b=other)

    def __neg__(self):
        return Term("neg", a=self)

    def compile(self):
        """Returns (ops, face_fields, cell_fields) with fields deduplicated by identity."""
        ops, ff, cf = [], [], []

        def fidx(lst, arr):
            for i, a in enumerate(lst):
                if a is arr:
                    return i
            lst.append(arr)
            return len(lst) - 1

        def consecutive(lst, arrs):   # fields an opcode reads as a run of slots
            lst.extend(arrs)
            return len(lst) - len(arrs)

        def walk(t):
            k = t.kind
            if k in ("add", "sub"):
                walk(t.kw["a"])
                walk(t.kw["b"])
                ops.append((OP_ADD if k == "add" else OP_SUB, 0, 0, 0, 0.0))
            elif k == "neg":
                walk(t.kw["a"])
                ops.append((OP_NEG, 0, 0, 0, 0.0))
            elif k == "laplacian":
                sc = t.kw["scheme"]
                if isinstance(sc, Corrected):
                    ops.append((OP_LAPLACIAN_CORRECTED, fidx(ff, t.kw["k"]), fidx(cf, sc.gradients), 0,
                                sc.reduction))
                else:
                    ops.append((OP_LAPLACIAN_ORTHOGONAL, fidx(ff, t.kw["k"]), 0, 0, 0.0))
            elif k == "convection":
                sc = t.kw["scheme"]
                if isinstance(sc, Upwind):
                    ops.append((OP_CONVECTION_UPWIND, fidx(ff, t.kw["flux"]), fidx(ff, sc.flux), 0, 0.0))
                elif isinstance(sc, Linear):
                    ops.append((OP_CONVECTION_LINEAR, fidx(ff, t.kw["flux"]), 0, 0, 0.0))
                else:
                    ops.append((OP_CONVECTION_LIMITED_LINEAR, fidx(ff, t.kw["flux"]), fidx(ff, sc.flux),
                                fidx(ff, sc.limiters), 0.0))
            elif k == "divergence":
                ops.append((OP_DIVERGENCE, fidx(ff, t.kw["flux"]), 0, 0, 0.0))
            elif k == "source":
                ops.append((OP_SOURCE, fidx(cf, t.kw["f"]), 0, 0, 0.0))
            elif k == "linear_source":
                ops.append((OP_LINEAR_SOURCE, fidx(cf, t.kw["c"]), 0, 0, 0.0))
            elif k == "time" and isinstance(t.kw["scheme"], Backward):
                e = t.kw["scheme"]
                a0, a1 = fidx(cf, e.previous), fidx(cf, e.previous_2)
                d = -1 if e.densities is None else consecutive(cf, e.densities)
                ops.append((OP_TIME_BACKWARD, a0, a1, d, e.dt))
                ops.append((OP_PARAM, 0, 0, 0, e.dt_2))
            elif k == "user":
                a0 = -1 if t.kw["face_terms"] is None else consecutive(ff, t.kw["face_terms"])
                a1 = -1 if t.kw["cell_terms"] is None else consecutive(cf, t.kw["cell_terms"])
                ops.append((OP_USER_TERM, a0, a1, 0, 0.0))
            elif k == "time":
                e = t.kw["scheme"]
                a1 = a2 = -1
                if e.density is not None:
                    a1 = fidx(cf, np.ascontiguousarray(e.density, dtype=np.float64))
                    a2 = fidx(cf, np.ascontiguousarray(e.density_previous, dtype=np.float64))
                ops.append((OP_TIME_EULER, fidx(cf, e.previous), a1, a2, e.dt))
            else:
                raise ValueError(k)

        walk(self)
        return ops, ff, cf


def laplacian(scheme, diffusivity):
    """laplacian(scheme, &k_f) (src/fvm/terms/laplacian.rs:31-46); Orthogonal or Corrected."""
    if not isinstance(scheme, (Orthogonal, Corrected)):
        raise FvgError(9, "face-normal-gradient schemes: Orthogonal, Corrected")
    return Term("laplacian", k=_field(diffusivity), scheme=scheme)


def convection(scheme, flux):
    """convection(scheme, &flux) (src/fvm/terms/convection.rs:32-45); Upwind, Linear or
    LimitedLinear."""
    if not isinstance(scheme, (Upwind, Linear, LimitedLinear)):
        raise FvgError(9, "face-interpolation schemes: Upwind, Linear, LimitedLinear")
    return Term("convection", flux=_field(flux), scheme=scheme)


def divergence(flux):
    return Term("divergence", flux=np.ascontiguousarray(flux, dtype=np.float64))


def source(f):
    """source(&f) (src/fvm/terms/source.rs:30-39)."""
    return Term("source", f=np.ascontiguousarray(f, dtype=np.float64))


def linear_source(c):
    return Term("linear_source", c=np.ascontiguousarray(c, dtype=np.float64))


def time(scheme):
    """time(scheme) (src/fvm/terms/time.rs:19-23); Euler or Backward."""
    if not isinstance(scheme, (Euler, Backward)):
        raise FvgError(9, "time schemes: Euler, Backward")
    return Term("time", scheme=scheme)


def user_term(face_terms=None, cell_terms=None):
    """Any other `impl Term` (src/fvm/terms/mod.rs:24-40) whose values the caller has
    evaluated: face_terms = (li, lj, r) with one entry per face in face order (what
    Term::face_terms returns for that face, before the area scaling of assembly.rs),
    cell_terms = (l, r) with one entry per cell (Term::cell_terms, before the volume
    scaling).  Either may be None: the trait's zero default.  Each array may be a
    DistributedVector, which stays on the device."""
    if face_terms is not None:
        if len(face_terms) != 3:
            raise FvgError(1, "face_terms is (li, lj, r)")
        face_terms = tuple(_field(a) for a in face_terms)
    if cell_terms is not None:
        if len(cell_terms) != 2:
            raise FvgError(1, "cell_terms is (l, r)")
        cell_terms = tuple(_field(a) for a in cell_terms)
    return Term("user", face_terms=face_terms, cell_terms=cell_terms)


class FaceConstraints:
    """Boundary closure per boundary face: value_j = bndi * value_i + bnde (src/fvm/bcs/mod.rs)."""

    def __init__(self, bndi, bnde):
        self.bndi = np.ascontiguousarray(bndi, dtype=np.float64)
        self.bnde = np.ascontiguousarray(bnde, dtype=np.float64)

    def as_bc(self):
        return self


def dirichlet(values):
    v = np.ascontiguousarray(values, dtype=np.float64)
    return FaceConstraints(np.zeros_like(v), v)


def assemble(equation, boundary_condition, mesh, lhs=None, rhs=None):
    """fvm::assembly::assemble (src/fvm/assembly.rs:11-79) on the GPU.

    boundary_condition: FaceConstraints, a (bndi, bnde) pair of arrays over the
    boundary faces in face order, or a callable taking the mesh geometry dict and
    the boundary face ids and returning such a pair.  lhs/rhs may be passed to reuse
    device storage from an earlier assembly on the same mesh.
    """
    ops, ff, cf = equation.compile()
    bfaces = mesh.boundary_faces()
    bc = boundary_condition
    if callable(bc) and not isinstance(bc, FaceConstraints):
        bc = bc(mesh.geometry(), bfaces)
    if not isinstance(bc, FaceConstraints):
        bc = FaceConstraints(*bc)
    return assemble_program(mesh, ops, ff, cf, bc.bndi, bc.bnde, lhs, rhs)


class Assembler:
    """A compiled equation with its fields and boundary closure resident on the device;
    run() re-assembles into the given lhs/rhs without host transfers."""

    def __init__(self, mesh, ops, face_fields, cell_fields, bndi, bnde):
        ctx = mesh.ctx
        self.mesh = mesh
        nb = mesh.n_boundary_faces
        bndi = np.ascontiguousarray(bndi, dtype=np.float64)
        bnde = np.ascontiguousarray(bnde, dtype=np.float64)
        if bndi.shape != (nb,) or bnde.shape != (nb,):
            raise FvgError(1, "boundary closure arrays must have one entry per boundary face")
        dev = lambda a: a if isinstance(a, DistributedVector) else DistributedVector.from_data(ctx, a)
        self.fvecs = [dev(a) for a in face_fields]
        self.cvecs = [dev(a) for a in cell_fields]
        self.bi = DistributedVector.from_data(ctx, bndi) if nb else None
        self.be = DistributedVector.from_data(ctx, bnde) if nb else None
        self.ops = list(ops)
        self._prog = (TermOp * len(self.ops))(*[TermOp(*o) for o in self.ops])
        self._fv = (P * max(1, len(self.fvecs)))(*[v.h for v in self.fvecs])
        self._cv = (P * max(1, len(self.cvecs)))(*[v.h for v in self.cvecs])

    def run(self, lhs, rhs, sync=True, ctx=None):
        """Assemble on ctx's stream (the mesh's context by default)."""
        m = self.mesh
        c = ctx or m.ctx
        check(lib().fvg_assemble(c.h, m.h, C.cast(self._prog, P), len(self.ops), C.cast(self._fv, P),
                                 len(self.fvecs), C.cast(self._cv, P), len(self.cvecs),
                                 self.bi.h if self.bi else None, self.be.h if self.be else None, lhs.h, rhs.h))
        if sync:
            c.sync()


def assemble_program(mesh, ops, face_fields, cell_fields, bndi, bnde, lhs=None, rhs=None):
    """assemble() with the equation already in postfix form: ops is a list of
    (opcode, arg0, arg1, arg2, scalar) over the given face and cell field arrays."""
    asm = Assembler(mesh, ops, face_fields, cell_fields, bndi, bnde)
    if lhs is None:
        lhs = DistributedMatrix.from_mesh(mesh.ctx, mesh)
    if rhs is None:
        rhs = DistributedVector.from_size(mesh.ctx, mesh.n_cells)
    asm.run(lhs, rhs)
    return lhs, rhs


# ---------------------------------------------------------------------------------
# fvm tools (src/fvm/tools)
# ---------------------------------------------------------------------------------

def _closure_vecs(mesh, boundary_condition, ncomp):
    """The boundary closure as device vectors: bc_i per boundary face, bc_e per boundary
    face and component (cached on the FaceConstraints object)."""
    bc = boundary_condition
    if not isinstance(bc, FaceConstraints):
        bc = FaceConstraints(*bc)
    nb = mesh.n_boundary_faces
    if nb == 0:
        return None, None
    key = ("_dev", id(mesh.ctx), ncomp)
    cache = bc.__dict__.setdefault("_dev", {})
    if key not in cache:
        if bc.bndi.shape != (nb,) or bc.bnde.size != nb * ncomp:
            raise FvgError(1, "boundary closure: bc_i per boundary face, bc_e per boundary face and component")
        cache[key] = (DistributedVector.from_data(mesh.ctx, bc.bndi),
                      DistributedVector.from_data(mesh.ctx, bc.bnde.ravel()))
    return cache[key]


def compute_gradients(gradients, values, boundary_condition, mesh, ncomp=1):
    """compute_gradients(&mut gradients, &values, bc, &mesh) (tools/gradients.rs:41-79) on
    the device: values is n_cells x ncomp, gradients n_cells x ncomp x dim."""
    bi, be = _closure_vecs(mesh, boundary_condition, ncomp)
    check(lib().fvg_compute_gradients(mesh.ctx.h, mesh.h, int(ncomp), values.h, bi.h if bi else None,
                                      be.h if be else None, gradients.h))


def compute_limiters(limiters_out, values, gradients, limiter_scheme, fgradinterp_scheme, fngrad_scheme,
                     boundary_condition, mesh, ncomp=1):
    """compute_limiters(&mut limiters, &values, &gradients, limiter, fgradinterp, fngrad,
    bc, &mesh) (tools/limiters.rs:47-126) on the device.  The gradient interpolation is
    Linear or Upwind(&flux); the face-normal gradient Orthogonal or Corrected."""
    ctx = mesh.ctx
    keep = []

    def dev(a):
        if a is None or isinstance(a, DistributedVector):
            return a
        v = DistributedVector.from_data(ctx, np.ascontiguousarray(a, dtype=np.float64).ravel())
        keep.append(v)
        return v

    spec = LimiterSpec()
    spec.limiter, spec.k = limiter_scheme.kind, limiter_scheme.k
    if isinstance(fgradinterp_scheme, Upwind):
        spec.interp, spec.flux = INTERP_UPWIND, dev(fgradinterp_scheme.flux).h
    elif isinstance(fgradinterp_scheme, Linear):
        spec.interp = INTERP_LINEAR
    else:
        raise FvgError(9, "gradient interpolation schemes: Linear, Upwind")
    if isinstance(fngrad_scheme, Corrected):
        spec.fngrad, spec.reduction, spec.fn_grads = FNGRAD_CORRECTED, fngrad_scheme.reduction, \
            dev(fngrad_scheme.gradients).h
    elif isinstance(fngrad_scheme, Orthogonal):
        spec.fngrad = FNGRAD_ORTHOGONAL
    else:
        raise FvgError(9, "face-normal-gradient schemes: Orthogonal, Corrected")
    bi, be = _closure_vecs(mesh, boundary_condition, ncomp)
    check(lib().fvg_compute_limiters(ctx.h, mesh.h, int(ncomp), values.h, gradients.h, C.byref(spec),
                                     bi.h if bi else None, be.h if be else None, limiters_out.h))
    if keep:
        ctx.sync()


# ---- SIMPLE helpers (src/fvm/tools/hbya.rs) ----

def _interp_spec(scheme, dev):
    spec = InterpSpec()
    if isinstance(scheme, Linear):
        spec.interp = INTERP_LINEAR
    elif isinstance(scheme, Upwind):
        spec.interp, spec.flux = INTERP_UPWIND, dev(scheme.flux).h
    elif isinstance(scheme, LimitedLinear):
        spec.interp, spec.flux, spec.limiters = INTERP_LIMITED_LINEAR, dev(scheme.flux).h, dev(scheme.limiters).h
    else:
        raise FvgError(9, "face interpolation schemes: Linear, Upwind, LimitedLinear")
    return spec


def _devs(ctx):
    keep = []

    def dev(a):
        if a is None or isinstance(a, DistributedVector):
            return a
        v = DistributedVector.from_data(ctx, np.ascontiguousarray(a, dtype=np.float64).ravel())
        keep.append(v)
        return v

    return dev, keep


def compute_hbya_ainv(hbya, ainv, velocity, mlhs, mrhs, mesh):
    """compute_hbya_ainv(&mut hbya, &mut ainv, &velocity, &mlhs, &mrhs, &mesh)
    (tools/hbya.rs:15-50) on the device.  hbya, velocity and mrhs are Vector<dim> cell
    fields (n_cells x dim); mlhs is the momentum matrix, one row per cell."""
    ctx = mesh.ctx
    dev, keep = _devs(ctx)
    check(lib().fvg_compute_hbya_ainv(ctx.h, mesh.h, mlhs.h, dev(mrhs).h, dev(velocity).h, hbya.h, ainv.h))
    if keep:
        ctx.sync()


def intepolate_hbya_ainv_faces(hbyan_face, ainv_face, hbya, ainv, velocity_bc, pressure_bc, hbya_interp_scheme,
                               ainv_interp_scheme, mesh):
    """intepolate_hbya_ainv_faces (tools/hbya.rs:56-102, the reference's spelling) on the
    device.  velocity_bc: FaceConstraints with bndi per boundary face and bnde per boundary
    face and component; pressure_bc: FaceConstraints of the pressure."""
    ctx = mesh.ctx
    dev, keep = _devs(ctx)
    hs = _interp_spec(hbya_interp_scheme, dev)
    a_s = _interp_spec(ainv_interp_scheme, dev)
    ubi, ube = _closure_vecs(mesh, velocity_bc, mesh.dim)
    pbi, pbe = _closure_vecs(mesh, pressure_bc, 1)
    h = lambda v: v.h if v is not None else None
    check(lib().fvg_interpolate_hbya_ainv_faces(ctx.h, mesh.h, dev(hbya).h, dev(ainv).h, C.byref(hs), C.byref(a_s),
                                                h(ubi), h(ube), h(pbi), h(pbe), hbyan_face.h, ainv_face.h))
    if keep:
        ctx.sync()


interpolate_hbya_ainv_faces = intepolate_hbya_ainv_faces


def correct_phi(phi, hbyan_face, ainv_face, pressure, fngrad_scheme, pressure_bc, mesh):
    """correct_phi (tools/hbya.rs:108-157) on the device; fngrad_scheme Orthogonal or
    Corrected(&pressure_grad, reduction)."""
    ctx = mesh.ctx
    dev, keep = _devs(ctx)
    if isinstance(fngrad_scheme, Corrected):
        kind, grads, red = FNGRAD_CORRECTED, dev(fngrad_scheme.gradients).h, fngrad_scheme.reduction
    elif isinstance(fngrad_scheme, Orthogonal):
        kind, grads, red = FNGRAD_ORTHOGONAL, None, 0.0
    else:
        raise FvgError(9, "face-normal-gradient schemes: Orthogonal, Corrected")
    pbi, pbe = _closure_vecs(mesh, pressure_bc, 1)
    h = lambda v: v.h if v is not None else None
    check(lib().fvg_correct_phi(ctx.h, mesh.h, dev(hbyan_face).h, dev(ainv_face).h, dev(pressure).h, kind, grads,
                                red, h(pbi), h(pbe), phi.h))
    if keep:
        ctx.sync()


def correct_velocity(velocity, hbya, ainv, pressure_gradient, mesh):
    """correct_velocity (tools/hbya.rs:161-179) on the device."""
    ctx = mesh.ctx
    dev, keep = _devs(ctx)
    check(lib().fvg_correct_velocity(ctx.h, mesh.h, dev(hbya).h, dev(ainv).h, dev(pressure_gradient).h,
                                     velocity.h))
    if keep:
        ctx.sync()


def compute_wall_distance(wall_distance, mesh, wall_patches, max_iterations, tolerance, solver_options=None):
    """compute_wall_distance(&mut wall_distance, &mesh, wall_patches, max_iterations,
    tolerance, solver_options) (tools/wall_distance.rs:8-119) on the device.
    wall_patches: patch ids, or None for every patch whose name contains "wall" and not
    "slip" (wall_distance.rs:17-27; the mesh must carry patch names).  wall_distance is a
    DistributedVector holding the initial guess and, on return, the distance.  Returns
    (number of solves, LinearSolverInfo of the last one)."""
    if wall_patches is None:
        ids = np.unique(mesh.face_patch[mesh.face_patch >= 0])
        names = {int(p): mesh.patch_name(int(p)) for p in ids}
        if any(v is None for v in names.values()):
            raise FvgError(1, "the mesh carries no patch names: pass wall_patches")
        wall_patches = [p for p, name in names.items() if "wall" in name and "slip" not in name]
    bpatch = mesh.face_patch[mesh.face_patch >= 0]
    wall = np.ascontiguousarray(np.isin(bpatch, np.asarray(list(wall_patches), dtype=np.int64)), dtype=np.uint8)
    opts = (solver_options or LinearSolverOptions()).c()
    info = SolverInfo()
    n = C.c_int32()
    check(lib().fvg_compute_wall_distance(mesh.ctx.h, mesh.h, _ptr(wall) if wall.size else None,
                                          int(max_iterations), float(tolerance), C.byref(opts), wall_distance.h,
                                          C.byref(info), C.byref(n)))
    return n.value, LinearSolverInfo("ConjugateGradient", info.iterations, info.initial_residual,
                                     info.final_residual)


# ---------------------------------------------------------------------------------
# preconditioners
# ---------------------------------------------------------------------------------

class Preconditioner:
    def __init__(self, ctx, handle, n):
        self.ctx = ctx
        self.h = handle
        self.n = n

    def precondition(self, solution, rhs):
        check(lib().fvg_precondition(self.ctx.h, self.h, solution.h, rhs.h))

    def sweep_layout(self):
        """Shape of the blocked sweep layouts: lower_blocks, upper_blocks (0: not built),
        levels, lower_rows / lower_fetch (most rows of a block in one level, most values it
        fetches from outside for one level), lower_fetch_total, upper_rows, upper_fetch,
        lower_slots / upper_slots (dependency slots per row: 4 or 8), mirrored (1: the upper
        layout is the lower one reversed; FP64 solves then iterate in position order),
        block_levels (levels of the blocked lower layout, more than `levels` where it delays
        dependencies between blocks; 0: not built)."""
        info = np.zeros(12, dtype=np.int32)
        check(lib().fvg_prec_sweep_layout(self.h, info.ctypes.data, 12))
        keys = ("lower_blocks", "upper_blocks", "levels", "lower_rows", "lower_fetch", "lower_fetch_total",
                "upper_rows", "upper_fetch", "lower_slots", "upper_slots", "mirrored", "block_levels")
        return dict(zip(keys, (int(v) for v in info)))

    def refactor_async(self, matrix):
        """Enqueue the refactorisation; status() reports pivot / symmetry failures later."""
        check(lib().fvg_prec_refactor_async(self.ctx.h, self.h, matrix.h))

    def status(self, ctx=None):
        """Raise the factorisation's failure, if any, of the last refactor_async (reads the
        flags on ctx's stream, the preconditioner's context by default)."""
        check(lib().fvg_prec_status((ctx or self.ctx).h, self.h))

    def refactor(self, matrix):
        check(lib().fvg_prec_refactor(self.ctx.h, self.h, matrix.h))
        self.ctx.sync()

    def factor_pattern(self):
        """(row_ptr, col) of the factor pattern: the matrix's at level 0, with the fill above."""
        n, nnz = C.c_int64(), C.c_int64()
        check(lib().fvg_prec_factor_pattern(self.h, C.byref(n), C.byref(nnz), None, None))
        rp = np.empty(n.value + 1, np.int32)
        ci = np.empty(nnz.value, np.int32)
        check(lib().fvg_prec_factor_pattern(self.h, None, None, _ptr(rp), _ptr(ci)))
        return rp, ci

    def factor_values(self):
        nnz = C.c_int64()
        check(lib().fvg_prec_factor_pattern(self.h, None, C.byref(nnz), None, None))
        v = np.empty(nnz.value)
        d = np.empty(self.n)
        check(lib().fvg_prec_factor_download(self.ctx.h, self.h, _ptr(v), _ptr(d)))
        return v, d

    def levels(self):
        lo, up = C.c_int32(), C.c_int32()
        check(lib().fvg_prec_levels(self.h, C.byref(lo), C.byref(up)))
        return lo.value, up.value

    def nonpositive_pivots(self):
        c = C.c_int64()
        check(lib().fvg_prec_nonpositive_pivots(self.ctx.h, self.h, C.byref(c)))
        return c.value

    def close(self):
        if getattr(self, "h", None) and self.ctx.h:
            lib().fvg_prec_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        self.close()


class IdentityPreconditioner(Preconditioner):
    @classmethod
    def new(cls, ctx, n):
        h = P()
        check(lib().fvg_identity_create(ctx.h, n, C.byref(h)))
        return cls(ctx, h, n)


class DiagonalPreconditioner(Preconditioner):
    @classmethod
    def from_matrix(cls, matrix):
        h = P()
        check(lib().fvg_diagonal_create(matrix.ctx.h, matrix.h, C.byref(h)))
        return cls(matrix.ctx, h, matrix.nrows)


class IncompleteCholesky(Preconditioner):
    """IncompleteCholesky::from_matrix(&A, level) (incomplete_cholesky.rs:205-211)."""

    @classmethod
    def from_matrix(cls, matrix, level=0):
        h = P()
        check(lib().fvg_ic_create(matrix.ctx.h, matrix.h, int(level), C.byref(h)))
        p = cls(matrix.ctx, h, matrix.nrows)
        p._nnz = matrix.nnz
        return p


class IncompleteLowerUpper(Preconditioner):
    """IncompleteLowerUpper::from_matrix(&A, level) (incomplete_lower_upper.rs:266-282)."""

    @classmethod
    def from_matrix(cls, matrix, level=0):
        h = P()
        check(lib().fvg_ilu_create(matrix.ctx.h, matrix.h, int(level), C.byref(h)))
        p = cls(matrix.ctx, h, matrix.nrows)
        p._nnz = matrix.nnz
        return p


# ---------------------------------------------------------------------------------
# solvers
# ---------------------------------------------------------------------------------

@dataclass
class LinearSolverOptions:
    """LinearSolverOptions (src/linalg/solvers/mod.rs:20-36)."""
    relative_tolerance: float = 1e-6
    absolute_tolerance: float = 1e-6
    max_iterations: int = 1000
    precision: int = PRECISION_FP64
    inner_tolerance: float = 0.0

    def c(self):
        return SolverOptions(self.relative_tolerance, self.absolute_tolerance, self.max_iterations, self.precision,
                             0, self.inner_tolerance)


@dataclass
class LinearSolverInfo:
    """LinearSolverInfo (src/linalg/solvers/mod.rs:11-17)."""
    solver_identifier: str
    iterations: int
    initial_residual: float
    final_residual: float
    history: np.ndarray = field(default_factory=lambda: np.zeros(0))
    restarts: int = 0
    outer_iterations: int = 0
    true_residual: float = 0.0
    status: int = 0
    flags: int = 0

    def __str__(self):
        return (f"(solver={self.solver_identifier}, iterations={self.iterations}, "
                f"initial_residual={self.initial_residual:.3e}, final_residual={self.final_residual:.3e})")


def _solve(fn, name, solution, lhs, rhs, preconditioner, options):
    opts = options.c()
    info = SolverInfo()
    cap = max(2, options.max_iterations + 2)
    hist = np.zeros(cap)
    check(fn(lhs.ctx.h, lhs.h, preconditioner.h, rhs.h, solution.h, C.byref(opts), C.byref(info), _ptr(hist), cap))
    hl = min(info.history_len, cap)
    return LinearSolverInfo(name, info.iterations, info.initial_residual, info.final_residual, hist[:hl].copy(),
                            info.restarts, info.outer_iterations, info.true_residual, info.status, info.flags)


def conjugate_gradient(solution, lhs, rhs, preconditioner, options=None):
    """conjugate_gradient (src/linalg/solvers/conjugate_gradient.rs:12-112)."""
    return _solve(lib().fvg_cg, "ConjugateGradient", solution, lhs, rhs, preconditioner,
                  options or LinearSolverOptions())


def conjugate_gradient_noalloc(solution, lhs, rhs, preconditioner, options=None):
    """conjugate_gradient in FP64 with every device buffer in the library's static scratch
    array (fvg_cg_noalloc): nothing is allocated."""
    return _solve(lib().fvg_cg_noalloc, "ConjugateGradient", solution, lhs, rhs, preconditioner,
                  options or LinearSolverOptions())


def bi_conjugate_gradient_stab_noalloc(solution, lhs, rhs, preconditioner, options=None):
    """bi_conjugate_gradient_stab in FP64 with every device buffer in the library's static
    scratch array (fvg_bicgstab_noalloc): nothing is allocated."""
    return _solve(lib().fvg_bicgstab_noalloc, "BiConjugateGradientStab", solution, lhs, rhs, preconditioner,
                  options or LinearSolverOptions())


def noalloc_scratch_bytes():
    v = C.c_uint64()
    check(lib().fvg_noalloc_scratch_bytes(C.byref(v)))
    return v.value


def bi_conjugate_gradient_stab(solution, lhs, rhs, preconditioner, options=None):
    """bi_conjugate_gradient_stab (src/linalg/solvers/bi_conjugate_gradient_stab.rs:13-179)."""
    return _solve(lib().fvg_bicgstab, "BiConjugateGradientStab", solution, lhs, rhs, preconditioner,
                  options or LinearSolverOptions())


# ---------------------------------------------------------------------------------
# batched solves (no reference counterpart)
# ---------------------------------------------------------------------------------

SOLVER_CG = 0
SOLVER_BICGSTAB = 1


class Batch:
    """Many independent systems solved by one launch.

    ``matrices[k]`` / ``preconditioners[k]`` stay owned by the caller and must outlive
    the batch.  ``solver`` is ``"cg"`` or ``"bicgstab"``; ``options.precision`` selects the
    FP64 reference algorithm (one workgroup per system) or mixed precision (each system
    on ceil(n / rows_per_group) workgroups, see set_cluster).  Each system's result does
    not depend on the rest of the batch.
    """

    def __init__(self, ctx, matrices, preconditioners, solver="cg", options=None):
        if len(matrices) != len(preconditioners) or not matrices:
            raise ValueError("one preconditioner per matrix, at least one system")
        self.ctx = ctx
        self.options = options or LinearSolverOptions()
        self.solver = SOLVER_CG if solver == "cg" else SOLVER_BICGSTAB
        self.n = len(matrices)
        self._keep = (list(matrices), list(preconditioners))
        am = (P * self.n)(*[m.h for m in matrices])
        pm = (P * self.n)(*[p.h for p in preconditioners])
        h = P()
        opts = self.options.c()
        check(lib().fvg_batch_create(ctx.h, self.n, am, pm, self.solver, C.byref(opts), C.byref(h)))
        self.h = h
        self._inflight = None
        self.rows_per_group = 4096 if self.options.precision == PRECISION_MIXED else 0
        self.max_workers = 0
        self.buckets = ([], [])   # size buckets of set_cluster_buckets

    def set_cluster(self, rows_per_group=4096, max_workers=0):
        """Mixed precision: workgroups per system = ceil(n / rows_per_group) (0: one per
        system); max_workers bounds the resident workgroups of a launch (0: all that fit)."""
        check(lib().fvg_batch_set_cluster(self.ctx.h, self.h, int(rows_per_group), int(max_workers)))
        self.rows_per_group = int(rows_per_group) if self.options.precision == PRECISION_MIXED else 0
        self.max_workers = int(max_workers)

    def set_cluster_buckets(self, max_rows, rows_per_group):
        """Mixed precision: a system of at most max_rows[i] rows (ascending) runs on
        ceil(n / rows_per_group[i]) workgroups; larger systems keep set_cluster's rows.
        Empty lists remove the buckets."""
        lim = np.ascontiguousarray(max_rows, dtype=np.int64)
        rows = np.ascontiguousarray(rows_per_group, dtype=np.int32)
        if lim.size != rows.size:
            raise ValueError("one rows_per_group per bucket")
        check(lib().fvg_batch_set_cluster_buckets(self.ctx.h, self.h, int(lim.size), _ptr(lim), _ptr(rows)))
        self.buckets = (lim.tolist(), rows.tolist())

    @property
    def workers(self):
        """Workgroups a mixed-precision launch keeps resident (0 for one workgroup per system)."""
        r, w = C.c_int32(), C.c_int32()
        check(lib().fvg_batch_get_cluster(self.h, C.byref(r), C.byref(w)))
        return w.value

    def reserve(self, capacity):
        """Size the per-system storage for up to `capacity` systems now."""
        check(lib().fvg_batch_reserve(self.ctx.h, self.h, int(capacity)))

    def rebind(self, matrices, preconditioners):
        """Point this (finished) batch at another set of systems of the same solver."""
        if len(matrices) != len(preconditioners) or not matrices:
            raise ValueError("one preconditioner per matrix, at least one system")
        n = len(matrices)
        am = (P * n)(*[m.h for m in matrices])
        pm = (P * n)(*[p.h for p in preconditioners])
        check(lib().fvg_batch_rebind(self.ctx.h, self.h, n, am, pm))
        self.n = n
        self._keep = (list(matrices), list(preconditioners))

    @property
    def threads(self):
        t = C.c_int32()
        check(lib().fvg_batch_get_threads(self.h, C.byref(t)))
        return t.value

    @threads.setter
    def threads(self, value):
        check(lib().fvg_batch_set_threads(self.ctx.h, self.h, int(value)))

    def _time_solve(self, rhs, solutions, repeats):
        import time as _time
        best = None
        for _ in range(repeats):
            for x in solutions:
                x.fill(0.0)
            self.ctx.sync()
            t = _time.perf_counter()
            self.solve(rhs, solutions, history=False)
            el = _time.perf_counter() - t
            best = el if best is None else min(best, el)
        return best

    def tune(self, rhs, solutions, candidates=None, repeats=1):
        """Autotune the launch shape: time a solve of this batch from x = 0 with each
        candidate and keep the fastest.  One workgroup per system (FP64): the workgroup
        size (rows per tile = threads, one row per thread; candidates 1024, 512, 256).
        Mixed precision: rows per workgroup of each system's cluster (candidates 2048,
        4096, 8192, 16384).  Overwrites the solutions.  Returns {candidate: seconds}."""
        times = {}
        if self.rows_per_group > 0:
            for rows in candidates or (2048, 4096, 8192, 16384):
                self.set_cluster(rows, self.max_workers)
                times[rows] = self._time_solve(rhs, solutions, repeats)
            self.set_cluster(min(times, key=times.get), self.max_workers)
            return times
        for th in candidates or (1024, 512, 256):
            self.threads = th
            times[th] = self._time_solve(rhs, solutions, repeats)
        self.threads = min(times, key=times.get)
        return times

    def tune_buckets(self, rhs, solutions, limits=(8192, 32768), candidates=(2048, 4096, 8192, 16384), repeats=1):
        """Mixed precision: autotune the rows per workgroup per system size bucket.  The
        systems of each bucket (at most limits[0] rows, at most limits[1], ..., the rest)
        are timed as a batch of their own with each candidate, from x = 0, and the fastest
        is kept for the bucket; the last bucket's choice becomes the batch's rows_per_group.
        Overwrites the solutions.  Returns {bucket upper limit (None: the rest): {candidate:
        seconds}}."""
        if self.rows_per_group == 0:
            raise ValueError("bucket tuning applies to mixed-precision clusters")
        mats, precs = self._keep
        sizes = [m.nrows for m in mats]
        solver = "cg" if self.solver == SOLVER_CG else "bicgstab"
        times, chosen, lo = {}, [], 0
        for hi in list(limits) + [None]:
            idx = [k for k, n in enumerate(sizes) if n > lo and (hi is None or n <= hi)]
            lo = hi if hi is not None else lo
            if not idx:
                chosen.append(chosen[-1] if chosen else self.rows_per_group)
                continue
            sub = Batch(self.ctx, [mats[k] for k in idx], [precs[k] for k in idx], solver, self.options)
            try:
                t = {}
                for rows in candidates:
                    sub.set_cluster(rows, self.max_workers)
                    t[rows] = sub._time_solve([rhs[k] for k in idx], [solutions[k] for k in idx], repeats)
            finally:
                sub.close()
            times[hi] = t
            chosen.append(min(t, key=t.get))
        self.set_cluster_buckets(list(limits), chosen[:-1])
        self.set_cluster(chosen[-1], self.max_workers)
        return times

    def set_phase_timing(self, enable=True):
        check(lib().fvg_batch_set_phase_timing(self.ctx.h, self.h, 1 if enable else 0))

    def phase_times(self):
        """Microseconds per system in SpMV, lower sweeps, upper sweeps and the whole solve (n x 4)."""
        out = np.zeros((self.n, 4))
        check(lib().fvg_batch_phase_times(self.ctx.h, self.h, _ptr(out)))
        return out

    def launch(self, rhs, solutions, ctx=None):
        """Enqueue the solves (solutions: initial guesses in, answers out); returns at once.
        The batch starts after the work already enqueued on ctx's stream (the batch's own
        context by default), where value refreshes are enqueued too."""
        if len(rhs) != self.n or len(solutions) != self.n:
            raise ValueError("one right-hand side and one solution per system")
        bv = (P * self.n)(*[v.h for v in rhs])
        xv = (P * self.n)(*[v.h for v in solutions])
        check(lib().fvg_batch_launch((ctx or self.ctx).h, self.h, bv, xv))
        self._inflight = (rhs, solutions)

    def poll(self):
        """Per-system completion of the launched batch: (done flags, infos of the done ones)."""
        flags = np.zeros(self.n, np.int32)
        infos = (SolverInfo * self.n)()
        check(lib().fvg_batch_poll(self.ctx.h, self.h, _ptr(flags), infos))
        name = "ConjugateGradient" if self.solver == SOLVER_CG else "BiConjugateGradientStab"
        out = {}
        for k in np.nonzero(flags)[0]:
            i = infos[int(k)]
            out[int(k)] = LinearSolverInfo(name, i.iterations, i.initial_residual, i.final_residual, np.zeros(0),
                                           i.restarts, i.outer_iterations, i.true_residual, i.status, i.flags)
        return flags.astype(bool), out

    def done(self):
        d = C.c_int32()
        check(lib().fvg_batch_query(self.ctx.h, self.h, C.byref(d)))
        return bool(d.value)

    def finish(self, history=True):
        """Wait for the launched solves; returns one LinearSolverInfo per system."""
        infos = (SolverInfo * self.n)()
        cap = max(2, self.options.max_iterations + 2) if self.options.precision == PRECISION_FP64 else 64
        hists = [np.zeros(cap) for _ in range(self.n)] if history else None
        hp = (P * self.n)(*[_ptr(hh) for hh in hists]) if history else None
        check(lib().fvg_batch_finish(self.ctx.h, self.h, infos, hp, cap if history else 0))
        self._inflight = None
        name = "ConjugateGradient" if self.solver == SOLVER_CG else "BiConjugateGradientStab"
        out = []
        for k in range(self.n):
            i = infos[k]
            hl = min(i.history_len, cap)
            out.append(LinearSolverInfo(name, i.iterations, i.initial_residual, i.final_residual,
                                        hists[k][:hl].copy() if history else np.zeros(0), i.restarts,
                                        i.outer_iterations, i.true_residual, i.status, i.flags))
        return out

    def solve(self, rhs, solutions, history=True):
        self.launch(rhs, solutions)
        return self.finish(history)

    def close(self):
        if getattr(self, "h", None):
            if self._inflight is not None:
                lib().fvg_batch_finish(self.ctx.h, self.h, None, None, 0)
            lib().fvg_batch_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
