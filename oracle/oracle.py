"""ctypes binding of the CPU FP64 oracle (oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by the
CPU-baseline leg of bench.py.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

OP = dict(LAPLACIAN_ORTHOGONAL=1, CONVECTION_UPWIND=2, SOURCE=3, LINEAR_SOURCE=4, TIME_EULER=5,
          DIVERGENCE=6, LAPLACIAN_CORRECTED=7, CONVECTION_LINEAR=8, CONVECTION_LIMITED_LINEAR=9,
          TIME_BACKWARD=10, USER_TERM=11, PARAM=15, ADD=16, SUB=17, NEG=18)
# limiter schemes (schemes/limiters/mod.rs) and gradient interpolation schemes
LIMITER = dict(limited_linear=0, minmod=1, vanleer=2, vanalbada=3, sweby=4, superbee=5)
INTERP = dict(linear=0, upwind=1)


class TermOp(C.Structure):
    _fields_ = [("op", C.c_int32), ("arg0", C.c_int32), ("arg1", C.c_int32), ("arg2", C.c_int32),
                ("scalar", C.c_double)]


def build():
    subprocess.check_call(["make", "-s", "-B", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        P = C.c_void_p
        L.orc_last_error.restype = C.c_char_p
        L.orc_mesh_new.restype = P
        L.orc_mesh_new.argtypes = [C.c_int, C.c_int64, P, C.c_int64, P, P, P, C.c_int64, P, P]
        L.orc_mesh_free.argtypes = [P]
        L.orc_mesh_nnz.restype = C.c_int64
        L.orc_mesh_nnz.argtypes = [P]
        L.orc_mesh_geometry.argtypes = [P] * 8
        L.orc_assemble.restype = C.c_int
        L.orc_assemble.argtypes = [P, P, C.c_int, P, P, P, P, P, P, P, P]
        L.orc_spmv.argtypes = [C.c_int64, P, P, P, P, P]
        L.orc_mesh_grad_coefficients.argtypes = [P, P, P]
        L.orc_gradients.restype = C.c_int
        L.orc_gradients.argtypes = [P, C.c_int, P, P, P, P]
        L.orc_limiters.restype = C.c_int
        L.orc_limiters.argtypes = [P, C.c_int, P, P, C.c_int, C.c_double, C.c_int, P, P, C.c_double, P, P, P]
        L.orc_matrix_inverse.restype = C.c_int
        L.orc_matrix_inverse.argtypes = [C.c_int, P, P]
        L.orc_prec_new.restype = P
        L.orc_prec_new.argtypes = [C.c_int, C.c_int64, P, P, P, C.c_int, C.POINTER(C.c_int)]
        L.orc_prec_free.argtypes = [P]
        L.orc_prec_factor_nnz.restype = C.c_int64
        L.orc_prec_factor_nnz.argtypes = [P]
        L.orc_prec_factor.restype = C.c_int
        L.orc_prec_factor.argtypes = [P, P, P, P, P]
        L.orc_precondition.argtypes = [P, C.c_int64, P, P]
        L.orc_solve.restype = C.c_int
        L.orc_solve.argtypes = [C.c_int, C.c_int64, P, P, P, P, P, P, C.c_double, C.c_double, C.c_int64, P, P,
                                C.c_int64]
        for name, args in (("orc_hbya_ainv", [P, C.c_int64, P, P, P, P, P, P, P]),
                           ("orc_interpolate_hbya_ainv_faces", [P, P, P, C.c_int, P, P, C.c_int, P, P, P, P, P, P,
                                                                P, P]),
                           ("orc_correct_phi", [P, P, P, P, P, C.c_double, P, P, P]),
                           ("orc_correct_velocity", [P, P, P, P, P]),
                           ("orc_wall_face_viscosity", [P, P, P]),
                           ("orc_wall_normalize", [P, P, P]),
                           ("orc_wall_distance", [P, P, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int64, P,
                                                  P])):
            getattr(L, name).restype = C.c_int
            getattr(L, name).argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _err():
    return lib().orc_last_error().decode()


class OracleMesh:
    """Mesh built with add_node/add_face/add_cell and Mesh::compute (compute.rs:12-261)."""

    def __init__(self, m):
        self.dim = int(m["dim"])
        self._keep = [np.ascontiguousarray(m["nodes"], dtype=np.float64),
                      np.ascontiguousarray(m["face_ptr"], dtype=np.int64),
                      np.ascontiguousarray(m["face_nodes"], dtype=np.int32),
                      np.ascontiguousarray(m["face_patch"], dtype=np.int32),
                      np.ascontiguousarray(m["cell_ptr"], dtype=np.int64),
                      np.ascontiguousarray(m["cell_faces"], dtype=np.int32)]
        nodes, fptr, fn, fp, cptr, cf = self._keep
        self.n_faces = fptr.shape[0] - 1
        self.n_cells = cptr.shape[0] - 1
        self.n_bnd = int((fp >= 0).sum())
        h = lib().orc_mesh_new(self.dim, nodes.shape[0], _p(nodes), self.n_faces, _p(fptr), _p(fn), _p(fp),
                               self.n_cells, _p(cptr), _p(cf))
        if not h:
            raise RuntimeError("oracle mesh: " + _err())
        self.h = h
        self.nnz = int(lib().orc_mesh_nnz(h))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_mesh_free(self.h)
            self.h = None

    def geometry(self):
        F, N, D = self.n_faces, self.n_cells, self.dim
        g = dict(face_area=np.empty(F), face_center=np.empty((F, D)), face_normal=np.empty((F, D)),
                 owner=np.empty(F, np.int32), neighbor=np.empty(F, np.int32), cell_volume=np.empty(N),
                 cell_center=np.empty((N, D)))
        lib().orc_mesh_geometry(self.h, *[_p(g[k]) for k in ("face_area", "face_center", "face_normal", "owner",
                                                               "neighbor", "cell_volume", "cell_center")])
        return g

    def grad_coefficients(self):
        """Mesh::compute's least-squares gradient coefficients (compute.rs:154-200): one
        DIM-vector per cell-face listing (cell_faces order) and one per cell."""
        D = self.dim
        face = np.empty((self._keep[5].shape[0], D))
        diag = np.empty((self.n_cells, D))
        lib().orc_mesh_grad_coefficients(self.h, _p(face), _p(diag))
        return face, diag

    def gradients(self, values, bc_i, bc_e):
        """compute_gradients (tools/gradients.rs:41-79).  values: (n,) or (n, nc);
        bc_e: (n_bnd,) or (n_bnd, nc).  Returns (n, DIM) or (n, nc, DIM)."""
        v = np.ascontiguousarray(values, dtype=np.float64)
        nc = 1 if v.ndim == 1 else v.shape[1]
        out = np.empty((self.n_cells, nc, self.dim))
        bi = np.ascontiguousarray(bc_i, dtype=np.float64)
        be = np.ascontiguousarray(bc_e, dtype=np.float64)
        lib().orc_gradients(self.h, nc, _p(v), _p(bi), _p(be), _p(out))
        return out[:, 0, :] if v.ndim == 1 else out

    def limiters(self, values, grads, limiter="limited_linear", k=1.0, interp="upwind", flux=None,
                 corrected_grads=None, reduction=1.0, bc_i=None, bc_e=None):
        """compute_limiters (tools/limiters.rs:47-126) with the given limiter scheme,
        gradient interpolation (linear / upwind on `flux`) and face-normal gradient
        (Orthogonal, or Corrected on `corrected_grads` with `reduction`)."""
        v = np.ascontiguousarray(values, dtype=np.float64)
        nc = 1 if v.ndim == 1 else v.shape[1]
        g = np.ascontiguousarray(grads, dtype=np.float64)
        fl = None if flux is None else np.ascontiguousarray(flux, dtype=np.float64)
        cg = None if corrected_grads is None else np.ascontiguousarray(corrected_grads, dtype=np.float64)
        bi = np.ascontiguousarray(bc_i, dtype=np.float64)
        be = np.ascontiguousarray(bc_e, dtype=np.float64)
        out = np.empty(self.n_faces)
        rc = lib().orc_limiters(self.h, nc, _p(v), _p(g), LIMITER[limiter], k, INTERP[interp], _p(fl), _p(cg),
                                reduction, _p(bi), _p(be), _p(out))
        if rc:
            raise RuntimeError("oracle limiters: " + _err())
        return out

    # ---- SIMPLE helpers (tools/hbya.rs) and wall distance (tools/wall_distance.rs) ----
    # An interpolation scheme is "linear", ("upwind", flux) or ("limited_linear", flux,
    # limiters); Vector fields are (n, dim).

    @staticmethod
    def _interp(spec):
        if spec == "linear" or spec == ("linear",):
            return 0, None, None
        if spec[0] == "upwind":
            return 1, np.ascontiguousarray(spec[1], dtype=np.float64), None
        if spec[0] == "limited_linear":
            return 2, np.ascontiguousarray(spec[1], dtype=np.float64), np.ascontiguousarray(spec[2],
                                                                                             dtype=np.float64)
        raise ValueError(spec)

    def hbya_ainv(self, row_ptr, col, val, mrhs, velocity):
        """compute_hbya_ainv (hbya.rs:15-50).  Returns (hbya (n, dim), ainv (n,))."""
        n = self.n_cells
        rp, ci = np.ascontiguousarray(row_ptr, np.int32), np.ascontiguousarray(col, np.int32)
        va = np.ascontiguousarray(val, np.float64)
        mr = np.ascontiguousarray(mrhs, np.float64)
        ve = np.ascontiguousarray(velocity, np.float64)
        hb, ai = np.empty((n, self.dim)), np.empty(n)
        if lib().orc_hbya_ainv(self.h, n, _p(rp), _p(ci), _p(va), _p(mr), _p(ve), _p(hb), _p(ai)):
            raise RuntimeError("oracle hbya_ainv: " + _err())
        return hb, ai

    def interpolate_hbya_ainv_faces(self, hbya, ainv, hbya_interp, ainv_interp, u_bc_i, u_bc_e, p_bc_i, p_bc_e):
        """intepolate_hbya_ainv_faces (hbya.rs:56-102).  Returns (hbyan_face, ainv_face)."""
        hk, hf, hl = self._interp(hbya_interp)
        ak, af, al = self._interp(ainv_interp)
        arr = lambda a: np.ascontiguousarray(a, np.float64)
        hb, ai, ubi, ube, pbi, pbe = (arr(a) for a in (hbya, ainv, u_bc_i, u_bc_e, p_bc_i, p_bc_e))
        hn, af_out = np.empty(self.n_faces), np.empty(self.n_faces)
        if lib().orc_interpolate_hbya_ainv_faces(self.h, _p(hb), _p(ai), hk, _p(hf), _p(hl), ak, _p(af), _p(al),
                                                 _p(ubi), _p(ube), _p(pbi), _p(pbe), _p(hn), _p(af_out)):
            raise RuntimeError("oracle interpolate_hbya_ainv_faces: " + _err())
        return hn, af_out

    def correct_phi(self, hbyan_face, ainv_face, pressure, p_bc_i, p_bc_e, corrected_grads=None, reduction=1.0):
        """correct_phi (hbya.rs:108-157) with Orthogonal, or Corrected on corrected_grads."""
        arr = lambda a: None if a is None else np.ascontiguousarray(a, np.float64)
        hn, af, pr, cg, pbi, pbe = (arr(a) for a in (hbyan_face, ainv_face, pressure, corrected_grads, p_bc_i,
                                                     p_bc_e))
        phi = np.empty(self.n_faces)
        lib().orc_correct_phi(self.h, _p(hn), _p(af), _p(pr), _p(cg), reduction, _p(pbi), _p(pbe), _p(phi))
        return phi

    def correct_velocity(self, hbya, ainv, pressure_gradient):
        """correct_velocity (hbya.rs:161-179).  Returns velocity (n, dim)."""
        arr = lambda a: np.ascontiguousarray(a, np.float64)
        hb, ai, pg = arr(hbya), arr(ainv), arr(pressure_gradient)
        v = np.empty((self.n_cells, self.dim))
        lib().orc_correct_velocity(self.h, _p(hb), _p(ai), _p(pg), _p(v))
        return v

    def wall_face_viscosity(self, grads):
        """wall_distance.rs:92-103 (p = 3): |Linear interpolation of grads| per face."""
        g = np.ascontiguousarray(grads, np.float64)
        out = np.empty(self.n_faces)
        lib().orc_wall_face_viscosity(self.h, _p(g), _p(out))
        return out

    def wall_normalize(self, grads, dist):
        """wall_distance.rs:107-116 (p = 3)."""
        g = np.ascontiguousarray(grads, np.float64)
        d = np.array(dist, dtype=np.float64)
        lib().orc_wall_normalize(self.h, _p(g), _p(d))
        return d

    def wall_distance(self, wall, max_iterations, tolerance, rtol, atol, max_iter, initial=None):
        """compute_wall_distance (wall_distance.rs:8-119); wall: bool per boundary face.
        Returns (distance, {solves, iterations, initial_residual, final_residual})."""
        w = np.ascontiguousarray(wall, dtype=np.uint8)
        d = np.zeros(self.n_cells) if initial is None else np.array(initial, dtype=np.float64)
        info = np.zeros(4)
        if lib().orc_wall_distance(self.h, _p(w), int(max_iterations), tolerance, rtol, atol, int(max_iter), _p(d),
                                   _p(info)):
            raise RuntimeError("oracle wall_distance: " + _err())
        return d, dict(solves=int(info[0]), iterations=int(info[1]), initial_residual=info[2],
                       final_residual=info[3])

    def assemble(self, program, face_fields, cell_fields, bc_i, bc_e):
        """program: list of (op, arg0, arg1, arg2, scalar). Returns (row_ptr, col, values, rhs)."""
        ops = (TermOp * max(1, len(program)))(*[TermOp(*t) for t in program])
        ff = [np.ascontiguousarray(a, dtype=np.float64) for a in face_fields]
        cf = [np.ascontiguousarray(a, dtype=np.float64) for a in cell_fields]
        ffp = (C.c_void_p * max(1, len(ff)))(*[a.ctypes.data for a in ff])
        cfp = (C.c_void_p * max(1, len(cf)))(*[a.ctypes.data for a in cf])
        bi = np.ascontiguousarray(bc_i, dtype=np.float64)
        be = np.ascontiguousarray(bc_e, dtype=np.float64)
        rp = np.empty(self.n_cells + 1, np.int32)
        ci = np.empty(self.nnz, np.int32)
        va = np.empty(self.nnz)
        rhs = np.empty(self.n_cells)
        rc = lib().orc_assemble(self.h, C.cast(ops, C.c_void_p), len(program), C.cast(ffp, C.c_void_p),
                                C.cast(cfp, C.c_void_p), _p(bi), _p(be), _p(rp), _p(ci), _p(va), _p(rhs))
        if rc:
            raise RuntimeError("oracle assemble: " + _err())
        return rp, ci, va, rhs


def matrix_inverse(a):
    """Matrix::inv (matrix.rs:320-383) of a 2x2 or 3x3 matrix; None where it returns Err."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    out = np.empty_like(a)
    return None if lib().orc_matrix_inverse(a.shape[0], _p(a), _p(out)) else out


def spmv(row_ptr, col, val, x):
    n = row_ptr.shape[0] - 1
    y = np.empty(n)
    x = np.ascontiguousarray(x, dtype=np.float64)
    lib().orc_spmv(n, _p(row_ptr), _p(col), _p(val), _p(x), _p(y))
    return y


class ZeroPivot(RuntimeError):
    pass


class NotSymmetric(RuntimeError):
    pass


class OraclePrec:
    """kind: 'identity', 'diagonal', 'ic', 'ilu' (preconditioners/*.rs)."""

    KINDS = dict(identity=0, diagonal=1, ic=2, ilu=3)

    def __init__(self, kind, row_ptr, col, val, level=0):
        self.kind = kind
        self.n = row_ptr.shape[0] - 1
        self._keep = (np.ascontiguousarray(row_ptr, np.int32), np.ascontiguousarray(col, np.int32),
                      np.ascontiguousarray(val, np.float64))
        st = C.c_int(0)
        h = lib().orc_prec_new(self.KINDS[kind], self.n, *[_p(a) for a in self._keep], level, C.byref(st))
        if not h:
            if st.value == 4:
                raise ZeroPivot(_err())
            if st.value == 5:
                raise NotSymmetric(_err())
            raise RuntimeError("oracle preconditioner: " + _err())
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_prec_free(self.h)
            self.h = None

    def factor(self):
        nnz = int(lib().orc_prec_factor_nnz(self.h))
        rp = np.empty(self.n + 1, np.int32)
        ci = np.empty(nnz, np.int32)
        va = np.empty(nnz)
        inv = np.empty(self.n)
        lib().orc_prec_factor(self.h, _p(rp), _p(ci), _p(va), _p(inv))
        return rp, ci, va, inv

    def apply(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.zeros(self.n)
        lib().orc_precondition(self.h, self.n, _p(r), _p(z))
        return z


def solve(solver, row_ptr, col, val, prec, b, x0=None, rtol=1e-8, atol=0.0, max_iter=1000):
    """solver: 'cg' or 'bicgstab'. Returns (x, info dict)."""
    n = row_ptr.shape[0] - 1
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    out = np.zeros(5)
    hist = np.zeros(max_iter + 2)
    rc = lib().orc_solve(0 if solver == "cg" else 1, n, _p(row_ptr), _p(col), _p(val), prec.h, _p(b), _p(x), rtol,
                         atol, max_iter, _p(out), _p(hist), hist.shape[0])
    if rc:
        raise RuntimeError("oracle solve: " + _err())
    hl = int(out[4])
    return x, dict(iterations=int(out[0]), initial_residual=float(out[1]), final_residual=float(out[2]),
                   restarts=int(out[3]), history=hist[:hl].copy())
