"""CPU FP64 oracle for the H1-HDiv hot path -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the CPU-baseline leg of bench.py
may import this module, and only as the checker.  The product path never
imports, links or calls anything under oracle/.

Layers restated here (reference files read as text):
  * element integrands: oracle/mhd_oracle.c (src/weakforms.jl:211-312, 670-733)
  * Gridap ``SparseMatrixAssembler`` semantics (src/main.jl:218-224): the CSC
    pattern is the union over cells of every (row, col) of every touched field
    block (uu, up, uj, pu, jj, jphi, ju, phij), Dirichlet rows/cols dropped,
    duplicates summed, rows sorted inside columns, explicit zeros kept.
  * Newton with a direct solve per step: ``_solver(::Val{:julia})``
    (src/main.jl:181-186) = ``NewtonSolver(LUSolver(); maxiter, rtol=1e-6)``.
    The iterative ``Badia2024Solver`` (src/Solvers/badia2024.jl:2-48) is not
    restated: it converges to the solution of the same discrete system, which
    the direct solve fixes.
  * Post-processing norms of ``_hunt`` (src/Applications/hunt.jl:247-260).
  * Skeleton stabilisation a_Lambda (src/weakforms.jl:685-688) in numpy
    (``skeleton_system``): interior faces found from the cells' shared
    face-centre velocity nodes, face points matched by physical position (up
    to the periodic translation), surface measure |dx/ds x dx/dt|.

oracle/cpu_baseline.py + cpu_step.c price one Newton step on the host CPU
(timing only, for the benchmark's CPU baseline).

The direct solve uses scipy's SuperLU.  A singular Jacobian (constant pressure
mode of a fully Dirichlet velocity; constant potential when j.n is prescribed
on the whole boundary) is made regular by pinning one dof of each null field
(``null_fields`` tells which); the pinned equation is implied by the others,
so every other unknown is unchanged and the pinned field is compared modulo
its constant.  A pin on a field without a null mode drops an equation that is
not implied, and Newton stalls at its residual.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Dict, Optional

import numpy as np
import scipy.sparse as sps
import scipy.sparse.linalg as spla

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "_build", "liborc.so")
NDOF = 129
OFF_P, OFF_J, OFF_PHI = 81, 85, 121
TYPES = np.array([0] * 81 + [1] * 4 + [2] * 36 + [3] * 8)
TOUCHED = np.array([[1, 1, 1, 0], [1, 0, 0, 0], [1, 0, 1, 1], [0, 0, 1, 0]], dtype=bool)


def build() -> str:
    subprocess.check_call(["make", "-s", "-C", _HERE], stdout=subprocess.DEVNULL)
    return _LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = ctypes.CDLL(_LIB)
    return _lib


class OrcProblem(ctypes.Structure):
    _fields_ = [
        ("ncells", ctypes.c_int32), ("nfree", ctypes.c_int32), ("ndir", ctypes.c_int32),
        ("convection", ctypes.c_int32),
        ("cell_xyz", ctypes.c_void_p), ("cell_dofs", ctypes.c_void_p), ("rt_sign", ctypes.c_void_p),
        ("dir_vals", ctypes.c_void_p),
        ("alpha", ctypes.c_double), ("beta", ctypes.c_double), ("gamma", ctypes.c_double),
        ("sigma", ctypes.c_double), ("B", ctypes.c_double * 3),
        ("f0", ctypes.c_double * 3), ("fx", ctypes.c_double * 9),
        ("g0", ctypes.c_double * 3), ("gx", ctypes.c_double * 9),
        ("zeta_u", ctypes.c_double), ("zeta_j", ctypes.c_double),
        ("nfaces", ctypes.c_int32), ("pad_", ctypes.c_int32),
        ("face_cell", ctypes.c_void_p), ("face_lface", ctypes.c_void_p),
        ("face_tau", ctypes.c_void_p), ("face_cw", ctypes.c_void_p), ("face_jw", ctypes.c_void_p),
    ]


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p).value


CONVECTION = {"none": 0, "picard": 1, "newton": 2}


class Oracle:
    """The oracle bound to one problem: geometry, dof tables and physics.

    ``problem`` is a dict with keys cell_xyz (C,8,3), cell_dofs (C,129),
    rt_sign (C,6), dir_vals, nfree, field sizes ``nfields`` (u,p,j,phi),
    physics (alpha, beta, gamma, sigma, B, f0, fx, g0, gx, zeta_u, zeta_j,
    convection) and thin-wall faces (face_cell, face_lface, face_tau,
    face_cw, face_jw)."""

    def __init__(self, problem: Dict):
        self.pb = problem
        self.keep = []
        c = lambda a, dt: self.keep.append(np.ascontiguousarray(a, dtype=dt)) or self.keep[-1]
        self.cell_xyz = c(problem["cell_xyz"], np.float64)
        self.cell_dofs = c(problem["cell_dofs"], np.int32)
        self.rt_sign = c(problem["rt_sign"], np.int8)
        self.dir_vals = c(problem.get("dir_vals", np.zeros(1)), np.float64)
        if self.dir_vals.size == 0:
            self.dir_vals = c(np.zeros(1), np.float64)
        self.ncells = self.cell_dofs.shape[0]
        self.nfree = int(problem["nfree"])
        fc = c(problem.get("face_cell", np.zeros(0)), np.int32)
        fl = c(problem.get("face_lface", np.zeros(0)), np.int8)
        ft = c(problem.get("face_tau", np.zeros(0)), np.float64)
        fcw = c(problem.get("face_cw", np.zeros(0)), np.float64)
        fjw = c(problem.get("face_jw", np.zeros(0)), np.float64)
        self.nfaces = fc.size
        P = OrcProblem()
        P.ncells = self.ncells
        P.nfree = self.nfree
        P.ndir = self.dir_vals.size
        P.convection = CONVECTION[problem.get("convection", "newton")]
        P.cell_xyz, P.cell_dofs, P.rt_sign, P.dir_vals = map(_ptr, (self.cell_xyz, self.cell_dofs, self.rt_sign, self.dir_vals))
        for k in ("alpha", "beta", "gamma", "sigma", "zeta_u", "zeta_j"):
            setattr(P, k, float(problem.get(k, {"sigma": 1.0}.get(k, 0.0))))
        for k, n in (("B", 3), ("f0", 3), ("fx", 9), ("g0", 3), ("gx", 9)):
            v = np.asarray(problem.get(k, np.zeros(n)), dtype=np.float64).ravel()
            getattr(P, k)[:] = list(v)
        P.nfaces = self.nfaces
        P.face_cell, P.face_lface, P.face_tau, P.face_cw, P.face_jw = map(_ptr, (fc, fl, ft, fcw, fjw))
        self.P = P
        L = lib()
        for fn in ("orc_cell_residuals", "orc_cell_jacobians"):
            getattr(L, fn).argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
        for fn in ("orc_face_residuals", "orc_face_jacobians"):
            getattr(L, fn).argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_cell_norms.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
        L.orc_point_values.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        assert L.orc_struct_size() == ctypes.sizeof(OrcProblem)
        self._pattern = None
        self.stab_mu = float(problem.get("stab_mu", 0.0))
        self._skel = None

    def skeleton(self):
        """(K (nf, 54, 54), dofs (nf, 3, 54)) of the skeleton stabilisation, or None."""
        if self.stab_mu == 0.0:
            return None
        if self._skel is None:
            self._skel = skeleton_system(self.cell_xyz, self.cell_dofs, self.stab_mu)
        return self._skel

    # ---- element level ----------------------------------------------------
    def cell_residuals(self, x, c0=0, c1=None):
        c1 = self.ncells if c1 is None else c1
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros((c1 - c0, NDOF))
        lib().orc_cell_residuals(ctypes.byref(self.P), _ptr(x), c0, c1, _ptr(out))
        return out

    def cell_jacobians(self, x, c0=0, c1=None):
        c1 = self.ncells if c1 is None else c1
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros((c1 - c0, NDOF, NDOF))
        lib().orc_cell_jacobians(ctypes.byref(self.P), _ptr(x), c0, c1, _ptr(out))
        return out

    def face_residuals(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros((self.nfaces, 36))
        if self.nfaces:
            lib().orc_face_residuals(ctypes.byref(self.P), _ptr(x), _ptr(out))
        return out

    def face_jacobians(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros((self.nfaces, 36, 36))
        if self.nfaces:
            lib().orc_face_jacobians(ctypes.byref(self.P), _ptr(x), _ptr(out))
        return out

    def set_dirichlet(self, dir_vals):
        self.dir_vals[:] = dir_vals

    # ---- global assembly (Gridap SparseMatrixAssembler semantics) ---------
    def residual(self, x):
        r = np.zeros(self.nfree)
        re = self.cell_residuals(x)
        cd = self.cell_dofs
        m = cd >= 0
        np.add.at(r, cd[m], re[m])
        if self.nfaces:
            rf = self.face_residuals(x)
            fd = cd[self.keep_face_cell(), OFF_J:OFF_J + 36]
            m = fd >= 0
            np.add.at(r, fd[m], rf[m])
        sk = self.skeleton()
        if sk is not None:
            K, dofs = sk
            xv = np.where(dofs >= 0, np.asarray(x)[np.maximum(dofs, 0)], self.dir_vals[np.maximum(-dofs - 1, 0)])
            rs = np.einsum("fab,fcb->fca", K, xv)
            m = dofs >= 0
            np.add.at(r, dofs[m], rs[m])
        return r

    def keep_face_cell(self):
        return np.asarray(self.pb.get("face_cell", np.zeros(0)), dtype=np.int64)

    def _coo(self, x=None, values=True, chunk=256):
        cd = self.cell_dofs.astype(np.int64)
        mask = TOUCHED[TYPES[:, None], TYPES[None, :]]
        rows, cols, vals = [], [], []
        for c0 in range(0, self.ncells, chunk):
            c1 = min(self.ncells, c0 + chunk)
            Ke = self.cell_jacobians(x, c0, c1) if values else None
            R = np.broadcast_to(cd[c0:c1, :, None], (c1 - c0, NDOF, NDOF))
            C = np.broadcast_to(cd[c0:c1, None, :], (c1 - c0, NDOF, NDOF))
            m = (R >= 0) & (C >= 0) & mask[None]
            rows.append(R[m]); cols.append(C[m])
            vals.append(Ke[m] if values else np.zeros(int(m.sum())))
        if self.nfaces:
            fc = self.keep_face_cell()
            fd = cd[fc, OFF_J:OFF_J + 36]
            Kf = self.face_jacobians(x) if values else None
            R = np.broadcast_to(fd[:, :, None], (fc.size, 36, 36))
            C = np.broadcast_to(fd[:, None, :], (fc.size, 36, 36))
            m = (R >= 0) & (C >= 0)
            rows.append(R[m]); cols.append(C[m])
            vals.append(Kf[m] if values else np.zeros(int(m.sum())))
        sk = self.skeleton()
        if sk is not None:
            K, dofs = sk
            nf = K.shape[0]
            R = np.broadcast_to(dofs[:, :, :, None], (nf, 3, 54, 54))
            C = np.broadcast_to(dofs[:, :, None, :], (nf, 3, 54, 54))
            m = (R >= 0) & (C >= 0)
            rows.append(R[m]); cols.append(C[m])
            vals.append(np.broadcast_to(K[:, None], (nf, 3, 54, 54))[m] if values else np.zeros(int(m.sum())))
        return np.concatenate(rows), np.concatenate(cols), np.concatenate(vals)

    def jacobian(self, x) -> sps.csc_matrix:
        r, c, v = self._coo(x)
        A = sps.csc_matrix((v, (r, c)), shape=(self.nfree, self.nfree))
        A.sum_duplicates()
        A.has_sorted_indices = False
        A.sort_indices()
        return A

    def pattern(self):
        """(colptr, rowval) of the CSC, 0-based, int64."""
        if self._pattern is None:
            r, c, _ = self._coo(values=False)
            key = np.unique(c * np.int64(self.nfree) + r)
            cols = key // self.nfree
            rows = key % self.nfree
            colptr = np.zeros(self.nfree + 1, dtype=np.int64)
            np.add.at(colptr, cols + 1, 1)
            self._pattern = (np.cumsum(colptr), rows.astype(np.int64))
        return self._pattern

    def jacobian_on_pattern(self, x) -> np.ndarray:
        """nzval in the CSC order of ``pattern()`` (explicit zeros kept)."""
        colptr, rowval = self.pattern()
        A = self.jacobian(x)
        # A may drop nothing (scipy keeps explicit zeros after sum_duplicates)
        assert A.nnz == rowval.size, (A.nnz, rowval.size)
        assert np.array_equal(A.indptr, colptr) and np.array_equal(A.indices, rowval)
        return A.data.copy()

    # ---- norms ------------------------------------------------------------
    def cell_norms(self, x, nq=4):
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros((self.ncells, 6))
        lib().orc_cell_norms(ctypes.byref(self.P), _ptr(x), nq, _ptr(out))
        return out

    def point_values(self, x, cells, xi):
        x = np.ascontiguousarray(x, dtype=np.float64)
        cells = np.ascontiguousarray(cells, dtype=np.int32)
        xi = np.ascontiguousarray(xi, dtype=np.float64)
        out = np.zeros((cells.size, 17))
        lib().orc_point_values(ctypes.byref(self.P), _ptr(x), cells.size, _ptr(cells), _ptr(xi), _ptr(out))
        return out


# ---------------------------------------------------------------------------
# Skeleton stabilisation  a_Lambda(u, v) = int_Lambda (1/2) mu h^2 [grad u] : [grad v]
# (src/weakforms.jl:685-688), Lambda the interior faces, h = |face|^(1/2)
# (get_cell_size of the skeleton triangulation: measure^(1/dim) with dim 2,
# src/geometry.jl:150-158).

def _gauss3():
    x, w = np.polynomial.legendre.leggauss(3)
    return (x + 1.0) / 2.0, w / 2.0


def _lagrange2(t):
    """Quadratic Lagrange basis on the nodes 0, 1/2, 1 and its derivative."""
    L = np.stack([2.0 * (t - 0.5) * (t - 1.0), -4.0 * t * (t - 1.0), 2.0 * t * (t - 0.5)], -1)
    dL = np.stack([4.0 * t - 3.0, 4.0 - 8.0 * t, 4.0 * t - 1.0], -1)
    return L, dL


def q2_ref_gradients(xi):
    """Reference gradients (n, 27, 3) of the Q2 node functions a = i + 3j + 9k."""
    L, dL = zip(*[_lagrange2(xi[:, d]) for d in range(3)])
    out = np.zeros((xi.shape[0], 27, 3))
    for a in range(27):
        i, j, k = a % 3, (a // 3) % 3, a // 9
        out[:, a, 0] = dL[0][:, i] * L[1][:, j] * L[2][:, k]
        out[:, a, 1] = L[0][:, i] * dL[1][:, j] * L[2][:, k]
        out[:, a, 2] = L[0][:, i] * L[1][:, j] * dL[2][:, k]
    return out


def trilinear_map(X, xi):
    """Points (n, 3) and Jacobians J[i][l] = dx_i / dxi_l (n, 3, 3) of cells X (n, 8, 3) at xi (n, 3)."""
    n = xi.shape[0]
    N = np.ones((n, 8))
    dN = np.ones((n, 8, 3))
    for v in range(8):
        b = [(v >> d) & 1 for d in range(3)]
        for d in range(3):
            f = xi[:, d] if b[d] else 1.0 - xi[:, d]
            df = 1.0 if b[d] else -1.0
            N[:, v] *= f
            for l in range(3):
                dN[:, v, l] *= df if l == d else f
    return np.einsum("nv,nvi->ni", N, X), np.einsum("nvi,nvl->nil", X, dN)


def _face_ref_points(lf):
    """3x3 Gauss points and weights on local face lf (z-, z+, y-, y+, x-, x+)."""
    g, w = _gauss3()
    d = 2 - lf // 2
    t = [a for a in range(3) if a != d]
    P = np.zeros((9, 3))
    P[:, d] = float(lf & 1)
    P[:, t[0]] = np.tile(g, 3)
    P[:, t[1]] = np.repeat(g, 3)
    return P, np.outer(w, w).ravel(), t


def interior_faces(cell_dofs):
    """(cell, local face) pairs (nf, 2, 2) of the faces two cells share: the
    faces whose centre velocity node (local node with coordinate 0 or 2 along
    the face axis, 1 along the others) carries the same dof in both cells."""
    centre = []
    for lf in range(6):
        d = 2 - lf // 2
        c = [1, 1, 1]
        c[d] = 2 if lf & 1 else 0
        centre.append(c[0] + 3 * c[1] + 9 * c[2])
    key = cell_dofs[:, centre].astype(np.int64).ravel()
    order = np.argsort(key, kind="stable")
    ks = key[order]
    same = np.nonzero(ks[1:] == ks[:-1])[0]
    a, b = order[same], order[same + 1]
    return np.stack([np.stack([a // 6, a % 6], 1), np.stack([b // 6, b % 6], 1)], 1)


def skeleton_system(cell_xyz, cell_dofs, mu):
    """Face matrices K (nf, 54, 54) of one velocity component over the 27
    velocity nodes of each side (first side, then second), and their dofs
    (nf, 3, 54) per component; a node shared by both sides appears twice
    (the assembly sums them)."""
    pairs = interior_faces(cell_dofs)
    nf = pairs.shape[0]
    X = cell_xyz.reshape(-1, 8, 3)
    K = np.zeros((nf, 54, 54))
    dofs = np.zeros((nf, 3, 54), dtype=np.int64)
    for f in range(nf):
        (c0, l0), (c1, l1) = pairs[f]
        P0, w, t = _face_ref_points(l0)
        P1, _, _ = _face_ref_points(l1)
        x0, J0 = trilinear_map(np.repeat(X[c0][None], 9, 0), P0)
        x1, J1 = trilinear_map(np.repeat(X[c1][None], 9, 0), P1)
        # the second side's points in the first side's order (translation: periodic images)
        x1s = x1 + (x0.mean(0) - x1.mean(0))
        dist = np.linalg.norm(x0[:, None] - x1s[None], axis=2)
        m = dist.argmin(1)
        scale = np.linalg.norm(x0[0] - x0[8])
        assert (dist[np.arange(9), m] < 1e-9 * scale).all() and np.unique(m).size == 9, "face points do not match"
        dA = w * np.linalg.norm(np.cross(J0[:, :, t[0]], J0[:, :, t[1]]), axis=1)
        g0 = np.einsum("qli,qal->qai", np.linalg.inv(J0), q2_ref_gradients(P0))
        g1 = np.einsum("qli,qal->qai", np.linalg.inv(J1[m]), q2_ref_gradients(P1[m]))
        G = np.concatenate([g0, -g1], axis=1)                      # (9, 54, 3) jumps
        K[f] = 0.5 * mu * dA.sum() * np.einsum("q,qai,qbi->ab", dA, G, G)
        for b in range(3):
            dofs[f, b, :27] = cell_dofs[c0, b * 27:(b + 1) * 27]
            dofs[f, b, 27:] = cell_dofs[c1, b * 27:(b + 1) * 27]
    return K, dofs


# ---------------------------------------------------------------------------
# Solvers

def null_pins(spaces, pin_p=True, pin_phi=False):
    """Indices of the dofs pinned to remove constant-pressure/potential modes:
    the first free dof of p (and of phi).  ``spaces`` is a Spaces object (any
    multifield numbering, read from its field slices) or a dict of field sizes
    in the consecutive (u, p, j, phi) numbering."""
    if hasattr(spaces, "field_slices"):
        first = {f: s.start for f, s in spaces.field_slices().items()}
    else:
        first, off = {}, 0
        for f in ("u", "p", "j", "phi"):
            first[f] = off
            off += spaces[f]
    out = []
    if pin_p:
        out.append(first["p"])
    if pin_phi:
        out.append(first["phi"])
    return out


def null_fields(orc: Oracle, spaces, x=None, tol=1e-12):
    """The fields among p and phi whose constant is a null vector of the
    Jacobian at x (A c = 0 for c = 1 on that field's dofs: the dP1 and dQ1
    bases both sum to one), i.e. the fields ``null_pins`` should pin.  The
    constant pressure is null with u Dirichlet on every wall; the constant
    potential only where j.n is prescribed on the whole boundary."""
    x = np.zeros(orc.nfree) if x is None else x
    A = orc.jacobian(x)
    scale = abs(A).max()
    out = []
    for f in ("p", "phi"):
        c = np.zeros(orc.nfree)
        c[spaces.field_slices()[f]] = 1.0
        if np.abs(A @ c).max() <= tol * scale * max(1, np.sqrt(c.sum())):
            out.append(f)
    return out


def direct_solve(A: sps.csc_matrix, b: np.ndarray, pins=()) -> np.ndarray:
    """Solve A x = b with SuperLU; rows/cols of ``pins`` replaced by identity."""
    if pins:
        A = A.tolil(copy=True)
        b = b.copy()
        for k in pins:
            A[k, :] = 0.0
            A[:, k] = 0.0
            A[k, k] = 1.0
            b[k] = 0.0
        A = A.tocsc()
    lu = spla.splu(A, permc_spec="COLAMD")
    return lu.solve(b)


def newton_direct(orc: Oracle, x0: np.ndarray, maxiter=10, rtol=1e-6, atol=1e-12, pins=(), log=None):
    """Newton with a direct solve per step (``_solver(::Val{:julia})``).

    Stops when ||R|| <= max(atol, rtol ||R0||) like GridapSolvers' NewtonSolver."""
    x = x0.copy()
    r = orc.residual(x)
    r0 = np.linalg.norm(r)
    hist = [r0]
    for it in range(maxiter):
        if hist[-1] <= max(atol, rtol * r0):
            break
        A = orc.jacobian(x)
        dx = direct_solve(A, -r, pins)
        x += dx
        r = orc.residual(x)
        hist.append(np.linalg.norm(r))
    if log is not None:
        log["residuals"] = hist
    return x
