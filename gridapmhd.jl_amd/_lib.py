"""ctypes binding of libmhd's C ABI (include/mhd.h).

This is the Python-side counterpart of the ``ccall`` glue a Julia caller
uses; it passes plain pointers and sizes and owns nothing on the device.
The library is required: every entry point raises if it cannot be loaded.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmhd.so")

c_int32, c_int64, c_double, c_void_p = ctypes.c_int32, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p


class MhdMesh(ctypes.Structure):
    _fields_ = [("ncells", c_int32), ("nfaces_tw", c_int32), ("cell_xyz", c_void_p),
                ("cell_verts", c_void_p), ("nverts", c_int32), ("pad_", c_int32),
                ("tw_face_cell", c_void_p), ("tw_face_lface", c_void_p), ("tw_tau", c_void_p),
                ("tw_cw", c_void_p), ("tw_jw", c_void_p)]


class MhdSkeleton(ctypes.Structure):
    _fields_ = [("nfaces", c_int32), ("pad_", c_int32), ("face_cell", c_void_p), ("face_lface", c_void_p),
                ("face_perm", c_void_p), ("mu", c_void_p)]


class MhdSpaces(ctypes.Structure):
    _fields_ = [("nfree", c_int32 * 4), ("ndir", c_int32), ("block_style", c_int32), ("cell_dofs", c_void_p),
                ("rt_sign", c_void_p), ("dir_vals", c_void_p)]


class MhdParams(ctypes.Structure):
    _fields_ = [("alpha", c_double), ("beta", c_double), ("gamma", c_double), ("sigma", c_double),
                ("B", c_double * 3), ("f0", c_double * 3), ("fx", c_double * 9), ("g0", c_double * 3),
                ("gx", c_double * 9), ("zeta_u", c_double), ("zeta_j", c_double), ("convection", c_int32),
                ("pad_", c_int32)]


class MhdSolverOpts(ctypes.Structure):
    _fields_ = [("max_newton", c_int32), ("restart", c_int32), ("max_krylov", c_int32), ("precision", c_int32),
                ("rtol", c_double), ("atol", c_double), ("ls_rtol", c_double), ("inner_iters", c_int32),
                ("verbose", c_int32)]


class MhdSolveInfo(ctypes.Structure):
    _fields_ = [("newton_iters", c_int32), ("krylov_iters", c_int32), ("converged", c_int32), ("flags", c_int32),
                ("residual0", c_double), ("residual", c_double), ("t_residual", c_double),
                ("t_jacobian", c_double), ("t_setup", c_double), ("t_solve", c_double),
                ("newton_residuals", c_double * 64), ("krylov_per_step", c_int32 * 64)]


EXPORTS = {
    "mhd_create": (ctypes.c_int, [ctypes.POINTER(c_void_p), ctypes.POINTER(MhdMesh), ctypes.POINTER(MhdSpaces),
                                  ctypes.POINTER(MhdParams), c_int32]),
    "mhd_create_stabilized": (ctypes.c_int, [ctypes.POINTER(c_void_p), ctypes.POINTER(MhdMesh),
                                             ctypes.POINTER(MhdSpaces), ctypes.POINTER(MhdParams),
                                             ctypes.POINTER(MhdSkeleton), c_int32]),
    "mhd_destroy": (None, [c_void_p]),
    "mhd_last_error": (ctypes.c_char_p, [c_void_p]),
    "mhd_set_params": (ctypes.c_int, [c_void_p, ctypes.POINTER(MhdParams)]),
    "mhd_set_dirichlet": (ctypes.c_int, [c_void_p, c_void_p]),
    "mhd_sizes": (ctypes.c_int, [c_void_p, ctypes.POINTER(c_int64), ctypes.POINTER(c_int64)]),
    "mhd_set_assembly_mode": (ctypes.c_int, [c_void_p, c_int32]),
    "mhd_pattern": (ctypes.c_int, [c_void_p, c_void_p, c_void_p]),
    "mhd_residual": (ctypes.c_int, [c_void_p, c_void_p, c_void_p]),
    "mhd_jacobian": (ctypes.c_int, [c_void_p, c_void_p, c_void_p]),
    "mhd_cell_integrals": (ctypes.c_int, [c_void_p, c_void_p, c_void_p]),
    "mhd_residual_d": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "mhd_residual_d_noalloc": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "mhd_jacobian_d": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "mhd_csr_arrays": (ctypes.c_int, [c_void_p, ctypes.POINTER(c_void_p), ctypes.POINTER(c_void_p)]),
    "mhd_spmv_d": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "mhd_newton_solve": (ctypes.c_int, [c_void_p, c_void_p, ctypes.POINTER(MhdSolverOpts),
                                        ctypes.POINTER(MhdSolveInfo)]),
    "mhd_default_solver_opts": (None, [ctypes.POINTER(MhdSolverOpts)]),
    "mhd_precond_setup": (ctypes.c_int, [c_void_p, ctypes.POINTER(MhdSolverOpts)]),
    "mhd_dresidual_dparam": (ctypes.c_int, [c_void_p, c_void_p, c_int32, c_void_p]),
    "mhd_adjoint_solve": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, c_void_p, ctypes.POINTER(MhdSolverOpts),
                                         ctypes.POINTER(MhdSolveInfo)]),
    "mhd_linear_solve": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, c_void_p, ctypes.POINTER(MhdSolverOpts),
                                        ctypes.POINTER(MhdSolveInfo)]),
    "mhd_attach_coarse": (ctypes.c_int, [c_void_p, c_void_p, c_int64, c_int64, c_void_p, c_void_p, c_void_p, c_void_p,
                                         c_void_p, c_void_p, c_int32, c_double, c_int32]),
    "mhd_set_pivot_threshold": (ctypes.c_int, [c_void_p, c_double]),
    "mhd_smoother_apply_d": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "mhd_smoother_info": (ctypes.c_int, [c_void_p, ctypes.POINTER(c_int64), ctypes.POINTER(c_int64),
                                         ctypes.POINTER(c_int64), ctypes.POINTER(c_int64)]),
    "mhd_op_counts": (ctypes.c_int, [c_void_p, c_void_p, ctypes.c_int]),
}

# slots of mhd_op_counts (include/mhd.h)
OP_COUNT_NAMES = ("spmv_nnz_fp64", "spmv_nnz_fp32", "patch_entries_fp64", "patch_entries_fp32",
                  "patch_entries_bf16", "dense_entries", "vector_bytes", "setup_cells", "setup_patch_dim3",
                  "setup_dense_dim3", "outer_iters")

_lib: Optional[ctypes.CDLL] = None


class MhdError(RuntimeError):
    pass


def load(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load libmhd and bind every exported symbol.  Raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise MhdError(f"libmhd not built: {path} is missing (run __graft_entry__.build())")
    lib = ctypes.CDLL(path)
    for name, (res, args) in EXPORTS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(c_void_p).value


CONVECTION = {"none": 0, "picard": 1, "newton": 2}
COARSE_STEP = {"full": 0, "minres": 1}


def make_params(phys: dict) -> MhdParams:
    p = MhdParams()
    p.alpha = float(phys.get("alpha", 1.0))
    p.beta = float(phys.get("beta", 1.0))
    p.gamma = float(phys.get("gamma", 1.0))
    p.sigma = float(phys.get("sigma", 1.0))
    for k, n in (("B", 3), ("f0", 3), ("fx", 9), ("g0", 3), ("gx", 9)):
        v = np.asarray(phys.get(k, np.zeros(n)), dtype=np.float64).ravel()
        getattr(p, k)[:] = list(v)
    p.zeta_u = float(phys.get("zeta_u", 0.0))
    p.zeta_j = float(phys.get("zeta_j", 0.0))
    conv = phys.get("convection", "newton")
    p.convection = CONVECTION[conv] if isinstance(conv, str) else int(conv)
    return p


class Context:
    """One libmhd context: the operator of one problem on one device."""

    def __init__(self, cell_xyz, cell_dofs, rt_sign, nfree4, dir_vals, phys, faces=None, device=0,
                 cell_verts=None, block_style=0, skeleton=None):
        self.lib = load()
        self._keep = []
        k = lambda a, dt: self._keep.append(np.ascontiguousarray(a, dtype=dt)) or self._keep[-1]
        xyz = k(cell_xyz, np.float64).reshape(-1)
        dofs = k(cell_dofs, np.int32)
        sign = k(rt_sign, np.int8)
        dv = k(dir_vals if len(dir_vals) else np.zeros(1), np.float64)
        faces = faces or {}
        fc = k(faces.get("cell", np.zeros(0)), np.int32)
        fl = k(faces.get("lface", np.zeros(0)), np.int8)
        ft = k(faces.get("tau", np.zeros(0)), np.float64)
        fcw = k(faces.get("cw", np.zeros(0)), np.float64)
        fjw = k(faces.get("jw", np.zeros(0)), np.float64)
        cv = k(cell_verts, np.int32) if cell_verts is not None else None
        m = MhdMesh(ncells=dofs.shape[0], nfaces_tw=fc.size, cell_xyz=_ptr(xyz), tw_face_cell=_ptr(fc),
                    tw_face_lface=_ptr(fl), tw_tau=_ptr(ft), tw_cw=_ptr(fcw), tw_jw=_ptr(fjw),
                    cell_verts=_ptr(cv), nverts=int(cv.max()) + 1 if cv is not None else 0)
        s = MhdSpaces()
        s.block_style = int(block_style)
        s.nfree[:] = [int(v) for v in nfree4]
        s.ndir = int(len(dir_vals))
        s.cell_dofs, s.rt_sign, s.dir_vals = _ptr(dofs), _ptr(sign), _ptr(dv)
        self.params = make_params(phys)
        h = c_void_p()
        if skeleton is not None:
            # skeleton = dict(cell (nf, 2), lface (nf, 2), perm (nf, 9), mu (nf,))
            sc = k(skeleton["cell"], np.int32)
            sk = MhdSkeleton(nfaces=sc.shape[0], face_cell=_ptr(sc), face_lface=_ptr(k(skeleton["lface"], np.int8)),
                             face_perm=_ptr(k(skeleton["perm"], np.int8)), mu=_ptr(k(skeleton["mu"], np.float64)))
            rc = self.lib.mhd_create_stabilized(ctypes.byref(h), ctypes.byref(m), ctypes.byref(s),
                                                ctypes.byref(self.params), ctypes.byref(sk), device)
        else:
            rc = self.lib.mhd_create(ctypes.byref(h), ctypes.byref(m), ctypes.byref(s), ctypes.byref(self.params),
                                     device)
        self.h = h
        if rc != 0:
            msg = self.lib.mhd_last_error(h).decode() if h.value else "allocation failed"
            self.close()
            raise MhdError(f"mhd_create failed ({rc}): {msg}")
        n, nnz = c_int64(), c_int64()
        self._check(self.lib.mhd_sizes(self.h, ctypes.byref(n), ctypes.byref(nnz)), "mhd_sizes")
        self.n, self.nnz = n.value, nnz.value
        self.ncells = int(dofs.shape[0])

    def _check(self, rc, what):
        if rc < 0:
            raise MhdError(f"{what} failed ({rc}): {self.lib.mhd_last_error(self.h).decode()}")
        return rc

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.mhd_destroy(self.h)
            self.h = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- host-buffer entry points ------------------------------------------
    def pattern(self):
        colptr = np.empty(self.n + 1, dtype=np.int64)
        rowval = np.empty(self.nnz, dtype=np.int64)
        self._check(self.lib.mhd_pattern(self.h, _ptr(colptr), _ptr(rowval)), "mhd_pattern")
        return colptr, rowval

    def residual(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        r = np.empty(self.n)
        self._check(self.lib.mhd_residual(self.h, _ptr(x), _ptr(r)), "mhd_residual")
        return r

    def jacobian(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        nz = np.empty(self.nnz)
        self._check(self.lib.mhd_jacobian(self.h, _ptr(x), _ptr(nz)), "mhd_jacobian")
        return nz

    def cell_integrals(self, x):
        """(ncells, 6): int u.u, int grad u:grad u, int j.j, int (div j)^2, int p, |K|."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.empty((self.ncells, 6))
        self._check(self.lib.mhd_cell_integrals(self.h, _ptr(x), _ptr(out)), "mhd_cell_integrals")
        return out

    def set_assembly_mode(self, mode: int):
        """2 = element buffer + row assembly (default), 1 = colour classes,
        0 = atomic scatter."""
        self._check(self.lib.mhd_set_assembly_mode(self.h, int(mode)), "mhd_set_assembly_mode")

    def set_params(self, phys):
        self.params = make_params(phys)
        self._check(self.lib.mhd_set_params(self.h, ctypes.byref(self.params)), "mhd_set_params")

    def set_dirichlet(self, dir_vals):
        dv = np.ascontiguousarray(dir_vals, dtype=np.float64)
        self._check(self.lib.mhd_set_dirichlet(self.h, _ptr(dv)), "mhd_set_dirichlet")

    # ---- device-pointer entry points (pointers as ints) ---------------------
    def residual_d(self, d_x: int, d_r: int, stream: int = 0):
        return self._check(self.lib.mhd_residual_d(self.h, d_x, d_r, stream or None), "mhd_residual_d")

    def residual_d_noalloc(self, d_x: int, d_r: int, stream: int = 0):
        """mhd_residual_d through the library's fixed device workspace (no allocation)."""
        return self._check(self.lib.mhd_residual_d_noalloc(self.h, d_x, d_r, stream or None),
                           "mhd_residual_d_noalloc")

    def jacobian_d(self, d_x: int, d_val: int = 0, stream: int = 0):
        """CSR values into d_val; d_val = 0 assembles into the context's own
        storage (the Jacobian mhd_precond_setup factors)."""
        return self._check(self.lib.mhd_jacobian_d(self.h, d_x, d_val or None, stream or None), "mhd_jacobian_d")

    def spmv_d(self, d_x: int, d_y: int, stream: int = 0):
        return self._check(self.lib.mhd_spmv_d(self.h, d_x, d_y, stream or None), "mhd_spmv_d")

    def default_opts(self, **kw) -> MhdSolverOpts:
        o = MhdSolverOpts()
        self.lib.mhd_default_solver_opts(ctypes.byref(o))
        for k, v in kw.items():
            setattr(o, k, v)
        return o

    def attach_coarse(self, coarse: "Context", P, R, coarse_iters: int = 0, smoother_weight: float = 0.175,
                      coarse_step: str = "minres"):
        """Hand `coarse` (with the prolongation P and the state interpolation R,
        scipy CSR) to this context.  coarse_step: "full" or "minres"
        (minimal-residual scaling of the coarse correction)."""
        rowptr = np.ascontiguousarray(P.indptr, dtype=np.int64)
        col = np.ascontiguousarray(P.indices, dtype=np.int32)
        val = np.ascontiguousarray(P.data, dtype=np.float64)
        rrow = np.ascontiguousarray(R.indptr, dtype=np.int64)
        rcol = np.ascontiguousarray(R.indices, dtype=np.int32)
        rval = np.ascontiguousarray(R.data, dtype=np.float64)
        self._check(self.lib.mhd_attach_coarse(self.h, coarse.h, P.shape[0], P.shape[1], _ptr(rowptr), _ptr(col),
                                               _ptr(val), _ptr(rrow), _ptr(rcol), _ptr(rval), int(coarse_iters),
                                               float(smoother_weight), COARSE_STEP[coarse_step]),
                    "mhd_attach_coarse")
        coarse.h = c_void_p()          # owned by this context now
        self._coarse = coarse

    def dresidual_dparam(self, x, param: int):
        """dR/dtheta at x: param 0 = gamma (Lorentz term), 1 = cw (thin walls)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.empty(self.n)
        self._check(self.lib.mhd_dresidual_dparam(self.h, _ptr(x), int(param), _ptr(out)), "mhd_dresidual_dparam")
        return out

    def linear_solve(self, x, rhs, opts: Optional[MhdSolverOpts] = None):
        """s = A(x)^-1 rhs on the GPU (forward preconditioner); returns (s, info, rc)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        sol = np.empty(self.n)
        opts = opts or self.default_opts()
        info = MhdSolveInfo()
        rc = self._check(self.lib.mhd_linear_solve(self.h, _ptr(x), _ptr(rhs), _ptr(sol), ctypes.byref(opts),
                                                   ctypes.byref(info)), "mhd_linear_solve")
        return sol, info, rc

    def adjoint_solve(self, x, rhs, opts: Optional[MhdSolverOpts] = None):
        """lambda = A(x)^-T rhs on the GPU; returns (lambda, info, rc)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        lam = np.empty(self.n)
        opts = opts or self.default_opts()
        info = MhdSolveInfo()
        rc = self._check(self.lib.mhd_adjoint_solve(self.h, _ptr(x), _ptr(rhs), _ptr(lam), ctypes.byref(opts),
                                                    ctypes.byref(info)), "mhd_adjoint_solve")
        return lam, info, rc

    def precond_setup(self, opts: MhdSolverOpts):
        return self._check(self.lib.mhd_precond_setup(self.h, ctypes.byref(opts)), "mhd_precond_setup")

    def smoother_apply_d(self, d_r: int, d_y: int, stream: int = 0):
        """y += smoother(r); float64 vectors after a precision-0 set-up, float32 otherwise."""
        return self._check(self.lib.mhd_smoother_apply_d(self.h, d_r, d_y, stream or None), "mhd_smoother_apply_d")

    def set_pivot_threshold(self, tau: float):
        """Diagonal-pivot threshold of the patch inversions (0 .. 1, see mhd.h)."""
        self._check(self.lib.mhd_set_pivot_threshold(self.h, float(tau)), "mhd_set_pivot_threshold")

    def smoother_info(self, pivoted=False):
        """(patches, padded patch size, sum of size^2) [, patches inverted with partial pivoting]."""
        a, b, c_, d = c_int64(), c_int64(), c_int64(), c_int64()
        self._check(self.lib.mhd_smoother_info(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c_),
                                               ctypes.byref(d)), "mhd_smoother_info")
        return (a.value, b.value, c_.value, d.value) if pivoted else (a.value, b.value, c_.value)

    def op_counts(self, reset=False):
        """Operation counts of the solver work since the last reset (dict, see mhd_op_counts)."""
        out = np.zeros(len(OP_COUNT_NAMES))
        self._check(self.lib.mhd_op_counts(self.h, _ptr(out), int(bool(reset))), "mhd_op_counts")
        return dict(zip(OP_COUNT_NAMES, out.tolist()))

    def newton_solve(self, x0, opts: Optional[MhdSolverOpts] = None):
        x = np.array(x0, dtype=np.float64, copy=True)
        if opts is None:
            opts = MhdSolverOpts()
            self.lib.mhd_default_solver_opts(ctypes.byref(opts))
        info = MhdSolveInfo()
        rc = self._check(self.lib.mhd_newton_solve(self.h, _ptr(x), ctypes.byref(opts), ctypes.byref(info)),
                         "mhd_newton_solve")
        return x, info, rc


def context_with_levels(sp, phys, levels=1, coarse_iters=0, smoother_weight=0.175, coarse_step="minres",
                        device=0, coarsen=(0, 1)) -> Context:
    """Fine context plus `levels - 1` semi-coarsened levels attached for the
    [u,j] V-cycle (badia2024 numbering, nested x/y coarsening, the same
    boundary conditions on every level).  `levels="auto"` picks the depth
    from the mesh; `coarse_iters = 0` solves the coarsest level exactly;
    `coarse_step` is "full" or "minres" (see mhd_attach_coarse); `coarsen`
    the halved axes ((0, 1) = x and y, the reference's hierarchy; (0,) keeps y)."""
    from . import fespaces, multigrid
    if getattr(sp.mesh, "nc", None) is None:
        # unstructured (gmsh) meshes have no semi-coarsened hierarchy here
        if levels not in ("auto", 1):
            raise ValueError("the [u,j] V-cycle needs a Cartesian mesh hierarchy; use levels=1 on unstructured meshes")
        levels = 1
    coarsen = tuple(coarsen)
    if levels == "auto":
        levels = multigrid.auto_levels(sp.mesh, axes=coarsen)
    ctx = context_from_spaces(sp, phys, device=device)
    if levels <= 1:
        return ctx
    bc = sp.bc_spec
    # the coarse spaces and transfers depend on the mesh only: kept on the
    # fine spaces for further contexts of the same problem
    caches = sp.__dict__.setdefault("_multigrid_cache", {})
    cache = caches.setdefault(coarsen, {"chain": [sp], "transfers": {}})
    chain = cache["chain"]
    while len(chain) < levels:
        mesh = multigrid.coarsen_mesh(chain[-1].mesh, axes=coarsen)
        chain.append(fespaces.setup_fe_spaces(mesh, bc["u_tags"], bc["j_tags"], bc["thin_walls"], style=sp.style))
    ctxs = [ctx] + [context_from_spaces(c, phys, device=device) for c in chain[1:levels]]
    # attach from the coarsest upwards so each attached context already carries its chain
    for lev in range(levels - 1, 0, -1):
        if lev not in cache["transfers"]:
            cache["transfers"][lev] = multigrid.prolongation(chain[lev - 1], chain[lev])
        P, R = cache["transfers"][lev]
        ctxs[lev - 1].attach_coarse(ctxs[lev], P, R, coarse_iters, smoother_weight, coarse_step)
    return ctx


def context_from_spaces(sp, phys, dir_vals=None, device=0) -> Context:
    """Context for a ``fespaces.Spaces`` object (thin-wall faces concatenated).
    ``phys["stab_mu"]`` > 0 adds the skeleton stabilisation with that mu on
    every interior face."""
    tw = sp.thin_wall_faces
    faces = None
    if tw:
        faces = dict(cell=np.concatenate([t["cell"] for t in tw]), lface=np.concatenate([t["lface"] for t in tw]),
                     tau=np.concatenate([np.full(t["cell"].size, t["tau"]) for t in tw]),
                     cw=np.concatenate([np.full(t["cell"].size, t["cw"]) for t in tw]),
                     jw=np.concatenate([np.full(t["cell"].size, t["jw"]) for t in tw]))
    nf = [sp.nfree["u"], sp.nfree["p"], sp.nfree["j"], sp.nfree["phi"]]
    dv = sp.dir_vals if dir_vals is None else dir_vals
    skel = None
    mu = float(phys.get("stab_mu", 0.0))
    if mu != 0.0:
        from . import fespaces
        skel = dict(fespaces.skeleton_faces(sp))
        skel["mu"] = np.full(skel["cell"].shape[0], mu)
    return Context(sp.mesh.cell_xyz(), sp.cell_dofs, sp.rt_sign, nf, dv, phys, faces, device,
                   cell_verts=sp.cell_verts, block_style=1 if sp.style == "badia2024" else 0, skeleton=skel)
