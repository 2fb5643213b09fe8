"""``main(params)`` of the reference, on libmhd.

Mirrors src/main.jl:122-169: parameters -> FE spaces -> operator -> initial
guess -> solver -> solve -> optional residual / Jacobian assembly, with the
reference's phase names in ``output["timings"]`` (``fe_spaces``,
``solver_setup``, ``solve``, ``residual``, ``jacobian``).  The operator and
the Newton / FGMRES / block-preconditioner solve run in libmhd on the GPU;
this module only prepares plain arrays for its C ABI.
"""
from __future__ import annotations

import dataclasses
import time
from typing import Any, Dict, Optional

import numpy as np

from . import _lib, fespaces
from .meshers import CartesianMesh
from .parameters import add_default_params, physics

PRECISION = {"fp64": 0, "mixed": 1, "mixed_bf16": 2, "mixed_fp16": 3}


@dataclasses.dataclass
class MultiFieldFEFunction:
    """Free and Dirichlet values of (u, p, j, phi) on their spaces."""

    spaces: fespaces.Spaces
    free_values: np.ndarray
    dirichlet_values: np.ndarray

    def field(self, name: str) -> np.ndarray:
        return self.free_values[self.spaces.field_slices()[name]]

    def __iter__(self):
        return iter(self.field(f) for f in ("u", "p", "j", "phi"))


class PTimer:
    """Named phase timer (the reference's PTimer tic!/toc!, src/main.jl:131-164)."""

    def __init__(self):
        self.data: Dict[str, float] = {}
        self._t = None

    def tic(self):
        self._t = time.perf_counter()

    def toc(self, name):
        self.data[name] = self.data.get(name, 0.0) + time.perf_counter() - self._t


def _values_fn(values):
    if values is None:
        return None
    if callable(values):
        return values
    v = np.asarray(values, float)
    return lambda x: np.broadcast_to(v, (x.shape[0], v.size)).copy()


def setup_fe_spaces(params) -> fespaces.Spaces:
    """``setup_fe_spaces`` (src/fespaces.jl:13-46) + the solver's multifield style."""
    mesh: CartesianMesh = params["model"]
    bcs = params["bcs"]
    thin = [dict(domain=tw["domain"], cw=tw["cw"], tau=tw["τ"], jw=tw["jw"]) for tw in (bcs["thin_wall"] or [])]
    sp = fespaces.setup_fe_spaces(mesh, bcs["u"]["tags"], bcs["j"]["tags"], thin, style="badia2024")
    # Dirichlet values (TrialFESpace(V, u_bc), src/fespaces.jl:116-124)
    dv = np.zeros(sp.ndir["u"] + sp.ndir["j"])
    for f, sl in (("u", slice(0, sp.ndir["u"])), ("j", slice(sp.ndir["u"], None))):
        vals = bcs[f].get("values")
        if vals is None:
            continue
        if isinstance(vals, (list, tuple)) and len(vals) and (callable(vals[0]) or np.ndim(vals[0]) == 1):
            # one value per Dirichlet tag
            dv[sl] = fespaces.dirichlet_values_by_tag(sp, f, bcs[f]["tags"], [_values_fn(v) for v in vals])
        else:
            fn = _values_fn(vals)
            dv[sl] = (fespaces.interpolate_u if f == "u" else fespaces.interpolate_j)(sp, fn)[1]
    sp.dir_vals = dv
    return sp


def _fe_operator(sp, params) -> _lib.Context:
    """``_fe_operator`` (src/main.jl:207-233): the operator lives in libmhd,
    with the [u,j] V-cycle hierarchy of the solver parameters attached."""
    s = params["solver"]
    return _lib.context_with_levels(sp, physics(params), levels=s.get("levels", 1),
                                    coarse_iters=int(s.get("coarse_iters", 0)),
                                    smoother_weight=float(s.get("smoother_weight", 0.175)),
                                    coarse_step=s.get("coarse_step", "minres"))


def _solver_opts(ctx, params):
    s = params["solver"]
    return ctx.default_opts(max_newton=int(s["niter"]), max_krylov=int(s["niter_ls"]), rtol=float(s["rtol"]),
                            atol=float(s["atol"]), ls_rtol=float(s["rtol"]) / 10.0,
                            precision=PRECISION[s["precision"]], restart=int(s["restart"]),
                            inner_iters=int(s["inner_iters"]), verbose=int(s.get("verbose", 0)))


def initial_guess(x0, sp, ctx, params):
    """``initial_guess`` (src/main.jl:285-315): :zero, a dict of fields, or :solve."""
    if isinstance(x0, str) and x0 == "zero":
        return np.zeros(sp.ndofs)
    if isinstance(x0, dict):
        x, _ = fespaces.interpolate(sp, u=_values_fn(x0.get("u")), p=_values_fn(x0.get("p")),
                                    j=_values_fn(x0.get("j")), phi=_values_fn(x0.get("φ")))
        return x
    if isinstance(x0, str) and x0 == "solve":
        if params["fluid"]["convection"] == "none":
            raise ValueError("Convection must be enabled to use initial guess :solve")
        phys = physics(params)
        ctx.set_params(dict(phys, convection="none"))
        x, _, _ = ctx.newton_solve(np.zeros(sp.ndofs), _solver_opts(ctx, params))
        ctx.set_params(phys)
        return x
    raise ValueError(f"unknown initial guess {x0!r}")


def main(_params: Dict[str, Any], output: Optional[Dict[str, Any]] = None):
    """Solve the inductionless MHD problem; returns (xh, full_params, output)."""
    output = {} if output is None else output
    params = add_default_params(_params)
    t = PTimer()
    t.tic()
    sp = setup_fe_spaces(params)
    t.toc("fe_spaces")
    t.tic()
    ctx = None
    try:
        if params["debug"]:
            # random state on the free dofs (src/main.jl:136-142; the Julia stream is not reproduced)
            x = np.random.default_rng(1234).random(sp.ndofs)
            t.toc("solve")
        else:
            ctx = _fe_operator(sp, params)
            x = initial_guess(params["x0"], sp, ctx, params)
            if params["solve"]:
                opts = _solver_opts(ctx, params)
                t.toc("solver_setup")
                t.tic()
                x, info, rc = ctx.newton_solve(x, opts)
                params["solver"]["solver_postpro"](info, output)
                output["converged"] = bool(info.converged)
                output["solver_times"] = dict(residual=info.t_residual, jacobian=info.t_jacobian,
                                              setup=info.t_setup, solve=info.t_solve)
                t.toc("solve")
            if params["res_assemble"]:
                t.tic()
                output["residual"] = ctx.residual(x)
                t.toc("residual")
            if params["jac_assemble"]:
                t.tic()
                output["jacobian_nzval"] = ctx.jacobian(x)
                t.toc("jacobian")
            output["cell_integrals"] = ctx.cell_integrals(x)
    finally:
        if ctx is not None:
            ctx.close()
    output["timings"] = t.data
    xh = MultiFieldFEFunction(sp, x, sp.dir_vals)
    return xh, params, output
