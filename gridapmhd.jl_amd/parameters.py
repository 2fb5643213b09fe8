"""Parameter schema of ``GridapMHD.main`` restricted to the hot path.

Mirrors src/parameters.jl: ``add_default_params`` (:85-137), the
mandatory/optional/unused-key checks (:3-76, here one ``KeySchema`` per
dict, with the same error messages), ``params_fluid`` (:693-723),
``params_bcs`` and its sub-dicts (:777-947), ``params_fespaces`` (:350-393)
and ``default_solver_params`` (:219-287).  Keys keep the reference's names
(``"α"``, ``"ζᵤ"``, ``"thin_wall"``, ``"τ"``...).  Inputs outside the
hexahedral H1-HDiv path (other formulations, solids, transient,
continuation, multigrid) raise ``NotImplementedError``.
"""
from __future__ import annotations

import copy
from typing import Any, Dict

import numpy as np


class ParamsError(ValueError):
    pass


REQUIRED = object()   # the value of a mandatory key in a KeySchema


class KeySchema:
    """The keys one parameter dict may hold, each with its default value or
    ``REQUIRED``.  ``validate`` returns a copy with the defaults filled in and
    raises ParamsError -- with the reference's messages, which callers match
    on -- for a value that is not a dict, a missing mandatory key and, unless
    params[:check_valid] is false, a key the schema does not know."""

    def __init__(self, spec: Dict[str, Any]):
        self.spec = dict(spec)

    def validate(self, sub, params=None, where=""):
        if not isinstance(sub, dict):
            raise ParamsError(f"The params{where} has to be a dict")
        missing = [k for k, d in self.spec.items() if d is REQUIRED and k not in sub]
        if missing:
            raise ParamsError(f"Key :{missing[0]} is a mandatory key in params{where}, but it is not provided.")
        out = {k: d for k, d in self.spec.items() if d is not REQUIRED}
        out.update(sub)
        if (out if params is None else params).get("check_valid", True):
            unknown = [k for k in out if k not in self.spec]
            if unknown:
                raise ParamsError(f"Key :{unknown[0]} is not a valid key in params{where}. "
                                  f"Set params[:check_valid] = false to ignore invalid keys.")
        return out

    def validate_each(self, value, params, where):
        """A dict or a list of dicts (the weak-BC entries), as a list."""
        if isinstance(value, dict):
            return [self.validate(value, params, where)]
        if isinstance(value, (list, tuple)):
            return [self.validate(v, params, f"{where}[{i}]") for i, v in enumerate(value)]
        raise ParamsError(f"The value params{where} has to be a dict or a list of dicts.")


# ---------------------------------------------------------------------------
def default_solver_params(name: str) -> Dict[str, Any]:
    """``default_solver_params(::Val{S})``.  ``"hip"`` is this backend; the
    Badia2024 name maps onto it (same preconditioner, on the GPU)."""
    if name in ("hip", "badia2024"):
        return {
            "solver": name,
            "matrix_type": "SparseMatrixCSC{Float64,Int64}",
            "vector_type": "Vector{Float64}",
            "solver_postpro": hip_postpro,
            "niter": 20,          # Newton iterations
            "niter_ls": 400,      # FGMRES iterations per Newton step
            "rtol": 1e-6,
            "atol": 1e-8,
            "precision": "fp64",  # "fp64" | "mixed" | "mixed_bf16" | "mixed_fp16"
            "restart": 200,
            "inner_iters": 3,     # flexible-GMRES steps of the inexact [u,j] solve
            "levels": "auto",     # [u,j] V-cycle levels (x/y semi-coarsening); 1 = Schwarz only
            "coarse_iters": 0,    # coarsest-level solve: 0 = exact (dense inverse), k = k Schwarz-GMRES steps
            "smoother_weight": 0.175,
            "coarse_step": "minres",  # coarse correction step: "minres" (robust) | "full"
            "block_solvers": ["hip_vertex_schwarz", "mass_inverse", "mass_inverse"],
        }
    raise NotImplementedError(f"solver {name!r}: this backend provides 'hip' (and 'badia2024')")


def hip_postpro(info, output):
    """``BlockSolver_postpro`` (src/parameters.jl:323-338) for the GPU solver."""
    output["nls_iters"] = int(info.newton_iters)
    output["ls_iters"] = int(info.krylov_iters)
    output["nls_residuals"] = [float(info.newton_residuals[k]) for k in range(min(64, info.newton_iters + 1))]
    output["ls_iters_per_step"] = [int(info.krylov_per_step[k]) for k in range(min(64, info.newton_iters))]


def params_solver(params):
    s = params["solver"]
    if isinstance(s, str):
        return default_solver_params(s)
    if "solver" not in s:
        raise ParamsError("params[:solver] needs the key :solver")
    out = dict(default_solver_params(s["solver"]))
    out.update(s)
    return out


def params_fespaces(params):
    fe = params.get("fespaces") or {}
    order_u = fe.get("order_u", 2) if isinstance(fe, dict) else 2
    fe = KeySchema({"order_u": 2, "order_j": order_u, "p_constraint": None, "φ_constraint": None, "rt_scaling": None,
                    "fluid_disc": "Qk_dPkm1", "current_disc": "RT"}).validate(fe, params, "[:fespaces]")
    if fe["order_u"] != 2 or fe["order_j"] != 2:
        raise NotImplementedError("the GPU kernels implement order_u = order_j = 2 (Q2 / RT1)")
    if fe["fluid_disc"] != "Qk_dPkm1" or fe["current_disc"] != "RT":
        raise NotImplementedError("the GPU kernels implement the :Qk_dPkm1 / :RT hexahedral pair (H1HDiv)")
    if fe["p_constraint"] is not None or fe["φ_constraint"] is not None or fe["rt_scaling"] is not None:
        raise NotImplementedError("p / φ constraints and rt_scaling are not supported")
    fe["formulation"] = "H1HDiv"
    fe["order_p"] = fe["order_φ"] = 1
    fe["k"] = 2
    # quadrature degree q = max(2(k-1), 3k-1, 2k, k+k, 2(k-1)) = 5  (src/parameters.jl:382-388)
    fe["q"] = max(2 * (2 - 1), 3 * 2 - 1, 2 * 2, 2 + 2, 2 * (1 - 1))
    return fe


def params_fluid(params):
    fluid = KeySchema({"domain": REQUIRED, "α": REQUIRED, "β": REQUIRED, "γ": REQUIRED, "B": REQUIRED,
                       "f": (0.0, 0.0, 0.0), "σ": 1.0, "ζᵤ": 0.0, "ζⱼ": 0.0, "g": (0.0, 0.0, 0.0), "divg": 0.0,
                       "convection": "newton", "μ": 2 * 3}).validate(params["fluid"], params, "[:fluid]")
    if fluid["convection"] not in ("none", "newton", "picard"):
        raise ParamsError("fluid[:convection] must be one of none, newton, picard")
    dom = fluid["domain"]
    if dom is not None:
        # a tag is accepted when it covers every cell (the fluid is the whole mesh)
        model = params.get("model")
        groups = getattr(model, "groups", None)
        cells_in = groups is not None and dom in groups and all(
            tuple(e) in groups[dom] for e in getattr(model, "cell_entity", []))
        if not cells_in:
            raise NotImplementedError("fluid sub-domains (solid regions) are outside the hot path")
    return fluid


def params_bcs(params):
    # the H1-HDiv formulation makes :j mandatory (src/parameters.jl:778-786)
    bcs = KeySchema({"u": REQUIRED, "j": REQUIRED, "φ": [], "t": [], "thin_wall": [],
                     "stabilization": []}).validate(params["bcs"], params, "[:bcs]")
    dirichlet = KeySchema({"tags": REQUIRED, "values": None})
    bcs["u"] = dirichlet.validate(bcs["u"], params, "[:bcs][:u]")
    if bcs["j"] != []:
        bcs["j"] = dirichlet.validate(bcs["j"], params, "[:bcs][:j]")
    else:
        bcs["j"] = {"tags": [], "values": None}
    if bcs["φ"] != []:
        raise NotImplementedError("weak potential BCs (res_φ_bcs) are not part of the GPU path")
    if bcs["t"] != []:
        raise NotImplementedError("Boundary tranction not yet implemented")  # as src/weakforms.jl:164-166
    if bcs["stabilization"] != []:
        # src/parameters.jl:968-977: the domain defaults to the fluid's
        fluid_dom = params["fluid"]["domain"]
        bcs["stabilization"] = KeySchema({"domain": fluid_dom, "μ": REQUIRED}).validate_each(
            bcs["stabilization"], params, "[:bcs][:stabilization]")
        for st in bcs["stabilization"]:
            if st["domain"] != fluid_dom:
                raise NotImplementedError("skeleton stabilisation on a domain other than the fluid's")
    if bcs["thin_wall"] != []:
        bcs["thin_wall"] = KeySchema({"domain": REQUIRED, "cw": REQUIRED, "τ": REQUIRED, "jw": 0.0}).validate_each(
            bcs["thin_wall"], params, "[:bcs][:thin_wall]")
    return bcs


def add_default_params(_params: Dict[str, Any]) -> Dict[str, Any]:
    """``add_default_params`` (src/parameters.jl:85-137)."""
    params = KeySchema({"ptimer": None, "debug": False, "solve": True, "res_assemble": False, "jac_assemble": False,
                        "model": REQUIRED, "fluid": REQUIRED, "solid": None, "bcs": REQUIRED, "fespaces": None,
                        "solver": REQUIRED, "multigrid": None, "check_valid": True, "transient": None,
                        "continuation": None, "x0": "zero"}).validate(_params)
    for k in ("solid", "multigrid", "transient", "continuation"):
        if params[k] is not None:
            raise NotImplementedError(f"params[:{k}] is outside the stationary H1-HDiv hot path")
    params = dict(params)
    params["fespaces"] = params_fespaces(params)
    params["fluid"] = params_fluid(params)
    params["bcs"] = params_bcs(params)
    params["solver"] = params_solver(params)
    return params


def duplicate_params(params):
    """``duplicate_params`` (src/parameters.jl:148-158)."""
    return {k: duplicate_params(v) if isinstance(v, dict) else v for k, v in params.items()}


def fluid_physics(fluid) -> Dict[str, Any]:
    """The fluid dict as the flat parameter set of the C ABI."""
    def vec_or_affine(v):
        # constant vector, or (v0, M) meaning v0 + M x
        if isinstance(v, tuple) and len(v) == 2 and np.ndim(v[1]) == 2:
            return np.asarray(v[0], float), np.asarray(v[1], float)
        return np.asarray(v, float), np.zeros((3, 3))
    f0, fx = vec_or_affine(fluid["f"])
    g0, gx = vec_or_affine(fluid["g"])
    return dict(alpha=float(fluid["α"]), beta=float(fluid["β"]), gamma=float(fluid["γ"]), sigma=float(fluid["σ"]),
                B=np.asarray(fluid["B"], float), f0=f0, fx=fx, g0=g0, gx=gx, zeta_u=float(fluid["ζᵤ"]),
                zeta_j=float(fluid["ζⱼ"]), convection=fluid["convection"])


def physics(params) -> Dict[str, Any]:
    """The flat parameter set of the C ABI for processed ``params``: the fluid
    terms plus the skeleton stabilisation weight (the entries' mu summed: each
    covers the whole fluid skeleton)."""
    phys = fluid_physics(params["fluid"])
    phys["stab_mu"] = float(sum(float(st["μ"]) for st in (params["bcs"].get("stabilization") or [])))
    return phys
