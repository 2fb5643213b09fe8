"""Applications built on ``main``: the Hunt duct and the thin-walled duct.

``hunt`` mirrors ``_hunt`` (src/Applications/hunt.jl:41-304) for the
sequential H1-HDiv case: non-dimensionalisation (:cfd or :mhd), the
Cartesian duct mesh with the reference's tags, noslip walls, insulating side
walls, the post-processed norms uh_l2, uh_h1, jh_l2 (with the rescaling
u = u0 ubar, j = sigma u0 B0 jbar) and the errors eu_h1, eu_l2, ej_l2 against
the analytic solution (``analytic.analytical_hunt_u/j``).  ``thin_wall_duct`` is the fully developed
duct with thin conducting walls on all four walls, the pattern of
src/Applications/FullyDevelopedFlow.jl:96-194 used by the benchmarks.
``expansion`` mirrors ``_expansion`` (src/Applications/expansion.jl:40-258)
on its gmsh hex meshes.
"""
from __future__ import annotations

import math
import os

import numpy as np

from . import analytic, mesh_io, meshers, postprocess
from .main import main


def hunt(nc=(4, 4), nz=3, ν=1.0, ρ=1.0, σ=1.0, B=(0.0, 10.0, 0.0), f=(0.0, 0.0, 1.0), ζᵤ=0.0, ζⱼ=0.0, L=1.0,
         u0=1.0, B0=None, order=2, order_j=None, formulation="cfd", convection="none", BL_adapted=True,
         kmap_x=1, kmap_y=1, solver="hip", initial_value="zero", domain=(-1.0, 1.0, -1.0, 1.0, 0.0, 0.1),
         debug=False, solve=True, res_assemble=False, jac_assemble=False, nsums=10, errors=True, μ=0.0):
    B0 = float(np.linalg.norm(B)) if B0 is None else B0
    order_j = order if order_j is None else order_j
    info = {}
    Re = u0 * L / ν
    Ha = B0 * L * math.sqrt(σ / (ρ * ν))
    N = Ha ** 2 / Re
    fbar = (L / (ρ * u0 ** 2)) * np.asarray(f, float)
    Bbar = np.asarray(B, float) / B0
    if formulation == "cfd":
        α, β, γ = 1.0, 1.0 / Re, N
    elif formulation == "mhd":
        α, β, γ = 1.0 / N, 1.0 / Ha ** 2, 1.0
        fbar = fbar / N
    else:
        raise ValueError("Unknown formulation")
    model = meshers.hunt_generate_base_mesh(nc, L=L, Ha=Ha, kmap_x=kmap_x, kmap_y=kmap_y, BL_adapted=BL_adapted,
                                            nz=nz, domain=domain)
    params = {
        "debug": debug, "solve": solve, "res_assemble": res_assemble, "jac_assemble": jac_assemble,
        "model": model,
        "fluid": {"domain": None, "α": α, "β": β, "γ": γ, "f": tuple(fbar), "B": tuple(Bbar), "ζᵤ": ζᵤ, "ζⱼ": ζⱼ,
                  "convection": convection},
        "fespaces": {"order_u": order, "order_j": order_j},
        "bcs": {"u": {"tags": "noslip"}, "j": {"tags": "insulating"}},
        "solver": solver,
        "x0": initial_value,
    }
    if μ > 0:
        # skeleton stabilisation weight (src/Applications/hunt.jl:52, 191-193)
        params["bcs"]["stabilization"] = {"μ": μ}
    xh, fullparams, out = main(params, output=info)
    # the debug mode builds the spaces only (no operator on the GPU): no norms then
    has_norms = "cell_integrals" in out
    if has_norms:
        ci = out["cell_integrals"].sum(axis=0)
        # rescale: uh = u0 ubar, jh = sigma u0 B0 jbar (src/Applications/hunt.jl:213-218)
        info["uh_l2"] = u0 * math.sqrt(ci[0])
        info["uh_h1"] = u0 * math.sqrt(ci[0] + ci[1])
        info["jh_l2"] = σ * u0 * B0 * math.sqrt(ci[2])
        info["divj_l2"] = σ * u0 * B0 * math.sqrt(ci[3])
    if errors and solve and has_norms:
        # eu = u - uh, ej = j - jh against the analytic Hunt solution with nsums and
        # 2 nsums modes (src/Applications/hunt.jl:237-254); L = 1 here, so the
        # physical and the computational meshes coincide.
        mu, grad_pz = ρ * ν, -float(f[2]) / ρ
        for suffix, ns in (("", nsums), ("_ref", 2 * nsums)):
            e = postprocess.hunt_errors(xh.spaces, xh.free_values, L, mu, σ, grad_pz, Ha, ns, u_scale=u0,
                                        j_scale=σ * u0 * B0)
            info["eu_h1" + suffix], info["eu_l2" + suffix], info["ej_l2" + suffix] = e["eu_h1"], e["eu_l2"], e["ej_l2"]
        info["nsums"] = nsums
    sp = xh.spaces
    info.update(ncells=sp.ncells, ndofs=sp.ndofs, ndofs_u=sp.nfree["u"], ndofs_p=sp.nfree["p"],
                ndofs_j=sp.nfree["j"], ndofs_φ=sp.nfree["phi"], Re=Re, Ha=Ha)
    return info, xh


def thin_wall_duct(nc=(64, 64, 8), Ha=100.0, Re=10.0, cw=0.01, τ=100.0, domain=(-1.0, 1.0, -1.0, 1.0, 0.0, 0.3),
                   convection="newton", ζ=1.0, map=None, solver="hip", **solver_kw):
    """Fully developed duct, thin conducting walls on all four walls, f = e_z, B = e_y (:cfd scaling)."""
    model = meshers.hunt_generate_base_mesh(nc[:2], nz=nc[2], domain=domain, Ha=Ha, map=map or (lambda x: x))
    s = {"solver": solver}
    s.update(solver_kw)
    params = {
        "model": model,
        "fluid": {"domain": None, "α": 1.0, "β": 1.0 / Re, "γ": Ha * Ha / Re, "f": (0.0, 0.0, 1.0),
                  "B": (0.0, 1.0, 0.0), "ζᵤ": ζ, "ζⱼ": ζ, "convection": convection},
        "bcs": {"u": {"tags": "noslip"}, "j": {"tags": []}, "thin_wall": [{"domain": "walls", "cw": cw, "τ": τ}]},
        "solver": s,
    }
    return main(params)


def fully_developed_problem(nc=(3, 3), Ha=10.0, dir_B=(0.0, 1.0, 0.0), b=1.0, cw_Ha=0.0, cw_s=0.0, τ_Ha=100.0,
                            τ_s=100.0, strech_fine=False, ζ=1.0, nz=3):
    """Mesh, boundary conditions and :mhd-scaled fluid parameters of
    ``_FullyDeveloped`` (src/Applications/FullyDevelopedFlow.jl:44-194):
    duct (-b, b) x (-1, 1) x (0, 0.1), periodic in z, Smolentsev stretching
    towards the walls, no-slip everywhere, Hartmann walls (y = +-1) and side
    walls (x = +-b) either insulating (cw = 0: j.n = 0) or thin conducting
    walls with conductance ratio cw and penalty tau.  Returns (params, model)."""
    Re = 1.0
    N = Ha * Ha / Re
    dB = np.asarray(dir_B, float) / np.linalg.norm(dir_B)
    s_Ha = math.sqrt(Ha / (Ha - 1.0))
    s_side = math.sqrt(math.sqrt(Ha) / (math.sqrt(Ha) - 1.0))
    fx = (s_Ha, s_Ha) if strech_fine else (s_side, s_Ha)

    def map1(x):
        y = meshers.strech_mhd(x, (0.0, -b, 0.0, -1.0), fx, (0, 1))
        return meshers.strech_mhd(y, (0.0, b, 0.0, 1.0), fx, (0, 1))

    model = meshers.cartesian_model((-b, b, -1.0, 1.0, 0.0, 0.1), (nc[0], nc[1], nz), (False, False, True), map1)
    meshers.hunt_add_tags(model)
    insulated, thin = [], []
    for tag, cw, tau in (("Ha_walls", cw_Ha, τ_Ha), ("side_walls", cw_s, τ_s)):
        if cw == 0.0:
            insulated.append(tag)
        else:
            thin.append({"domain": tag, "cw": cw, "τ": tau})
    params = {
        "model": model,
        "fluid": {"domain": None, "α": 1.0 / N, "β": 1.0 / Ha ** 2, "γ": 1.0, "f": (0.0, 0.0, 1.0), "B": tuple(dB),
                  "ζᵤ": ζ, "ζⱼ": ζ},
        "bcs": {"u": {"tags": ["Ha_walls", "side_walls"]}, "j": {"tags": insulated}, "thin_wall": thin},
    }
    return params, model


def fully_developed(nc=(3, 3), Ha=10.0, dir_B=(0.0, 1.0, 0.0), b=1.0, cw_Ha=0.0, cw_s=0.0, τ_Ha=100.0, τ_s=100.0,
                    nsums=10, strech_fine=False, ζ=1.0, solver="hip", **solver_kw):
    """``_FullyDeveloped`` (src/Applications/FullyDevelopedFlow.jl:44-270) on
    libmhd: solves the duct and post-processes the mean velocity uh_0, the
    pressure-gradient coefficient kp = 1/uh_0 against the correlation
    (kp_shercliff_cartesian for insulating walls, kp_tillac otherwise) and,
    with insulating side walls, the relative L2 distance of u_h to the
    general Hunt solution with conducting Hartmann walls
    (analytical_general_hunt_u, d_b = cw_Ha)."""
    params, model = fully_developed_problem(nc, Ha, dir_B, b, cw_Ha, cw_s, τ_Ha, τ_s, strech_fine, ζ)
    s = {"solver": solver}
    s.update(solver_kw)
    params["solver"] = s
    info = {}
    xh, full, out = main(params, output=info)
    sp = xh.spaces
    uh_0 = postprocess.mean_uz(sp, xh.free_values)
    kp = 1.0 / uh_0
    if cw_s == 0.0 and cw_Ha == 0.0:
        kp_a = analytic.kp_shercliff_cartesian(b, Ha)
    else:
        kp_a = analytic.kp_tillac(b, Ha, cw_s, cw_Ha)
    info.update(uh_0=uh_0, kp=kp, kp_a=kp_a, dev_kp=100.0 * abs(kp_a - kp) / max(kp_a, kp), nc=nc, Ha=Ha,
                ncells=sp.ncells, ndofs=sp.ndofs, ndofs_u=sp.nfree["u"], ndofs_p=sp.nfree["p"],
                ndofs_j=sp.nfree["j"], ndofs_φ=sp.nfree["phi"])
    if cw_s == 0.0:
        info["e_u_l2_rel"] = postprocess.general_hunt_error(sp, xh.free_values, b, cw_Ha, -1.0, Ha, nsums)
    return info, xh


def setup_expansion_mesh_tags(model) -> None:
    """``setup_expansion_mesh_tags!`` (src/Applications/expansion.jl:357-380)."""
    tags = set(model.groups)
    solid = {"wall_interior", "wall_exterior", "wall_volume"} <= tags
    if "wall" not in tags:
        if not solid:
            raise ValueError("the mesh has neither a 'wall' tag nor the solid wall tags")
        model.add_tag_from_tags("wall", ["wall_interior", "wall_exterior", "wall_volume"])
    if "boundary" not in tags:
        model.add_tag_from_tags("boundary", ["inlet", "outlet", "wall_exterior"] if solid else
                                ["inlet", "outlet", "wall"])
    if "interior" not in tags:
        model.add_tag_from_tags("interior", ["fluid", "wall_volume", "wall_interior"] if solid else ["fluid"])


def u_inlet(inlet, Ha, Z, b):
    """Inlet profiles with mean velocity 1 in the outlet channel
    (src/Applications/expansion.jl:383-408)."""
    if inlet == "parabolic":
        return lambda X: np.stack([36.0 * Z * (X[:, 1] - 1 / Z) * (X[:, 1] + 1 / Z) * (X[:, 2] - b * Z) *
                                   (X[:, 2] + b * Z), 0 * X[:, 0], 0 * X[:, 0]], 1)
    if inlet == "shercliff":
        kp = analytic.kp_shercliff_cartesian(b * Z, Ha / Z)

        def u(X):
            # the duct's own frame: axis x -> z, y scaled by Z, z scaled by Z
            Y = np.stack([X[:, 2] * Z, X[:, 1] * Z, X[:, 0]], 1)
            ux = analytic.analytical_general_hunt_u(b * Z, 0.0, -Z * kp, Ha / Z, 200, Y)[:, 2]
            ux = np.where((np.abs(Y[:, 0]) > b * Z) | (np.abs(Y[:, 1]) > 1.0), 0.0, ux)
            return np.stack([ux, 0 * ux, 0 * ux], 1)
        return u
    if inlet == "constant":
        return lambda X: np.tile([Z, 0.0, 0.0], (X.shape[0], 1))
    raise ValueError(f"unknown inlet {inlet!r}")


def expansion_params(mesh="710", mesh_dir=None, solver="hip", formulation="mhd", Z=4.0, b=0.2, N=1.0, Ha=1.0,
                     cw=0.028, τ=100.0, ζᵤ=0.0, ζⱼ=0.0, inlet="parabolic", initial_value="zero",
                     solid_coupling="none", niter=None, convection="newton", debug=False):
    """The ``params`` of ``_expansion`` (src/Applications/expansion.jl:40-220)
    for ``main``: gmsh mesh and tags, non-dimensionalisation, inlet profile,
    wall and electric boundary conditions, initial guess."""
    if solid_coupling not in ("none", "thin_wall"):
        raise NotImplementedError(f"solid_coupling={solid_coupling!r}: solid regions are not built")
    if formulation not in ("cfd", "mhd"):
        raise ValueError("Unknown formulation")
    msh = mesh if os.path.exists(str(mesh)) else os.path.join(mesh_dir or ".", f"Expansion_{mesh}.msh")
    model = mesh_io.read_gmsh(msh)
    setup_expansion_mesh_tags(model)
    Re = Ha ** 2 / N
    α, β, γ = (1.0, 1.0 / Re, N) if formulation == "cfd" else (1.0 / N, 1.0 / Ha ** 2, 1.0)
    u_in = u_inlet(inlet, Ha, Z, b)
    zero = (0.0, 0.0, 0.0)
    bcs = {"u": {"tags": ["inlet", "wall"], "values": [u_in, zero]}}
    if solid_coupling == "none":
        bcs["j"] = {"tags": ["wall", "inlet", "outlet"], "values": [zero, zero, zero]}
    else:
        bcs["j"] = {"tags": ["inlet", "outlet"], "values": [zero, zero]}
        bcs["thin_wall"] = {"τ": τ, "cw": cw, "domain": ["wall"]}
    x0 = {"u": u_in, "j": zero, "p": 0.0, "φ": 0.0} if initial_value == "inlet" else initial_value
    solver = dict(solver) if isinstance(solver, dict) else solver
    if niter is not None:
        solver = {"solver": solver} if isinstance(solver, str) else solver
        solver["niter"] = niter
    return {
        "debug": debug, "model": model, "x0": x0, "bcs": bcs, "solver": solver,
        "fluid": {"domain": "fluid", "α": α, "β": β, "γ": γ, "f": zero, "B": (0.0, 1.0, 0.0), "ζᵤ": ζᵤ, "ζⱼ": ζⱼ,
                  "convection": convection},
    }


def expansion(mesh="710", mesh_dir=None, title="Expansion", vtk=False, path=".", **kw):
    """``expansion`` / ``_expansion`` (src/Applications/expansion.jl:1-258) on
    a gmsh 4.1 hex mesh: ``mesh`` is a file path, or the suffix of
    ``Expansion_<mesh>.msh`` in ``mesh_dir``; the other keywords are those of
    ``expansion_params``.  Returns (info, xh)."""
    params = expansion_params(mesh=mesh, mesh_dir=mesh_dir, **kw)
    Ha, N = kw.get("Ha", 1.0), kw.get("N", 1.0)
    Re = Ha ** 2 / N
    info = {}
    xh, fullparams, out = main(params, output=info)
    if vtk:
        info["vtk"] = postprocess.write_vtk(os.path.join(path, title), xh)
    sp = xh.spaces
    info.update(ncells=sp.ncells, ndofs=sp.ndofs, ndofs_u=sp.nfree["u"], ndofs_p=sp.nfree["p"],
                ndofs_j=sp.nfree["j"], ndofs_φ=sp.nfree["phi"], Ha=Ha, N=N, Re=Re, mesh=str(mesh), title=title)
    return info, xh
