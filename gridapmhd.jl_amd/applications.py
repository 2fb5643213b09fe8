"""Applications built on ``main``: the Hunt duct and the thin-walled duct.

``hunt`` mirrors ``_hunt`` (src/Applications/hunt.jl:41-304) for the
sequential H1-HDiv case: non-dimensionalisation (:cfd or :mhd), the
Cartesian duct mesh with the reference's tags, noslip walls, insulating side
walls, and the post-processed norms uh_l2, uh_h1, jh_l2 (with the rescaling
u = u0 ubar, j = sigma u0 B0 jbar).  ``thin_wall_duct`` is the fully developed
duct with thin conducting walls on all four walls, the pattern of
src/Applications/FullyDevelopedFlow.jl:96-194 used by the benchmarks.
"""
from __future__ import annotations

import math

import numpy as np

from . import meshers
from .main import main


def hunt(nc=(4, 4), nz=3, ν=1.0, ρ=1.0, σ=1.0, B=(0.0, 10.0, 0.0), f=(0.0, 0.0, 1.0), ζᵤ=0.0, ζⱼ=0.0, L=1.0,
         u0=1.0, B0=None, order=2, order_j=None, formulation="cfd", convection="none", BL_adapted=True,
         kmap_x=1, kmap_y=1, solver="hip", initial_value="zero", domain=(-1.0, 1.0, -1.0, 1.0, 0.0, 0.1),
         debug=False, solve=True, res_assemble=False, jac_assemble=False):
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
    xh, fullparams, out = main(params, output=info)
    ci = out["cell_integrals"].sum(axis=0)
    # rescale: uh = u0 ubar, jh = sigma u0 B0 jbar (src/Applications/hunt.jl:213-218)
    info["uh_l2"] = u0 * math.sqrt(ci[0])
    info["uh_h1"] = u0 * math.sqrt(ci[0] + ci[1])
    info["jh_l2"] = σ * u0 * B0 * math.sqrt(ci[2])
    info["divj_l2"] = σ * u0 * B0 * math.sqrt(ci[3])
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
