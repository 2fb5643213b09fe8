"""Adjoint sensitivities of the pressure drop with respect to (Ha, cw).

The reference has no adjoint (no rrule, no ChainRules); this module serves
the differentiation configuration: a developing duct along z with an inlet
profile at z = 0, a do-nothing outlet at z = L (the inlet/outlet pattern of
src/Applications/channel.jl:251-271 turned to z), j.n = 0 on both ends and
thin conducting walls on the four side walls.  The objective is the pressure
drop  J = mean_{z=0} p - mean_{z=L} p  (face-area means of the dP1 field).

Gradient:  A(x)^T lambda = dJ/dx,   dJ/dtheta = - lambda . dR/dtheta,
with dR/dHa = (2 Ha / Re) dR/dgamma (:cfd scaling, gamma = Ha^2 / Re) and
dR/dcw the thin-wall term; A^T solve and dR/dtheta run in libmhd.
``rrule`` returns the value and a pullback, the shape of a ChainRules rrule.
"""
from __future__ import annotations

import numpy as np

from . import _lib, fespaces, meshers

_G = (0.5 - np.sqrt(0.15), 0.5, 0.5 + np.sqrt(0.15))
_W = (5.0 / 18.0, 8.0 / 18.0, 5.0 / 18.0)


def inlet_duct(nc, Ha, Re, cw, tau=100.0, U=1.0, domain=(-1.0, 1.0, -1.0, 1.0, 0.0, 0.3), convection="newton",
               zeta=1.0, style="badia2024", jitter=0.0, seed=20261202):
    """Spaces + physics of the developing thin-walled duct; inlet u = (0, 0, U (1-x^2)(1-y^2)).
    `jitter` > 0 moves the interior nodes by U[-jitter h, jitter h]^3 (seeded):
    non-affine cells, with the inlet and outlet planes left flat."""
    mesh = meshers.cartesian_model(domain, nc, (False, False, False))
    meshers.hunt_add_tags(mesh)
    if jitter:
        meshers.jitter_nodes(mesh, jitter, seed)
    sp = fespaces.setup_fe_spaces(mesh, ["walls", "inlet"], "ends",
                                  [dict(domain="walls", cw=cw, tau=tau, jw=0.0)], style=style)
    prof = lambda x: np.stack([0 * x[:, 0], 0 * x[:, 0], U * (1 - x[:, 0] ** 2) * (1 - x[:, 1] ** 2)], 1)
    sp.dir_vals[:sp.ndir["u"]] = fespaces.interpolate_u(sp, prof)[1]
    phys = dict(alpha=1.0, beta=1.0 / Re, gamma=Ha * Ha / Re, sigma=1.0, B=np.array([0.0, 1.0, 0.0]),
                f0=np.zeros(3), convection=convection, zeta_u=zeta, zeta_j=zeta)
    return sp, phys


def pressure_drop_functional(sp) -> np.ndarray:
    """g with J = g . x: face means of p on z = z0 minus on z = z1."""
    mesh = sp.mesh
    nx, ny, nz = mesh.nc
    xyz = mesh.cell_xyz()
    g = np.zeros(sp.ndofs)
    for side, cells, sgn in ((0, np.arange(nx * ny), 1.0), (1, np.arange((nz - 1) * nx * ny, nz * nx * ny), -1.0)):
        acc = np.zeros((cells.size, 4))
        area = 0.0
        for a in range(3):
            for b in range(3):
                xi = np.array([[_G[a], _G[b], float(side)]])
                N = np.array([1 - xi[0, 0] - xi[0, 1] - xi[0, 2], xi[0, 0], xi[0, 1], xi[0, 2]])
                J = fespaces.map_points(xyz[cells], xi)[1][:, 0]          # (c, 3, 3)
                dA = np.linalg.norm(np.cross(J[:, :, 0], J[:, :, 1]), axis=1) * _W[a] * _W[b]
                acc += dA[:, None] * N[None, :]
                area += dA.sum()
        pd = sp.cell_dofs[cells, 81:85]
        np.add.at(g, pd.ravel(), (sgn * acc / area).ravel())
    return g


def solve_state(ctx, sp, opts):
    x, info, rc = ctx.newton_solve(np.zeros(sp.ndofs), opts)
    if rc != 0:
        raise RuntimeError(f"state solve did not converge (rc={rc})")
    return x


def gradient(ctx, sp, x, Ha, Re, opts=None, info_out=None):
    """(J, dJ/dHa, dJ/dcw) at the converged state x.  `info_out` (a dict)
    receives the adjoint solve's iteration count and relative residual."""
    g = pressure_drop_functional(sp)
    lam, info, rc = ctx.adjoint_solve(x, g, opts)
    if info_out is not None:
        info_out.update(krylov=info.krylov_iters, rc=rc, rel_residual=info.residual / max(info.residual0, 1e-300))
    if rc != 0:
        raise RuntimeError(f"adjoint solve did not converge (rc={rc})")
    dR_dgamma = ctx.dresidual_dparam(x, 0)
    dR_dcw = ctx.dresidual_dparam(x, 1)
    J = float(g @ x)
    return J, -(2.0 * Ha / Re) * float(lam @ dR_dgamma), -float(lam @ dR_dcw)


def tangent_gradient(ctx, sp, x, Ha, Re, opts=None, info_out=None):
    """(J, dJ/dHa, dJ/dcw) by forward sensitivities: s = A^-1 (-dR/dtheta),
    dJ/dtheta = g . s, one linear solve per parameter with the forward
    preconditioner (the same numbers as the adjoint form up to the solve
    tolerance)."""
    g = pressure_drop_functional(sp)
    out = []
    for param in (0, 1):
        s, info, rc = ctx.linear_solve(x, -ctx.dresidual_dparam(x, param), opts)
        if info_out is not None:
            info_out[f"param{param}"] = dict(krylov=info.krylov_iters, rc=rc,
                                             rel_residual=info.residual / max(info.residual0, 1e-300))
        if rc != 0:
            raise RuntimeError(f"tangent solve {param} did not converge (rc={rc})")
        out.append(float(g @ s))
    return float(g @ x), (2.0 * Ha / Re) * out[0], out[1]


def rrule(nc, Ha, Re, cw, opts_kw=None, levels=1, coarse_iters=0, coarse_step="minres", **case_kw):
    """J(Ha, cw) and its pullback  dJ -> (dJ * dJ/dHa, dJ * dJ/dcw).  With
    `levels` > 1 (or "auto") the state and the adjoint solves both use the
    [u,j] V-cycle (the adjoint on the transposed operators of every level)."""
    sp, phys = inlet_duct(nc, Ha, Re, cw, **case_kw)
    ctx = _lib.context_with_levels(sp, phys, levels=levels, coarse_iters=coarse_iters, coarse_step=coarse_step)
    try:
        opts = ctx.default_opts(**(opts_kw or {}))
        x = solve_state(ctx, sp, opts)
        try:
            J, dHa, dcw = gradient(ctx, sp, x, Ha, Re, opts)
        except RuntimeError:
            # the transposed solve stalls on larger meshes (measured from
            # 16 x 16 x 8 up); the forward sensitivities give the same gradient
            J, dHa, dcw = tangent_gradient(ctx, sp, x, Ha, Re, opts)
    finally:
        ctx.close()
    return J, (lambda dJ: (dJ * dHa, dJ * dcw))
