"""Problem builders shared by the tests: mesh + dof tables + physics, in the
form the oracle and the C-ABI both take (the same inputs on both sides)."""
from __future__ import annotations

import math
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import gridapmhd_amd as g  # noqa: E402


def problem_dict(sp, phys, dir_vals=None):
    """Flatten Spaces + physics into the oracle's problem dict."""
    tw = sp.thin_wall_faces
    fc = np.concatenate([t["cell"] for t in tw]) if tw else np.zeros(0, np.int32)
    fl = np.concatenate([t["lface"] for t in tw]) if tw else np.zeros(0, np.int8)
    ft = np.concatenate([np.full(t["cell"].size, t["tau"]) for t in tw]) if tw else np.zeros(0)
    fcw = np.concatenate([np.full(t["cell"].size, t["cw"]) for t in tw]) if tw else np.zeros(0)
    fjw = np.concatenate([np.full(t["cell"].size, t["jw"]) for t in tw]) if tw else np.zeros(0)
    d = dict(cell_xyz=sp.mesh.cell_xyz(), cell_dofs=sp.cell_dofs, rt_sign=sp.rt_sign,
             dir_vals=sp.dir_vals if dir_vals is None else dir_vals, nfree=sp.ndofs,
             face_cell=fc, face_lface=fl, face_tau=ft, face_cw=fcw, face_jw=fjw)
    d.update(phys)
    return d


def cfd_physics(Ha, Re, B=(0.0, 1.0, 0.0), f=(0.0, 0.0, 1.0), convection="newton", zeta_u=0.0, zeta_j=0.0):
    """Option 1 (:cfd) non-dimensionalisation of src/Applications/hunt.jl:117-128."""
    return dict(alpha=1.0, beta=1.0 / Re, gamma=Ha * Ha / Re, sigma=1.0, B=np.array(B, float),
                f0=np.array(f, float), convection=convection, zeta_u=zeta_u, zeta_j=zeta_j)


def hunt_case(cx, Ha=50.0, nz=3, BL_adapted=False, domain=(-1.0, 1.0, -1.0, 1.0, 0.0, 0.1), convection="newton",
              style="consecutive"):
    """Hunt flow: noslip walls, insulating side walls, B along y, f = e_z."""
    mesh = g.meshers.hunt_generate_base_mesh((cx, cx), Ha=Ha, BL_adapted=BL_adapted, nz=nz, domain=domain)
    sp = g.fespaces.setup_fe_spaces(mesh, "noslip", "insulating", style=style)
    return sp, cfd_physics(Ha, 1.0, convection=convection)


def duct_thin_wall_case(nc, Ha, Re, cw, tau=100.0, domain=(-1.0, 1.0, -1.0, 1.0, 0.0, 0.3),
                        convection="newton", map=None, zeta_u=0.0, zeta_j=0.0, style="consecutive", jitter=0.0):
    """Duct with thin conducting walls on all four walls (no j Dirichlet);
    `jitter` moves interior nodes by U[-jitter h, jitter h]^3 (seed 20261202),
    which makes the cell maps non-affine."""
    mesh = g.meshers.hunt_generate_base_mesh(nc[:2], nz=nc[2], domain=domain, Ha=Ha, map=map or (lambda x: x))
    if jitter:
        g.meshers.jitter_nodes(mesh, jitter, 20261202)
    sp = g.fespaces.setup_fe_spaces(mesh, "noslip", [], [dict(domain="walls", cw=cw, tau=tau, jw=0.0)], style=style)
    return sp, cfd_physics(Ha, Re, convection=convection, zeta_u=zeta_u, zeta_j=zeta_j)


def manufactured_case(nc=(3, 3, 3), B=(0.0, 1.0, 0.0), alpha=1.0, beta=1.0, gamma=1.0, sigma=1.0,
                      jitter=0.0, convection="newton"):
    """Stationary manufactured solution lying in the FE space
    (src/Applications/transient.jl:262-270, 331-344 at d/dt = 0):
    u = (y, x, 0), j = (y, 1, 0), p = 0, phi = 1, with
    f = alpha (u.grad)u - beta lap u + grad p - gamma j x B,
    g = j + sigma grad phi - sigma u x B.  Dirichlet u and j on all walls."""
    mesh = g.meshers.cartesian_model((-1.0, 1.0, -1.0, 1.0, 0.0, 0.3), nc, (False, False, True))
    g.meshers.hunt_add_tags(mesh)
    if jitter:
        g.meshers.jitter_nodes(mesh, jitter, 20261202)
    sp = g.fespaces.setup_fe_spaces(mesh, "walls", "walls")
    Bv = np.array(B, float)
    u = lambda x: np.stack([x[:, 1], x[:, 0], 0 * x[:, 0]], 1)
    j = lambda x: np.stack([x[:, 1], np.ones_like(x[:, 0]), 0 * x[:, 0]], 1)
    xs, dv = g.fespaces.interpolate(sp, u=u, p=lambda x: 0 * x[:, 0], j=j, phi=lambda x: np.ones_like(x[:, 0]))
    # f(x) = f0 + F x ; (u.grad)u = (x, y, 0); j x B with j = (y,1,0)
    F = np.zeros((3, 3))
    f0 = np.zeros(3)
    if convection != "none":
        F[0, 0] += alpha
        F[1, 1] += alpha
    # -(gamma) (j x B): j x B = (jy Bz - jz By, jz Bx - jx Bz, jx By - jy Bx), j = (y, 1, 0)
    # x-comp: Bz ; y-comp: -y Bz ; z-comp: y By - Bx
    f0[0] += -gamma * Bv[2]
    F[1, 1] += gamma * Bv[2]
    F[2, 1] += -gamma * Bv[1]
    f0[2] += gamma * Bv[0]
    # g = j - sigma u x B ; u x B with u = (y, x, 0): (x Bz, -y Bz, y By - x Bx)
    G = np.zeros((3, 3))
    g0 = np.array([0.0, 1.0, 0.0])
    G[0, 1] += 1.0
    G[0, 0] += -sigma * Bv[2]
    G[1, 1] += sigma * Bv[2]
    G[2, 1] += -sigma * Bv[1]
    G[2, 0] += sigma * Bv[0]
    phys = dict(alpha=alpha, beta=beta, gamma=gamma, sigma=sigma, B=Bv, f0=f0, fx=F, g0=g0, gx=G,
                convection=convection)
    return sp, phys, xs, dv


def field_errors(sp, x, xref):
    """Relative errors per field; the pressure is compared modulo its constant
    (the constant of dP1 is the all-ones coefficient vector of every cell)."""
    out = {}
    for f, s in sp.field_slices().items():
        a, b = x[s].copy(), xref[s].copy()
        if f == "p":
            a -= a.mean()
            b -= b.mean()
        nb = max(np.linalg.norm(b), 1e-6 * np.linalg.norm(xref))
        out[f] = np.linalg.norm(a - b) / nb
    xa, xb = x.copy(), xref.copy()
    ps = sp.field_slices()["p"]
    xa[ps] -= xa[ps].mean()
    xb[ps] -= xb[ps].mean()
    out["all"] = np.linalg.norm(xa - xb) / np.linalg.norm(xb)
    return out


# ---------------------------------------------------------------------------
# Oracle fixtures: cases whose oracle direct solves take minutes (8 x 8 x 8
# cells, ~30K dofs; SuperLU factorisations of ~90 s each) are solved once by
# tests/golden/make_oracle_fixtures.py and stored under tests/golden/.  The
# CPU suite re-checks each stored vector against the oracle (residual, and
# the transposed system for adjoints), which takes a second.

FIXTURE_DIR = os.path.join(ROOT, "tests", "golden")

# duct 8 x 8 x 8 with the physics of the 64 x 64 x 8 benchmark duct
C2_PARAMS = dict(nc=(8, 8, 8), Ha=100.0, Re=10.0, cw=0.01, convection="newton")
# developing duct (inlet profile, outlet, thin walls): pressure-drop adjoint
ADJOINT_CASES = {
    "c4_adjoint_8x8x8": dict(nc=(8, 8, 8), Ha=100.0, Re=10.0, cw=0.01, jitter=0.0),
    "c5_adjoint_jitter_cw0.001": dict(nc=(6, 6, 6), Ha=500.0, Re=50.0, cw=0.001, jitter=0.2),
    "c5_adjoint_jitter_cw0.01": dict(nc=(6, 6, 6), Ha=500.0, Re=50.0, cw=0.01, jitter=0.2),
    "c5_adjoint_jitter_cw0.1": dict(nc=(6, 6, 6), Ha=500.0, Re=50.0, cw=0.1, jitter=0.2),
}


def c2_params_case(style="badia2024"):
    return duct_thin_wall_case(style=style, **C2_PARAMS)


def adjoint_case(name, zeta=None):
    """Spaces + physics of a developing-duct adjoint case.  The oracle solves
    the plain method (zeta = 0) on affine cells; on jittered cells the
    div-div terms do not vanish at the discrete solution, so the oracle then
    solves the same augmented problem as the GPU (zeta = 1)."""
    from gridapmhd_amd import adjoint as adj
    kw = dict(ADJOINT_CASES[name])
    jitter = kw.pop("jitter")
    if zeta is None:
        zeta = 1.0 if jitter else 0.0
    return adj.inlet_duct(kw["nc"], kw["Ha"], kw["Re"], kw["cw"], zeta=zeta, jitter=jitter)


EXPANSION_MSH = os.path.join(FIXTURE_DIR, "Expansion_710.msh")
# the expansion application (src/Applications/expansion.jl) on the reference's
# 240-cell gmsh mesh: insulating walls, and thin conducting walls
EXPANSION_CASES = {
    "expansion_710": dict(Ha=1.0, N=1.0, solid_coupling="none"),
    "expansion_710_thin_wall": dict(Ha=10.0, N=10.0, cw=0.028, solid_coupling="thin_wall"),
}


def expansion_case(name):
    """(params, spaces, physics) of an expansion case as main() builds them."""
    import importlib
    from gridapmhd_amd import applications, parameters
    m = importlib.import_module("gridapmhd_amd.main")
    params = applications.expansion_params(mesh=EXPANSION_MSH, **EXPANSION_CASES[name])
    full = parameters.add_default_params(params)
    sp = m.setup_fe_spaces(full)
    return params, sp, parameters.physics(full)


def load_fixture(name):
    with np.load(os.path.join(FIXTURE_DIR, f"oracle_{name}.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}
