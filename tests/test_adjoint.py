"""Oracle adjoint of the pressure drop (CPU): transpose solve + dR/dtheta,
checked against central finite differences of full Newton solves."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

import cases


def _solve(orc_mod, nc, Ha, Re, cw):
    from gridapmhd_amd import adjoint as adj
    sp, phys = adj.inlet_duct(nc, Ha, Re, cw, zeta=0.0)
    orc = orc_mod.Oracle(cases.problem_dict(sp, phys))
    x = orc_mod.newton_direct(orc, np.zeros(sp.ndofs), maxiter=20, rtol=1e-14, atol=1e-15)
    return sp, phys, orc, x


def oracle_gradient(orc_mod, nc, Ha, Re, cw):
    from gridapmhd_amd import adjoint as adj
    sp, phys, orc, x = _solve(orc_mod, nc, Ha, Re, cw)
    g = adj.pressure_drop_functional(sp)
    lam = spla.spsolve(orc.jacobian(x).T.tocsc(), g)
    # the residual is affine in gamma and in cw: exact differences
    p1, p0 = dict(phys, gamma=1.0), dict(phys, gamma=0.0)
    dRg = orc_mod.Oracle(cases.problem_dict(sp, p1)).residual(x) - orc_mod.Oracle(cases.problem_dict(sp, p0)).residual(x)
    b1, b0 = cases.problem_dict(sp, phys), cases.problem_dict(sp, phys)
    b1["face_cw"] = np.ones_like(b1["face_cw"])
    b0["face_cw"] = np.zeros_like(b0["face_cw"])
    dRc = orc_mod.Oracle(b1).residual(x) - orc_mod.Oracle(b0).residual(x)
    return float(g @ x), -(2 * Ha / Re) * float(lam @ dRg), -float(lam @ dRc)


def test_oracle_adjoint_matches_finite_differences(orc_mod):
    from gridapmhd_amd import adjoint as adj
    nc, Ha, Re, cw = (4, 4, 3), 20.0, 5.0, 0.05
    J, dHa, dcw = oracle_gradient(orc_mod, nc, Ha, Re, cw)

    def J_of(Ha_, cw_):
        sp, _, _, x = _solve(orc_mod, nc, Ha_, Re, cw_)
        return adj.pressure_drop_functional(sp) @ x

    h = 1e-4
    fd_Ha = (J_of(Ha * (1 + h), cw) - J_of(Ha * (1 - h), cw)) / (2 * h * Ha)
    fd_cw = (J_of(Ha, cw * (1 + h)) - J_of(Ha, cw * (1 - h))) / (2 * h * cw)
    assert abs(dHa / fd_Ha - 1) < 1e-6, (dHa, fd_Ha)
    assert abs(dcw / fd_cw - 1) < 1e-6, (dcw, fd_cw)
    assert J > 0  # flow from the inlet to the outlet needs a positive pressure drop
