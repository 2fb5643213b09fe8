"""The stored oracle solutions (tests/golden/oracle_*.npz, written by
make_oracle_fixtures.py) still solve the oracle's equations: the residual of
the stored state and, for the adjoint cases, the transposed-system residual of
the stored adjoint and the gradient recomputed from it.  Each check takes one
residual or Jacobian assembly instead of the minutes of direct solves that
made the fixture."""
import json

import numpy as np
import pytest

import cases


def _oracle(orc_mod, sp, phys):
    return orc_mod.Oracle(cases.problem_dict(sp, phys))


def test_c2_params_fixture_solves_the_oracle(orc_mod):
    fx = cases.load_fixture("c2_params_8x8x8")
    sp, phys = cases.c2_params_case()
    orc = _oracle(orc_mod, sp, phys)
    x = fx["x"]
    r0 = np.linalg.norm(orc.residual(np.zeros(sp.ndofs)))
    rel = np.linalg.norm(orc.residual(x)) / r0
    assert rel < 1e-12, rel
    # a planted mistake (a perturbed state) is seen by the same check
    assert np.linalg.norm(orc.residual(x * (1 + 1e-9))) / r0 > 10 * rel


@pytest.mark.parametrize("name", list(cases.ADJOINT_CASES))
def test_adjoint_fixture_solves_the_oracle(name, orc_mod):
    from gridapmhd_amd import adjoint as adj
    fx = cases.load_fixture(name)
    meta = json.loads(str(fx["meta"]))
    kw = cases.ADJOINT_CASES[name]
    sp, phys = cases.adjoint_case(name)
    orc = _oracle(orc_mod, sp, phys)
    x, lam = fx["x"], fx["lam"]
    r0 = np.linalg.norm(orc.residual(np.zeros(sp.ndofs)))
    assert np.linalg.norm(orc.residual(x)) / r0 < 1e-12
    g = adj.pressure_drop_functional(sp)
    A = orc.jacobian(x)
    # backward error of the transposed solve: the residual against the size of
    # the terms it cancels (lambda reaches ~1e2 and A ~ gamma, while |g| ~ 0.2)
    scale = np.linalg.norm(abs(A).T @ np.abs(lam))
    assert np.linalg.norm(A.T @ lam - g) / scale < 1e-14
    # dR/dgamma and dR/dcw: the residual is affine in both
    dRg = _oracle(orc_mod, sp, dict(phys, gamma=1.0)).residual(x) - _oracle(orc_mod, sp, dict(phys, gamma=0.0)).residual(x)
    b1, b0 = cases.problem_dict(sp, phys), cases.problem_dict(sp, phys)
    b1["face_cw"] = np.ones_like(b1["face_cw"])
    b0["face_cw"] = np.zeros_like(b0["face_cw"])
    dRc = orc_mod.Oracle(b1).residual(x) - orc_mod.Oracle(b0).residual(x)
    Ha, Re = kw["Ha"], kw["Re"]
    assert abs(float(g @ x) / meta["J"] - 1) < 1e-13
    assert abs(-(2 * Ha / Re) * float(lam @ dRg) / meta["dJ_dHa"] - 1) < 1e-12
    assert abs(-float(lam @ dRc) / meta["dJ_dcw"] - 1) < 1e-12


@pytest.mark.parametrize("name", list(cases.EXPANSION_CASES))
def test_expansion_fixture_solves_the_oracle(name, orc_mod):
    fx = cases.load_fixture(name)
    _, sp, phys = cases.expansion_case(name)
    orc = orc_mod.Oracle(cases.problem_dict(sp, phys))
    x = fx["x"]
    r0 = np.linalg.norm(orc.residual(np.zeros(sp.ndofs)))
    rel = np.linalg.norm(orc.residual(x)) / r0
    assert rel < 1e-12, rel
    assert np.linalg.norm(orc.residual(x * (1 + 1e-9))) / r0 > 10 * rel
