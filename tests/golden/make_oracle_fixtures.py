"""Solve the oracle once for the cases its direct solver needs minutes for,
and store the results as fixtures (tests/golden/oracle_<name>.npz).

* c2_params_8x8x8: the thin-wall duct at the 64 x 64 x 8 benchmark's physics
  (Ha = 100, Re = 10, cw = 0.01, Newton convection) on 8 x 8 x 8 cells:
  Newton with a direct solve per step (the plain method, zeta = 0).
* the expansion cases (tests/cases.py EXPANSION_CASES): the expansion
  application on the reference's 240-cell gmsh mesh, Newton + direct solve.
* the developing-duct adjoint cases (tests/cases.py ADJOINT_CASES): the
  state (Newton + direct solve), the adjoint lambda = A(x)^-T dJ/dx on the
  same factorisation of A(x) with iterative refinement, and dJ/d(Ha, cw) = -lambda . dR/d(Ha, cw) with
  dR/dgamma and dR/dcw as exact differences (R is affine in both).

Each fixture is re-checked against the oracle by tests/test_fixtures.py
(residual of the stored state, transposed-system residual of the stored
adjoint), so a fixture cannot drift from the oracle unnoticed.
Run: python tests/golden/make_oracle_fixtures.py [name ...]   (about 2-10 min per case)
"""
import json
import os
import sys
import time

import numpy as np
import scipy.sparse as sps
import scipy.sparse.linalg as spla

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "tests"), os.path.join(ROOT, "oracle")):
    sys.path.insert(0, p)

import cases  # noqa: E402
import oracle  # noqa: E402


def pinned(A, pins):
    m = np.ones(A.shape[0])
    m[list(pins)] = 0.0
    return (sps.diags(m) @ A @ sps.diags(m) + sps.diags(1.0 - m)).tocsc()


def newton(orc, x, pins=(), rtol=1e-13, atol=1e-15, maxiter=25):
    r = orc.residual(x)
    hist = [float(np.linalg.norm(r))]
    lu = None
    for _ in range(maxiter):
        if hist[-1] <= max(atol, rtol * hist[0]):
            break
        A = orc.jacobian(x)
        if pins:
            A = pinned(A, pins)
            r = r.copy()
            r[list(pins)] = 0.0
        t = time.perf_counter()
        lu = spla.splu(A, permc_spec="COLAMD")
        x = x - lu.solve(r)
        r = orc.residual(x)
        hist.append(float(np.linalg.norm(r)))
        print(f"    newton |R| = {hist[-1]:.3e}  (factorisation {time.perf_counter() - t:.0f} s)", flush=True)
    return x, hist


def make_c2_params():
    sp, phys = cases.c2_params_case()
    orc = oracle.Oracle(cases.problem_dict(sp, phys))
    x, hist = newton(orc, np.zeros(sp.ndofs), pins=oracle.null_pins(sp))
    meta = dict(case="c2_params_8x8x8", **{k: v for k, v in cases.C2_PARAMS.items()}, zeta=0.0,
                newton_residuals=hist, ndofs=sp.ndofs)
    return dict(x=x, meta=np.array(json.dumps(meta)))


def make_adjoint(name):
    from gridapmhd_amd import adjoint as adj
    kw = cases.ADJOINT_CASES[name]
    sp, phys = cases.adjoint_case(name)
    orc = oracle.Oracle(cases.problem_dict(sp, phys))
    x, hist = newton(orc, np.zeros(sp.ndofs))
    g = adj.pressure_drop_functional(sp)
    A = orc.jacobian(x)
    lu = spla.splu(A, permc_spec="COLAMD")
    lam = lu.solve(g, trans="T")
    # iterative refinement on the same factorisation (the system's condition
    # number reaches ~1e9 on the jittered meshes)
    for _ in range(3):
        lam = lam + lu.solve(g - A.T @ lam, trans="T")
    dRg = (oracle.Oracle(cases.problem_dict(sp, dict(phys, gamma=1.0))).residual(x)
           - oracle.Oracle(cases.problem_dict(sp, dict(phys, gamma=0.0))).residual(x))
    b1, b0 = cases.problem_dict(sp, phys), cases.problem_dict(sp, phys)
    b1["face_cw"] = np.ones_like(b1["face_cw"])
    b0["face_cw"] = np.zeros_like(b0["face_cw"])
    dRc = oracle.Oracle(b1).residual(x) - oracle.Oracle(b0).residual(x)
    Ha, Re = kw["Ha"], kw["Re"]
    J = float(g @ x)
    dHa = -(2.0 * Ha / Re) * float(lam @ dRg)
    dcw = -float(lam @ dRc)
    meta = dict(case=name, **kw, zeta=float(phys["zeta_u"]), newton_residuals=hist, ndofs=sp.ndofs, J=J,
                dJ_dHa=dHa, dJ_dcw=dcw)
    print(f"    J = {J!r}  dJ/dHa = {dHa!r}  dJ/dcw = {dcw!r}", flush=True)
    return dict(x=x, lam=lam, meta=np.array(json.dumps(meta)))


def make_expansion(name):
    _, sp, phys = cases.expansion_case(name)
    orc = oracle.Oracle(cases.problem_dict(sp, phys))
    # the outlet leaves u free (no pressure mode); j.n = 0 on the whole
    # boundary leaves the constant potential free unless thin walls carry current
    pin_phi = cases.EXPANSION_CASES[name]["solid_coupling"] == "none"
    x, hist = newton(orc, np.zeros(sp.ndofs), pins=oracle.null_pins(sp, pin_p=False, pin_phi=pin_phi))
    meta = dict(case=name, **cases.EXPANSION_CASES[name], newton_residuals=hist, ndofs=sp.ndofs)
    return dict(x=x, meta=np.array(json.dumps(meta)))


def main(names):
    oracle.build()
    todo = names or ["c2_params_8x8x8"] + list(cases.ADJOINT_CASES) + list(cases.EXPANSION_CASES)
    for name in todo:
        t = time.perf_counter()
        print(f"[{name}]", flush=True)
        if name == "c2_params_8x8x8":
            data = make_c2_params()
        elif name in cases.EXPANSION_CASES:
            data = make_expansion(name)
        else:
            data = make_adjoint(name)
        np.savez_compressed(os.path.join(HERE, f"oracle_{name}.npz"), **data)
        print(f"[{name}] done in {time.perf_counter() - t:.0f} s", flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])
