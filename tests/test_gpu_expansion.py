"""The expansion application on the reference's unstructured gmsh mesh
(meshes/Expansion_710.msh, 240 hexahedra) through main(params) on the GPU,
against the oracle's stored Newton + direct solutions of the same discrete
problem (tests/golden/oracle_expansion_*.npz): insulating walls (Ha = 1) and
thin conducting walls (Ha = 10, cw = 0.028).  The GPU solve carries the
augmented-Lagrangian weights zeta = 1, which vanish at the discrete solution
on these affine cells; the oracle solves the plain method.  Unstructured
meshes have no semi-coarsened hierarchy here, so the [u,j] block is
preconditioned by the vertex-star Schwarz smoother alone."""
import json
import sys

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _errors(sp, x, xref, free_phi):
    out = {}
    for f, s in sp.field_slices().items():
        a, b = x[s].copy(), xref[s].copy()
        if f == "phi" and free_phi:
            a -= a.mean()
            b -= b.mean()
        out[f] = float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
    xa, xb = x.copy(), xref.copy()
    if free_phi:
        s = sp.field_slices()["phi"]
        xa[s] -= xa[s].mean()
        xb[s] -= xb[s].mean()
    out["all"] = float(np.linalg.norm(xa - xb) / np.linalg.norm(xb))
    return out


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("name", list(cases.EXPANSION_CASES))
def test_expansion_through_main_matches_oracle(name, precision, tmp_path):
    import gridapmhd_amd as g
    from gridapmhd_amd import applications, postprocess
    _say(f"[expansion] case={name} precision={precision}")
    fx = cases.load_fixture(name)
    params = applications.expansion_params(mesh=cases.EXPANSION_MSH, ζᵤ=1.0, ζⱼ=1.0, **cases.EXPANSION_CASES[name])
    params["solver"] = {"solver": "hip", "precision": precision, "rtol": 1e-11, "atol": 1e-15, "niter": 12,
                        "niter_ls": 3000, "restart": 300, "inner_iters": 8}
    xh, full, out = g.main(params)
    sp = xh.spaces
    free_phi = cases.EXPANSION_CASES[name]["solid_coupling"] == "none"
    err = _errors(sp, xh.free_values, fx["x"], free_phi)
    _say(f"  converged={out['converged']} newton={out['nls_iters']} krylov={out.get('ls_iters_per_step')} "
         + " ".join(f"{k}={v:.2e}" for k, v in err.items()))
    assert out["converged"]
    assert err["all"] < (1e-10 if precision == "fp64" else 1e-8), err
    # the solution fields written as VTK (every cell's 8 vertices)
    path = postprocess.write_vtk(str(tmp_path / name), xh)
    assert open(path).read().count("<DataArray") == 3 + 1 + 5
