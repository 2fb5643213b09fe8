"""End to end on the GPU: hunt() -> main(params) -> libmhd, against the
reference's own published Hunt solutions (tests/golden/gadi_hunt.json)."""
import json
import os
import sys

import pytest

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "gadi_hunt.json")))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_hunt_matches_reference_norms(precision):
    import gridapmhd_amd as g
    print(f"[main] hunt cx=10 Ha=50 precision={precision}", flush=True)
    solver = {"solver": "hip", "rtol": 1e-11, "atol": 1e-14, "niter_ls": 2000, "precision": precision}
    info, xh = g.hunt(nc=(10, 10), B=(0.0, 50.0, 0.0), ζᵤ=1.0, ζⱼ=1.0, BL_adapted=False, solver=solver)
    ref = [r for r in GOLD if r["cx"] == 10 and r["Ha"] == 50.0][0]
    assert info["ndofs"] == ref["ndofs"] and info["converged"]
    for k in ("uh_l2", "uh_h1", "jh_l2"):
        rel = abs(info[k] / ref[k] - 1.0)
        print(f"  {k}: {info[k]:.15g} vs {ref[k]:.15g} rel {rel:.2e}", flush=True)
        assert rel < 1e-9, (k, rel)
    assert info["divj_l2"] < 1e-9 * info["jh_l2"]


def test_thin_wall_duct_main_converges():
    import gridapmhd_amd as g
    print("[main] thin_wall_duct 8x8x3 mixed", flush=True)
    xh, full, out = g.thin_wall_duct(nc=(8, 8, 3), Ha=10.0, Re=1.0, cw=0.01, convection="none",
                                     precision="mixed", rtol=1e-10, atol=1e-14, niter_ls=1000)
    assert out["converged"] and out["nls_iters"] >= 1 and out["ls_iters"] > 0
    assert set(out["timings"]) >= {"fe_spaces", "solver_setup", "solve"}


def test_hunt_analytic_errors_match_reference():
    """eu_h1 / ej_l2 of the GPU Hunt solution against the analytic series equal
    the reference's published values (cx = 10, Ha = 50, nsums = 500)."""
    import gridapmhd_amd as g
    print("[main] hunt cx=10 Ha=50 analytic errors", flush=True)
    ref = [r for r in GOLD if r["cx"] == 10 and r["Ha"] == 50.0][0]
    solver = {"solver": "hip", "rtol": 1e-12, "atol": 1e-16, "niter_ls": 2000, "precision": "fp64"}
    info, _ = g.hunt(nc=(10, 10), B=(0.0, 50.0, 0.0), ζᵤ=1.0, ζⱼ=1.0, BL_adapted=False, solver=solver,
                     nsums=ref["nsums"])
    for k in ("eu_h1", "ej_l2"):
        rel = abs(info[k] / ref[k] - 1.0)
        print(f"  {k}: {info[k]:.15g} vs {ref[k]:.15g} rel {rel:.2e}", flush=True)
        assert rel < 1e-9, (k, rel)


def test_fully_developed_thin_wall_duct():
    """FullyDeveloped duct (Ha = 10, thin conducting Hartmann walls cw = 0.1
    with the reference's default penalty tau = 100, insulating side walls)
    through main(params) on the GPU: on 8 x 8 x 3 cells the mean velocity and
    the distance to analytical_general_hunt_u equal the oracle's (CPU direct
    solve: uh_0 = 4.256257, e_u = 3.536e-3); on 16 x 16 x 3 the distance is
    that of the tau = 100 penalty limit (about 1e-2).  The h -> 0,
    tau -> infinity convergence to the analytic solution is pinned on the CPU
    (tests/test_oracle.py)."""
    import gridapmhd_amd as g
    errs = {}
    for n in (8, 16):
        info, _ = g.fully_developed(nc=(n, n), Ha=10.0, cw_Ha=0.1, τ_Ha=100.0, nsums=50, rtol=1e-8, atol=1e-16,
                                    niter=8, niter_ls=3000, precision="fp64")
        errs[n] = info["e_u_l2_rel"]
        print(f"[main] fully developed {n}x{n}x3 cw_Ha=0.1 tau=100: e_u={errs[n]:.4e} uh_0={info['uh_0']:.8f} "
              f"kp={info['kp']:.5f} kp_tillac={info['kp_a']:.5f} converged={info['converged']} "
              f"ls={info['ls_iters_per_step']}", flush=True)
        assert info["converged"]
        if n == 8:
            assert abs(info["uh_0"] / 4.256257 - 1) < 1e-6 and abs(errs[8] / 3.536e-3 - 1) < 1e-3
    assert 5e-3 < errs[16] < 2e-2, errs


def test_initial_guesses_through_main():
    """initial_guess (src/main.jl:285-315) on the GPU: :zero, a Dict of fields
    interpolated into the spaces, and :solve (a Stokes solve with convection
    off, then Newton) all reach the same Newton solution of a convective duct;
    :solve starts Newton closer (smaller first residual).  The walls carry a
    lid-like tangential velocity (zero on x = +-1 and y = -1, (1 - x^2)/2 along x
    on y = 1), so the flow is not unidirectional and convection does not vanish."""
    import gridapmhd_amd as g
    import numpy as np

    def lid(X):
        v = np.zeros((X.shape[0], 3))
        v[:, 0] = 0.25 * (1.0 - X[:, 0] ** 2) * (1.0 + X[:, 1])
        return v

    def params(x0):
        model = g.meshers.hunt_generate_base_mesh((8, 8), nz=4, domain=(-1.0, 1.0, -1.0, 1.0, 0.0, 0.3), Ha=20.0,
                                                  map=lambda x: x)
        return {"model": model, "x0": x0,
                "fluid": {"domain": None, "α": 1.0, "β": 1.0 / 20.0, "γ": 400.0 / 20.0, "f": (0.0, 0.0, 1.0),
                          "B": (0.0, 1.0, 0.0), "ζᵤ": 1.0, "ζⱼ": 1.0, "convection": "newton"},
                "bcs": {"u": {"tags": "noslip", "values": lid}, "j": {"tags": []},
                        "thin_wall": [{"domain": "walls", "cw": 0.05, "τ": 100.0}]},
                "solver": {"solver": "hip", "rtol": 1e-16, "atol": 1e-12, "niter_ls": 2000, "precision": "fp64"}}

    out = {}
    sols = {}
    guesses = {"zero": "zero", "solve": "solve",
               "dict": {"u": (0.0, 0.0, 0.5), "p": 0.0, "j": (0.0, 0.0, 0.0), "φ": 0.0}}
    for name, x0 in guesses.items():
        xh, full, o = g.main(params(x0))
        out[name], sols[name] = o, xh.free_values
        print(f"[main] initial guess {name}: converged={o['converged']} nls={o['nls_iters']} "
              f"residuals={['%.2e' % r for r in o['nls_residuals']]}", flush=True)
        assert o["converged"]
    sp = xh.spaces
    ps = sp.field_slices()["p"]
    for name in ("solve", "dict"):
        a, b = sols[name].copy(), sols["zero"].copy()
        a[ps] -= a[ps].mean()
        b[ps] -= b[ps].mean()
        assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-9, name
    assert out["solve"]["nls_residuals"][0] < out["zero"]["nls_residuals"][0]
