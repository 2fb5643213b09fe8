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
