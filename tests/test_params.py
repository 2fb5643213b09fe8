"""Parameter schema of main(params) (src/parameters.jl) -- CPU only."""
import numpy as np
import pytest

import gridapmhd_amd as g


def _params(**over):
    m = g.meshers.hunt_generate_base_mesh((3, 3), Ha=10.0, BL_adapted=False)
    p = {"model": m, "fluid": {"domain": None, "α": 1.0, "β": 1.0, "γ": 100.0, "B": (0, 1, 0), "f": (0, 0, 1)},
         "bcs": {"u": {"tags": "noslip"}, "j": {"tags": "insulating"}}, "solver": "hip"}
    p.update(over)
    return p


def test_defaults_filled():
    full = g.add_default_params(_params())
    assert full["fluid"]["σ"] == 1.0 and full["fluid"]["ζᵤ"] == 0.0 and full["fluid"]["convection"] == "newton"
    assert full["fespaces"]["q"] == 5 and full["fespaces"]["formulation"] == "H1HDiv"
    assert full["solver"]["niter"] == 20 and full["x0"] == "zero" and full["check_valid"] is True
    assert full["bcs"]["thin_wall"] == []


def test_mandatory_and_unknown_keys():
    p = _params()
    del p["bcs"]
    with pytest.raises(ValueError, match="mandatory"):
        g.add_default_params(p)
    p = _params()
    p["fluid"]["ζ"] = 0.0   # the key FullyDevelopedFlow.jl:155 passes; rejected like the reference
    with pytest.raises(ValueError, match="not a valid key"):
        g.add_default_params(p)
    p = _params(check_valid=False)
    p["fluid"]["ζ"] = 0.0
    g.add_default_params(p)


def test_out_of_scope_inputs_raise():
    with pytest.raises(NotImplementedError):
        g.add_default_params(_params(fespaces={"order_u": 3}))
    with pytest.raises(NotImplementedError):
        g.add_default_params(_params(solver="julia"))
    p = _params()
    p["bcs"]["t"] = [{"domain": "walls", "value": 1.0}]
    with pytest.raises(NotImplementedError):
        g.add_default_params(p)


def test_thin_wall_params():
    p = _params()
    p["bcs"] = {"u": {"tags": "noslip"}, "j": {"tags": []},
                "thin_wall": {"domain": "walls", "cw": 0.01, "τ": 100.0}}
    full = g.add_default_params(p)
    assert full["bcs"]["thin_wall"] == [{"domain": "walls", "cw": 0.01, "τ": 100.0, "jw": 0.0}]
    assert full["bcs"]["j"]["tags"] == []


def test_stabilization_params():
    """params[:bcs][:stabilization] (src/parameters.jl:950-977): mu mandatory,
    the domain defaults to the fluid's; the flat physics carries the summed mu."""
    p = _params()
    p["bcs"]["stabilization"] = {"μ": 0.3}
    full = g.add_default_params(p)
    assert full["bcs"]["stabilization"] == [{"μ": 0.3, "domain": None}]
    assert g.parameters.physics(full)["stab_mu"] == 0.3
    p = _params()
    p["bcs"]["stabilization"] = [{"μ": 0.3}, {"μ": 0.2, "domain": None}]
    assert abs(g.parameters.physics(g.add_default_params(p))["stab_mu"] - 0.5) < 1e-15
    assert g.parameters.physics(g.add_default_params(_params()))["stab_mu"] == 0.0
    p = _params()
    p["bcs"]["stabilization"] = {"domain": None}
    with pytest.raises(ValueError, match="mandatory"):
        g.add_default_params(p)
    p = _params()
    p["bcs"]["stabilization"] = {"μ": 0.3, "domain": "noslip"}
    with pytest.raises(NotImplementedError):
        g.add_default_params(p)


def test_main_debug_mode_builds_spaces_only():
    xh, full, out = g.main(_params(debug=True))
    sp = xh.spaces
    assert sp.style == "badia2024" and xh.free_values.size == sp.ndofs
    assert "fe_spaces" in out["timings"]


def test_hunt_debug_mode_reports_sizes_only():
    """hunt(debug=true) runs main's debug path (spaces and a random state, no
    operator): the dof counts are reported, the solution norms are not."""
    info, xh = g.hunt(nc=(3, 3), B=(0.0, 10.0, 0.0), BL_adapted=False, debug=True, μ=0.2)
    assert info["ndofs"] == xh.spaces.ndofs and "uh_l2" not in info
