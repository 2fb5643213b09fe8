"""Null modes of the discrete problem and the oracle's pins.

The oracle's direct solve pins one dof of each field whose constant is a
null vector of the Jacobian (``oracle.null_fields``): the pressure when u is
prescribed on the whole boundary, the potential when j.n is.  Pinning a field
that is not null drops an equation that the others do not imply, and the
Newton iteration stalls at that equation's residual -- which the Hunt duct
shows (insulating tags on part of its boundary only: the potential is
determined).  CPU only."""
import numpy as np

import cases


def _closed_box(g, u_tags, j_tags):
    m = g.meshers.cartesian_model((-1.0, 1.0, -1.0, 1.0, 0.0, 0.3), (3, 3, 3), (False, False, False))
    g.meshers.hunt_add_tags(m)
    return g.fespaces.setup_fe_spaces(m, u_tags, j_tags, style="badia2024")


def test_null_fields(orc_mod):
    import gridapmhd_amd as g
    phys = cases.cfd_physics(10.0, 1.0)
    sp = _closed_box(g, ["walls", "ends"], ["walls", "ends"])
    assert orc_mod.null_fields(orc_mod.Oracle(cases.problem_dict(sp, phys)), sp) == ["p", "phi"]
    sp = _closed_box(g, "walls", "walls")          # open ends: u and j.n free there
    assert orc_mod.null_fields(orc_mod.Oracle(cases.problem_dict(sp, phys)), sp) == []
    sp, phys = cases.hunt_case(4, Ha=20.0, style="badia2024")
    assert orc_mod.null_fields(orc_mod.Oracle(cases.problem_dict(sp, phys)), sp) == ["p"]


def test_pins_from_null_fields_reach_the_residual_floor(orc_mod):
    sp, phys = cases.hunt_case(4, Ha=20.0, style="badia2024")
    phys = dict(phys, zeta_u=1.0, zeta_j=1.0)
    orc = orc_mod.Oracle(cases.problem_dict(sp, phys))
    nf = orc_mod.null_fields(orc, sp)
    good, bad = {}, {}
    orc_mod.newton_direct(orc, np.zeros(sp.ndofs), rtol=1e-15, atol=1e-16, log=good,
                          pins=orc_mod.null_pins(sp, pin_p="p" in nf, pin_phi="phi" in nf))
    orc_mod.newton_direct(orc, np.zeros(sp.ndofs), rtol=1e-15, atol=1e-16, log=bad,
                          pins=orc_mod.null_pins(sp, pin_p=True, pin_phi=True))
    r0 = good["residuals"][0]
    assert min(good["residuals"]) < 1e-13 * r0
    assert min(bad["residuals"]) > 1e-9 * r0      # the dropped phi equation is not implied
