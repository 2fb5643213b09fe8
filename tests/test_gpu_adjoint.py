"""GPU adjoint (transpose FGMRES + dR/dtheta kernels, the rrule of the
differentiation configuration) against the oracle adjoint and against
central finite differences of GPU Newton solves (h = 1e-6 relative)."""
import numpy as np
import pytest

import cases
from test_adjoint import oracle_gradient

pytestmark = pytest.mark.gpu


def test_gpu_adjoint_matches_oracle_and_fd(orc_mod):
    import gridapmhd_amd as g
    from gridapmhd_amd import adjoint as adj
    nc, Ha, Re, cw = (4, 4, 3), 20.0, 5.0, 0.05
    print("[adjoint] developing duct 4x4x3", flush=True)
    okw = dict(rtol=1e-12, atol=1e-15, ls_rtol=1e-12, max_krylov=2000)
    J, pullback = adj.rrule(nc, Ha, Re, cw, opts_kw=okw)
    dHa, dcw = pullback(1.0)
    Jo, dHa_o, dcw_o = oracle_gradient(orc_mod, nc, Ha, Re, cw)
    print(f"  J={J:.12g} oracle {Jo:.12g}; dHa {dHa:.10g} oracle {dHa_o:.10g}; dcw {dcw:.10g} oracle {dcw_o:.10g}",
          flush=True)
    assert abs(J / Jo - 1) < 1e-9
    assert abs(dHa / dHa_o - 1) < 1e-8 and abs(dcw / dcw_o - 1) < 1e-8

    def J_of(Ha_, cw_):
        J_, _ = adj.rrule(nc, Ha_, Re, cw_, opts_kw=okw)
        return J_

    h = 1e-6
    fd_Ha = (J_of(Ha * (1 + h), cw) - J_of(Ha * (1 - h), cw)) / (2 * h * Ha)
    fd_cw = (J_of(Ha, cw * (1 + h)) - J_of(Ha, cw * (1 - h))) / (2 * h * cw)
    print(f"  FD: dHa {fd_Ha:.10g} dcw {fd_cw:.10g}", flush=True)
    assert abs(dHa / fd_Ha - 1) < 1e-6 and abs(dcw / fd_cw - 1) < 1e-6


def test_gpu_adjoint_multilevel_matches_oracle(orc_mod):
    """The adjoint through the V-cycle hierarchy: every level carries its
    transposed Jacobian, the transfers are unchanged."""
    from gridapmhd_amd import adjoint as adj
    nc, Ha, Re, cw = (8, 8, 3), 20.0, 5.0, 0.05
    print("[adjoint] developing duct 8x8x3, 2 levels", flush=True)
    # the transposed V-cycle converges in ~40 FGMRES iterations to a floor
    # near 2e-12 relative at this size: ask for 1e-11
    okw = dict(rtol=1e-12, atol=1e-15, ls_rtol=1e-11, max_krylov=2000, inner_iters=3, restart=200)
    J, pullback = adj.rrule(nc, Ha, Re, cw, opts_kw=okw, levels=2)
    dHa, dcw = pullback(1.0)
    Jo, dHa_o, dcw_o = oracle_gradient(orc_mod, nc, Ha, Re, cw)
    print(f"  J={J:.12g} oracle {Jo:.12g}; dHa {dHa:.10g} oracle {dHa_o:.10g}; dcw {dcw:.10g} oracle {dcw_o:.10g}",
          flush=True)
    assert abs(J / Jo - 1) < 1e-9
    assert abs(dHa / dHa_o - 1) < 1e-8 and abs(dcw / dcw_o - 1) < 1e-8


def test_gpu_tangent_matches_oracle(orc_mod):
    """Forward sensitivities (mhd_linear_solve) give the same gradient."""
    import gridapmhd_amd as g
    from gridapmhd_amd import adjoint as adj
    nc, Ha, Re, cw = (8, 8, 3), 20.0, 5.0, 0.05
    print("[tangent] developing duct 8x8x3, 2 levels", flush=True)
    sp, phys = adj.inlet_duct(nc, Ha, Re, cw)
    ctx = g._lib.context_with_levels(sp, phys, levels=2)
    try:
        o = ctx.default_opts(rtol=1e-12, atol=1e-15, ls_rtol=1e-11, max_krylov=2000, inner_iters=3, restart=200)
        x = adj.solve_state(ctx, sp, o)
        J, dHa, dcw = adj.tangent_gradient(ctx, sp, x, Ha, Re, o)
    finally:
        ctx.close()
    Jo, dHa_o, dcw_o = oracle_gradient(orc_mod, nc, Ha, Re, cw)
    print(f"  J={J:.12g} oracle {Jo:.12g}; dHa {dHa:.10g} oracle {dHa_o:.10g}; dcw {dcw:.10g} oracle {dcw_o:.10g}",
          flush=True)
    assert abs(J / Jo - 1) < 1e-9
    assert abs(dHa / dHa_o - 1) < 1e-8 and abs(dcw / dcw_o - 1) < 1e-8


def _fixture_gradient(name):
    import json
    fx = cases.load_fixture(name)
    meta = json.loads(str(fx["meta"]))
    return fx, meta


@pytest.mark.parametrize("name", list(cases.ADJOINT_CASES))
def test_gpu_adjoint_configs_match_oracle_and_fd(name):
    """The differentiation configuration's parameters (Ha = 100, Re = 10,
    cw = 0.01 on 8 x 8 x 8 cells) and the combined configuration's adjoint leg
    (jittered, non-affine cells, Ha = 500, Re = 50, cw = 0.001 / 0.01 / 0.1 on
    6 x 6 x 6 cells): the GPU adjoint gradient against the oracle's adjoint
    (fixture, 1e-8) and against central differences of GPU Newton solves
    (relative step 1e-6, bar 1e-6)."""
    from gridapmhd_amd import adjoint as adj
    kw = cases.ADJOINT_CASES[name]
    fx, meta = _fixture_gradient(name)
    nc, Ha, Re, cw, jit = kw["nc"], kw["Ha"], kw["Re"], kw["cw"], kw["jitter"]
    levels = 2 if nc[0] % 2 == 0 and nc[0] >= 8 else 1
    print(f"[adjoint] {name}: developing duct {nc} Ha={Ha:g} Re={Re:g} cw={cw:g} jitter={jit:g} levels={levels}",
          flush=True)
    okw = dict(rtol=1e-13, atol=1e-16, ls_rtol=1e-11, max_krylov=3000, inner_iters=3, restart=200, max_newton=20)
    J, pullback = adj.rrule(nc, Ha, Re, cw, opts_kw=okw, levels=levels, jitter=jit)
    dHa, dcw = pullback(1.0)
    Jo, dHa_o, dcw_o = meta["J"], meta["dJ_dHa"], meta["dJ_dcw"]
    print(f"  J={J:.12g} oracle {Jo:.12g}; dHa {dHa:.10g} oracle {dHa_o:.10g}; dcw {dcw:.10g} oracle {dcw_o:.10g}",
          flush=True)
    assert abs(J / Jo - 1) < 1e-9
    assert abs(dHa / dHa_o - 1) < 1e-8 and abs(dcw / dcw_o - 1) < 1e-8

    def J_of(Ha_, cw_):
        J_, _ = adj.rrule(nc, Ha_, Re, cw_, opts_kw=okw, levels=levels, jitter=jit)
        return J_

    h = 1e-6
    fd_Ha = (J_of(Ha * (1 + h), cw) - J_of(Ha * (1 - h), cw)) / (2 * h * Ha)
    fd_cw = (J_of(Ha, cw * (1 + h)) - J_of(Ha, cw * (1 - h))) / (2 * h * cw)
    print(f"  FD: dHa {fd_Ha:.10g} ({abs(dHa / fd_Ha - 1):.1e}) dcw {fd_cw:.10g} ({abs(dcw / fd_cw - 1):.1e})",
          flush=True)
    assert abs(dHa / fd_Ha - 1) < 1e-6 and abs(dcw / fd_cw - 1) < 1e-6
