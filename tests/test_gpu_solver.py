"""GPU Newton + FGMRES + Badia2024 block preconditioner against the oracle's
Newton + direct solve (the plain method) on the same mesh and dof tables.

The GPU path runs with augmented-Lagrangian weights zeta_u = zeta_j = 1: those
terms vanish at the discrete solution (div j_h = 0 and Pi_p div u_h = 0 hold
exactly), so both paths solve the same discrete problem.  Bar: the solution
within 1e-10 relative of the oracle (FP64), within 1e-8 (mixed precision);
the pressure is compared modulo its constant.
"""
import sys

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _opts(g, precision, rtol=1e-12, ls_rtol=1e-12):
    o = g._lib.MhdSolverOpts()
    g._lib.load().mhd_default_solver_opts(__import__("ctypes").byref(o))
    o.precision = precision
    o.rtol = rtol
    o.atol = 1e-14
    o.ls_rtol = ls_rtol
    o.max_krylov = 600
    return o


SOLVE_CASES = {
    "c1_stokes_thin_wall": dict(nc=(8, 8, 3), Ha=10.0, Re=1.0, cw=0.01, convection="none"),
    "newton_thin_wall": dict(nc=(6, 6, 3), Ha=50.0, Re=20.0, cw=0.05, convection="newton"),
}


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("name", list(SOLVE_CASES))
def test_newton_solve_matches_direct(name, precision, orc_mod):
    import gridapmhd_amd as g
    kw = SOLVE_CASES[name]
    _say(f"[solver] case={name} precision={precision}")
    sp, phys = cases.duct_thin_wall_case(style="badia2024", **kw)
    # oracle: plain method (zeta = 0), Newton + direct solve
    orc = orc_mod.Oracle(cases.problem_dict(sp, phys))
    pins = orc_mod.null_pins(sp.nfree)
    x_ref = orc_mod.newton_direct(orc, np.zeros(sp.ndofs), maxiter=20, rtol=1e-13, atol=1e-15, pins=pins)
    gphys = dict(phys, zeta_u=1.0, zeta_j=1.0)
    ctx = g._lib.context_from_spaces(sp, gphys)
    try:
        x, info, rc = ctx.newton_solve(np.zeros(sp.ndofs), _opts(g, precision))
        _say(f"  newton={info.newton_iters} krylov={info.krylov_iters} |R|={info.residual:.3e} rc={rc}")
        err = cases.field_errors(sp, x, x_ref)
        _say("  errors " + " ".join(f"{k}={v:.2e}" for k, v in err.items()))
        assert rc == 0 and info.converged
        bar = 1e-10 if precision == 0 else 1e-8
        assert err["all"] < bar, err
    finally:
        ctx.close()


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("nc,levels", [((8, 8, 3), 2), ((16, 16, 3), 3)])
def test_multilevel_solve_matches_direct(nc, levels, precision, orc_mod):
    """The [u,j] V-cycle (semi-coarsened levels attached through mhd_attach_coarse)."""
    import gridapmhd_amd as g
    _say(f"[solver] multilevel nc={nc} levels={levels} precision={precision}")
    sp, phys = cases.duct_thin_wall_case(nc, Ha=30.0, Re=10.0, cw=0.02, convection="newton", style="badia2024")
    orc = orc_mod.Oracle(cases.problem_dict(sp, phys))
    x_ref = orc_mod.newton_direct(orc, np.zeros(sp.ndofs), maxiter=20, rtol=1e-13, atol=1e-15,
                                  pins=orc_mod.null_pins(sp.nfree))
    ctx = g._lib.context_with_levels(sp, dict(phys, zeta_u=1.0, zeta_j=1.0), levels=levels, coarse_iters=10)
    try:
        o = _opts(g, precision)
        o.inner_iters = 3
        x, info, rc = ctx.newton_solve(np.zeros(sp.ndofs), o)
        err = cases.field_errors(sp, x, x_ref)
        _say(f"  newton={info.newton_iters} krylov={info.krylov_iters} rc={rc} err={err['all']:.2e}")
        assert rc == 0 and info.converged
        assert err["all"] < (1e-10 if precision == 0 else 1e-8), err
    finally:
        ctx.close()
