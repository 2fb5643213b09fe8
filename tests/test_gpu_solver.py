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
    pins = orc_mod.null_pins(sp)
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
@pytest.mark.parametrize("nc,levels,coarse_iters,step", [((8, 8, 3), 2, 10, "full"), ((16, 16, 3), 3, 10, "full"),
                                                          ((16, 16, 3), 3, 0, "full"), ((16, 16, 3), 3, 0, "minres")])
def test_multilevel_solve_matches_direct(nc, levels, coarse_iters, step, precision, orc_mod):
    """The [u,j] V-cycle (semi-coarsened levels attached through mhd_attach_coarse);
    coarse_iters = 0 solves the coarsest level exactly (dense inverse); the
    coarse correction enters with a full or a minimal-residual step."""
    import gridapmhd_amd as g
    _say(f"[solver] multilevel nc={nc} levels={levels} coarse_iters={coarse_iters} step={step} precision={precision}")
    sp, phys = cases.duct_thin_wall_case(nc, Ha=30.0, Re=10.0, cw=0.02, convection="newton", style="badia2024")
    orc = orc_mod.Oracle(cases.problem_dict(sp, phys))
    x_ref = orc_mod.newton_direct(orc, np.zeros(sp.ndofs), maxiter=20, rtol=1e-13, atol=1e-15,
                                  pins=orc_mod.null_pins(sp))
    ctx = g._lib.context_with_levels(sp, dict(phys, zeta_u=1.0, zeta_j=1.0), levels=levels,
                                     coarse_iters=coarse_iters, coarse_step=step)
    try:
        o = _opts(g, precision)
        o.inner_iters = 3
        x, info, rc = ctx.newton_solve(np.zeros(sp.ndofs), o)
        err = cases.field_errors(sp, x, x_ref)
        _say(f"  newton={info.newton_iters} krylov={info.krylov_iters} rc={rc} err={err['all']:.2e}")
        assert rc == 0 and info.converged
        assert err["all"] < (1e-10 if precision == 0 else 1e-8), err
        if coarse_iters == 0:
            # one dense coarsest inverse per Newton step, by block Gauss-Jordan
            # (its residual check passed: a pivoted-LU redo would count twice)
            coarsest = sp._multigrid_cache[(0, 1)]["chain"][levels - 1]
            n1c = coarsest.nfree["u"] + coarsest.nfree["j"]
            assert ctx.op_counts()["setup_dense_dim3"] == info.newton_iters * float(n1c) ** 3
    finally:
        ctx.close()


# Features of the larger configurations at oracle-sized meshes: tanh node
# stretching (beta = 3) with Ha = 1000 and cw = 0.05, and jittered
# (non-affine) cells with Ha = 500, Re = 50 at the three wall conductances.
FEATURE_CASES = {
    "tanh_ha1000": dict(nc=(8, 8, 3), Ha=1000.0, Re=10.0, cw=0.05, tanh=3.0),
    "jitter_ha500_cw0.001": dict(nc=(8, 8, 3), Ha=500.0, Re=50.0, cw=0.001, jitter=0.2),
    "jitter_ha500_cw0.01": dict(nc=(8, 8, 3), Ha=500.0, Re=50.0, cw=0.01, jitter=0.2),
    "jitter_ha500_cw0.1": dict(nc=(8, 8, 3), Ha=500.0, Re=50.0, cw=0.1, jitter=0.2),
}


_FEATURE_REF = {}
_FEATURE_REPRO = {}


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("name", list(FEATURE_CASES))
def test_feature_cases_match_direct(name, precision, orc_mod):
    """Two-level V-cycle with the exact coarsest solve and the minimal-residual
    coarse step (on 4 x 4 coarse cells the Hartmann layers of Ha = 500-1000
    are far from resolved; on jittered meshes the coarse level is not even
    nested, so the prolongation is an approximation: both only affect the
    iteration count)."""
    import gridapmhd_amd as g
    kw = dict(FEATURE_CASES[name])
    _say(f"[solver] feature case={name} precision={precision}")
    beta = kw.pop("tanh", None)
    if beta:
        kw["map"] = g.meshers.tanh_map(beta)
    sp, phys = cases.duct_thin_wall_case(convection="newton", style="badia2024", **kw)
    jitter = bool(kw.get("jitter"))
    # On non-affine cells the divergence of a mapped RT field is not in the
    # mapped dQ1 space, so the zeta_j div-div term does not vanish at the
    # discrete solution: the oracle then solves the same augmented problem.
    zeta = 1.0 if jitter else 0.0
    orc = orc_mod.Oracle(cases.problem_dict(sp, dict(phys, zeta_u=zeta, zeta_j=zeta)))
    # On non-affine cells Gauss quadrature of div v is inexact, so a constant
    # pressure is no longer an exact null mode: the system is nonsingular and
    # pinning a pressure dof would drop an equation that does not hold.
    pins = () if jitter else orc_mod.null_pins(sp)
    if name not in _FEATURE_REF:
        _FEATURE_REF[name] = orc_mod.newton_direct(orc, np.zeros(sp.ndofs), maxiter=20, rtol=1e-13, atol=1e-15,
                                                   pins=pins)
    x_ref = _FEATURE_REF[name]
    if kw["Ha"] >= 1000.0 and name not in _FEATURE_REPRO:
        # the oracle's own Newton restarted from its solution perturbed by 1e-9
        xp = x_ref + 1e-9 * np.linalg.norm(x_ref) / np.sqrt(x_ref.size) * \
            np.random.default_rng(1).standard_normal(x_ref.size)
        x2 = orc_mod.newton_direct(orc, xp, maxiter=20, rtol=1e-13, atol=1e-15, pins=pins)
        _FEATURE_REPRO[name] = cases.field_errors(sp, x2, x_ref)
        _say(f"  oracle restarted from a 1e-9 perturbation moves by {_FEATURE_REPRO[name]}")
    ctx = g._lib.context_with_levels(sp, dict(phys, zeta_u=1.0, zeta_j=1.0), levels=2, coarse_iters=0,
                                     coarse_step="minres")
    try:
        # Inexact Newton down to the FP64 residual floor: these cases amplify
        # a relative residual of 1e-12 into ~5e-9 solution error, so the
        # Newton target is 1e-14 with cheap 1e-6 linear solves per step.
        o = _opts(g, precision, rtol=1e-14, ls_rtol=1e-6)
        o.atol = 1e-18
        o.max_newton = 10
        o.inner_iters = 10
        o.restart = 300
        o.max_krylov = 600
        x, info, rc = ctx.newton_solve(np.zeros(sp.ndofs), o)
        err = cases.field_errors(sp, x, x_ref)
        relres = info.residual / info.residual0
        _say(f"  patches inverted with partial pivoting: {ctx.smoother_info(pivoted=True)[3]}")
        _say(f"  newton={info.newton_iters} krylov={info.krylov_iters} rc={rc} relres={relres:.1e} "
             + " ".join(f"{k}={v:.2e}" for k, v in err.items()))
        assert rc >= 0 and relres < 1e-12
        bar = 1e-10 if precision == 0 else 1e-8
        if kw["Ha"] >= 1000.0:
            # Ha = 1000 on 8 x 8 x 3 stretched cells: cond_1 of the pinned
            # Jacobian is about 1e13 (test_oracle.py), so the FP64 solution
            # itself is determined only to what the oracle's own Newton,
            # restarted from a 1e-9 perturbation, reproduces: u, j, phi to
            # 1e-10 .. 2e-10, the pressure to 3e-3 of |p| (3e-9 of |x|).
            # u, j, phi: the standard bar, or ten times that reproducibility
            # where it is larger.  p: 1e-6 of |x| (measured: 1e-8 FP64,
            # 7e-8 mixed, both at a 5e-15 relative residual).
            rep = _FEATURE_REPRO[name]
            for f in ("u", "j", "phi"):
                assert err[f] < max(bar, 10.0 * rep[f]), (f, err, rep)
            assert err["all"] < 1e-6, (err, rep)
        else:
            assert err["all"] < bar, err
    finally:
        ctx.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_solves_repeat_bitwise(precision):
    """Two Newton solves of the same problem in one process give the same bits:
    the residual, the Jacobian and the smoother sum their contributions in
    fixed orders (no atomics on the default paths)."""
    import gridapmhd_amd as g
    _say(f"[solver] repeatability precision={precision}")
    sp, phys = cases.duct_thin_wall_case((16, 16, 3), Ha=30.0, Re=10.0, cw=0.02, convection="newton",
                                         style="badia2024")
    ctx = g._lib.context_with_levels(sp, dict(phys, zeta_u=1.0, zeta_j=1.0), levels=3, coarse_iters=0,
                                     coarse_step="full")
    try:
        x0 = np.random.default_rng(11).standard_normal(sp.ndofs) * 0.1
        r1, r2 = ctx.residual(x0), ctx.residual(x0)
        assert np.array_equal(r1, r2)
        o = _opts(g, precision)
        o.inner_iters = 4
        xa, ia, rca = ctx.newton_solve(np.zeros(sp.ndofs), o)
        xb, ib, rcb = ctx.newton_solve(np.zeros(sp.ndofs), o)
        _say(f"  krylov {ia.krylov_iters} / {ib.krylov_iters}  |R| {ia.residual!r} / {ib.residual!r}")
        assert rca == 0 and rcb == 0
        assert ia.krylov_iters == ib.krylov_iters
        assert np.array_equal(xa, xb)
    finally:
        ctx.close()


def _vertex_patches(sp):
    """The vertex-star patches of the [u,j] block (badia2024 numbering): for
    every vertex, the free u / j dofs whose support lies in the cells around it
    (a Q2 node of a cell within one lattice step of the vertex; an RT1 face dof
    of a face through the vertex, or an interior dof)."""
    n1 = sp.nfree["u"] + sp.nfree["j"]
    lists = {}
    cd = sp.cell_dofs
    for cc in range(sp.ncells):
        for lv in range(8):
            vi, vj, vk = lv & 1, (lv >> 1) & 1, lv >> 2
            L = lists.setdefault(int(sp.cell_verts[cc, lv]), set())
            for a in range(27):
                i, j, k = a % 3, (a // 3) % 3, a // 9
                if abs(i - 2 * vi) > 1 or abs(j - 2 * vj) > 1 or abs(k - 2 * vk) > 1:
                    continue
                for comp in range(3):
                    g = int(cd[cc, comp * 27 + a])
                    if 0 <= g < n1:
                        L.add(g)
            for r in range(36):
                if r < 24:
                    f = r >> 2
                    d, side = 2 - f // 2, f & 1
                    if (vi, vj, vk)[d] != side:
                        continue
                g = int(cd[cc, 85 + r])
                if 0 <= g < n1:
                    L.add(g)
    return [np.array(sorted(L)) for _, L in sorted(lists.items()) if L]


@pytest.mark.parametrize("precision", [0, 1, 3])
def test_smoother_matches_patch_inverses(precision, orc_mod):
    """The vertex-star Schwarz smoother (mhd_smoother_apply_d after
    mhd_precond_setup: the patch matrices inverted on the GPU by in-register
    Gauss-Jordan) against y = sum_P R_P^T A_P^-1 R_P r computed on the CPU from
    the oracle's Jacobian with numpy inverses.  Newton convection at a random
    state (non-symmetric patches), thin walls, a jittered mesh."""
    import gridapmhd_amd as g
    from devmem import DeviceArray
    _say(f"[solver] smoother vs numpy patch inverses, precision={precision}")
    sp, phys = cases.duct_thin_wall_case((4, 4, 3), Ha=30.0, Re=10.0, cw=0.05, convection="newton",
                                         style="badia2024", zeta_u=1.0, zeta_j=1.0, jitter=0.15)
    rng = np.random.default_rng(9)
    x = 0.3 * rng.standard_normal(sp.ndofs)
    n1 = sp.nfree["u"] + sp.nfree["j"]
    A = orc_mod.Oracle(cases.problem_dict(sp, phys)).jacobian(x).tocsr()[:n1, :n1]
    patches = _vertex_patches(sp)
    r = rng.standard_normal(n1)
    y_ref = np.zeros(n1)
    for ids in patches:
        Ap = A[ids][:, ids].toarray()
        y_ref[ids] += np.linalg.solve(Ap, r[ids])
    ctx = g._lib.context_from_spaces(sp, phys)
    try:
        ctx.jacobian(x)                       # into the context: the Jacobian the set-up inverts
        ctx.precond_setup(ctx.default_opts(precision=precision))
        npatch, pdim, _ = ctx.smoother_info()
        assert npatch == len(patches) and pdim == max(len(p) for p in patches)
        dt = np.float64 if precision == 0 else np.float32
        rr = DeviceArray.from_numpy(r, dt)
        yy = DeviceArray.from_numpy(np.zeros(n1), dt)
        ctx.smoother_apply_d(rr.ptr, yy.ptr)
        y = yy.numpy().astype(np.float64)
    finally:
        ctx.close()
    err = np.linalg.norm(y - y_ref) / np.linalg.norm(y_ref)
    _say(f"  {len(patches)} patches (max {pdim} dofs): rel err {err:.2e}")
    # FP16 blocks (precision 3) carry 11 significant bits: 4.9e-4 per entry
    assert err < {0: 1e-12, 1: 1e-5, 3: 5e-3}[precision]
