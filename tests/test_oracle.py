"""Pins of the CPU oracle (runs without a GPU).

What each test fixes, and the mistake it would catch:
  * test_hunt_norms_match_reference: the full discrete problem (spaces,
    quadrature, weak form, BCs, dof numbering as a whole) against the
    reference's own published Hunt solutions (tests/golden/gadi_hunt.json,
    extracted from analysis/gadi/results/2023_04/*/summary.csv).  Any wrong
    sign, dropped term or basis error moves uh_l2/uh_h1/jh_l2 far above 1e-10.
  * test_manufactured_residual_zero: an in-space exact solution of
    src/Applications/transient.jl:262-270 gives R = 0 for all terms at once
    (convection, Lorentz force, Ohm's law, g and f sources, Dirichlet lifts).
  * test_jacobian_is_derivative: central differences of the residual.
  * test_block_identities: pu = up^T, phij = jphi^T / sigma,
    ju = -(sigma / gamma) uj^T, uu symmetric without convection.
  * test_closed_forms_unit_cube: tensor-product 1-D stiffness / mass
    matrices for the Q2 Laplacian, the RT mass and the div coupling on the
    unit cube, integrated exactly with numpy polynomials.
"""
import json
import os

import numpy as np
import numpy.polynomial.polynomial as P
import pytest

import cases

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "gadi_hunt.json")))


def _gold(cx, Ha):
    return [r for r in GOLD if r["cx"] == cx and r["Ha"] == Ha]


@pytest.mark.parametrize("cx,Ha", [(10, 50.0), (15, 500.0)])
def test_hunt_norms_match_reference(cx, Ha, orc_mod):
    sp, phys = cases.hunt_case(cx, Ha=Ha)
    # Hunt's :cfd scaling: B = (0, Ha, 0) -> Bbar = e_y, N = Ha^2 / Re (src/Applications/hunt.jl:117-136)
    ref = _gold(cx, Ha)
    assert ref, (cx, Ha)
    r = ref[0]
    assert (sp.nfree["u"], sp.nfree["p"], sp.nfree["j"], sp.nfree["phi"]) == (
        r["ndofs_u"], r["ndofs_p"], r["ndofs_j"], r["ndofs_phi"])
    orc = orc_mod.Oracle(cases.problem_dict(sp, phys))
    x = orc_mod.newton_direct(orc, np.zeros(sp.ndofs), pins=orc_mod.null_pins(sp))
    n = orc.cell_norms(x, nq=4).sum(0)
    uh_l2, uh_h1 = np.sqrt(n[0]), np.sqrt(n[0] + n[1])
    jh_l2 = Ha * np.sqrt(n[2])  # jh = sigma u0 B0 jbar with B0 = Ha (src/Applications/hunt.jl:217)
    for name, val in (("uh_l2", uh_l2), ("uh_h1", uh_h1), ("jh_l2", jh_l2)):
        rel = abs(val / r[name] - 1.0)
        assert rel < 1e-10, (name, val, r[name], rel)
    # div j_h = 0 pointwise (div RT1 lies in dQ1 on affine cells)
    assert np.sqrt(n[3]) < 1e-10 * np.sqrt(n[2]) * 100


def test_dof_counts_match_reference():
    import gridapmhd_amd as g
    seen = set()
    for r in GOLD:
        key = (r["cx"], r["Ha"])
        if key in seen or r["cx"] > 100:
            continue
        seen.add(key)
        mesh = g.meshers.hunt_generate_base_mesh((r["cx"], r["cx"]), Ha=r["Ha"], BL_adapted=False)
        sp = g.fespaces.setup_fe_spaces(mesh, "noslip", "insulating")
        assert sp.ndofs == r["ndofs"], (key, sp.ndofs, r["ndofs"])
        assert sp.nfree["p"] == r["ndofs_p"] and sp.nfree["phi"] == r["ndofs_phi"]


@pytest.mark.parametrize("B,conv", [((0.0, 1.0, 0.0), "newton"), ((0.3, 1.0, -0.2), "newton"),
                                    ((0.5, -0.7, 0.9), "picard"), ((0.0, 0.0, 1.0), "none")])
def test_manufactured_residual_zero(B, conv, orc_mod):
    sp, phys, xs, dv = cases.manufactured_case(nc=(4, 3, 3), B=B, alpha=0.7, beta=1.3, gamma=2.1, sigma=0.8,
                                               convection=conv)
    orc = orc_mod.Oracle(cases.problem_dict(sp, phys, dv))
    r = orc.residual(xs)
    assert np.max(np.abs(r)) < 1e-13
    # the residual is not trivially zero: perturbing the source breaks it
    phys2 = dict(phys, f0=np.asarray(phys["f0"]) + 1e-3)
    r2 = cases.problem_dict(sp, phys2, dv)
    assert np.max(np.abs(orc_mod.Oracle(r2).residual(xs))) > 1e-6


@pytest.mark.parametrize("zeta", [(0.0, 0.0), (2.0, 3.0)])
def test_jacobian_is_derivative(zeta, orc_mod):
    sp, phys = cases.duct_thin_wall_case((4, 3, 3), Ha=30.0, Re=5.0, cw=0.07, convection="newton",
                                         zeta_u=zeta[0], zeta_j=zeta[1],
                                         map=__import__("gridapmhd_amd").meshers.tanh_map(2.0))
    orc = orc_mod.Oracle(cases.problem_dict(sp, phys))
    rng = np.random.default_rng(3)
    x = 0.3 * rng.standard_normal(sp.ndofs)
    A = orc.jacobian(x)
    for _ in range(2):
        d = rng.standard_normal(sp.ndofs)
        h = 1e-6
        fd = (orc.residual(x + h * d) - orc.residual(x - h * d)) / (2 * h)
        assert np.linalg.norm(A @ d - fd) < 1e-7 * np.linalg.norm(fd)


def test_picard_drops_only_the_newton_term(orc_mod):
    sp, phys = cases.duct_thin_wall_case((3, 3, 3), Ha=10.0, Re=3.0, cw=0.1, convection="newton")
    rng = np.random.default_rng(5)
    x = rng.standard_normal(sp.ndofs)
    A_newton = orc_mod.Oracle(cases.problem_dict(sp, phys)).jacobian(x)
    A_picard = orc_mod.Oracle(cases.problem_dict(sp, dict(phys, convection="picard"))).jacobian(x)
    A_none = orc_mod.Oracle(cases.problem_dict(sp, dict(phys, convection="none"))).jacobian(x)
    # the Picard part is linear in u: doubling u doubles (A_picard - A_none)
    sl = sp.field_slices()["u"]
    x2 = x.copy()
    x2[sl] *= 2.0
    A_picard2 = orc_mod.Oracle(cases.problem_dict(sp, dict(phys, convection="picard"))).jacobian(x2)
    assert abs(((A_picard2 - A_none) - 2 * (A_picard - A_none))).max() < 1e-12 * abs(A_picard).max()
    assert abs(A_newton - A_picard).max() > 1e-6


def test_block_identities(orc_mod):
    sp, phys = cases.duct_thin_wall_case((3, 3, 3), Ha=20.0, Re=4.0, cw=0.05, convection="none")
    phys = dict(phys, sigma=0.6)
    A = orc_mod.Oracle(cases.problem_dict(sp, phys)).jacobian(np.zeros(sp.ndofs)).toarray()
    s = sp.field_slices()
    up, pu = A[s["u"], s["p"]], A[s["p"], s["u"]]
    uj, ju = A[s["u"], s["j"]], A[s["j"], s["u"]]
    jf, fj = A[s["j"], s["phi"]], A[s["phi"], s["j"]]
    uu = A[s["u"], s["u"]]
    assert np.abs(pu - up.T).max() < 1e-13 * np.abs(up).max()
    assert np.abs(fj - jf.T / phys["sigma"]).max() < 1e-13 * np.abs(fj).max()
    assert np.abs(ju + (phys["sigma"] / phys["gamma"]) * uj.T).max() < 1e-13 * np.abs(ju).max()
    assert np.abs(uu - uu.T).max() < 1e-13 * np.abs(uu).max()
    assert np.abs(A[s["p"], s["p"]]).max() == 0 and np.abs(A[s["phi"], s["phi"]]).max() == 0


# ---------------------------------------------------------------------------
# closed forms on the unit cube

def _one_cell_problem(phys):
    xyz = np.array([[(v >> 0) & 1, (v >> 1) & 1, (v >> 2) & 1] for v in range(8)], dtype=float)[None]
    return dict(cell_xyz=xyz, cell_dofs=np.arange(129, dtype=np.int32)[None], rt_sign=np.ones((1, 6), np.int8),
                dir_vals=np.zeros(1), nfree=129, **phys)


def _lagr(nodes, k):
    c = np.array([1.0])
    for m, xm in enumerate(nodes):
        if m != k:
            c = P.polymul(c, np.array([-xm, 1.0]) / (nodes[k] - xm))
    return c


def _int01(c):
    ci = P.polyint(c)
    return P.polyval(1.0, ci) - P.polyval(0.0, ci)


def test_closed_forms_unit_cube(orc_mod):
    beta, sigma = 1.7, 0.9
    phys = dict(alpha=1.0, beta=beta, gamma=0.0, sigma=sigma, B=np.zeros(3), convection="none")
    K = orc_mod.Oracle(_one_cell_problem(phys)).cell_jacobians(np.zeros(129))[0]
    q2n = [0.0, 0.5, 1.0]
    L = [_lagr(q2n, k) for k in range(3)]
    M1 = np.array([[_int01(P.polymul(L[a], L[b])) for b in range(3)] for a in range(3)])
    K1 = np.array([[_int01(P.polymul(P.polyder(L[a]), P.polyder(L[b]))) for b in range(3)] for a in range(3)])
    assert np.allclose(M1, np.array([[4, 2, -1], [2, 16, 2], [-1, 2, 4]]) / 30, atol=1e-15)
    assert np.allclose(K1, np.array([[7, -8, 1], [-8, 16, -8], [1, -8, 7]]) / 3, atol=1e-14)
    # index a = i + 3j + 9k -> kron(Z, kron(Y, X))
    lap = np.kron(M1, np.kron(M1, K1)) + np.kron(M1, np.kron(K1, M1)) + np.kron(K1, np.kron(M1, M1))
    for c in range(3):
        blk = K[c * 27:(c + 1) * 27, c * 27:(c + 1) * 27]
        assert np.abs(blk - beta * lap).max() < 1e-13
    assert np.abs(K[0:27, 27:54]).max() < 1e-14
    # RT mass: dofs of axis d are L_m(x_d) G(x_t1) G(x_t2); jj = M1 (x) MG (x) MG per axis
    g0, g1 = 0.5 - 0.5 / np.sqrt(3), 0.5 + 0.5 / np.sqrt(3)
    G = [_lagr([g0, g1], 0), _lagr([g0, g1], 1)]
    MG = np.array([[_int01(P.polymul(G[a], G[b])) for b in range(2)] for a in range(2)])
    l1 = [np.array([1.0, -1.0]), np.array([0.0, 1.0])]

    def rt_dofs(d):
        # local RT indices of axis d ordered by (m, g1, g2): faces at m = 0, 2, interior at m = 1
        face_lo, face_hi = {0: 4, 1: 2, 2: 0}[d], {0: 5, 1: 3, 2: 1}[d]
        out = {}
        for k in range(4):
            out[(0, k % 2, k // 2)] = 4 * face_lo + k
            out[(1, k % 2, k // 2)] = 24 + 4 * d + k
            out[(2, k % 2, k // 2)] = 4 * face_hi + k
        return out

    jj = K[85:121, 85:121]
    for d in range(3):
        ids = rt_dofs(d)
        for (m, a1, a2), r in ids.items():
            for (mm, b1, b2), s in ids.items():
                ref = M1[m, mm] * MG[a1, b1] * MG[a2, b2]
                assert abs(jj[r, s] - ref) < 1e-14, (d, r, s)
        for e in range(3):
            if e != d:
                for r in ids.values():
                    for s in rt_dofs(e).values():
                        assert abs(jj[r, s]) < 1e-15
    # j-phi coupling: -sigma int chi_v div psi_r, div psi_r = L_m'(x_d) G G
    jf = K[85:121, 121:129]
    for d in range(3):
        t1, t2 = {0: (1, 2), 1: (0, 2), 2: (0, 1)}[d]
        for (m, a1, a2), r in rt_dofs(d).items():
            for v in range(8):
                bits = [(v >> k) & 1 for k in range(3)]
                ref = -sigma * (_int01(P.polymul(l1[bits[d]], P.polyder(L[m]))) *
                                _int01(P.polymul(l1[bits[t1]], G[a1])) * _int01(P.polymul(l1[bits[t2]], G[a2])))
                assert abs(jf[r, v] - ref) < 1e-14, (d, r, v)


@pytest.mark.parametrize("cx,Ha", [(10, 50.0), (15, 500.0)])
def test_hunt_analytic_errors_match_reference(cx, Ha, orc_mod):
    """eu_h1 and ej_l2 of the oracle's Hunt solution against the analytic
    series (src/Applications/hunt.jl:237-254, 372-457) equal the reference's
    own published values: this pins the analytic functions, the evaluation of
    u_h, grad u_h and the Piola-mapped j_h, and the 4-point measure."""
    from gridapmhd_amd import postprocess
    sp, phys = cases.hunt_case(cx, Ha=Ha)
    r = _gold(cx, Ha)[0]
    orc = orc_mod.Oracle(cases.problem_dict(sp, phys))
    x = orc_mod.newton_direct(orc, np.zeros(sp.ndofs), pins=orc_mod.null_pins(sp))
    e = postprocess.hunt_errors(sp, x, 1.0, 1.0, 1.0, -1.0, Ha, r["nsums"], u_scale=1.0, j_scale=Ha)
    for name in ("eu_h1", "ej_l2"):
        rel = abs(e[name] / r[name] - 1.0)
        assert rel < 1e-9, (name, e[name], r[name], rel)
    # a truncated series (fewer modes) is told apart at this precision
    e5 = postprocess.hunt_errors(sp, x, 1.0, 1.0, 1.0, -1.0, Ha, 5, u_scale=1.0, j_scale=Ha)
    assert abs(e5["eu_h1"] / r["eu_h1"] - 1.0) > 1e-6


def test_thin_wall_duct_converges_to_general_hunt(orc_mod):
    """The thin-wall term against the closed-form fully developed flow with
    thin conducting Hartmann walls and insulating side walls
    (analytical_general_hunt_u, src/Applications/FullyDevelopedFlow.jl:284-316):
    the FullyDeveloped duct at Ha = 10, cw = 0.1, tau = 1e4 on 4 x 4 x 3 and
    8 x 8 x 3 stretched cells.  The distance shrinks with h (measured 6.5e-2 ->
    1.3e-2, then 4.8e-3 on 12 x 12 x 3); a thin-wall term with the wrong cw
    lands elsewhere (the insulating duct, cw = 0, is 30 % away)."""
    import importlib
    from gridapmhd_amd import applications, postprocess
    from gridapmhd_amd.parameters import add_default_params, fluid_physics
    main_mod = importlib.import_module("gridapmhd_amd.main")

    def err(nc, cw, cw_model):
        params, _ = applications.fully_developed_problem(nc, 10.0, cw_Ha=cw, τ_Ha=1e4, ζ=0.0)
        params["solver"] = {"solver": "hip"}
        p = add_default_params(params)
        sp = main_mod.setup_fe_spaces(p)
        orc = orc_mod.Oracle(cases.problem_dict(sp, fluid_physics(p["fluid"])))
        x = orc_mod.newton_direct(orc, np.zeros(sp.ndofs), maxiter=5, rtol=1e-12, atol=1e-15,
                                  pins=orc_mod.null_pins(sp, pin_phi=(cw == 0.0)))
        return postprocess.general_hunt_error(sp, x, 1.0, cw_model, -1.0, 10.0, 50)
    e4, e8 = err((4, 4), 0.1, 0.1), err((8, 8), 0.1, 0.1)
    assert e8 < 0.5 * e4 and e8 < 2e-2, (e4, e8)
    assert err((4, 4), 0.1, 0.0) > 0.2


def test_pinned_jacobian_condition_at_ha1000(orc_mod):
    """Condition of the oracle's pinned Jacobian at the solution of the
    Ha = 1000, tanh(beta = 3) thin-wall duct on 8 x 8 x 3 cells (block
    numbering): cond_1 = |A|_1 |A^-1|_1 with |A^-1|_1 from scipy's onenormest
    on the LU solve.  With the first pressure dof pinned it is about 1e13
    (so a 1e-16 residual bounds the error only to about 1e-3, and the
    solution is reproducible to about 3e-9 of |x| in practice).  Pinning the
    first dof of the wrong field leaves the constant-pressure mode: cond_1
    then exceeds 1e17, which is what a numbering-blind pin did."""
    import scipy.sparse.linalg as spla
    import gridapmhd_amd as g
    sp, phys = cases.duct_thin_wall_case((8, 8, 3), 1000.0, 10.0, 0.05, map=g.meshers.tanh_map(3.0),
                                         style="badia2024")
    orc = orc_mod.Oracle(cases.problem_dict(sp, dict(phys)))
    pins = orc_mod.null_pins(sp)
    assert pins == [sp.field_slices()["p"].start]
    x = orc_mod.newton_direct(orc, np.zeros(sp.ndofs), maxiter=20, rtol=1e-13, atol=1e-15, pins=pins)
    A = orc.jacobian(x).tocsc()

    def cond1(pin):
        keep = np.ones(sp.ndofs, bool)
        keep[pin] = False
        Ak = A[keep][:, keep].tocsc()
        lu = spla.splu(Ak)
        n = Ak.shape[0]
        inv = spla.LinearOperator((n, n), matvec=lu.solve, rmatvec=lambda v: lu.solve(v, trans="T"))
        return spla.norm(Ak, 1) * spla.onenormest(inv)

    c = cond1(pins[0])
    assert 1e11 < c < 1e14, c
    assert cond1(sp.field_slices()["j"].start) > 1e17
