"""GPU parity of the skeleton stabilisation (src/weakforms.jl:685-688,
added to the residual and Jacobian at :227-229, :248-250):

* pattern bit-exact, residual and Jacobian of libmhd (mhd_create_stabilized)
  against the oracle's numpy restatement, on a periodic Hunt duct, a
  tanh-stretched thin-wall duct with augmented-Lagrangian terms, a jittered
  (non-affine) duct and an unstructured box whose cells are shuffled and
  rotated (face Gauss points permuted between the two cells of a face);
  bars as for the cell terms: 1e-12;
* the allocation-free residual carries the term too;
* hunt(mu = ...) -> main(params) -> Newton-FGMRES with the term on every
  level of the V-cycle, FP64 and mixed precision: the oracle's residual
  (with the term) at the GPU solution is below 1e-12 |R0|, and every field
  agrees with the oracle's Newton + direct solve of the same (augmented)
  problem to 1e-10 (FP64) / 1e-8 (mixed), the pressure modulo its constant
  (phi, the weakly determined field, to ten times that);
* mhd_create_stabilized rejects malformed face lists on the host."""
import sys

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu

MU = 0.4


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _row_scaled_err(rowval, a, b, n):
    rmax = np.zeros(n)
    np.maximum.at(rmax, rowval, np.abs(b))
    rmax[rmax == 0] = 1.0
    return np.max(np.abs(a - b) / rmax[rowval])


def _unstructured_box():
    import gridapmhd_amd as g
    from test_mesh_io import _box, _scramble
    hm = _scramble(g.mesh_io.hexmesh_from_cartesian(_box()), 3)
    sp = g.fespaces.setup_fe_spaces(hm, "noslip", "ends", [dict(domain="insulating", cw=0.05, tau=100.0, jw=0.0)],
                                    style="badia2024")
    return sp, cases.cfd_physics(10.0, 5.0)


CASES = {
    "hunt_periodic": lambda: cases.hunt_case(4, Ha=50.0),
    "thin_wall_tanh_al": lambda: cases.duct_thin_wall_case((4, 4, 3), Ha=30.0, Re=5.0, cw=0.05, zeta_u=1.0,
                                                           zeta_j=1.0,
                                                           map=__import__("gridapmhd_amd").meshers.tanh_map(3.0)),
    "jitter_newton": lambda: cases.duct_thin_wall_case((4, 4, 3), Ha=100.0, Re=10.0, cw=0.01, jitter=0.2),
    "unstructured_box": _unstructured_box,
}


@pytest.mark.parametrize("name", list(CASES))
def test_skeleton_assembly_matches_oracle(name, orc_mod):
    import gridapmhd_amd as g
    _say(f"[skeleton] assembly case={name}")
    sp, phys = CASES[name]()
    phys = dict(phys, stab_mu=MU)
    orc = orc_mod.Oracle(cases.problem_dict(sp, phys))
    rng = np.random.default_rng(99)
    x = 0.1 * rng.standard_normal(sp.ndofs)
    dv = sp.dir_vals.copy()
    ctx = g._lib.context_from_spaces(sp, phys)
    base = g._lib.context_from_spaces(sp, dict(phys, stab_mu=0.0))
    try:
        colptr, rowval = ctx.pattern()
        ocp, orv = orc.pattern()
        assert np.array_equal(colptr - 1, ocp) and np.array_equal(rowval - 1, orv)
        assert ctx.nnz > base.nnz
        r_gpu, r_orc = ctx.residual(x), orc.residual(x)
        err_r = np.max(np.abs(r_gpu - r_orc)) / np.linalg.norm(r_orc)
        # the term's own size: it must not be lost in the tolerance
        r_skel = np.linalg.norm(r_orc - orc_mod.Oracle(cases.problem_dict(sp, dict(phys, stab_mu=0.0))).residual(x))
        _say(f"  residual rel err {err_r:.3e} (|R_skeleton|/|R| = {r_skel / np.linalg.norm(r_orc):.2e})")
        assert err_r < 1e-12 and r_skel > 1e-3 * np.linalg.norm(r_orc)
        nz_gpu, nz_orc = ctx.jacobian(x), orc.jacobian_on_pattern(x)
        err_j = _row_scaled_err(orv, nz_gpu, nz_orc, sp.ndofs)
        _say(f"  jacobian row-scaled err {err_j:.3e} nnz={nz_gpu.size} (without the term {base.nnz})")
        assert err_j < 1e-12
        # allocation-free residual: the same bits as mhd_residual_d
        from devmem import DeviceArray
        dx = DeviceArray.from_numpy(x)
        r1 = DeviceArray.from_numpy(np.zeros(sp.ndofs))
        r2 = DeviceArray.from_numpy(np.full(sp.ndofs, np.nan))
        ctx.residual_d(dx.ptr, r1.ptr)
        ctx.residual_d_noalloc(dx.ptr, r2.ptr)
        assert np.array_equal(r1.numpy(), r2.numpy()) and np.array_equal(r1.numpy(), r_gpu)
        assert np.array_equal(dv, sp.dir_vals)
    finally:
        ctx.close()
        base.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_hunt_with_skeleton_through_main_matches_oracle(precision, orc_mod):
    import gridapmhd_amd as g
    _say(f"[skeleton] hunt 6x6x3 mu={MU} two levels precision={precision}")
    solver = {"solver": "hip", "rtol": 1e-12, "atol": 1e-15, "niter_ls": 2000, "precision": precision,
              "levels": 2, "restart": 300}
    # the augmented-Lagrangian weights Badia2024's Schur approximations rely on
    info, xh = g.hunt(nc=(6, 6), B=(0.0, 20.0, 0.0), ζᵤ=1.0, ζⱼ=1.0, convection="newton", BL_adapted=False,
                      solver=solver, errors=False, μ=MU)
    sp = xh.spaces
    assert info["converged"], {k: v for k, v in info.items() if "iter" in k or "resid" in k}
    phys = dict(cases.cfd_physics(20.0, 1.0, zeta_u=1.0, zeta_j=1.0), stab_mu=MU)
    orc = orc_mod.Oracle(cases.problem_dict(sp, phys))
    x = xh.free_values
    # the oracle's residual (with the term) vanishes at the GPU solution
    r0 = np.linalg.norm(orc.residual(np.zeros(sp.ndofs)))
    rg = np.linalg.norm(orc.residual(x))
    # and the term is not negligible at the solution
    rs = np.linalg.norm(orc.residual(x) - orc_mod.Oracle(cases.problem_dict(sp, dict(phys, stab_mu=0.0))).residual(x))
    # the oracle's Newton + direct solve; the Jacobian's one null mode is the
    # constant pressure (its smallest singular values: 2e-19, then 1e-5): phi
    # is determined, since j.n is free on part of the boundary
    pins = orc_mod.null_pins(sp, pin_p=True, pin_phi=False)
    xref = orc_mod.newton_direct(orc, np.zeros(sp.ndofs), rtol=1e-14, atol=1e-15, pins=pins)
    errs = {}
    for f in ("u", "j", "phi"):
        s_ = sp.field_slices()[f]
        errs[f] = np.linalg.norm(x[s_] - xref[s_]) / np.linalg.norm(xref[s_])
    ps = sp.field_slices()["p"]
    errs["p"] = np.linalg.norm((x[ps] - x[ps].mean()) - (xref[ps] - xref[ps].mean())) / np.linalg.norm(xref)
    _say(f"  newton={info['nls_iters']} krylov={info.get('ls_iters')} |R_oracle(x)|/|R0| = {rg / r0:.2e} "
         f"|R_skeleton(x)|/|R0| = {rs / r0:.2e} vs oracle direct: " + " ".join(f"{k}={v:.2e}" for k, v in errs.items()))
    assert rg < 1e-12 * r0 and rs > 1e-6 * r0
    bar = 1e-10 if precision == "fp64" else 1e-8
    # phi is the weakly determined field (the Jacobian's smallest nonzero
    # singular values, 1e-5, belong to it): at a 5e-13 relative residual its
    # error is measured at 2e-10, so its bar is ten times the others
    assert errs["u"] < bar and errs["j"] < bar and errs["p"] < bar and errs["phi"] < 10 * bar, errs


def test_skeleton_face_list_is_checked():
    import gridapmhd_amd as g
    _say("[skeleton] malformed face lists")
    sp, phys = cases.hunt_case(4, Ha=50.0)
    sk = dict(g.fespaces.skeleton_faces(sp))
    nf = sk["cell"].shape[0]
    bad = []
    p = sk["perm"].copy()
    p[3, 1] = p[3, 0]
    bad.append((dict(sk, perm=p), "not a permutation"))
    c = sk["cell"].copy()
    c[5, 1] = c[5, 0]
    bad.append((dict(sk, cell=c), "glued to itself"))
    lf = sk["lface"].copy()
    lf[7, 1] = (lf[7, 1] + 2) % 6
    bad.append((dict(sk, lface=lf), "do not share"))
    c = sk["cell"].copy()
    c[2, 0] = sp.ncells
    bad.append((dict(sk, cell=c), "out of range"))
    nf4 = [sp.nfree["u"], sp.nfree["p"], sp.nfree["j"], sp.nfree["phi"]]
    for skel, msg in bad:
        with pytest.raises(g._lib.MhdError, match=msg):
            g._lib.Context(sp.mesh.cell_xyz(), sp.cell_dofs, sp.rt_sign, nf4, sp.dir_vals, phys,
                           cell_verts=sp.cell_verts, skeleton=dict(skel, mu=np.full(nf, MU)))
