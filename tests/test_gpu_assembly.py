"""GPU parity of residual / Jacobian assembly against the CPU oracle.

Every test drives libmhd through its C ABI (ctypes) and compares with the
oracle on the same mesh, dof tables, parameters and state.  Tolerances:
residual entries 1e-12 relative to the residual norm, Jacobian entries 1e-12
relative to the largest entry of their row (the bar for matrix values),
pattern bit-exact.
"""
import sys

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _row_scaled_err(colptr, rowval, a, b, n):
    # rows of a CSC matrix: scale each entry by the max |entry| of its row
    rows = rowval
    rmax = np.zeros(n)
    np.maximum.at(rmax, rows, np.abs(b))
    rmax[rmax == 0] = 1.0
    return np.max(np.abs(a - b) / rmax[rows])


CASES = {
    "c1_thin_wall_stokes": lambda: cases.duct_thin_wall_case((8, 8, 3), Ha=10.0, Re=1.0, cw=0.01, convection="none"),
    "hunt_newton": lambda: cases.hunt_case(6, Ha=50.0),
    "thin_wall_newton_al": lambda: cases.duct_thin_wall_case((5, 4, 3), Ha=100.0, Re=10.0, cw=0.05,
                                                             convection="newton", zeta_u=3.0, zeta_j=2.0),
    "thin_wall_picard_tanh": lambda: cases.duct_thin_wall_case((6, 6, 3), Ha=30.0, Re=5.0, cw=0.1,
                                                               convection="picard",
                                                               map=__import__("gridapmhd_amd").meshers.tanh_map(3.0)),
    # non-affine cells: interior nodes jittered by +-0.2 h (the combined configuration's mesh)
    "jitter_newton_al": lambda: cases.duct_thin_wall_case((5, 5, 3), Ha=500.0, Re=50.0, cw=0.001,
                                                          convection="newton", zeta_u=1.0, zeta_j=1.0, jitter=0.2),
}


@pytest.mark.parametrize("name", list(CASES))
def test_assembly_matches_oracle(name, orc_mod):
    import gridapmhd_amd as g
    _say(f"[assembly] case={name}")
    sp, phys = CASES[name]()
    pb = cases.problem_dict(sp, phys)
    orc = orc_mod.Oracle(pb)
    rng = np.random.default_rng(1234)
    x = 0.1 * rng.standard_normal(sp.ndofs)
    ctx = g._lib.context_from_spaces(sp, phys)
    try:
        colptr, rowval = ctx.pattern()
        ocp, orv = orc.pattern()
        assert np.array_equal(colptr - 1, ocp)
        assert np.array_equal(rowval - 1, orv)
        r_gpu = ctx.residual(x)
        r_orc = orc.residual(x)
        err_r = np.max(np.abs(r_gpu - r_orc)) / np.linalg.norm(r_orc)
        _say(f"  residual rel err {err_r:.3e}")
        assert err_r < 1e-12
        nz_gpu = ctx.jacobian(x)
        nz_orc = orc.jacobian_on_pattern(x)
        err_j = _row_scaled_err(ocp, orv, nz_gpu, nz_orc, sp.ndofs)
        _say(f"  jacobian row-scaled err {err_j:.3e} nnz={nz_gpu.size}")
        assert err_j < 1e-12
    finally:
        ctx.close()


def test_manufactured_residual_vanishes():
    """R(I_h x*) = 0 for the in-space manufactured solution (no oracle needed)."""
    import gridapmhd_amd as g
    _say("[assembly] manufactured")
    sp, phys, xs, dv = cases.manufactured_case(nc=(4, 3, 3), B=(0.3, 1.0, -0.2))
    ctx = g._lib.context_from_spaces(sp, phys, dir_vals=dv)
    try:
        r = ctx.residual(xs)
        assert np.max(np.abs(r)) < 1e-12
    finally:
        ctx.close()


@pytest.mark.parametrize("name", ["hunt_newton", "thin_wall_newton_al"])
def test_scatter_modes_agree(name):
    """Row assembly (default) and colour classes are bitwise reproducible; the
    three scatter modes agree up to summation order."""
    import gridapmhd_amd as g
    _say(f"[assembly] scatter modes, {name}")
    sp, phys = CASES[name]()
    rng = np.random.default_rng(7)
    x = 0.1 * rng.standard_normal(sp.ndofs)
    ctx = g._lib.context_from_spaces(sp, phys)
    try:
        out = {}
        for mode in (2, 1, 0):
            ctx.set_assembly_mode(mode)
            out[mode] = ctx.jacobian(x)
            if mode != 0:
                assert np.array_equal(out[mode], ctx.jacobian(x))
        scale = np.max(np.abs(out[2]))
        for mode in (1, 0):
            err = np.max(np.abs(out[2] - out[mode])) / scale
            _say(f"  mode {mode} vs 2: max diff / max entry {err:.3e}")
            assert err < 1e-14
        with pytest.raises(g._lib.MhdError):
            ctx.set_assembly_mode(5)
    finally:
        ctx.close()


def test_face_terms_touch_face_rows_only_and_repeat():
    """Thin-wall faces on jittered (non-affine) cells: the face terms (GPU with
    minus without the wall faces) are nonzero only in the rows of the faces'
    own normal-flux dofs, so no entry receives two face contributions, and the
    residual and the Jacobian repeat bit for bit."""
    import gridapmhd_amd as g
    _say("[assembly] face rows, jittered 5x5x3")
    sp, phys = cases.duct_thin_wall_case((5, 5, 3), Ha=500.0, Re=50.0, cw=0.01, convection="newton",
                                         zeta_u=1.0, zeta_j=1.0, jitter=0.2)
    rng = np.random.default_rng(3)
    x = 0.1 * rng.standard_normal(sp.ndofs)
    ctx = g._lib.context_from_spaces(sp, phys)
    sp0 = cases.duct_thin_wall_case((5, 5, 3), Ha=500.0, Re=50.0, cw=0.01, convection="newton",
                                    zeta_u=1.0, zeta_j=1.0, jitter=0.2)[0]
    sp0.thin_wall_faces = []
    ctx0 = g._lib.context_from_spaces(sp0, phys)
    try:
        r1, r2 = ctx.residual(x), ctx.residual(x)
        j1, j2 = ctx.jacobian(x), ctx.jacobian(x)
        assert np.array_equal(r1, r2) and np.array_equal(j1, j2)
        face_rows = set()
        for tw in sp.thin_wall_faces:
            for c, lf in zip(tw["cell"], tw["lface"]):
                ids = sp.cell_dofs[c, 85 + 4 * lf:85 + 4 * lf + 4]
                face_rows.update(int(i) for i in ids if i >= 0)
        dr = r1 - ctx0.residual(x)
        rows = set(np.nonzero(dr)[0].tolist())
        _say(f"  {len(rows)} residual rows carry face terms, {len(face_rows)} face dofs")
        assert rows <= face_rows and len(rows) > 0
    finally:
        ctx.close()
        ctx0.close()


@pytest.mark.parametrize("name", ["stretched_box", "graded_3x2x2"])
def test_gpu_thin_wall_matches_closed_form(name):
    """The closed-form pin of the thin-wall term (tests/test_thin_wall.py)
    through the GPU face kernels: s* . R_face(j) against exact face integrals."""
    import gridapmhd_amd as g
    import test_thin_wall as tw
    _say(f"[assembly] thin-wall closed form, {name}")
    tau, cw, jw = 37.0, 0.173, 0.61
    kw = tw.CASES[name]
    sp = tw._case(tau=tau, cw=cw, jw=jw, **kw)
    rng = np.random.default_rng(3)
    J, S = tw._rt_field(rng), tw._rt_field(rng)
    xj, _ = g.fespaces.interpolate(sp, j=lambda X: np.stack([tw._eval(c, X) for c in J], 1))
    xs, _ = g.fespaces.interpolate(sp, j=lambda X: np.stack([tw._eval(c, X) for c in S], 1))
    phys = dict(alpha=0.0, beta=0.0, gamma=0.0, sigma=1.0, B=np.zeros(3), convection="none")
    ctx = g._lib.context_from_spaces(sp, phys)
    faces = sp.thin_wall_faces
    sp.thin_wall_faces = []
    ctx0 = g._lib.context_from_spaces(sp, phys)
    sp.thin_wall_faces = faces
    try:
        js = sp.field_slices()["j"]
        lhs = float(xs[js] @ (ctx.residual(xj) - ctx0.residual(xj))[js])
    finally:
        ctx.close()
        ctx0.close()
    rhs = tw._closed_form(J, S, tau, cw, jw, kw["bounds"])
    _say(f"  GPU {lhs!r} closed form {rhs!r}")
    assert abs(lhs - rhs) <= 1e-11 * max(1.0, abs(rhs))
