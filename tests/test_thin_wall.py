"""Closed-form pin of the oracle's thin-wall face term.

res_thin_wall (src/weakforms.jl:721-726, evaluated at x) and jac_thin_wall
(:728-733) integrate  tau (s.n) (j.n - jw + cw n.(grad j).n)  over the wall
faces.  The residual is linear in the test function, so for any s* in the
discrete RT1 space  s* . R_face(j) = int_walls tau (s*.n)(j.n - jw + cw d_n j_n).
With j and s* polynomial fields of RT1 (component d of degree 2 in x_d and 1
in the other two coordinates) on axis-aligned boxes, both sides are exact:
the left from the oracle's assembled residual (with minus without the wall
faces), the right from exact monomial integrals over every wall face.  The
boxes are stretched (h_x != h_y != h_z) and the meshes graded, so a missing
1/h in the push-forward of grad j, a dropped cw, a wrong tau or jw, or a
wrong normal sign gives a different number; the planted variants below show
each of those mistakes is detected.
"""
import numpy as np
import pytest

import cases

DEG = 5


def _poly(rng, shape):
    """Random polynomial with monomials x^a y^b z^c, (a, b, c) < shape."""
    C = np.zeros((DEG, DEG, DEG))
    C[:shape[0], :shape[1], :shape[2]] = rng.uniform(-1.0, 1.0, shape)
    return C


def _eval(C, x):
    a = np.arange(DEG)
    return np.einsum("abc,na,nb,nc->n", C, x[:, 0:1] ** a, x[:, 1:2] ** a, x[:, 2:3] ** a)


def _deriv(C, axis):
    D = np.zeros_like(C)
    sl = [slice(None)] * 3
    for k in range(1, DEG):
        src = list(sl)
        dst = list(sl)
        src[axis] = k
        dst[axis] = k - 1
        D[tuple(dst)] = k * C[tuple(src)]
    return D


def _mul(A, B):
    out = np.zeros((2 * DEG - 1,) * 3)
    for a, b, c in zip(*np.nonzero(A)):
        out[a:a + DEG, b:b + DEG, c:c + DEG] += A[a, b, c] * B
    return out


def _face_integral(P, axis, value, box):
    """Exact integral of the polynomial P (coefficient cube) over the face
    x_axis = value of the box ((x0, x1), (y0, y1), (z0, z1))."""
    n = P.shape[0]
    e = np.arange(n)
    t1, t2 = [a for a in range(3) if a != axis]
    # collapse the normal coordinate
    Q = np.moveaxis(P, axis, 0)
    Q = np.tensordot(value ** e, Q, axes=(0, 0))        # (n, n) over (t1, t2) in increasing axis order
    (a0, a1), (b0, b1) = box[t1], box[t2]
    It1 = (a1 ** (e + 1) - a0 ** (e + 1)) / (e + 1)
    It2 = (b1 ** (e + 1) - b0 ** (e + 1)) / (e + 1)
    return float(It1 @ Q @ It2)


def _closed_form(J, S, tau, cw, jw, bounds, wall_axes=(0, 1), pushforward=True):
    """sum over wall faces of int tau (s.n)(j.n - jw + cw n.(grad j).n) dA."""
    total = 0.0
    for axis in wall_axes:
        for side, value in ((-1.0, bounds[axis][0]), (1.0, bounds[axis][1])):
            h = bounds[axis][1] - bounds[axis][0]
            dJ = _deriv(J[axis], axis) * (1.0 if pushforward else h)
            inner = side * J[axis] + cw * dJ
            inner[0, 0, 0] -= jw
            total += tau * _face_integral(_mul(side * S[axis], inner), axis, value, bounds)
    return total


def _rt_field(rng):
    shapes = [(3, 2, 2), (2, 3, 2), (2, 2, 3)]
    return [_poly(rng, s) for s in shapes]


def _case(nc, bounds, grade, tau, cw, jw):
    import gridapmhd_amd as g
    (x0, x1), (y0, y1), (z0, z1) = bounds
    mp = None
    if grade:
        # per-axis affine grading keeps every cell an axis-aligned box
        def mp(X):
            Y = X.copy()
            for a, (lo, hi) in enumerate(bounds):
                t = (X[:, a] - lo) / (hi - lo)
                Y[:, a] = lo + (hi - lo) * (t + 0.3 * t * (1 - t))
            return Y
    mesh = g.meshers.cartesian_model((x0, x1, y0, y1, z0, z1), nc, (False, False, False), map=mp)
    g.meshers.hunt_add_tags(mesh)
    sp = g.fespaces.setup_fe_spaces(mesh, "walls", [], [dict(domain="walls", cw=cw, tau=tau, jw=jw)])
    return sp


CASES = {
    "unit_cube": dict(nc=(1, 1, 1), bounds=((0.0, 1.0), (0.0, 1.0), (0.0, 1.0)), grade=False),
    "stretched_box": dict(nc=(1, 1, 1), bounds=((-0.3, 0.9), (0.1, 0.35), (0.0, 2.0)), grade=False),
    "graded_3x2x2": dict(nc=(3, 2, 2), bounds=((-1.0, 1.0), (-0.5, 0.7), (0.0, 0.3)), grade=True),
}


@pytest.mark.parametrize("name", list(CASES))
def test_thin_wall_residual_matches_closed_form(name, orc_mod):
    tau, cw, jw = 37.0, 0.173, 0.61
    kw = CASES[name]
    sp = _case(tau=tau, cw=cw, jw=jw, **kw)
    rng = np.random.default_rng(3)
    J, S = _rt_field(rng), _rt_field(rng)
    import gridapmhd_amd as g
    xj, _ = g.fespaces.interpolate(sp, j=lambda X: np.stack([_eval(c, X) for c in J], 1))
    xs, _ = g.fespaces.interpolate(sp, j=lambda X: np.stack([_eval(c, X) for c in S], 1))
    phys = dict(alpha=0.0, beta=0.0, gamma=0.0, sigma=1.0, B=np.zeros(3), convection="none")
    with_f = orc_mod.Oracle(cases.problem_dict(sp, phys))
    pb = cases.problem_dict(sp, phys)
    for k in ("face_cell", "face_lface", "face_tau", "face_cw", "face_jw"):
        pb[k] = pb[k][:0]
    without = orc_mod.Oracle(pb)
    js = sp.field_slices()["j"]
    r_face = with_f.residual(xj) - without.residual(xj)
    lhs = float(xs[js] @ r_face[js])
    rhs = _closed_form(J, S, tau, cw, jw, kw["bounds"])
    assert abs(lhs - rhs) <= 1e-11 * max(1.0, abs(rhs)), (lhs, rhs)
    # the face Jacobian: K x = R(x) + jw-term (the residual is affine in j)
    K_face = with_f.jacobian(xj) - without.jacobian(xj)
    lhs_k = float(xs[js] @ (K_face @ xj)[js])
    rhs_k = _closed_form(J, S, tau, cw, 0.0, kw["bounds"])
    assert abs(lhs_k - rhs_k) <= 1e-11 * max(1.0, abs(rhs_k)), (lhs_k, rhs_k)
    # planted mistakes in the closed form are told apart from the oracle
    planted = {
        "cw dropped": _closed_form(J, S, tau, 0.0, jw, kw["bounds"]),
        "tau halved": _closed_form(J, S, tau / 2, cw, jw, kw["bounds"]),
        "jw dropped": _closed_form(J, S, tau, cw, 0.0, kw["bounds"]),
        "walls on y only": _closed_form(J, S, tau, cw, jw, kw["bounds"], wall_axes=(1,)),
    }
    if name == "stretched_box":
        planted["reference-cell derivative"] = _closed_form(J, S, tau, cw, jw, kw["bounds"], pushforward=False)
    for what, v in planted.items():
        assert abs(lhs - v) > 1e-6 * max(1.0, abs(rhs)), what
