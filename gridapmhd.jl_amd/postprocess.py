"""Host-side post-processing of a solution: the fields at quadrature points
and the error norms of the Hunt application against its analytic solution.

``_hunt`` (src/Applications/hunt.jl:237-260) measures eu = u - u_h and
ej = j - j_h with ``Measure(Omega, 2 (order + 1))`` = tensor Gauss-Legendre
with 4 points per direction on every cell, and the H1 error
sqrt(int grad(eu) : grad(eu) + eu . eu).  This module evaluates u_h, grad u_h
(Q2) and j_h (RT1, contravariant Piola map with the face signs) at those
points from the free and Dirichlet values, exactly as the kernels define the
local bases.  It runs once after a solve and is not part of the hot path.
"""
from __future__ import annotations

import numpy as np

from . import analytic, fespaces, multigrid


def gauss_legendre01(n):
    p, w = np.polynomial.legendre.leggauss(n)
    return 0.5 * (p + 1.0), 0.5 * w


def tensor_rule(n):
    p, w = gauss_legendre01(n)
    xi = np.array([(p[i], p[j], p[k]) for k in range(n) for j in range(n) for i in range(n)])
    wt = np.array([w[i] * w[j] * w[k] for k in range(n) for j in range(n) for i in range(n)])
    return xi, wt


def _dq2_1d(t):
    return np.stack([4 * t - 3, 4 - 8 * t, 4 * t - 1], -1)


def q2_basis_grad(xi):
    """(P, 27, 3) reference gradients of the Q2 shape functions."""
    L = [multigrid._q2_1d(xi[:, a]) for a in range(3)]
    D = [_dq2_1d(xi[:, a]) for a in range(3)]
    gx = np.einsum("nk,nj,ni->nkji", L[2], L[1], D[0]).reshape(-1, 27)
    gy = np.einsum("nk,nj,ni->nkji", L[2], D[1], L[0]).reshape(-1, 27)
    gz = np.einsum("nk,nj,ni->nkji", D[2], L[1], L[0]).reshape(-1, 27)
    return np.stack([gx, gy, gz], -1)


def local_values(sp, x_free, dir_vals=None):
    """(ncells, 129) local dof values (free values and Dirichlet values)."""
    dv = sp.dir_vals if dir_vals is None else dir_vals
    cd = sp.cell_dofs.astype(np.int64)
    out = np.empty(cd.shape)
    m = cd >= 0
    out[m] = np.asarray(x_free)[cd[m]]
    out[~m] = np.asarray(dv)[-cd[~m] - 1]
    return out


def eval_fields(sp, x_free, xi, dir_vals=None, cells=None):
    """u_h, grad u_h, j_h and the physical points at reference points xi (P, 3)
    of every cell: x (C, P, 3), u (C, P, 3), G (C, P, 3, 3) with
    G[..., i, k] = d u_k / d x_i, j (C, P, 3), det J (C, P)."""
    cells = np.arange(sp.ncells) if cells is None else np.asarray(cells)
    loc = local_values(sp, x_free, dir_vals)[cells]
    xyz = sp.mesh.cell_xyz()[cells]
    X, J = fespaces.map_points(xyz, xi)                       # (C, P, 3), (C, P, 3, 3)
    det = np.linalg.det(J)
    Jinv = np.linalg.inv(J)                                   # d xi_l / d x_i = Jinv[l, i]
    N = multigrid.q2_basis(xi)                                # (P, 27)
    dN = q2_basis_grad(xi)                                    # (P, 27, 3)
    uc = loc[:, :81].reshape(-1, 3, 27)                       # (C, comp, node)
    u = np.einsum("pa,cka->cpk", N, uc)
    gref = np.einsum("pal,cka->cplk", dN, uc)                 # d u_k / d xi_l
    G = np.einsum("cpli,cplk->cpik", Jinv, gref)
    psi = multigrid.rt_basis(xi)                              # (P, 36, 3)
    sg = np.ones((cells.size, 36))
    sg[:, :24] = np.repeat(sp.rt_sign[cells].astype(np.float64), 4, axis=1)
    jc = loc[:, 85:121] * sg
    jref = np.einsum("prl,cr->cpl", psi, jc)
    j = np.einsum("cpil,cpl->cpi", J, jref) / det[..., None]
    return X, u, G, j, det


def _unique_xy(X):
    """Unique (x, y) columns of points (the duct fields do not depend on z)."""
    key = np.round(X[..., :2].reshape(-1, 2), 12)
    uniq, inv = np.unique(key, axis=0, return_inverse=True)
    return uniq, inv.reshape(-1)


def hunt_errors(sp, x_free, L, mu, sigma, grad_pz, Ha, nsums, u_scale=1.0, j_scale=1.0, nq=4, chunk=4096):
    """eu_h1, eu_l2, ej_l2 of u_h = u_scale * ubar_h and j_h = j_scale * jbar_h
    against analytical_hunt_u/j with nsums modes (src/Applications/hunt.jl:237-254),
    on the physical mesh: the computational mesh scaled by L (the reference's
    _warp), so points scale by L, gradients by 1/L and volumes by L^3."""
    xi, w = tensor_rule(nq)
    acc = np.zeros(3)
    for c0 in range(0, sp.ncells, chunk):
        cells = np.arange(c0, min(sp.ncells, c0 + chunk))
        X, u, G, j, det = eval_fields(sp, x_free, xi, cells=cells)
        X, G, det = L * X, G / L, det * L ** 3
        xy, inv = _unique_xy(X)
        pts = np.concatenate([xy, np.zeros((xy.shape[0], 1))], 1)
        ua, Ga = analytic.analytical_hunt_u(L, L, mu, grad_pz, Ha, nsums, pts, grad=True)
        ja = analytic.analytical_hunt_j(L, L, sigma, mu, grad_pz, Ha, nsums, pts)
        shp = X.shape[:2]
        eu = ua[inv].reshape(*shp, 3) - u_scale * u
        eG = Ga[inv].reshape(*shp, 3, 3) - u_scale * G
        ej = ja[inv].reshape(*shp, 3) - j_scale * j
        dV = np.abs(det) * w[None, :]
        acc += [np.sum(dV * np.einsum("cpk,cpk->cp", eu, eu)),
                np.sum(dV * np.einsum("cpik,cpik->cp", eG, eG)),
                np.sum(dV * np.einsum("cpk,cpk->cp", ej, ej))]
    return dict(eu_h1=float(np.sqrt(acc[0] + acc[1])), eu_l2=float(np.sqrt(acc[0])), ej_l2=float(np.sqrt(acc[2])))


def mean_uz(sp, x_free, nq=4):
    """Cell-integrated mean of u_z over the mesh, int u_z / |Omega|
    (uh_0 of src/Applications/FullyDevelopedFlow.jl:232)."""
    xi, w = tensor_rule(nq)
    X, u, G, j, det = eval_fields(sp, x_free, xi)
    dV = np.abs(det) * w[None, :]
    return float(np.sum(dV * u[..., 2]) / np.sum(dV))


def general_hunt_error(sp, x_free, l, d_b, grad_pz, Ha, nsums, nq=4):
    """Relative L2 error of u_h against analytical_general_hunt_u."""
    xi, w = tensor_rule(nq)
    X, u, G, j, det = eval_fields(sp, x_free, xi)
    xy, inv = _unique_xy(X)
    pts = np.concatenate([xy, np.zeros((xy.shape[0], 1))], 1)
    ua = analytic.analytical_general_hunt_u(l, d_b, grad_pz, Ha, nsums, pts)[inv].reshape(u.shape)
    dV = np.abs(det) * w[None, :]
    e = np.sum(dV * np.einsum("cpk,cpk->cp", ua - u, ua - u))
    n = np.sum(dV * np.einsum("cpk,cpk->cp", ua, ua))
    return float(np.sqrt(e / n))


def write_vtk(path, xh):
    """``writevtk`` of the solution fields (src/Applications/hunt.jl:263-273,
    expansion.jl:246-258) as a VTK unstructured grid: u, j, p, phi and div u
    at the 8 vertices of every cell (each cell keeps its own points, so the
    discontinuous fields are written as they are)."""
    from . import mesh_io
    sp, x = xh.spaces, xh.free_values
    corners = np.array([[v & 1, (v >> 1) & 1, (v >> 2) & 1] for v in range(8)], dtype=np.float64)
    X, u, G, j, _ = eval_fields(sp, x, corners)
    loc = local_values(sp, x)
    pb = np.stack([1 - corners.sum(1), corners[:, 0], corners[:, 1], corners[:, 2]], 1)   # (8, 4) dP1 basis
    p = loc[:, 81:85] @ pb.T
    phi = loc[:, 121:129]
    divu = np.trace(G, axis1=2, axis2=3)
    return mesh_io.write_vtu(path, sp.mesh.cell_xyz(), point_data={"uh": u, "jh": j, "ph": p, "phi": phi,
                                                                     "div_uh": divu})
