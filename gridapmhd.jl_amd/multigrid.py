"""Mesh hierarchies and [u, j] prolongations for the geometric coarse-grid
correction of the [u, j] block solve.

The hierarchy semi-coarsens x and y by 2 and keeps z, the refinement pattern
of the reference's GMG hierarchy (``nrefs = (2, 2, 1)``,
src/Meshers/hunt_mesher.jl:140-155; solver src/Solvers/gmg.jl:323-355).
Coarse nodes are taken from the fine mesh, so the coarse cells are unions of
fine cells; on meshes of axis-aligned boxes (uniform or graded) the fine
spaces contain the coarse ones.  The prolongation P (fine [u; j] free dofs x
coarse [u; j] free dofs) is then the exact embedding: every fine dof
functional is applied to every coarse basis function, at the coarse reference
point of the fine dof's physical position (Newton inversion of the coarse
trilinear map; on non-nested, jittered meshes this interpolates the coarse
field at the fine dof points).
  * u (Q2, nodal): coarse Q2 basis evaluated at the fine node.
  * j (RT1, nodal flux at Gauss points): the coarse field pushed forward with
    its Piola map, evaluated at the fine dof point, pulled back with the fine
    cell's Piola map, axis component of the fine dof.
"""
from __future__ import annotations

from typing import List, Tuple

import numpy as np
import scipy.sparse as sps

from . import fespaces
from .meshers import CartesianMesh

G2 = fespaces.G2


def coarsen_mesh(mesh: CartesianMesh, axes=(0, 1)) -> CartesianMesh:
    """Coarse mesh: every second node along the ``axes`` (x and y by default,
    the reference's hierarchy; ``(0,)`` keeps y, e.g. to keep Hartmann layers
    resolved on every level), all nodes in z."""
    nx, ny, nz = mesh.nc
    fx, fy = (2 if 0 in axes else 1), (2 if 1 in axes else 1)
    if nx % fx or ny % fy:
        raise ValueError("cell counts must be even along the coarsened axes")
    X = mesh.node_xyz.reshape(nz + 1, ny + 1, nx + 1, 3)
    Xc = X[:, ::fy, ::fx, :].reshape(-1, 3).copy()
    cm = CartesianMesh((nx // fx, ny // fy, nz), mesh.domain, mesh.isperiodic, Xc, dict(mesh.tags))
    return cm


def _ratios(fine: CartesianMesh, coarse: CartesianMesh):
    return fine.nc[0] // coarse.nc[0], fine.nc[1] // coarse.nc[1]


def _q2_1d(t):
    return np.stack([(2 * t - 1) * (t - 1), 4 * t * (1 - t), t * (2 * t - 1)], -1)


def _lin_g(t):
    g0, g1 = G2
    return np.stack([(g1 - t) / (g1 - g0), (t - g0) / (g1 - g0)], -1)


def q2_basis(xi):
    """(N, 27) Q2 shape values, node a = i + 3j + 9k."""
    Lx, Ly, Lz = _q2_1d(xi[:, 0]), _q2_1d(xi[:, 1]), _q2_1d(xi[:, 2])
    return np.einsum("nk,nj,ni->nkji", Lz, Ly, Lx).reshape(-1, 27)


def rt_basis(xi):
    """(N, 36, 3) nodal RT1 reference fields (same local order as the kernels)."""
    L = [_q2_1d(xi[:, a]) for a in range(3)]
    G = [_lin_g(xi[:, a]) for a in range(3)]
    out = np.zeros((xi.shape[0], 36, 3))
    for r in range(36):
        if r < 24:
            f, k = divmod(r, 4)
            d = 2 - f // 2
            m = 2 if f % 2 else 0
        else:
            d, k = divmod(r - 24, 4)
            m = 1
        t1, t2 = {0: (1, 2), 1: (0, 2), 2: (0, 1)}[d]
        out[:, r, d] = L[d][:, m] * G[t1][:, k % 2] * G[t2][:, k // 2]
    return out


def _trilinear(cell_xyz, xi, points=True):
    """Points and Jacobians J[n, i, l] = dx_i / dxi_l of trilinear cell maps:
    cell_xyz (n, 8, 3), xi (n, 3).  Vertex v = bx + 2 by + 4 bz."""
    C = np.ascontiguousarray(cell_xyz.transpose(1, 2, 0))   # (8, 3, n)
    t = [(1.0 - xi[:, a], np.ascontiguousarray(xi[:, a])) for a in range(3)]
    n = xi.shape[0]
    x = np.zeros((3, n)) if points else None
    J = np.zeros((3, 3, n))
    for v in range(8):
        bx, by, bz = v & 1, (v >> 1) & 1, (v >> 2) & 1
        fx, fy, fz = t[0][bx], t[1][by], t[2][bz]
        fyz, fxz, fxy = fy * fz, fx * fz, fx * fy
        if not bx:
            fyz = -fyz
        if not by:
            fxz = -fxz
        if not bz:
            fxy = -fxy
        X = C[v]
        if points:
            x += (fx * (fy * fz)) * X
        J[:, 0] += fyz * X
        J[:, 1] += fxz * X
        J[:, 2] += fxy * X
    return (x.T if points else None), J.transpose(2, 0, 1)


def _solve3(J, r):
    """J^-1 r for batches of 3 x 3 matrices (rows of the inverse = cross
    products of the columns over the determinant)."""
    c0, c1, c2 = J[:, :, 0], J[:, :, 1], J[:, :, 2]
    k0, k1, k2 = np.cross(c1, c2), np.cross(c2, c0), np.cross(c0, c1)
    det = np.einsum("ni,ni->n", c0, k0)
    return np.stack([np.einsum("ni,ni->n", k0, r), np.einsum("ni,ni->n", k1, r),
                     np.einsum("ni,ni->n", k2, r)], 1) / det[:, None]


def invert_map(cell_xyz, X, xi0=None, iters=30, tol=1e-12):
    """Reference coordinates of physical points X (n, 3) in cells cell_xyz
    (n, 8, 3): Newton iterations on the trilinear map, each point until its
    update is below `tol` (the error after that update is of order tol^2;
    affine cells converge in one step)."""
    xi = np.full(X.shape, 0.5) if xi0 is None else xi0.copy()
    act = np.arange(X.shape[0])
    for _ in range(iters):
        x, J = _trilinear(cell_xyz[act], xi[act])
        d = _solve3(J, X[act] - x)
        xi[act] += d
        act = act[np.max(np.abs(d), axis=1) >= tol]
        if act.size == 0:
            break
    return xi


def _locate(coarse: CartesianMesh, fine: CartesianMesh, fine_cells: np.ndarray, fine_xi: np.ndarray):
    """Coarse cell and coarse reference coordinates of fine reference points.
    The fine cell (i, j, k) lies in coarse cell (i//2, j//2, k); the coarse
    reference point is found from the physical point (the inverse of the
    coarse trilinear map), which is exact on graded axis-aligned meshes and
    is the natural interpolation on non-nested (jittered) ones."""
    nx, ny, _ = fine.nc
    rx, ry = _ratios(fine, coarse)
    fi = fine_cells % nx
    fj = (fine_cells // nx) % ny
    fk = fine_cells // (nx * ny)
    cnx, cny = coarse.nc[0], coarse.nc[1]
    cc = fi // rx + cnx * (fj // ry + cny * fk)
    X, _ = _trilinear(fine.cell_xyz()[fine_cells], fine_xi)
    guess = fine_xi.copy()
    guess[:, 0] = (fine_xi[:, 0] + (fi % rx)) / rx
    guess[:, 1] = (fine_xi[:, 1] + (fj % ry)) / ry
    xi = invert_map(coarse.cell_xyz()[cc], X, guess)
    return cc, xi


def prolongation(sp_f: fespaces.Spaces, sp_c: fespaces.Spaces) -> Tuple[sps.csr_matrix, sps.csr_matrix]:
    """P (n1_fine x n1_coarse, CSR) on the [u; j] blocks (badia2024 numbering),
    and the state interpolation R (coarse free u x fine free u, CSR)."""
    assert sp_f.style == "badia2024" and sp_c.style == "badia2024"
    fine, coarse = sp_f.mesh, sp_c.mesh
    n1f = sp_f.nfree["u"] + sp_f.nfree["j"]
    n1c = sp_c.nfree["u"] + sp_c.nfree["j"]
    cdf = sp_f.cell_dofs.astype(np.int64)
    cdc = sp_c.cell_dofs.astype(np.int64)
    rows, cols, vals = [], [], []
    # ---- u: one representative (cell, local node) per fine free node
    ids = cdf[:, 0:27]
    cells, loc = np.nonzero(ids >= 0)
    g = ids[cells, loc]
    _, first = np.unique(g, return_index=True)
    cells, loc, g = cells[first], loc[first], g[first]
    xi_f = fespaces.Q2_OFFSETS[loc] / 2.0
    cc, xi_c = _locate(coarse, fine, cells, xi_f)
    N = q2_basis(xi_c)                                   # (n, 27)
    for comp in range(3):
        gfine = cdf[cells, comp * 27 + loc]              # fine free id of this component
        gcoarse = cdc[cc][:, comp * 27:(comp + 1) * 27]  # (n, 27)
        m = gcoarse >= 0
        r = np.broadcast_to(gfine[:, None], gcoarse.shape)[m]
        rows.append(r); cols.append(gcoarse[m]); vals.append(N[m])
    # ---- j: one representative (cell, local dof) per fine free j dof
    off_j = 85
    ids = cdf[:, off_j:off_j + 36]
    cells, loc = np.nonzero(ids >= 0)
    g = ids[cells, loc]
    _, first = np.unique(g, return_index=True)
    cells, loc, g = cells[first], loc[first], g[first]
    pts, axis = fespaces._rt_dof_points()
    xi_f = pts[loc]
    cc, xi_c = _locate(coarse, fine, cells, xi_f)
    # Piola maps of the fine and coarse cells at the point
    Jf = _jac_at(fine.cell_xyz()[cells], xi_f)
    Jc = _jac_at(coarse.cell_xyz()[cc], xi_c)
    detf, detc = np.linalg.det(Jf), np.linalg.det(Jc)
    # fine dof = axis comp of det_f Jf^-1 (Jc psi_c / det_c)
    Mmat = np.matmul(np.linalg.inv(Jf), Jc) * (detf / detc)[:, None, None]         # (n, 3, 3)
    psi = rt_basis(xi_c)                                                             # (n, 36, 3)
    row_axis = Mmat[np.arange(len(cells)), axis[loc], :]                             # (n, 3)
    W = (psi * row_axis[:, None, :]).sum(-1)                                         # (n, 36)
    sgn_c = np.ones((len(cells), 36))
    sgn_c[:, :24] = np.repeat(sp_c.rt_sign[cc].astype(float), 4, axis=1)
    sgn_f = np.where(loc < 24, sp_f.rt_sign[cells, np.minimum(loc, 23) // 4].astype(float), 1.0)
    W = W * sgn_c * sgn_f[:, None]
    gcoarse = cdc[cc][:, off_j:off_j + 36]
    m = (gcoarse >= 0) & (np.abs(W) > 1e-14)
    rows.append(np.broadcast_to(g[:, None], gcoarse.shape)[m]); cols.append(gcoarse[m]); vals.append(W[m])
    r = np.concatenate(rows); c = np.concatenate(cols); v = np.concatenate(vals)
    keep = np.abs(v) > 1e-14
    P = sps.csr_matrix((v[keep], (r[keep], c[keep])), shape=(n1f, n1c))
    P.sum_duplicates()
    # ---- state transfer: the fine u field at every coarse free u node
    R = state_interpolation(sp_f, sp_c)
    return P, R


def _jac_at(cell_xyz, xi):
    """Jacobians of the trilinear maps of cells[k] at points xi[k] -> (n, 3, 3)."""
    return _trilinear(cell_xyz, xi, points=False)[1]


def state_interpolation(sp_f: fespaces.Spaces, sp_c: fespaces.Spaces) -> sps.csr_matrix:
    """R (coarse free u dofs x fine free u dofs, CSR): the fine Q2 velocity
    evaluated at the physical position of every coarse u node (the state the
    coarse Jacobian is assembled at).  A coarse node of a graded mesh that is
    not a fine node (the midpoint of a coarse edge) gets the fine field there,
    not the value of the nearest fine node.  Fine Dirichlet values are not
    part of the free state; the coarse Dirichlet values are the coarse
    level's own."""
    fine, coarse = sp_f.mesh, sp_c.mesh
    cdf = sp_f.cell_dofs.astype(np.int64)
    cdc = sp_c.cell_dofs.astype(np.int64)
    ids = cdc[:, 0:27]
    cells, loc = np.nonzero(ids >= 0)
    g = ids[cells, loc]
    _, first = np.unique(g, return_index=True)
    cells, loc = cells[first], loc[first]
    X, _ = _trilinear(coarse.cell_xyz()[cells], fespaces.Q2_OFFSETS[loc] / 2.0)
    # the fine cell of the coarse cell (I, J, K) holding X: fine cells (rx I + a, ry J + b, K)
    cnx, cny = coarse.nc[0], coarse.nc[1]
    nx, ny = fine.nc[0], fine.nc[1]
    rx, ry = _ratios(fine, coarse)
    I, J, K = cells % cnx, (cells // cnx) % cny, cells // (cnx * cny)
    best_cell = np.zeros(cells.size, dtype=np.int64)
    best_xi = np.zeros((cells.size, 3))
    best_out = np.full(cells.size, np.inf)
    fxyz = fine.cell_xyz()
    for a in range(rx):
        for b in range(ry):
            fc = (rx * I + a) + nx * ((ry * J + b) + ny * K)
            xi = invert_map(fxyz[fc], X)
            out = np.max(np.maximum(-xi, xi - 1.0), axis=1)
            take = out < best_out - 1e-12
            best_cell[take], best_xi[take], best_out[take] = fc[take], xi[take], out[take]
    N = q2_basis(best_xi)                                    # (n, 27)
    rows, cols, vals = [], [], []
    for comp in range(3):
        gc = cdc[cells, comp * 27 + loc]
        gf = cdf[best_cell][:, comp * 27:(comp + 1) * 27]
        m = gf >= 0
        rows.append(np.broadcast_to(gc[:, None], gf.shape)[m])
        cols.append(gf[m])
        vals.append(N[m])
    r, c, v = np.concatenate(rows), np.concatenate(cols), np.concatenate(vals)
    keep = np.abs(v) > 1e-14
    R = sps.csr_matrix((v[keep], (r[keep], c[keep])), shape=(sp_c.nfree["u"], sp_f.nfree["u"]))
    R.sum_duplicates()
    return R


def auto_levels(mesh: CartesianMesh, min_cells: int = 136, min_xy: int = 8, axes=(0, 1)) -> int:
    """Number of levels: halve the ``axes`` while their cell counts are even
    and >= `min_xy` and the level still has more than `min_cells` cells (about
    6,000 [u, j] dofs, small enough for the exact coarsest solve)."""
    n = list(mesh.nc)
    levels = 1
    while all(n[a] % 2 == 0 and n[a] >= min_xy for a in axes) and n[0] * n[1] * n[2] > min_cells:
        for a in axes:
            n[a] //= 2
        levels += 1
    return levels


def hierarchy(sp: fespaces.Spaces, nlevels: int, u_tags, j_tags, thin_walls) -> List[fespaces.Spaces]:
    """[fine, coarse1, ...] spaces with the same boundary conditions."""
    out = [sp]
    mesh = sp.mesh
    for _ in range(nlevels - 1):
        mesh = coarsen_mesh(mesh)
        out.append(fespaces.setup_fe_spaces(mesh, u_tags, j_tags, thin_walls, style=sp.style))
    return out
