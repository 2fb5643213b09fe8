"""The [u, j] prolongation and the coarse-state interpolation of the V-cycle
hierarchy (multigrid.py) on uniform and graded meshes.

On meshes of axis-aligned boxes the fine Q2 / RT1 spaces contain the coarse
ones, so P must reproduce every coarse-space field exactly: the fine
interpolant of a field of the coarse space equals P times its coarse
interpolant.  Likewise R (the fine velocity at the coarse u nodes) must
reproduce the coarse interpolant of a field of the fine space.  The test
fields are global polynomials of the spaces' degrees (u: degree 2 per axis;
j_d: degree 2 in x_d, 1 in the other coordinates), which lie in both spaces
on any box mesh."""
import numpy as np
import pytest


def _mesh(kind):
    import gridapmhd_amd as g
    dom = (-1.0, 1.0, -1.0, 1.0, 0.0, 0.3)
    if kind == "uniform":
        return g.meshers.cartesian_model(dom, (8, 6, 3), (False, False, False))
    if kind == "tanh3":
        return g.meshers.cartesian_model(dom, (8, 8, 3), (False, False, False), map=g.meshers.tanh_map(3.0))
    if kind == "tanh2_xy":
        return g.meshers.cartesian_model(dom, (8, 8, 2), (False, False, False),
                                         map=g.meshers.tanh_map(2.0, axes=(0, 1)))
    if kind == "hunt_smolentsev":
        return g.meshers.hunt_generate_base_mesh((8, 8), Ha=50.0, BL_adapted=True)
    raise ValueError(kind)


def _u(X, zw=1.0):
    x, y, z = X[:, 0], X[:, 1], zw * X[:, 2]
    return np.stack([x * x + 0.3 * y * z - 0.2, x * y + 0.7 * z * z + x, y * y - x * z + 0.1 * x * y * z], 1)


def _j(X, zw=1.0):
    x, y, z = X[:, 0], X[:, 1], zw * X[:, 2]
    return np.stack([0.4 * x * x * y + z - 0.3 * x * z + 1.0, y * y * x - 0.2 * y + 0.5 * z * x,
                     z * z + 0.6 * x * y - 0.1 * z * y], 1)


@pytest.mark.parametrize("axes", [(0, 1), (0,)])
@pytest.mark.parametrize("kind", ["uniform", "tanh3", "tanh2_xy", "hunt_smolentsev"])
def test_prolongation_reproduces_coarse_fields(kind, axes):
    """Both hierarchies: x and y halved (the reference's), and x only."""
    import gridapmhd_amd as g
    fine = _mesh(kind)
    g.meshers.hunt_add_tags(fine)
    coarse = g.multigrid.coarsen_mesh(fine, axes=axes)
    assert coarse.nc == (fine.nc[0] // 2, fine.nc[1] // (2 if 1 in axes else 1), fine.nc[2])
    spf = g.fespaces.setup_fe_spaces(fine, [], [], style="badia2024")
    spc = g.fespaces.setup_fe_spaces(coarse, [], [], style="badia2024")
    P, R = g.multigrid.prolongation(spf, spc)
    # a periodic direction needs periodic fields: the Hunt mesh is periodic in z
    zw = 0.0 if fine.isperiodic[2] else 1.0
    u, j = (lambda X: _u(X, zw)), (lambda X: _j(X, zw))
    xf, _ = g.fespaces.interpolate(spf, u=u, j=j)
    xc, _ = g.fespaces.interpolate(spc, u=u, j=j)
    n1f = spf.nfree["u"] + spf.nfree["j"]
    n1c = spc.nfree["u"] + spc.nfree["j"]
    err = np.abs(P @ xc[:n1c] - xf[:n1f]).max() / np.abs(xf[:n1f]).max()
    assert err < 1e-13, err
    nuf, nuc = spf.nfree["u"], spc.nfree["u"]
    err_r = np.abs(R @ xf[:nuf] - xc[:nuc]).max() / np.abs(xc[:nuc]).max()
    assert err_r < 1e-13, err_r
    # the graded meshes are graded (a midpoint rule for the coarse reference
    # coordinate, exact only on uniform meshes, would fail the checks above)
    if kind != "uniform":
        fy = np.unique(np.round(fine.node_xyz.reshape(-1, 3)[:, 1], 12))
        assert np.ptp(np.diff(fy)) > 1e-3


def test_prolongation_reproduces_linear_fields_on_jittered_cells():
    """On non-affine (jittered) cells the meshes are not nested, but linear
    velocities lie in every isoparametric Q2 space: P reproduces them (the mesh
    is not periodic: a linear field is not)."""
    import gridapmhd_amd as g
    fine = g.meshers.cartesian_model((-1.0, 1.0, -1.0, 1.0, 0.0, 0.3), (8, 8, 3), (False, False, False))
    g.meshers.hunt_add_tags(fine)
    g.meshers.jitter_nodes(fine, 0.2, 20261202)
    coarse = g.multigrid.coarsen_mesh(fine)
    spf = g.fespaces.setup_fe_spaces(fine, [], [], style="badia2024")
    spc = g.fespaces.setup_fe_spaces(coarse, [], [], style="badia2024")
    P, R = g.multigrid.prolongation(spf, spc)
    lin = lambda X: np.stack([X[:, 0] + 2 * X[:, 1] - X[:, 2], 0.5 * X[:, 0] - X[:, 2], X[:, 1] + 0.3], 1)
    xf, _ = g.fespaces.interpolate(spf, u=lin)
    xc, _ = g.fespaces.interpolate(spc, u=lin)
    nuf, nuc = spf.nfree["u"], spc.nfree["u"]
    Pu = P[:nuf, :nuc]
    assert np.abs(Pu @ xc[:nuc] - xf[:nuf]).max() < 1e-12
    assert np.abs(R @ xf[:nuf] - xc[:nuc]).max() < 1e-12
