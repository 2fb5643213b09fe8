"""Host-side logic (no GPU): mesher tags, dof numbering, interpolation, the
C-ABI library's exported symbols, and bench/entry-point plumbing."""
import ctypes
import os
import re

import numpy as np
import pytest

import cases

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_c_abi_exports_every_declared_symbol():
    """libmhd.so loads and exports every function include/mhd.h declares."""
    import gridapmhd_amd as g
    header = open(os.path.join(ROOT, "include", "mhd.h")).read()
    declared = set(re.findall(r"\b(mhd_[a-z_0-9]+)\s*\(", header))
    assert len(declared) >= 15
    lib = ctypes.CDLL(g._lib.LIB_PATH)
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, missing
    # the Python binding covers all of them too
    assert declared <= set(g._lib.EXPORTS), declared - set(g._lib.EXPORTS)


def test_struct_layouts_match_header():
    import gridapmhd_amd as g
    # sizes that the C compiler gives the ABI structs (LP64)
    assert ctypes.sizeof(g._lib.MhdMesh) == 4 + 4 + 8 + 8 + 4 + 4 + 5 * 8
    assert ctypes.sizeof(g._lib.MhdSpaces) == 16 + 4 + 4 + 3 * 8
    assert ctypes.sizeof(g._lib.MhdParams) == 8 * (4 + 3 + 3 + 9 + 3 + 9 + 2) + 8
    assert ctypes.sizeof(g._lib.MhdSolverOpts) == 4 * 4 + 3 * 8 + 2 * 4


def test_hunt_tags_and_walls():
    import gridapmhd_amd as g
    m = g.meshers.hunt_generate_base_mesh((4, 4), Ha=50.0, BL_adapted=False)
    # box entities: x- face 25, y+ face 24, the (x-, y-) edge along z is 17
    I = np.array([0, 4, 8, 0, 3])
    J = np.array([3, 8, 3, 0, 3])
    K = np.array([1, 1, 2, 1, 0])
    ents = m.lattice_entity(I, J, K)
    assert list(ents) == [25, 24, 26, 17, 27]
    assert list(m.tagged("insulating", I, J, K)) == [True, False, True, False, False]
    assert list(m.tagged("noslip", I, J, K)) == [True, True, True, True, False]


def test_periodic_identification_and_counts():
    import gridapmhd_amd as g
    m = g.meshers.hunt_generate_base_mesh((3, 2), nz=3, Ha=10.0, BL_adapted=False)
    sp = g.fespaces.setup_fe_spaces(m, "noslip", "insulating")
    # periodic z: 4*3*3 vertices, no vertex plane at z = top
    assert sp.cell_verts.max() + 1 == 4 * 3 * 3
    # every free dof appears in some cell; L2 fields are cell-major
    cd = sp.cell_dofs
    free = np.unique(cd[cd >= 0])
    assert free.size == sp.ndofs and free[0] == 0 and free[-1] == sp.ndofs - 1
    off = sp.offsets
    assert np.array_equal(cd[:, 81], off["p"] + 4 * np.arange(sp.ncells))
    # top-layer cells share their z+ face dofs with the bottom layer's z- faces
    ncxy = 3 * 2
    top = np.arange(2 * ncxy, 3 * ncxy)
    bot = np.arange(0, ncxy)
    assert np.array_equal(cd[top, 85 + 4:85 + 8], cd[bot, 85:89])


def test_thin_wall_faces_and_dofs():
    import gridapmhd_amd as g
    m = g.meshers.hunt_generate_base_mesh((64, 64), nz=8, domain=(-1, 1, -1, 1, 0, 0.3), Ha=100.0, map=lambda x: x)
    sp = g.fespaces.setup_fe_spaces(m, "noslip", [], [dict(domain="walls", cw=0.01, tau=100.0)])
    assert len(sp.thin_wall_faces[0]["cell"]) == 4 * 64 * 8
    assert sp.ndofs == 1_953_840 + 4 * 2 * 64 * 8
    sp2 = g.fespaces.setup_fe_spaces(m, "noslip", "insulating", style="badia2024")
    assert sp2.ndofs == 1_953_840
    o = sp2.offsets
    assert o["u"] == 0 and o["j"] == sp2.nfree["u"] and o["p"] == o["j"] + sp2.nfree["j"]


def test_interpolation_is_exact_for_space_fields(orc_mod):
    """Interpolating fields of the discrete spaces and evaluating them back
    reproduces them (u in Q2, j in RT1, p in P1, phi in Q1) on a jittered mesh
    for u / p / phi and an affine mesh for j."""
    import gridapmhd_amd as g
    for jitter in (0.0, 0.15):
        # non-periodic box: the fields below are not periodic in z
        m = g.meshers.cartesian_model((-1, 1, -1, 1, 0, 0.5), (3, 3, 3), (False, False, False))
        g.meshers.hunt_add_tags(m)
        if jitter:
            g.meshers.jitter_nodes(m, jitter, 11)
        sp = g.fespaces.setup_fe_spaces(m, "walls", [])
        u = lambda x: np.stack([x[:, 0] * x[:, 1], 1 + x[:, 2] ** 2, x[:, 0] - 2 * x[:, 1]], 1)
        j = lambda x: np.stack([x[:, 0] ** 2 + x[:, 1], x[:, 1] * 3 - 1, 2 * x[:, 2] + x[:, 0]], 1)
        p = lambda x: 1 + 2 * x[:, 0] - x[:, 1] + 0.5 * x[:, 2]
        phi = lambda x: x[:, 0] * x[:, 1] * x[:, 2]
        xs, dv = g.fespaces.interpolate(sp, u=u, p=p, j=j, phi=phi)
        orc = orc_mod.Oracle(cases.problem_dict(sp, dict(alpha=1, beta=1, gamma=1, B=np.zeros(3)), dv))
        rng = np.random.default_rng(0)
        cells = rng.integers(0, sp.ncells, 40).astype(np.int32)
        xi = rng.random((40, 3))
        vals = orc.point_values(xs, cells, xi)
        X, _ = g.fespaces.map_points(sp.mesh.cell_xyz()[cells], xi)
        Xp = np.stack([X[k, k] for k in range(40)])
        if jitter == 0.0:
            # u = (xy, 1 + z^2, x - 2y) is in Q2 only on affine cells
            assert np.abs(vals[:, 0:3] - u(Xp)).max() < 1e-12
            assert np.abs(vals[:, 3:6] - j(Xp)).max() < 1e-12
            assert np.abs(vals[:, 6] - p(Xp)).max() < 1e-12
        assert np.abs(vals[:, 7] - phi(Xp)).max() < 2e-2 if jitter else np.abs(vals[:, 7] - phi(Xp)).max() < 1e-12


def test_bench_refuses_multi_gpu():
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_splitmix_matches_reference_sequence():
    import importlib.util
    spec = importlib.util.spec_from_file_location("bench", os.path.join(ROOT, "bench.py"))
    b = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(b)
    import gridapmhd_amd as g
    # vectorised generator == scalar SplitMix64 of the mesher (first outputs, seed 1234)
    ref = g.meshers._splitmix64(5, 1234)
    u = (ref >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    assert np.array_equal(b.splitmix64_uniform(5, 1234), u)


def test_bl_adapted_hunt_map():
    """The Smolentsev-stretched Hunt mesh (BL_adapted, src/Meshers/hunt_mesher.jl:5-46):
    the map keeps the walls and the centre lines, is odd in x and y and
    monotone, leaves z alone, and clusters cells at the walls so that the
    Hartmann layers (width 1/Ha at y = +-L) and the side layers (1/sqrt(Ha) at
    x = +-L) hold whole cells; the kmap = 1 map is the identity."""
    import gridapmhd_amd as g
    Ha, L = 50.0, 1.0
    m = g.meshers.hunt_stretch_map(L, Ha, BL_adapted=True)
    t = np.linspace(-L, L, 33)
    P = np.stack([t, t[::-1], np.linspace(0.0, 0.1, 33)], 1)
    Q = m(P)
    assert np.allclose(Q[[0, 16, 32], :2], P[[0, 16, 32], :2], atol=1e-14)
    assert np.allclose(m(-P)[:, :2], -Q[:, :2], atol=1e-14) and np.array_equal(Q[:, 2], P[:, 2])
    assert (np.diff(Q[:, 0]) > 0).all() and (np.diff(Q[:, 1]) < 0).all()
    mesh = g.meshers.hunt_generate_base_mesh((16, 16), Ha=Ha, BL_adapted=True)
    xs = np.unique(np.round(mesh.node_xyz[:, 0], 13))
    ys = np.unique(np.round(mesh.node_xyz[:, 1], 13))
    assert ys[1] - ys[0] < 1.0 / Ha and ys[-1] - ys[-2] < 1.0 / Ha
    assert xs[1] - xs[0] < 1.0 / np.sqrt(Ha) and xs[-1] - xs[-2] < 1.0 / np.sqrt(Ha)
    assert (ys[1] - ys[0]) < (xs[1] - xs[0]) < 2.0 * L / 16
    ident = g.meshers.hunt_stretch_map(L, Ha, kmap_x=1, kmap_y=1, BL_adapted=False)
    assert np.allclose(ident(P), P, atol=1e-15)
