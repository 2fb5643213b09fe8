"""The CPU Newton-step baseline computes what it prices: its assembly by
binary search into the CSC pattern equals the oracle's sparse assembly, and
its sparse, patch, dense and vector kernels equal numpy / scipy on the same
operands.  (The baseline is timed only; these checks keep its work honest.)"""
import os
import sys

import numpy as np
import pytest
import scipy.sparse as sps

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import cases  # noqa: E402
import oracle  # noqa: E402


@pytest.fixture(scope="module")
def lib():
    oracle.build()
    import cpu_baseline as cb
    L, isa = cb.load()
    return L


def _p(a):
    return a.ctypes.data_as(__import__("ctypes").c_void_p).value


def test_assembly_matches_oracle(lib):
    import ctypes
    sp, phys = cases.duct_thin_wall_case((3, 3, 3), 30.0, 5.0, 0.02)
    orc = oracle.Oracle(cases.problem_dict(sp, dict(phys)))
    x = np.random.default_rng(3).uniform(-1, 1, sp.ndofs)
    colptr, rowval = orc.pattern()
    nz = np.zeros(rowval.size)
    miss = lib.orc_assemble_csc(ctypes.byref(orc.P), _p(x), 0, orc.ncells, _p(colptr), _p(rowval), _p(nz))
    assert miss == 0
    # the oracle's Jacobian minus its thin-wall face part
    A = orc.jacobian(x)
    fc = orc.keep_face_cell()
    fd = orc.cell_dofs[fc, oracle.OFF_J:oracle.OFF_J + 36].astype(np.int64)
    Kf = orc.face_jacobians(x)
    R = np.broadcast_to(fd[:, :, None], Kf.shape)
    C = np.broadcast_to(fd[:, None, :], Kf.shape)
    m = (R >= 0) & (C >= 0)
    Af = sps.csc_matrix((Kf[m], (R[m], C[m])), shape=A.shape)
    B = sps.csc_matrix((nz, rowval, colptr), shape=A.shape)
    assert abs(A - Af - B).max() <= 1e-12 * abs(A).max()
    r = np.zeros(sp.ndofs)
    lib.orc_assemble_residual(ctypes.byref(orc.P), _p(x), 0, orc.ncells, _p(r))
    re = orc.cell_residuals(x)
    cd = orc.cell_dofs
    rr = np.zeros(sp.ndofs)
    np.add.at(rr, cd[cd >= 0], re[cd >= 0])
    assert np.abs(r - rr).max() <= 1e-12 * np.abs(rr).max()


def test_kernels_match_numpy(lib):
    rng = np.random.default_rng(5)
    A = sps.random(300, 300, density=0.05, random_state=1, format="csr") + sps.eye(300)
    A = A.tocsr()
    A.sort_indices()
    rp, ci = A.indptr.astype(np.int64), A.indices.astype(np.int32)
    x = rng.uniform(-1, 1, 300)
    y = np.zeros(300)
    lib.cpu_spmv_f64(300, _p(rp), _p(ci), _p(A.data), _p(x), _p(y))
    assert np.abs(y - A @ x).max() < 1e-13
    v32, x32, y32 = A.data.astype(np.float32), x.astype(np.float32), np.zeros(300, np.float32)
    lib.cpu_spmv_f32(300, _p(rp), _p(ci), _p(v32), _p(x32), _p(y32))
    assert np.abs(y32 - A @ x).max() < 1e-4
    npch, PS = 5, 17
    idx = rng.integers(0, 300, npch * PS).astype(np.int32)
    M = rng.uniform(-1, 1, (npch, PS, PS))
    part = np.zeros(npch * PS)
    lib.cpu_patch_apply_f64(npch, PS, _p(idx), _p(M), _p(x), _p(part))
    ref = np.einsum("pij,pj->pi", M, x[idx.reshape(npch, PS)]).ravel()
    assert np.abs(part - ref).max() < 1e-13
    M32, part32 = M.astype(np.float32), np.zeros(npch * PS, np.float32)
    lib.cpu_patch_apply_f32(npch, PS, _p(idx), _p(M32), _p(x32), _p(part32))
    assert np.abs(part32 - ref).max() < 1e-5
    D = M + PS * np.eye(PS)
    Di = D.copy()
    assert lib.cpu_patch_invert_f64(npch, PS, _p(Di)) == 0
    assert np.abs(Di - np.linalg.inv(D)).max() < 1e-14
    m = 40
    G = rng.uniform(-1, 1, (m, m))
    xg, yg = rng.uniform(-1, 1, m), np.zeros(m)
    lib.cpu_dense_matvec(m, _p(G), _p(xg), _p(yg))
    assert np.abs(yg - G @ xg).max() < 1e-13
    V = rng.uniform(-1, 1, (7, 300))
    d = np.zeros(7)
    lib.cpu_mdot_f64(300, 7, _p(V), _p(x), _p(d))
    assert np.abs(d - V @ x).max() < 1e-12


def test_compose_prices_every_count():
    import cpu_baseline as cb
    rt = {"residual_per_cell_s": 1.0, "jacobian_per_cell_s": 10.0, "face_jacobians_s": 0.5,
          "patch_invert_per_dim3_s": 100.0, "dense_invert_per_dim3_s": 1000.0,
          "spmv_per_nnz_s": {"fp64": 1e4, "fp32": 1e5}, "patch_apply_per_entry_s": {"fp64": 1e6, "fp32": 1e7},
          "dense_per_entry_s": 1e8, "vector_bytes_per_s": 1e-9}
    counts = dict(spmv_nnz_fp64=1, spmv_nnz_fp32=1, patch_entries_fp64=1, patch_entries_fp32=1,
                  patch_entries_bf16=0, dense_entries=1, vector_bytes=1, setup_cells=1, setup_patch_dim3=1,
                  setup_dense_dim3=1, outer_iters=1)
    t, parts = cb.compose(rt, counts, ncells=2, residuals_per_step=3)
    assert parts == {"residual": 6.0, "jacobian": 20.5, "setup": 1110.0, "solve": 1e4 + 1e5 + 1e6 + 1e7 + 1e8 + 1e9}
    assert t == sum(parts.values())
