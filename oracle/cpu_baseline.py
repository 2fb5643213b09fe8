"""CPU Newton-step baseline -- TEST INFRASTRUCTURE ONLY.

Only the CPU-baseline leg of bench.py imports this module.  It prices one
FP64 Newton step of the benchmark duct on the host cores, the work the
reference's CPU path does per step (src/main.jl:173-190 with the Badia2024
solver, src/Solvers/badia2024.jl:2-48 and the GMG of src/Solvers/gmg.jl):

  residual    cell residuals assembled into a vector (oracle element kernels)
  jacobian    cell Jacobians assembled into the CSC pattern by binary search,
              plus the thin-wall face Jacobians
  set-up      coarse-level Jacobians, Gauss-Jordan patch inversions, the
              dense coarsest inversion (LAPACK through numpy)
  solve       sparse products, patch-inverse products, dense coarsest
              products and Krylov vector work

Each operation is timed on a bounded sample at the real sizes (the element
kernels on a prefix of the cells, the sparse products on the duct's own CSC
pattern, the patch kernels on a few thousand 225 x 225 patches) and scaled
by the operation counts of the GPU solve (mhd_op_counts): the CPU does
exactly the work of the same algorithm, at its own per-operation rates.  The
element kernels are the oracle's restatement (mhd_oracle.c) built at -O3 for
the host's vector ISA; every kernel runs under OpenMP on all threads.
"""
from __future__ import annotations

import ctypes
import os
import time

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))


def _cpu_flags():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("flags"):
                return set(line.split(":", 1)[1].split())
    except OSError:
        pass
    return set()


def host_info():
    model = "unknown"
    try:
        model = [l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo") if l.startswith("model name")][0]
    except (OSError, IndexError):
        pass
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return dict(model=model, nproc=avail, cpu_count=os.cpu_count(), threads=threads)


def load():
    """The widest build of libcpustep the host supports (AVX-512 or AVX2)."""
    v4 = {"avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl"} <= _cpu_flags()
    path = os.path.join(_HERE, "_build", "libcpustep_v4.so" if v4 else "libcpustep_v3.so")
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built (make -C oracle)")
    L = ctypes.CDLL(path)
    vp, i64, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int
    L.orc_assemble_residual.argtypes = [vp, vp, i32, i32, vp]
    L.orc_assemble_csc.argtypes = [vp, vp, i32, i32, vp, vp, vp]
    L.orc_assemble_csc.restype = i64
    L.orc_face_jacobians.argtypes = [vp, vp, vp]
    for f in ("cpu_spmv_f64", "cpu_spmv_f32"):
        getattr(L, f).argtypes = [i64, vp, vp, vp, vp, vp]
    for f in ("cpu_patch_apply_f64", "cpu_patch_apply_f32"):
        getattr(L, f).argtypes = [i32, i32, vp, vp, vp, vp]
    L.cpu_patch_invert_f64.argtypes = [i32, i32, vp]
    L.cpu_dense_matvec.argtypes = [i64, vp, vp, vp]
    L.cpu_mdot_f64.argtypes = [i64, i32, vp, vp, vp]
    return L, ("x86-64-v4" if v4 else "x86-64-v3")


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p).value


def _best(fn, reps):
    fn()
    ts = []
    for _ in range(reps):
        t = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t)
    return min(ts)


def rates(orc, colptr, rowval, n1, nfree, log=print, cell_budget_s=6.0, patch_dim=225):
    """Per-operation CPU times on the host.  ``orc`` is an oracle.Oracle bound
    to the problem, (colptr, rowval) its 0-based CSC pattern (structurally
    symmetric: the same arrays are a CSR of the pattern)."""
    L, isa = load()
    P = ctypes.byref(orc.P)
    rng = np.random.default_rng(7)
    x = rng.uniform(-1, 1, nfree)
    ncells = orc.ncells
    out = {"isa": isa}
    # element kernels: a prefix of the cells, sized to the budget
    r = np.zeros(nfree)
    n0 = min(ncells, 64)
    t = time.perf_counter()
    L.orc_assemble_residual(P, _p(x), 0, n0, _p(r))
    dt = time.perf_counter() - t
    nr = int(min(ncells, max(n0, n0 * 0.2 * cell_budget_s / max(dt, 1e-6))))
    t = time.perf_counter()
    L.orc_assemble_residual(P, _p(x), 0, nr, _p(r))
    out["residual_per_cell_s"] = (time.perf_counter() - t) / nr
    nz = np.zeros(rowval.size)
    t = time.perf_counter()
    miss = L.orc_assemble_csc(P, _p(x), 0, n0, _p(colptr), _p(rowval), _p(nz))
    dt = time.perf_counter() - t
    if miss:
        raise RuntimeError(f"{miss} cell-matrix entries outside the CSC pattern")
    nj = int(min(ncells, max(n0, n0 * 0.8 * cell_budget_s / max(dt, 1e-6))))
    t = time.perf_counter()
    L.orc_assemble_csc(P, _p(x), 0, nj, _p(colptr), _p(rowval), _p(nz))
    out["jacobian_per_cell_s"] = (time.perf_counter() - t) / nj
    out["cells_sampled"] = dict(residual=nr, jacobian=nj, of=ncells)
    del nz
    nf = orc.nfaces
    if nf:
        kf = np.zeros((nf, 36, 36))
        t = time.perf_counter()
        L.orc_face_jacobians(P, _p(x), _p(kf))
        out["face_jacobians_s"] = time.perf_counter() - t
        del kf
    else:
        out["face_jacobians_s"] = 0.0
    log(f"cpu: element kernels {out['residual_per_cell_s'] * 1e3:.3f} ms/cell residual, "
        f"{out['jacobian_per_cell_s'] * 1e3:.3f} ms/cell Jacobian (CSC), {nr}/{nj} of {ncells} cells")
    # sparse products on the duct's own pattern
    n = colptr.size - 1
    rp = np.ascontiguousarray(colptr, dtype=np.int64)
    ci = rowval.astype(np.int32)
    nnz = ci.size
    v64 = rng.uniform(-1, 1, nnz)
    xv, yv = rng.uniform(-1, 1, n), np.zeros(n)
    t64 = _best(lambda: L.cpu_spmv_f64(n, _p(rp), _p(ci), _p(v64), _p(xv), _p(yv)), 4)
    del v64
    v32 = rng.uniform(-1, 1, nnz).astype(np.float32)
    xf, yf = xv.astype(np.float32), np.zeros(n, np.float32)
    t32 = _best(lambda: L.cpu_spmv_f32(n, _p(rp), _p(ci), _p(v32), _p(xf), _p(yf)), 4)
    del v32, ci
    out["spmv_per_nnz_s"] = {"fp64": t64 / nnz, "fp32": t32 / nnz}
    log(f"cpu: spmv {t64 * 1e3:.1f} ms FP64 ({nnz * 12 / t64 / 1e9:.0f} GB/s), {t32 * 1e3:.1f} ms FP32, nnz={nnz}")
    # patch-inverse products and inversions
    PS = patch_dim
    npa = 2048
    idx = rng.integers(0, n1, size=npa * PS).astype(np.int32)
    A = rng.uniform(-1, 1, (npa, PS, PS))
    rr = rng.uniform(-1, 1, n1)
    part = np.zeros(npa * PS)
    tp64 = _best(lambda: L.cpu_patch_apply_f64(npa, PS, _p(idx), _p(A), _p(rr), _p(part)), 3)
    A32 = A.astype(np.float32)
    del A
    rr32, part32 = rr.astype(np.float32), np.zeros(npa * PS, np.float32)
    tp32 = _best(lambda: L.cpu_patch_apply_f32(npa, PS, _p(idx), _p(A32), _p(rr32), _p(part32)), 3)
    del A32
    out["patch_apply_per_entry_s"] = {"fp64": tp64 / (npa * PS * PS), "fp32": tp32 / (npa * PS * PS)}
    npi = 512
    B = rng.uniform(-1, 1, (npi, PS, PS)) + PS * np.eye(PS)
    t = time.perf_counter()
    bad = L.cpu_patch_invert_f64(npi, PS, _p(B))
    tgj = time.perf_counter() - t
    if bad:
        raise RuntimeError("zero pivot in the patch inversion sample")
    out["patch_invert_per_dim3_s"] = tgj / (npi * PS ** 3)
    del B
    log(f"cpu: patch products {tp64 * 1e3:.1f} ms FP64 / {tp32 * 1e3:.1f} ms FP32 for {npa} patches, "
        f"inversion {tgj / npi * 1e3:.2f} ms/patch")
    # dense coarsest products and inversion, Krylov vector work
    m = 4096
    D = rng.uniform(-1, 1, (m, m))
    xd, yd = rng.uniform(-1, 1, m), np.zeros(m)
    td = _best(lambda: L.cpu_dense_matvec(m, _p(D), _p(xd), _p(yd)), 5)
    out["dense_per_entry_s"] = td / (m * m)
    mi = 2048
    Di = D[:mi, :mi] + mi * np.eye(mi)
    ti = _best(lambda: np.linalg.inv(Di), 2)
    out["dense_invert_per_dim3_s"] = ti / mi ** 3
    del D, Di
    K = 30
    V = rng.uniform(-1, 1, (K, n))
    dots = np.zeros(K)
    tv = _best(lambda: L.cpu_mdot_f64(n, K, _p(V), _p(xv), _p(dots)), 3)
    out["vector_bytes_per_s"] = (K + 1) * n * 8 / tv
    del V
    log(f"cpu: dense {td * 1e3:.2f} ms ({m}^2 matvec), inverse {ti:.3f} s ({mi}^2), "
        f"vector work {out['vector_bytes_per_s'] / 1e9:.0f} GB/s")
    return out


def compose(rt, counts, ncells, residuals_per_step):
    """Seconds of one FP64 Newton step on the CPU: the measured per-operation
    times of ``rates`` times the per-step operation counts of the GPU solve."""
    parts = {
        "residual": residuals_per_step * ncells * rt["residual_per_cell_s"],
        "jacobian": ncells * rt["jacobian_per_cell_s"] + rt["face_jacobians_s"],
        "setup": (counts["setup_cells"] * rt["jacobian_per_cell_s"]
                  + counts["setup_patch_dim3"] * rt["patch_invert_per_dim3_s"]
                  + counts["setup_dense_dim3"] * rt["dense_invert_per_dim3_s"]),
        "solve": (counts["spmv_nnz_fp64"] * rt["spmv_per_nnz_s"]["fp64"]
                  + counts["spmv_nnz_fp32"] * rt["spmv_per_nnz_s"]["fp32"]
                  + counts["patch_entries_fp64"] * rt["patch_apply_per_entry_s"]["fp64"]
                  + counts["patch_entries_fp32"] * rt["patch_apply_per_entry_s"]["fp32"]
                  + counts["dense_entries"] * rt["dense_per_entry_s"]
                  + counts["vector_bytes"] / rt["vector_bytes_per_s"]),
    }
    return sum(parts.values()), parts
