"""The allocation-free residual entry point (mhd_residual_d_noalloc) against
mhd_residual_d on the same inputs, through the C ABI: two calls in a row on
different contexts, the second with the smaller mesh, each compared bit for
bit (the same kernels sum in the same fixed order), with thin-wall faces and
jittered cells on the first."""
import sys

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_noalloc_residual_matches_residual_d():
    import gridapmhd_amd as g
    from devmem import DeviceArray
    problems = [
        ("jittered thin-wall duct 12x10x4", cases.duct_thin_wall_case((12, 10, 4), Ha=500.0, Re=50.0, cw=0.01,
                                                                      convection="newton", zeta_u=1.0, zeta_j=1.0,
                                                                      jitter=0.2)),
        ("hunt 6x6x3", cases.hunt_case(6, Ha=50.0)),
    ]
    sizes = []
    for label, (sp, phys) in problems:
        _say(f"[noalloc] {label}: ncells={sp.ncells} ndofs={sp.ndofs}")
        ctx = g._lib.context_from_spaces(sp, phys)
        try:
            xh = 0.1 * np.random.default_rng(sp.ncells).standard_normal(sp.ndofs)
            x = DeviceArray.from_numpy(xh)
            r_ref = DeviceArray.from_numpy(np.zeros(sp.ndofs))
            r_na = DeviceArray.from_numpy(np.full(sp.ndofs, np.nan))
            ctx.residual_d(x.ptr, r_ref.ptr)
            ctx.residual_d_noalloc(x.ptr, r_na.ptr)
            a, b = r_na.numpy(), r_ref.numpy()
            _say(f"  max |noalloc - residual_d| = {np.max(np.abs(a - b)):.3e}, |r| = {np.linalg.norm(b):.3e}")
            assert np.array_equal(a, b)
            # the host-buffer residual agrees too
            assert np.array_equal(ctx.residual(xh), b)
            sizes.append(sp.ncells)
        finally:
            ctx.close()
    assert sizes[1] < sizes[0]
