"""phg_spsi3pre_noalloc (Spsi3pre staged through the library's static device
array, no allocation) against phg_spsi3pre (staging buffer allocated on
first use), through the C-ABI (this file is collected last): the same kernels on the same inputs, so the
scaled lhsK must agree bit for bit.  Two calls in a row share the one
scratch array, the second on a smaller mesh, then a third on the first
mesh again; a mesh whose lhsK does not fit the array is refused.
"""
import ctypes

import numpy as np
import pytest

from phasta_amd import mesh, solver

pytestmark = pytest.mark.gpu
MAX_BLOCKS = (64 << 20) // (25 * 8)


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def inputs(case, gpu):
    """Factored BDiag and unscaled lhsK of the predicted state."""
    y, ac = case.yold.copy(order="F"), np.zeros((case.nshg, 5), order="F")
    _, _, bd, lk = gpu.elmgmrs(y, ac)
    bd_lu, _ = gpu.i3lu(bd, None, "LU_Fact ")
    return bd_lu, lk


def noalloc(gpu, bd_lu, lk):
    out = np.asfortranarray(lk.copy())
    rc = solver.lib().phg_spsi3pre_noalloc(gpu.h, _p(np.asfortranarray(bd_lu)), _p(out))
    if rc:
        raise solver.PhgError(rc, solver.lib().phg_last_error().decode())
    return out


def test_noalloc_spsi3pre_matches_the_allocating_entry_point():
    big = mesh.box_case(16, mach=0.3, kspace=10)
    small = mesh.box_case(4, mach=0.3, kspace=10, grade=1.1)
    assert big.nnz_tot <= MAX_BLOCKS and small.nnz_tot < big.nnz_tot
    g_big = solver.GpuSolver(big)
    g_small = solver.GpuSolver(small)
    try:
        bd_b, lk_b = inputs(big, g_big)
        bd_s, lk_s = inputs(small, g_small)
        want_b = g_big.spsi3pre(bd_b, lk_b)
        want_s = g_small.spsi3pre(bd_s, lk_s)
        print(f"[noalloc] box16 nnz {big.nnz_tot}, box4 graded nnz {small.nnz_tot}", flush=True)
        got_b = noalloc(g_big, bd_b, lk_b)
        got_s = noalloc(g_small, bd_s, lk_s)  # smaller input, same scratch
        again = noalloc(g_big, bd_b, lk_b)    # the last call's contents do not leak
        assert np.array_equal(got_b, want_b)
        assert np.array_equal(got_s, want_s)
        assert np.array_equal(again, want_b)
    finally:
        g_big.close()
        g_small.close()


def test_noalloc_spsi3pre_refuses_a_mesh_beyond_the_scratch():
    case = mesh.box_case(32, mach=0.3, kspace=10)
    assert case.nnz_tot > MAX_BLOCKS
    gpu = solver.GpuSolver(case)
    try:
        bd = np.zeros((case.nshg, 5, 5), order="F")
        lk = np.zeros((25, case.nnz_tot), order="F")
        with pytest.raises(solver.PhgError) as ei:
            noalloc(gpu, bd, lk)
        assert ei.value.code == 1 and "scratch" in str(ei.value)
    finally:
        gpu.close()


def test_noalloc_spsi3pre_refuses_a_mixed_context():
    case = mesh.box_case(4, mach=0.3, kspace=10)
    gpu = solver.GpuSolver(case, precision=solver.PHG_MIXED)
    try:
        bd = np.zeros((case.nshg, 5, 5), order="F")
        lk = np.zeros((25, case.nnz_tot), order="F")
        with pytest.raises(solver.PhgError) as ei:
            noalloc(gpu, bd, lk)
        assert ei.value.code == 1
    finally:
        gpu.close()
