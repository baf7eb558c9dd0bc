"""Discontinuity capturing (iDC 1, 2, 3) on the GPU: e3dc.f's DC factor
and flux, with e3mtrx.f's A0DC, A0inv and dV/dY and e3tau.f's rTLS, raLS
and giju (called at e3.f:217-222), against the CPU oracle (bit for bit
equal to the reference's compiled routines) and the reference fixture.

* ElmGMRs res, BDiag, lhsK for iDC 1, 2, 3 on 4^3, 16^3 and the 8^3 graded
  box at a strongly perturbed state (the DC factor is active at most
  points), 1e-12; DC moves the residual and the tangent by far more than
  the bar;
* against the fixture made by the reference's own compiled ElmGMRs with
  iDC 1, 2, 3 (tests/golden/ref_dc_box3.npz), 1e-12;
* SolGMRs with DC: Dy at 1e-12 and the same Krylov count;
* the FP64 tangent product (the mixed path's refinement operator) with
  DC against the oracle's SparseAp + bc3per on its assembled lhsK, 1e-12;
* the mixed-precision path with DC against FP64: residual 1e-13,
  tangent and BDiag 2e-6;
* iDC outside 0..3 is refused.
"""
import os

import numpy as np
import pytest

import pyoracle
from phasta_amd import mesh, solver

pytestmark = pytest.mark.gpu
TOL = 1e-12
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module", params=[(4, 1.0), (16, 1.0), (8, 1.1)],
                ids=["box4", "box16", "box8_graded"])
def state(request):
    """A box at a perturbed predicted state (velocity, pressure and
    temperature jittered by 5 %, large accelerations)."""
    N, grade = request.param
    case = mesh.box_case(N, mach=0.5, kspace=10, jitter=0.1, grade=grade)
    print(f"[dc] box{N} grade {grade}: nshg={case.nshg}", flush=True)
    rng = np.random.default_rng(N)
    y, ac = pyoracle.Oracle(case).itr_predict_bc(case.yold)
    y = np.asfortranarray(y * (1 + 0.05 * rng.standard_normal(y.shape)))
    ac = np.asfortranarray(ac + 1e3 * rng.standard_normal(ac.shape))
    return case, y, ac


@pytest.mark.parametrize("idc", [1, 2, 3])
def test_dc_elmgmrs_against_oracle(state, idc):
    case, y, ac = state
    prm = dict(case.params, idc=idc)
    ref = pyoracle.Oracle(case, params=prm).elmgmrs(y, ac)
    base = pyoracle.Oracle(case).elmgmrs(y, ac)
    gpu = solver.GpuSolver(case, params=prm)
    try:
        out = gpu.elmgmrs(y, ac)
    finally:
        gpu.close()
    for name, k in (("res", 0), ("BDiag", 2), ("lhsK", 3)):
        e = rel(out[k], ref[k])
        print(f"[dc] iDC {idc} {name}: GPU vs oracle {e:.2e}, DC term {rel(ref[k], base[k]):.1e}",
              flush=True)
        assert e < TOL, name
        assert rel(ref[k], base[k]) > 1e3 * TOL, name


def test_dc_elmgmrs_against_reference_golden():
    g = np.load(os.path.join(GOLD, "ref_dc_box3.npz"))
    case = mesh.box_case(3, mach=0.5, kspace=10, jitter=0.1)
    y, ac = np.asfortranarray(g["y"]), np.asfortranarray(g["ac"])
    for idc in (1, 2, 3):
        gpu = solver.GpuSolver(case, params=dict(case.params, idc=idc))
        try:
            res, _, bd, lk = gpu.elmgmrs(y, ac)
        finally:
            gpu.close()
        assert rel(res, g[f"res{idc}"]) < TOL, idc
        assert rel(bd, g[f"BDiag{idc}"]) < TOL, idc
        assert rel(lk, g[f"lhsK{idc}"]) < TOL, idc


@pytest.mark.parametrize("idc", [1, 3])
def test_dc_solgmrs_against_oracle(state, idc):
    case, y, ac = state
    prm = dict(case.params, idc=idc)
    ref = pyoracle.Oracle(case, params=prm).solgmrs(y, ac)
    gpu = solver.GpuSolver(case, params=prm)
    try:
        out = gpu.solgmrs(y, ac)
    finally:
        gpu.close()
    assert out["iKs"] == ref["iKs"]
    for k in ("res", "BDiag", "Dy"):
        assert rel(out[k], ref[k]) < TOL, k


@pytest.mark.parametrize("prec", [solver.PHG_FP64, solver.PHG_MIXED], ids=["fp64", "mixed"])
def test_dc_tangent_product_against_oracle(state, prec):
    case, y, ac = state
    prm = dict(case.params, idc=1)
    _, _, _, lk = pyoracle.Oracle(case, params=prm).elmgmrs(y, ac)
    orc = pyoracle.Oracle(case, params=prm)
    gpu = solver.GpuSolver(case, params=prm, precision=prec)
    try:
        p = np.asfortranarray(np.random.default_rng(3).standard_normal((case.nshg, 5)))
        with pytest.raises(solver.PhgError):
            gpu.tangent_ap(y, p)  # no assembly has formed the DC factors yet
        gpu.elmgmrs(y, ac)
        q = p.copy(order="F")
        orc.sparse_ap(lk, q)
        orc.bc3per(q)
        assert rel(gpu.tangent_ap(y, p), q) < TOL
    finally:
        gpu.close()


@pytest.mark.parametrize("idc", [1, 2, 3])
def test_dc_mixed_against_fp64(state, idc):
    case, y, ac = state
    prm = dict(case.params, idc=idc)
    g64 = solver.GpuSolver(case, params=prm)
    gmx = solver.GpuSolver(case, params=prm, precision=solver.PHG_MIXED)
    try:
        r64, _, bd64, lk64 = g64.elmgmrs(y, ac)
        rmx, _, bdmx, lkmx = gmx.elmgmrs(y, ac)
    finally:
        g64.close()
        gmx.close()
    assert rel(rmx, r64) < 1e-13
    assert rel(bdmx, bd64) < 2e-6 and rel(lkmx, lk64) < 2e-6


def test_dc_flag_outside_the_reference_is_refused(state):
    case, y, ac = state
    gpu = solver.GpuSolver(case, params=dict(case.params, idc=4))
    try:
        with pytest.raises(solver.PhgError) as ei:
            gpu.elmgmrs(y, ac)
        assert ei.value.code == 1
    finally:
        gpu.close()
