"""The solver's statistics and abort rules on the GPU (rstat.f:95-116,
rstatCheck.f:36-79, tnanq.f:10-28), and the FP64 tangent product the
mixed path refines with, against the CPU oracle.

* totres: ||res||/nshg and ||rmes||/nshg as rstat forms them (single
  precision nshg), against the oracle's, 1e-12.
* tnanq: a NaN in the state reaches res and BDiag; the solve returns
  PHG_ENAN (the reference's error() abort).
* rstatCheck: from the second Newton iteration of a time step (iter > 1)
  a residual more than 1e4 times the last one returns PHG_EDIVERGE; the
  first iteration of a step is not tested.  A resting uniform state
  (residual at rounding level) followed by the real state provides the
  jump.
* phg_tangent_ap: bc3per(A p(iper)) with the FP64 tangent formed element
  by element equals the oracle's SparseAp + bc3per on its assembled lhsK
  (after bc3LHS), 1e-12, in FP64 and in mixed-precision contexts.
"""
import numpy as np
import pytest

import pyoracle
from phasta_amd import mesh, solver

pytestmark = pytest.mark.gpu
TOL = 1e-12


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module", params=[(4, 1.0), (16, 1.0), (8, 1.1)],
                ids=["box4", "box16", "box8_graded"])
def case(request):
    N, grade = request.param
    c = mesh.box_case(N, mach=0.3, kspace=10, grade=grade)
    print(f"[rstat] box{N} grade {grade}: nshg={c.nshg}", flush=True)
    return c


def rest_state(case):
    y = np.zeros((case.nshg, 5), order="F")
    y[:, 3], y[:, 4] = 1.0e5, 300.0
    return y


def test_totres_matches_oracle(case):
    orc = pyoracle.Oracle(case)
    y, ac = orc.itr_predict_bc(case.yold)
    ref = orc.solgmrs(y, ac, iter=1)
    gpu = solver.GpuSolver(case)
    try:
        out = gpu.solgmrs(y, ac, iter=1)
    finally:
        gpu.close()
    for k in (0, 1):
        assert abs(out["totres"][k] - ref["totres"][k]) <= TOL * ref["totres"][k], k
    assert out["reslast"] == pytest.approx(orc.reslast, rel=TOL)


def test_nan_in_the_state_returns_enan(case):
    orc = pyoracle.Oracle(case)
    y, ac = orc.itr_predict_bc(case.yold)
    bad = y.copy(order="F")
    bad[case.nshg // 2, 3] = np.nan
    gpu = solver.GpuSolver(case)
    try:
        with pytest.raises(solver.PhgError) as ei:
            gpu.solgmrs(bad, ac, iter=1)
        assert ei.value.code == 10
        # the context stays usable: the next call is the oracle's solve
        # (box16 at K = 10 stops unconverged in the reference too)
        ref = orc.solgmrs(y, ac, iter=1)
        out = gpu.solgmrs(y, ac, iter=1)
        assert out["iKs"] == ref["iKs"] and out["converged"] == ref["converged"]
        assert rel(out["Dy"], ref["Dy"]) < TOL
    finally:
        gpu.close()


def test_divergence_abort_from_the_second_newton_iteration(case):
    orc = pyoracle.Oracle(case)
    y0, ac0 = orc.itr_predict_bc(rest_state(case))
    y, ac = orc.itr_predict_bc(case.yold)
    for prec in (solver.PHG_FP64, solver.PHG_MIXED):
        gpu = solver.GpuSolver(case, precision=prec)
        try:
            quiet = gpu.solgmrs(y0, ac0, iter=1)["reslast"]
            with pytest.raises(solver.PhgError) as ei:
                gpu.solgmrs(y, ac, iter=2)
            assert ei.value.code == 11
            gpu.solgmrs(y0, ac0, iter=1)
            out = gpu.solgmrs(y, ac, iter=1)  # first iteration of a step: no test
            assert out["reslast"] > 1e4 * quiet
        finally:
            gpu.close()


def test_resident_stepping_resets_the_iteration_count(case):
    """phg_predict starts a time step (iter = 0): its first Newton
    iteration is not tested against the previous step's residual."""
    orc = pyoracle.Oracle(case)
    y, ac = orc.itr_predict_bc(case.yold)
    ref = orc.solgmrs(y, ac, iter=1)
    gpu = solver.GpuSolver(case)
    try:
        gpu.upload_state(rest_state(case))
        gpu.predict()
        st = gpu.newton_iteration()
        quiet = st["reslast"]
        gpu.upload_state(case.yold)
        gpu.predict()
        st = gpu.newton_iteration()
        # not aborted, and the same solve as the oracle's first iteration
        assert st["reslast"] > 1e4 * quiet
        assert st["iKs"] == ref["iKs"] and st["converged"] == ref["converged"]
    finally:
        gpu.close()


@pytest.mark.parametrize("prec", [solver.PHG_FP64, solver.PHG_MIXED], ids=["fp64", "mixed"])
def test_tangent_product_equals_oracle_sparse_ap(case, prec):
    orc = pyoracle.Oracle(case)
    y, ac = orc.itr_predict_bc(case.yold)
    rng = np.random.default_rng(5)
    y = np.asfortranarray(y * (1 + 0.01 * rng.standard_normal(y.shape)))
    _, _, _, lk = orc.elmgmrs(y, ac)
    gpu = solver.GpuSolver(case, precision=prec)
    try:
        for seed in (1, 2):
            p = np.asfortranarray(np.random.default_rng(seed).standard_normal((case.nshg, 5)))
            q = p.copy(order="F")
            orc.sparse_ap(lk, q)
            orc.bc3per(q)
            g = gpu.tangent_ap(y, p)
            e = rel(g, q)
            print(f"[rstat] tangent product rel err {e:.2e}", flush=True)
            assert e < TOL
    finally:
        gpu.close()
