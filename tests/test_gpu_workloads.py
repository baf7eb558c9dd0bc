"""The workloads at their full sizes on the GPU.

* 128^3 y-graded box (12.6 M tets, 2.1 M nodes): one FP64 Newton
  iteration against the oracle (res, BDiag, the scaled lhsK, Dy at 1e-12,
  the same Krylov count), and the mixed-precision form against FP64: the
  residual (double in both, 1e-13), the float tangent and BDiag (2e-6),
  and its update checked on the FP64 system the way SolGMRs' tolerance
  is defined (solgmr.f:499-519): ||L^-1 (b - A Dy)|| <= etol ||L^-1 b||
  with A the FP64 tangent and L the FP64 block-diagonal factors.  The
  distances of both updates from a tightly converged FP64 solve are
  printed beside their distance from each other.
* 256^3 box with a 1.08 boundary layer from a first cell of h/100
  (100.7 M tets, 17 M nodes): five time steps of three Newton iterations
  with the mixed form and the learned preconditioner (block-diagonal
  GMRES as its fallback); every Newton iteration must meet the tolerance
  on the FP64 system (the refinement residual is formed with the FP64
  tangent) and the state must stay finite.
"""
import os
import time

import numpy as np
import pytest

import pyoracle
from phasta_amd import config, mesh, solver

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
WEIGHTS = os.path.join(ROOT, "phasta_amd", "weights", "mpnn_box64.npz")
TOL = 1e-12


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def box128():
    case = mesh.named_case("box128")
    print(f"[workload] box128: nshg={case.nshg} nel={case.nel}", flush=True)
    orc = pyoracle.Oracle(case)
    acold = np.zeros((case.nshg, 5), order="F")
    y, ac = orc.itr_predict_bc(case.yold, acold)
    t0 = time.perf_counter()
    ref = orc.solgmrs(y, ac, iter=1)
    print(f"[workload] box128 oracle Newton iteration {time.perf_counter() - t0:.1f} s "
          f"(iKs {ref['iKs']})", flush=True)
    return case, orc, y, ac, ref


@pytest.mark.timeout(600)
def test_box128_fp64_newton_iteration_against_oracle(box128):
    case, orc, y, ac, ref = box128
    gpu = solver.GpuSolver(case)
    try:
        gpu.diagnostics(orth=True)
        out = gpu.solgmrs(y, ac, want_lhs=True, iter=1)
    finally:
        gpu.close()
    print(f"[workload] box128 FP64 (MGS) Krylov basis orthogonality loss "
          f"{out['orth_loss']:.2e}", flush=True)
    errs = {k: rel(out[k], ref[k]) for k in ("res", "rmes", "BDiag", "lhsK", "Dy")}
    print(f"[workload] box128 FP64 vs oracle: " +
          ", ".join(f"{k} {v:.2e}" for k, v in errs.items()) +
          f"; iKs {out['iKs']} vs {ref['iKs']}", flush=True)
    assert out["iKs"] == ref["iKs"]
    for k, v in errs.items():
        assert v < TOL, k


@pytest.mark.timeout(600)
def test_box128_mixed_against_fp64(box128):
    case, orc, y, ac, ref = box128
    g64 = solver.GpuSolver(case)
    gmx = solver.GpuSolver(case, precision=solver.PHG_MIXED)
    try:
        r64, _, bd64, lk64 = g64.elmgmrs(y, ac)
        rmx, _, bdmx, lkmx = gmx.elmgmrs(y, ac)
        e_res, e_bd, e_lk = rel(rmx, r64), rel(bdmx, bd64), rel(lkmx, lk64)
        del lk64, lkmx
        gmx.diagnostics(orth=True)
        out = gmx.solgmrs(y, ac, iter=1)
        dy = out["Dy"]
        # the FP64 system: A from the FP64 tangent, L from the oracle's
        # factored FP64 BDiag, b the raw residual
        adv = g64.tangent_ap(y, dy)
        r = np.asfortranarray(ref["rmes"] - adv)
        orc.i3lu(ref["BDiag"], r, "forward ")
        epsnrm = case.params["etol"] * np.linalg.norm(ref["res"])
        ratio = np.linalg.norm(r) / epsnrm
        e_dy = rel(dy, ref["Dy"])
        # a tightly converged FP64 solve of the same system for scale
        tight = solver.GpuSolver(case, params=dict(case.params, Kspace=60, etol=1e-8))
        try:
            t_out = tight.solgmrs(y, ac, iter=1)
        finally:
            tight.close()
        print(f"[workload] box128 mixed vs FP64: res {e_res:.2e}, BDiag {e_bd:.2e}, lhsK "
              f"{e_lk:.2e}, Dy {e_dy:.2e} (iKs {out['iKs']} vs {ref['iKs']}, cycles "
              f"{out['refine_cycles']}); FP64-system residual / epsnrm {ratio:.4f}; from the "
              f"tight solve ({t_out['iKs']} iterations, |e|/unorm0 "
              f"{t_out['echeck'] / t_out['unorm0']:.1e}): mixed "
              f"{rel(dy, t_out['Dy']):.2e}, FP64 {rel(ref['Dy'], t_out['Dy']):.2e}", flush=True)
    finally:
        g64.close()
        gmx.close()
    assert e_res < 1e-13
    assert e_bd < 2e-6 and e_lk < 2e-6
    assert out["converged"]
    assert ratio <= 1.0 + 1e-5
    assert e_dy < 1e-5
    # classical Gram-Schmidt keeps the basis orthogonal to this level here
    print(f"[workload] box128 mixed Krylov basis orthogonality loss {out['orth_loss']:.2e}",
          flush=True)
    assert out["orth_loss"] < 1e-8


@pytest.mark.timeout(900)
def test_box256_time_steps_mixed_learned():
    spec = config.CASES["box256"]
    t0 = time.perf_counter()
    case = mesh.named_case("box256")
    print(f"[workload] box256: nshg={case.nshg} nel={case.nel} "
          f"(mesh {time.perf_counter() - t0:.1f} s)", flush=True)
    t0 = time.perf_counter()
    gpu = solver.GpuSolver(case, precision=solver.PHG_MIXED)
    try:
        w = np.load(WEIGHTS)
        gpu.mpnn_init(w["dscale"])
        gpu.mpnn_set_weights(w["weights"])
        gpu.mpnn_enable(use_pc=1)
        print(f"[workload] box256 set-up {time.perf_counter() - t0:.1f} s, device bytes "
              f"{gpu.info()['dev_bytes'] / 1e9:.1f} GB", flush=True)
        gpu.upload_state(case.yold)
        for step in range(spec["steps"]):
            gpu.predict()
            for k in range(spec["newton"]):
                t1 = time.perf_counter()
                st = gpu.newton_iteration()
                print(f"[workload] box256 step {step + 1} iteration {k + 1}: iKs {st['iKs']}, "
                      f"cycles {st['refine_cycles']}, pc_fallback {st['pc_fallback']}, "
                      f"converged {st['converged']}, {time.perf_counter() - t1:.3f} s",
                      flush=True)
                assert st["converged"]
            gpu.update()
        yo, _ = gpu.download_old()
        assert np.isfinite(yo).all()
    finally:
        gpu.close()
