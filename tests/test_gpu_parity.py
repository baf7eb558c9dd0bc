"""GPU parity of the FP64 path against the CPU oracle, through the C-ABI.

Bar: FP64 results within 1e-12 normwise relative error of the oracle (the
reference's operation order; the GPU sums element contributions per block in
element order and regroups the tangent algebraically, so agreement is to
rounding, not bit-exact), and the same Krylov iteration count.  Cases: 4^3
and 16^3 boxes, and an 8^3 box y-graded 1.1 from the walls (the stretching
of the 128^3 workload).
"""
import os

import numpy as np
import pytest

import pyoracle
from phasta_amd import config, mesh, solver

pytestmark = pytest.mark.gpu
TOL = 1e-12


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


@pytest.fixture(scope="module", params=[(4, 1.0), (16, 1.0), (8, 1.1)],
                ids=["box4", "box16", "box8_graded"])
def pair(request):
    N, grade = request.param
    case = mesh.box_case(N, mach=0.3, kspace=10, grade=grade)
    print(f"[parity] box{N} grade {grade}: nshg={case.nshg} nel={case.nel}", flush=True)
    orc = pyoracle.Oracle(case)
    gpu = solver.GpuSolver(case)
    yield case, orc, gpu
    gpu.close()


def test_elmgmrs(pair):
    case, orc, gpu = pair
    y, ac = orc.itr_predict_bc(case.yold)
    rng = np.random.default_rng(1)
    ac = ac + 1e3 * rng.standard_normal(ac.shape)  # exercise the time terms
    ac = np.asfortranarray(ac)
    r_res, r_rmes, r_bd, r_lk = orc.elmgmrs(y, ac)
    g_res, g_rmes, g_bd, g_lk = gpu.elmgmrs(y, ac)
    assert rel(g_res, r_res) < TOL
    assert rel(g_bd, r_bd) < TOL
    assert rel(g_lk, r_lk) < TOL
    assert not g_rmes.any()


def test_solgmrs(pair):
    case, orc, gpu = pair
    y, ac = orc.itr_predict_bc(case.yold)
    ref = orc.solgmrs(y, ac)
    out = gpu.solgmrs(y, ac, want_lhs=True)
    assert out["iKs"] == ref["iKs"]
    for k in ("res", "rmes", "BDiag", "Dy", "lhsK"):
        assert rel(out[k], ref[k]) < TOL, k
    assert abs(out["unorm0"] - ref["unorm0"]) <= TOL * ref["unorm0"]


def test_sparse_ap_i3lu_spsi3pre_bc3per(pair):
    case, orc, gpu = pair
    y, ac = orc.itr_predict_bc(case.yold)
    _, _, bd, lk = orc.elmgmrs(y, ac)
    rng = np.random.default_rng(2)
    p = np.asfortranarray(rng.standard_normal((case.nshg, 5)))
    ref = p.copy(order="F")
    orc.sparse_ap(lk, ref)
    assert rel(gpu.sparse_ap(lk, p), ref) < TOL
    # transpose: (q, K p) = (K^T q, p) for the operator of the GMRES
    # matvec, K p = bc3per(SparseAp(p))
    q = np.asfortranarray(rng.standard_normal((case.nshg, 5)))
    kp = gpu.bc3per(gpu.sparse_ap(lk, p))
    ktq = gpu.sparse_ap_t(lk, q)
    lhs_ip, rhs_ip = np.sum(q * kp), np.sum(ktq * p)
    assert abs(lhs_ip - rhs_ip) <= 1e-12 * (np.linalg.norm(q) * np.linalg.norm(kp))
    # i3LU factor / forward / backward
    d_ref = bd.copy(order="F")
    orc.i3lu(d_ref, None, "LU_Fact ")
    d_gpu, _ = gpu.i3lu(bd, None, "LU_Fact ")
    assert rel(d_gpu, d_ref) < TOL
    for code in ("forward ", "backward"):
        r_ref = p.copy(order="F")
        orc.i3lu(d_ref, r_ref, code)
        _, r_gpu = gpu.i3lu(d_ref, p, code)
        assert rel(r_gpu, r_ref) < TOL, code
    # Spsi3pre
    lk_ref = lk.copy(order="F")
    orc.spsi3pre(d_ref, lk_ref)
    assert rel(gpu.spsi3pre(d_ref, lk), lk_ref) < TOL
    # bc3per
    b_ref = p.copy(order="F")
    orc.bc3per(b_ref)
    assert np.array_equal(gpu.bc3per(p), b_ref)



def test_lhs_reuse_across_calls(pair):
    """LHSupd > 1 (itrdrv.f:456-457): a call with lhs = iprec = 1 forms and
    keeps the factored BDiag and scaled lhsK; the next call, lhs = iprec = 0,
    assembles the residual only and solves with the kept ones
    (solgmr.f:471,495).  Same 1e-12 bar against the oracle given the same
    persistent arrays."""
    case, orc, gpu = pair
    n, nnz = case.nshg, case.nnz_tot
    acold = np.zeros((n, 5), order="F")
    y, ac = orc.itr_predict_bc(case.yold, acold)
    bd = np.zeros((n, 5, 5), order="F")
    lk = np.zeros((25, nnz), order="F")
    ref1 = orc.solgmrs(y, ac, bd=bd, lk=lk, lhs=1, iprec=1)
    gpu.set_params(lhs=1, iprec=1)
    out1 = gpu.solgmrs(y, ac)
    assert rel(out1["Dy"], ref1["Dy"]) < TOL
    y2, ac2 = orc.itr_correct_bc(y, ac, case.yold, acold, ref1["Dy"])
    ref2 = orc.solgmrs(y2, ac2, bd=bd, lk=lk, lhs=0, iprec=0)
    try:
        gpu.set_params(lhs=0, iprec=0)
        out2 = gpu.solgmrs(y2, ac2, want_lhs=True)
    finally:
        gpu.set_params(lhs=1, iprec=1)
    assert out2["iKs"] == ref2["iKs"]
    for k in ("res", "Dy"):
        assert rel(out2[k], ref2[k]) < TOL, k
    assert rel(out2["BDiag"], bd) < TOL     # the factored BDiag of call 1
    assert rel(out2["lhsK"], lk) < TOL      # the scaled lhsK of call 1

def test_device_newton_step(pair):
    case, orc, gpu = pair
    acold = np.zeros((case.nshg, 5), order="F")
    y_ref, ac_ref, hist = orc.newton_step(case.yold, acold, newton=2)
    gpu.upload_state(case.yold, acold)
    gpu.predict()
    its = [gpu.newton_iteration() for _ in range(2)]
    y, ac, _ = gpu.download_state()
    assert [h["iKs"] for h in its] == [h["iKs"] for h in hist]
    assert rel(y, y_ref) < TOL
    assert rel(ac, ac_ref) < 1e-10  # ac = (y - yold)/dt amplifies y's rounding


def test_device_time_steps(pair):
    """Two time steps of predict / 2 Newton iterations / update on HBM."""
    case, orc, gpu = pair
    acold = np.zeros((case.nshg, 5), order="F")
    yo_ref, aco_ref, hist = orc.time_steps(case.yold, acold, steps=2, newton=2)
    gpu.upload_state(case.yold, acold)
    its = []
    for _ in range(2):
        gpu.predict()
        its += [gpu.newton_iteration() for _ in range(2)]
        gpu.update()
    yo, aco = gpu.download_old()
    assert [h["iKs"] for h in its] == [h["iKs"] for h in hist]
    assert rel(yo, yo_ref) < TOL
    assert rel(aco, aco_ref) < 1e-10


@pytest.mark.parametrize("idiff", [1, 3])
def test_flux_reconstruction(pair, idiff):
    """idiff 1 / 3 (elmgmr.f:361-391): the reconstructed diffusive flux and
    inverted lumped mass (phg_asiq: AsIq + qpbc), ElmGMRs's outputs with the
    flux divergence in the least-squares residual, and two device Newton
    iterations, against the oracle with the same flag."""
    case, orc, gpu = pair
    o = pyoracle.Oracle(case, params=dict(case.params, idiff=idiff))
    y, ac = o.itr_predict_bc(case.yold)
    rng = np.random.default_rng(6)
    ac = np.asfortranarray(ac + 1e3 * rng.standard_normal(ac.shape))
    q_ref, rm_ref = o.asiq(y)
    r_res, _, r_bd, r_lk = o.elmgmrs(y, ac)
    acold = np.zeros((case.nshg, 5), order="F")
    y_ref, _, hist = o.newton_step(case.yold, acold, newton=2)
    try:
        gpu.set_params(idiff=idiff)
        q, rm = gpu.asiq(y)
        g_res, _, g_bd, g_lk = gpu.elmgmrs(y, ac)
        gpu.upload_state(case.yold, acold)
        gpu.predict()
        its = [gpu.newton_iteration() for _ in range(2)]
        y_g, _, _ = gpu.download_state()
    finally:
        gpu.set_params(idiff=0)
    print(f"[parity] idiff {idiff}: q {rel(q, q_ref):.2e} rmass {rel(rm, rm_ref):.2e} "
          f"res {rel(g_res, r_res):.2e}", flush=True)
    assert rel(q, q_ref) < TOL and rel(rm, rm_ref) < TOL
    assert rel(g_res, r_res) < TOL
    assert rel(g_bd, r_bd) < TOL and rel(g_lk, r_lk) < TOL
    assert [h["iKs"] for h in its] == [h["iKs"] for h in hist]
    assert rel(y_g, y_ref) < TOL


def test_gpu_flux_reconstruction_against_reference_golden():
    """ElmGMRs's residual with idiff 1 and 3 against the reference's
    compiled routines (tests/golden/ref_idiff_box3.npz); idiff 2 refused."""
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                             "ref_idiff_box3.npz"))
    case = mesh.box_case(3, mach=0.5, kspace=10, jitter=0.1)
    gpu = solver.GpuSolver(case)
    try:
        for idiff in (1, 3):
            gpu.set_params(idiff=idiff)
            res, _, bd, lk = gpu.elmgmrs(g["y"], g["ac"])
            assert rel(res, g[f"res{idiff}"]) < TOL, idiff
            assert rel(bd, g["BDiag"]) < TOL and rel(lk, g["lhsK"]) < TOL
        gpu.set_params(idiff=2)
        with pytest.raises(solver.PhgError):
            gpu.elmgmrs(g["y"], g["ac"])
    finally:
        gpu.close()


def test_device_time_steps_generalized_alpha(pair):
    """Two time steps at rho_inf = 0.5 (itrSetup's almi, alfi, gami,
    itrPC.f:21-25): predictor, the tangent's time factor almi/(gami alfi),
    corrector and update against the oracle with the same factors, the
    second step from a nonzero acold.  The predictors the reference leaves
    undefined (ipred 2-4) are refused."""
    case, orc, gpu = pair
    alpha = config.time_integration(0.5)
    o = pyoracle.Oracle(case, params=dict(case.params, **alpha))
    acold = np.zeros((case.nshg, 5), order="F")
    yo_ref, aco_ref, hist = o.time_steps(case.yold, acold, steps=2, newton=2)
    try:
        gpu.set_params(**alpha)
        gpu.upload_state(case.yold, acold)
        its = []
        for _ in range(2):
            gpu.predict()
            its += [gpu.newton_iteration() for _ in range(2)]
            gpu.update()
        yo, aco = gpu.download_old()
        for ip in (2, 3, 4):
            gpu.set_params(ipred=ip)
            with pytest.raises(solver.PhgError):
                gpu.predict()
    finally:
        gpu.set_params(**config.time_integration(-1.0))
    assert [h["iKs"] for h in its] == [h["iKs"] for h in hist]
    assert rel(yo, yo_ref) < TOL
    assert rel(aco, aco_ref) < 1e-10


def test_gpu_generalized_alpha_predictor_against_reference_golden():
    """itrPredict + itrBC at rho_inf = 0.5 from a nonzero acold against the
    reference's compiled routines (tests/golden/ref_alpha_box3.npz)."""
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                             "ref_alpha_box3.npz"))
    case = mesh.box_case(3, mach=0.3, kspace=10)
    gpu = solver.GpuSolver(case, params=dict(case.params, **config.time_integration(0.5)))
    try:
        gpu.upload_state(case.yold, g["acold"])
        gpu.predict()
        y, ac, _ = gpu.download_state()
    finally:
        gpu.close()
    assert rel(y, g["y"]) < TOL
    assert rel(ac, g["ac"]) < TOL


def test_restart_file_resumes_the_device_state(tmp_path):
    """A restart file (Write_Restart layout, phastaIO) written from the
    device state after a time step and read into a fresh context continues
    exactly as the original context does: the next time step agrees bit
    for bit."""
    from phasta_amd import phastaio
    case = mesh.box_case(8, mach=0.3, kspace=10, grade=1.1)
    acold = np.zeros((case.nshg, 5), order="F")
    g1 = solver.GpuSolver(case)
    g2 = solver.GpuSolver(case)
    try:
        g1.upload_state(case.yold, acold)
        g1.predict()
        g1.newton_iteration()
        g1.update()
        yo, aco = g1.download_old()
        path = tmp_path / phastaio.restart_name(1)
        phastaio.write_restart(path, yo, aco, 1)
        q, a, step = phastaio.read_restart(path)
        assert step == 1 and np.array_equal(q, yo) and np.array_equal(a, aco)
        g2.upload_state(q, a)
        out = []
        for g in (g1, g2):
            g.predict()
            st = g.newton_iteration()
            g.update()
            out.append((st["iKs"],) + g.download_old())
    finally:
        g1.close()
        g2.close()
    assert out[0][0] == out[1][0]
    assert np.array_equal(out[0][1], out[1][1]) and np.array_equal(out[0][2], out[1][2])


def test_case_from_phasta_files_against_oracle(tmp_path):
    """A case read from PHASTA files (a geombc and a restart written in the
    phastaIO POSIX format, top wall moving at 10 m/s, 8^3 graded box),
    stepped on the GPU (two time steps of two Newton iterations) against
    the oracle on the same files: every update and the final state at
    1e-12 with the same Krylov counts; the GPU's restart file matches the
    oracle's under checkphasta's rule."""
    from phasta_amd import phastaio
    case = mesh.box_case(8, mach=0.3, kspace=10, grade=1.1)
    top = np.isclose(case.x[:, 1], 1.0)
    case.BC[top, 2] = 10.0
    geo = tmp_path / phastaio.geombc_name()
    rst = tmp_path / phastaio.restart_name(0)
    phastaio.write_geombc(geo, case)
    phastaio.write_restart(rst, case.yold, np.zeros_like(case.yold), 0)
    c2 = phastaio.case_from_geombc(geo, case.params, restart=rst)
    q, a, _ = phastaio.read_restart(rst)
    yo_ref, aco_ref, hist = pyoracle.Oracle(c2).time_steps(q, a, steps=2, newton=2)
    gpu = solver.GpuSolver(c2)
    its = []
    try:
        gpu.upload_state(q, a)
        for _ in range(2):
            gpu.predict()
            for _ in range(2):
                its.append(gpu.newton_iteration()["iKs"])
            gpu.update()
        yo, aco = gpu.download_old()
    finally:
        gpu.close()
    assert its == [h["iKs"] for h in hist]
    assert rel(yo, yo_ref) < TOL and rel(aco, aco_ref) < TOL
    assert np.allclose(yo[top, 0], 10.0)
    r_gpu, r_orc = tmp_path / "gpu" / phastaio.restart_name(2), tmp_path / "orc" / phastaio.restart_name(2)
    r_gpu.parent.mkdir()
    r_orc.parent.mkdir()
    phastaio.write_restart(r_gpu, yo, aco, 2)
    phastaio.write_restart(r_orc, yo_ref, aco_ref, 2)
    worst, ok = phastaio.compare_restarts([r_gpu], [r_orc])
    assert ok and worst <= 1e-10 * np.max(np.abs(yo_ref))


@pytest.mark.parametrize("N", [2, 3])
def test_gpu_against_reference_golden(N):
    """GPU vs the reference's own compiled routines (committed fixtures)."""
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", f"ref_box{N}.npz")
    g = np.load(gold)
    case = mesh.box_case(N, mach=0.3, kspace=10)
    gpu = solver.GpuSolver(case)
    try:
        res, rmes, bd, lk = gpu.elmgmrs(g["y"], g["ac"])
        assert rel(res, g["res"]) < TOL
        assert rel(bd, g["BDiag"]) < TOL
        assert rel(lk, g["lhsK"]) < TOL
        lu, _ = gpu.i3lu(g["BDiag"], None, "LU_Fact ")
        assert rel(lu, g["BDiag_lu"]) < TOL
        assert rel(gpu.sparse_ap(g["lhsK"], g["p"]), g["Ap"]) < TOL
        assert rel(gpu.spsi3pre(g["BDiag_lu"], g["lhsK"]), g["lhsK_scaled"]) < TOL
        # itrUpdate with the fixtures' alpha factors
        gpu.set_params(almi=0.75, alfi=0.6, gami=0.65)
        gpu.upload_state(case.yold, g["acold"])
        gpu.upload_current(g["y_corr"], g["ac_corr"])
        gpu.update()
        yo, aco = gpu.download_old()
        assert rel(yo, g["yold_upd"]) < TOL
        assert rel(aco, g["acold_upd"]) < TOL
    finally:
        gpu.close()


def test_box64_newton_iterations_against_oracle():
    """The 64^3 workload (1.57 M tets, Mach 0.5, GMRES(20)) for its three
    Newton iterations: each FP64 update on the GPU against the oracle's,
    to the same 1e-12 bar and Krylov counts, then the corrected state; the
    first mixed-precision update against the FP64 one."""
    case = mesh.named_case("box64")
    newton = config.CASES["box64"]["newton"]
    print(f"[parity] box64: nshg={case.nshg} nel={case.nel}, {newton} Newton iterations",
          flush=True)
    orc = pyoracle.Oracle(case)
    acold = np.zeros((case.nshg, 5), order="F")
    y_ref, _, hist = orc.newton_step(case.yold, acold, newton=newton)
    gpu = solver.GpuSolver(case)
    try:
        gpu.upload_state(case.yold, acold)
        gpu.predict()
        for k in range(newton):
            st = gpu.newton_iteration()
            y, _, dy = gpu.download_state()
            e = rel(dy, hist[k]["Dy"])
            print(f"[parity] box64 iteration {k + 1}: Dy rel err {e:.2e} "
                  f"(iKs {st['iKs']} vs {hist[k]['iKs']})", flush=True)
            assert st["iKs"] == hist[k]["iKs"]
            assert st["converged"] == hist[k]["converged"]
            assert e < TOL
            if k == 0:
                dy64 = dy
        assert rel(y, y_ref) < TOL
    finally:
        gpu.close()
    gpu = solver.GpuSolver(case, precision=solver.PHG_MIXED)
    try:
        gpu.upload_state(case.yold, acold)
        gpu.predict()
        stmx = gpu.newton_iteration()
        _, _, dymx = gpu.download_state()
    finally:
        gpu.close()
    emx = rel(dymx, dy64)
    print(f"[parity] box64 mixed vs FP64 first update {emx:.2e} (iKs {stmx['iKs']})", flush=True)
    assert stmx["converged"] and emx < 1e-5


def test_box64_newton_iteration_with_flux_reconstruction():
    """idiff = 1 at the 64^3 workload's size (1.57 M tets): one FP64
    Newton iteration on the GPU against the oracle's, 1e-12 and the same
    Krylov count."""
    case = mesh.named_case("box64")
    prm = dict(case.params, idiff=1)
    print(f"[parity] box64 idiff 1: nshg={case.nshg}", flush=True)
    orc = pyoracle.Oracle(case, params=prm)
    acold = np.zeros((case.nshg, 5), order="F")
    y, ac = orc.itr_predict_bc(case.yold, acold)
    ref = orc.solgmrs(y, ac)
    gpu = solver.GpuSolver(case, params=prm)
    try:
        gpu.upload_state(case.yold, acold)
        gpu.predict()
        st = gpu.newton_iteration()
        _, _, dy = gpu.download_state()
    finally:
        gpu.close()
    e = rel(dy, ref["Dy"])
    print(f"[parity] box64 idiff 1 Dy rel err {e:.2e} (iKs {st['iKs']} vs {ref['iKs']})", flush=True)
    assert st["iKs"] == ref["iKs"]
    assert e < TOL
