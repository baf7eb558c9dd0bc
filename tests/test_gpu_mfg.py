"""Matrix-free GMRES (SolMFG: solmfg.f, au1mfg.f, au2mfg.f, itrfdi.f,
elmmfg.f, itrres.f) on the GPU against the CPU oracle (bit for bit equal to
the reference's compiled routines) and the reference fixture.

Bars:
* ElmMFG's res, modified residual rmes and BDiag, and ItrRes: 1e-12, as
  every FP64 assembly (summation order differs, nothing else);
* SolMFG's finite-difference products divide residual differences by the
  interval eGMRES, which multiplies the rounding differences between two
  summation orders by 1/eGMRES.  The bars are set against the oracle's own
  sensitivity to rounding (its answer for y moved by one ulp), measured in
  each test: Au1MFG within ten times it (and 1e-5 / 1e-7 absolute), Dy at
  itrFDI's interval within five times it (about 1 %: the reference's
  interval is estimated along the unnormalised residual), Dy at 1e4 times
  the interval within 1e-5 of the oracle's and of SolGMRs' Dy;
* the interval itself (itrFDI) within 1e-5;
* with the matrix tau (itau = 10) the same bars: ElmMFG's BDiag is then
  e3bdg_nd's, mis-indexed entries included (against the oracle and the
  reference fixture ref_mfg_mtau_box3.npz);
* iDC, itau 2 and 3, lhs = 1 and the mixed-precision context are refused;
* a PHG_FP64_MATRIX_FREE context (no lhsK, as itrdrv.f:163-185 allocates
  for impl 2x) runs the same kernels as a PHG_FP64 one: its SolMFG results
  are bitwise equal, it holds at least the matrix's bytes less, and the
  entry points that need lhsK refuse it.
"""
import ctypes
import os

import numpy as np
import pytest

import pyoracle
from phasta_amd import mesh, solver

pytestmark = pytest.mark.gpu
TOL = 1e-12
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def mfg_params(case, **over):
    """The flags itrdrv sets for SolMFG (lhs = 0, iprec = 1; itrdrv.f:494-495)."""
    prm = dict(case.params, lhs=0, iprec=1)
    prm.update(over)
    return prm


@pytest.fixture(scope="module", params=[(4, 1.0), (16, 1.0), (8, 1.1)],
                ids=["box4", "box16", "box8_graded"])
def state(request):
    """A jittered box at a perturbed predicted state, with itrBC applied
    again: SolMFG's differences re-impose the Dirichlet values on every
    perturbed point (au1mfg.f:70), so a base state that broke them would
    add a constant of the size of the residual over eGMRES to every
    product (the solver's states always come from itrBC)."""
    N, grade = request.param
    case = mesh.box_case(N, mach=0.5, kspace=10, jitter=0.1, grade=grade)
    print(f"[mfg] box{N} grade {grade}: nshg={case.nshg}", flush=True)
    rng = np.random.default_rng(N)
    o = pyoracle.Oracle(case)
    y, ac = o.itr_predict_bc(case.yold)
    y = np.asfortranarray(y * (1 + 0.05 * rng.standard_normal(y.shape)))
    ac = np.asfortranarray(ac + 1e3 * rng.standard_normal(ac.shape))
    assert pyoracle.lib().orc_itr_bc(ctypes.byref(o.m), pyoracle._p(y), pyoracle._p(ac)) == 0
    return case, y, ac


@pytest.mark.parametrize("idiff,itau", [(0, 0), (1, 0), (0, 10), (1, 10)])
def test_elmmfg_against_oracle(state, idiff, itau):
    case, y, ac = state
    prm = mfg_params(case, idiff=idiff, itau=itau)
    ref = pyoracle.Oracle(case, params=prm).elmmfg(y, ac)
    gpu = solver.GpuSolver(case, params=prm)
    try:
        out = gpu.elmmfg(y, ac)
    finally:
        gpu.close()
    for name, a, b in zip(("res", "rmes", "BDiag"), out, ref):
        e = rel(a, b)
        print(f"[mfg] idiff {idiff} itau {itau} ElmMFG {name}: {e:.2e}", flush=True)
        assert e < TOL, name


def test_mfg_pieces_against_reference_golden():
    g = np.load(os.path.join(GOLD, "ref_mfg_box3.npz"))
    case = mesh.box_case(3, mach=0.5, kspace=10, jitter=0.1)
    y, ac = np.asfortranarray(g["y"]), np.asfortranarray(g["ac"])
    gpu = solver.GpuSolver(case, params=mfg_params(case))
    try:
        res, rmes, bd = gpu.elmmfg(y, ac)
        assert rel(res, g["res"]) < TOL
        assert rel(rmes, g["rmes"]) < TOL
        assert rel(bd, g["BDiag"]) < TOL
        assert rel(gpu.itrres(g["yp"], y, ac), g["rmes_p"]) < TOL
        assert rel(gpu.itrres(g["yp"], y, ac, iabres=True), g["rmes_abs"]) < TOL
        _, ypre = gpu.i3lu(g["BDiag_lu"], np.asfortranarray(g["y"][:, [3, 0, 1, 2, 4]]),
                           "product ")
        assert rel(ypre, g["ypre"]) < TOL
        gpu.set_params(idiff=1)
        res1, rmes1, bd1 = gpu.elmmfg(y, ac)
        assert rel(res1, g["res_idiff"]) < TOL
        assert rel(rmes1, g["rmes_idiff"]) < TOL
        assert rel(bd1, g["BDiag_idiff"]) < TOL
    finally:
        gpu.close()


def test_mfg_matrix_tau_pieces_against_reference_golden():
    g = np.load(os.path.join(GOLD, "ref_mfg_mtau_box3.npz"))
    case = mesh.box_case(3, mach=0.5, kspace=10, jitter=0.1)
    y, ac = np.asfortranarray(g["y"]), np.asfortranarray(g["ac"])
    gpu = solver.GpuSolver(case, params=mfg_params(case, itau=10))
    try:
        res, rmes, bd = gpu.elmmfg(y, ac)
        assert rel(res, g["res"]) < TOL
        assert rel(rmes, g["rmes"]) < TOL
        assert rel(bd, g["BDiag"]) < TOL
        assert rel(gpu.itrres(g["yp"], y, ac), g["rmes_p"]) < TOL
    finally:
        gpu.close()


@pytest.mark.parametrize("itau", [0, 10])
def test_itrres_against_oracle_accumulates(state, itau):
    case, y, ac = state
    prm = mfg_params(case, itau=itau)
    rng = np.random.default_rng(5)
    yp = np.asfortranarray(y * (1 + 0.02 * rng.standard_normal(y.shape)))
    acc = np.asfortranarray(rng.standard_normal(y.shape))
    o = pyoracle.Oracle(case, params=prm)
    gpu = solver.GpuSolver(case, params=prm)
    try:
        for iabres in (False, True):
            ref = o.itrres(yp, y, ac, rmes=acc.copy(order="F"), iabres=iabres)
            assert rel(gpu.itrres(yp, y, ac, rmes=acc, iabres=iabres), ref) < TOL, iabres
    finally:
        gpu.close()


def ulp_twin(case, y, ac, seed=99):
    """y moved by about one ulp (relative 1e-15, Dirichlet values kept): the
    oracle's own answer on it measures how far rounding alone moves a
    finite-difference result."""
    o = pyoracle.Oracle(case)
    rng = np.random.default_rng(seed)
    y2 = np.asfortranarray(y * (1 + 1e-15 * rng.standard_normal(y.shape)))
    ac2 = ac.copy(order="F")
    assert pyoracle.lib().orc_itr_bc(ctypes.byref(o.m), pyoracle._p(y2), pyoracle._p(ac2)) == 0
    return y2


def base_point(o, y, ac):
    """SolMFG's preconditioned base point: factored BDiag, forward-reduced
    res and rmes, ypre = U new2old(y), and the interval of itrFDI."""
    res, rmes, bd = o.elmmfg(y, ac)
    lu = bd.copy(order="F")
    o.i3lu(lu, None, "LU_Fact ")
    r, m = res.copy(order="F"), rmes.copy(order="F")
    o.i3lu(lu, r, "forward ")
    o.i3lu(lu, m, "forward ")
    ypre = o.yshuffle(y, "new2old ")
    o.i3lu(lu, ypre, "product ")
    return lu, r, m, ypre, o.itrfdi(ypre, y, ac, m, r, lu)


@pytest.mark.parametrize("itau", [0, 10])
@pytest.mark.parametrize("scale", [1.0, 100.0], ids=["itrfdi_interval", "interval_x100"])
def test_au1mfg_against_oracle(state, scale, itau):
    """The one-sided product of a random direction (zero where res is: the
    constrained equations and periodic slaves): within ten times the
    oracle's own rounding sensitivity, measured 1e-8..1e-7 at itrFDI's
    interval and 1e-10..1e-9 at a hundred times it."""
    case, y, ac = state
    prm = mfg_params(case, itau=itau)
    o = pyoracle.Oracle(case, params=prm)
    lu, r, m, ypre, eg = base_point(o, y, ac)
    rng = np.random.default_rng(11)
    u = np.asfortranarray(rng.standard_normal(y.shape))
    u[r == 0.0] = 0.0
    ref = o.au1mfg(ypre, y, ac, m, lu, scale * eg, u)
    y2 = ulp_twin(case, y, ac)
    o2 = pyoracle.Oracle(case, params=prm)
    lu2, _, m2, ypre2, _ = base_point(o2, y2, ac)
    noise = rel(o2.au1mfg(ypre2, y2, ac, m2, lu2, scale * eg, u), ref)
    gpu = solver.GpuSolver(case, params=prm)
    try:
        out = gpu.au1mfg(ypre, y, ac, m, lu, scale * eg, u)
    finally:
        gpu.close()
    e = rel(out, ref)
    print(f"[mfg] Au1MFG x{scale:g}: GPU vs oracle {e:.2e}, oracle rounding {noise:.2e}",
          flush=True)
    assert e < 10 * noise + 1e-12
    assert e < (1e-5 if scale == 1.0 else 1e-7)


def test_au1mfg_against_reference_golden():
    g = np.load(os.path.join(GOLD, "ref_mfg_box3.npz"))
    case = mesh.box_case(3, mach=0.5, kspace=10, jitter=0.1)
    gpu = solver.GpuSolver(case, params=mfg_params(case))
    try:
        out = gpu.au1mfg(g["ypre"], g["y"], g["ac"], g["rmes_f"], g["BDiag_lu"],
                         float(g["eGMRES"]), g["u"])
    finally:
        gpu.close()
    e = rel(out, g["au1"])
    print(f"[mfg] Au1MFG vs the reference fixture: {e:.2e}", flush=True)
    assert e < 1e-6


@pytest.mark.parametrize("itau", [0, 10])
@pytest.mark.parametrize("fdi", [False, True], ids=["given_interval", "itrfdi"])
def test_solmfg_against_oracle(state, fdi, itau):
    """At itrFDI's interval, which the reference estimates along the
    unnormalised residual (solmfg.f:165-168) and applies to unit Krylov
    vectors, the products of Krylov directions keep two or three digits and
    the oracle's own Dy moves by about 1 % when y moves by one ulp: the GPU's
    Dy is held to five times that measured sensitivity.  The interval
    itself is a ratio of second differences taken at sqrt(epsM): the
    oracle's moves by 2e-7..1.3e-5 over three one-ulp moves of y (the upper
    end with the matrix tau on the 4^3 box); the GPU's is held to five times
    the largest of those and to 1e-4."""
    case, y, ac = state
    prm = mfg_params(case, itau=itau)
    ref = pyoracle.Oracle(case, params=prm).solmfg(y, ac, fdi=True)
    twin = pyoracle.Oracle(case, params=prm).solmfg(ulp_twin(case, y, ac), ac, fdi=False,
                                                   egmres=ref["eGMRES"])
    noise = rel(twin["Dy"], ref["Dy"])
    eg_noise = max(abs(pyoracle.Oracle(case, params=prm).solmfg(
        ulp_twin(case, y, ac, seed), ac, fdi=True)["eGMRES"] - ref["eGMRES"]) / ref["eGMRES"]
        for seed in (99, 7, 3))
    gpu = solver.GpuSolver(case, params=prm)
    try:
        out = gpu.solmfg(y, ac, egmres=0.0 if fdi else ref["eGMRES"], fdi=fdi)
    finally:
        gpu.close()
    e_eg = abs(out["eGMRES"] - ref["eGMRES"]) / ref["eGMRES"]
    e_dy = rel(out["Dy"], ref["Dy"])
    print(f"[mfg] SolMFG fdi={fdi}: eGMRES {out['eGMRES']:.6e} ({e_eg:.1e}), iKs "
          f"{out['iKs']}/{ref['iKs']}, Dy {e_dy:.2e} (oracle rounding {noise:.2e}, interval "
          f"{eg_noise:.1e})", flush=True)
    assert e_eg < 5 * eg_noise + 1e-9 and e_eg < 1e-4
    assert abs(out["iKs"] - ref["iKs"]) <= 1
    for k in ("res", "rmes", "BDiag"):
        assert rel(out[k], ref[k]) < TOL, k
    assert e_dy < 5 * noise + 1e-10


def test_solmfg_at_a_wide_interval_against_oracle_and_solgmrs(state):
    """At 1e4 times itrFDI's interval the products keep about eight digits:
    the GPU's Dy agrees with the oracle's to 1e-5 (the oracle's own rounding
    sensitivity there is about 1e-6), and with SolGMRs' Dy on the same state
    to 1e-5, the matrix-free solve reproducing the assembled one."""
    case, y, ac = state
    prm = mfg_params(case, Kspace=10)
    eg = 1e4 * pyoracle.Oracle(case, params=prm).solmfg(y, ac, fdi=True)["eGMRES"]
    ref = pyoracle.Oracle(case, params=prm).solmfg(y, ac, fdi=False, egmres=eg)
    gpu = solver.GpuSolver(case, params=prm)
    try:
        out = gpu.solmfg(y, ac, egmres=eg, fdi=False)
    finally:
        gpu.close()
    gm = solver.GpuSolver(case)
    try:
        g = gm.solgmrs(y, ac)
    finally:
        gm.close()
    e, eg_ = rel(out["Dy"], ref["Dy"]), rel(out["Dy"], g["Dy"])
    print(f"[mfg] SolMFG at 1e4 x eGMRES: vs oracle {e:.2e}, vs SolGMRs {eg_:.2e}", flush=True)
    assert out["iKs"] == ref["iKs"]
    assert e < 1e-5 and eg_ < 1e-5


def test_solmfg_matrix_tau_at_a_wide_interval_against_oracle(state):
    """The matrix tau at 1e4 times itrFDI's interval: Dy within 1e-5 of the
    oracle's, the same Krylov count."""
    case, y, ac = state
    prm = mfg_params(case, Kspace=10, itau=10)
    eg = 1e4 * pyoracle.Oracle(case, params=prm).solmfg(y, ac, fdi=True)["eGMRES"]
    ref = pyoracle.Oracle(case, params=prm).solmfg(y, ac, fdi=False, egmres=eg)
    gpu = solver.GpuSolver(case, params=prm)
    try:
        out = gpu.solmfg(y, ac, egmres=eg, fdi=False)
    finally:
        gpu.close()
    e = rel(out["Dy"], ref["Dy"])
    print(f"[mfg] SolMFG itau 10 at 1e4 x eGMRES: vs oracle {e:.2e}", flush=True)
    assert out["iKs"] == ref["iKs"]
    assert e < 1e-5


def test_restart_cycles_against_oracle():
    """nGMRES = 3 with Kspace 2: later cycles restart from Au2MFG's two-sided
    residual res - A Dy (au2mfg.f), whose interval epsM^(2/3) / ||Dy|| does
    not follow eGMRES; Dy within five times the oracle's own rounding
    sensitivity (measured 1e-4 .. 1e-3) and within 1e-2."""
    case = mesh.box_case(4, mach=0.5, kspace=2, jitter=0.1)
    prm = mfg_params(case, nGMRES=3)
    o = pyoracle.Oracle(case, params=prm)
    y, ac = o.itr_predict_bc(case.yold)
    eg = 1e4 * o.solmfg(y, ac, fdi=True)["eGMRES"]
    ref = pyoracle.Oracle(case, params=prm).solmfg(y, ac, fdi=False, egmres=eg)
    twin = pyoracle.Oracle(case, params=prm).solmfg(ulp_twin(case, y, ac), ac, fdi=False,
                                                   egmres=eg)
    noise = rel(twin["Dy"], ref["Dy"])
    gpu = solver.GpuSolver(case, params=prm)
    try:
        out = gpu.solmfg(y, ac, egmres=eg, fdi=False)
    finally:
        gpu.close()
    e = rel(out["Dy"], ref["Dy"])
    print(f"[mfg] restarts: lGMRES {out['lGMRES']}/{ref['lGMRES']} Dy {e:.2e} "
          f"(oracle rounding {noise:.2e})", flush=True)
    assert ref["lGMRES"] >= 1
    assert out["lGMRES"] == ref["lGMRES"] and out["iKs"] == ref["iKs"]
    assert e < 5 * noise + 1e-10 and e < 1e-2


def test_resident_newton_iterations_against_oracle():
    """Two time steps of two matrix-free Newton iterations on the resident
    state (itrdrv.f:398-652 with SolMFG); the interval is recomputed at the
    first iteration of each step, as at istep = 0 mod 20.  At the
    reference's interval: Dy within 5e-2 (about five times the rounding
    sensitivity measured above), y after itrCorrect by the same amount."""
    case = mesh.box_case(8, mach=0.5, kspace=10, jitter=0.1, grade=1.1)
    prm = mfg_params(case)
    o = pyoracle.Oracle(case, params=prm)
    yold, acold = case.yold, np.zeros((case.nshg, 5), order="F")
    gpu = solver.GpuSolver(case, params=prm)
    try:
        gpu.upload_state(yold, acold)
        for step in range(2):
            gpu.predict()
            y, ac = o.itr_predict_bc(yold, acold)
            eg = 0.0
            for it in range(2):
                ref = o.solmfg(y, ac, fdi=it == 0, egmres=eg)
                out = gpu.newton_iteration_mfg(egmres=eg, fdi=it == 0)
                eg = ref["eGMRES"]
                y, ac = o.itr_correct_bc(y, ac, yold, acold, ref["Dy"])
                yg, acg, dyg = gpu.download_state()
                e = rel(dyg, ref["Dy"])
                print(f"[mfg] step {step} iter {it}: iKs {out['iKs']}/{ref['iKs']} Dy {e:.2e} "
                      f"y {rel(yg, y):.2e}", flush=True)
                assert abs(out["iKs"] - ref["iKs"]) <= 1
                assert e < 5e-2
                # itrCorrect moves y by Dy: the same gap, in y's norm
                assert np.linalg.norm(yg - y) <= 5e-2 * np.linalg.norm(ref["Dy"])
                # continue both from the oracle's iterate
                gpu.upload_current(y, ac)
            gpu.update()
            yold, acold = o.itr_update_bc(yold, acold, y, ac)
    finally:
        gpu.close()


@pytest.mark.parametrize("over", [{"idc": 1}, {"itau": 2}, {"lhs": 1}], ids=str)
def test_flags_outside_the_matrix_free_path_are_refused(state, over):
    case, y, ac = state
    gpu = solver.GpuSolver(case, params=mfg_params(case, **over))
    try:
        with pytest.raises(solver.PhgError) as ei:
            gpu.solmfg(y, ac)
        assert ei.value.code == 1
    finally:
        gpu.close()


def test_mixed_precision_context_is_refused(state):
    case, y, ac = state
    gpu = solver.GpuSolver(case, params=mfg_params(case), precision=solver.PHG_MIXED)
    try:
        with pytest.raises(solver.PhgError) as ei:
            gpu.elmmfg(y, ac)
        assert ei.value.code == 1
    finally:
        gpu.close()


def test_matrix_free_context_equals_the_fp64_context(state):
    """Same kernels, same order: res, rmes, BDiag, eGMRES and Dy bitwise
    equal; two resident Newton iterations too.  The context holds no lhsK
    (26 doubles per CSR entry in the element-block layout, 25 in the matvec
    layout) and no off-diagonal owner lanes."""
    case, y, ac = state
    prm = mfg_params(case)
    outs, infos = [], []
    for prec in (solver.PHG_FP64, solver.PHG_FP64_MATRIX_FREE):
        gpu = solver.GpuSolver(case, params=prm, precision=prec)
        try:
            o = gpu.solmfg(y, ac, egmres=0.0, fdi=True)
            gpu.upload_state(case.yold, np.zeros((case.nshg, 5), order="F"))
            gpu.predict()
            its = [gpu.newton_iteration_mfg(egmres=0.0, fdi=True)["iKs"],
                   gpu.newton_iteration_mfg(fdi=False)["iKs"]]
            o["resident"] = gpu.download_state()
            o["its"] = its
            outs.append(o)
            infos.append(gpu.info()["dev_bytes"])
        finally:
            gpu.close()
    a, b = outs
    for k in ("res", "rmes", "BDiag", "Dy"):
        assert np.array_equal(a[k], b[k]), k
    assert a["eGMRES"] == b["eGMRES"] and a["iKs"] == b["iKs"] and a["its"] == b["its"]
    for u, v in zip(a["resident"], b["resident"]):
        assert np.array_equal(u, v)
    saved = infos[0] - infos[1]
    print(f"[mfg] device bytes {infos[0]} -> {infos[1]} ({saved} less)", flush=True)
    assert saved >= 8 * 51 * case.nnz_tot


def test_matrix_free_context_refuses_matrix_entry_points(state):
    case, y, ac = state
    n, nnz = case.nshg, case.nnz_tot
    gpu = solver.GpuSolver(case, params=mfg_params(case), precision=solver.PHG_FP64_MATRIX_FREE)
    try:
        calls = [lambda: gpu.solgmrs(y, ac),
                 lambda: gpu.sparse_ap(np.zeros((25, nnz), order="F"), np.zeros((n, 5), order="F")),
                 lambda: gpu.spsi3pre(np.zeros((n, 5, 5), order="F"), np.zeros((25, nnz), order="F"))]
        for call in calls:
            with pytest.raises(solver.PhgError) as ei:
                call()
            assert ei.value.code == 1
            assert "MATRIX_FREE" in str(ei.value)
        gpu.set_params(lhs=1)
        with pytest.raises(solver.PhgError) as ei:
            gpu.elmgmrs(y, ac)
        assert ei.value.code == 1
        # the matrix-free path still runs after the refusals
        gpu.set_params(lhs=0)
        assert gpu.solmfg(y, ac)["iKs"] >= 1
    finally:
        gpu.close()
