"""Writes reference-compiled golden vectors for the hot path.

Runs the reference's own Fortran routines (oracle/_ref/libphref.so, built by
oracle/ref/Makefile from the reference sources) on tiny box cases and stores
inputs and outputs as .npz fixtures next to this script.  Run from the repo
root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from phasta_amd import mesh  # noqa: E402
import pyref  # noqa: E402


UPDATE_ALPHA = dict(almi=0.75, alfi=0.6, gami=0.65)


def make(N, seed):
    case = mesh.box_case(N, mach=0.3, kspace=10)
    ref = pyref.Reference(case)
    y, ac = ref.predict_bc(case.yold)
    rng = np.random.default_rng(seed)
    ac = np.asfortranarray(ac + 1e3 * rng.standard_normal(ac.shape))
    res, rmes, bd, lk = ref.elmgmrs(y, ac)
    bd_lu, _ = ref.i3lu(bd, None, "LU_Fact ")
    p = np.asfortranarray(rng.standard_normal((case.nshg, 5)))
    ap = ref.sparse_ap(lk, p)
    lk_s = ref.spsi3pre(bd_lu, lk)
    dy = np.asfortranarray(1e-3 * rng.standard_normal((case.nshg, 5)))
    y1, ac1 = ref.correct_bc(y, ac, case.yold, np.zeros_like(ac), dy)
    # itrUpdate with distinct alpha factors, so that a swapped almi/alfi shows
    acold = np.asfortranarray(1e2 * rng.standard_normal((case.nshg, 5)))
    upd = dict(case.params, **UPDATE_ALPHA)
    yold1, acold1 = pyref.Reference(case, params=upd).update_bc(case.yold, acold, y1, ac1)
    np.savez_compressed(os.path.join(HERE, f"ref_box{N}.npz"), N=N, y=y, ac=ac, res=res,
                        BDiag=bd, lhsK=lk, BDiag_lu=bd_lu, p=p, Ap=ap, lhsK_scaled=lk_s,
                        dy=dy, y_corr=y1, ac_corr=ac1, acold=acold, yold_upd=yold1,
                        acold_upd=acold1)


# spectral radii for itrSetup, and the one the time-step fixtures use
RHOINF = (-1.0, 0.0, 0.25, 0.5, 0.75, 1.0, 1.5)
RHOINF_STEP = 0.5


def make_alpha(N, seed):
    """itrSetup's factors (almi, alfi, gami) and settled predictor for each
    RHOINF, asked for ipred 1 and 4, with y after its ipred = 4 start-up
    shift; then itrPredict + itrBC and itrCorrect + itrBC at RHOINF_STEP
    from a nonzero acold."""
    case = mesh.box_case(N, mach=0.3, kspace=10)
    rng = np.random.default_rng(seed)
    n = case.nshg
    acold = np.asfortranarray(1e3 * rng.standard_normal((n, 5)))
    ref = pyref.Reference(case)
    fac, ipo, shift = [], [], []
    for r in RHOINF:
        for ip in (1, 4):
            f, ipout, ys = ref.itr_setup(r, ip, case.yold, acold)
            fac.append(f)
            ipo.append(ipout)
            shift.append(np.abs(ys - case.yold).max())
    f = fac[2 * RHOINF.index(RHOINF_STEP)]
    prm = dict(case.params, almi=f[0], alfi=f[1], gami=f[2], ipred=1)
    r2 = pyref.Reference(case, params=prm)
    y, ac = r2.predict_bc(case.yold, acold)
    dy = np.asfortranarray(1e-3 * rng.standard_normal((n, 5)))
    y1, ac1 = r2.correct_bc(y, ac, case.yold, acold, dy)
    np.savez_compressed(os.path.join(HERE, f"ref_alpha_box{N}.npz"), N=N,
                        rhoinf=np.array(RHOINF), factors=np.array(fac).reshape(-1, 2, 3),
                        ipred=np.array(ipo).reshape(-1, 2), shift=np.array(shift).reshape(-1, 2),
                        acold=acold, y=y, ac=ac, dy=dy, y_corr=y1, ac_corr=ac1)


def make_idiff(N, seed):
    """ElmGMRs with the global reconstruction of the diffusive flux
    (idiff 1 and 3) on a jittered box at a perturbed state."""
    case = mesh.box_case(N, mach=0.5, kspace=10, jitter=0.1)
    rng = np.random.default_rng(seed)
    y, ac = pyref.Reference(case).predict_bc(case.yold)
    y = np.asfortranarray(y * (1 + 0.02 * rng.standard_normal(y.shape)))
    ac = np.asfortranarray(ac + 1e3 * rng.standard_normal(ac.shape))
    out = {}
    for idiff in (1, 3):
        res, rmes, bd, lk = pyref.Reference(case, params=dict(case.params, idiff=idiff)).elmgmrs(y, ac)
        out[f"res{idiff}"] = res
    out["BDiag"], out["lhsK"] = bd, lk
    np.savez_compressed(os.path.join(HERE, f"ref_idiff_box{N}.npz"), N=N, y=y, ac=ac, **out)


def make_dc(N, seed):
    """ElmGMRs with discontinuity capturing (iDC 1, 2, 3: e3dc.f with the
    A0DC, A0inv, dV/dY of e3mtrx.f and the rTLS, raLS, giju of e3tau.f) on
    a jittered box at a strongly perturbed state, so the DC factor is
    active at most points."""
    case = mesh.box_case(N, mach=0.5, kspace=10, jitter=0.1)
    rng = np.random.default_rng(seed)
    y, ac = pyref.Reference(case).predict_bc(case.yold)
    y = np.asfortranarray(y * (1 + 0.05 * rng.standard_normal(y.shape)))
    ac = np.asfortranarray(ac + 1e3 * rng.standard_normal(ac.shape))
    out = {}
    for idc in (1, 2, 3):
        res, rmes, bd, lk = pyref.Reference(case, params=dict(case.params, idc=idc)).elmgmrs(y, ac)
        out[f"res{idc}"], out[f"BDiag{idc}"], out[f"lhsK{idc}"] = res, bd, lk
    np.savez_compressed(os.path.join(HERE, f"ref_dc_box{N}.npz"), N=N, y=y, ac=ac, **out)


def make_mtau(N, seed):
    """ElmGMRs with the matrix tau of e3tau_nd (itau = 10: e3eig1/e3eig2
    eigen-decomposition, e3visc's compK diffusion correction, dY/dV) alone
    and with discontinuity capturing (iDC 1, whose giju then comes from the
    rotated dxidx), on a jittered box at a perturbed state."""
    case = mesh.box_case(N, mach=0.5, kspace=10, jitter=0.1)
    rng = np.random.default_rng(seed)
    y, ac = pyref.Reference(case).predict_bc(case.yold)
    y = np.asfortranarray(y * (1 + 0.05 * rng.standard_normal(y.shape)))
    ac = np.asfortranarray(ac + 1e3 * rng.standard_normal(ac.shape))
    out = {}
    for tag, extra in (("", {}), ("_dc", {"idc": 1})):
        prm = dict(case.params, itau=10, **extra)
        res, rmes, bd, lk = pyref.Reference(case, params=prm).elmgmrs(y, ac)
        out[f"res{tag}"], out[f"BDiag{tag}"], out[f"lhsK{tag}"] = res, bd, lk
    np.savez_compressed(os.path.join(HERE, f"ref_mtau_box{N}.npz"), N=N, y=y, ac=ac, **out)


def make_mfg(N, seed):
    """The matrix-free solver's pieces (lhs = 0, iprec = 1) at a perturbed
    state on a jittered box: ElmMFG (res, the modified residual rmes, e3bdg's
    BDiag), ItrRes of perturbed variables (plain and with iabres), the
    preconditioned base point ypre = U new2old(y) (i3LU 'product'), itrFDI's
    interval eGMRES, Au1MFG's one-sided and Au2MFG's two-sided products; and
    ElmMFG with idiff = 1."""
    case = mesh.box_case(N, mach=0.5, kspace=10, jitter=0.1)
    rng = np.random.default_rng(seed)
    y, ac = pyref.Reference(case).predict_bc(case.yold)
    y = np.asfortranarray(y * (1 + 0.05 * rng.standard_normal(y.shape)))
    ac = np.asfortranarray(ac + 1e3 * rng.standard_normal(ac.shape))
    prm = dict(case.params, lhs=0, iprec=1)
    R = pyref.Reference(case, params=prm)
    res, rmes, bd = R.elmmfg(y, ac)
    yp = np.asfortranarray(y * (1 + 0.01 * rng.standard_normal(y.shape)))
    rm_p = R.itrres(yp, y, ac)
    rm_abs = R.itrres(yp, y, ac, iabres=True)
    bd_lu, _ = R.i3lu(bd, None, "LU_Fact ")
    _, res_f = R.i3lu(bd_lu, res, "forward ")
    _, rmes_f = R.i3lu(bd_lu, rmes, "forward ")
    _, ypre = R.i3lu(bd_lu, R.yshuffle(y, "new2old "), "product ")
    eg = R.itrfdi(ypre, y, ac, rmes_f, res_f, bd_lu)
    u = np.asfortranarray(rng.standard_normal(y.shape))
    au1 = R.au1mfg(ypre, y, ac, rmes_f, res_f, bd_lu, eg, u)
    dy = np.asfortranarray(rng.standard_normal(y.shape))
    au2 = R.au2mfg(ypre, y, ac, rmes_f, res_f, dy, bd_lu)
    res1, rmes1, bd1 = pyref.Reference(case, params=dict(prm, idiff=1)).elmmfg(y, ac)
    np.savez_compressed(os.path.join(HERE, f"ref_mfg_box{N}.npz"), N=N, y=y, ac=ac, res=res,
                        rmes=rmes, BDiag=bd, yp=yp, rmes_p=rm_p, rmes_abs=rm_abs,
                        BDiag_lu=bd_lu, res_f=res_f, rmes_f=rmes_f, ypre=ypre, eGMRES=eg,
                        u=u, au1=au1, dy=dy, au2=au2, res_idiff=res1, rmes_idiff=rmes1,
                        BDiag_idiff=bd1)


def make_mfg_mtau(N, seed):
    """The matrix-free solver's pieces with the matrix tau (itau = 10): ElmMFG
    (res, rmes, and the block diagonal of e3bdg_nd), ItrRes of perturbed
    variables, itrFDI's interval and Au1MFG's product."""
    case = mesh.box_case(N, mach=0.5, kspace=10, jitter=0.1)
    rng = np.random.default_rng(seed)
    y, ac = pyref.Reference(case).predict_bc(case.yold)
    y = np.asfortranarray(y * (1 + 0.05 * rng.standard_normal(y.shape)))
    ac = np.asfortranarray(ac + 1e3 * rng.standard_normal(ac.shape))
    prm = dict(case.params, lhs=0, iprec=1, itau=10)
    R = pyref.Reference(case, params=prm)
    res, rmes, bd = R.elmmfg(y, ac)
    yp = np.asfortranarray(y * (1 + 0.01 * rng.standard_normal(y.shape)))
    rm_p = R.itrres(yp, y, ac)
    bd_lu, _ = R.i3lu(bd, None, "LU_Fact ")
    _, res_f = R.i3lu(bd_lu, res, "forward ")
    _, rmes_f = R.i3lu(bd_lu, rmes, "forward ")
    _, ypre = R.i3lu(bd_lu, R.yshuffle(y, "new2old "), "product ")
    eg = R.itrfdi(ypre, y, ac, rmes_f, res_f, bd_lu)
    u = np.asfortranarray(rng.standard_normal(y.shape))
    au1 = R.au1mfg(ypre, y, ac, rmes_f, res_f, bd_lu, eg, u)
    np.savez_compressed(os.path.join(HERE, f"ref_mfg_mtau_box{N}.npz"), N=N, y=y, ac=ac,
                        res=res, rmes=rmes, BDiag=bd, yp=yp, rmes_p=rm_p, eGMRES=eg, u=u,
                        au1=au1)


def make_franca(N, seed):
    """ElmGMRs with Franca's diagonal tau (itau = 1, e3tau.f:54-139) alone,
    with discontinuity capturing (iDC 1) and with the flux reconstruction
    (idiff 1), and ElmMFG's pieces with it, at a perturbed state."""
    case = mesh.box_case(N, mach=0.5, kspace=10, jitter=0.1)
    rng = np.random.default_rng(seed)
    y, ac = pyref.Reference(case).predict_bc(case.yold)
    y = np.asfortranarray(y * (1 + 0.05 * rng.standard_normal(y.shape)))
    ac = np.asfortranarray(ac + 1e3 * rng.standard_normal(ac.shape))
    out = {}
    for tag, extra in (("", {}), ("_dc", {"idc": 1}), ("_idiff", {"idiff": 1})):
        prm = dict(case.params, itau=1, **extra)
        res, rmes, bd, lk = pyref.Reference(case, params=prm).elmgmrs(y, ac)
        out[f"res{tag}"], out[f"BDiag{tag}"], out[f"lhsK{tag}"] = res, bd, lk
    res, rmes, bd = pyref.Reference(case, params=dict(case.params, itau=1, lhs=0,
                                                      iprec=1)).elmmfg(y, ac)
    out.update(mfg_res=res, mfg_rmes=rmes, mfg_BDiag=bd)
    np.savez_compressed(os.path.join(HERE, f"ref_franca_box{N}.npz"), N=N, y=y, ac=ac, **out)


MAKERS = {
    "box": lambda: [make(N, seed) for N, seed in ((2, 11), (3, 12))],
    "alpha": lambda: make_alpha(3, 13),
    "idiff": lambda: make_idiff(3, 14),
    "dc": lambda: make_dc(3, 15),
    "mtau": lambda: make_mtau(3, 17),
    "mfg": lambda: make_mfg(3, 19),
    "mfg_mtau": lambda: make_mfg_mtau(3, 21),
    "franca": lambda: make_franca(3, 23),
}

if __name__ == "__main__":
    # python tests/golden/make_golden.py [name ...]: all fixtures, or the named ones
    for name in sys.argv[1:] or list(MAKERS):
        MAKERS[name]()
        print("wrote", name)
