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


if __name__ == "__main__":
    for N, seed in ((2, 11), (3, 12)):
        make(N, seed)
        print("wrote", f"ref_box{N}.npz")
