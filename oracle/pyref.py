"""ctypes bindings of oracle/_ref/libphref.so: the reference's own compiled
ElmGMRs, e3, SparseAp, Spsi3pre, bc3per, i3LU, itrPredict/itrCorrect/
itrBC and the matrix-free ElmMFG, ItrRes, itrFDI, Au1MFG, Au2MFG and
yshuffle behind the test harness oracle/ref/phref_harness.f.  Test
infrastructure only.

The library keeps its state in the reference's COMMON blocks, so one case
is active at a time (phref_setup replaces it).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_PATH = os.path.join(_HERE, "_ref", "libphref.so")
_LIB = None
_i64 = ctypes.c_longlong
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)

PARAM_ORDER = ("Rgas", "gamma", "gamma1", "pr", "mu", "dtsfct", "taucfct", "temper",
               "almi", "alfi", "gami", "Dtgl", "etol", "epsM")
IPARAM_ORDER = ("Kspace", "nGMRES", "minIters", "lhs", "iprec", "ipred", "idiff", "idc",
                "itau")


def available():
    return os.path.exists(_PATH)


def lib():
    global _LIB
    if _LIB is None:
        if not available():
            raise RuntimeError("oracle/_ref not built (make -C oracle/ref)")
        _LIB = ctypes.CDLL(_PATH, mode=os.RTLD_LAZY | os.RTLD_LOCAL)
    return _LIB


def _p(a, t=_dp):
    return a.ctypes.data_as(t)


class Reference:
    def __init__(self, case, params=None, ibksiz=64):
        self.case = case
        prm = dict(case.params if params is None else params)
        self._keep = [np.asfortranarray(a) for a in
                      (case.x, case.ien, case.ienb, case.iBC, case.BC, case.iper,
                       case.colm, case.rowp)]
        x, ien, ienb, iBC, BC, iper, colm, rowp = self._keep
        d = np.array([prm[k] for k in PARAM_ORDER], dtype=np.float64)
        i = np.array([prm[k] for k in IPARAM_ORDER], dtype=np.int32)
        self._d, self._i = d, i
        lib().phref_setup(_i64(case.nshg), _i64(case.nel), _i64(case.nelb),
                          _i64(case.nnz_tot), ctypes.c_int(BC.shape[1]), ctypes.c_int(ibksiz),
                          _p(x), _p(ien, _ip), _p(ienb, _ip), _p(iBC, _ip), _p(BC),
                          _p(iper, _ip), _p(colm, _ip), _p(rowp, _ip), _p(d), _p(i, _ip))

    def predict_bc(self, yold, acold=None):
        n = self.case.nshg
        acold = np.zeros((n, 5), order="F") if acold is None else np.asfortranarray(acold)
        yold = np.asfortranarray(yold)
        y = np.zeros((n, 5), order="F")
        ac = np.zeros((n, 5), order="F")
        lib().phref_predict_bc(_p(yold), _p(acold), _p(y), _p(ac))
        return y, ac

    def itr_setup(self, rhoinf, ipred, y, acold):
        """itrSetup (itrPC.f:6-37): ((almi, alfi, gami), ipred, y after
        the start-up shift)."""
        y = np.array(y, order="F")
        fac = np.zeros(3)
        ip = ctypes.c_int()
        lib().phref_itrsetup(ctypes.c_double(rhoinf), ctypes.c_int(ipred), _p(y),
                             _p(np.asfortranarray(acold)), _p(fac), ctypes.byref(ip))
        return tuple(fac), ip.value, y

    def correct_bc(self, y, ac, yold, acold, dy):
        y, ac = y.copy(order="F"), ac.copy(order="F")
        lib().phref_correct_bc(_p(y), _p(ac), _p(np.asfortranarray(yold)),
                               _p(np.asfortranarray(acold)), _p(np.asfortranarray(dy)))
        return y, ac

    def update_bc(self, yold, acold, y, ac):
        yold, acold = yold.copy(order="F"), acold.copy(order="F")
        lib().phref_update_bc(_p(yold), _p(acold), _p(np.asfortranarray(y)),
                              _p(np.asfortranarray(ac)))
        return yold, acold

    def elmgmrs(self, y, ac):
        n, nnz = self.case.nshg, self.case.nnz_tot
        res = np.zeros((n, 5), order="F")
        rmes = np.zeros((n, 5), order="F")
        bd = np.zeros((n, 5, 5), order="F")
        lk = np.zeros((25, nnz), order="F")
        lib().phref_elmgmrs(_p(np.asfortranarray(y)), _p(np.asfortranarray(ac)), _p(res),
                            _p(rmes), _p(bd), _p(lk))
        return res, rmes, bd, lk

    def sparse_ap(self, lhsK, p):
        p = np.array(p, order="F")
        lib().phref_sparseap(_p(np.asfortranarray(lhsK)), _p(p))
        return p

    def spsi3pre(self, bdiag_lu, lhsK):
        lk = np.array(lhsK, order="F")
        lib().phref_spsi3pre(_p(np.asfortranarray(bdiag_lu)), _p(lk))
        return lk

    def bc3per(self, res):
        res = np.array(res, order="F")
        lib().phref_bc3per(_p(res), ctypes.c_int(res.shape[1]))
        return res

    def e3(self, yl, acl, xl):
        rl = np.zeros((1, 4, 5), order="F")
        eg = np.zeros((1, 20, 20), order="F")
        yl = np.asfortranarray(np.asarray(yl, dtype=np.float64).reshape(1, 4, 5))
        acl = np.asfortranarray(np.asarray(acl, dtype=np.float64).reshape(1, 4, 5))
        xl = np.asfortranarray(np.asarray(xl, dtype=np.float64).reshape(1, 4, 3))
        lib().phref_e3(_p(yl), _p(acl), _p(xl), _p(rl), _p(eg))
        return rl[0], eg[0]

    def i3lu(self, diag, r, code):
        codes = {"LU_Fact ": 0, "forward ": 1, "backward": 2, "product ": 3}
        n = self.case.nshg
        diag = np.array(diag, order="F")
        r = np.zeros((n, 5), order="F") if r is None else np.array(r, order="F")
        lib().phref_i3lu(_p(diag), _p(r), ctypes.c_int(codes[code]))
        return diag, r

    # --- matrix-free GMRES pieces ---------------------------------------
    def elmmfg(self, y, ac, bd=None):
        n = self.case.nshg
        res = np.zeros((n, 5), order="F")
        rmes = np.zeros((n, 5), order="F")
        bd = np.zeros((n, 5, 5), order="F") if bd is None else np.array(bd, order="F")
        lib().phref_elmmfg(_p(np.asfortranarray(y)), _p(np.asfortranarray(ac)), _p(res),
                           _p(rmes), _p(bd))
        return res, rmes, bd

    def itrres(self, yp, yc, ac, rmes=None, iabres=False):
        n = self.case.nshg
        rmes = np.zeros((n, 5), order="F") if rmes is None else np.array(rmes, order="F")
        lib().phref_itrres(_p(np.asfortranarray(yp)), _p(np.asfortranarray(yc)),
                           _p(np.asfortranarray(ac)), _p(rmes), ctypes.c_int(int(iabres)))
        return rmes

    def itrfdi(self, ypre, y, ac, rmes, u, bd_lu):
        eg = ctypes.c_double(0.0)
        lib().phref_itrfdi(*(_p(np.asfortranarray(a)) for a in (ypre, y, ac, rmes, u, bd_lu)),
                           ctypes.byref(eg))
        return eg.value

    def au1mfg(self, ypre, y, ac, rmes, res, bd_lu, egmres, u):
        u = np.array(u, order="F", dtype=np.float64)
        lib().phref_au1mfg(*(_p(np.asfortranarray(a)) for a in (ypre, y, ac, rmes, res)),
                           _p(u), _p(np.asfortranarray(bd_lu)), ctypes.c_double(egmres))
        return u

    def au2mfg(self, ypre, y, ac, rmes, res, dy, bd_lu):
        u = np.zeros((self.case.nshg, 5), order="F")
        lib().phref_au2mfg(*(_p(np.asfortranarray(a)) for a in (ypre, y, ac, rmes, res, dy)),
                           _p(u), _p(np.asfortranarray(bd_lu)))
        return u

    def yshuffle(self, r, code):
        r = np.array(r, order="F", dtype=np.float64)
        lib().phref_yshuffle(_p(r), ctypes.c_int(1 if code == "new2old " else 0))
        return r
