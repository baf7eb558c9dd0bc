"""ctypes bindings of the CPU oracle (test infrastructure only).

Imported by tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline
leg as the checker; never by the product package.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_i64 = ctypes.c_longlong
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class OrcMesh(ctypes.Structure):
    _fields_ = [("nshg", _i64), ("nel", _i64), ("nelb", _i64), ("nnz_tot", _i64),
                ("ndofBC", ctypes.c_int), ("ibksiz", ctypes.c_int),
                ("x", _dp), ("ien", _ip), ("ienb", _ip), ("iBC", _ip), ("BC", _dp),
                ("iper", _ip), ("colm", _ip), ("rowp", _ip)]


class OrcParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in
                ("Rgas", "gamma", "gamma1", "pr", "mu", "dtsfct", "taucfct", "temper",
                 "almi", "alfi", "gami", "Dtgl", "etol", "epsM")] + \
               [(n, ctypes.c_int) for n in ("Kspace", "nGMRES", "minIters", "lhs", "iprec", "ipred",
                                               "idiff", "iter", "idc", "itau")]


class OrcStats(ctypes.Structure):
    _fields_ = [("iKs", ctypes.c_int), ("lGMRES", ctypes.c_int), ("ntotGM", ctypes.c_int),
                ("converged", ctypes.c_int), ("unorm0", ctypes.c_double),
                ("epsnrm", ctypes.c_double), ("echeck", ctypes.c_double),
                ("totres", ctypes.c_double * 2), ("reslast", ctypes.c_double),
                ("nnan", _i64), ("nlarge", _i64)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "_build", "libphoracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        _LIB = ctypes.CDLL(path)
    return _LIB


def set_threads(n):
    """Threads of the oracle's OpenMP loops (results do not depend on it);
    returns the count in effect."""
    return lib().orc_set_threads(int(n))


def _p(a, t=_dp):
    return a.ctypes.data_as(t)


def make_params(d):
    p = OrcParams()
    for k, v in d.items():
        setattr(p, k, v)
    return p


class Oracle:
    """Holds one case's arrays and exposes the restated routines."""

    def __init__(self, case, params=None, ibksiz=64):
        self.case = case
        self.params = dict(case.params if params is None else params)
        m = OrcMesh()
        m.nshg, m.nel, m.nelb, m.nnz_tot = case.nshg, case.nel, case.nelb, case.nnz_tot
        m.ndofBC, m.ibksiz = case.BC.shape[1], ibksiz
        m.x, m.ien, m.ienb = _p(case.x), _p(case.ien, _ip), _p(case.ienb, _ip)
        m.iBC, m.BC, m.iper = _p(case.iBC, _ip), _p(case.BC), _p(case.iper, _ip)
        m.colm, m.rowp = _p(case.colm, _ip), _p(case.rowp, _ip)
        self.m = m
        self.reslast = 0.0
        self.ntotGM = 0

    def p(self, **over):
        d = dict(self.params)
        d.update(over)
        return make_params(d)

    # --- element level -------------------------------------------------
    def e3(self, yl, acl, xl, lhs=True):
        rl = np.zeros((4, 5))
        eg = np.zeros((20, 20))
        yl, acl, xl = (np.ascontiguousarray(a, dtype=np.float64) for a in (yl, acl, xl))
        lib().orc_e3(ctypes.byref(self.p()), _p(yl), _p(acl), _p(xl), _p(rl), _p(eg),
                     ctypes.c_int(1 if lhs else 0))
        return rl, eg

    def e3b(self, yl, xl):
        rl = np.zeros((4, 5))
        yl, xl = (np.ascontiguousarray(a, dtype=np.float64) for a in (yl, xl))
        lib().orc_e3b(ctypes.byref(self.p()), _p(yl), _p(xl), _p(rl))
        return rl

    # --- nodal/global level -------------------------------------------
    def itr_predict_bc(self, yold, acold=None):
        """itrPredict + itrBC (itrdrv.f:398-399)."""
        n = self.case.nshg
        acold = np.zeros((n, 5), order="F") if acold is None else acold
        y = np.zeros((n, 5), order="F")
        ac = np.zeros((n, 5), order="F")
        L = lib()
        rc = L.orc_itr_predict(_i64(n), ctypes.byref(self.p()), _p(yold), _p(acold), _p(y),
                               _p(ac))
        if rc:
            raise RuntimeError(f"orc_itr_predict rc={rc}")
        rc = L.orc_itr_bc(ctypes.byref(self.m), _p(y), _p(ac))
        if rc:
            raise RuntimeError(f"orc_itr_bc rc={rc}")
        return y, ac

    def asiq(self, y):
        """AsIq + qpbc (elmgmr.f:361-391): (qres (nshg,12), 1/rmass)."""
        n = self.case.nshg
        qres = np.zeros((n, 12), order="F")
        rmass = np.zeros(n)
        rc = lib().orc_asiq(ctypes.byref(self.m), ctypes.byref(self.p()),
                            _p(np.asfortranarray(y)), _p(qres), _p(rmass))
        if rc:
            raise RuntimeError(f"orc_asiq rc={rc}")
        return qres, rmass

    def elmgmrs(self, y, ac):
        n, nnz = self.case.nshg, self.case.nnz_tot
        res = np.zeros((n, 5), order="F")
        rmes = np.zeros((n, 5), order="F")
        bd = np.zeros((n, 5, 5), order="F")
        lk = np.zeros((25, nnz), order="F")
        rc = lib().orc_elmgmrs(ctypes.byref(self.m), ctypes.byref(self.p()), _p(y), _p(ac),
                               _p(res), _p(rmes), _p(bd), _p(lk))
        if rc:
            raise RuntimeError(f"orc_elmgmrs rc={rc}")
        return res, rmes, bd, lk

    def solgmrs(self, y, ac, bd=None, lk=None, **over):
        """SolGMRs; bd and lk, when given, are the caller's persistent BDiag
        and lhsK (Fortran order, updated in place): with lhs = iprec = 0 the
        reference reuses the factored BDiag and scaled lhsK of the last call
        that formed them (itrdrv.f:456-457, solgmr.f:471,495)."""
        n, nnz = self.case.nshg, self.case.nnz_tot
        res = np.zeros((n, 5), order="F")
        rmes = np.zeros((n, 5), order="F")
        if bd is None:
            bd = np.zeros((n, 5, 5), order="F")
        if lk is None:
            lk = np.zeros((25, nnz), order="F")
        assert bd.flags.f_contiguous and lk.flags.f_contiguous
        dy = np.zeros((n, 5), order="F")
        st = OrcStats()
        st.reslast = self.reslast
        st.ntotGM = self.ntotGM
        rc = lib().orc_solgmrs(ctypes.byref(self.m), ctypes.byref(self.p(**over)), _p(y),
                               _p(ac), _p(res), _p(rmes), _p(bd), _p(lk), _p(dy),
                               ctypes.byref(st))
        if rc:
            raise RuntimeError(f"orc_solgmrs rc={rc}")
        self.reslast = st.reslast
        self.ntotGM = st.ntotGM
        return dict(res=res, rmes=rmes, BDiag=bd, lhsK=lk, Dy=dy, iKs=st.iKs,
                    ntotGM=st.ntotGM, converged=bool(st.converged), unorm0=st.unorm0,
                    epsnrm=st.epsnrm, echeck=st.echeck, totres=tuple(st.totres))

    # --- matrix-free GMRES (SolMFG) ------------------------------------
    def _chk(self, rc, name):
        if rc:
            raise RuntimeError(f"{name} rc={rc}")

    def elmmfg(self, y, ac, bd=None, **over):
        """ElmMFG: (res, rmes, BDiag) after bc3Res / bc3BDg; bd, when given,
        is the caller's persistent BDiag (kept when iprec = 0)."""
        n = self.case.nshg
        res = np.zeros((n, 5), order="F")
        rmes = np.zeros((n, 5), order="F")
        if bd is None:
            bd = np.zeros((n, 5, 5), order="F")
        self._chk(lib().orc_elmmfg(ctypes.byref(self.m), ctypes.byref(self.p(**over)),
                                   _p(np.asfortranarray(y)), _p(np.asfortranarray(ac)),
                                   _p(res), _p(rmes), _p(bd)), "orc_elmmfg")
        return res, rmes, bd

    def itrres(self, yp, yc, ac, rmes=None, iabres=False, **over):
        """ItrRes: rmes (accumulated; zeros if None) += the modified residual
        of the variables yp at the state yc."""
        n = self.case.nshg
        rmes = np.zeros((n, 5), order="F") if rmes is None else rmes
        assert rmes.flags.f_contiguous
        self._chk(lib().orc_itrres(ctypes.byref(self.m), ctypes.byref(self.p(**over)),
                                   _p(np.asfortranarray(yp)), _p(np.asfortranarray(yc)),
                                   _p(np.asfortranarray(ac)), _p(rmes),
                                   ctypes.c_int(1 if iabres else 0)), "orc_itrres")
        return rmes

    def yshuffle(self, r, code):
        r = np.array(r, order="F", dtype=np.float64)
        lib().orc_yshuffle(_i64(r.shape[0]), _p(r), ctypes.c_int(1 if code == "new2old " else 0))
        return r

    def itrfdi(self, ypre, y, ac, rmes, u, bd_lu, **over):
        eg = ctypes.c_double(0.0)
        self._chk(lib().orc_itrfdi(ctypes.byref(self.m), ctypes.byref(self.p(**over)),
                                   _p(np.asfortranarray(ypre)), _p(np.asfortranarray(y)),
                                   _p(np.asfortranarray(ac)), _p(np.asfortranarray(rmes)),
                                   _p(np.asfortranarray(u)), _p(np.asfortranarray(bd_lu)),
                                   ctypes.byref(eg)), "orc_itrfdi")
        return eg.value

    def au1mfg(self, ypre, y, ac, rmes, bd_lu, egmres, u, **over):
        u = np.array(u, order="F", dtype=np.float64)
        self._chk(lib().orc_au1mfg(ctypes.byref(self.m), ctypes.byref(self.p(**over)),
                                   _p(np.asfortranarray(ypre)), _p(np.asfortranarray(y)),
                                   _p(np.asfortranarray(ac)), _p(np.asfortranarray(rmes)),
                                   _p(np.asfortranarray(bd_lu)), ctypes.c_double(egmres),
                                   _p(u)), "orc_au1mfg")
        return u

    def au2mfg(self, ypre, y, ac, res, dy, bd_lu, **over):
        u = np.zeros((self.case.nshg, 5), order="F")
        self._chk(lib().orc_au2mfg(ctypes.byref(self.m), ctypes.byref(self.p(**over)),
                                   _p(np.asfortranarray(ypre)), _p(np.asfortranarray(y)),
                                   _p(np.asfortranarray(ac)), _p(np.asfortranarray(res)),
                                   _p(np.asfortranarray(dy)), _p(np.asfortranarray(bd_lu)),
                                   _p(u)), "orc_au2mfg")
        return u

    def solmfg(self, y, ac, egmres=0.0, fdi=True, bd=None, **over):
        """SolMFG.  egmres: the interval held from an earlier call (used when
        fdi is False); returns the dict of solgmrs plus eGMRES."""
        n = self.case.nshg
        res = np.zeros((n, 5), order="F")
        rmes = np.zeros((n, 5), order="F")
        if bd is None:
            bd = np.zeros((n, 5, 5), order="F")
        dy = np.zeros((n, 5), order="F")
        eg = ctypes.c_double(egmres)
        st = OrcStats()
        st.ntotGM = self.ntotGM
        self._chk(lib().orc_solmfg(ctypes.byref(self.m), ctypes.byref(self.p(**over)),
                                   _p(np.asfortranarray(y)), _p(np.asfortranarray(ac)),
                                   _p(res), _p(rmes), _p(bd), _p(dy), ctypes.byref(eg),
                                   ctypes.c_int(1 if fdi else 0), ctypes.byref(st)),
                  "orc_solmfg")
        self.ntotGM = st.ntotGM
        return dict(res=res, rmes=rmes, BDiag=bd, Dy=dy, iKs=st.iKs, lGMRES=st.lGMRES,
                    ntotGM=st.ntotGM, converged=bool(st.converged), unorm0=st.unorm0,
                    epsnrm=st.epsnrm, echeck=st.echeck, totres=tuple(st.totres),
                    eGMRES=eg.value)

    def itr_correct_bc(self, y, ac, yold, acold, dy):
        n = self.case.nshg
        y, ac = y.copy(order="F"), ac.copy(order="F")
        L = lib()
        L.orc_itr_correct(_i64(n), ctypes.byref(self.p()), _p(y), _p(ac), _p(yold),
                          _p(acold), _p(dy))
        rc = L.orc_itr_bc(ctypes.byref(self.m), _p(y), _p(ac))
        if rc:
            raise RuntimeError(f"orc_itr_bc rc={rc}")
        return y, ac

    def itr_update_bc(self, yold, acold, y, ac):
        """itrUpdate + itrBC on (yold, acold) (itrdrv.f:651-652)."""
        n = self.case.nshg
        yold, acold = yold.copy(order="F"), acold.copy(order="F")
        L = lib()
        L.orc_itr_update(_i64(n), ctypes.byref(self.p()), _p(yold), _p(acold),
                         _p(np.asfortranarray(y)), _p(np.asfortranarray(ac)))
        rc = L.orc_itr_bc(ctypes.byref(self.m), _p(yold), _p(acold))
        if rc:
            raise RuntimeError(f"orc_itr_bc rc={rc}")
        return yold, acold

    def i3lu(self, diag, r, code):
        codes = {"LU_Fact ": 0, "forward ": 1, "backward": 2, "product ": 3}
        rp = _p(r) if r is not None else None
        rc = lib().orc_i3lu(_i64(diag.shape[0]), _p(diag), rp, ctypes.c_int(codes[code]))
        if rc:
            raise RuntimeError("bad i3LU code")

    def spsi3pre(self, bdiag_lu, lhsK):
        c = self.case
        lib().orc_spsi3pre(_i64(c.nshg), _p(bdiag_lu), _p(lhsK), _p(c.colm, _ip),
                           _p(c.rowp, _ip))

    def sparse_ap(self, lhsK, p):
        c = self.case
        lib().orc_sparse_ap(_i64(c.nshg), _p(c.iper, _ip), _p(c.colm, _ip),
                            _p(c.rowp, _ip), _p(lhsK), _p(p))

    def bc3per(self, res):
        c = self.case
        lib().orc_bc3per(_i64(c.nshg), _p(c.iBC, _ip), _p(c.iper, _ip), _p(res),
                         ctypes.c_int(res.shape[1]))

    def newton_step(self, yold, acold, newton=1):
        """One time step: predictor + BC, then `newton` solve/correct pairs."""
        y, ac = self.itr_predict_bc(yold, acold)
        hist = []
        for k in range(newton):
            out = self.solgmrs(y, ac, iter=k + 1)  # itrdrv.f:409,434
            y, ac = self.itr_correct_bc(y, ac, yold, acold, out["Dy"])
            hist.append(out)
        return y, ac, hist

    def time_steps(self, yold, acold, steps, newton=1):
        """`steps` time steps of itrdrv.f:398-652: predict, Newton loop,
        itrUpdate + itrBC.  Returns (yold, acold, history)."""
        hist = []
        for _ in range(steps):
            y, ac, h = self.newton_step(yold, acold, newton)
            yold, acold = self.itr_update_bc(yold, acold, y, ac)
            hist += h
        return yold, acold, hist
