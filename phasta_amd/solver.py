"""Host-side mirror of the reference's solver interface over ``libphgpu.so``.

``GpuSolver`` exposes the reference's entry points with their argument
meaning and array layouts (Fortran-ordered numpy arrays, 1-based
connectivity):

* ``elmgmrs(y, ac)``  -> res, rmes, BDiag, lhsK        (ElmGMRs, elmgmr.f:280)
* ``solgmrs(y, ac)``  -> res, rmes, BDiag, lhsK, Dy    (SolGMRs, solgmr.f:368)
* ``sparse_ap(lhsK, p)``, ``i3lu(Diag, r, code)``, ``spsi3pre(BDiag, lhsK)``,
  ``bc3per(res)``                                       (sparseap.f, i3lu.f,
  spsi3pre.f, bc3per.f)
* ``predict()``, ``newton_iteration()`` on the device-resident state
  (itrPredict/itrBC and SolGMRs + itrCorrect/itrBC, itrdrv.f:398-599).

Errors from the library raise ``PhgError`` with the library's message, the
counterpart of the reference's ``error()`` abort.
"""
import ctypes

import numpy as np

from ._native import load

_i64 = ctypes.c_longlong
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)

PHG_FP64, PHG_MIXED, PHG_FP64_MATRIX_FREE = 0, 1, 2


class PhgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libphgpu error {code}: {msg}")
        self.code = code


class PhgParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in
                ("Rgas", "gamma", "gamma1", "pr", "mu", "dtsfct", "taucfct", "temper",
                 "almi", "alfi", "gami", "Dtgl", "etol", "epsM")] + \
               [(n, ctypes.c_int) for n in ("Kspace", "nGMRES", "minIters", "lhs", "iprec", "ipred",
                                               "idiff", "iter", "idc", "itau")]


class PhgMesh(ctypes.Structure):
    _fields_ = [("nshg", _i64), ("nel", _i64), ("nelb", _i64), ("nnz_tot", _i64),
                ("ndofBC", ctypes.c_int), ("reserved", ctypes.c_int),
                ("x", _dp), ("ien", _ip), ("ienb", _ip), ("iBC", _ip), ("BC", _dp),
                ("iper", _ip), ("colm", _ip), ("rowp", _ip)]


class PhgStats(ctypes.Structure):
    _fields_ = [("iKs", ctypes.c_int), ("lGMRES", ctypes.c_int), ("ntotGM", ctypes.c_int),
                ("converged", ctypes.c_int), ("unorm0", ctypes.c_double),
                ("epsnrm", ctypes.c_double), ("echeck", ctypes.c_double),
                ("totres", ctypes.c_double * 2), ("reslast", ctypes.c_double),
                ("nnan", _i64), ("nlarge", _i64), ("ms_assembly", ctypes.c_double),
                ("ms_solve", ctypes.c_double), ("ms_total", ctypes.c_double),
                ("ms_spmv", ctypes.c_double), ("ms_mgs", ctypes.c_double),
                ("n_spmv", ctypes.c_int), ("n_mgs", ctypes.c_int),
                ("refine_cycles", ctypes.c_int), ("pc_fallback", ctypes.c_int),
                ("orth_loss", ctypes.c_double), ("ms_point", ctypes.c_double),
                ("ms_blocks", ctypes.c_double)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}
        d["totres"] = tuple(self.totres)
        d["converged"] = bool(self.converged)
        return d


# Every symbol declared in include/phgpu.h.
EXPORTS = ("phg_init", "phg_finalize", "phg_set_params", "phg_last_error", "phg_info",
           "phg_elmgmrs", "phg_solgmrs", "phg_state_upload", "phg_state_download",
           "phg_state_download_old", "phg_state_upload_cur", "phg_predict",
           "phg_newton_iteration", "phg_update", "phg_elmmfg", "phg_itrres", "phg_au1mfg",
           "phg_solmfg",
           "phg_newton_iteration_mfg", "phg_asiq", "phg_sparse_ap", "phg_sparse_ap_t",
           "phg_tangent_ap",
           "phg_i3lu", "phg_spsi3pre", "phg_spsi3pre_noalloc", "phg_bc3per",
           "phg_bench_spmv", "phg_bench_assembly", "phg_bench_copy", "phg_sync",
           "phg_diagnostics",
           "phg_mpnn_param_count", "phg_mpnn_init", "phg_mpnn_set_weights",
           "phg_mpnn_get_weights", "phg_mpnn_enable", "phg_mpnn_apply",
           "phg_mpnn_train_step", "phg_mpnn_last_pair", "phg_mpnn_share",
           "phg_mpnn_dataset", "phg_mpnn_store_pair", "phg_mpnn_train_pair",
           "phg_mpnn_train_krylov")


def lib():
    L = load("libphgpu.so")
    if not getattr(L, "_typed", False):
        L.phg_last_error.restype = ctypes.c_char_p
        L.phg_init.argtypes = [ctypes.POINTER(PhgMesh), ctypes.POINTER(PhgParams),
                               ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
        for name in EXPORTS:
            getattr(L, name)
        L._typed = True
    return L


def _p(a, t=_dp):
    if a is None:
        return None
    return a.ctypes.data_as(t)


def _check(rc):
    if rc:
        raise PhgError(rc, lib().phg_last_error().decode())


def _f64(a, shape):
    a = np.asarray(a, dtype=np.float64)
    if a.shape != shape:
        raise ValueError(f"expected shape {shape}, got {a.shape}")
    return np.asfortranarray(a)


def make_params(d):
    p = PhgParams()
    for k, v in d.items():
        setattr(p, k, v)
    return p


class GpuSolver:
    """One device context for one mesh (phg_init ... phg_finalize)."""

    def __init__(self, case, params=None, precision=PHG_FP64):
        self.case = case
        self.params = dict(case.params if params is None else params)
        m = PhgMesh()
        m.nshg, m.nel, m.nelb, m.nnz_tot = case.nshg, case.nel, case.nelb, case.nnz_tot
        m.ndofBC = case.BC.shape[1]
        self._keep = [np.asfortranarray(a) for a in
                      (case.x, case.ien, case.ienb, case.iBC, case.BC, case.iper,
                       case.colm, case.rowp)]
        x, ien, ienb, iBC, BC, iper, colm, rowp = self._keep
        m.x, m.ien, m.ienb, m.iBC = _p(x), _p(ien, _ip), _p(ienb, _ip), _p(iBC, _ip)
        m.BC, m.iper, m.colm, m.rowp = _p(BC), _p(iper, _ip), _p(colm, _ip), _p(rowp, _ip)
        self._prm = make_params(self.params)
        h = ctypes.c_void_p()
        _check(lib().phg_init(ctypes.byref(m), ctypes.byref(self._prm), precision,
                              ctypes.byref(h)))
        self.h = h
        self._mesh = m
        self.precision = precision

    def close(self):
        if getattr(self, "h", None):
            lib().phg_finalize(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, **over):
        self.params.update(over)
        self._prm = make_params(self.params)
        _check(lib().phg_set_params(self.h, ctypes.byref(self._prm)))

    def info(self):
        nc, nb, by = ctypes.c_int(), ctypes.c_int(), _i64()
        _check(lib().phg_info(self.h, ctypes.byref(nc), ctypes.byref(nb), ctypes.byref(by)))
        return dict(ncolors=nc.value, ncolors_b=nb.value, dev_bytes=by.value)

    # ---- reference-shaped entry points (host arrays) ----
    def elmgmrs(self, y, ac, want_lhs=True):
        n, nnz = self.case.nshg, self.case.nnz_tot
        y, ac = _f64(y, (n, 5)), _f64(ac, (n, 5))
        res = np.zeros((n, 5), order="F")
        rmes = np.zeros((n, 5), order="F")
        bd = np.zeros((n, 5, 5), order="F")
        lk = np.zeros((25, nnz), order="F") if want_lhs else None
        _check(lib().phg_elmgmrs(self.h, _p(y), _p(ac), _p(res), _p(rmes), _p(bd), _p(lk)))
        return res, rmes, bd, lk

    def solgmrs(self, y, ac, want_lhs=False, iter=None):
        """SolGMRs on host arrays.  iter: the /timdat/ Newton iteration of
        the time step (the divergence abort applies from iter 2 on); None
        keeps the context's value."""
        if iter is not None:
            self.set_params(iter=int(iter))
        n, nnz = self.case.nshg, self.case.nnz_tot
        y, ac = _f64(y, (n, 5)), _f64(ac, (n, 5))
        res = np.zeros((n, 5), order="F")
        rmes = np.zeros((n, 5), order="F")
        bd = np.zeros((n, 5, 5), order="F")
        dy = np.zeros((n, 5), order="F")
        lk = np.zeros((25, nnz), order="F") if want_lhs else None
        st = PhgStats()
        _check(lib().phg_solgmrs(self.h, _p(y), _p(ac), _p(res), _p(rmes), _p(bd), _p(lk),
                                 _p(dy), ctypes.byref(st)))
        out = st.as_dict()
        out.update(res=res, rmes=rmes, BDiag=bd, lhsK=lk, Dy=dy)
        return out

    # ---- matrix-free GMRES (SolMFG) ----
    def elmmfg(self, y, ac):
        """ElmMFG: (res, rmes, BDiag); BDiag only with iprec != 0."""
        n = self.case.nshg
        y, ac = _f64(y, (n, 5)), _f64(ac, (n, 5))
        res = np.zeros((n, 5), order="F")
        rmes = np.zeros((n, 5), order="F")
        bd = np.zeros((n, 5, 5), order="F")
        _check(lib().phg_elmmfg(self.h, _p(y), _p(ac), _p(res), _p(rmes), _p(bd)))
        return res, rmes, bd

    def itrres(self, yp, yc, ac, rmes=None, iabres=False):
        """ItrRes: rmes (zeros if None) += the modified residual of yp at
        the state yc."""
        n = self.case.nshg
        rmes = np.zeros((n, 5), order="F") if rmes is None else _f64(rmes, (n, 5)).copy(order="F")
        _check(lib().phg_itrres(self.h, _p(_f64(yp, (n, 5))), _p(_f64(yc, (n, 5))),
                                _p(_f64(ac, (n, 5))), _p(rmes), int(bool(iabres))))
        return rmes

    def au1mfg(self, ypre, y, ac, rmes, bd_lu, egmres, u):
        """Au1MFG: the one-sided difference product of u at the state y."""
        n = self.case.nshg
        u = _f64(u, (n, 5)).copy(order="F")
        _check(lib().phg_au1mfg(self.h, _p(_f64(ypre, (n, 5))), _p(_f64(y, (n, 5))),
                                _p(_f64(ac, (n, 5))), _p(_f64(rmes, (n, 5))),
                                _p(_f64(bd_lu, (n, 5, 5))), ctypes.c_double(egmres), _p(u)))
        return u

    def solmfg(self, y, ac, egmres=0.0, fdi=True):
        """SolMFG on host arrays; returns the stats dict with res, rmes,
        BDiag, Dy and the interval eGMRES."""
        n = self.case.nshg
        y, ac = _f64(y, (n, 5)), _f64(ac, (n, 5))
        res = np.zeros((n, 5), order="F")
        rmes = np.zeros((n, 5), order="F")
        bd = np.zeros((n, 5, 5), order="F")
        dy = np.zeros((n, 5), order="F")
        eg = ctypes.c_double(egmres)
        st = PhgStats()
        _check(lib().phg_solmfg(self.h, _p(y), _p(ac), _p(res), _p(rmes), _p(bd), _p(dy),
                                ctypes.byref(eg), int(bool(fdi)), ctypes.byref(st)))
        out = st.as_dict()
        out.update(res=res, rmes=rmes, BDiag=bd, Dy=dy, eGMRES=eg.value)
        return out

    def newton_iteration_mfg(self, egmres=None, fdi=False):
        """SolMFG + itrCorrect + itrBC on the resident state.  egmres: the
        interval to use (None: the context's); returns the stats dict with
        the interval used."""
        eg = ctypes.c_double(0.0 if egmres is None else egmres)
        st = PhgStats()
        _check(lib().phg_newton_iteration_mfg(self.h, None if egmres is None else ctypes.byref(eg),
                                              int(bool(fdi)), ctypes.byref(st)))
        out = st.as_dict()
        if egmres is not None:
            out["eGMRES"] = eg.value
        return out

    def asiq(self, y):
        """AsIq + qpbc (idiff 1 or 3): (qres (nshg,12), 1/lumped mass)."""
        n = self.case.nshg
        qres = np.zeros((n, 12), order="F")
        rmass = np.zeros(n)
        _check(lib().phg_asiq(self.h, _p(_f64(y, (n, 5))), _p(qres), _p(rmass)))
        return qres, rmass

    def sparse_ap(self, lhsK, p):
        p = _f64(p, (self.case.nshg, 5)).copy(order="F")
        lk = None if lhsK is None else _f64(lhsK, (25, self.case.nnz_tot))
        _check(lib().phg_sparse_ap(self.h, _p(lk), _p(p)))
        return p

    def tangent_ap(self, y, p):
        """bc3per(A p(iper)) with the FP64 tangent A of state y, formed
        element by element (no stored matrix)."""
        n = self.case.nshg
        p = _f64(p, (n, 5)).copy(order="F")
        _check(lib().phg_tangent_ap(self.h, _p(_f64(y, (n, 5))), _p(p)))
        return p

    def sparse_ap_t(self, lhsK, p):
        p = _f64(p, (self.case.nshg, 5)).copy(order="F")
        lk = None if lhsK is None else _f64(lhsK, (25, self.case.nnz_tot))
        _check(lib().phg_sparse_ap_t(self.h, _p(lk), _p(p)))
        return p

    def i3lu(self, diag, r, code):
        n = self.case.nshg
        diag = _f64(diag, (n, 5, 5)).copy(order="F")
        r = None if r is None else _f64(r, (n, 5)).copy(order="F")
        _check(lib().phg_i3lu(self.h, _p(diag), _p(r), code.encode()))
        return diag, r

    def spsi3pre(self, bdiag_lu, lhsK):
        bd = _f64(bdiag_lu, (self.case.nshg, 5, 5))
        lk = _f64(lhsK, (25, self.case.nnz_tot)).copy(order="F")
        _check(lib().phg_spsi3pre(self.h, _p(bd), _p(lk)))
        return lk

    def bc3per(self, res):
        res = np.asfortranarray(np.array(res, dtype=np.float64))
        _check(lib().phg_bc3per(self.h, _p(res), ctypes.c_int(res.shape[1])))
        return res

    # ---- device-resident stepping ----
    def upload_state(self, yold, acold=None):
        n = self.case.nshg
        yold = _f64(yold, (n, 5))
        acold = np.zeros((n, 5), order="F") if acold is None else _f64(acold, (n, 5))
        _check(lib().phg_state_upload(self.h, _p(yold), _p(acold)))

    def predict(self):
        _check(lib().phg_predict(self.h))

    def newton_iteration(self):
        st = PhgStats()
        _check(lib().phg_newton_iteration(self.h, ctypes.byref(st)))
        return st.as_dict()

    def update(self):
        """itrUpdate + itrBC on the resident (yold, acold)."""
        _check(lib().phg_update(self.h))

    def upload_current(self, y, ac):
        n = self.case.nshg
        _check(lib().phg_state_upload_cur(self.h, _p(_f64(y, (n, 5))), _p(_f64(ac, (n, 5)))))

    def download_old(self):
        n = self.case.nshg
        yold = np.zeros((n, 5), order="F")
        acold = np.zeros((n, 5), order="F")
        _check(lib().phg_state_download_old(self.h, _p(yold), _p(acold)))
        return yold, acold

    def download_state(self):
        n = self.case.nshg
        y = np.zeros((n, 5), order="F")
        ac = np.zeros((n, 5), order="F")
        dy = np.zeros((n, 5), order="F")
        _check(lib().phg_state_download(self.h, _p(y), _p(ac), _p(dy)))
        return y, ac, dy

    def sync(self):
        _check(lib().phg_sync(self.h))

    def diagnostics(self, orth=False):
        """orth: measure each GMRES cycle's basis orthogonality into
        the stats' orth_loss."""
        _check(lib().phg_diagnostics(self.h, 1 if orth else 0))

    def bench_spmv(self, reps=20):
        ms, by = ctypes.c_double(), ctypes.c_double()
        _check(lib().phg_bench_spmv(self.h, reps, ctypes.byref(ms), ctypes.byref(by)))
        return ms.value, by.value

    # ---- learned preconditioner (k_mpnn.hip) ----
    def mpnn_init(self, dscale):
        m = self._mesh
        ds = np.ascontiguousarray(dscale, dtype=np.float32)
        if ds.shape != (25,):
            raise ValueError("dscale must hold 25 floats")
        _check(lib().phg_mpnn_init(self.h, ctypes.byref(m), ds.ctypes.data_as(ctypes.c_void_p)))

    def mpnn_set_weights(self, flat):
        w = np.ascontiguousarray(flat, dtype=np.float32)
        _check(lib().phg_mpnn_set_weights(self.h, w.ctypes.data_as(ctypes.c_void_p), len(w)))

    def mpnn_get_weights(self):
        w = np.zeros(lib().phg_mpnn_param_count(), dtype=np.float32)
        _check(lib().phg_mpnn_get_weights(self.h, w.ctypes.data_as(ctypes.c_void_p), len(w)))
        return w

    def mpnn_enable(self, use_pc=1, keep_pair=False):
        """use_pc: directions per GMRES cycle preconditioned by the network
        (0 off; a value >= Kspace: every direction)."""
        _check(lib().phg_mpnn_enable(self.h, int(use_pc), int(keep_pair)))

    def mpnn_apply(self, bdiag_raw, v):
        n = self.case.nshg
        bd, v = _f64(bdiag_raw, (n, 5, 5)), _f64(v, (n, 5))
        z = np.zeros((n, 5), order="F")
        _check(lib().phg_mpnn_apply(self.h, _p(bd), _p(v), _p(z)))
        return z

    def mpnn_train_step(self, bdiag_raw, v, t, lr=0.0, want_grad=False):
        n = self.case.nshg
        bd, v, t = _f64(bdiag_raw, (n, 5, 5)), _f64(v, (n, 5)), _f64(t, (n, 5))
        loss = ctypes.c_double()
        g = np.zeros(lib().phg_mpnn_param_count(), dtype=np.float32) if want_grad else None
        gp = g.ctypes.data_as(ctypes.c_void_p) if want_grad else None
        _check(lib().phg_mpnn_train_step(self.h, _p(bd), _p(v), _p(t), ctypes.c_float(lr),
                                         ctypes.byref(loss), gp))
        return (loss.value, g) if want_grad else loss.value

    def mpnn_last_pair(self):
        n = self.case.nshg
        bd = np.zeros((n, 5, 5), order="F")
        b = np.zeros((n, 5), order="F")
        z = np.zeros((n, 5), order="F")
        _check(lib().phg_mpnn_last_pair(self.h, _p(bd), _p(b), _p(z)))
        return bd, b, z

    def mpnn_share(self, src):
        _check(lib().phg_mpnn_share(self.h, src.h))

    def mpnn_dataset(self, capacity):
        _check(lib().phg_mpnn_dataset(self.h, int(capacity)))

    def mpnn_store_pair(self, slot):
        _check(lib().phg_mpnn_store_pair(self.h, int(slot)))

    def mpnn_train_pair(self, slot, lr, want_loss=True):
        loss = ctypes.c_double()
        _check(lib().phg_mpnn_train_pair(self.h, int(slot), ctypes.c_float(lr),
                                         ctypes.byref(loss) if want_loss else None))
        return loss.value if want_loss else None

    def mpnn_train_krylov(self, nvec, lr, want_grad=False):
        """Mean residual loss over Krylov vectors 1..nvec of the last solve
        (one Adam step per vector when lr > 0); with want_grad also the
        gradient of the last vector's loss."""
        loss = ctypes.c_double()
        g = np.zeros(lib().phg_mpnn_param_count(), dtype=np.float32) if want_grad else None
        gp = g.ctypes.data_as(ctypes.c_void_p) if want_grad else None
        _check(lib().phg_mpnn_train_krylov(self.h, int(nvec), ctypes.c_float(lr),
                                           ctypes.byref(loss), gp))
        return (loss.value, g) if want_grad else loss.value

    def bench_copy(self, reps=10):
        ms, by = ctypes.c_double(), ctypes.c_double()
        _check(lib().phg_bench_copy(self.h, reps, ctypes.byref(ms), ctypes.byref(by)))
        return ms.value, by.value

    def bench_assembly(self, reps=3):
        ms = ctypes.c_double()
        _check(lib().phg_bench_assembly(self.h, reps, ctypes.byref(ms)))
        return ms.value


# Parameter block of the learned preconditioner, in the library's order:
# W0 (32 x 64), b0, W1 (320 x 64), b1, W2, b2, W3, b3, Wout (64 x 5), bout,
# every matrix row-major W[k][c] (input k, output c).
MPNN_SHAPES = (("W0", (32, 64)), ("b0", (64,)), ("W1", (320, 64)), ("b1", (64,)),
               ("W2", (320, 64)), ("b2", (64,)), ("W3", (320, 64)), ("b3", (64,)),
               ("Wout", (64, 5)), ("bout", (5,)))


def mpnn_flatten(params):
    return np.concatenate([np.asarray(params[k], dtype=np.float32).ravel()
                           for k, _ in MPNN_SHAPES])


def mpnn_unflatten(flat):
    out, o = {}, 0
    for k, shp in MPNN_SHAPES:
        m = int(np.prod(shp))
        out[k] = np.asarray(flat[o:o + m], dtype=np.float32).reshape(shp)
        o += m
    return out
