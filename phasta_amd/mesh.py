"""Synthetic box cases in the reference's array layouts.

The arrays mirror what PHASTA holds after ``readnblk``/``genblk``/``genadj``
(see ``include/phmesh.h``): Fortran-ordered numpy arrays with 1-based
connectivity.  Generation runs in ``libphmesh.so`` (C++/OpenMP).
"""
import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import config
from ._native import load

_i64 = ctypes.c_longlong
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="F_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="F_CONTIGUOUS")


def _lib():
    lib = load("libphmesh.so")
    if not getattr(lib, "_typed", False):
        lib.phm_box_counts.argtypes = [ctypes.c_int, ctypes.POINTER(_i64),
                                       ctypes.POINTER(_i64), ctypes.POINTER(_i64)]
        lib.phm_box_mesh.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_ulonglong, _dp, _ip, _ip, _ip, _ip]
        lib.phm_box_mesh_layer.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_ulonglong, _dp, _ip, _ip,
                                           _ip, _ip]
        lib.phm_csr_count.argtypes = [_i64, _i64, _ip, _ip, ctypes.POINTER(_i64)]
        lib.phm_csr_fill.argtypes = [_i64, _i64, _ip, _ip, _ip]
        lib.phm_box_state.argtypes = [_i64, _dp, ctypes.c_double, ctypes.c_ulonglong,
                                      ctypes.c_double, ctypes.c_double, _dp]
        lib._typed = True
    return lib


@dataclass
class BoxCase:
    N: int
    mach: float
    x: np.ndarray          # (nshg,3)
    ien: np.ndarray        # (nel,4) 1-based
    ienb: np.ndarray       # (nelb,4) 1-based, wall face first
    iBC: np.ndarray        # (nshg,)
    BC: np.ndarray         # (nshg,ndofBC)
    iper: np.ndarray       # (nshg,) 1-based
    colm: np.ndarray       # (nshg+1,) 1-based row pointers
    rowp: np.ndarray       # (nnz_tot,) 1-based columns
    yold: np.ndarray       # (nshg,5) (u,v,w,p,T), before itrBC
    params: dict = field(default_factory=dict)

    @property
    def nshg(self):
        return self.x.shape[0]

    @property
    def nel(self):
        return self.ien.shape[0]

    @property
    def nelb(self):
        return self.ienb.shape[0]

    @property
    def nnz_tot(self):
        return self.rowp.shape[0]


def box_case(N, mach=0.3, jitter=0.1, grade=1.0, kspace=20, etol=1e-3,
             min_iters=0, mesh_seed=config.MESH_SEED,
             field_seed=config.FIELD_SEED, ndofBC=10, cfl=config.CFL, first=0.0):
    """first > 0: capped boundary-layer grading (first wall cell first/N,
    growing by `grade` until the uniform core); 0: geometric grading over
    each half channel."""
    lib = _lib()
    nshg, nel, nelb = _i64(), _i64(), _i64()
    if lib.phm_box_counts(N, ctypes.byref(nshg), ctypes.byref(nel), ctypes.byref(nelb)):
        raise ValueError(f"bad box size N={N}")
    nshg, nel, nelb = nshg.value, nel.value, nelb.value
    x = np.zeros((nshg, 3), order="F")
    ien = np.zeros((nel, 4), dtype=np.int32, order="F")
    ienb = np.zeros((nelb, 4), dtype=np.int32, order="F")
    iBC = np.zeros(nshg, dtype=np.int32)
    iper = np.zeros(nshg, dtype=np.int32)
    rc = lib.phm_box_mesh_layer(N, jitter, grade, first, mesh_seed, x, ien, ienb, iBC, iper)
    if rc:
        raise RuntimeError(f"phm_box_mesh failed ({rc})")
    colm = np.zeros(nshg + 1, dtype=np.int32)
    nnz = _i64()
    if lib.phm_csr_count(nshg, nel, ien, colm, ctypes.byref(nnz)):
        raise RuntimeError("phm_csr_count failed")
    rowp = np.zeros(nnz.value, dtype=np.int32)
    if lib.phm_csr_fill(nshg, nel, ien, colm, rowp):
        raise RuntimeError("phm_csr_fill failed")
    yold = np.zeros((nshg, 5), order="F")
    lib.phm_box_state(nshg, x, mach, field_seed, config.RGAS, config.GAMMA, yold)
    BC = np.zeros((nshg, ndofBC), order="F")
    params = config.solver_params(N, mach, kspace, etol=etol, min_iters=min_iters,
                                  grade=grade, cfl=cfl, first=first)
    return BoxCase(N=N, mach=mach, x=x, ien=ien, ienb=ienb, iBC=iBC, BC=BC,
                   iper=iper, colm=colm, rowp=rowp, yold=yold, params=params)


def box_state(case, mach, field_seed):
    """A new synthetic state y(nshg,5) on an existing box mesh."""
    y = np.zeros((case.nshg, 5), order="F")
    x = np.asfortranarray(case.x)
    if _lib().phm_box_state(case.nshg, x, mach, field_seed, config.RGAS, config.GAMMA, y):
        raise RuntimeError("phm_box_state failed")
    return y


def named_case(name, **over):
    c = dict(config.CASES[name])
    c.update(over)
    return box_case(c["N"], mach=c["mach"], jitter=c["jitter"], grade=c["grade"],
                    kspace=c["kspace"], cfl=c.get("cfl", config.CFL),
                    first=c.get("first", 0.0))


def expected_nnz(N):
    """nnzb = nshg + 2E for the Kuhn split, E = 3N(N+1)^2 + 3N^2(N+1) + N^3."""
    E = 3 * N * (N + 1) ** 2 + 3 * N * N * (N + 1) + N ** 3
    return (N + 1) ** 3 + 2 * E
