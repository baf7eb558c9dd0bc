"""PHASTA's POSIX phastaIO files through libphmesh.so (include/phastaio.h).

* ``write_restart`` / ``read_restart``: the restart.<step>.<part> file of
  Write_Restart (new_interface.c:240-330) and readnblk.f:368-445: the
  magic number, "solution" and "time derivative of solution", each an
  (nshg, ndof) column-major double block with header ints nshg, ndof, step.
* ``write_geombc`` / ``read_geombc``: the geombc.dat.<part> fields that
  readnblk.f:114-376, genblk.f and genbkb.f read for a linear-tet mesh:
  sizes, co-ordinates, interior and boundary connectivity (with the wall
  elements' nbc codes and values), bc mapping / codes / condition arrays
  and periodic masters; ``read_geombc`` returns the raw fields.
* ``gen_bc`` / ``case_from_geombc``: genBC1's conversion (genbc1.f) of
  the condition array into BC(nshg, ndofBC) for the boundary codes this
  path supports (temperature, pressure, all three velocity components,
  periodic), and a solver case (mesh arrays, CSR graph as genadj builds it,
  the state of an optional restart) from a geombc file.  genBC's wall-normal
  and slip-wall handling (genwnm, genotwn, slpwBC in genbc.f) serve codes
  the solver refuses and are not restated.
* ``compare_restarts``: checkphasta's test (checkphasta.cpp:96-132), the
  largest absolute difference of the "solution" fields against a tolerance.
"""
import ctypes
import os

import numpy as np

from ._native import load

_ip = ctypes.POINTER(ctypes.c_int)
MAGIC = 362436


class PhastaIOError(RuntimeError):
    pass


def lib():
    L = load("libphmesh.so")
    if not getattr(L, "_phastaio_typed", False):
        L.phastaio_last_error.restype = ctypes.c_char_p
        L._phastaio_typed = True
    return L


def _check(rc, what):
    if rc:
        raise PhastaIOError(f"{what}: {lib().phastaio_last_error().decode()} (code {rc})")


class File:
    """One open phastaIO file (openfile ... closefile)."""

    def __init__(self, path, mode, iotype="binary"):
        self.mode = mode
        self.iotype = iotype.encode()
        fd = ctypes.c_int(0)
        _check(lib().phastaio_openfile(os.fsencode(path), mode.encode(), ctypes.byref(fd)),
               f"openfile {path}")
        self.fd = fd

    def close(self):
        if self.fd.value:
            _check(lib().phastaio_closefile(ctypes.byref(self.fd), self.mode.encode()), "closefile")
            self.fd = ctypes.c_int(0)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def write(self, key, header, data=None, dtype="integer"):
        """writeheader + writedatablock (no block when data is None)."""
        h = np.ascontiguousarray(header, dtype=np.int32)
        nh = ctypes.c_int(h.size)
        arr = None
        if data is not None:
            arr = np.asarray(data, dtype=np.int32 if dtype == "integer" else np.float64)
            arr = np.asfortranarray(arr).ravel(order="F")
        nd = ctypes.c_int(0 if arr is None else arr.size)
        L = lib()
        _check(L.phastaio_writeheader(ctypes.byref(self.fd), key.encode(), h.ctypes.data_as(_ip),
                                      ctypes.byref(nh), ctypes.byref(nd), dtype.encode(),
                                      self.iotype), f"writeheader {key}")
        if arr is not None:
            _check(L.phastaio_writedatablock(ctypes.byref(self.fd), key.encode(),
                                             arr.ctypes.data_as(ctypes.c_void_p), ctypes.byref(nd),
                                             dtype.encode(), self.iotype), f"writedatablock {key}")

    def header(self, key, nints, dtype="integer"):
        """readheader: the first nints header integers (None if absent)."""
        v = np.zeros(max(nints, 1), dtype=np.int32)
        n = ctypes.c_int(nints)
        rc = lib().phastaio_readheader(ctypes.byref(self.fd), key.encode(), v.ctypes.data_as(_ip),
                                       ctypes.byref(n), dtype.encode(), self.iotype)
        if rc == 1:
            return None
        _check(rc, f"readheader {key}")
        return v[:nints]

    def block(self, key, n, dtype="integer"):
        """readdatablock of n items after the last header."""
        arr = np.zeros(n, dtype=np.int32 if dtype == "integer" else np.float64)
        nn = ctypes.c_int(n)
        _check(lib().phastaio_readdatablock(ctypes.byref(self.fd), key.encode(),
                                            arr.ctypes.data_as(ctypes.c_void_p), ctypes.byref(nn),
                                            dtype.encode(), self.iotype), f"readdatablock {key}")
        return arr


def restart_name(step, part=1):
    """posix_constructname + appendRank (phio_posix.cc:98-107)."""
    return f"restart.{step}.{part}"


def geombc_name(part=1):
    return f"geombc.dat.{part}"


def write_restart(path, y, ac, step, iotype="binary"):
    """Write_Restart (new_interface.c:240-330), one part."""
    y = np.asfortranarray(y, dtype=np.float64)
    ac = np.asfortranarray(ac, dtype=np.float64)
    nshg, ndof = y.shape
    with File(path, "write", iotype) as f:
        f.write("byteorder magic number", [1], [MAGIC])
        f.write("solution", [nshg, ndof, step], y, "double")
        f.write("time derivative of solution", [nshg, ndof, step], ac, "double")


def read_restart(path, iotype="binary"):
    """readnblk.f:378-445: (solution, time derivative or None, step)."""
    with File(path, "read", iotype) as f:
        h = f.header("solution", 3, "double")
        if h is None:
            raise PhastaIOError(f"{path}: no solution field")
        nshg, ndof, step = (int(v) for v in h)
        q = f.block("solution", nshg * ndof, "double").reshape((nshg, ndof), order="F")
        h = f.header("time derivative of solution", 3, "double")
        ac = None
        if h is not None:
            ac = f.block("time derivative of solution", int(h[0]) * int(h[1]),
                         "double").reshape((int(h[0]), int(h[1])), order="F")
    return q, ac, step


def write_geombc(path, case, ndof=5, iotype="binary"):
    """The geombc fields of a linear-tet mesh as readnblk.f, genblk.f and
    genbkb.f read them.  Nodes with nonzero iBC get a condition record
    holding the case's BC values in the input form gen_bc reads."""
    nshg, nel, nelb = case.nshg, case.nel, case.nelb
    ibc = np.asarray(case.iBC, dtype=np.int32)
    cons = np.nonzero(ibc)[0]
    nbc = np.zeros(nshg, dtype=np.int32)
    nbc[cons] = np.arange(1, cons.size + 1, dtype=np.int32)
    bcinp = _bc_to_input(ibc[cons], np.asarray(case.BC)[cons], ndof)
    with File(path, "write", iotype) as f:
        f.write("byteorder magic number", [1], [MAGIC])
        for key, v in (("number of nodes", nshg), ("number of modes", nshg),
                       ("number of interior elements", nel),
                       ("number of boundary elements", nelb),
                       ("maximum number of element nodes", 4),
                       ("number of interior tpblocks", 1), ("number of boundary tpblocks", 1),
                       ("number of nodes with Dirichlet BCs", cons.size),
                       ("number of shape functions", 4)):
            f.write(key, [v])
        f.write("co-ordinates", [nshg, 3], np.asfortranarray(case.x), "double")
        # genblk.f:90-99 header: neltp nenl ipordl nshl ijunk ijunk lcsyst
        f.write("connectivity interior linear tetrahedron", [nel, 4, 1, 4, 20, 20, 1],
                np.asfortranarray(case.ien))
        # genbkb.f header: neltp nenl ipordl nshl nshlb nenbl lcsyst numnbc
        f.write("connectivity boundary linear tetrahedron", [nelb, 4, 1, 4, 3, 3, 1, 6],
                np.asfortranarray(case.ienb))
        f.write("nbc codes linear tetrahedron", [nelb, 2], np.zeros((nelb, 2), np.int32))
        f.write("nbc values linear tetrahedron", [nelb, 6], np.zeros((nelb, 6)), "double")
        f.write("bc mapping array", [nshg], nbc)
        f.write("bc codes array", [cons.size], ibc[cons])
        f.write("boundary condition array", [cons.size * (ndof + 7)], bcinp, "double")
        f.write("periodic masters array", [nshg], np.asarray(case.iper, dtype=np.int32))


def read_geombc(path, iotype="binary"):
    """The raw geombc fields (readnblk.f:114-376): x (nshg,3), ien (nel,4)
    and ienb (nelb,4) 1-based, iBC rebuilt from the bc mapping and codes
    arrays, the condition array BCinp (numpbc, ndof+7) and iper."""
    out = {}
    with File(path, "read", iotype) as f:
        nshg = int(f.header("number of nodes", 1)[0])
        h = f.header("co-ordinates", 2, "double")
        out["x"] = f.block("co-ordinates", int(h[0]) * int(h[1]), "double").reshape(
            (int(h[0]), int(h[1])), order="F")
        h = f.header("connectivity interior linear tetrahedron", 7)
        out["ien"] = f.block("connectivity interior linear tetrahedron",
                             int(h[0]) * int(h[3])).reshape((int(h[0]), int(h[3])), order="F")
        h = f.header("connectivity boundary linear tetrahedron", 8)
        out["ienb"] = f.block("connectivity boundary linear tetrahedron",
                              int(h[0]) * int(h[3])).reshape((int(h[0]), int(h[3])), order="F")
        h = f.header("bc mapping array", 1)
        nbc = f.block("bc mapping array", int(h[0]))
        h = f.header("bc codes array", 1)
        codes = f.block("bc codes array", int(h[0]))
        h = f.header("boundary condition array", 1, "double")
        numpbc = codes.size
        bcinp = f.block("boundary condition array", int(h[0]), "double")
        out["BCinp"] = bcinp.reshape((numpbc, -1), order="F") if numpbc else bcinp
        h = f.header("periodic masters array", 1)
        out["iper"] = f.block("periodic masters array", int(h[0]))
    # readnblk.f:296-310: iBC(n) = iBCtmp(nBC(n)) where nBC(n) /= 0
    ibc = np.zeros(nshg, dtype=np.int32)
    m = nbc > 0
    ibc[m] = codes[nbc[m] - 1]
    out["iBC"] = ibc
    out["nBC"] = nbc
    return out


def _check_codes(ibc):
    vel = (ibc >> 3) & 7
    bad = (ibc & 1) | ((vel != 0) & (vel != 7)) | (ibc & ~(2 | 4 | 56 | 1024))
    if np.any(bad):
        n = int(np.nonzero(bad)[0][0])
        raise PhastaIOError(f"boundary code {int(ibc[n])} at node {n + 1} is not supported "
                            "(temperature, pressure, all three velocity components, periodic)")


def _bc_to_input(ibc, bc, ndof):
    """Condition records (numpbc, ndof+7) that gen_bc turns back into bc:
    column 2 temperature, 3 pressure, 4-6 velocity direction, 7 its
    magnitude (the columns genbc1.f reads)."""
    _check_codes(ibc)
    out = np.zeros((ibc.size, ndof + 7))
    if ibc.size == 0:
        return out
    out[:, 1] = np.where(ibc & 2, bc[:, 1], 0.0)
    out[:, 2] = np.where(ibc & 4, bc[:, 0], 0.0)
    v = bc[:, 2:5]
    mag = np.sqrt(np.sum(v * v, axis=1))
    v7 = (((ibc >> 3) & 7) == 7) & (mag > 0)
    out[:, 3] = 1.0
    out[v7, 3:6] = v[v7] / mag[v7, None]
    out[:, 6] = np.where(v7, mag, 0.0)
    return out


def gen_bc(ibc, nbc, bcinp, ndofBC=10):
    """genBC1 (genbc1.f) for the supported codes: BCtmp(n,:) =
    BCinp(nBC(n),:) (genbc.f:12-16); temperature BC(:,2) = BCtmp(:,2),
    pressure BC(:,1) = BCtmp(:,3), velocity code 7 BC(:,3:5) = BCtmp(:,7)
    BCtmp(:,4:6).  Other codes raise (the solver refuses them too)."""
    ibc = np.asarray(ibc, dtype=np.int32)
    _check_codes(ibc)
    nshg = ibc.size
    tmp = np.zeros((nshg, bcinp.shape[1] if bcinp.ndim == 2 else 12))
    m = nbc > 0
    if np.any(m):
        tmp[m] = bcinp[nbc[m] - 1]
    bc = np.zeros((nshg, ndofBC), order="F")
    t = (ibc & 2) != 0
    bc[t, 1] = tmp[t, 1]
    p = (ibc & 4) != 0
    bc[p, 0] = tmp[p, 2]
    v7 = ((ibc >> 3) & 7) == 7
    bc[v7, 2:5] = tmp[v7, 6:7] * tmp[v7, 3:6]
    return bc


def case_from_geombc(path, params, restart=None, ndofBC=10, iotype="binary"):
    """A solver case from a geombc file: mesh arrays, BC by gen_bc, the
    CSR graph of the interior connectivity (genadj.f), and yold from the
    "solution" field of a restart file (zeros without one)."""
    from . import mesh
    g = read_geombc(path, iotype)
    nshg = g["x"].shape[0]
    ien = np.asfortranarray(g["ien"], dtype=np.int32)
    lib = mesh._lib()
    colm = np.zeros(nshg + 1, dtype=np.int32)
    nnz = ctypes.c_longlong()
    if lib.phm_csr_count(nshg, ien.shape[0], ien, colm, ctypes.byref(nnz)):
        raise PhastaIOError("phm_csr_count failed")
    rowp = np.zeros(nnz.value, dtype=np.int32)
    if lib.phm_csr_fill(nshg, ien.shape[0], ien, colm, rowp):
        raise PhastaIOError("phm_csr_fill failed")
    bc = gen_bc(g["iBC"], g["nBC"], g["BCinp"], ndofBC)
    yold = np.zeros((nshg, 5), order="F")
    if restart is not None:
        q, _, _ = read_restart(restart, iotype)
        yold = np.asfortranarray(q[:, :5])
    return mesh.BoxCase(N=0, mach=0.0, x=np.asfortranarray(g["x"]), ien=ien,
                        ienb=np.asfortranarray(g["ienb"], dtype=np.int32),
                        iBC=g["iBC"], BC=bc, iper=g["iper"], colm=colm, rowp=rowp,
                        yold=yold, params=dict(params))


def compare_restarts(left, right, tolerance=1e-6):
    """checkphasta (checkphasta.cpp:96-132) over pairs of restart files:
    the largest |difference| of the solution fields, and whether it is
    within the tolerance."""
    worst = 0.0
    for lp, rp in zip(left, right):
        lq, _, _ = read_restart(lp)
        rq, _, _ = read_restart(rp)
        if lq.size != rq.size:
            raise PhastaIOError(f"solution sizes differ: {lq.size}, {rq.size}")
        worst = max(worst, float(np.max(np.abs(lq - rq))) if lq.size else 0.0)
    return worst, worst <= tolerance
