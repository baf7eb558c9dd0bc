"""The phastaIO file format (include/phastaio.h, phasta_amd/phastaio.py).

The reference's phastaIO.cc includes the CMake-generated FCMangle.h, so it
is not built here; the format is pinned by the byte layout its POSIX
writer produces (writeHeader phastaIO.cc:1624-1644: "<key> : < <bytes> >
", each header integer as "%d ", "\\n"; writeDataBlock :1819-1850: the raw
items then "\\n", or one "%d\\n" / "%f\\n" line per item in ASCII), by a
big-endian file assembled byte by byte, and by round trips.
"""
import os
import struct

import numpy as np
import pytest

import pyoracle
from phasta_amd import mesh, phastaio as pio

MAGIC = b"byteorder magic number : < 5 > 1 \n"


def test_binary_bytes_known_answer(tmp_path):
    p = tmp_path / "restart.7.1"
    y = np.arange(6, dtype=np.float64).reshape((2, 3), order="F") + 0.25
    pio.write_restart(p, y, -y, 7)
    data = p.read_bytes()
    body = y.ravel(order="F").tobytes()
    want = (MAGIC + struct.pack("<i", 362436) + b"\n"
            + b"solution : < 49 > 2 3 7 \n" + body + b"\n"
            + b"time derivative of solution : < 49 > 2 3 7 \n" + (-y).ravel(order="F").tobytes()
            + b"\n")
    assert data == want


def test_ascii_bytes_known_answer(tmp_path):
    p = tmp_path / "a.dat"
    with pio.File(p, "write", "ascii") as f:
        f.write("number of nodes", [12])
        f.write("solution", [2, 1, 3], [1.5, -2.0], "double")
        f.write("bc codes array", [2], [56, 1080])
    assert p.read_bytes() == (b"number of nodes : < 0 > 12 \n"
                              b"solution : < 2 > 2 1 3 \n1.500000\n-2.000000\n"
                              b"bc codes array : < 2 > 2 \n56\n1080\n")
    with pio.File(p, "read", "ascii") as f:
        h = f.header("solution", 3, "double")
        assert list(h) == [2, 1, 3]
        assert np.array_equal(f.block("solution", 2, "double"), [1.5, -2.0])
        assert list(f.header("bc codes", 1)) == [2]
        assert np.array_equal(f.block("bc codes array", 2), [56, 1080])


def test_restart_round_trip_and_search(tmp_path):
    rng = np.random.default_rng(0)
    y = np.asfortranarray(rng.standard_normal((57, 5)))
    ac = np.asfortranarray(rng.standard_normal((57, 5)))
    p = tmp_path / pio.restart_name(40)
    pio.write_restart(p, y, ac, 40)
    q, a, step = pio.read_restart(p)
    assert step == 40 and np.array_equal(q, y) and np.array_equal(a, ac)
    with pio.File(p, "read") as f:
        # out of order: the reader wraps to the start once
        assert list(f.header("time derivative of solution", 3, "double")) == [57, 5, 40]
        assert list(f.header("SOLUTION", 3, "double")) == [57, 5, 40]      # case-insensitive
        assert np.array_equal(f.block("solution", 57 * 5, "double"), y.ravel(order="F"))
        assert list(f.header("time derivative", 3, "double")) == [57, 5, 40]  # key prefix
        assert f.header("errors", 1) is None


def test_comments_and_unknown_records_are_skipped(tmp_path):
    p = tmp_path / "g.dat"
    with pio.File(p, "write") as f:
        lib = pio.lib()
        lib.phastaio_writestring(pio.ctypes.byref(f.fd), b"# PHASTA Input File Version 2.0\n")
        f.write("byteorder magic number", [1], [362436])
        f.write("some other field", [3], [1.0, 2.0, 3.0], "double")
        f.write("number of nodes", [9])
    with pio.File(p, "read") as f:
        assert list(f.header("number of nodes", 1)) == [9]
        assert list(f.header("some other field", 1, "double")) == [3]


def test_big_endian_file_is_swapped(tmp_path):
    p = tmp_path / "restart.3.1"
    vals = np.array([1.0, -2.5, 3.25, 1e-300])
    p.write_bytes(MAGIC + struct.pack(">i", 362436) + b"\n"
                  + b"solution : < 33 > 2 2 3 \n" + struct.pack(">4d", *vals) + b"\n")
    q, ac, step = pio.read_restart(p)
    assert step == 3 and ac is None
    assert np.array_equal(q.ravel(order="F"), vals)


def test_strict_mode_refuses_a_mismatched_block(tmp_path):
    p = tmp_path / "r.dat"
    pio.write_restart(p, np.ones((2, 1)), np.ones((2, 1)), 1)
    pio.lib().phastaio_togglestrictmode()
    try:
        with pio.File(p, "read") as f:
            f.header("solution", 3, "double")
            with pytest.raises(pio.PhastaIOError):
                f.block("time derivative of solution", 2, "double")
    finally:
        pio.lib().phastaio_togglestrictmode()


def test_geombc_round_trip(tmp_path):
    case = mesh.box_case(3, jitter=0.1)
    p = tmp_path / pio.geombc_name()
    pio.write_geombc(p, case)
    g = pio.read_geombc(p)
    assert np.array_equal(g["x"], case.x)
    assert np.array_equal(g["ien"], case.ien) and np.array_equal(g["ienb"], case.ienb)
    assert np.array_equal(g["iBC"], case.iBC) and np.array_equal(g["iper"], case.iper)
    assert g["BCinp"].shape == (int(np.count_nonzero(case.iBC)), 12)


def test_compare_restarts_tolerance(tmp_path):
    y = np.linspace(0.0, 1.0, 20).reshape((4, 5), order="F")
    a, b = tmp_path / "a.1", tmp_path / "b.1"
    pio.write_restart(a, y, y, 1)
    pio.write_restart(b, y + 5e-7, y, 1)
    d, ok = pio.compare_restarts([a], [b])
    assert ok and abs(d - 5e-7) < 1e-12
    d, ok = pio.compare_restarts([a], [b], tolerance=1e-7)
    assert not ok


def test_gen_bc_formulas():
    """genBC1 (genbc1.f) on hand-made condition records: temperature from
    column 2, pressure from column 3, velocity code 7 = magnitude (column
    7) times direction (columns 4-6); unsupported codes raise."""
    ibc = np.array([0, 2, 4, 56, 56 | 1024 | 2, 1024], dtype=np.int32)
    nbc = np.array([0, 1, 2, 3, 4, 5], dtype=np.int32)
    rec = np.zeros((5, 12))
    rec[0, 1] = 310.0
    rec[1, 2] = 1.2e5
    rec[2, 3:7] = [0.6, 0.8, 0.0, 5.0]
    rec[3, 1] = 290.0
    rec[3, 3:7] = [0.0, 0.0, 1.0, 2.0]
    bc = pio.gen_bc(ibc, nbc, rec)
    assert bc[1, 1] == 310.0 and bc[2, 0] == 1.2e5
    assert np.allclose(bc[3, 2:5], [3.0, 4.0, 0.0], rtol=0, atol=1e-15)
    assert bc[4, 1] == 290.0 and np.array_equal(bc[4, 2:5], [0.0, 0.0, 2.0])
    assert not bc[0].any() and not bc[5].any()
    with pytest.raises(pio.PhastaIOError):
        pio.gen_bc(np.array([8], dtype=np.int32), np.array([1], dtype=np.int32), rec[:1])


def test_case_from_geombc_reproduces_the_case(tmp_path):
    """geombc + restart written from a box case (top wall moving at
    10 m/s) and read back into a solver case: the same arrays, and the
    oracle's Newton step on it equals the one on the original, bit for bit."""
    case = mesh.box_case(4, jitter=0.1, kspace=10)
    top = np.isclose(case.x[:, 1], 1.0)
    case.BC[top, 2] = 10.0
    g = tmp_path / pio.geombc_name()
    r = tmp_path / pio.restart_name(0)
    pio.write_geombc(g, case)
    pio.write_restart(r, case.yold, np.zeros_like(case.yold), 0)
    c2 = pio.case_from_geombc(g, case.params, restart=r)
    for k in ("x", "ien", "ienb", "iBC", "iper", "colm", "rowp", "BC", "yold"):
        assert np.array_equal(getattr(c2, k), getattr(case, k)), k
    acold = np.zeros((case.nshg, 5), order="F")
    y1, ac1, h1 = pyoracle.Oracle(case).newton_step(case.yold, acold)
    y2, ac2, h2 = pyoracle.Oracle(c2).newton_step(c2.yold, acold)
    assert np.array_equal(y1, y2) and h1[0]["iKs"] == h2[0]["iKs"]
    assert np.allclose(y1[top, 0], 10.0)


def test_truncated_and_missing_files_raise(tmp_path):
    p = tmp_path / "restart.1.1"
    y = np.ones((50, 5))
    pio.write_restart(p, y, y, 1)
    data = p.read_bytes()
    p.write_bytes(data[: len(data) // 2])          # cut inside the solution block
    with pytest.raises(pio.PhastaIOError):
        pio.read_restart(p)
    with pytest.raises(pio.PhastaIOError):
        pio.read_restart(tmp_path / "restart.9.1")  # no such file
    with pytest.raises(pio.PhastaIOError):
        pio.File(tmp_path / "x", "modify")           # unknown mode
    (tmp_path / "empty").write_bytes(b"")
    with pio.File(tmp_path / "empty", "read") as f:
        assert f.header("solution", 3, "double") is None


def test_ascii_geombc_round_trip(tmp_path):
    case = mesh.box_case(2, jitter=0.0)
    p = tmp_path / "geombc.dat.1"
    pio.write_geombc(p, case, iotype="ascii")
    g = pio.read_geombc(p, iotype="ascii")
    # ASCII doubles carry six decimals ("%f", as the reference writes them)
    assert np.allclose(g["x"], case.x, atol=5e-7)
    assert np.array_equal(g["ien"], case.ien) and np.array_equal(g["iper"], case.iper)
    assert np.array_equal(g["iBC"], case.iBC)
