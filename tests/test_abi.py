"""The C-ABI library loads without a GPU and exports every declared symbol."""
import ctypes
import os
import re

from phasta_amd import _native, solver

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header, prefix=r"ph[gm]_"):
    txt = open(os.path.join(ROOT, "include", header)).read()
    return set(re.findall(r"^\s*(?:int|void|const char\*)\s+(" + prefix + r"\w+)\(", txt, re.M))


def test_phgpu_exports_every_declared_symbol():
    names = declared("phgpu.h")
    assert names == set(solver.EXPORTS)
    lib = _native.load("libphgpu.so")
    for n in names:
        assert hasattr(lib, n), n


def test_phmesh_exports_every_declared_symbol():
    lib = _native.load("libphmesh.so")
    for n in declared("phmesh.h"):
        assert hasattr(lib, n), n


def test_phastaio_exports_every_declared_symbol():
    names = declared("phastaio.h", "phastaio_")
    assert len(names) == 11
    lib = _native.load("libphmesh.so")
    for n in names:
        assert hasattr(lib, n), n


def test_every_header_is_covered():
    assert sorted(os.listdir(os.path.join(ROOT, "include"))) == ["phastaio.h", "phgpu.h",
                                                                 "phmesh.h"]


def test_last_error_is_callable_without_gpu():
    lib = solver.lib()
    assert isinstance(lib.phg_last_error(), bytes)
