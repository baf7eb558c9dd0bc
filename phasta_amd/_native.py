"""Loading of the in-tree native libraries.

``libphmesh.so`` is the host-only box-mesh generator (g++), ``libphgpu.so``
the HIP library behind the C-ABI declared in ``include/phgpu.h``.  Both are
built in-tree by ``__graft_entry__.build()``; nothing here falls back to a
Python or CPU implementation: a missing library raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}


def lib_path(name):
    return os.path.join(_HERE, name)


def load(name):
    """Load ``phasta_amd/<name>`` once; raise with a build hint if absent."""
    if name in _LIBS:
        return _LIBS[name]
    path = lib_path(name)
    if not os.path.exists(path):
        raise RuntimeError(
            f"{name} is not built ({path} missing); run "
            "`python -c 'import __graft_entry__ as g; g.build()'` first")
    lib = ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
    _LIBS[name] = lib
    return lib
