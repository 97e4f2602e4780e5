"""ctypes binding of the C-ABI library ``libgospl_hip.so``.

Every exported symbol is declared here with its argument types.  There is no
fallback: if the library is missing the import fails with an error that says
how to build it.
"""

from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgospl_hip.so")

_lib = None

P = ctypes.c_void_p
I64 = ctypes.c_int64
I32 = ctypes.c_int32
F64 = ctypes.c_double

# name -> (restype, argtypes).  Mirrors include/gospl_hip.h.
SIGNATURES = {
    "gospl_fv_geometry": (I32, [P, I64, P, P, I64, P, I64, P, I32, P, P, P, P, P]),
}


def _declare(lib):
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def load(path: str | None = None):
    """Load (once) and return the library handle."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            "libgospl_hip.so not found at %s: build it with "
            "`python -c 'import __graft_entry__ as g; g.build()'` "
            "(or `python -m gospl_amd.build`)" % p)
    lib = ctypes.CDLL(p)
    _declare(lib)
    if path is None:
        _lib = lib
    return lib
