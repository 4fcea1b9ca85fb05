"""Loads the in-tree HIP library (libcrmech.so) and binds the C ABI.

There is no fallback: if the library is missing or cannot be loaded, importing
the simulation layer raises, so a GPU run can never silently use another path.
"""
from __future__ import annotations

import ctypes as C
import os

from . import _abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcrmech.so")

_lib = None


class NativeLibraryMissing(RuntimeError):
    pass


def load() -> C.CDLL:
    """Returns the bound library, loading it on first use."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise NativeLibraryMissing(
            f"{LIB_PATH} is not built; run __graft_entry__.build() (hipcc --offload-arch=gfx950)"
        )
    lib = C.CDLL(LIB_PATH)
    for name, (restype, argtypes) in _abi.SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = restype
        fn.argtypes = argtypes
    _lib = lib
    return lib


def exported_symbols() -> dict:
    """Symbol -> bool (present) for every entry point the header declares."""
    lib = C.CDLL(LIB_PATH)
    return {name: hasattr(lib, name) for name in _abi.SIGNATURES}
