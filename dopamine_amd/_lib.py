"""Loader for the native library libdopamine_hip.so (HIP kernels + mesh setup + C ABI).

The library is built in-tree by ``__graft_entry__.build()`` (``make -C dopamine_amd/csrc``).
There is no Python or CPU fallback for anything the library does: if it is missing, importing
the product raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdopamine_hip.so")

_lib = None


def lib():
    """Return the loaded ctypes handle, loading it on first use (raises if not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                "(make -C dopamine_amd/csrc)")
        _lib = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
    return _lib
