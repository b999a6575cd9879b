"""Locates and loads the in-tree HIP library (never a CPU fallback)."""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libmci_amd.so")
_handle = None


def library_path():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"the HIP library is not built: {LIB_PATH} (run python -m mcintegratorplusplus_amd.build)")
    return LIB_PATH


def load():
    global _handle
    if _handle is None:
        _handle = ctypes.CDLL(library_path())
    return _handle
