"""The C-ABI library loads and exports every symbol include/mci_amd.h declares (no GPU needed)."""
import ctypes
import os
import re

from mcitest import ROOT


def declared():
    src = open(os.path.join(ROOT, "include", "mci_amd.h")).read()
    return sorted(set(re.findall(r"\b(mci_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    from mcintegratorplusplus_amd import _lib
    lib = ctypes.CDLL(_lib.library_path())
    names = declared()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_library_reports_errors_without_a_device():
    from mcintegratorplusplus_amd import _lib
    lib = ctypes.CDLL(_lib.library_path())
    lib.mci_last_error.restype = ctypes.c_char_p
    lib.mci_version.restype = ctypes.c_char_p
    assert b"gfx950" in lib.mci_version()
    h = ctypes.c_void_p()
    assert lib.mci_create(0, 1, ctypes.byref(h)) == 1  # invalid_argument before touching a device
    assert b"dimensions" in lib.mci_last_error()
