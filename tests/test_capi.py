"""CPU checks of the C ABI library: it loads (no GPU needed) and exports every
entry point include/tptfem.h declares; the Python binding covers them all."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "include", "tptfem.h")
LIB = os.path.join(ROOT, "tpt-fem_amd", "lib", "libtptfem.so")


def declared():
    src = open(HDR).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(tfm_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = declared()
    assert "tfm_assemble" in names and "tfm_system_solve" in names and len(names) >= 30


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    import tptfem

    assert sorted(tptfem.EXPORTED) == declared()
