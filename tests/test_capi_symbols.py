"""The C ABI library loads and exports every entry point include/plainmp_amd.h declares."""

import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "plainmp_amd.h")).read()
    return sorted(set(re.findall(r"\b(pmp_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_header_symbols():
    import plainmp_amd

    lib = ctypes.CDLL(plainmp_amd.LIBRARY_PATH)
    names = declared()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    lib.pmp_abi_version.restype = ctypes.c_int32
    assert lib.pmp_abi_version() == 1


def test_extension_matches_library():
    import plainmp_amd

    assert plainmp_amd._native.abi_version() == 1
