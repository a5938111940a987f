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
    assert lib.pmp_abi_version() == 2


def test_extension_matches_library():
    import plainmp_amd

    assert plainmp_amd._native.abi_version() == 2


def test_cloud_tile_sizes_and_argument_errors():
    """pmp_cloud_tile_sizes is host arithmetic (tiles of 64 points, super-tiles
    of 16 tiles, keys padded to a power of two >= 512); out-of-range counts
    and a null context are rejected with a message, before any GPU call."""
    import plainmp_amd

    lib = ctypes.CDLL(plainmp_amd.LIBRARY_PATH)
    lib.pmp_last_error.restype = ctypes.c_char_p
    i64 = ctypes.c_int64
    for n, tiles, supers, keys in [(1, 1, 1, 512), (64, 1, 1, 512), (65, 2, 1, 512), (1025, 17, 2, 2048),
                                   (1 << 20, 16384, 1024, 1 << 20)]:
        nt, ns, scr = i64(), i64(), i64()
        assert lib.pmp_cloud_tile_sizes(i64(n), ctypes.byref(nt), ctypes.byref(ns), ctypes.byref(scr)) == 0
        assert (nt.value, ns.value) == (tiles, supers)
        assert scr.value >= 8 * keys
    for n in (0, -5, 1 << 30):
        assert lib.pmp_cloud_tile_sizes(i64(n), None, None, None) != 0
        assert b"point count" in lib.pmp_last_error()
    buf = ctypes.c_void_p(1)
    assert lib.pmp_cloud_build_tiles(None, buf, i64(10), buf, buf, buf, buf, buf, buf) != 0
    assert b"null context" in lib.pmp_last_error()
