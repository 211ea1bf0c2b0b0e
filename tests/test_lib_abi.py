"""CPU checks of the drop-in boundary: libfamehip.so loads here (no GPU) and
exports every entry point that include/fame_hip.h declares."""
import ctypes
import os
import re

from fame_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "fame_hip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(fame_\w+)\s*\(", src, re.M)))


def test_header_matches_binding():
    assert declared() == sorted(_lib.EXPORTS)


def test_library_exports_every_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared():
        assert hasattr(lib, name), name


def test_missing_index_is_an_error_not_a_crash():
    lib = _lib.load()
    h = ctypes.c_void_p()
    rc = lib.fame_open(b"/nonexistent/index", 0, ctypes.byref(h))
    assert rc != 0
    if h.value:
        lib.fame_close(h.value)


def test_index_library_exports_every_symbol():
    """include/fame_index.h is the C ABI of libfameindex.so (the builder)."""
    from fame_amd import index as fidx
    src = open(os.path.join(ROOT, "include", "fame_index.h")).read()
    names = sorted(set(re.findall(r"^\s*int\s+(fame_\w+)\s*\(", src, re.M)))
    assert names == ["fame_build_index", "fame_build_index_fasta"]
    lib = ctypes.CDLL(fidx.LIB_PATH)
    for name in names:
        assert hasattr(lib, name), name
