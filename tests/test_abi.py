"""The C-ABI library loads on a host without a GPU and exports every entry
point that include/mvoro.h declares (no compute calls here)."""
import ctypes as C
import os
import re

from meshless_voro_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "mvoro.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = set(re.findall(r"\b(mv_[a-z0-9_]+)\s*\(", src))
    return sorted(names)


def test_header_matches_binding_list():
    assert set(declared_functions()) == set(_native.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(_native.LIB_PATH)
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_status_strings_without_gpu():
    L = _native.lib()
    for s in range(10):
        assert L.mv_status_string(s)
