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
    for s in range(14):
        assert L.mv_status_string(s) != b"unknown status"


def test_struct_layouts_match_header(tmp_path):
    """Field offsets and sizes of the ctypes mirrors equal those the C
    compiler gives the header's structs."""
    import subprocess
    structs = {"mv_options": _native.mv_options, "mv_stats": _native.mv_stats,
               "mv_model": _native.mv_model, "mv_train_config": _native.mv_train_config,
               "mv_device_view": _native.mv_device_view,
               "mv_cell_geometry": _native.mv_cell_geometry,
               "mv_noalloc_out": _native.mv_noalloc_out}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "mvoro.h"', "int main(void){"]
    lines.append('printf("scratch bytes %llu\\n", (unsigned long long)MV_NOALLOC_SCRATCH_BYTES);')
    lines.append('printf("scratch seeds %u\\n", MV_NOALLOC_MAX_SEEDS);')
    for sn, cls in structs.items():
        lines.append(f'printf("{sn} size %zu\\n", sizeof({sn}));')
        for f, _ in cls._fields_:
            lines.append(f'printf("{sn} {f} %zu\\n", offsetof({sn}, {f}));')
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    got = {}
    for line in out.splitlines():
        sn, f, v = line.split()
        got[(sn, f)] = int(v)
    assert got[("scratch", "bytes")] == _native.NOALLOC_SCRATCH_BYTES
    assert got[("scratch", "seeds")] == _native.NOALLOC_MAX_SEEDS
    for sn, cls in structs.items():
        assert got[(sn, "size")] == C.sizeof(cls), sn
        for f, _ in cls._fields_:
            assert got[(sn, f)] == getattr(cls, f).offset, (sn, f)


def test_download_threads_fall_back_to_the_calling_thread():
    """The host download's fan-out (copy_to_host): when a thread cannot be
    started, its share and every later one run on the calling thread, the
    started threads are joined, and the call returns MV_OK instead of
    letting an exception cross the C ABI (std::terminate would abort the
    process).  Exercised through the test hook without a GPU."""
    import numpy as np
    L = _native.lib()
    for T, fail_after in ((16, -1), (16, 0), (16, 5), (16, 15), (3, 1), (0, -1)):
        out = np.zeros(max(T, 1), np.int32)
        assert L.mv_test_run_shares(T, fail_after, out.ctypes.data) == 0
        assert out[:T].tolist() == list(range(1, T + 1))
    assert L.mv_test_run_shares(-1, -1, None) == 1
