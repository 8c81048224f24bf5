"""The host copies of a build's arrays (mv_cell_* / mv_face_* accessors).

Above 32 MiB of output the library copies through pinned staging on several
threads; these tests compare those bytes with one plain hipMemcpy per array
from the device view (mv_get_device_view) of the same handle, and a host
build (mv_build) with a device build (mv_build_device) of the same seeds.
"""
import ctypes as C

import numpy as np
import pytest

from meshless_voro_amd import _native, workloads as W

pytestmark = pytest.mark.gpu

# (view field, accessor, element ctype, numpy dtype, elements per cell/face/link)
ARRAYS = [
    ("cell_loc", "mv_cell_loc", C.c_double, np.float64, "3n"),
    ("cell_centroid", "mv_cell_centroid", C.c_double, np.float64, "3n"),
    ("cell_volume", "mv_cell_volume", C.c_double, np.float64, "n"),
    ("cell_safety_radius", "mv_cell_safety_radius", C.c_double, np.float64, "n"),
    ("cell_face_offset", "mv_cell_face_offset", C.c_uint64, np.uint64, "n"),
    ("cell_face_count", "mv_cell_face_count", C.c_uint64, np.uint64, "n"),
    ("cell_idx", "mv_cell_idx", C.c_uint64, np.uint64, "n"),
    ("face_left", "mv_face_left", C.c_uint64, np.uint64, "F"),
    ("face_right", "mv_face_right", C.c_int64, np.int64, "F"),
    ("face_area", "mv_face_area", C.c_double, np.float64, "F"),
    ("face_centroid", "mv_face_centroid", C.c_double, np.float64, "3F"),
    ("face_normal", "mv_face_normal", C.c_double, np.float64, "3F"),
    ("face_has_shift", "mv_face_has_shift", C.c_uint8, np.uint8, "F"),
    ("face_shift", "mv_face_shift", C.c_double, np.float64, "3F"),
    ("cell_face_connections", "mv_cell_face_connections", C.c_uint64, np.uint64, "NL"),
]


def hip_runtime():
    """The HIP runtime libmvoro.so is linked against (already loaded)."""
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipMemcpy.restype = C.c_int
    return hip


def counts(L, h):
    n, F, NL = L.mv_n_cells(h), L.mv_n_faces(h), L.mv_n_links(h)
    return {"n": n, "3n": 3 * n, "F": F, "3F": 3 * F, "NL": NL}


def device_arrays(L, h):
    hip = hip_runtime()
    v = _native.mv_device_view()
    L.mv_get_device_view(h, C.byref(v))
    cnt = counts(L, h)
    out = {}
    for field, _, _, dt, k in ARRAYS:
        a = np.empty(cnt[k], dtype=dt)
        src = getattr(v, field)
        if a.size:
            assert src
            assert hip.hipMemcpy(a.ctypes.data, src, a.nbytes, 2) == 0, field  # D2H
        out[field] = a
    return out


def host_arrays(L, h):
    cnt = counts(L, h)
    out = {}
    for field, acc, ct, dt, k in ARRAYS:
        p = getattr(L, acc)(h)
        if cnt[k]:
            assert p, acc
            out[field] = np.ctypeslib.as_array(C.cast(p, C.POINTER(ct)), shape=(cnt[k],)).copy()
        else:
            out[field] = np.empty(0, dtype=dt)
    return out


def output_bytes(arrs):
    return sum(a.nbytes for a in arrs.values())


def device_build(L, x, periodic):
    import torch
    xt = torch.from_numpy(np.ascontiguousarray(x)).to("cuda:0")
    h = C.c_void_p()
    a, w = np.zeros(3), np.ones(3)
    st = L.mv_build_device(C.c_void_p(xt.data_ptr()), len(x), None, a.ctypes.data,
                           w.ctypes.data, 3, int(periodic), None, C.byref(h))
    torch.cuda.synchronize()
    assert st == 0, st
    return h


def host_build(L, x, periodic):
    h = C.c_void_p()
    a, w = np.zeros(3), np.ones(3)
    xc = np.ascontiguousarray(x, dtype=np.float64)
    st = L.mv_build(xc.ctypes.data, len(x), None, a.ctypes.data, w.ctypes.data, 3,
                    int(periodic), None, C.byref(h))
    assert st == 0, st
    return h


@pytest.mark.parametrize("n,periodic", [(400000, True), (150000, False), (3000, True)])
def test_host_copies_equal_device_arrays(n, periodic):
    L = _native.lib()
    x = W.uniform(n, 3100 + n % 7)
    h = device_build(L, x, periodic)
    try:
        dev = device_arrays(L, h)
        host = host_arrays(L, h)
    finally:
        L.mv_free(h)
    if n >= 150000:
        assert output_bytes(dev) > 32 << 20  # the staged, threaded copy
    for field in dev:
        assert np.array_equal(host[field], dev[field]), field


def test_host_build_equals_device_build():
    L = _native.lib()
    x = W.uniform(300000, 3200)
    hd = device_build(L, x, True)
    hh = host_build(L, x, True)
    try:
        a, b = host_arrays(L, hd), host_arrays(L, hh)
    finally:
        L.mv_free(hd)
        L.mv_free(hh)
    for field in a:
        assert np.array_equal(a[field], b[field]), field
