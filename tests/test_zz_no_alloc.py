"""mv_build_device_noalloc: the allocation-free variant of mv_build_device
(no counterpart in the reference).  Its results, written to caller-owned
device arrays, equal those of mv_build_device on the same seeds -- checked
on two calls in a row, the second with a smaller input (the static scratch
keeps the previous call's contents) -- and the library makes no hipMalloc
call during them.  Inputs beyond the scratch's maximum, output arrays that
are too short and kept-cell builds are refused with their status codes."""
import ctypes as C

import numpy as np
import pytest

from meshless_voro_amd import _native, workloads as W
from meshless_voro_amd.voronoi import BuildOptions, MODE_LEARNED, MODE_PLAIN, Voronoi
from tests.helpers import REL_TOL

MV_E_ARG, MV_E_SCRATCH, MV_E_OUT_SMALL = 1, 12, 13
ZERO, ONE = np.zeros(3), np.ones(3)
CELL = [("cell_loc", 3, np.float64), ("cell_centroid", 3, np.float64),
        ("cell_volume", 1, np.float64), ("cell_safety_radius", 1, np.float64),
        ("cell_face_offset", 1, np.uint64), ("cell_face_count", 1, np.uint64),
        ("cell_idx", 1, np.uint64)]
FACE = [("face_left", 1, np.uint64), ("face_right", 1, np.int64), ("face_area", 1, np.float64),
        ("face_centroid", 3, np.float64), ("face_normal", 3, np.float64),
        ("face_has_shift", 1, np.uint8), ("face_shift", 3, np.float64)]
# the Voronoi attribute each output array is compared with
ATTR = {"cell_face_connections": "links"}


def test_refuses_inputs_beyond_the_scratch():
    """Needs no GPU: the size check comes before any device call."""
    L = _native.lib()
    out = _native.mv_noalloc_out()
    a, w = np.zeros(3), np.ones(3)
    n = _native.NOALLOC_MAX_SEEDS + 1
    st = L.mv_build_device_noalloc(C.c_void_p(16), n, None, a.ctypes.data, w.ctypes.data, 3, 0,
                                   None, C.byref(out))
    assert st == MV_E_SCRATCH
    assert b"scratch" in L.mv_status_string(st)
    opt = BuildOptions().to_c()
    opt.flags |= 2  # MV_FLAG_KEEP_CELLS
    st = L.mv_build_device_noalloc(C.c_void_p(16), 10, None, a.ctypes.data, w.ctypes.data, 3, 0,
                                   C.byref(opt), C.byref(out))
    assert st == MV_E_ARG


def _alloc_outputs(torch, n, F, NL):
    bufs = {}
    for name, k, dt in CELL:
        bufs[name] = torch.empty(k * n, dtype=getattr(torch, np.dtype(dt).name), device="cuda:0")
    for name, k, dt in FACE:
        bufs[name] = torch.empty(k * F, dtype=getattr(torch, np.dtype(dt).name), device="cuda:0")
    bufs["cell_face_connections"] = torch.empty(NL, dtype=torch.uint64, device="cuda:0")
    out = _native.mv_noalloc_out()
    for name, t in bufs.items():
        setattr(out, name, t.data_ptr())
    out.face_cap, out.link_cap = F, NL
    return bufs, out


def _check_same(bufs, ref, n):
    for name, t in bufs.items():
        k = dict((c[0], c[1]) for c in CELL + FACE).get(name, 1)
        got = t.cpu().numpy()
        want = getattr(ref, ATTR.get(name, name)).reshape(-1)
        assert got.shape == want.shape, name
        if got.dtype.kind == "f":
            lim = REL_TOL * np.maximum(np.abs(want), 1e-300)
            assert np.all(np.abs(got - want) <= lim), name
        else:
            assert np.array_equal(got, want), name


@pytest.mark.gpu
def test_noalloc_equals_build_device_two_calls():
    import torch
    L = _native.lib()
    cases = [(W.uniform(20000, 61), True, MODE_LEARNED), (W.uniform(3000, 62), False, MODE_PLAIN)]
    for x, periodic, mode in cases:
        n = len(x)
        print(f"case noalloc n={n} periodic={periodic} mode={mode}", flush=True)
        xt = torch.from_numpy(np.ascontiguousarray(x)).to("cuda:0")
        opt = BuildOptions(mode=mode).to_c()
        h = C.c_void_p()
        assert L.mv_build_device(C.c_void_p(xt.data_ptr()), n, None, ZERO.ctypes.data,
                                 ONE.ctypes.data, 3, int(periodic), C.byref(opt),
                                 C.byref(h)) == 0
        ref = Voronoi._from_handle(h)
        F, NL = len(ref.face_area), len(ref.links)
        # too short: nothing written, the counts come back
        small, out = _alloc_outputs(torch, n, F - 1, NL)
        torch.cuda.synchronize()
        before = L.mv_test_device_mallocs()
        st = L.mv_build_device_noalloc(C.c_void_p(xt.data_ptr()), n, None, ZERO.ctypes.data,
                                       ONE.ctypes.data, 3, int(periodic), C.byref(opt),
                                       C.byref(out))
        assert st == MV_E_OUT_SMALL
        assert (out.n_faces, out.n_links) == (F, NL)
        bufs, out = _alloc_outputs(torch, n, F, NL)
        torch.cuda.synchronize()
        st = L.mv_build_device_noalloc(C.c_void_p(xt.data_ptr()), n, None, ZERO.ctypes.data,
                                       ONE.ctypes.data, 3, int(periodic), C.byref(opt),
                                       C.byref(out))
        assert st == 0, L.mv_status_string(st)
        assert L.mv_test_device_mallocs() == before
        assert (out.n_faces, out.n_links) == (F, NL)
        assert 0 < out.scratch_used <= _native.NOALLOC_SCRATCH_BYTES
        print(f"  scratch used {out.scratch_used / 2**20:.1f} MiB", flush=True)
        _check_same(bufs, ref, n)


@pytest.mark.gpu
def test_noalloc_partial_and_empty():
    import torch
    L = _native.lib()
    x = W.uniform(5000, 63)
    m = W.partial_mask(5000, 0.1, 64).astype(np.uint8)
    xt = torch.from_numpy(x).to("cuda:0")
    mt = torch.from_numpy(m).to("cuda:0")
    ref = Voronoi.build_partial(x, m.astype(bool), ZERO, ONE, 3, True)
    bufs, out = _alloc_outputs(torch, 5000, len(ref.face_area), len(ref.links))
    st = L.mv_build_device_noalloc(C.c_void_p(xt.data_ptr()), 5000, C.c_void_p(mt.data_ptr()),
                                   ZERO.ctypes.data, ONE.ctypes.data, 3, 1, None, C.byref(out))
    assert st == 0
    _check_same(bufs, ref, 5000)
    out = _native.mv_noalloc_out()
    assert L.mv_build_device_noalloc(None, 0, None, ZERO.ctypes.data, ONE.ctypes.data, 3, 0,
                                     None, C.byref(out)) == 0
    assert (out.n_faces, out.n_links) == (0, 0)
