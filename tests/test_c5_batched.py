"""The 512-problem batched workload (workloads.batched_problems(512, 5):
n_p = floor(10^U(3,5)), about 10.6 M seeds; even problems uniform, odd ones
8 wrapped Gaussian clusters with sigma 0.05; periodic iff p % 4 < 2) in one
learned-mode mv_build_batch_device call, full and with a 10 % mask on every
problem (build_partial, src/voronoi.rs:181-190).

On all 512 problems (device arrays reduced with torch): the cell volumes sum
to the unit box within 1e-11, every face is linked once to its left cell and
once more to its right cell iff that is a cell without periodic shift
(Voronoi::finalize, src/voronoi.rs:241-263), and the per-cell link counts
add up to the link array.  A partial build gives every active cell the
bits of the full build and every inactive cell zeros (the per-cell
construction does not depend on the mask; ownership follows
src/voronoi/voronoi_cell.rs:53-73).  On the 8 smallest problems of each
class (uniform/clustered x periodic/rectangular), full and partial, the
connectivity equals the oracle's bit for bit and the integrals agree within
1e-12 relative."""
import ctypes as C

import numpy as np
import pytest

from meshless_voro_amd import _native, workloads as W
from meshless_voro_amd.voronoi import BuildOptions, MODE_LEARNED, VoronoiError, Voronoi
from oracle import oracle as O
from tests.helpers import assert_connectivity_equal, assert_integrals_close, closure_residual

pytestmark = pytest.mark.gpu

P = 512


def hip_memcpy():
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipMemcpy.restype = C.c_int
    return hip.hipMemcpy


@pytest.fixture(scope="module")
def c5():
    import torch
    probs = W.batched_problems(P, 5)
    sizes = np.array([len(x) for x, _ in probs], np.int64)
    offsets = np.zeros(P + 1, np.uint64)
    offsets[1:] = np.cumsum(sizes)
    xyz = np.ascontiguousarray(np.concatenate([x for x, _ in probs]))
    per = np.array([int(p) for _, p in probs], np.uint32)
    masks = [W.partial_mask(int(n), 0.1, 1000 + i) for i, n in enumerate(sizes)]
    mask = np.ascontiguousarray(np.concatenate(masks).astype(np.uint8))
    xt = torch.from_numpy(xyz).to("cuda:0")
    mt = torch.from_numpy(mask).to("cuda:0")
    print(f"case c5 P={P} seeds={int(offsets[-1])} periodic={int(per.sum())}", flush=True)
    return dict(probs=probs, sizes=sizes, offsets=offsets, xyz=xyz, per=per, masks=masks,
                mask=mask, xt=xt, mt=mt)


def batch(c5, use_mask):
    L = _native.lib()
    anchors = np.zeros((P, 3))
    widths = np.ones((P, 3))
    opt = BuildOptions(mode=MODE_LEARNED).to_c()
    hs = (C.c_void_p * P)()
    st = L.mv_build_batch_device(C.c_void_p(c5["xt"].data_ptr()), c5["offsets"].ctypes.data, P,
                                 C.c_void_p(c5["mt"].data_ptr()) if use_mask else None,
                                 anchors.ctypes.data, widths.ctypes.data, c5["per"].ctypes.data,
                                 3, C.byref(opt), hs)
    if st != 0:
        raise VoronoiError(st)
    return [C.c_void_p(hs[p]) for p in range(P)]


def device_tensors(h, fields):
    """Device copies (torch) of some arrays of a handle's device view."""
    import torch
    L = _native.lib()
    cp = hip_memcpy()
    v = _native.mv_device_view()
    L.mv_get_device_view(h, C.byref(v))
    n, F, NL = L.mv_n_cells(h), L.mv_n_faces(h), L.mv_n_links(h)
    spec = {"cell_volume": (n, torch.float64), "cell_centroid": (3 * n, torch.float64),
            "cell_face_count": (n, torch.int64), "face_right": (F, torch.int64),
            "face_has_shift": (F, torch.uint8), "cell_face_connections": (NL, torch.int64)}
    out = {}
    for f in fields:
        cnt, dt = spec[f]
        t = torch.empty(cnt, dtype=dt, device="cuda:0")
        if cnt:
            assert cp(C.c_void_p(t.data_ptr()), getattr(v, f), t.numel() * t.element_size(),
                      3) == 0
        out[f] = t
    return out, (n, F, NL)


def sample_problems(c5):
    """The 8 smallest problems of each (uniform/clustered x periodic/rect)
    class."""
    out = []
    for cls in range(4):
        ps = [p for p in range(P) if p % 4 == cls]
        out += sorted(ps, key=lambda p: c5["sizes"][p])[:8]
    return out


def test_c5_full_properties_and_oracle_sample(c5):
    import torch
    L = _native.lib()
    hs = batch(c5, False)
    try:
        worst = 0.0
        for p, h in enumerate(hs):
            t, (n, F, NL) = device_tensors(h, ("cell_volume", "cell_face_count", "face_right",
                                               "face_has_shift"))
            assert n == c5["sizes"][p]
            tot = float(t["cell_volume"].sum())
            worst = max(worst, abs(tot - 1.0))
            assert abs(tot - 1.0) <= 1e-11, (p, tot)
            assert bool((t["cell_volume"] > 0).all()), p
            linked_right = int(((t["face_right"] >= 0) & (t["face_has_shift"] == 0)).sum())
            assert NL == F + linked_right, p
            assert int(t["cell_face_count"].sum()) == NL, p
            if not c5["per"][p]:
                assert not bool(t["face_has_shift"].any())
        print(f"c5 full: max |sum vol - 1| = {worst:.3g}", flush=True)
        for p in sample_problems(c5):
            x, per = c5["probs"][p]
            g = Voronoi._from_handle(hs[p])
            hs[p] = None
            o = O.build(x, [0, 0, 0], [1, 1, 1], 3, bool(per))
            assert_connectivity_equal(g, o)
            assert_integrals_close(g, o)
            assert np.array_equal(g.cell_n_popped, o.cell_n_popped)
            assert closure_residual(g) < 1e-9
    finally:
        for h in hs:
            if h is not None:
                L.mv_free(h)
        torch.cuda.synchronize()


def test_c5_partial_equals_full_per_cell(c5):
    import torch
    L = _native.lib()
    full = batch(c5, False)
    part = batch(c5, True)
    try:
        for p in range(P):
            a, (n, F, NL) = device_tensors(full[p], ("cell_volume", "cell_centroid"))
            b, (n2, F2, NL2) = device_tensors(part[p], ("cell_volume", "cell_centroid",
                                                        "cell_face_count", "face_right",
                                                        "face_has_shift"))
            m = torch.from_numpy(c5["masks"][p]).to("cuda:0")
            assert n2 == n
            # active cells: the full build's bits; inactive cells: zeros
            assert torch.equal(b["cell_volume"][m], a["cell_volume"][m]), p
            assert torch.equal(b["cell_centroid"].view(-1, 3)[m],
                               a["cell_centroid"].view(-1, 3)[m]), p
            assert bool((b["cell_volume"][~m] == 0).all()), p
            linked_right = int(((b["face_right"] >= 0) & (b["face_has_shift"] == 0)).sum())
            assert NL2 == F2 + linked_right, p
            assert int(b["cell_face_count"].sum()) == NL2, p
        for p in sample_problems(c5):
            x, per = c5["probs"][p]
            g = Voronoi._from_handle(part[p])
            part[p] = None
            o = O.build(x, [0, 0, 0], [1, 1, 1], 3, bool(per), mask=c5["masks"][p])
            assert_connectivity_equal(g, o)
            assert_integrals_close(g, o)
    finally:
        for h in full + part:
            if h is not None:
                L.mv_free(h)
        torch.cuda.synchronize()
