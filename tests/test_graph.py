"""Captured builds (mv_graph_create / mv_graph_launch): a replay of the
recorded hipGraph gives the arrays of the ordinary build on the same seeds,
bit for bit -- on the seeds it was recorded with, after the seeds moved in
place, for many problems with masks, in 2D, over tight clusters (dense grid
cells, long lists) -- and a replay that needs more
than was recorded (clustered seeds written over uniform ones) runs the
ordinary build instead, with the same results."""
import numpy as np
import pytest

from meshless_voro_amd import BuildGraph, BuildOptions, Voronoi, workloads as W
from meshless_voro_amd.voronoi import MODE_LEARNED, VoronoiError

pytestmark = pytest.mark.gpu
Z, O = [0.0, 0.0, 0.0], [1.0, 1.0, 1.0]
ARRAYS = ("cell_loc", "cell_centroid", "cell_volume", "cell_safety_radius",
          "cell_face_offset", "cell_face_count", "cell_idx", "cell_n_popped", "face_left",
          "face_right", "face_area", "face_centroid", "face_normal", "face_has_shift",
          "face_shift", "links")


def same(a, b):
    for k in ARRAYS:
        x, y = getattr(a, k), getattr(b, k)
        assert x.shape == y.shape and np.array_equal(x, y), k


def moved(x, periodic, seed, eps=1e-3):
    rng = np.random.default_rng(seed)
    y = x + eps * rng.standard_normal(x.shape)
    return np.mod(y, 1.0) if periodic else np.clip(y, 0.0, np.nextafter(1.0, 0.0))


@pytest.mark.parametrize("mode,periodic", [(0, True), (MODE_LEARNED, True), (0, False),
                                           (MODE_LEARNED, False)])
def test_graph_replays_equal_builds(mode, periodic):
    import torch
    x = W.uniform(20000, 30 + mode + 2 * periodic)
    print(f"case uniform 20000 mode={mode} periodic={periodic}", flush=True)
    xt = torch.from_numpy(np.ascontiguousarray(x)).to("cuda:0")
    opt = BuildOptions(mode=mode)
    g = BuildGraph(xt, [(0, Z, O, periodic)], options=opt)
    try:
        vs, rep = g.launch()
        assert rep
        same(vs[0], Voronoi.build(x, Z, O, 3, periodic, options=opt))
        for step in range(3):
            x = moved(x, periodic, step)
            xt.copy_(torch.from_numpy(x))
            vs, rep = g.launch()
            assert rep, step
            same(vs[0], Voronoi.build(x, Z, O, 3, periodic, options=opt))
    finally:
        g.close()


def test_graph_many_problems_with_masks():
    import torch
    rng = np.random.default_rng(11)
    xs, probs, masks, off = [], [], [], 0
    for p in range(24):
        n = int(10 ** rng.uniform(2.0, 3.6))
        x = rng.random((n, 3)) if p % 2 == 0 else np.mod(
            rng.random((4, 3))[rng.integers(0, 4, n)] + 0.05 * rng.normal(size=(n, 3)), 1.0)
        xs.append(x)
        probs.append((off, Z, O, p % 4 < 2))
        masks.append(W.partial_mask(n, 0.3, 500 + p) if p % 3 == 0 else np.ones(n, bool))
        off += n
    xyz = np.ascontiguousarray(np.concatenate(xs))
    mask = np.concatenate(masks).astype(np.uint8)
    xt = torch.from_numpy(xyz).to("cuda:0")
    mt = torch.from_numpy(mask).to("cuda:0")
    opt = BuildOptions(mode=MODE_LEARNED)
    g = BuildGraph(xt, probs, mask=mt, options=opt)
    try:
        for step in range(2):
            print(f"case 24 problems, masks, step {step}", flush=True)
            vs, rep = g.launch()
            ref = Voronoi.build_batch([(x, Z, O, p[3]) for x, p in zip(xs, probs)],
                                      masks=masks, options=opt)
            for a, b in zip(vs, ref):
                same(a, b)
            xs = [moved(x, p[3], 40 + step, 1e-4) for x, p in zip(xs, probs)]
            xt.copy_(torch.from_numpy(np.ascontiguousarray(np.concatenate(xs))))
    finally:
        g.close()


@pytest.mark.parametrize("mode", [0, MODE_LEARNED])
def test_graph_replays_dense_cells_and_long_lists(mode):
    """Tight clusters: dense grid cells (sub-columns) and rim cells whose
    lists pass 1,024 entries.  The recorded build takes the planning build's
    passes, so a replay on the same seeds runs the recording."""
    import torch
    from tests.test_dense_cells import tight_clusters
    x = tight_clusters(24000, 61 + mode)
    print(f"case tight clusters 24000 mode={mode}", flush=True)
    xt = torch.from_numpy(np.ascontiguousarray(x)).to("cuda:0")
    opt = BuildOptions(mode=mode)
    ref = Voronoi.build(x, Z, O, 3, True, options=opt)
    assert ref.stats["dense_cells"] > 0 and ref.stats["knn_big_items"] > 0
    g = BuildGraph(xt, [(0, Z, O, True)], options=opt)
    try:
        for _ in range(2):
            vs, rep = g.launch()
            assert rep and g.last_violation() == 0
            same(vs[0], ref)
    finally:
        g.close()


def test_graph_launch_sees_writes_queued_before_it():
    """Seeds moved by device work queued just before the launch (no
    synchronisation in between): the replay reads the moved seeds."""
    import torch
    x = W.uniform(1 << 20, 71)
    xt = torch.from_numpy(np.ascontiguousarray(x)).to("cuda:0")
    g = BuildGraph(xt, [(0, Z, O, True)], options=BuildOptions(mode=MODE_LEARNED))
    try:
        g.launch()
        gen = torch.Generator(device="cuda:0")
        gen.manual_seed(5)
        d = 1e-4 * torch.randn(xt.shape, generator=gen, device="cuda:0", dtype=xt.dtype)
        torch.cuda.synchronize()
        # (a seed outside the box while the writes are in flight)
        xt.add_(d)
        xt.add_(2.0)
        xt.remainder_(1.0)
        vs, rep = g.launch()
        assert rep, g.last_violation()
        y = xt.cpu().numpy()
        same(vs[0], Voronoi.build(y, Z, O, 3, True, options=BuildOptions(mode=MODE_LEARNED)))
    finally:
        g.close()


def test_graph_2d():
    import torch
    x = W.uniform(5000, 41)
    x[:, 2] = 0.0
    xt = torch.from_numpy(np.ascontiguousarray(x)).to("cuda:0")
    g = BuildGraph(xt, [(0, Z, O, True)], dimensionality=2)
    try:
        vs, rep = g.launch()
        same(vs[0], Voronoi.build(x, Z, O, 2, True))
    finally:
        g.close()


def test_graph_falls_back_when_the_plan_is_too_small():
    """Clustered seeds written over the uniform ones the graph was recorded
    with need more redone cells and longer lists than recorded: the replay
    reports it and the ordinary build's results come back."""
    import torch
    x = W.uniform(20000, 50)
    xt = torch.from_numpy(np.ascontiguousarray(x)).to("cuda:0")
    g = BuildGraph(xt, [(0, Z, O, True)])
    try:
        y = W.clustered(20000, 51)
        xt.copy_(torch.from_numpy(np.ascontiguousarray(y)))
        vs, rep = g.launch()
        assert not rep
        same(vs[0], Voronoi.build(y, Z, O, 3, True))
        xt.copy_(torch.from_numpy(np.ascontiguousarray(x)))
        vs, rep = g.launch()
        assert rep
        same(vs[0], Voronoi.build(x, Z, O, 3, True))
    finally:
        g.close()


def test_graph_refuses_exact_order_and_other_clip_layouts():
    import torch
    x = W.uniform(1000, 52)
    xt = torch.from_numpy(np.ascontiguousarray(x)).to("cuda:0")
    for kw in (dict(exact_order=True), dict(cells_per_wave=2), dict(cells_per_wave=64)):
        with pytest.raises(VoronoiError):
            BuildGraph(xt, [(0, Z, O, True)], options=BuildOptions(**kw))
