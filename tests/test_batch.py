"""Many tessellations in one build (mv_build_batch): every problem's result
equals the single build of that problem, bit for bit -- uneven sizes, both
boundary kinds, clustered and uniform seeds, partial masks, empty problems,
learned mode, and the oracle on a few problems."""
import numpy as np
import pytest

from meshless_voro_amd import BuildOptions, Voronoi, workloads as W
from meshless_voro_amd.voronoi import MODE_LEARNED
from oracle import oracle as O
from tests.helpers import assert_connectivity_equal, assert_integrals_close

pytestmark = pytest.mark.gpu
ARRAYS = ("cell_loc", "cell_centroid", "cell_volume", "cell_safety_radius", "cell_face_offset",
          "cell_face_count", "cell_idx", "cell_n_popped", "face_left", "face_right", "face_area",
          "face_centroid", "face_normal", "face_has_shift", "face_shift", "links")


def same(a, b):
    for k in ARRAYS:
        assert np.array_equal(getattr(a, k), getattr(b, k)), k


def mixed_problems(P, seed, lo=1.0, hi=3.2):
    rng = np.random.default_rng(seed)
    out = []
    for p in range(P):
        n = int(10 ** rng.uniform(lo, hi))
        width = rng.uniform(0.5, 2.0, 3)
        anchor = rng.uniform(-1.0, 1.0, 3)
        if p % 2 == 0:
            x = anchor + rng.random((n, 3)) * width
        else:
            c = anchor + rng.random((4, 3)) * width
            x = c[rng.integers(0, 4, n)] + 0.05 * width * rng.normal(size=(n, 3))
            x = anchor + np.mod(x - anchor, width)
        out.append((x, anchor, width, p % 4 < 2))
    return out


@pytest.mark.parametrize("learned", [False, True])
def test_batch_equals_single_builds(learned):
    probs = mixed_problems(40, 7)
    opt = BuildOptions(mode=MODE_LEARNED) if learned else None
    print(f"case batch P={len(probs)} n={sum(len(x) for x, *_ in probs)} learned={learned}",
          flush=True)
    got = Voronoi.build_batch(probs, options=opt)
    assert len(got) == len(probs)
    for (x, a, w, per), g in zip(probs, got):
        same(g, Voronoi.build(x, a, w, 3, per))


def test_batch_partial_and_empty():
    probs = mixed_problems(12, 8)
    probs.insert(3, (np.zeros((0, 3)), [0, 0, 0], [1, 1, 1], False))
    probs.append((np.array([[0.5, 0.5, 0.5]]), [0, 0, 0], [1, 1, 1], True))
    masks = [None if i % 3 == 0 else W.partial_mask(len(x), 0.1, 100 + i)
             for i, (x, *_) in enumerate(probs)]
    got = Voronoi.build_batch(probs, masks=masks)
    for (x, a, w, per), mk, g in zip(probs, masks, got):
        ref = (Voronoi.build(x, a, w, 3, per) if mk is None
               else Voronoi.build_partial(x, mk, a, w, 3, per))
        same(g, ref)


def test_batch_against_oracle():
    probs = mixed_problems(6, 9, 2.0, 3.0)
    got = Voronoi.build_batch(probs)
    for (x, a, w, per), g in zip(probs, got):
        o = O.build(x, a, w, 3, per)
        assert_connectivity_equal(g, o)
        assert_integrals_close(g, o)


def test_batch_2d():
    rng = np.random.default_rng(10)
    probs = [(rng.random((int(rng.integers(5, 300)), 3)), [0, 0, 0], [1, 1, 1], i % 2 == 0)
             for i in range(9)]
    got = Voronoi.build_batch(probs, dimensionality=2)
    for (x, a, w, per), g in zip(probs, got):
        same(g, Voronoi.build(x, a, w, 2, per))


def test_batch_errors():
    probs = mixed_problems(3, 11)
    x = probs[1][0].copy()
    x[1] = x[0]
    probs[1] = (x,) + probs[1][1:]
    from meshless_voro_amd import VoronoiError
    with pytest.raises(VoronoiError):
        Voronoi.build_batch(probs)
