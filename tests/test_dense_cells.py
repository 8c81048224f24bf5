"""Dense grid cells (more than 256 seeds) split into sub-columns for the
neighbour gathers: builds over a few tight clusters, where the gathers read
those sub-columns, the long lists of the clusters' rim cells included, give
the oracle's connectivity and integrals in both boundary modes and both
modes, also in a box far from the origin with unequal sides; a batched build
over such problems equals the builds one by one; and a set with no dense
cell builds no sub-columns."""
import numpy as np
import pytest

from meshless_voro_amd import BuildOptions, Voronoi, workloads as W
from meshless_voro_amd.voronoi import MODE_LEARNED
from oracle import oracle as O
from tests.helpers import assert_connectivity_equal, assert_integrals_close

pytestmark = pytest.mark.gpu
ZERO, ONE = [0.0, 0.0, 0.0], [1.0, 1.0, 1.0]
ARRAYS = ("cell_loc", "cell_volume", "cell_centroid", "cell_safety_radius",
          "cell_face_offset", "cell_face_count", "cell_idx", "cell_n_popped", "face_left",
          "face_right", "face_area", "face_centroid", "face_normal", "face_has_shift",
          "face_shift", "links")


def tight_clusters(n, seed, nclusters=4, sigma=0.004, frac=0.75):
    """frac of n seeds in a few Gaussian clusters of width sigma (each puts
    thousands of seeds in a grid cell of a 2-seed-per-cell grid), the rest
    uniform; wrapped into the unit box."""
    rng = np.random.default_rng(seed)
    centres = rng.random((nclusters, 3))
    nc = int(frac * n)
    x = centres[rng.integers(0, nclusters, nc)] + sigma * rng.standard_normal((nc, 3))
    x = np.concatenate([np.mod(x, 1.0), rng.random((n - nc, 3))])
    x[x >= 1.0] = 0.0
    return x


@pytest.mark.parametrize("periodic", [True, False])
@pytest.mark.parametrize("mode", [0, MODE_LEARNED])
def test_dense_cells_match_oracle(periodic, mode):
    x = tight_clusters(24000, 300 + 2 * mode + periodic)
    print(f"case tight clusters 24000 periodic={periodic} mode={mode}", flush=True)
    v = Voronoi.build(x, ZERO, ONE, 3, periodic, options=BuildOptions(mode=mode))
    assert v.stats["dense_cells"] > 0
    # rim cells of the clusters read whole clusters (long lists)
    assert v.stats["knn_big_items"] > 0
    o = O.build(x, ZERO, ONE, 3, periodic)
    assert_connectivity_equal(v, o)
    assert_integrals_close(v, o)


@pytest.mark.parametrize("periodic", [True, False])
def test_dense_cells_in_a_shifted_stretched_box(periodic):
    """The sub-column boundaries and the gathers' bounds in absolute
    coordinates far from the origin, with unequal box sides."""
    anchor = np.array([-7.25, 3.5, 1000.0])
    width = np.array([0.5, 3.0, 0.02])
    x = anchor + width * tight_clusters(20000, 350 + periodic)
    print(f"case shifted box periodic={periodic}", flush=True)
    v = Voronoi.build(x, anchor, width, 3, periodic)
    assert v.stats["dense_cells"] > 0
    o = O.build(x, anchor, width, 3, periodic)
    assert_connectivity_equal(v, o)
    assert_integrals_close(v, o)


def test_dense_cells_batched_equals_one_by_one():
    probs = []
    for p in range(6):
        x = tight_clusters(6000 + 1500 * p, 400 + p, nclusters=2, sigma=0.003)
        probs.append((x, ZERO, ONE, p % 2 == 0))
    print("case 6 problems with dense cells", flush=True)
    b = Voronoi.build_batch(probs)
    for (x, lo, w, per), vb in zip(probs, b):
        v = Voronoi.build(x, lo, w, 3, per)
        assert v.stats["dense_cells"] > 0
        for k in ARRAYS:
            assert np.array_equal(getattr(v, k), getattr(vb, k)), k


def test_uniform_has_no_dense_cells():
    v = Voronoi.build(W.uniform(20000, 5), ZERO, ONE, 3, True)
    assert v.stats["dense_cells"] == 0
