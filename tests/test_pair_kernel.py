"""Two cells per wavefront (cells_per_wave = 2, k_clip_pair): the same arrays
as one cell per wavefront, bit for bit, and the oracle's connectivity, on
inputs that exercise its hand-over to the one-cell kernel -- cells that
outgrow 32 vertices (clustered seeds) and cells that read past entry 128 of
their list (learned first lists of up to 512 entries) -- in both boundary
modes, in 2D, for partial builds and batched problems, and with the same
error status on a degenerate input."""
import numpy as np
import pytest

from meshless_voro_amd import BuildOptions, Voronoi, workloads as W
from meshless_voro_amd.voronoi import MODE_LEARNED, VoronoiError
from oracle import oracle as O
from tests.helpers import assert_connectivity_equal, assert_integrals_close

pytestmark = pytest.mark.gpu
ZERO, ONE = [0.0, 0.0, 0.0], [1.0, 1.0, 1.0]
ARRAYS = ("cell_loc", "cell_volume", "cell_centroid", "cell_safety_radius",
          "cell_face_offset", "cell_face_count", "cell_idx", "cell_n_popped", "face_left",
          "face_right", "face_area", "face_centroid", "face_normal", "face_has_shift",
          "face_shift", "links")


def same(a, b):
    for k in ARRAYS:
        assert np.array_equal(getattr(a, k), getattr(b, k)), k


@pytest.mark.parametrize("periodic", [False, True])
def test_pair_equals_single_and_oracle_clustered(periodic):
    x = W.clustered(20000, 81 + periodic)
    print(f"case clustered 20000 periodic={periodic}", flush=True)
    one = Voronoi.build(x, ZERO, ONE, 3, periodic, options=BuildOptions(cells_per_wave=1))
    two = Voronoi.build(x, ZERO, ONE, 3, periodic, options=BuildOptions(cells_per_wave=2))
    same(two, one)
    # clustered cells outgrow 32 vertices: the hand-over path ran
    assert two.stats["pair_handovers"] > 100
    o = O.build(x, ZERO, ONE, 3, periodic)
    assert_connectivity_equal(two, o)
    assert_integrals_close(two, o)


@pytest.mark.parametrize("periodic", [False, True])
def test_pair_long_learned_lists(periodic):
    """First lists up to 512 entries: cells that read past entry 128 continue
    on the one-cell kernel from their saved state."""
    x = W.clustered(30000, 91 + periodic)
    print(f"case learned clustered 30000 periodic={periodic}", flush=True)
    kw = dict(mode=MODE_LEARNED, first_list_cap=512)
    one = Voronoi.build(x, ZERO, ONE, 3, periodic, options=BuildOptions(cells_per_wave=1, **kw))
    two = Voronoi.build(x, ZERO, ONE, 3, periodic, options=BuildOptions(cells_per_wave=2, **kw))
    same(two, one)
    assert two.stats["pair_handovers"] > 0
    assert np.max(two.cell_n_popped) > 129


def test_pair_uniform_2d_partial_and_small():
    cases = [("uniform 3000 periodic", W.uniform(3000, 5), True, 3, None),
             ("uniform 2d", W.uniform(4000, 6), False, 2, None),
             ("uniform 1d", W.uniform(300, 7), True, 1, None),
             ("partial 10%", W.uniform(20000, 8), True, 3, W.partial_mask(20000, 0.1, 9))]
    for n in (1, 2, 3, 7):
        cases.append((f"uniform {n}", W.uniform(n, 100 + n), n % 2 == 0, 3, None))
    for name, x, periodic, dim, mask in cases:
        print(f"case {name}", flush=True)
        r = []
        for cpw in (1, 2):
            o = BuildOptions(cells_per_wave=cpw)
            if mask is None:
                r.append(Voronoi.build(x, ZERO, ONE, dim, periodic, options=o))
            else:
                r.append(Voronoi.build_partial(x, mask, ZERO, ONE, dim, periodic, options=o))
        same(r[1], r[0])


def test_pair_k0_2_many_passes():
    x = W.uniform(5000, 12)
    r = [Voronoi.build(x, ZERO, ONE, 3, True, options=BuildOptions(cells_per_wave=c, k0=2))
         for c in (1, 2)]
    assert r[1].stats["passes"] > 2
    same(r[1], r[0])


def test_pair_batched():
    rng = np.random.default_rng(3)
    probs = []
    for p in range(24):
        n = int(10 ** rng.uniform(1.0, 3.5))
        x = rng.random((n, 3)) if p % 2 == 0 else np.mod(
            rng.random((4, 3))[rng.integers(0, 4, n)] + 0.05 * rng.normal(size=(n, 3)), 1.0)
        probs.append((x, ZERO, ONE, p % 4 < 2))
    a = Voronoi.build_batch(probs, options=BuildOptions(cells_per_wave=1))
    b = Voronoi.build_batch(probs, options=BuildOptions(cells_per_wave=2))
    for va, vb in zip(a, b):
        same(vb, va)


def test_pair_same_error_status():
    x = W.uniform(2000, 13)
    x = np.concatenate([x, x[:5]])  # duplicate seeds
    st = []
    for cpw in (1, 2):
        try:
            Voronoi.build(x, ZERO, ONE, 3, True, options=BuildOptions(cells_per_wave=cpw))
            st.append(0)
        except VoronoiError as e:
            st.append(e.status)
    assert st[0] != 0 and st[0] == st[1]
