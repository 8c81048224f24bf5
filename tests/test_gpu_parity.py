"""GPU parity: libmvoro (gfx950 HIP, through the C ABI) against the CPU
oracle on the same seeded inputs.  Connectivity must be bit-identical;
integrals must agree within the FP64 tolerance of helpers.REL_TOL
(relative 1e-12, near-zero values scaled by the cell size)."""
import json
import os

import numpy as np
import pytest

from meshless_voro_amd import BuildOptions, Voronoi, VoronoiError, workloads as W
from oracle import oracle as O
from tests.helpers import (assert_connectivity_equal, assert_integrals_close, closure_residual,
                           consistency_check, degenerate_fixture)

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
ZERO, ONE = [0.0, 0.0, 0.0], [1.0, 1.0, 1.0]


def both(x, anchor=ZERO, width=ONE, dim=3, periodic=False, mask=None, options=None):
    print(f"case n={len(x)} dim={dim} periodic={periodic} mask={mask is not None}", flush=True)
    if mask is None:
        g = Voronoi.build(x, anchor, width, dim, periodic, options=options)
    else:
        g = Voronoi.build_partial(x, mask, anchor, width, dim, periodic, options=options)
    o = O.build(x, anchor, width, dim, periodic, mask=mask)
    return g, o


def check(g, o):
    assert_connectivity_equal(g, o)
    assert_integrals_close(g, o)


@pytest.mark.parametrize("n", [1, 2, 3, 17, 1000, 20000])
@pytest.mark.parametrize("periodic", [False, True])
def test_uniform(n, periodic):
    x = W.uniform(n, 100 + n)
    g, o = both(x, periodic=periodic)
    check(g, o)
    assert np.array_equal(g.cell_n_popped, o.cell_n_popped)


def test_empty():
    g = Voronoi.build(np.zeros((0, 3)), ZERO, ONE)
    assert g.cell_volume.shape == (0,) and g.face_area.shape == (0,)


@pytest.mark.parametrize("count", [2, 3, 4, 10])
@pytest.mark.parametrize("periodic", [False, True])
def test_perfect_grid_exact_predicate(count, periodic):
    """perfect grids: every cell reaches the exact in-sphere path and exact
    zeros; face order among tied neighbours is (idx, image)."""
    x = W.grid(count)
    g, o = both(x, periodic=periodic)
    check(g, o)
    assert g.stats["exact_tests"] > 0
    assert np.allclose(g.cell_volume, 1.0 / count ** 3, rtol=1e-12)


def test_perturbed_grid_32():
    """32^3 perturbed grid (+-0.1 spacing), rectangular box."""
    x = W.perturbed_grid_32()
    g, o = both(x)
    check(g, o)


def test_grid_jitter_16():
    x = W.grid_jitter(16)
    for periodic in (False, True):
        g, o = both(x, periodic=periodic)
        check(g, o)


@pytest.mark.parametrize("periodic", [False, True])
def test_perfect_grid_64(periodic):
    """262,144-seed perfect grid (exact dyadic positions): every cell is a
    cube of side 1/64 decided through the exact predicate."""
    x = W.grid(64)
    g, o = both(x, periodic=periodic)
    check(g, o)
    assert np.allclose(g.cell_volume, 64.0 ** -3, rtol=1e-12)


@pytest.mark.parametrize("periodic", [False, True])
def test_grid_jitter_64_fails_like_the_oracle(periodic):
    """1% of a 64^3 perfect grid moved by 1e-12: the filtered and the exact
    decisions disagree about near-co-spherical vertices and the boundary
    walk of compute_boundary (convex_cell.rs:421-441) finds no extension;
    the reference panics there and the oracle reports it.  The GPU path
    must report the same error, not a different tessellation."""
    x = W.grid_jitter(64)
    print(f"case grid_jitter_64 periodic={periodic}", flush=True)
    with pytest.raises(O.OracleError) as eo:
        O.build(x, ZERO, ONE, 3, periodic)
    with pytest.raises(VoronoiError) as eg:
        Voronoi.build(x, ZERO, ONE, 3, periodic)
    assert eo.value.status == eg.value.status == 6


@pytest.mark.parametrize("periodic", [False, True])
def test_partial(periodic):
    x = W.uniform(5000, 7)
    m = W.partial_mask(5000, 0.1, 8)
    g, o = both(x, periodic=periodic, mask=m)
    check(g, o)
    assert np.all(g.cell_volume[~m] == 0.0) and np.all(g.cell_idx[~m] == 0)


def test_partial_single_cells():
    """src/voronoi.rs:1051-1085"""
    x = W.perturbed_grid(3, 0.9, seed=3)
    full = Voronoi.build(x, ZERO, ONE)
    for i in range(27):
        m = np.zeros(27, bool)
        m[i] = True
        part = Voronoi.build_partial(x, m, ZERO, ONE)
        assert part.cell_volume[i] == full.cell_volume[i]
        assert part.cell_face_count[i] == full.cell_face_count[i]


def test_clustered():
    x = W.clustered(30000, 33)
    g, o = both(x, periodic=True)
    check(g, o)


def test_reference_fixtures():
    r = json.load(open(os.path.join(GOLDEN, "reference_points.json")))
    z = r["test_non_perturbed_z"]
    check(*both(np.array(z["points"]), z["anchor"], z["width"]))
    d = r["degenerate_test"]
    x, a, w = degenerate_fixture(d)
    g, o = both(x, a, w)
    check(g, o)
    consistency_check(g)
    s = r["degenerate_test_simplified"]
    x, a, w = degenerate_fixture(s)
    m = np.zeros(27, bool)
    m[s["mask_true"]] = True
    check(*both(x, a, w, mask=m))


@pytest.mark.parametrize("dim", [1, 2])
def test_low_dimensions(dim):
    rng = np.random.default_rng(dim)
    x = rng.random((500, 3))
    for periodic in (False, True):
        g, o = both(x, dim=dim, periodic=periodic)
        check(g, o)


def test_small_k0_forces_redo_passes():
    """k0 = 2 makes every cell run out of neighbours and be redone with a list
    sized from its safety radius (several passes); output unchanged."""
    x = W.uniform(3000, 44)
    g, o = both(x, periodic=True, options=BuildOptions(k0=2))
    check(g, o)
    assert g.stats["passes"] >= 3 and g.stats["cells_redone"] >= 3000
    assert np.array_equal(g.cell_n_popped, o.cell_n_popped)


@pytest.mark.parametrize("k0", [3, 8, 40])
def test_small_k_max_reports_capacity_or_matches(k0):
    """A cap on the list length below what cells need returns
    MV_E_CAPACITY; a sufficient cap gives the oracle's result."""
    x = W.uniform(2000, 45)
    with pytest.raises(VoronoiError) as e:
        Voronoi.build(x, ZERO, ONE, 3, True, options=BuildOptions(k0=k0, k_max=12))
    assert e.value.status == 9
    g, o = both(x, periodic=True, options=BuildOptions(k0=k0, k_max=4096))
    check(g, o)


def test_errors():
    with pytest.raises(VoronoiError) as e:
        Voronoi.build([[0.5, 0.5, 0.5], [0.5, 0.5, 0.5]], ZERO, ONE)
    assert e.value.status in (2, 3)
    with pytest.raises(VoronoiError) as e:
        Voronoi.build([[0.5, 0.5, np.nan]], ZERO, ONE)
    assert e.value.status == 2
    with pytest.raises(VoronoiError) as e:
        Voronoi.build_partial(np.random.rand(4, 3), [True], ZERO, ONE)
    assert e.value.status == 1


def test_insphere_golden_on_device():
    import ctypes as C
    from meshless_voro_amd import _native
    vecs = json.load(open(os.path.join(GOLDEN, "insphere_vectors.json")))
    arr = np.ascontiguousarray(np.array([v["pts"] for v in vecs], dtype=np.int64).reshape(-1, 15))
    out = np.zeros(len(vecs), np.int8)
    st = _native.lib().mv_in_sphere_exact(arr.ctypes.data, len(vecs), out.ctypes.data)
    assert st == 0
    assert np.array_equal(out, np.array([v["sign"] for v in vecs], np.int8))


def test_insphere_random_against_oracle():
    import ctypes as C
    from meshless_voro_amd import _native
    rng = np.random.default_rng(99)
    base = rng.integers(1 << 40, 1 << 51, size=(20000, 1, 3))
    pts = base + rng.integers(-(1 << 12), 1 << 12, size=(20000, 5, 3))
    # make a third of them exactly co-spherical box corners
    lo = rng.integers(0, 1 << 50, size=(7000, 3))
    ext = rng.integers(1, 1 << 40, size=(7000, 3))
    for t in range(7000):
        corners = [[lo[t, i] + (ext[t, i] if (k >> i) & 1 else 0) for i in range(3)] for k in range(8)]
        pts[t] = np.array(corners[:5])
    arr = np.ascontiguousarray(pts.reshape(-1, 15).astype(np.int64))
    out = np.zeros(len(arr), np.int8)
    assert _native.lib().mv_in_sphere_exact(arr.ctypes.data, len(arr), out.ctypes.data) == 0
    want = np.array([O.in_sphere_exact(*p.reshape(5, 3).tolist()) for p in arr[:6000]], np.int8)
    assert np.array_equal(out[:6000], want)
    assert (out[:6000] == 0).sum() > 1000


def test_large_uniform_properties():
    """2,097,152 periodic uniform seeds: size-independent properties (box
    volume, closure of a sample of cells, face-link symmetry)."""
    x = W.uniform_2m()
    g = Voronoi.build(x, ZERO, ONE, 3, True)
    assert abs(g.cell_volume.sum() - 1.0) < 1e-9
    assert np.all(g.cell_volume > 0)
    # every internal non-periodic face is linked from both its cells
    internal = (g.face_right >= 0) & (g.face_has_shift == 0)
    assert g.links.shape[0] == g.face_area.shape[0] + int(internal.sum())
    # sample parity with the oracle on a 65,536-seed prefix of the stream
    xs = x[:65536]
    check(*both(xs, periodic=True))


@pytest.mark.parametrize("periodic", [False, True])
def test_exact_order_mode_is_bit_identical(periodic):
    """MV_FLAG_EXACT_ORDER: the thread-per-cell kernel sums every integral in
    the reference's order, so all outputs equal the oracle's bit for bit."""
    x = W.uniform(20000, 555)
    g, o = both(x, periodic=periodic, options=BuildOptions(exact_order=True))
    assert_connectivity_equal(g, o)
    for f in ("cell_volume", "cell_centroid", "face_area", "face_centroid", "face_normal",
              "cell_safety_radius", "face_shift"):
        assert np.array_equal(getattr(g, f), getattr(o, f)), f


def test_default_mode_deterministic():
    x = W.uniform(50000, 77)
    a = Voronoi.build(x, ZERO, ONE, 3, True)
    b = Voronoi.build(x, ZERO, ONE, 3, True)
    for f in ("cell_volume", "cell_centroid", "face_area", "face_centroid", "links"):
        assert np.array_equal(getattr(a, f), getattr(b, f)), f


def test_clusters_in_empty_space_need_long_lists():
    """Eight tight clusters with empty space between them (the odd problems
    of the many-problem workload): cells on a cluster's rim reach across the
    empty box, so their security radius holds nearly every seed (lists far
    beyond 32,768 entries with the periodic images)."""
    rng = np.random.default_rng(77)
    n = 40000
    c = rng.random((8, 3))
    x = np.mod(c[rng.integers(0, 8, n)] + 0.03 * rng.normal(size=(n, 3)), 1.0)
    x[x >= 1.0] = 0.0
    g, o = both(x, periodic=True)
    check(g, o)
    assert np.array_equal(g.cell_n_popped, o.cell_n_popped)
    assert o.cell_n_popped.max() > 32768
    # the rim cells went through the large-cell kernel (over 64 vertices)
    assert g.stats["big_cells"] > 0


@pytest.mark.parametrize("periodic", [False, True])
def test_dense_cluster_in_one_grid_cell(periodic):
    """8,000 seeds in a Gaussian cluster of sigma 2e-4 inside one grid cell
    (the grid is sized for the mean density): the cell's seeds are put in
    index order by the block sort (LDS chunks of 4,096, then merges), and
    the tessellation equals the oracle's."""
    rng = np.random.default_rng(2024)
    x = np.concatenate([rng.random((12000, 3)),
                        0.5 + 2e-4 * rng.normal(size=(8000, 3))])
    g, o = both(x, periodic=periodic)
    check(g, o)
    assert np.array_equal(g.cell_n_popped, o.cell_n_popped)
