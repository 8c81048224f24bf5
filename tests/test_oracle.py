"""Pins the CPU oracle (oracle/mvoro_oracle.c) against known answers, closed
forms, invariants and brute force, before it is trusted as the checker of
the GPU path.  The reference test each case restates is cited."""
import json
import math
import os

import numpy as np
import pytest

from oracle import oracle as O
from meshless_voro_amd import workloads as W
from tests.helpers import closure_residual, consistency_check, degenerate_fixture

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
ZERO, ONE = [0.0, 0.0, 0.0], [1.0, 1.0, 1.0]


# ---------------------------------------------------------------- primitives
def test_signed_volume():
    """src/geometry.rs:368-379"""
    v0, v1, v2, v3 = ZERO, [1, 0, 0], [0, 1, 0], [0, 0, 1]
    assert O.signed_volume_tet(v0, v1, v2, v3) == 1.0 / 6.0
    assert O.signed_volume_tet(v0, v2, v1, v3) == -1.0 / 6.0


def test_signed_area():
    """src/geometry.rs:382-402"""
    v0, v1, v2, t = ZERO, [1, 0, 0], [0, 1, 0], [0, 0, 1]
    assert O.signed_area_tri(v0, v1, v2, t) == 0.5
    assert O.signed_area_tri(v0, v2, v1, t) == -0.5
    assert O.signed_area_tri(v0, v1, v2, [10.0, 10.0, 1.0]) == 0.5


def test_simple_cycle_known_answer():
    """src/simple_cycle.rs:238-263: the cycle is [2, 4, 6, 5]."""
    tris = [(2, 4, 1), (1, 5, 2), (5, 1, 3), (5, 3, 6), (3, 4, 6), (4, 3, 1)]
    assert O.simple_cycle(tris, 7) == [2, 4, 6, 5]


def test_simple_cycle_single_triangle():
    assert O.simple_cycle([(0, 1, 2)], 3) == [0, 1, 2]


def test_insphere_golden_vectors():
    """in_sphere_test_exact (src/geometry.rs:271-314) against Python ints."""
    vecs = json.load(open(os.path.join(GOLDEN, "insphere_vectors.json")))
    signs = set()
    for v in vecs:
        s = O.in_sphere_exact(*v["pts"])
        assert s == v["sign"], v
        signs.add(s)
    assert signs == {-1, 0, 1}


def test_insphere_planted_mistake_is_caught():
    """A swapped argument flips the sign of a non-degenerate case: the
    golden vectors distinguish the argument order."""
    vecs = json.load(open(os.path.join(GOLDEN, "insphere_vectors.json")))
    flipped = sum(O.in_sphere_exact(v["pts"][0], v["pts"][2], v["pts"][1], v["pts"][3],
                                    v["pts"][4]) != v["sign"] for v in vecs if v["sign"] != 0)
    assert flipped > 0


def test_iloc_golden_vectors():
    """SimulationBoundary::iloc (src/voronoi/boundary.rs:49-64)."""
    for v in json.load(open(os.path.join(GOLDEN, "iloc_vectors.json"))):
        assert O.iloc(v["anchor"], v["width"], v["periodic"], v["dim"], v["loc"]) == v["iloc"]


def test_insphere_equivalence():
    """src/geometry.rs:510-547 (seeded): the plane test of the bisector,
    applied to the circumcentre of a tet, has the sign of the exact in-sphere
    test on snapped coordinates."""
    rng = np.random.default_rng(1234)
    checked = 0
    for _ in range(4000):
        tet = rng.random((4, 3))
        v1, v2, v3 = tet[0] - tet[1], tet[0] - tet[2], tet[0] - tet[3]
        nrm = np.cross(v1, v2)
        nrm /= np.linalg.norm(nrm)
        if nrm @ v3 < 0:
            continue
        # the reference builds these planes through tet[0] + 0.5 * v_i
        vert = O.intersect_planes(v1 / np.linalg.norm(v1), tet[0] + 0.5 * v1,
                                  v2 / np.linalg.norm(v2), tet[0] + 0.5 * v2,
                                  v3 / np.linalg.norm(v3), tet[0] + 0.5 * v3)
        ngb = rng.random(3)
        v4 = tet[0] - ngb
        clip = O.halfspace_clip(v4 / np.linalg.norm(v4), tet[0] + 0.5 * v4, vert)
        pts = [O.iloc(ZERO, ONE, False, 3, p) for p in (tet[0], tet[1], tet[2], tet[3], ngb)]
        assert clip == O.in_sphere_exact(*pts)
        checked += 1
    assert checked > 1000


def test_intersect_planes_closed_form():
    rng = np.random.default_rng(5)
    for _ in range(200):
        n = rng.normal(size=(3, 3))
        n /= np.linalg.norm(n, axis=1)[:, None]
        p = rng.random((3, 3))
        got = np.array(O.intersect_planes(n[0], p[0], n[1], p[1], n[2], p[2]))
        want = np.linalg.solve(n, np.einsum("ij,ij->i", n, p))
        assert np.allclose(got, want, rtol=1e-9, atol=1e-9)
    assert O.intersect_planes([1, 0, 0], ZERO, [1, 0, 0], ONE, [0, 0, 1], ZERO) is None


# ------------------------------------------------------ neighbour ordering
def _brute_stream(x, width, q, periodic, dim=3):
    rng_i = range(-1, 2) if periodic else range(0, 1)
    rng_j = range(-1, 2) if periodic and dim >= 2 else range(0, 1)
    rng_k = range(-1, 2) if periodic and dim == 3 else range(0, 1)
    ent = []
    for i in rng_i:
        for j in rng_j:
            for k in rng_k:
                s = np.array([i * width[0], j * width[1], k * width[2]], np.float64)
                qs = x[q] + s
                d = qs[None, :] - x
                key = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
                img = (i + 1) * 9 + (j + 1) * 3 + (k + 1)
                ent += [(float(key[g]), g, img) for g in range(x.shape[0])]
    ent.sort()
    return ent


@pytest.mark.parametrize("periodic", [False, True])
def test_neighbour_stream_matches_brute_force(periodic):
    """rtree_nn.rs:10-36, 124-164: ascending key; (idx, image) among ties."""
    rng = np.random.default_rng(11)
    x = rng.random((300, 3))
    for q in (0, 7, 150, 299):
        idx, img, key = O.neighbour_stream(x, ONE, q, 200, 3, periodic)
        ref = _brute_stream(x, ONE, q, periodic)[:200]
        assert [(float(a), int(b), int(c)) for a, b, c in zip(key, idx, img)] == ref


def test_neighbour_stream_ties_on_grid():
    x = W.grid(4)
    idx, img, key = O.neighbour_stream(x, ONE, 21, 60, 3, True)
    ref = _brute_stream(x, ONE, 21, True)[:60]
    assert [(float(a), int(b), int(c)) for a, b, c in zip(key, idx, img)] == ref


# -------------------------------------------------------- whole builds
def test_single_cell():
    """src/voronoi.rs:902-909; the six box faces sit at their face centres,
    with the box planes' inward normals (src/voronoi/boundary.rs:32-39)."""
    v = O.build([[0.5, 0.5, 0.5]], ZERO, ONE)
    assert v.cell_volume[0] == pytest.approx(1.0, abs=1e-15)
    assert len(v.face_area) == 6 and np.allclose(v.face_area, 1.0)
    assert np.allclose(v.cell_centroid[0], [0.5, 0.5, 0.5], rtol=0, atol=1e-15)
    centres = {(0.0, 0.5, 0.5), (1.0, 0.5, 0.5), (0.5, 0.0, 0.5), (0.5, 1.0, 0.5),
               (0.5, 0.5, 0.0), (0.5, 0.5, 1.0)}
    got = {tuple(float(round(c, 12)) for c in f) for f in v.face_centroid}
    assert got == centres
    # the normal points from the face towards the generator
    for f in range(6):
        assert np.dot(v.face_normal[f], [0.5, 0.5, 0.5] - v.face_centroid[f]) == pytest.approx(0.5)


def test_single_cell_off_centre_box():
    """One generator off the centre of a non-cubic box: the cell is the box,
    so volume, centroid and the face centroids are the box's, whatever the
    generator position (src/voronoi.rs:902-909 generalised)."""
    a, w = np.array([1.0, -2.0, 0.5]), np.array([2.0, 3.0, 0.25])
    v = O.build([[1.3, -0.1, 0.6]], a, w)
    assert v.cell_volume[0] == pytest.approx(1.5, rel=1e-14)
    mid = a + 0.5 * w
    assert np.allclose(v.cell_centroid[0], mid, rtol=0, atol=1e-14)
    want = []
    for ax in range(3):
        for end in (a[ax], a[ax] + w[ax]):
            c = mid.copy()
            c[ax] = end
            want.append(c)
    got = v.face_centroid
    for c in want:
        assert np.min(np.max(np.abs(got - c), axis=1)) < 1e-14
    areas = sorted(float(x) for x in v.face_area)
    assert areas == pytest.approx(sorted([w[1] * w[2]] * 2 + [w[0] * w[2]] * 2 +
                                         [w[0] * w[1]] * 2), rel=1e-14)


def test_two_cells():
    """src/voronoi.rs:911-930"""
    v = O.build([[0.3, 0.4, 0.25], [0.7, 0.6, 0.75]], ZERO, ONE)
    assert v.cell_volume == pytest.approx([0.5, 0.5], abs=1e-14)


@pytest.mark.parametrize("count", [2, 3, 4])
def test_regular_grid_cells(count):
    """test_eight_cells / test_27_cells / test_64_cells (src/voronoi.rs:1008-
    1039): perfect grids reach the exact predicate; every cell is a cube."""
    x = W.grid(count)
    v = O.build(x, ZERO, ONE)
    h = 1.0 / count
    assert np.allclose(v.cell_volume, h ** 3, rtol=1e-12)
    assert np.allclose(v.cell_centroid, x, rtol=0, atol=1e-14)
    assert np.allclose(v.face_area, h * h, rtol=1e-12)
    _grid_face_closed_form(v, x, h)
    assert v.exact_tests > 0


def _grid_face_closed_form(v, x, h):
    """Perfect grid: every face is the h x h square half a spacing from its
    left generator, so its centroid is loc[left] - (h / 2) * normal (the
    normal points towards the left generator), it lies on the axis through
    the generator, and an internal face's centroid is the midpoint of its two
    generators (src/voronoi/voronoi_face.rs:36-43)."""
    left = v.face_left.astype(np.int64)
    want = x[left] - 0.5 * h * v.face_normal
    assert np.allclose(v.face_centroid, want, rtol=0, atol=1e-14)
    # the normal is an axis vector
    assert np.allclose(np.sort(np.abs(v.face_normal), 1), [0.0, 0.0, 1.0], atol=0)
    internal = (v.face_right >= 0) & (v.face_has_shift == 0)
    right = v.face_right[internal].astype(np.int64)
    mid = 0.5 * (x[left[internal]] + x[right])
    assert np.allclose(v.face_centroid[internal], mid, rtol=0, atol=1e-14)


def test_regular_grid_closed_form_periodic():
    x = W.grid(6)
    v = O.build(x, ZERO, ONE, periodic=True)
    assert np.allclose(v.cell_volume, 1.0 / 216, rtol=1e-12)
    assert np.allclose(v.cell_centroid, x, rtol=0, atol=1e-14)
    # every cell has exactly 6 face links
    assert np.all(v.cell_face_count == 6)
    _grid_face_closed_form(v, x, 1.0 / 6)
    # a periodic face is built by both of its cells, each in its own frame
    per = v.face_has_shift != 0
    assert per.any()
    assert np.allclose(v.face_area, 1.0 / 36, rtol=1e-12)


def test_125_cells():
    """src/voronoi.rs:1041-1049"""
    v = O.build(W.perturbed_grid(5, 0.5, seed=1), ZERO, ONE)
    consistency_check(v)


def test_3_d():
    """src/voronoi.rs:1112-1122: 15^3, pert 0.95, box 2."""
    x = W.perturbed_grid(15, 0.95, seed=2, width=(2.0, 2.0, 2.0))
    v = O.build(x, ZERO, [2.0, 2.0, 2.0])
    assert v.cell_volume.shape[0] == x.shape[0]
    consistency_check(v)
    assert closure_residual(v) < 1e-10


def test_partial():
    """src/voronoi.rs:1051-1085: a one-cell partial build reproduces that
    cell's volume and face count; the others have volume 0."""
    x = W.perturbed_grid(3, 0.9, seed=3)
    full = O.build(x, ZERO, ONE)
    for i in range(27):
        m = np.zeros(27, bool)
        m[i] = True
        part = O.build(x, ZERO, ONE, mask=m)
        assert part.cell_volume[i] == full.cell_volume[i]
        assert part.cell_face_count[i] == full.cell_face_count[i]
        assert np.all(part.cell_volume[np.arange(27) != i] == 0.0)


def test_degenerate_large_coordinates():
    """degenerate_test (src/voronoi.rs:1157-1165)"""
    r = json.load(open(os.path.join(GOLDEN, "reference_points.json")))["degenerate_test"]
    x, a, w = degenerate_fixture(r)
    v = O.build(x, a, w)
    consistency_check(v)


def test_degenerate_simplified_partial():
    """degenerate_test_simplified (src/voronoi.rs:1167-1184)"""
    r = json.load(open(os.path.join(GOLDEN, "reference_points.json")))["degenerate_test_simplified"]
    x, a, w = degenerate_fixture(r)
    m = np.zeros(27, bool)
    m[r["mask_true"]] = True
    v = O.build(x, a, w, mask=m)
    consistency_check(v)


def test_non_perturbed_z():
    """test_non_perturbed_z (src/voronoi.rs:1247-1296): 41 fixed points."""
    r = json.load(open(os.path.join(GOLDEN, "reference_points.json")))["test_non_perturbed_z"]
    v = O.build(r["points"], r["anchor"], r["width"])
    consistency_check(v)


def test_five_cells_2d():
    """src/voronoi.rs:968-1006"""
    d = math.sqrt(0.1)
    x = [[0.5, 0.5, 0.5], [0.5 - d, 0.5 - d, 0.5], [0.5 - d, 0.5 + d, 0.5],
         [0.5 + d, 0.5 + d, 0.5], [0.5 + d, 0.5 - d, 0.5]]
    v = O.build(x, ZERO, ONE, dimensionality=2)
    assert v.cell_volume == pytest.approx([0.2] * 5, abs=1e-12)


def test_4_cells_2d_periodic():
    """src/voronoi.rs:932-966"""
    x = [[0.4, 0.3, 0.0], [1.6, 0.2, 0.0], [0.6, 0.8, 0.0], [1.4, 0.7, 0.0]]
    v = O.build(x, ZERO, [2.0, 1.0, 1.0], dimensionality=2, periodic=True)
    consistency_check(v, dim=2)
    assert v.cell_volume.sum() == pytest.approx(2.0, rel=1e-12)


def test_2_d():
    """src/voronoi.rs:1087-1110: 25^2 perturbed plane, periodic, sum = 4."""
    rs = W.Stream(9)
    count, pert = 25, 0.95
    n = np.arange(count * count)
    ij = np.stack([n // count, n % count], 1).astype(np.float64)
    u = rs.uniform(2 * len(n)).reshape(-1, 2) - 0.5
    xy = (ij + 0.5 + pert * u) * 2.0 / count + 2.0
    x = np.concatenate([xy, np.full((len(n), 1), 2.0 + 0.5 * count * 1.0 / count)], 1)
    v = O.build(x, [2.0, 2.0, 2.0], [2.0, 2.0, 1.0], dimensionality=2, periodic=True)
    assert v.cell_volume.sum() == pytest.approx(4.0, abs=1e-10)


def test_1_d():
    x = np.zeros((10, 3))
    x[:, 0] = (np.arange(10) + 0.5) / 10
    v = O.build(x, ZERO, ONE, dimensionality=1)
    assert np.allclose(v.cell_volume, 0.1, rtol=1e-12)
    consistency_check(v, dim=1)


@pytest.mark.parametrize("periodic", [False, True])
def test_scipy_brute_force(periodic):
    """Volumes, centroids, face areas and face centroids of every 7th cell
    against Qhull (scipy HalfspaceIntersection + ConvexHull) on 150 random
    points.  Each face the oracle reports for a cell (as left, or as the
    right of an unshifted face) is matched to the hull triangles whose
    outward normal is the face's (normals of distinct faces of a convex cell
    differ); the face's area and centroid are the area-weighted sums over
    those triangles, and the cell centroid is the volume-weighted sum over
    the tetrahedra from an interior point to every hull triangle."""
    from scipy.spatial import ConvexHull, HalfspaceIntersection
    rng = np.random.default_rng(21)
    x = rng.random((150, 3))
    v = O.build(x, ZERO, ONE, periodic=periodic)
    imgs = [np.array([i, j, k], float) for i in (-1, 0, 1) for j in (-1, 0, 1)
            for k in (-1, 0, 1)] if periodic else [np.zeros(3)]
    pts = np.concatenate([x + s for s in imgs])
    nface = 0
    for c in range(0, 150, 7):
        q = x[c]
        hs = []
        for p in pts:
            d = p - q
            if np.dot(d, d) == 0:
                continue
            hs.append(np.concatenate([d, [-np.dot(d, 0.5 * (p + q))]]))
        if not periodic:
            for ax in range(3):
                e = np.zeros(3)
                e[ax] = 1
                hs.append(np.concatenate([-e, [0.0]]))
                hs.append(np.concatenate([e, [-1.0]]))
        hsi = HalfspaceIntersection(np.array(hs), q)
        hull = ConvexHull(hsi.intersections)
        assert v.cell_volume[c] == pytest.approx(hull.volume, rel=1e-9)
        tri = hull.points[hull.simplices]                      # (T, 3, 3)
        cr = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
        t_area = 0.5 * np.linalg.norm(cr, axis=1)
        t_cen = tri.mean(axis=1)
        # cell centroid from tets (interior point m, triangle)
        m = hull.points[hull.vertices].mean(axis=0)
        t_vol = np.abs(np.einsum("ij,ij->i", tri[:, 0] - m, cr)) / 6.0
        cen = np.sum(t_vol[:, None] * (tri.sum(axis=1) + m), 0) / (4.0 * t_vol.sum())
        assert np.allclose(v.cell_centroid[c], cen, rtol=0, atol=1e-9)
        # faces of cell c; the face normal points towards the left generator
        own = np.nonzero(v.face_left == c)[0]
        other = np.nonzero((v.face_right == c) & (v.face_has_shift == 0))[0]
        out_n = np.concatenate([-v.face_normal[own], v.face_normal[other]])
        fids = np.concatenate([own, other])
        claimed = np.zeros(len(tri), bool)
        for f, nrm in zip(fids, out_n):
            sel = np.linalg.norm(hull.equations[:, :3] - nrm, axis=1) < 1e-7
            claimed |= sel
            if v.face_area[f] < 1e-6:
                continue  # Qhull may merge a sliver into its neighbour
            a = t_area[sel].sum()
            assert v.face_area[f] == pytest.approx(a, rel=1e-8)
            fc = np.sum(t_area[sel, None] * t_cen[sel], 0) / a
            assert np.allclose(v.face_centroid[f], fc, rtol=0, atol=1e-9)
            nface += 1
        # every hull triangle of a face of non-negligible size was claimed
        assert t_area[~claimed].sum() < 1e-6
    assert nface > 200


def test_perturbed_grid_32_invariants():
    """32^3 perturbed grid (+-0.1 spacing), rectangular: box volume and
    closure."""
    x = W.perturbed_grid_32()
    v = O.build(x, ZERO, ONE)
    tot = float(np.sum(v.cell_volume))
    assert abs(tot - 1.0) <= 1e-11
    assert v.cell_n_popped.max() <= 40


def test_duplicate_points_are_an_error():
    x = [[0.2, 0.2, 0.2], [0.5, 0.5, 0.5], [0.5, 0.5, 0.5]]
    with pytest.raises(O.OracleError):
        O.build(x, ZERO, ONE)


def test_flop_count_per_uniform_cell():
    """The algorithmic FP64 count per cell of the uniform periodic workload
    (used as the roofline numerator): stable to within a few percent."""
    x = W.uniform(20000, 2)
    v = O.build(x, ZERO, ONE, periodic=True, flops=True)
    per_cell = v.flops / x.shape[0]
    assert 2.0e4 < per_cell < 4.0e4


# the cells at which the jittered 64^3 grid fails (index in the built order),
# per (periodic, permutation seed); found by full oracle builds
_JITTER_FAILS = {(False, None): 57798, (False, 1): 14149, (False, 2): 11141,
                 (True, None): 44063, (True, 1): 14149, (True, 2): 23604}


@pytest.mark.parametrize("periodic", [False, True])
@pytest.mark.parametrize("perm_seed", [None, 1, 2])
def test_grid_jitter_64_boundary_walk_fails(periodic, perm_seed):
    """The 64^3 grid with 1 % of the seeds moved by 1e-12 makes the
    reference algorithm's boundary walk fail (compute_boundary's panic,
    convex_cell.rs:428; status 6) in both boundary modes, and not because of
    the tie order among equal keys: under two random permutations of the
    seeds the build fails as well (at another cell).  The identity order in
    the rectangular box runs the whole build; the other cases build the
    failing cell alone (the per-cell construction does not depend on the
    mask)."""
    x = W.grid_jitter(64)
    if perm_seed is not None:
        x = x[np.random.default_rng(perm_seed).permutation(len(x))]
    cell = _JITTER_FAILS[(periodic, perm_seed)]
    mask = None
    if (periodic, perm_seed) != (False, None):
        mask = np.zeros(len(x), bool)
        mask[cell] = True
    with pytest.raises(O.OracleError) as e:
        O.build(x, ZERO, ONE, periodic=periodic, mask=mask)
    assert e.value.status == 6
    assert e.value.cell == cell
    # the cells before it in a partial build of a few neighbours succeed
    m = np.zeros(len(x), bool)
    m[max(0, cell - 3):cell] = True
    O.build(x, ZERO, ONE, periodic=periodic, mask=m)
