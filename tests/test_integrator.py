"""Custom integral hooks (VoronoiIntegrator): the GPU instantiations of the
reference's example integrals and host-folded custom integrals against the
built-in integrals of Voronoi.build, bit for bit (the reference's
test_integrator, src/voronoi.rs:1187-1245, plus larger and 2D cases)."""
import numpy as np
import pytest

from meshless_voro_amd import BuildOptions, Voronoi, workloads as W
from meshless_voro_amd.integrator import (AreaCentroidIntegral, AreaIntegral,
                                          VolumeCentroidIntegral, VolumeIntegral,
                                          VoronoiIntegrator)
from meshless_voro_amd.voronoi import MODE_LEARNED

pytestmark = pytest.mark.gpu
ZERO, ONE = [0.0, 0.0, 0.0], [1.0, 1.0, 1.0]


def canon_faces(v):
    """Voronoi faces in (left, plane order) = array order."""
    return v.face_left, v.face_right, v.face_area, v.face_centroid


def test_reference_test_integrator():
    """src/voronoi.rs:1187-1245: one active cell of a perturbed 5^3 grid in a
    periodic box of width 2; the integrator's face and cell integrals equal
    the Voronoi build's, exactly."""
    anchor, width = [0.0, 0.0, 0.0], [2.0, 2.0, 2.0]
    x = W.perturbed_grid(5, 0.95, seed=11, anchor=anchor, width=width)
    mask = np.zeros(125, bool)
    mask[62] = True
    v = Voronoi.build_partial(x, mask, anchor, width, 3, True)
    it = VoronoiIntegrator.build(x, mask, anchor, width, 3, True)
    faces = it.compute_face_integrals(AreaCentroidIntegral)
    cells = it.compute_cell_integrals(VolumeCentroidIntegral)
    assert len(faces) == len(v.face_area)
    for i, f in enumerate(faces):
        assert f.integral().area == v.face_area[i]
        assert np.array_equal(f.integral().centroid, v.face_centroid[i])
        assert f.left() == v.face_left[i]
        assert (f.right() if f.right() is not None else -1) == v.face_right[i]
    assert len(cells) == 1
    assert cells[0].volume == v.cell_volume[62]
    assert np.array_equal(cells[0].centroid, v.cell_centroid[62])


@pytest.mark.parametrize("periodic", [False, True])
@pytest.mark.parametrize("exact_order", [False, True])
def test_gpu_integrals_equal_build(periodic, exact_order):
    x = W.uniform(20000, 300 + periodic)
    opt = BuildOptions(exact_order=exact_order)
    print(f"case integrator n={len(x)} periodic={periodic} exact_order={exact_order}",
          flush=True)
    v = Voronoi.build(x, ZERO, ONE, 3, periodic, options=opt)
    it = VoronoiIntegrator.build(x, None, ZERO, ONE, 3, periodic, options=opt)
    assert it.n_active() == len(x)
    c = it.compute_cell_integrals_array(VolumeCentroidIntegral)
    assert np.array_equal(c[:, 0], v.cell_volume)
    assert np.array_equal(c[:, 1:], v.cell_centroid)
    cv = it.compute_cell_integrals_array(VolumeIntegral)
    assert np.array_equal(cv[:, 0], v.cell_volume)
    # symmetric faces = the faces of Voronoi.build, in the same order
    left, right, hs, shift, plane, vals = it.compute_face_integrals_array(
        AreaCentroidIntegral, symmetric=True)
    assert np.array_equal(left, v.face_left)
    assert np.array_equal(right, v.face_right)
    assert np.array_equal(hs, v.face_has_shift)
    assert np.array_equal(shift, v.face_shift)
    assert np.array_equal(vals[:, 0], v.face_area)
    assert np.array_equal(vals[:, 1:], v.face_centroid)
    # all faces of every cell: each internal face twice, with equal areas
    l2, r2, hs2, sh2, p2, a2 = it.compute_face_integrals_array(AreaIntegral, symmetric=False)
    internal = (right >= 0) & (hs == 0)
    assert len(l2) == len(left) + int(np.sum(internal))
    key = {(int(a), int(b), int(p)): val for a, b, p, val in zip(l2, r2, p2, a2[:, 0])}
    by_pair = {}
    for a, b, val, h in zip(l2, r2, a2[:, 0], hs2):
        if b >= 0 and not h:
            by_pair[(int(a), int(b))] = val
    for a, b, val, h in zip(left, right, vals[:, 0], hs):
        if b >= 0 and not h:
            other = by_pair[(int(b), int(a))]
            # the two cells sum different tets: equal up to rounding on the
            # scale of a face (h^2), which tiny faces cannot resolve relatively
            assert abs(other - val) <= 1e-10 * max(abs(val), len(x) ** (-2.0 / 3.0))
    assert len(key) == len(l2)


class MyVolumeCentroid(VolumeCentroidIntegral):
    """A custom (host-folded) integral with the reference's formulas."""


class MyAreaCentroid(AreaCentroidIntegral):
    pass


class Mass:
    """CellIntegralWithData: density (per generator) x volume."""

    def __init__(self, rho):
        self.rho, self.mass = rho, 0.0

    @classmethod
    def init_with_data(cls, cell, rho):
        return cls(rho)

    def collect(self, v0, v1, v2, gen):
        from meshless_voro_amd.integrator import _signed_volume_tet
        self.mass += self.rho * _signed_volume_tet(v0, v1, v2, gen)

    def finalize(self):
        return self


@pytest.mark.parametrize("periodic", [False, True])
def test_custom_integrals_fold_the_same_tets(periodic):
    """Custom integrals receive the reference's tets in the reference's
    order: a host re-implementation of the example integrals reproduces the
    GPU's bits, and per-cell data reaches init_with_data."""
    x = W.uniform(600, 400 + periodic)
    mask = W.partial_mask(600, 0.5, 401)
    it = VoronoiIntegrator.build(x, mask, ZERO, ONE, 3, periodic)
    gpu = it.compute_cell_integrals(VolumeCentroidIntegral)
    host = it.compute_cell_integrals(MyVolumeCentroid)
    assert len(gpu) == len(host) == int(mask.sum())
    for a, b in zip(gpu, host):
        assert a.volume == b.volume and np.array_equal(a.centroid, b.centroid)
    for sym in (False, True):
        fg = (it.compute_face_integrals_sym if sym else it.compute_face_integrals)(
            AreaCentroidIntegral)
        fh = (it.compute_face_integrals_sym if sym else it.compute_face_integrals)(MyAreaCentroid)
        assert len(fg) == len(fh)
        for a, b in zip(fg, fh):
            assert (a.left(), a.right()) == (b.left(), b.right())
            assert a.integral().area == b.integral().area
            assert np.array_equal(a.integral().centroid, b.integral().centroid)
    rho = np.linspace(1.0, 2.0, 600)
    m = it.compute_cell_integrals_with_data(Mass, rho)
    act = np.nonzero(mask)[0]
    for k, c in enumerate(act):
        assert abs(m[k].mass - rho[c] * gpu[k].volume) <= 1e-13 * abs(m[k].mass)


def test_tets_count_and_volume():
    x = W.uniform(3000, 410)
    it = VoronoiIntegrator.build(x, None, ZERO, ONE, 3, True)
    v012, plane, off = it.tets()
    assert off[-1] == len(v012) == len(plane)
    assert len(v012) % 6 == 0
    # the tets tile the periodic box
    g = np.repeat(x, np.diff(off.astype(np.int64)), axis=0)
    # signed_volume_tet (geometry.rs:82-88): det(v1 - v0, v2 - v0, gen - v0) / 6
    e1, e2, e3 = v012[:, 1] - v012[:, 0], v012[:, 2] - v012[:, 0], g - v012[:, 0]
    vol = np.einsum("ij,ij->i", e3, np.cross(e1, e2)) / 6.0
    assert abs(vol.sum() - 1.0) < 1e-12


@pytest.mark.parametrize("dim", [2, 1])
def test_low_dimension_faces(dim):
    rng = np.random.default_rng(420 + dim)
    x = rng.random((400, 3))
    for periodic in (False, True):
        v = Voronoi.build(x, ZERO, ONE, dim, periodic)
        it = VoronoiIntegrator.build(x, None, ZERO, ONE, dim, periodic)
        left, right, hs, shift, plane, vals = it.compute_face_integrals_array(
            AreaCentroidIntegral, symmetric=True)
        assert np.array_equal(left, v.face_left)
        assert np.array_equal(right, v.face_right)
        assert np.array_equal(vals[:, 0], v.face_area)
        c = it.compute_cell_integrals_array(VolumeCentroidIntegral)
        assert np.array_equal(c[:, 0], v.cell_volume)


def test_learned_mode_keeps_the_same_cells():
    x = W.uniform(30000, 430)
    a = VoronoiIntegrator.build(x, None, ZERO, ONE, 3, True)
    b = VoronoiIntegrator.build(x, None, ZERO, ONE, 3, True,
                                options=BuildOptions(mode=MODE_LEARNED))
    assert np.array_equal(a.compute_cell_integrals_array(), b.compute_cell_integrals_array())
    fa = a.compute_face_integrals_array(symmetric=True)
    fb = b.compute_face_integrals_array(symmetric=True)
    for p, q in zip(fa, fb):
        assert np.array_equal(p, q)


def test_large_cells_keep_the_same_integrals():
    """Clusters with empty space between them: the rim cells exceed the
    wavefront kernel's 64 vertices and are built one per wavefront by the
    large-cell kernel, which also keeps their half-spaces and vertices; the
    hook path still equals the built-in integrals bit for bit."""
    rng = np.random.default_rng(78)
    n = 12000
    c = rng.random((8, 3))
    x = np.mod(c[rng.integers(0, 8, n)] + 0.03 * rng.normal(size=(n, 3)), 1.0)
    x[x >= 1.0] = 0.0
    v = Voronoi.build(x, ZERO, ONE, 3, True)
    assert v.stats["big_cells"] > 0
    it = VoronoiIntegrator.build(x, None, ZERO, ONE, 3, True)
    c = it.compute_cell_integrals_array(VolumeCentroidIntegral)
    assert np.array_equal(c[:, 0], v.cell_volume)
    assert np.array_equal(c[:, 1:], v.cell_centroid)
    left, right, hs, shift, plane, vals = it.compute_face_integrals_array(
        AreaCentroidIntegral, symmetric=True)
    assert np.array_equal(left, v.face_left)
    assert np.array_equal(right, v.face_right)
    assert np.array_equal(vals[:, 0], v.face_area)
