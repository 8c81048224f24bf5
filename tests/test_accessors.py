"""The Python accessors of VoronoiCell / VoronoiFace
(src/voronoi/voronoi_cell.rs:107-165, src/voronoi/voronoi_face.rs:72-131)
against closed forms on perfect grids: cell (i, j, k) of an n^3 grid has the
6-connected grid cells as neighbours, a boundary face per coordinate at 0 or
n-1 in a rectangular box, and a periodic face per such coordinate in a
periodic box (whose neighbours across it are excluded from neighbour_ids).
The CPU cases wrap the oracle's arrays in the same Voronoi class; the GPU
cases read a build of libmvoro."""
import itertools

import numpy as np
import pytest

from meshless_voro_amd import Voronoi, workloads as W
from meshless_voro_amd.voronoi import Dimensionality
from oracle import oracle as O

ZERO, ONE = [0.0, 0.0, 0.0], [1.0, 1.0, 1.0]
FIELDS = ("cell_loc", "cell_centroid", "cell_volume", "cell_safety_radius", "cell_face_offset",
          "cell_face_count", "cell_idx", "cell_n_popped", "face_left", "face_right", "face_area",
          "face_centroid", "face_normal", "face_has_shift", "face_shift", "links")


def from_oracle(o, anchor, width, periodic):
    v = object.__new__(Voronoi)
    for f in FIELDS:
        setattr(v, f, getattr(o, f))
    v._anchor, v._width = np.array(anchor, float), np.array(width, float)
    v._dimensionality, v._periodic = Dimensionality.ThreeD, periodic
    v.stats = {}
    return v


def grid_expectations(n, periodic):
    ids = {}
    for i, j, k in itertools.product(range(n), repeat=3):
        me = (i * n + j) * n + k
        nb, edge = set(), 0
        for ax in range(3):
            for s in (-1, 1):
                c = [i, j, k]
                c[ax] += s
                if 0 <= c[ax] < n:
                    nb.add((c[0] * n + c[1]) * n + c[2])
                else:
                    edge += 1
        ids[me] = (nb, edge)
    return ids


def check_accessors(v, n, periodic):
    exp = grid_expectations(n, periodic)
    assert v.periodic() == periodic and v.dimensionality() == 3
    assert np.array_equal(v.anchor(), ZERO) and np.array_equal(v.width(), ONE)
    cells = v.cells()
    faces = v.faces()
    assert len(cells) == n ** 3
    for c in cells:
        me = int(v.cell_idx[c._i])
        nb, edge = exp[me]
        ids = list(c.neighbour_ids(v))
        assert sorted(ids) == sorted(nb), me
        fs = list(c.faces(v))
        assert len(fs) == c.face_count() == 6
        assert np.array_equal(c.face_indices(v), v.links[c.face_connections_offset():
                                                         c.face_connections_offset() + 6])
        n_bnd = sum(f.is_boundary() for f in fs)
        n_per = sum(f.is_periodic() for f in fs)
        if periodic:
            assert n_bnd == 0 and n_per == edge
        else:
            assert n_per == 0 and n_bnd == edge
        assert c.volume() == pytest.approx(1.0 / n ** 3, rel=1e-12)
        assert np.allclose(c.centroid(), c.loc(), rtol=0, atol=1e-14)
        assert c.safety_radius() > 0.0
    for f in faces:
        assert f.is_boundary() == (f.right() is None)
        assert f.is_periodic() == (f.shift() is not None)
        if f.is_periodic():
            # the shift brings the right generator next to the left one
            r = v.cell_loc[f.right()] + f.shift()
            assert np.linalg.norm(r - v.cell_loc[f.left()]) == pytest.approx(1.0 / n)
        # the normal points from the face towards the left generator
        assert np.dot(f.normal(), v.cell_loc[f.left()] - f.centroid()) == pytest.approx(0.5 / n)
        assert f.area() == pytest.approx(1.0 / n ** 2, rel=1e-12)


@pytest.mark.parametrize("periodic", [False, True])
def test_accessors_on_oracle_arrays(periodic):
    x = W.grid(4)
    v = from_oracle(O.build(x, ZERO, ONE, periodic=periodic), ZERO, ONE, periodic)
    check_accessors(v, 4, periodic)


def test_neighbour_ids_symmetric_on_random_points():
    """j is a neighbour of i exactly when i is one of j (non-periodic faces
    are built once and linked to both cells)."""
    x = np.random.default_rng(4).random((400, 3))
    v = from_oracle(O.build(x, ZERO, ONE, periodic=True), ZERO, ONE, True)
    nb = {c._i: set(c.neighbour_ids(v)) for c in v.cells()}
    for i, s in nb.items():
        for j in s:
            assert i in nb[j]
    assert sum(len(s) for s in nb.values()) > 400 * 10


@pytest.mark.gpu
@pytest.mark.parametrize("periodic", [False, True])
def test_accessors_on_gpu_build(periodic):
    x = W.grid(5)
    v = Voronoi.build(x, ZERO, ONE, 3, periodic)
    check_accessors(v, 5, periodic)
