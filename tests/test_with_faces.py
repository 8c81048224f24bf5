"""VoronoiIntegrator::with_faces on the GPU (mv_with_faces, mv_face_vertices,
mv_kept_cells, mv_voronoi_from_integrator) against the oracle's restatement
of ConvexCell::with_faces / DecompositionWithFaces
(src/voronoi/convex_cell.rs:89-129, 461-543), plus the reference's
test_extract_vertices (tests/test_extract_vertices.rs) and
test_get_face_vertices_cuboid (convex_cell.rs:806-822).

Integer outputs (duals, face lists, offsets) must equal the oracle's bit for
bit; positions are compared within 1e-12 of the box size."""
import numpy as np
import pytest

from meshless_voro_amd import Voronoi, VoronoiError, workloads as W
from meshless_voro_amd.integrator import (AreaCentroidIntegral, AreaIntegral,
                                          VolumeCentroidIntegral, VolumeIntegral,
                                          VoronoiIntegrator)
from oracle import oracle as O

pytestmark = pytest.mark.gpu
ZERO, ONE = [0.0, 0.0, 0.0], [1.0, 1.0, 1.0]
POS_TOL = 1e-12


def kept_equal_oracle(it, T, scale):
    k = it._kept()
    act = it.active_cells().astype(np.int64)
    # the oracle numbers every generator; keep the active ones
    po = T.plane_offset.astype(np.int64)
    vo = T.vert_offset.astype(np.int64)
    assert np.array_equal(np.diff(k["poff"].astype(np.int64)), (po[1:] - po[:-1])[act])
    assert np.array_equal(np.diff(k["voff"].astype(np.int64)), (vo[1:] - vo[:-1])[act])
    psel = np.concatenate([np.arange(po[c], po[c + 1]) for c in act]) if len(act) else []
    vsel = np.concatenate([np.arange(vo[c], vo[c + 1]) for c in act]) if len(act) else []
    assert np.array_equal(k["right"], T.plane_right[psel])
    assert np.array_equal(k["has_shift"], T.plane_has_shift[psel])
    assert np.array_equal(k["dual"], T.vert_dual[vsel])
    for a, b in ((k["n"], T.plane_n[psel]), (k["p"], T.plane_p[psel]),
                 (k["shift"], T.plane_shift[psel]), (k["loc"], T.vert_loc[vsel])):
        assert np.max(np.abs(a - b), initial=0.0) <= POS_TOL * scale
    return act


def faces_equal_oracle(it, T, act):
    cfo, fp, fc, fo, fv = it.face_vertex_arrays()
    fo_t = T.face_offset.astype(np.int64)
    nf = np.diff(cfo.astype(np.int64))
    assert np.array_equal(nf, (fo_t[1:] - fo_t[:-1])[act])
    fsel = np.concatenate([np.arange(fo_t[c], fo_t[c + 1]) for c in act])
    assert np.array_equal(fp, T.face_plane[fsel])
    assert np.array_equal(fc, T.face_count[fsel])
    # vertex lists: global offsets here, cell-relative in the oracle
    cell_of_face = np.repeat(act, nf)
    want = np.concatenate([
        T.conn[int(T.conn_offset[c]) + int(o):int(T.conn_offset[c]) + int(o) + int(m)]
        for c, o, m in zip(cell_of_face, T.face_vertex_offset[fsel], T.face_count[fsel])])
    got = np.concatenate([fv[int(o):int(o) + int(m)] for o, m in zip(fo, fc)])
    assert np.array_equal(got, want)
    return len(fp)


CASES = {
    "uniform": lambda: (W.uniform(3000, 901), None),
    "partial": lambda: (W.uniform(3000, 902), W.partial_mask(3000, 0.3, 903)),
    "grid": lambda: (W.grid(5), None),
    "clustered": lambda: (W.clustered(6000, 904), None),
}


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("periodic", [False, True])
def test_face_lists_equal_oracle(case, periodic):
    x, mask = CASES[case]()
    print(f"case with_faces {case} n={len(x)} periodic={periodic}", flush=True)
    it = VoronoiIntegrator.build(x, mask, ZERO, ONE, 3, periodic)
    T = O.build_tets(x, ZERO, ONE, 3, periodic, mask=mask, with_faces=True)
    act = kept_equal_oracle(it, T, 1.0)
    it.with_faces()
    assert faces_equal_oracle(it, T, act) >= 6 * len(act)
    # fan tets: the oracle's DecompositionWithFaces
    v, p, off = it.tets()
    tsel = np.concatenate([np.arange(int(T.tet_offset[c]), int(T.tet_offset[c + 1]))
                           for c in act])
    assert np.array_equal(p, T.tet_plane[tsel])
    assert np.max(np.abs(v - T.tet_v[tsel])) <= POS_TOL


def test_large_cells_face_lists():
    """Rim cells of clusters in empty space exceed 64 vertices (large-cell
    kernel); their faces still equal the oracle's."""
    rng = np.random.default_rng(79)
    n = 6000
    c = rng.random((8, 3))
    x = np.mod(c[rng.integers(0, 8, n)] + 0.03 * rng.normal(size=(n, 3)), 1.0)
    x[x >= 1.0] = 0.0
    print("case with_faces large cells", flush=True)
    it = VoronoiIntegrator.build(x, None, ZERO, ONE, 3, True)
    T = O.build_tets(x, ZERO, ONE, 3, True, with_faces=True)
    assert int(np.diff(T.vert_offset.astype(np.int64)).max()) > 64
    act = kept_equal_oracle(it, T, 1.0)
    it.with_faces()
    faces_equal_oracle(it, T, act)


def test_cuboid_reference_test():
    """convex_cell.rs:806-822 through the integrator: the one cell of a
    single generator is the box, 6 faces of 4 distinct vertices each."""
    it = VoronoiIntegrator.build(np.array([[2.0, 2.0, 2.0]]), None, [1.0] * 3, [2.0] * 3, 3,
                                 False).with_faces()
    cell = it.get_cell_at(0)
    assert cell.face_count() == 6
    for i in range(6):
        fv = cell.face_vertices(i)
        assert len(fv) == 4 and len(set(fv.tolist())) == 4
        for vi in fv:
            assert i in cell.vertices[vi].dual
    assert cell.face_vertices(0).tolist() == [0, 4, 5, 1]
    assert cell.compute_cell_integral(VolumeIntegral).volume == 8.0


def test_extract_vertices_reference_test():
    """tests/test_extract_vertices.rs: 1000 generators in a box of width 3;
    Voronoi::from the integrator equals the build before with_faces, and
    the fan decomposition after it gives the same volumes, areas and
    centroids (1e-10 / 1e-5 as the reference asserts)."""
    size = [3.0] * 3
    x = W.uniform(1000, 42) * 3.0
    it = VoronoiIntegrator.build(x, None, ZERO, size, 3, False)
    v2 = Voronoi.from_integrator(it)
    ref = Voronoi.build(x, ZERO, size, 3, False)
    for f in ("cell_volume", "cell_centroid", "face_area", "face_centroid", "face_normal",
              "face_left", "face_right", "links"):
        assert np.array_equal(getattr(v2, f), getattr(ref, f)), f
    volumes = it.compute_cell_integrals(VolumeIntegral)
    areas = it.compute_face_integrals_sym(AreaIntegral)
    it.with_faces()
    raw = []
    for cell_id, cell in enumerate(it.cells_iter()):
        assert cell.idx == cell_id
        assert np.array_equal(cell.loc, v2.cell_loc[cell_id])
        vs = [v.loc - cell.loc for v in cell.vertices]
        faces = [cell.face_vertices(f).tolist() for f in range(int(v2.cell_face_count[cell_id]))]
        raw.append((vs, faces))
    assert len(raw) == 1000
    v3 = Voronoi.from_integrator(it)
    volumes2 = it.compute_cell_integrals(VolumeIntegral)
    areas2 = it.compute_face_integrals_sym(AreaIntegral)
    for a, b in zip(volumes, volumes2):
        assert abs(a.volume - b.volume) <= 1e-10
    assert len(areas) == len(areas2)
    for a, b in zip(areas, areas2):
        assert (a.left(), a.right()) == (b.left(), b.right())
        assert abs(a.integral().area - b.integral().area) <= 1e-10
    assert np.array_equal(v2.cell_face_count, v3.cell_face_count)
    assert np.allclose(v3.cell_volume, v2.cell_volume, rtol=0, atol=1e-10)
    assert np.allclose(v3.cell_centroid, v2.cell_centroid, rtol=0, atol=1e-5)
    assert np.allclose(v3.face_area, v2.face_area, rtol=0, atol=1e-10)
    assert np.allclose(v3.face_centroid, v2.face_centroid, rtol=0, atol=1e-5)
    assert np.array_equal(v3.face_normal, v2.face_normal)
    # tighter than the reference: the two decompositions agree to rounding
    assert np.allclose(v3.cell_volume, v2.cell_volume, rtol=1e-12, atol=0)
    assert abs(v3.cell_volume.sum() - 27.0) < 1e-11


@pytest.mark.parametrize("periodic", [False, True])
def test_fan_integrals_fold_the_same_tets(periodic):
    """GPU fan integrals equal a host fold of the exported fan tets bit for
    bit (same tets, same order); discard_faces restores the integrals
    without faces."""
    x = W.uniform(800, 910 + periodic)
    mask = W.partial_mask(800, 0.5, 912)
    it = VoronoiIntegrator.build(x, mask, ZERO, ONE, 3, periodic)
    before = it.compute_cell_integrals_array(VolumeCentroidIntegral)
    fb = it.compute_face_integrals_array(AreaCentroidIntegral, symmetric=True)
    it.with_faces()
    gpu = it.compute_cell_integrals(VolumeCentroidIntegral)

    class HostVC(VolumeCentroidIntegral):
        pass

    host = it.compute_cell_integrals(HostVC)
    for a, b in zip(gpu, host):
        assert a.volume == b.volume and np.array_equal(a.centroid, b.centroid)
    assert np.allclose([a.volume for a in gpu], before[:, 0], rtol=1e-12, atol=0)
    fg = it.compute_face_integrals_sym(AreaCentroidIntegral)

    class HostAC(AreaCentroidIntegral):
        pass

    fh = it.compute_face_integrals_sym(HostAC)
    for a, b in zip(fg, fh):
        assert a.integral().area == b.integral().area
    # per-cell ConvexCell integrals equal the integrator's
    for k, cell in enumerate(it.cells_iter()):
        if k % 37 == 0:
            c = cell.compute_cell_integral(VolumeCentroidIntegral)
            assert c.volume == gpu[k].volume
    it.discard_faces()
    again = it.compute_cell_integrals_array(VolumeCentroidIntegral)
    assert np.array_equal(again, before)
    fa = it.compute_face_integrals_array(AreaCentroidIntegral, symmetric=True)
    for p, q in zip(fa, fb):
        assert np.array_equal(p, q)


def test_convex_cell_face_accessors():
    x = W.uniform(500, 920)
    it = VoronoiIntegrator.build(x, None, ZERO, ONE, 3, True).with_faces()
    v = Voronoi.build(x, ZERO, ONE, 3, True)
    cell = it.get_cell_at(123)
    assert cell.idx == 123
    nb = sorted(cell.neighbour(f) for f in range(cell.face_count()))
    faces = v.face_left == 123
    mine = v.face_right[faces].tolist()
    theirs = v.face_left[(v.face_right == 123) & (v.face_has_shift == 0)].tolist()
    assert nb == sorted(mine + theirs)
    for f in range(cell.face_count()):
        fv = cell.face_vertices(f)
        assert cell.face_vertex_count(f) == len(fv) >= 3
        pl = cell.clipping_plane(f)
        # every face vertex lies on its plane
        for vi in fv:
            assert abs(np.dot(pl.n, cell.vertices[vi].loc - pl.p)) < 1e-12


def test_with_faces_needs_3d():
    x = np.random.default_rng(5).random((50, 3))
    it = VoronoiIntegrator.build(x, None, ZERO, ONE, 2, False)
    with pytest.raises(VoronoiError) as e:
        it.with_faces()
    assert e.value.status == 11


def test_batch_problems_with_faces():
    """Integrators of a batched build (one per problem, sharing the build's
    device arenas) give the face lists, fan integrals and Voronoi of each
    problem's own build."""
    probs = [(W.uniform(700, 930), ZERO, ONE, True),
             (W.clustered(900, 931), ZERO, ONE, False),
             (W.uniform(500, 932) * 2.0, ZERO, [2.0] * 3, True)]
    masks = [None, W.partial_mask(900, 0.4, 933), None]
    print("case with_faces batch of 3", flush=True)
    its = VoronoiIntegrator.build_batch(probs, masks=masks)
    assert len(its) == 3
    for (x, a, w, per), mk, it in zip(probs, masks, its):
        one = VoronoiIntegrator.build(x, mk, a, w, 3, per)
        assert np.array_equal(it.active_cells(), one.active_cells())
        it.with_faces()
        one.with_faces()
        for p, q in zip(it.face_vertex_arrays(), one.face_vertex_arrays()):
            assert np.array_equal(p, q)
        assert np.array_equal(it.compute_cell_integrals_array(), one.compute_cell_integrals_array())
        vb, vo = Voronoi.from_integrator(it), Voronoi.from_integrator(one)
        for f in ("cell_volume", "cell_centroid", "face_area", "face_centroid", "face_left",
                  "face_right", "links"):
            assert np.array_equal(getattr(vb, f), getattr(vo, f)), f
