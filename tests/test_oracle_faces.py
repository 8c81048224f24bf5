"""The oracle's restatement of ConvexCell::with_faces
(src/voronoi/convex_cell.rs:461-543) and of DecompositionWithFaces
(convex_cell.rs:89-129), pinned on the CPU by:

  - the reference's own cuboid test (convex_cell.rs:806-822) and a
    hand-walked known answer for it;
  - an independent restatement of the face order: the vertices of a
    half-space sorted by angle around their mean, counter-clockwise seen
    from the generator (the side the inward normal points to), rotated to
    start at the lowest vertex index, which is where the reference's walk
    starts (its list is collected in vertex order);
  - the reference's test_extract_vertices (tests/test_extract_vertices.rs):
    fan tets give the volumes and face areas of the decomposition without
    faces (here within 1e-12 relative instead of the reference's 1e-10).
"""
import numpy as np
import pytest

from oracle import oracle as O
from meshless_voro_amd import workloads as W


def face_segments(T, i):
    """(plane, vertex indices within the cell) for each face of cell i."""
    cb = int(T.conn_offset[i])
    for f in range(int(T.face_offset[i]), int(T.face_offset[i + 1])):
        o, m = int(T.face_vertex_offset[f]), int(T.face_count[f])
        yield int(T.face_plane[f]), T.conn[cb + o:cb + o + m]


def test_cuboid_reference_test():
    """convex_cell.rs:806-822: the box cell of one generator has 6 faces of
    4 distinct vertices, each on its face's half-space.  The walk from
    vertex 0 of the +x half-space (dual (2, 5, 0)) goes to the vertex on
    planes 0 and dual[(2 + 1) % 3] = 2, i.e. vertex 4 (4, 2, 0), and so on."""
    T = O.build_tets(np.array([[2.0, 2.0, 2.0]]), [1.0] * 3, [2.0] * 3, 3, False,
                     with_faces=True)
    faces = list(face_segments(T, 0))
    assert [p for p, _ in faces] == list(range(6))
    for p, vs in faces:
        assert len(vs) == 4 and len(set(vs.tolist())) == 4
        for v in vs:
            assert p in T.vert_dual[v]
    assert faces[0][1].tolist() == [0, 4, 5, 1]
    # 12 fan tets tiling the 2 x 2 x 2 box around the generator
    g = np.array([2.0, 2.0, 2.0])
    e1 = T.tet_v[:, 1] - T.tet_v[:, 0]
    e2 = T.tet_v[:, 2] - T.tet_v[:, 0]
    vol = np.einsum("ij,ij->i", g - T.tet_v[:, 0], np.cross(e1, e2)) / 6.0
    assert len(vol) == 12 and np.all(vol > 0) and abs(vol.sum() - 8.0) < 1e-14


def angle_sorted_face(loc, n, vs):
    """Vertices vs (indices into loc) of a convex face with inward normal n,
    counter-clockwise seen from the side n points to, from the lowest index."""
    P = loc[vs]
    c = P.mean(axis=0)
    u = P[0] - c
    u /= np.linalg.norm(u)
    w = np.cross(n, u)
    ang = np.arctan2((P - c) @ w, (P - c) @ u)
    order = [int(vs[k]) for k in np.argsort(ang)]
    s = order.index(min(order))
    return order[s:] + order[:s]


@pytest.mark.parametrize("periodic", [False, True])
def test_face_order_equals_angle_sort(periodic):
    x = W.uniform(400, 21 + periodic)
    T = O.build_tets(x, [0.0] * 3, [1.0] * 3, 3, periodic, with_faces=True)
    nfaces = 0
    for i in range(T.n_cells):
        vb, pb = int(T.vert_offset[i]), int(T.plane_offset[i])
        loc = T.vert_loc[vb:int(T.vert_offset[i + 1])]
        dual = T.vert_dual[vb:int(T.vert_offset[i + 1])]
        planes = [p for p, _ in face_segments(T, i)]
        # faces: exactly the half-spaces holding a vertex, in half-space order
        assert planes == sorted(set(dual.ravel().tolist()))
        for p, vs in face_segments(T, i):
            on = np.nonzero((dual == p).any(axis=1))[0]
            assert sorted(vs.tolist()) == on.tolist()
            assert vs.tolist() == angle_sorted_face(loc, T.plane_n[pb + p], on)
            nfaces += 1
    assert nfaces > 400 * 12


def test_planted_reversal_is_caught():
    """A walk that took dual[(pos + 2) % 3] would list every face clockwise;
    the angle-sort restatement rejects that order."""
    x = W.uniform(50, 23)
    T = O.build_tets(x, [0.0] * 3, [1.0] * 3, 3, True, with_faces=True)
    vb, pb = int(T.vert_offset[0]), int(T.plane_offset[0])
    loc = T.vert_loc[vb:int(T.vert_offset[1])]
    p, vs = next(face_segments(T, 0))
    rev = [int(vs[0])] + [int(v) for v in vs[1:][::-1]]
    assert rev != angle_sorted_face(loc, T.plane_n[pb + p], np.sort(vs))


@pytest.mark.parametrize("periodic", [False, True])
def test_fan_integrals_equal_decomposition_without_faces(periodic):
    """tests/test_extract_vertices.rs:104-132: volumes, face areas and face
    centroids through the fan equal those of the projection tets."""
    x = W.uniform(1000, 42 + periodic) * 3.0
    a, w = [0.0] * 3, [3.0] * 3
    A = O.build_tets(x, a, w, 3, periodic)
    B = O.build_tets(x, a, w, 3, periodic, with_faces=True)

    def per_cell(T):
        g = np.repeat(x, np.diff(T.tet_offset.astype(np.int64)), axis=0)
        v0, v1, v2 = T.tet_v[:, 0], T.tet_v[:, 1], T.tet_v[:, 2]
        vol = np.einsum("ij,ij->i", g - v0, np.cross(v1 - v0, v2 - v0)) / 6.0
        cell = np.add.reduceat(vol, T.tet_offset[:-1].astype(np.int64))
        # face areas: |e1 x e2| / 2, signed by the side of the generator
        cr = np.cross(v1 - v0, v2 - v0)
        area = 0.5 * np.linalg.norm(cr, axis=1) * np.sign(np.einsum("ij,ij->i", g - v0, cr))
        cid = np.repeat(np.arange(len(x)), np.diff(T.tet_offset.astype(np.int64)))
        key = cid * 1000 + T.tet_plane
        ks, inv = np.unique(key, return_inverse=True)
        fa = np.bincount(inv, weights=area)
        # face centroid: area-weighted mean of the triangles' centroids
        tc = (v0 + v1 + v2) / 3.0
        fc = np.stack([np.bincount(inv, weights=area * tc[:, i]) for i in range(3)], 1)
        return cell, ks, fa, fc / np.where(fa > 0, fa, 1.0)[:, None]

    ca, ka, fa, fca = per_cell(A)
    cb, kb, fb, fcb = per_cell(B)
    assert np.allclose(cb, ca, rtol=1e-12, atol=0)
    assert abs(cb.sum() - 27.0) < 1e-11
    assert np.array_equal(ka, kb)
    assert np.allclose(fb, fa, rtol=1e-10, atol=1e-13)
    big = fa > 1e-8
    assert np.allclose(fcb[big], fca[big], rtol=0, atol=1e-9)


def test_with_faces_needs_3d():
    x = np.random.default_rng(3).random((20, 3))
    with pytest.raises(O.OracleError) as e:
        O.build_tets(x, [0.0] * 3, [1.0] * 3, 2, False, with_faces=True)
    assert e.value.status == 11
