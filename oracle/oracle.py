"""ctypes view of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by bench.py's cpu_baseline leg, always as the checker or the timed CPU
baseline, never as the product path.  The oracle restates the per-cell
construction of the meshless_voronoi crate v0.7.2 (see mvoro_oracle.c for the
file:line citations).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))


def compile_lib(force: bool = False) -> None:
    """Compile liboracle.so / liboracle_flops.so with the Makefile here.

    make rebuilds a library whose sources are newer, so an edited oracle is
    never tested stale.  Where make is absent (a box that only runs the
    prebuilt libraries) existing libraries are used as they are."""
    libs = [os.path.join(_HERE, n) for n in ("liboracle.so", "liboracle_flops.so")]
    try:
        subprocess.run(["make", "-s", "-B" if force else "all"], cwd=_HERE, check=True)
    except FileNotFoundError:
        if not all(os.path.exists(p) for p in libs):
            raise


class _OrcVoronoi(C.Structure):
    _fields_ = [
        ("status", C.c_int),
        ("status_cell", C.c_int64),
        ("anchor", C.c_double * 3),
        ("width", C.c_double * 3),
        ("dim", C.c_int),
        ("periodic", C.c_int),
        ("n_cells", C.c_size_t),
        ("n_faces", C.c_size_t),
        ("n_links", C.c_size_t),
        ("cell_loc", C.POINTER(C.c_double)),
        ("cell_centroid", C.POINTER(C.c_double)),
        ("cell_volume", C.POINTER(C.c_double)),
        ("cell_safety_radius", C.POINTER(C.c_double)),
        ("cell_face_offset", C.POINTER(C.c_uint64)),
        ("cell_face_count", C.POINTER(C.c_uint64)),
        ("cell_idx", C.POINTER(C.c_uint64)),
        ("face_left", C.POINTER(C.c_uint64)),
        ("face_right", C.POINTER(C.c_int64)),
        ("face_area", C.POINTER(C.c_double)),
        ("face_centroid", C.POINTER(C.c_double)),
        ("face_normal", C.POINTER(C.c_double)),
        ("face_has_shift", C.POINTER(C.c_uint8)),
        ("face_shift", C.POINTER(C.c_double)),
        ("links", C.POINTER(C.c_uint64)),
        ("cell_n_popped", C.POINTER(C.c_uint32)),
        ("cell_n_planes", C.POINTER(C.c_uint32)),
        ("cell_n_vertices", C.POINTER(C.c_uint32)),
        ("cell_exhausted", C.POINTER(C.c_uint8)),
        ("flops", C.c_uint64),
        ("exact_tests", C.c_uint64),
    ]


class _OrcTets(C.Structure):
    _fields_ = [
        ("status", C.c_int),
        ("n_cells", C.c_size_t),
        ("tet_offset", C.POINTER(C.c_uint64)),
        ("tet_v", C.POINTER(C.c_double)),
        ("tet_plane", C.POINTER(C.c_int32)),
        ("plane_offset", C.POINTER(C.c_uint64)),
        ("plane_n", C.POINTER(C.c_double)),
        ("plane_p", C.POINTER(C.c_double)),
        ("plane_right", C.POINTER(C.c_int64)),
        ("plane_has_shift", C.POINTER(C.c_uint8)),
        ("plane_shift", C.POINTER(C.c_double)),
        ("vert_offset", C.POINTER(C.c_uint64)),
        ("vert_loc", C.POINTER(C.c_double)),
        ("vert_dual", C.POINTER(C.c_int32)),
    ]


class _OrcFaces(C.Structure):
    _fields_ = [
        ("status", C.c_int),
        ("n_cells", C.c_size_t),
        ("face_offset", C.POINTER(C.c_uint64)),
        ("face_plane", C.POINTER(C.c_int32)),
        ("face_count", C.POINTER(C.c_uint32)),
        ("face_vertex_offset", C.POINTER(C.c_uint64)),
        ("conn_offset", C.POINTER(C.c_uint64)),
        ("conn", C.POINTER(C.c_uint32)),
        ("tet_offset", C.POINTER(C.c_uint64)),
        ("tet_v", C.POINTER(C.c_double)),
        ("tet_plane", C.POINTER(C.c_int32)),
    ]


_libs: dict = {}

_D3 = C.c_double * 3
_I3 = C.c_int64 * 3


def lib(flops: bool = False):
    name = "liboracle_flops.so" if flops else "liboracle.so"
    if name not in _libs:
        path = os.path.join(_HERE, name)
        if not os.path.exists(path):
            compile_lib()
        L = C.CDLL(path)
        L.orc_build.restype = C.c_int
        L.orc_build.argtypes = [
            C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p, C.c_void_p,
            C.c_int, C.c_int, C.c_int, C.POINTER(C.POINTER(_OrcVoronoi))]
        L.orc_free.argtypes = [C.POINTER(_OrcVoronoi)]
        L.orc_build_tets.restype = C.c_int
        L.orc_build_tets.argtypes = [
            C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p, C.c_void_p,
            C.c_int, C.c_int, C.c_int, C.POINTER(C.POINTER(_OrcTets))]
        L.orc_free_tets.argtypes = [C.POINTER(_OrcTets)]
        L.orc_with_faces.restype = C.c_int
        L.orc_with_faces.argtypes = [C.POINTER(_OrcTets), C.c_int,
                                     C.POINTER(C.POINTER(_OrcFaces))]
        L.orc_free_faces.argtypes = [C.POINTER(_OrcFaces)]
        L.orc_in_sphere_exact.restype = C.c_int
        L.orc_in_sphere_exact.argtypes = [C.c_void_p] * 5
        L.orc_iloc.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
        L.orc_signed_volume_tet.restype = C.c_double
        L.orc_signed_volume_tet.argtypes = [C.c_void_p] * 4
        L.orc_signed_area_tri.restype = C.c_double
        L.orc_signed_area_tri.argtypes = [C.c_void_p] * 4
        L.orc_intersect_planes.restype = C.c_int
        L.orc_intersect_planes.argtypes = [C.c_void_p] * 7
        L.orc_halfspace_clip.restype = C.c_int
        L.orc_halfspace_clip.argtypes = [C.c_void_p] * 3
        L.orc_simple_cycle.restype = C.c_int
        L.orc_simple_cycle.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p]
        L.orc_neighbour_stream.restype = C.c_size_t
        L.orc_neighbour_stream.argtypes = [
            C.c_void_p, C.c_size_t, C.c_void_p, C.c_int, C.c_int, C.c_size_t,
            C.c_size_t, C.c_void_p, C.c_void_p, C.c_void_p]
        _libs[name] = L
    return _libs[name]


def _arr(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


class OracleError(RuntimeError):
    def __init__(self, status, cell):
        super().__init__(f"oracle status {status} at cell {cell}")
        self.status = status
        self.cell = cell


class OracleVoronoi:
    """Result of the oracle's Voronoi::build, copied into numpy arrays."""

    def __init__(self, V: _OrcVoronoi):
        n, F, L = V.n_cells, V.n_faces, V.n_links
        self.anchor = np.array(V.anchor[:])
        self.width = np.array(V.width[:])
        self.dimensionality = V.dim
        self.periodic = bool(V.periodic)
        self.cell_loc = _arr(V.cell_loc, 3 * n, np.float64).reshape(n, 3)
        self.cell_centroid = _arr(V.cell_centroid, 3 * n, np.float64).reshape(n, 3)
        self.cell_volume = _arr(V.cell_volume, n, np.float64)
        self.cell_safety_radius = _arr(V.cell_safety_radius, n, np.float64)
        self.cell_face_offset = _arr(V.cell_face_offset, n, np.uint64)
        self.cell_face_count = _arr(V.cell_face_count, n, np.uint64)
        self.cell_idx = _arr(V.cell_idx, n, np.uint64)
        self.face_left = _arr(V.face_left, F, np.uint64)
        self.face_right = _arr(V.face_right, F, np.int64)
        self.face_area = _arr(V.face_area, F, np.float64)
        self.face_centroid = _arr(V.face_centroid, 3 * F, np.float64).reshape(F, 3)
        self.face_normal = _arr(V.face_normal, 3 * F, np.float64).reshape(F, 3)
        self.face_has_shift = _arr(V.face_has_shift, F, np.uint8)
        self.face_shift = _arr(V.face_shift, 3 * F, np.float64).reshape(F, 3)
        self.links = _arr(V.links, L, np.uint64)
        self.cell_n_popped = _arr(V.cell_n_popped, n, np.uint32)
        self.cell_n_planes = _arr(V.cell_n_planes, n, np.uint32)
        self.cell_n_vertices = _arr(V.cell_n_vertices, n, np.uint32)
        self.cell_exhausted = _arr(V.cell_exhausted, n, np.uint8)
        self.flops = int(V.flops)
        self.exact_tests = int(V.exact_tests)


def _prep(xyz, mask, anchor, width):
    xyz = np.ascontiguousarray(np.asarray(xyz, dtype=np.float64).reshape(-1, 3))
    m = None
    if mask is not None:
        m = np.ascontiguousarray(np.asarray(mask, dtype=np.uint8))
    a = np.ascontiguousarray(np.asarray(anchor, dtype=np.float64).reshape(3))
    w = np.ascontiguousarray(np.asarray(width, dtype=np.float64).reshape(3))
    return xyz, m, a, w


def build(xyz, anchor, width, dimensionality=3, periodic=False, mask=None,
          nthreads=0, flops=False) -> OracleVoronoi:
    """Oracle Voronoi::build / build_partial (voronoi.rs:151-190)."""
    xyz, m, a, w = _prep(xyz, mask, anchor, width)
    L = lib(flops)
    out = C.POINTER(_OrcVoronoi)()
    st = L.orc_build(xyz.ctypes.data, xyz.shape[0],
                     m.ctypes.data if m is not None else None,
                     a.ctypes.data, w.ctypes.data, int(dimensionality),
                     int(bool(periodic)), int(nthreads), C.byref(out))
    try:
        if st != 0:
            raise OracleError(st, out.contents.status_cell if out else -1)
        return OracleVoronoi(out.contents)
    finally:
        if out:
            L.orc_free(out)


class OracleTets:
    """Per-cell decomposition (oriented tets in decomposition order) and
    clipping planes, as the reference's VoronoiIntegrator sees them.  With
    with_faces=True the tets are those of DecompositionWithFaces
    (convex_cell.rs:89-129) and the face lists of ConvexCell::with_faces
    (convex_cell.rs:461-543) are attached: face_offset (per cell),
    face_plane, face_count, face_vertex_offset (relative to the cell's
    conn_offset), conn (vertex indices within the cell)."""

    def __init__(self, T: _OrcTets, F=None):
        n = T.n_cells
        self.n_cells = n
        self.plane_offset = _arr(T.plane_offset, n + 1, np.uint64)
        npl = int(self.plane_offset[-1])
        self.plane_n = _arr(T.plane_n, 3 * npl, np.float64).reshape(npl, 3)
        self.plane_p = _arr(T.plane_p, 3 * npl, np.float64).reshape(npl, 3)
        self.plane_right = _arr(T.plane_right, npl, np.int64)
        self.plane_has_shift = _arr(T.plane_has_shift, npl, np.uint8)
        self.plane_shift = _arr(T.plane_shift, 3 * npl, np.float64).reshape(npl, 3)
        self.vert_offset = _arr(T.vert_offset, n + 1, np.uint64)
        nv = int(self.vert_offset[-1])
        self.vert_loc = _arr(T.vert_loc, 3 * nv, np.float64).reshape(nv, 3)
        self.vert_dual = _arr(T.vert_dual, 3 * nv, np.int32).reshape(nv, 3)
        src = T if F is None else F
        self.tet_offset = _arr(src.tet_offset, n + 1, np.uint64)
        nt = int(self.tet_offset[-1])
        self.tet_v = _arr(src.tet_v, 9 * nt, np.float64).reshape(nt, 3, 3)
        self.tet_plane = _arr(src.tet_plane, nt, np.int32)
        self.with_faces = F is not None
        if F is not None:
            self.face_offset = _arr(F.face_offset, n + 1, np.uint64)
            nf = int(self.face_offset[-1])
            self.face_plane = _arr(F.face_plane, nf, np.int32)
            self.face_count = _arr(F.face_count, nf, np.uint32)
            self.face_vertex_offset = _arr(F.face_vertex_offset, nf, np.uint64)
            self.conn_offset = _arr(F.conn_offset, n + 1, np.uint64)
            self.conn = _arr(F.conn, int(self.conn_offset[-1]), np.uint32)


def build_tets(xyz, anchor, width, dimensionality=3, periodic=False, mask=None,
               with_faces=False):
    xyz, m, a, w = _prep(xyz, mask, anchor, width)
    L = lib()
    out = C.POINTER(_OrcTets)()
    fout = C.POINTER(_OrcFaces)()
    st = L.orc_build_tets(xyz.ctypes.data, xyz.shape[0],
                          m.ctypes.data if m is not None else None,
                          a.ctypes.data, w.ctypes.data, int(dimensionality),
                          int(bool(periodic)), 1, C.byref(out))
    try:
        if st != 0:
            raise OracleError(st, -1)
        if not with_faces:
            return OracleTets(out.contents)
        st = L.orc_with_faces(out, int(dimensionality), C.byref(fout))
        if st != 0:
            raise OracleError(st, -1)
        return OracleTets(out.contents, fout.contents)
    finally:
        if fout:
            L.orc_free_faces(fout)
        if out:
            L.orc_free_tets(out)


def in_sphere_exact(a, b, c, d, v) -> int:
    args = [_I3(*[int(x) for x in p]) for p in (a, b, c, d, v)]
    return lib().orc_in_sphere_exact(*[C.byref(x) for x in args])


def iloc(anchor, width, periodic, dim, loc):
    out = _I3()
    lib().orc_iloc(C.byref(_D3(*anchor)), C.byref(_D3(*width)), int(periodic),
                   int(dim), C.byref(_D3(*loc)), C.byref(out))
    return [out[0], out[1], out[2]]


def signed_volume_tet(v0, v1, v2, v3):
    return lib().orc_signed_volume_tet(*[C.byref(_D3(*p)) for p in (v0, v1, v2, v3)])


def signed_area_tri(v0, v1, v2, t):
    return lib().orc_signed_area_tri(*[C.byref(_D3(*p)) for p in (v0, v1, v2, t)])


def intersect_planes(n0, p0, n1, p1, n2, p2):
    out = _D3()
    ok = lib().orc_intersect_planes(*[C.byref(_D3(*p)) for p in (n0, p0, n1, p1, n2, p2)],
                                    C.byref(out))
    return (out[0], out[1], out[2]) if ok else None


def halfspace_clip(n, p, v) -> int:
    return lib().orc_halfspace_clip(*[C.byref(_D3(*x)) for x in (n, p, v)])


def simple_cycle(tris, capacity):
    t = np.ascontiguousarray(np.asarray(tris, dtype=np.int32).reshape(-1, 3))
    out = np.zeros(capacity + 4, dtype=np.int32)
    n = lib().orc_simple_cycle(t.ctypes.data, t.shape[0], int(capacity), out.ctypes.data)
    return None if n < 0 else out[:n].tolist()


def neighbour_stream(xyz, width, query, max_out, dimensionality=3, periodic=False):
    xyz = np.ascontiguousarray(np.asarray(xyz, dtype=np.float64).reshape(-1, 3))
    w = np.ascontiguousarray(np.asarray(width, dtype=np.float64).reshape(3))
    idx = np.zeros(max_out, np.uint64)
    img = np.zeros(max_out, np.int32)
    key = np.zeros(max_out, np.float64)
    n = lib().orc_neighbour_stream(xyz.ctypes.data, xyz.shape[0], w.ctypes.data,
                                   int(dimensionality), int(bool(periodic)), int(query),
                                   int(max_out), idx.ctypes.data, img.ctypes.data,
                                   key.ctypes.data)
    return idx[:n], img[:n], key[:n]
