"""VoronoiIntegrator: custom cell and face integrals over kept cells.

Mirrors the reference's `VoronoiIntegrator` (src/voronoi.rs:527-829) and the
`CellIntegral` / `FaceIntegral` traits (src/voronoi/integrals.rs:24-259).  A
build with MV_FLAG_KEEP_CELLS keeps every active cell's half-spaces and
vertices on the GPU.  The reference's example integrals (VolumeIntegral,
VolumeCentroidIntegral, AreaIntegral, AreaCentroidIntegral) run as GPU
kernels (instantiations of include/mvoro_integrals.hpp, which C++/HIP callers
instantiate with their own device integrals).  Any other Python class with
the trait's methods

    init(cell)  or  init(cell, plane)   (classmethod returning an instance;
                                         `cell` is the generator index)
    collect(v0, v1, v2, gen)
    finalize()  (returns the finished integral)

receives the same oriented tets, in the same order, decomposed on the GPU
(mv_tets) and folded on the host.
"""
from __future__ import annotations

import ctypes as C
from typing import List, Optional, Sequence

import numpy as np

from . import _native
from .voronoi import BuildOptions, Dimensionality, VoronoiError, _vec3


def _check(st: int) -> None:
    if st != 0:
        raise VoronoiError(st)


# ------------------------------------------------------- example integrals
class VolumeCentroidIntegral:
    """integrals.rs:69-106"""
    kind = 1
    width = 4

    def __init__(self):
        self.volume = 0.0
        self.centroid = np.zeros(3)

    @classmethod
    def init(cls, cell, *args):
        return cls()

    def collect(self, v0, v1, v2, gen):
        v = _signed_volume_tet(v0, v1, v2, gen)
        self.volume += v
        self.centroid = self.centroid + v * (((v0 + v1) + v2) + gen)

    def finalize(self):
        self.centroid = self.centroid * (0.25 / self.volume if self.volume > 0 else 0.0)
        return self

    @classmethod
    def _from_row(cls, r):
        o = cls()
        o.volume = float(r[0])
        o.centroid = np.array(r[1:4])
        return o


class VolumeIntegral:
    """integrals.rs:124-142"""
    kind = 0
    width = 1

    def __init__(self):
        self.volume = 0.0

    @classmethod
    def init(cls, cell, *args):
        return cls()

    def collect(self, v0, v1, v2, gen):
        self.volume += _signed_volume_tet(v0, v1, v2, gen)

    def finalize(self):
        return self

    @classmethod
    def _from_row(cls, r):
        o = cls()
        o.volume = float(r[0])
        return o


class AreaCentroidIntegral:
    """integrals.rs:277-322"""
    kind = 1
    width = 4

    def __init__(self):
        self.area = 0.0
        self.centroid = np.zeros(3)

    @classmethod
    def init(cls, cell, plane, *args):
        return cls()

    def collect(self, v0, v1, v2, gen):
        a = _signed_area_tri(v0, v1, v2, gen)
        self.area += a
        self.centroid = self.centroid + a * ((v0 + v1) + v2)

    def finalize(self):
        self.centroid = self.centroid * (1.0 / (3.0 * self.area) if self.area > 0 else 0.0)
        return self

    @classmethod
    def _from_row(cls, r):
        o = cls()
        o.area = float(r[0])
        o.centroid = np.array(r[1:4])
        return o


class AreaIntegral:
    """integrals.rs:341-361"""
    kind = 0
    width = 1

    def __init__(self):
        self.area = 0.0

    @classmethod
    def init(cls, cell, plane, *args):
        return cls()

    def collect(self, v0, v1, v2, gen):
        self.area += _signed_area_tri(v0, v1, v2, gen)

    def finalize(self):
        return self

    @classmethod
    def _from_row(cls, r):
        o = cls()
        o.area = float(r[0])
        return o


_CELL_BUILTIN = (VolumeCentroidIntegral, VolumeIntegral)
_FACE_BUILTIN = (AreaCentroidIntegral, AreaIntegral)


def _dot(a, b):
    return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]


def _cross(a, b):
    return np.array([a[1] * b[2] - b[1] * a[2], a[2] * b[0] - b[2] * a[0],
                     a[0] * b[1] - b[0] * a[1]])


def _signed_volume_tet(v0, v1, v2, v3):
    """geometry.rs:82-88 (det = c . (a x b), over 6)"""
    return _dot(v3 - v0, _cross(v1 - v0, v2 - v0)) / 6.0


def _signed_area_tri(v0, v1, v2, t):
    """geometry.rs:90-100"""
    n = 0.5 * _cross(v1 - v0, v2 - v0)
    s = _dot(t - v0, n)
    sign = -1.0 if (s < 0 or (s == 0 and np.signbit(s))) else 1.0
    return float(np.sqrt(_dot(n, n))) * sign


class FaceIntegrator:
    """integrals.rs:196-252"""

    def __init__(self, left, right, shift, integral):
        self._left, self._right, self._shift, self._integral = left, right, shift, integral

    def left(self) -> int:
        return self._left

    def right(self) -> Optional[int]:
        return self._right

    def shift(self):
        return self._shift

    def integral(self):
        return self._integral


class VoronoiIntegrator:
    """The reference's VoronoiIntegrator<WithoutFaces> (src/voronoi.rs:527-829)."""

    def __init__(self):
        raise TypeError("use VoronoiIntegrator.build")

    @classmethod
    def build(cls, generators, mask: Optional[Sequence[bool]], anchor, width,
              dimensionality=Dimensionality.ThreeD, periodic: bool = False,
              options: Optional[BuildOptions] = None) -> "VoronoiIntegrator":
        """src/voronoi.rs:549-621"""
        L = _native.lib()
        xyz = np.ascontiguousarray(np.asarray(generators, dtype=np.float64).reshape(-1, 3))
        n = xyz.shape[0]
        m = None
        if mask is not None:
            m = np.ascontiguousarray(np.asarray(mask, dtype=np.uint8))
            if m.shape[0] < n:
                raise VoronoiError(1)
        opt = (options or BuildOptions()).to_c()
        opt.flags |= _native.MV_FLAG_KEEP_CELLS
        a, w = _vec3(anchor), _vec3(width)
        h = C.c_void_p()
        _check(L.mv_build(xyz.ctypes.data if n else None, n,
                          m.ctypes.data if m is not None else None, a.ctypes.data,
                          w.ctypes.data, int(dimensionality), int(bool(periodic)),
                          C.byref(opt), C.byref(h)))
        return cls._wrap(h, xyz, m, int(dimensionality))

    @classmethod
    def _wrap(cls, h, xyz, m, d) -> "VoronoiIntegrator":
        self = object.__new__(cls)
        self._h = h
        self._n = xyz.shape[0]
        self._mask = m
        self._loc = xyz.copy()
        self._dim = d
        if d < 3:
            self._loc[:, 2] = 0.0
        if d < 2:
            self._loc[:, 1] = 0.0
        self._active = None
        self._faces_on = False
        self._cache = {}
        return self

    @classmethod
    def build_batch(cls, problems, dimensionality=Dimensionality.ThreeD, masks=None,
                    options: Optional[BuildOptions] = None) -> "List[VoronoiIntegrator]":
        """One integrator per problem of a batched build (mv_build_batch with
        MV_FLAG_KEEP_CELLS): `problems` is a sequence of (generators,
        anchor, width, periodic), `masks` one mask or None per problem.  Each
        equals VoronoiIntegrator.build of its problem alone."""
        from .voronoi import _build_batch_handles
        if len(problems) == 0:
            return []
        hs = _build_batch_handles(problems, dimensionality, masks, options,
                                  _native.MV_FLAG_KEEP_CELLS)
        out = []
        for p, h in enumerate(hs):
            xyz = np.ascontiguousarray(np.asarray(problems[p][0], dtype=np.float64).reshape(-1, 3))
            m = None
            if masks is not None and masks[p] is not None:
                m = np.ascontiguousarray(np.asarray(masks[p], dtype=np.uint8))
            out.append(cls._wrap(h, xyz, m, int(dimensionality)))
        return out

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _native.lib().mv_free(h)
            self._h = None

    # -- bookkeeping ------------------------------------------------------
    def n_active(self) -> int:
        return int(_native.lib().mv_n_active(self._h))

    def active_cells(self) -> np.ndarray:
        """Generator indices of the built cells, ascending."""
        if self._active is None:
            if self._mask is None:
                self._active = np.arange(self._n, dtype=np.uint64)
            else:
                self._active = np.nonzero(self._mask[:self._n])[0].astype(np.uint64)
        return self._active

    def tets(self):
        """(v012 [K, 3, 3], plane [K], cell_offset [n_active + 1]) -- the
        oriented tets of ConvexCell::decompose for every active cell (the
        fans of DecompositionWithFaces after with_faces)."""
        key = ("tets", self._faces_on)
        if key not in self._cache:
            self._cache[key] = self._tets()
        return self._cache[key]

    def _tets(self):
        L = _native.lib()
        cnt = C.c_size_t(0)
        _check(L.mv_tets(self._h, C.byref(cnt), None, None, None))
        K = cnt.value
        v = np.zeros((K, 3, 3))
        p = np.zeros(K, np.int32)
        off = np.zeros(self.n_active() + 1, np.uint64)
        _check(L.mv_tets(self._h, C.byref(cnt), v.ctypes.data if K else None,
                         p.ctypes.data if K else None, off.ctypes.data))
        return v, p, off

    # -- cell integrals -----------------------------------------------------
    def compute_cell_integrals_array(self, integral=VolumeCentroidIntegral) -> np.ndarray:
        """GPU path of compute_cell_integrals for the example integrals:
        n_active x (1 | 4) doubles (volume[, centroid])."""
        na = self.n_active()
        out = np.zeros((na, integral.width))
        _check(_native.lib().mv_cell_integrals(self._h, integral.kind,
                                               out.ctypes.data if na else None))
        return out

    def compute_cell_integrals(self, integral) -> List:
        """src/voronoi.rs:661-668"""
        return self.compute_cell_integrals_with_data(integral, None)

    def compute_cell_integrals_with_data(self, integral, extra_data) -> List:
        """src/voronoi.rs:670-695 (extra_data indexed by generator)"""
        if integral in _CELL_BUILTIN and extra_data is None:
            return [integral._from_row(r) for r in self.compute_cell_integrals_array(integral)]
        v, p, off = self.tets()
        out = []
        for i, c in enumerate(self.active_cells()):
            c = int(c)
            acc = (integral.init(c) if extra_data is None
                   else integral.init_with_data(c, extra_data[c]))
            gen = self._loc[c]
            for k in range(int(off[i]), int(off[i + 1])):
                acc.collect(v[k, 0], v[k, 1], v[k, 2], gen)
            out.append(acc.finalize())
        return out

    # -- face integrals -------------------------------------------------------
    def _faces_gpu(self, integral, symmetric):
        L = _native.lib()
        cnt = C.c_size_t(0)
        _check(L.mv_face_integrals(self._h, integral.kind, int(symmetric), C.byref(cnt),
                                   None, None, None, None, None, None))
        F = cnt.value
        left = np.zeros(F, np.uint64)
        right = np.zeros(F, np.int64)
        hs = np.zeros(F, np.uint8)
        shift = np.zeros((F, 3))
        plane = np.zeros(F, np.int32)
        vals = np.zeros((F, integral.width))
        if F:
            _check(L.mv_face_integrals(self._h, integral.kind, int(symmetric), C.byref(cnt),
                                       left.ctypes.data, right.ctypes.data, hs.ctypes.data,
                                       shift.ctypes.data, plane.ctypes.data, vals.ctypes.data))
        return left, right, hs, shift, plane, vals

    def compute_face_integrals_array(self, integral=AreaCentroidIntegral, symmetric=False):
        """GPU path: (left, right (-1 = None), has_shift, shift, plane, values)."""
        return self._faces_gpu(integral, symmetric)

    def compute_face_integrals(self, integral) -> List[FaceIntegrator]:
        """src/voronoi.rs:697-716"""
        return self._faces(integral, False, None)

    def compute_face_integrals_sym(self, integral) -> List[FaceIntegrator]:
        """src/voronoi.rs:718-735"""
        return self._faces(integral, True, None)

    def compute_face_integrals_with_data(self, integral, extra_data) -> List[FaceIntegrator]:
        """src/voronoi.rs:737-767 (extra_data indexed by generator)"""
        return self._faces(integral, False, extra_data)

    def compute_face_integrals_sym_with_data(self, integral, extra_data) -> List[FaceIntegrator]:
        """src/voronoi.rs:769-805"""
        return self._faces(integral, True, extra_data)

    def _faces(self, integral, symmetric, extra_data):
        builtin = integral in _FACE_BUILTIN and extra_data is None
        left, right, hs, shift, plane, vals = self._faces_gpu(
            integral if integral in _FACE_BUILTIN else AreaIntegral, symmetric)

        def wrap(f, value):
            return FaceIntegrator(int(left[f]), int(right[f]) if right[f] >= 0 else None,
                                  shift[f].copy() if hs[f] else None, value)

        if builtin:
            return [wrap(f, integral._from_row(vals[f])) for f in range(len(left))]
        # custom integrals: the GPU numbered the faces (cell, half-space);
        # fold each face's tets on the host in the decomposition order
        v, p, off = self.tets()
        pos = {int(c): i for i, c in enumerate(self.active_cells())}
        out = []
        for f in range(len(left)):
            c, pl = int(left[f]), int(plane[f])
            i = pos[c]
            acc = (integral.init(c, pl) if extra_data is None
                   else integral.init_with_data(c, pl, extra_data[c]))
            gen = self._loc[c]
            for k in range(int(off[i]), int(off[i + 1])):
                if int(p[k]) == pl:
                    acc.collect(v[k, 0], v[k, 1], v[k, 2], gen)
            out.append(wrap(f, acc.finalize()))
        return out

    # -- faces and cells (VoronoiIntegrator<WithFaces>, ConvexCell) ----------
    def with_faces(self) -> "VoronoiIntegrator":
        """VoronoiIntegrator::with_faces (src/voronoi.rs:625-643): sorts the
        vertices of every face of every kept cell counter-clockwise on the
        GPU (mv_with_faces).  The reference consumes the integrator and
        returns its WithFaces form; here the integrator is converted in
        place and returned.  Integrals and tets from then on use the fan
        decomposition.  Raises VoronoiError(11) outside 3D."""
        _check(_native.lib().mv_with_faces(self._h))
        self._faces_on = True
        return self

    def discard_faces(self) -> "VoronoiIntegrator":
        """ConvexCell::discard_faces (src/voronoi/convex_cell.rs:667-673)
        for every cell: back to the decomposition without faces."""
        _check(_native.lib().mv_discard_faces(self._h))
        self._faces_on = False
        return self

    def has_faces(self) -> bool:
        return self._faces_on

    def _kept(self):
        """Half-spaces and vertices of every active cell (mv_kept_cells)."""
        if "kept" not in self._cache:
            L = _native.lib()
            P, V = C.c_size_t(0), C.c_size_t(0)
            na = self.n_active()
            _check(L.mv_kept_cells(self._h, C.byref(P), C.byref(V), *([None] * 9)))
            k = {"poff": np.zeros(na + 1, np.uint64), "voff": np.zeros(na + 1, np.uint64),
                 "n": np.zeros((P.value, 3)), "p": np.zeros((P.value, 3)),
                 "right": np.zeros(P.value, np.int64), "has_shift": np.zeros(P.value, np.uint8),
                 "shift": np.zeros((P.value, 3)), "loc": np.zeros((V.value, 3)),
                 "dual": np.zeros((V.value, 3), np.int32)}
            if na:
                _check(L.mv_kept_cells(
                    self._h, C.byref(P), C.byref(V), k["poff"].ctypes.data,
                    k["voff"].ctypes.data, k["n"].ctypes.data, k["p"].ctypes.data,
                    k["right"].ctypes.data, k["has_shift"].ctypes.data,
                    k["shift"].ctypes.data, k["loc"].ctypes.data, k["dual"].ctypes.data))
            self._cache["kept"] = k
        return self._cache["kept"]

    def face_vertex_arrays(self):
        """(cell_face_offset [n_active + 1], face_plane, face_vertex_count,
        face_vertex_offset, face_vertices) of mv_face_vertices: the faces of
        active cell i are cell_face_offset[i] .. cell_face_offset[i+1];
        face f lists face_vertex_count[f] vertex indices (within its cell)
        at face_vertices[face_vertex_offset[f]:]."""
        if not self._faces_on:
            raise TypeError("face vertices need with_faces()")
        if "faces" not in self._cache:
            L = _native.lib()
            nf, ne = C.c_size_t(0), C.c_size_t(0)
            _check(L.mv_face_vertices(self._h, C.byref(nf), C.byref(ne), *([None] * 5)))
            cfo = np.zeros(self.n_active() + 1, np.uint64)
            fp = np.zeros(max(nf.value, 1), np.int32)
            fc = np.zeros(max(nf.value, 1), np.uint32)
            fo = np.zeros(max(nf.value, 1), np.uint64)
            fv = np.zeros(max(ne.value, 1), np.uint32)
            _check(L.mv_face_vertices(self._h, C.byref(nf), C.byref(ne), cfo.ctypes.data,
                                      fp.ctypes.data, fc.ctypes.data, fo.ctypes.data,
                                      fv.ctypes.data))
            self._cache["faces"] = (cfo, fp[:nf.value], fc[:nf.value], fo[:nf.value],
                                    fv[:ne.value])
        return self._cache["faces"]

    def get_cell_at(self, index: int) -> Optional["ConvexCell"]:
        """src/voronoi.rs:652-660: the ConvexCell of generator `index`, or
        None for a generator without one (partial construction)."""
        act = self.active_cells()
        k = int(np.searchsorted(act, index))
        if k >= len(act) or int(act[k]) != index:
            return None
        return ConvexCell(self, k)

    def cells_iter(self):
        """src/voronoi.rs:662-670: the ConvexCells in generator order."""
        for k in range(len(self.active_cells())):
            yield ConvexCell(self, k)


class Plane:
    """geometry.rs:21-27"""

    def __init__(self, n, p):
        self.n, self.p = n, p


class HalfSpace:
    """half_space.rs:7-19 (public fields: plane, right_idx, shift)"""

    def __init__(self, n, p, right, shift):
        self.plane = Plane(n, p)
        self.right_idx = right
        self.shift = shift

    def normal(self):
        return self.plane.n


class Vertex:
    """convex_cell.rs:14-26 (public fields: loc, dual)"""

    def __init__(self, loc, dual):
        self.loc, self.dual = loc, dual


class ConvexCell:
    """ConvexCell (src/voronoi/convex_cell.rs:247-268) of a kept cell: a view
    of the integrator's k-th active cell, with the face accessors of
    ConvexCell<WithFaces> (:675-722) when the integrator has faces."""

    def __init__(self, it: VoronoiIntegrator, k: int):
        self._it, self._k = it, k
        self.idx = int(it.active_cells()[k])
        self.loc = it._loc[self.idx].copy()
        kp = it._kept()
        p0, p1 = int(kp["poff"][k]), int(kp["poff"][k + 1])
        v0, v1 = int(kp["voff"][k]), int(kp["voff"][k + 1])
        self.clipping_planes = [
            HalfSpace(kp["n"][j].copy(), kp["p"][j].copy(),
                      int(kp["right"][j]) if kp["right"][j] >= 0 else None,
                      kp["shift"][j].copy() if kp["has_shift"][j] else None)
            for j in range(p0, p1)]
        self.vertices = [Vertex(kp["loc"][j].copy(), [int(d) for d in kp["dual"][j]])
                         for j in range(v0, v1)]
        self._faces = None
        if it.has_faces():
            cfo, fp, fc, fo, fv = it.face_vertex_arrays()
            self._faces = [(int(fp[f]), fv[int(fo[f]):int(fo[f]) + int(fc[f])])
                           for f in range(int(cfo[k]), int(cfo[k + 1]))]

    # ConvexCell<WithFaces> (convex_cell.rs:675-722)
    def _face(self, face_idx):
        if self._faces is None:
            raise TypeError("face accessors need VoronoiIntegrator.with_faces()")
        return self._faces[face_idx]

    def face_count(self) -> int:
        if self._faces is None:
            raise TypeError("face accessors need VoronoiIntegrator.with_faces()")
        return len(self._faces)

    def clipping_plane(self, face_idx: int) -> Plane:
        return self.clipping_planes[self._face(face_idx)[0]].plane

    def neighbour(self, face_idx: int) -> Optional[int]:
        return self.clipping_planes[self._face(face_idx)[0]].right_idx

    def shift(self, face_idx: int):
        return self.clipping_planes[self._face(face_idx)[0]].shift

    def face_vertex_count(self, face_idx: int) -> int:
        return len(self._face(face_idx)[1])

    def face_vertices(self, face_idx: int) -> np.ndarray:
        return self._face(face_idx)[1]

    # integrals of this cell (convex_cell.rs:583-664), folded on the host
    # over the tets the GPU decomposed
    def _tets(self):
        v, p, off = self._it.tets()
        a, b = int(off[self._k]), int(off[self._k + 1])
        return v[a:b], p[a:b]

    def compute_cell_integral(self, integral, extra_data=None):
        v, _ = self._tets()
        acc = (integral.init(self.idx) if extra_data is None
               else integral.init_with_data(self.idx, extra_data))
        for t in range(len(v)):
            acc.collect(v[t, 0], v[t, 1], v[t, 2], self.loc)
        return acc.finalize()

    def _face_integrals(self, integral, extra_data, mask):
        v, p = self._tets()
        dim = self._it._dim
        acc = {}
        for t in range(len(v)):
            pl = int(p[t])
            hs = self.clipping_planes[pl]
            if not Dimensionality(dim).vector_is_valid(hs.normal()):
                continue
            if pl not in acc:
                r = hs.right_idx
                if (mask is not None and hs.shift is None and r is not None and r < self.idx
                        and mask[r]):
                    continue
                acc[pl] = (integral.init(self.idx, pl) if extra_data is None
                           else integral.init_with_data(self.idx, pl, extra_data))
            acc[pl].collect(v[t, 0], v[t, 1], v[t, 2], self.loc)
        out = []
        for pl in sorted(acc):
            hs = self.clipping_planes[pl]
            out.append(FaceIntegrator(self.idx, hs.right_idx, hs.shift, acc[pl].finalize()))
        return out

    def compute_face_integrals(self, integral, extra_data=None) -> List[FaceIntegrator]:
        """convex_cell.rs:598-620"""
        return self._face_integrals(integral, extra_data, None)

    def compute_face_integrals_sym(self, integral, extra_data=None,
                                   mask=None) -> List[FaceIntegrator]:
        """convex_cell.rs:622-664 (mask: cell_is_active of the build)"""
        if mask is None:
            mask = np.ones(self._it._n, bool)
        return self._face_integrals(integral, extra_data, mask)
