"""Host-side mirror of the meshless_voronoi crate's construction API, over
the C ABI of libmvoro (include/mvoro.h).

Names, argument meaning and error behaviour follow the reference:
  Voronoi.build / build_partial ......... src/voronoi.rs:151-190
  Voronoi accessors ..................... src/voronoi.rs:317-359
  VoronoiCell ........................... src/voronoi/voronoi_cell.rs:107-165
  VoronoiFace ........................... src/voronoi/voronoi_face.rs:241-300
  Dimensionality ........................ src/voronoi.rs:30-50
Where the reference panics, this module raises VoronoiError with the
library's status code.  Results are also exposed as whole numpy arrays
(the SoA layout the library returns), which is the efficient way to read a
large tessellation.
"""
from __future__ import annotations

import ctypes as C
from enum import IntEnum
from typing import Iterator, Optional, Sequence

import numpy as np

from . import _native

__all__ = ["Dimensionality", "Voronoi", "VoronoiCell", "VoronoiFace", "VoronoiError",
           "BuildOptions", "BuildGraph", "MODE_PLAIN", "MODE_LEARNED"]

MODE_PLAIN = 0
MODE_LEARNED = 1


class Dimensionality(IntEnum):
    """src/voronoi.rs:30-50"""
    OneD = 1
    TwoD = 2
    ThreeD = 3

    def vector_is_valid(self, v) -> bool:
        if self == Dimensionality.OneD:
            return v[1] == 0.0 and v[2] == 0.0
        if self == Dimensionality.TwoD:
            return v[2] == 0.0
        return True


class VoronoiError(RuntimeError):
    """A build failed where the reference would panic."""

    def __init__(self, status: int):
        msg = _native.lib().mv_status_string(int(status)).decode()
        super().__init__(f"mvoro status {status}: {msg}")
        self.status = int(status)


class BuildOptions:
    """mv_options: construction mode and neighbour-list parameters."""

    def __init__(self, mode: int = MODE_PLAIN, k0: int = 27, k_max: int = 4194304,
                 lanes: int = 0, occupancy16: int = 0, exact_order: bool = False,
                 model=None, clip_blocks_per_cu: int = 0, length_order: bool = True,
                 coop_max_planes: int = 0, coop_max_vertices: int = 0,
                 cells_per_wave: int = 1, first_list_cap: int = 0, knn_min_tile: int = 0,
                 pipeline: int = 0):
        self.mode, self.k0, self.k_max, self.lanes = mode, k0, k_max, lanes
        self.occupancy16 = occupancy16
        self.exact_order = exact_order
        self.clip_blocks_per_cu = clip_blocks_per_cu
        self.length_order = length_order
        # capacities of the large-cell kernel before a cell moves to the wide
        # kernel (0: the kernel's own; lower values force the wide path)
        self.coop_max_planes, self.coop_max_vertices = coop_max_planes, coop_max_vertices
        # main lists: 1 cell per wavefront or 64 (one per lane); learned
        # first-pass list cap (0: 256); smallest gather LDS tile (0: 512)
        self.cells_per_wave = cells_per_wave
        self.first_list_cap = first_list_cap
        self.knn_min_tile = knn_min_tile
        # first pass in chunks, the next chunk's lists beside the clipping:
        # chunks | (list blocks per CU << 8); 0 = off
        self.pipeline = pipeline
        # learned mode: a meshless_voro_amd.learned.Model (None: the
        # committed checkpoint)
        self.model = model

    def to_c(self) -> _native.mv_options:
        o = _native.mv_options()
        _native.lib().mv_default_options(C.byref(o))
        o.mode, o.k0, o.k_max, o.lanes = self.mode, self.k0, self.k_max, self.lanes
        o.grid_occupancy16 = self.occupancy16
        o.flags = (_native.MV_FLAG_EXACT_ORDER if self.exact_order else 0) | (
            0 if self.length_order else _native.MV_FLAG_NO_LENGTH_ORDER)
        o.clip_blocks_per_cu = self.clip_blocks_per_cu
        o.coop_max_planes = self.coop_max_planes
        o.coop_max_vertices = self.coop_max_vertices
        o.cells_per_wave = self.cells_per_wave
        o.first_list_cap = self.first_list_cap
        o.knn_min_tile = self.knn_min_tile
        o.pipeline = self.pipeline
        if self.mode == MODE_LEARNED:
            from . import learned
            m = self.model if self.model is not None else learned.default_model()
            o.model = C.addressof(m.raw)
            self._keep = m  # the struct must outlive the call
        return o


def _vec3(v) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(v, dtype=np.float64).reshape(3))
    return a


def _as_array(ptr, n, ctype, dtype, shape=None):
    if n == 0:
        out = np.zeros(0, dtype=dtype)
    else:
        arr = np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ctype)), shape=(n,))
        out = np.array(arr, dtype=dtype, copy=True)
    if shape is not None:
        out = out.reshape(shape)
    return out


class Voronoi:
    """The main Voronoi struct (src/voronoi.rs:117-126)."""

    def __init__(self):
        raise TypeError("use Voronoi.build or Voronoi.build_partial")

    @classmethod
    def _from_handle(cls, h: C.c_void_p) -> "Voronoi":
        self = object.__new__(cls)
        L = _native.lib()
        try:
            n = L.mv_n_cells(h)
            F = L.mv_n_faces(h)
            NL = L.mv_n_links(h)
            a = (C.c_double * 3)()
            w = (C.c_double * 3)()
            L.mv_anchor(h, a)
            L.mv_width(h, w)
            self._anchor = np.array(a[:])
            self._width = np.array(w[:])
            self._dimensionality = Dimensionality(L.mv_dimensionality(h))
            self._periodic = bool(L.mv_periodic(h))
            d, u64, i64, u8, u32 = C.c_double, C.c_uint64, C.c_int64, C.c_uint8, C.c_uint32
            self.cell_loc = _as_array(L.mv_cell_loc(h), 3 * n, d, np.float64, (n, 3))
            self.cell_centroid = _as_array(L.mv_cell_centroid(h), 3 * n, d, np.float64, (n, 3))
            self.cell_volume = _as_array(L.mv_cell_volume(h), n, d, np.float64)
            self.cell_safety_radius = _as_array(L.mv_cell_safety_radius(h), n, d, np.float64)
            self.cell_face_offset = _as_array(L.mv_cell_face_offset(h), n, u64, np.uint64)
            self.cell_face_count = _as_array(L.mv_cell_face_count(h), n, u64, np.uint64)
            self.cell_idx = _as_array(L.mv_cell_idx(h), n, u64, np.uint64)
            self.cell_n_popped = _as_array(L.mv_cell_n_popped(h), n, u32, np.uint32)
            self.face_left = _as_array(L.mv_face_left(h), F, u64, np.uint64)
            self.face_right = _as_array(L.mv_face_right(h), F, i64, np.int64)
            self.face_area = _as_array(L.mv_face_area(h), F, d, np.float64)
            self.face_centroid = _as_array(L.mv_face_centroid(h), 3 * F, d, np.float64, (F, 3))
            self.face_normal = _as_array(L.mv_face_normal(h), 3 * F, d, np.float64, (F, 3))
            self.face_has_shift = _as_array(L.mv_face_has_shift(h), F, u8, np.uint8)
            self.face_shift = _as_array(L.mv_face_shift(h), 3 * F, d, np.float64, (F, 3))
            self.links = _as_array(L.mv_cell_face_connections(h), NL, u64, np.uint64)
            st = _native.mv_stats()
            L.mv_get_stats(h, C.byref(st))
            self.stats = {k: getattr(st, k) for k, _ in _native.mv_stats._fields_}
        finally:
            L.mv_free(h)
        return self

    @classmethod
    def _build(cls, generators, mask, anchor, width, dimensionality, periodic,
               options: Optional[BuildOptions]):
        xyz = np.ascontiguousarray(np.asarray(generators, dtype=np.float64).reshape(-1, 3))
        n = xyz.shape[0]
        m = None
        if mask is not None:
            m = np.ascontiguousarray(np.asarray(mask).astype(np.uint8).reshape(-1))
            if m.shape[0] < n:
                raise VoronoiError(1)  # index out of bounds in the reference
        a, w = _vec3(anchor), _vec3(width)
        L = _native.lib()
        opt = (options or BuildOptions()).to_c()
        h = C.c_void_p()
        st = L.mv_build(xyz.ctypes.data if n else None, n,
                        m.ctypes.data if m is not None else None,
                        a.ctypes.data, w.ctypes.data, int(Dimensionality(dimensionality)),
                        int(bool(periodic)), C.byref(opt), C.byref(h))
        if st != 0:
            raise VoronoiError(st)
        return cls._from_handle(h)

    @classmethod
    def build(cls, generators, anchor, width, dimensionality=Dimensionality.ThreeD,
              periodic: bool = False, options: Optional[BuildOptions] = None) -> "Voronoi":
        """Voronoi::build (src/voronoi.rs:151-159)."""
        return cls._build(generators, None, anchor, width, dimensionality, periodic, options)

    @classmethod
    def build_partial(cls, generators, mask: Sequence[bool], anchor, width,
                      dimensionality=Dimensionality.ThreeD, periodic: bool = False,
                      options: Optional[BuildOptions] = None) -> "Voronoi":
        """Voronoi::build_partial (src/voronoi.rs:181-190)."""
        return cls._build(generators, mask, anchor, width, dimensionality, periodic, options)

    @classmethod
    def from_integrator(cls, integrator) -> "Voronoi":
        """From<&VoronoiIntegrator<M>> for Voronoi (src/voronoi.rs:493-512):
        cells and faces integrated from the integrator's kept cells in their
        current decomposition (fans after with_faces), with the build's
        connectivity (mv_voronoi_from_integrator)."""
        h = C.c_void_p()
        st = _native.lib().mv_voronoi_from_integrator(integrator._h, C.byref(h))
        if st != 0:
            raise VoronoiError(st)
        return cls._from_handle(h)

    @classmethod
    def build_batch(cls, problems, dimensionality=Dimensionality.ThreeD, masks=None,
                    options: Optional[BuildOptions] = None) -> "list[Voronoi]":
        """Many independent builds in one pass (mv_build_batch): `problems` is
        a sequence of (generators, anchor, width, periodic); `masks` (optional)
        one mask or None per problem.  Returns one Voronoi per problem, equal
        to Voronoi.build / build_partial of that problem alone."""
        P = len(problems)
        if P == 0:
            return []
        hs = _build_batch_handles(problems, dimensionality, masks, options, 0)
        return [cls._from_handle(h) for h in hs]

    # --- accessors (src/voronoi.rs:317-359) ---
    def anchor(self) -> np.ndarray:
        return self._anchor.copy()

    def width(self) -> np.ndarray:
        return self._width.copy()

    def cells(self) -> "list[VoronoiCell]":
        return [VoronoiCell(self, i) for i in range(self.cell_volume.shape[0])]

    def faces(self) -> "list[VoronoiFace]":
        return [VoronoiFace(self, i) for i in range(self.face_area.shape[0])]

    def cell_face_connections(self) -> np.ndarray:
        return self.links

    def dimensionality(self) -> int:
        return int(self._dimensionality)

    def periodic(self) -> bool:
        return self._periodic


def _build_batch_handles(problems, dimensionality, masks, options, flags):
    """mv_build_batch over (generators, anchor, width, periodic) problems:
    one handle per problem (options' flags OR `flags`)."""
    P = len(problems)
    xs = [np.ascontiguousarray(np.asarray(g, dtype=np.float64).reshape(-1, 3))
          for g, _, _, _ in problems]
    offsets = np.zeros(P + 1, np.uint64)
    offsets[1:] = np.cumsum([len(x) for x in xs])
    xyz = np.ascontiguousarray(np.concatenate(xs, 0)) if offsets[-1] else np.zeros((0, 3))
    anchors = np.ascontiguousarray(np.stack([_vec3(a) for _, a, _, _ in problems]))
    widths = np.ascontiguousarray(np.stack([_vec3(w) for _, _, w, _ in problems]))
    per = np.ascontiguousarray(np.array([int(bool(p)) for _, _, _, p in problems], np.uint32))
    m = None
    if masks is not None and any(mk is not None for mk in masks):
        parts = []
        for x, mk in zip(xs, masks):
            if mk is None:
                parts.append(np.ones(len(x), np.uint8))
            else:
                mk = np.asarray(mk).astype(np.uint8).reshape(-1)
                if mk.shape[0] < len(x):
                    raise VoronoiError(1)
                parts.append(mk[:len(x)])
        m = np.ascontiguousarray(np.concatenate(parts)) if parts else None
    L = _native.lib()
    opt = (options or BuildOptions()).to_c()
    opt.flags |= flags
    hs = (C.c_void_p * P)()
    st = L.mv_build_batch(xyz.ctypes.data if offsets[-1] else None, offsets.ctypes.data, P,
                          m.ctypes.data if m is not None else None, anchors.ctypes.data,
                          widths.ctypes.data, per.ctypes.data,
                          int(Dimensionality(dimensionality)), C.byref(opt), hs)
    if st != 0:
        raise VoronoiError(st)
    return [C.c_void_p(hs[p]) for p in range(P)]


class BuildGraph:
    """A captured build (mv_graph_create / mv_graph_launch; no counterpart in
    the reference): the build of P problems over a device array of seeds,
    recorded once as a hipGraph and replayed on the array's current
    contents.  `xyz` is a contiguous float64 tensor of shape (n, 3) on the
    GPU (the caller may rewrite it in place between launches and must keep
    it alive), `problems` a sequence of (offset, anchor, width, periodic)
    per problem with offsets ascending from 0, `mask` an optional uint8 GPU
    tensor of n entries.  launch() returns (list of Voronoi, replayed):
    replayed is False when the replay needed more than was recorded and the
    ordinary build ran instead; the results are the same either way."""

    def __init__(self, xyz, problems, dimensionality=Dimensionality.ThreeD, mask=None,
                 options: Optional[BuildOptions] = None):
        self._h = None
        self._xyz, self._mask = xyz, mask
        n = int(xyz.shape[0])
        P = len(problems)
        self.P = P
        offsets = np.zeros(P + 1, np.uint64)
        offsets[:P] = [int(o) for o, _, _, _ in problems]
        offsets[P] = n
        anchors = np.ascontiguousarray(np.stack([_vec3(a) for _, a, _, _ in problems]))
        widths = np.ascontiguousarray(np.stack([_vec3(w) for _, _, w, _ in problems]))
        per = np.ascontiguousarray(np.array([int(bool(p)) for _, _, _, p in problems], np.uint32))
        opt = (options or BuildOptions()).to_c()
        h = C.c_void_p()
        st = _native.lib().mv_graph_create(
            C.c_void_p(xyz.data_ptr()), offsets.ctypes.data, P,
            C.c_void_p(mask.data_ptr()) if mask is not None else None, anchors.ctypes.data,
            widths.ctypes.data, per.ctypes.data, int(Dimensionality(dimensionality)),
            C.byref(opt), C.byref(h))
        if st != 0:
            raise VoronoiError(st)
        self._h = h

    def launch(self):
        hs = (C.c_void_p * self.P)()
        rep = C.c_uint32(0)
        st = _native.lib().mv_graph_launch(self._h, hs, C.byref(rep))
        if st != 0:
            raise VoronoiError(st)
        return [Voronoi._from_handle(C.c_void_p(hs[p])) for p in range(self.P)], bool(rep.value)

    def last_violation(self) -> int:
        """Why the last launch ran the ordinary build (0: it replayed): bits
        of mv_graph_last_violation."""
        return int(_native.lib().mv_graph_last_violation(self._h))

    def close(self):
        if self._h is not None:
            _native.lib().mv_graph_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class VoronoiCell:
    """View of one cell (src/voronoi/voronoi_cell.rs:107-165)."""

    __slots__ = ("_v", "_i")

    def __init__(self, v: Voronoi, i: int):
        self._v, self._i = v, i

    def loc(self):
        return self._v.cell_loc[self._i]

    def centroid(self):
        return self._v.cell_centroid[self._i]

    def volume(self) -> float:
        return float(self._v.cell_volume[self._i])

    def safety_radius(self) -> float:
        return float(self._v.cell_safety_radius[self._i])

    def face_connections_offset(self) -> int:
        return int(self._v.cell_face_offset[self._i])

    def face_count(self) -> int:
        return int(self._v.cell_face_count[self._i])

    def face_indices(self, voronoi: Voronoi) -> np.ndarray:
        o = self.face_connections_offset()
        return voronoi.links[o:o + self.face_count()]

    def faces(self, voronoi: Voronoi) -> Iterator["VoronoiFace"]:
        return (VoronoiFace(voronoi, int(i)) for i in self.face_indices(voronoi))

    def neighbour_ids(self, voronoi: Voronoi) -> Iterator[int]:
        """Neighbours across internal non-periodic faces
        (voronoi_cell.rs:141-153)."""
        me = int(voronoi.cell_idx[self._i])
        for i in self.face_indices(voronoi):
            if voronoi.face_has_shift[i] or voronoi.face_right[i] < 0:
                continue
            left = int(voronoi.face_left[i])
            yield int(voronoi.face_right[i]) if left == me else left


class VoronoiFace:
    """View of one face (src/voronoi/voronoi_face.rs:241-300)."""

    __slots__ = ("_v", "_i")

    def __init__(self, v: Voronoi, i: int):
        self._v, self._i = v, i

    def left(self) -> int:
        return int(self._v.face_left[self._i])

    def right(self) -> Optional[int]:
        r = int(self._v.face_right[self._i])
        return None if r < 0 else r

    def area(self) -> float:
        return float(self._v.face_area[self._i])

    def centroid(self):
        return self._v.face_centroid[self._i]

    def normal(self):
        """Normal of the left cell's half-space: it points from the face
        toward the left generator (voronoi_face.rs:26, convex_cell.rs:340)."""
        return self._v.face_normal[self._i]

    def shift(self):
        return self._v.face_shift[self._i] if self._v.face_has_shift[self._i] else None

    def is_periodic(self) -> bool:
        return bool(self._v.face_has_shift[self._i])

    def is_boundary(self) -> bool:
        return self._v.face_right[self._i] < 0
