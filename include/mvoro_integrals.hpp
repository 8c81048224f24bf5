// mvoro_integrals.hpp -- custom cell and face integrals over the cells a
// build kept on the device (MV_FLAG_KEEP_CELLS): the device-side counterpart
// of the reference's CellIntegral / FaceIntegral traits
// (src/voronoi/integrals.rs:24-33, 144-156) and of
// VoronoiIntegrator::compute_cell_integrals / compute_face_integrals(_sym)
// (src/voronoi.rs:661-767, src/voronoi/convex_cell.rs:583-664).
//
// HIP C++ (gfx950), header-only: include it in your own .hip file (with
// -I include -I meshless_voro_amd/csrc), get the device view with
// mv_get_cell_geometry, and launch mv::cell_integrals<I> /
// mv::face_integrals<I>.  The built-in integrals of libmvoro
// (mv_cell_integrals / mv_face_integrals) are instantiations of these same
// templates.
//
// An integral type I supplies, as __device__ members,
//   cell:  void init(const mv::CellGeom &g, uint64_t cell, const void *user);
//   face:  void init(const mv::CellGeom &g, uint64_t cell, int plane, const void *user);
//   both:  void collect(mv::d3 v0, mv::d3 v1, mv::d3 v2, mv::d3 gen);
//          void finalize();
// `user` is the pointer passed to the launch (per-cell data, as the
// reference's *_with_data variants).  collect receives the cell's oriented
// tets in the reference's order: vertex order, then the six tets
// (P_t, P_{t-1 mod 6}, vertex) of each vertex, where P_{2k} projects the
// generator on the vertex's k-th plane and P_{2k+1} on the line of planes
// k+1 and k (DecompositionWithoutFaces, convex_cell.rs:131-187); after
// mv_with_faces, face by face the fan (v0, vk, vk+1), k = 1 .. m-2, of the
// face's counter-clockwise vertices (DecompositionWithFaces,
// convex_cell.rs:89-129).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mv_geom.hpp"

namespace mv {

// Device view of the kept cells (layout of mv_cell_geometry in mvoro.h).
struct CellGeom {
  const double4 *pl_a;     // half-space: n.xyz, d
  const double4 *pl_b;     // half-space: p.xyz, reference word (~0 = box)
  const double4 *vx;       // vertex: loc.xyz, packed dual (bits of w)
  const uint64_t *pbase;   // first half-space of each cell (by generator)
  const uint64_t *vbase;   // first vertex of each cell
  const uint32_t *np;      // half-spaces per cell (0: not built)
  const uint32_t *nv;      // vertices per cell
  const double *cell_loc;  // generator position per cell (3 per cell)
  double width[3];         // periodic box width (image shifts)
  int dim;                 // dimensionality 1, 2, 3
  uint64_t idx_base;       // generator index of the problem's first cell
                           // in the reference words (batched builds)
  // faces (mv_with_faces; null without): per generator its first face and
  // face count; per face its half-space, vertex count and the offset of
  // its vertex indices (within the cell, counter-clockwise) in face_vtx
  const uint64_t *fbase;
  const uint32_t *nf;
  const int32_t *face_plane;
  const uint32_t *face_nvert;
  const uint64_t *face_voff;
  const uint32_t *face_vtx;
};

constexpr uint64_t KEPT_BOX_REF = ~0ull;

__device__ __forceinline__ d3 cell_generator(const CellGeom &G, uint64_t c) {
  return mk3(G.cell_loc[3 * c], G.cell_loc[3 * c + 1], G.cell_loc[3 * c + 2]);
}

// Right generator of half-space p of cell c (-1 for a box face) and its
// periodic shift (HalfSpace::right_idx / shift, half_space.rs:7-19).
__device__ __forceinline__ int64_t kept_right(const CellGeom &G, uint64_t c, int p,
                                              bool *has_shift, d3 *shift) {
  const uint64_t ref = (uint64_t)__double_as_longlong(G.pl_b[G.pbase[c] + p].w);
  *has_shift = false;
  *shift = mk3(0., 0., 0.);
  if (ref == KEPT_BOX_REF) return -1;
  const uint32_t img = (uint32_t)(ref & 31u);
  const int i = (int)(img / 9) - 1, j = (int)((img / 3) % 3) - 1, k = (int)(img % 3) - 1;
  if (i != 0 || j != 0 || k != 0) {
    *has_shift = true;
    *shift = mk3(-((double)i * G.width[0]), -((double)j * G.width[1]),
                 -((double)k * G.width[2]));
  }
  return (int64_t)((ref >> 32) - G.idx_base);
}

// Dimensionality::vector_is_valid of a half-space normal
// (src/voronoi.rs:35-50, convex_cell.rs:594-596)
__device__ __forceinline__ bool kept_plane_valid(const CellGeom &G, uint64_t c, int p) {
  const double4 a = G.pl_a[G.pbase[c] + p];
  if (G.dim == 1) return a.y == 0. && a.z == 0.;
  if (G.dim == 2) return a.z == 0.;
  return true;
}

// The cell's oriented tets in the reference's order: f(v0, v1, v2, plane).
// Returns false where a projection's determinant is zero (the reference
// panics there).
template <class F>
__device__ bool for_each_tet(const CellGeom &G, uint64_t c, F &f) {
  const double4 *V = G.vx + G.vbase[c];
  if (G.face_plane) {
    // DecompositionWithFaces::next (convex_cell.rs:104-128)
    const uint64_t fb = G.fbase[c];
    const uint32_t nfc = G.nf[c];
    for (uint32_t k = 0; k < nfc; k++) {
      const uint32_t *fv = G.face_vtx + G.face_voff[fb + k];
      const uint32_t m = G.face_nvert[fb + k];
      const int p = G.face_plane[fb + k];
      const d3 v0 = xyz(V[fv[0]]);
      for (uint32_t j = 1; j + 1 < m; j++) f(v0, xyz(V[fv[j]]), xyz(V[fv[j + 1]]), p);
    }
    return true;
  }
  const double4 *pa = G.pl_a + G.pbase[c];
  const double4 *pb = G.pl_b + G.pbase[c];
  const uint32_t nv = G.nv[c];
  const d3 q = cell_generator(G, c);
  for (uint32_t v = 0; v < nv; v++) {
    const double4 vv = V[v];
    const uint32_t du = (uint32_t)__double_as_longlong(vv.w);
    const d3 vl = xyz(vv);
    const int dl[3] = {(int)(du & 0xFFu), (int)((du >> 8) & 0xFFu), (int)((du >> 16) & 0xFFu)};
    d3 prev;
    if (!project_onto_intersection(xyz(pa[dl[0]]), pa[dl[0]].w, xyz(pa[dl[2]]), pa[dl[2]].w,
                                   q, &prev))
      return false;
    for (int k = 0; k < 3; k++) {
      const int pi = dl[k], pn = dl[k == 2 ? 0 : k + 1];
      const d3 pface = plane_project_onto(xyz(pa[pi]), xyz(pb[pi]), q);
      d3 pedge;
      if (!project_onto_intersection(xyz(pa[pn]), pa[pn].w, xyz(pa[pi]), pa[pi].w, q, &pedge))
        return false;
      f(pface, prev, vl, pi);  // tet 2k
      f(pedge, pface, vl, pi); // tet 2k + 1
      prev = pedge;
    }
  }
  return true;
}

// compute_cell_integrals: out[i] for cells[i] (one thread per cell).
template <class I>
__global__ void __launch_bounds__(256)
    cell_integrals(CellGeom G, const uint64_t *__restrict__ cells, uint64_t n,
                   const void *user, I *__restrict__ out, unsigned int *err) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t c = cells[i];
  const d3 q = cell_generator(G, c);
  I acc;
  acc.init(G, c, user);
  auto f = [&](d3 v0, d3 v1, d3 v2, int) { acc.collect(v0, v1, v2, q); };
  if (!for_each_tet(G, c, f)) atomicOr(err, 1u);
  acc.finalize();
  out[i] = acc;
}

// One face integral with the reference's bookkeeping (FaceIntegrator,
// src/voronoi/integrals.rs:196-252).
template <class I>
struct FaceIntegrator {
  uint64_t left;
  int64_t right;       // -1: boundary face
  double shift[3];     // 0 when has_shift == 0
  uint32_t has_shift;
  int32_t plane;       // half-space index within the cell
  I integral;
};

// Which half-spaces of cell c get a face integral, numbered in plane order:
// face_local[pbase + p] = rank or -1; returns the count.  symmetric: skip
// faces shared with an active cell of smaller index
// (compute_face_integrals_sym, convex_cell.rs:631-664).
__device__ __forceinline__ uint32_t kept_face_ranks(const CellGeom &G, uint64_t c,
                                                    int symmetric, const uint8_t *mask,
                                                    int32_t *face_local) {
  const uint32_t np = G.np[c], nv = G.nv[c];
  int32_t *fl = face_local + G.pbase[c];
  for (uint32_t p = 0; p < np; p++) fl[p] = 0;
  const double4 *V = G.vx + G.vbase[c];
  for (uint32_t v = 0; v < nv; v++) {
    const uint32_t du = (uint32_t)__double_as_longlong(V[v].w);
    fl[du & 0xFFu] = 1;
    fl[(du >> 8) & 0xFFu] = 1;
    fl[(du >> 16) & 0xFFu] = 1;
  }
  uint32_t r = 0;
  for (uint32_t p = 0; p < np; p++) {
    bool want = fl[p] != 0 && kept_plane_valid(G, c, (int)p);
    if (want && symmetric) {
      bool hs;
      d3 sh;
      const int64_t right = kept_right(G, c, (int)p, &hs, &sh);
      if (right >= 0 && !hs && (uint64_t)right < c && (mask == nullptr || mask[right]))
        want = false;
    }
    fl[p] = want ? (int32_t)r++ : -1;
  }
  return r;
}

__global__ void __launch_bounds__(256)
    face_counts(CellGeom G, const uint64_t *__restrict__ cells, uint64_t n, int symmetric,
                const uint8_t *__restrict__ mask, int32_t *__restrict__ face_local,
                uint64_t *__restrict__ counts) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  counts[i] = kept_face_ranks(G, cells[i], symmetric, mask, face_local);
}

// compute_face_integrals(_sym): faces of cells[i] at out[offsets[i] ...] in
// plane order (one thread per cell; face_local from face_counts).
template <class I>
__global__ void __launch_bounds__(256)
    face_integrals(CellGeom G, const uint64_t *__restrict__ cells, uint64_t n,
                   const int32_t *__restrict__ face_local,
                   const uint64_t *__restrict__ offsets, const void *user,
                   FaceIntegrator<I> *__restrict__ out, unsigned int *err) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t c = cells[i];
  const uint32_t np = G.np[c];
  const int32_t *fl = face_local + G.pbase[c];
  FaceIntegrator<I> *o = out + offsets[i];
  for (uint32_t p = 0; p < np; p++) {
    const int32_t r = fl[p];
    if (r < 0) continue;
    bool hs;
    d3 sh;
    o[r].left = c;
    o[r].right = kept_right(G, c, (int)p, &hs, &sh);
    o[r].has_shift = hs ? 1u : 0u;
    o[r].shift[0] = sh.x;
    o[r].shift[1] = sh.y;
    o[r].shift[2] = sh.z;
    o[r].plane = (int32_t)p;
    o[r].integral.init(G, c, (int)p, user);
  }
  const d3 q = cell_generator(G, c);
  auto f = [&](d3 v0, d3 v1, d3 v2, int p) {
    const int32_t r = fl[p];
    if (r >= 0) o[r].integral.collect(v0, v1, v2, q);
  };
  if (!for_each_tet(G, c, f)) atomicOr(err, 1u);
  for (uint32_t p = 0; p < np; p++) {
    const int32_t r = fl[p];
    if (r >= 0) o[r].integral.finalize();
  }
}

// ---- the reference's example integrals (src/voronoi/integrals.rs) ----

// VolumeCentroidIntegral (integrals.rs:69-106)
struct VolumeCentroidIntegral {
  double volume, centroid[3];
  __device__ void init(const CellGeom &, uint64_t, const void *) {
    volume = 0.;
    centroid[0] = centroid[1] = centroid[2] = 0.;
  }
  __device__ void collect(d3 v0, d3 v1, d3 v2, d3 gen) {
    const double v = signed_volume_tet(v0, v1, v2, gen);
    volume += v;
    const d3 s = scl3(v, add3(add3(add3(v0, v1), v2), gen));
    centroid[0] += s.x;
    centroid[1] += s.y;
    centroid[2] += s.z;
  }
  __device__ void finalize() {
    const double nrm = volume > 0. ? 0.25 / volume : 0.;
    centroid[0] *= nrm;
    centroid[1] *= nrm;
    centroid[2] *= nrm;
  }
};

// VolumeIntegral (integrals.rs:124-142)
struct VolumeIntegral {
  double volume;
  __device__ void init(const CellGeom &, uint64_t, const void *) { volume = 0.; }
  __device__ void collect(d3 v0, d3 v1, d3 v2, d3 gen) {
    volume += signed_volume_tet(v0, v1, v2, gen);
  }
  __device__ void finalize() {}
};

// AreaCentroidIntegral (integrals.rs:277-322)
struct AreaCentroidIntegral {
  double area, centroid[3];
  __device__ void init(const CellGeom &, uint64_t, int, const void *) {
    area = 0.;
    centroid[0] = centroid[1] = centroid[2] = 0.;
  }
  __device__ void collect(d3 v0, d3 v1, d3 v2, d3 gen) {
    const double a = signed_area_tri(v0, v1, v2, gen);
    area += a;
    const d3 s = scl3(a, add3(add3(v0, v1), v2));
    centroid[0] += s.x;
    centroid[1] += s.y;
    centroid[2] += s.z;
  }
  __device__ void finalize() {
    const double nrm = area > 0. ? 1. / (3. * area) : 0.;
    centroid[0] *= nrm;
    centroid[1] *= nrm;
    centroid[2] *= nrm;
  }
};

// AreaIntegral (integrals.rs:341-361)
struct AreaIntegral {
  double area;
  __device__ void init(const CellGeom &, uint64_t, int, const void *) { area = 0.; }
  __device__ void collect(d3 v0, d3 v1, d3 v2, d3 gen) {
    area += signed_area_tri(v0, v1, v2, gen);
  }
  __device__ void finalize() {}
};

}  // namespace mv
