// mv_faces.hpp -- face vertex lists of the kept cells and their export.
//
// ConvexCell::with_faces (src/voronoi/convex_cell.rs:461-505): the vertices
// of every half-space, collected in vertex order, then sorted by the
// reference's walk (sort_face_vertices, :508-543).  One thread per active
// cell; a cell's faces are written at the offsets of two exclusive scans
// (faces and vertex entries over the active cells), so the flattened lists
// are in (cell, half-space) order like the reference's per-cell Vecs.  The
// walk runs in place on the output list, as the reference's swaps do on its
// Vec; a face has 5-6 vertices on average and its list stays in L1/L2.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mvoro_integrals.hpp"

namespace mv {

__device__ __forceinline__ uint32_t kept_dual(double4 v) {
  return (uint32_t)__double_as_longlong(v.w);
}
__device__ __forceinline__ uint32_t fdual_at(uint32_t du, int k) {
  return (du >> (8 * k)) & 0xFFu;
}
// Vertex::plane_idx (convex_cell.rs:50-59)
__device__ __forceinline__ int fdual_pos(uint32_t du, uint32_t p) {
  return fdual_at(du, 0) == p ? 0 : fdual_at(du, 1) == p ? 1 : fdual_at(du, 2) == p ? 2 : -1;
}

// Pass 1: vertices per half-space (the lengths of face_vertex_connections,
// convex_cell.rs:470-475), faces (half-spaces holding any) and entries
// (three per vertex) per active cell.  pcnt is indexed like the kept
// half-spaces (pbase + plane).
__global__ void __launch_bounds__(256)
    k_wf_count(CellGeom G, const uint64_t *__restrict__ cells, uint64_t n,
               uint32_t *__restrict__ pcnt, uint64_t *__restrict__ nfaces,
               uint64_t *__restrict__ nentries) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t c = cells[i];
  const uint32_t np = G.np[c], nv = G.nv[c];
  uint32_t *cnt = pcnt + G.pbase[c];
  for (uint32_t p = 0; p < np; p++) cnt[p] = 0u;
  const double4 *V = G.vx + G.vbase[c];
  for (uint32_t v = 0; v < nv; v++) {
    const uint32_t du = kept_dual(V[v]);
    cnt[fdual_at(du, 0)]++;
    cnt[fdual_at(du, 1)]++;
    cnt[fdual_at(du, 2)]++;
  }
  uint32_t f = 0;
  for (uint32_t p = 0; p < np; p++) f += cnt[p] != 0u ? 1u : 0u;
  nfaces[i] = f;
  nentries[i] = 3ull * nv;
}

// Pass 2: face records in half-space order (convex_cell.rs:481-497), the
// vertex lists in vertex order, then the walk of sort_face_vertices on each
// list: from the first (lowest) vertex, the next edge's plane is the one
// after the face's plane in the current vertex's dual; the first later
// vertex holding it is swapped into the next slot.  err bit 0: no such
// vertex (the reference's assert, :538-541), or a face of fewer than three
// vertices, whose fan would index past the face (:113-118).
__global__ void __launch_bounds__(256)
    k_wf_fill(CellGeom G, const uint64_t *__restrict__ cells, uint64_t n,
              const uint64_t *__restrict__ foff, const uint64_t *__restrict__ eoff,
              uint32_t *__restrict__ pcnt, uint64_t *__restrict__ fbase,
              uint32_t *__restrict__ nf, int32_t *__restrict__ face_plane,
              uint32_t *__restrict__ face_nvert, uint64_t *__restrict__ face_voff,
              uint32_t *__restrict__ face_vtx, unsigned int *err) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t c = cells[i];
  const uint32_t np = G.np[c], nv = G.nv[c];
  uint32_t *cnt = pcnt + G.pbase[c];
  const uint64_t F = foff[i], E = eoff[i];
  uint32_t run = 0, f = 0;
  for (uint32_t p = 0; p < np; p++) {
    const uint32_t m = cnt[p];
    if (!m) continue;
    face_plane[F + f] = (int32_t)p;
    face_nvert[F + f] = m;
    face_voff[F + f] = E + run;
    cnt[p] = run;  // from here on: the face's fill cursor
    run += m;
    f++;
  }
  fbase[c] = F;
  nf[c] = f;
  const double4 *V = G.vx + G.vbase[c];
  uint32_t *ev = face_vtx + E;
  for (uint32_t v = 0; v < nv; v++) {
    const uint32_t du = kept_dual(V[v]);
    ev[cnt[fdual_at(du, 0)]++] = v;
    ev[cnt[fdual_at(du, 1)]++] = v;
    ev[cnt[fdual_at(du, 2)]++] = v;
  }
  for (uint32_t g = 0; g < f; g++) {
    const uint32_t m = face_nvert[F + g];
    const uint32_t p = (uint32_t)face_plane[F + g];
    uint32_t *vi = face_vtx + face_voff[F + g];
    if (m < 3u) {
      atomicOr(err, 1u);
      continue;
    }
    uint32_t du = kept_dual(V[vi[0]]);
    uint32_t nxt = fdual_at(du, (fdual_pos(du, p) + 1) % 3);
    for (uint32_t cur = 1; cur + 1 < m;) {
      uint32_t t = cur;
      for (; t < m; t++) {
        du = kept_dual(V[vi[t]]);
        if (fdual_at(du, 0) == nxt || fdual_at(du, 1) == nxt || fdual_at(du, 2) == nxt) break;
      }
      if (t >= m) {
        atomicOr(err, 1u);
        break;
      }
      nxt = fdual_at(du, (fdual_pos(du, p) + 1) % 3);
      const uint32_t sw = vi[cur];
      vi[cur] = vi[t];
      vi[t] = sw;
      cur++;
    }
  }
}

// Per active cell: its half-space and vertex counts (for the export's
// offsets).
__global__ void __launch_bounds__(256)
    k_kept_counts(CellGeom G, const uint64_t *__restrict__ cells, uint64_t n,
                  uint64_t *__restrict__ np_out, uint64_t *__restrict__ nv_out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  np_out[i] = G.np[cells[i]];
  nv_out[i] = G.nv[cells[i]];
}

// ConvexCell::clipping_planes / vertices of every active cell, compacted
// (one thread per active cell).
__global__ void __launch_bounds__(256)
    k_kept_export(CellGeom G, const uint64_t *__restrict__ cells, uint64_t n,
                  const uint64_t *__restrict__ poff, const uint64_t *__restrict__ voff,
                  double *__restrict__ pn, double *__restrict__ pp,
                  int64_t *__restrict__ pright, uint8_t *__restrict__ phs,
                  double *__restrict__ pshift, double *__restrict__ vloc,
                  int32_t *__restrict__ vdual) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t c = cells[i];
  const uint32_t np = G.np[c], nv = G.nv[c];
  const double4 *pa = G.pl_a + G.pbase[c];
  const double4 *pb = G.pl_b + G.pbase[c];
  for (uint32_t p = 0; p < np; p++) {
    const uint64_t o = poff[i] + p;
    const double4 a = pa[p], b = pb[p];
    pn[3 * o] = a.x;
    pn[3 * o + 1] = a.y;
    pn[3 * o + 2] = a.z;
    pp[3 * o] = b.x;
    pp[3 * o + 1] = b.y;
    pp[3 * o + 2] = b.z;
    bool hs;
    d3 sh;
    pright[o] = kept_right(G, c, (int)p, &hs, &sh);
    phs[o] = hs ? 1 : 0;
    pshift[3 * o] = sh.x;
    pshift[3 * o + 1] = sh.y;
    pshift[3 * o + 2] = sh.z;
  }
  const double4 *V = G.vx + G.vbase[c];
  for (uint32_t v = 0; v < nv; v++) {
    const uint64_t o = voff[i] + v;
    const double4 x = V[v];
    vloc[3 * o] = x.x;
    vloc[3 * o + 1] = x.y;
    vloc[3 * o + 2] = x.z;
    const uint32_t du = kept_dual(x);
    vdual[3 * o] = (int32_t)fdual_at(du, 0);
    vdual[3 * o + 1] = (int32_t)fdual_at(du, 1);
    vdual[3 * o + 2] = (int32_t)fdual_at(du, 2);
  }
}

// Voronoi from the integrator: cell and face integrals of the kept cells
// scattered into a Voronoi's per-generator and per-face arrays.
__global__ void __launch_bounds__(256)
    k_scatter_cell_integrals(const uint64_t *__restrict__ cells, uint64_t n,
                             const VolumeCentroidIntegral *__restrict__ in,
                             double *__restrict__ vol, double *__restrict__ cen) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t c = cells[i];
  vol[c] = in[i].volume;
  cen[3 * c] = in[i].centroid[0];
  cen[3 * c + 1] = in[i].centroid[1];
  cen[3 * c + 2] = in[i].centroid[2];
}

__global__ void __launch_bounds__(256)
    k_scatter_face_integrals(uint64_t F, const FaceIntegrator<AreaCentroidIntegral> *__restrict__ in,
                             double *__restrict__ area, double *__restrict__ cen) {
  const uint64_t f = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= F) return;
  area[f] = in[f].integral.area;
  cen[3 * f] = in[f].integral.centroid[0];
  cen[3 * f + 1] = in[f].integral.centroid[1];
  cen[3 * f + 2] = in[f].integral.centroid[2];
}

}  // namespace mv
