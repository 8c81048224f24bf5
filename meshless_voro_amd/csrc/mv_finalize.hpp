// mv_finalize.hpp -- face compaction and cell -> face links.
//
// flatten! + Voronoi::finalize (src/voronoi.rs:225-226, 241-263): faces are
// ordered by left cell, then by plane index within the cell (the staging
// area holds them in completion order, a scan of per-cell face counts gives
// the final offsets); each face is linked to its left cell and, when it has
// a right cell and no periodic shift, to the right cell; every cell's links
// are ascending in face index.
//
// A face without shift is built by the lower of its two cells (or by the
// active one of a partial build), so a cell's links are the faces of lower
// cells that name it (ascending in face index because ascending in left
// cell), then its own faces.  The clip kernels count the first kind per
// cell (link_in), the count's old value being the face's slot among that
// cell's links; k_place_faces puts both links in place; k_sort_links orders
// each cell's claimed slots.
#pragma once
#include "mv_common.hpp"

namespace mv {

struct FinalArgs {
  // staging
  const uint32_t *st_cell;
  const int64_t *st_right;
  const uint8_t *st_img;
  const uint32_t *st_lslot;  // slot of a linked face in its right cell's links
  const double *st_area, *st_cen, *st_norm;
  uint64_t n_staged;
  const unsigned long long *n_staged_dev;  // the count on the device (n_staged bounds it)
  const uint64_t *cell_fbase;
  const uint64_t *face_off;  // exclusive scan of cell_nf
  const BoxParams *boxes;    // per problem (periodic width for the shifts)
  const uint32_t *prob;      // problem of each generator (nullptr: one problem)
  // final faces
  uint64_t *face_left;
  int64_t *face_right;
  double *face_area, *face_centroid, *face_normal, *face_shift;
  uint8_t *face_has_shift;
  const uint32_t *link_in;   // links per cell from lower cells' faces
  const uint64_t *link_off;  // exclusive scan of cell_nf + link_in
  uint64_t *links;
};

__global__ void k_place_faces(FinalArgs A) {
  size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= A.n_staged || (A.n_staged_dev && o >= *A.n_staged_dev)) return;
  uint32_t c = A.st_cell[o];
  size_t f = A.face_off[c] + (o - A.cell_fbase[c]);
  int64_t right = A.st_right[o];
  uint8_t img = A.st_img[o];
  A.face_left[f] = c;
  A.face_right[f] = right;
  A.face_area[f] = A.st_area[o];
  for (int k = 0; k < 3; k++) {
    A.face_centroid[3 * f + k] = A.st_cen[3 * o + k];
    A.face_normal[3 * f + k] = A.st_norm[3 * o + k];
  }
  d3 s;
  const double *width = A.boxes[A.prob ? A.prob[c] : 0u].width;
  bool has = right >= 0 && image_shift(img, width, &s);
  if (!has) s = mk3(0, 0, 0);
  A.face_has_shift[f] = has ? 1 : 0;
  A.face_shift[3 * f] = s.x;
  A.face_shift[3 * f + 1] = s.y;
  A.face_shift[3 * f + 2] = s.z;
  A.links[A.link_off[c] + A.link_in[c] + (o - A.cell_fbase[c])] = f;
  if (right >= 0 && !has) A.links[A.link_off[right] + A.st_lslot[o]] = f;
}

__global__ void k_link_count(const uint32_t *__restrict__ cell_nf,
                             const uint32_t *__restrict__ link_in, size_t n,
                             uint64_t *__restrict__ link_cnt) {
  size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n) link_cnt[c] = (uint64_t)cell_nf[c] + link_in[c];
}

// ascending order of each cell's links from lower cells (claimed in any
// order)
__global__ void k_sort_links(const uint64_t *__restrict__ link_off,
                             const uint32_t *__restrict__ link_in, size_t n,
                             uint64_t *__restrict__ links,
                             const unsigned long long *__restrict__ dNL) {
  size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  uint64_t b = link_off[c], e = b + link_in[c];
  // dNL: the link total on the device (zero when the counts are not to be
  // trusted)
  if (dNL && e > *dNL) return;
  const uint32_t m = (uint32_t)(e - b);
  if (m <= 1) return;
  if (m <= 16) {
    // a short run (the common case, ~8): every load issued at once, an
    // insertion sort in registers, the run stored back
    uint64_t r[16];
#pragma unroll
    for (uint32_t i = 0; i < 16; i++) r[i] = i < m ? links[b + i] : ~0ull;
#pragma unroll
    for (int i = 1; i < 16; i++) {
#pragma unroll
      for (int j = i; j > 0; j--) {
        const uint64_t lo = r[j - 1] < r[j] ? r[j - 1] : r[j];
        const uint64_t hi = r[j - 1] < r[j] ? r[j] : r[j - 1];
        r[j - 1] = lo;
        r[j] = hi;
      }
    }
#pragma unroll
    for (uint32_t i = 0; i < 16; i++)
      if (i < m) links[b + i] = r[i];
    return;
  }
  for (uint64_t i = b + 1; i < e; i++) {
    uint64_t v = links[i];
    uint64_t j = i;
    while (j > b && links[j - 1] > v) {
      links[j] = links[j - 1];
      j--;
    }
    links[j] = v;
  }
}

// VoronoiCell records: loc and idx of active cells (Default for inactive)
__global__ void k_cell_records(const double4 *__restrict__ gpos, size_t n,
                               const uint8_t *__restrict__ mask,
                               const uint64_t *__restrict__ goff,
                               const uint32_t *__restrict__ prob,
                               double *__restrict__ cell_loc,
                               uint64_t *__restrict__ cell_idx) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  bool act = mask == nullptr || mask[i];
  double4 p = gpos[i];
  cell_loc[3 * i] = act ? p.x : 0.;
  cell_loc[3 * i + 1] = act ? p.y : 0.;
  cell_loc[3 * i + 2] = act ? p.z : 0.;
  const uint64_t base = prob ? goff[prob[i]] : 0ull;
  cell_idx[i] = act ? (uint64_t)i - base : 0;
}

// Many problems: every index above is global over the concatenated
// problems; these turn each problem's faces and links into the ones a
// build of that problem alone returns (generator indices, face indices and
// link offsets relative to the problem's first generator, face and link).
__global__ void k_problem_bases(const uint64_t *__restrict__ goff, uint32_t P,
                                const uint64_t *__restrict__ face_off, uint64_t F,
                                const uint64_t *__restrict__ link_off, uint64_t L,
                                size_t n, uint64_t *__restrict__ foff,
                                uint64_t *__restrict__ loff,
                                const unsigned long long *__restrict__ dF,
                                const unsigned long long *__restrict__ dL) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > P) return;
  // totals on the device when given
  if (dF) F = *dF;
  if (dL) L = *dL;
  const uint64_t g = goff[p];
  foff[p] = g < n ? face_off[g] : F;
  loff[p] = g < n ? link_off[g] : L;
}
__global__ void k_localize_faces(uint64_t *__restrict__ face_left,
                                 int64_t *__restrict__ face_right, uint64_t nf,
                                 const uint64_t *__restrict__ goff,
                                 const uint32_t *__restrict__ prob,
                                 const unsigned long long *__restrict__ dn) {
  const uint64_t f = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf || (dn && f >= *dn)) return;
  const uint64_t base = goff[prob[face_left[f]]];
  face_left[f] -= base;
  if (face_right[f] >= 0) face_right[f] -= (int64_t)base;
}
__global__ void k_localize_links(uint64_t *__restrict__ link_off,
                                 const uint64_t *__restrict__ link_cnt, size_t n,
                                 const uint32_t *__restrict__ prob,
                                 const uint64_t *__restrict__ foff,
                                 const uint64_t *__restrict__ loff,
                                 uint64_t *__restrict__ links,
                                 const unsigned long long *__restrict__ dNL) {
  const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const uint32_t p = prob[c];
  const uint64_t b = link_off[c], e = b + link_cnt[c];
  if (dNL && e > *dNL) return;
  for (uint64_t i = b; i < e; i++) links[i] -= foff[p];
  link_off[c] = b - loff[p];
}

__global__ void k_exact_batch(const int64_t *__restrict__ t, size_t n,
                              int8_t *__restrict__ out) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t *p = t + 15 * i;
  out[i] = (int8_t)in_sphere_exact(p, p + 3, p + 6, p + 9, p + 12);
}

}  // namespace mv
