// mv_grid.hpp -- generator preparation, counting-sort uniform grid, scans
// and the fixed-length neighbour-list kernel.
//
// The R-tree of the reference (src/rtree_nn.rs:6-36) becomes a uniform
// grid: generators are counting-sorted by cell, and the neighbour stream is
// produced in the same ascending key order as the reference's
// nearest-neighbour iterators, with (idx, image) deciding among equal keys.
#pragma once
#include "mv_common.hpp"

namespace mv {

// Generator::new (src/voronoi/generator.rs:13-24): copy each seed, zero the
// unused components; per-block bounding-box partials.
__global__ void k_prepare(const double *__restrict__ xyz, size_t n, int dim,
                          double4 *__restrict__ gpos,
                          double *__restrict__ bbox_part) {
  __shared__ double smin[3][256], smax[3][256];
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  double mn[3] = {INFINITY, INFINITY, INFINITY};
  double mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  if (i < n) {
    double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
    if (dim == 1) {
      y = 0.;
      z = 0.;
    } else if (dim == 2) {
      z = 0.;
    }
    gpos[i] = make_double4(x, y, z, 0.);
    mn[0] = mx[0] = x;
    mn[1] = mx[1] = y;
    mn[2] = mx[2] = z;
  }
  for (int c = 0; c < 3; c++) {
    smin[c][threadIdx.x] = mn[c];
    smax[c][threadIdx.x] = mx[c];
  }
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < (unsigned)s)
      for (int c = 0; c < 3; c++) {
        smin[c][threadIdx.x] = fmin(smin[c][threadIdx.x], smin[c][threadIdx.x + s]);
        smax[c][threadIdx.x] = fmax(smax[c][threadIdx.x], smax[c][threadIdx.x + s]);
      }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int c = 0; c < 3; c++) {
      bbox_part[6 * blockIdx.x + c] = smin[c][0];
      bbox_part[6 * blockIdx.x + 3 + c] = smax[c][0];
    }
}

__global__ void k_bbox_final(const double *__restrict__ part, int nb,
                             double *__restrict__ out) {
  __shared__ double s[6][256];
  int t = threadIdx.x;
  double v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  for (int b = t; b < nb; b += blockDim.x)
    for (int c = 0; c < 3; c++) {
      v[c] = fmin(v[c], part[6 * b + c]);
      v[3 + c] = fmax(v[3 + c], part[6 * b + 3 + c]);
    }
  for (int c = 0; c < 6; c++) s[c][t] = v[c];
  __syncthreads();
  for (int st = blockDim.x / 2; st > 0; st >>= 1) {
    if (t < st)
      for (int c = 0; c < 3; c++) {
        s[c][t] = fmin(s[c][t], s[c][t + st]);
        s[3 + c][t] = fmax(s[3 + c][t], s[3 + c][t + st]);
      }
    __syncthreads();
  }
  if (t == 0)
    for (int c = 0; c < 6; c++) out[c] = s[c][0];
}

__device__ __forceinline__ int grid_coord(const GridParams &g, int c,
                                          double x) {
  double f = (x - g.lo[c]) * g.inv_h[c];
  if (!(f >= 0.)) return 0;
  if (f >= (double)g.G[c]) return g.G[c] - 1;
  return (int)f;
}

__global__ void k_grid_count(const double4 *__restrict__ gpos, size_t n,
                             GridParams g, uint32_t *__restrict__ cellid,
                             uint32_t *__restrict__ counts) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double4 p = gpos[i];
  int a = grid_coord(g, 0, p.x), b = grid_coord(g, 1, p.y),
      c = grid_coord(g, 2, p.z);
  uint32_t id = (uint32_t)(((size_t)a * g.G[1] + b) * g.G[2] + c);
  cellid[i] = id;
  atomicAdd(&counts[id], 1u);
}

__global__ void k_grid_scatter(const uint32_t *__restrict__ cellid, size_t n,
                               const uint32_t *__restrict__ cell_start,
                               uint32_t *__restrict__ fill,
                               uint32_t *__restrict__ order) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t c = cellid[i];
  uint32_t pos = cell_start[c] + atomicAdd(&fill[c], 1u);
  order[pos] = (uint32_t)i;
}

// make each grid cell's list ascending in generator index (deterministic
// slot order)
__global__ void k_grid_sort_cells(const uint32_t *__restrict__ cell_start,
                                  size_t ncell, uint32_t *__restrict__ order) {
  size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  uint32_t b = cell_start[c], e = cell_start[c + 1];
  for (uint32_t i = b + 1; i < e; i++) {
    uint32_t v = order[i];
    uint32_t j = i;
    while (j > b && order[j - 1] > v) {
      order[j] = order[j - 1];
      j--;
    }
    order[j] = v;
  }
}

// sorted positions; w carries the generator index bits
__global__ void k_grid_gather(const double4 *__restrict__ gpos,
                              const uint32_t *__restrict__ order, size_t n,
                              double4 *__restrict__ spos,
                              uint32_t *__restrict__ slot_of) {
  size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  uint32_t i = order[s];
  double4 p = gpos[i];
  p.w = __longlong_as_double((long long)i);
  spos[s] = p;
  slot_of[i] = (uint32_t)s;
}

// ---------------------------------------------------------------------
// exclusive scan (three kernels: block scan, scan of block sums, add)
// ---------------------------------------------------------------------
constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

template <typename TIn, typename TOut>
__global__ void k_scan_tile(const TIn *__restrict__ in, size_t n,
                            TOut *__restrict__ out,
                            TOut *__restrict__ tile_sums) {
  __shared__ TOut s[SCAN_THREADS];
  size_t base = (size_t)blockIdx.x * SCAN_TILE + (size_t)threadIdx.x * SCAN_ITEMS;
  TOut v[SCAN_ITEMS];
  TOut acc = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; k++) {
    size_t i = base + k;
    TOut x = i < n ? (TOut)in[i] : (TOut)0;
    v[k] = acc;
    acc += x;
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int d = 1; d < SCAN_THREADS; d <<= 1) {
    TOut y = threadIdx.x >= (unsigned)d ? s[threadIdx.x - d] : (TOut)0;
    __syncthreads();
    s[threadIdx.x] += y;
    __syncthreads();
  }
  TOut excl = s[threadIdx.x] - acc;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; k++) {
    size_t i = base + k;
    if (i < n) out[i] = v[k] + excl;
  }
  if (threadIdx.x == SCAN_THREADS - 1) tile_sums[blockIdx.x] = s[threadIdx.x];
}

template <typename T>
__global__ void k_scan_add(T *__restrict__ out, size_t n,
                           const T *__restrict__ tile_offsets) {
  size_t i = (size_t)blockIdx.x * SCAN_TILE + threadIdx.x;
  T off = tile_offsets[blockIdx.x];
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; k++, i += SCAN_THREADS)
    if (i < n) out[i] += off;
}

// ---------------------------------------------------------------------
// k nearest stream entries (fixed length L) per work item.
//
// Order: ascending (key, idx, image), key = wrapping_distance_2 of
// src/rtree_nn.rs:138-146 (for non-periodic builds the reference's
// distance_squared, src/rtree_nn.rs:124-134, gives the same bits).
// Grid search: Chebyshev shells of grid cells around the cell of each
// shifted query q + s, one shell at a time for every image, until the
// L-th best key is below the squared distance to every unvisited cell.
// The list lives in global memory, interleaved by work item:
// entry j of item w is at [j * nwork + w].
// ---------------------------------------------------------------------
__device__ __forceinline__ bool entry_less(double ka, uint64_t pa, double kb,
                                           uint64_t pb) {
  return ka < kb || (ka == kb && pa < pb);
}

struct ImageView {
  double qs[3];
  int C[3];
};

__device__ __forceinline__ void image_center(const GridParams &g,
                                             const double *q, const double *w,
                                             int i, int j, int k,
                                             ImageView *iv) {
  int s[3] = {i, j, k};
  for (int c = 0; c < 3; c++) {
    iv->qs[c] = q[c] + (double)s[c] * w[c];
    double f = floor((iv->qs[c] - g.lo[c]) * g.inv_h[c]);
    f = fmax(-1.0e8, fmin(1.0e8, f));
    iv->C[c] = g.inv_h[c] == 0. ? 0 : (int)f;
  }
}

// distance from the shifted query to the unvisited part of the grid after
// the cells within Chebyshev radius r of its cell were visited; +inf when
// nothing is left.
__device__ __forceinline__ double image_bound(const GridParams &g,
                                              const ImageView &iv, int r) {
  double lb = INFINITY;
  for (int c = 0; c < 3; c++) {
    if (iv.C[c] - r > 0) {
      double face = g.lo[c] + (double)(iv.C[c] - r) * g.h[c];
      lb = fmin(lb, iv.qs[c] - face);
    }
    if (iv.C[c] + r < g.G[c] - 1) {
      double face = g.lo[c] + (double)(iv.C[c] + r + 1) * g.h[c];
      lb = fmin(lb, face - iv.qs[c]);
    }
  }
  if (lb == INFINITY) return lb;
  return fmax(0.0, lb - g.margin);
}

__global__ void __launch_bounds__(256)
    k_knn(const double4 *__restrict__ spos,
          const uint32_t *__restrict__ cell_start, GridParams g,
          BoxParams box, const uint32_t *__restrict__ work, uint32_t nwork,
          uint32_t L, double *__restrict__ kkey, uint64_t *__restrict__ kpk,
          uint32_t *__restrict__ kcnt) {
  uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nwork) return;
  const uint32_t slot = work[w];
  const double4 qq = spos[slot];
  const double q[3] = {qq.x, qq.y, qq.z};
  uint32_t cnt = 0;
  double worst_k = INFINITY;
  uint64_t worst_p = ~0ull;
  const size_t NW = nwork;
  for (int r = 0;; r++) {
    bool any_left = false;
    double lb_next = INFINITY;
    for (int i = g.img_lo[0]; i <= g.img_hi[0]; i++)
      for (int j = g.img_lo[1]; j <= g.img_hi[1]; j++)
        for (int k = g.img_lo[2]; k <= g.img_hi[2]; k++) {
          ImageView iv;
          image_center(g, q, box.width, i, j, k, &iv);
          double lb_prev = r == 0 ? 0.0 : image_bound(g, iv, r - 1);
          if (lb_prev == INFINITY) continue;  // image fully visited
          double lb_r = image_bound(g, iv, r);
          if (lb_r < INFINITY) any_left = true;
          lb_next = fmin(lb_next, lb_r);
          if (cnt == L && lb_prev * lb_prev > worst_k * (1.0 + 1e-12))
            continue;  // nothing in shell r can enter the list
          const uint32_t img = (uint32_t)((i + 1) * 9 + (j + 1) * 3 + (k + 1));
          int a0 = max(0, iv.C[0] - r), a1 = min(g.G[0] - 1, iv.C[0] + r);
          int b0 = max(0, iv.C[1] - r), b1 = min(g.G[1] - 1, iv.C[1] + r);
          int c0 = max(0, iv.C[2] - r), c1 = min(g.G[2] - 1, iv.C[2] + r);
          for (int a = a0; a <= a1; a++) {
            int da = abs(a - iv.C[0]);
            for (int b = b0; b <= b1; b++) {
              int db = abs(b - iv.C[1]);
              bool full = max(da, db) == r;
              int cl[2];
              int ncl = 0;
              if (!full) {
                int ca = iv.C[2] - r, cb = iv.C[2] + r;
                if (ca >= c0 && ca <= c1) cl[ncl++] = ca;
                if (cb != ca && cb >= c0 && cb <= c1) cl[ncl++] = cb;
              }
              int cstart = full ? c0 : 0, cend = full ? c1 : ncl - 1;
              for (int ci = cstart; ci <= cend; ci++) {
                int c = full ? ci : cl[ci];
                size_t cell = ((size_t)a * g.G[1] + b) * g.G[2] + c;
                uint32_t t0 = cell_start[cell], t1 = cell_start[cell + 1];
                for (uint32_t t = t0; t < t1; t++) {
                  double4 gp = spos[t];
                  double d0 = iv.qs[0] - gp.x, d1 = iv.qs[1] - gp.y,
                         d2 = iv.qs[2] - gp.z;
                  double key = (d0 * d0 + d1 * d1) + d2 * d2;
                  uint64_t gidx = (uint64_t)__double_as_longlong(gp.w);
                  uint64_t pk = (gidx << 32) | ((uint64_t)t << 5) | img;
                  if (cnt == L && !entry_less(key, pk, worst_k, worst_p))
                    continue;
                  uint32_t pos = cnt < L ? cnt : L - 1;
                  while (pos > 0) {
                    double pkey = kkey[(size_t)(pos - 1) * NW + w];
                    uint64_t ppk = kpk[(size_t)(pos - 1) * NW + w];
                    if (!entry_less(key, pk, pkey, ppk)) break;
                    kkey[(size_t)pos * NW + w] = pkey;
                    kpk[(size_t)pos * NW + w] = ppk;
                    pos--;
                  }
                  kkey[(size_t)pos * NW + w] = key;
                  kpk[(size_t)pos * NW + w] = pk;
                  if (cnt < L) cnt++;
                  if (cnt == L) {
                    worst_k = kkey[(size_t)(L - 1) * NW + w];
                    worst_p = kpk[(size_t)(L - 1) * NW + w];
                  }
                }
              }
            }
          }
        }
    if (!any_left) break;
    if (cnt == L && worst_k < lb_next * lb_next * (1.0 - 1e-12)) break;
  }
  kcnt[w] = cnt;
}

// work-list compaction by flag (ascending order kept)
__global__ void k_flag_state(const uint8_t *__restrict__ st, uint32_t n,
                             uint8_t want, uint32_t *__restrict__ flag) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flag[i] = st[i] == want ? 1u : 0u;
}
__global__ void k_compact(const uint32_t *__restrict__ in,
                          const uint32_t *__restrict__ flag,
                          const uint32_t *__restrict__ pos, uint32_t n,
                          uint32_t *__restrict__ out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && flag[i]) out[pos[i]] = in[i];
}
// initial work list: slots of the active generators
__global__ void k_flag_active(const double4 *__restrict__ spos, size_t n,
                              const uint8_t *__restrict__ mask,
                              uint32_t *__restrict__ flag,
                              uint32_t *__restrict__ ident) {
  size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  uint32_t i = (uint32_t)__double_as_longlong(spos[s].w);
  flag[s] = (mask == nullptr || mask[i]) ? 1u : 0u;
  ident[s] = (uint32_t)s;
}

}  // namespace mv
