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
// unused components.
__global__ void k_prepare(const double *__restrict__ xyz, size_t n, int dim,
                          double4 *__restrict__ gpos) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
  if (dim == 1) {
    y = 0.;
    z = 0.;
  } else if (dim == 2) {
    z = 0.;
  }
  gpos[i] = make_double4(x, y, z, 0.);
}

// problem of every generator (one block row per problem)
__global__ void k_problem_of(const uint64_t *__restrict__ goff, uint32_t P,
                             uint32_t *__restrict__ prob) {
  const uint32_t p = blockIdx.y;
  if (p >= P) return;
  for (uint64_t i = goff[p] + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < goff[p + 1];
       i += (uint64_t)gridDim.x * blockDim.x)
    prob[i] = p;
}

// bounding box of each problem: block (s, p) reduces chunk s of problem p,
// then one block per problem reduces the chunks
constexpr uint32_t BBOX_CHUNK = 65536;
__global__ void __launch_bounds__(256)
    k_bbox_part(const double4 *__restrict__ gpos, const uint64_t *__restrict__ goff,
                uint32_t S, double *__restrict__ part) {
  __shared__ double sm[6][256];
  const uint32_t p = blockIdx.y, sidx = blockIdx.x;
  const uint64_t b = goff[p] + (uint64_t)sidx * BBOX_CHUNK;
  const uint64_t e = min(goff[p + 1], b + BBOX_CHUNK);
  double v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  for (uint64_t i = b + threadIdx.x; i < e; i += blockDim.x) {
    const double4 g = gpos[i];
    v[0] = fmin(v[0], g.x);
    v[1] = fmin(v[1], g.y);
    v[2] = fmin(v[2], g.z);
    v[3] = fmax(v[3], g.x);
    v[4] = fmax(v[4], g.y);
    v[5] = fmax(v[5], g.z);
  }
  for (int c = 0; c < 6; c++) sm[c][threadIdx.x] = v[c];
  __syncthreads();
  for (int st = blockDim.x / 2; st > 0; st >>= 1) {
    if (threadIdx.x < (unsigned)st)
      for (int c = 0; c < 3; c++) {
        sm[c][threadIdx.x] = fmin(sm[c][threadIdx.x], sm[c][threadIdx.x + st]);
        sm[3 + c][threadIdx.x] = fmax(sm[3 + c][threadIdx.x], sm[3 + c][threadIdx.x + st]);
      }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int c = 0; c < 6; c++) part[((size_t)p * S + sidx) * 6 + c] = sm[c][0];
}
__global__ void k_bbox_final(const double *__restrict__ part, uint32_t S, uint32_t P,
                             double *__restrict__ out) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  double v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  for (uint32_t s2 = 0; s2 < S; s2++)
    for (int c = 0; c < 3; c++) {
      v[c] = fmin(v[c], part[((size_t)p * S + s2) * 6 + c]);
      v[3 + c] = fmax(v[3 + c], part[((size_t)p * S + s2) * 6 + 3 + c]);
    }
  for (int c = 0; c < 6; c++) out[6 * p + c] = v[c];
}

__device__ __forceinline__ int grid_coord(const GridParams &g, int c,
                                          double x) {
  double f = (x - g.lo[c]) * g.inv_h[c];
  if (!(f >= 0.)) return 0;
  if (f >= (double)g.G[c]) return g.G[c] - 1;
  return (int)f;
}

__global__ void k_grid_count(const double4 *__restrict__ gpos, size_t n,
                             const GridParams *__restrict__ grids,
                             const uint32_t *__restrict__ coff,
                             const uint32_t *__restrict__ prob,
                             uint32_t *__restrict__ cellid,
                             uint32_t *__restrict__ counts) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t pr = prob ? prob[i] : 0u;
  const GridParams &g = grids[pr];
  double4 p = gpos[i];
  int a = grid_coord(g, 0, p.x), b = grid_coord(g, 1, p.y),
      c = grid_coord(g, 2, p.z);
  uint32_t id = coff[pr] + (uint32_t)(((size_t)a * g.G[1] + b) * g.G[2] + c);
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

// The order of a grid cell's seeds: by zkey (the top 32 bits of an
// order-preserving map of z), then by generator index.  Cells along z are
// consecutive and seeds fall in cells monotonically in z, so each row of
// cells along z is then ordered by zkey, and the gather can cut a long row
// to the ball's chord with two binary searches (the key is truncated, so
// the cut keeps every seed whose z lies in the chord).  The order is
// deterministic; nothing else depends on it (lists are sorted by key and
// generator index after the gather).
__device__ __forceinline__ uint32_t zkey(double z) {
  const uint64_t b = (uint64_t)__double_as_longlong(z);
  return (uint32_t)(((b >> 63) ? ~b : (b | 0x8000000000000000ull)) >> 32);
}
__device__ __forceinline__ uint64_t zkey_idx(const double4 *gpos, uint32_t i) {
  return ((uint64_t)zkey(gpos[i].z) << 32) | i;
}

// A thread sorts a cell of up to GRID_SMALL seeds by insertion; larger cells
// (dense clusters put tens of thousands of seeds in one cell) are listed
// for k_grid_sort_big.
constexpr uint32_t GRID_SMALL = 32;
__global__ void k_grid_sort_cells(const uint32_t *__restrict__ cell_start, size_t ncell,
                                  const double4 *__restrict__ gpos,
                                  uint32_t *__restrict__ order,
                                  uint32_t *__restrict__ big_list,
                                  unsigned int *__restrict__ big_cnt) {
  size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  uint32_t b = cell_start[c], e = cell_start[c + 1];
  if (e - b > GRID_SMALL) {
    big_list[atomicAdd(big_cnt, 1u)] = (uint32_t)c;
    return;
  }
  for (uint32_t i = b + 1; i < e; i++) {
    const uint32_t v = order[i];
    const uint64_t kv = zkey_idx(gpos, v);
    uint32_t j = i;
    while (j > b && zkey_idx(gpos, order[j - 1]) > kv) {
      order[j] = order[j - 1];
      j--;
    }
    order[j] = v;
  }
}

// One block per large cell (of the list k_grid_sort_cells made): the
// cell's (zkey, index) words in chunks of GS_CHUNK sorted by a bitonic
// network in LDS, then runs of width w merged pairwise between two scratch
// ranges.  The words are distinct, so a word's place in the merged run is
// its rank in its own run plus the number of smaller words in the partner
// run (a binary search), with no merge path to share.  The sorted indices
// go back to `order`.
constexpr uint32_t GS_CHUNK = 4096;
constexpr int GS_THREADS = 1024;
__global__ void __launch_bounds__(GS_THREADS)
    k_grid_sort_big(const uint32_t *__restrict__ cell_start,
                    const uint32_t *__restrict__ big_list,
                    const unsigned int *__restrict__ big_cnt,
                    const double4 *__restrict__ gpos, uint32_t *__restrict__ order,
                    uint64_t *__restrict__ ta, uint64_t *__restrict__ tb) {
  __shared__ uint64_t s[GS_CHUNK];
  const uint32_t tid = threadIdx.x;
  const uint32_t nbig = *big_cnt;
  for (uint32_t k = blockIdx.x; k < nbig; k += gridDim.x) {
    const uint32_t c = big_list[k];
    const uint32_t b = cell_start[c], m = cell_start[c + 1] - b;
    for (uint32_t c0 = 0; c0 < m; c0 += GS_CHUNK) {
      const uint32_t len = min(GS_CHUNK, m - c0);
      uint32_t n2 = 1;
      while (n2 < len) n2 <<= 1;
      for (uint32_t i = tid; i < n2; i += GS_THREADS)
        s[i] = i < len ? zkey_idx(gpos, order[b + c0 + i]) : ~0ull;
      __syncthreads();
      for (uint32_t kk = 2; kk <= n2; kk <<= 1) {
        for (uint32_t j = kk >> 1; j > 0; j >>= 1) {
          for (uint32_t p = tid; p < (n2 >> 1); p += GS_THREADS) {
            const uint32_t i = ((p & ~(j - 1)) << 1) | (p & (j - 1));
            const uint64_t x = s[i], y = s[i | j];
            if ((x > y) == ((i & kk) == 0)) {
              s[i] = y;
              s[i | j] = x;
            }
          }
          __syncthreads();
        }
      }
      for (uint32_t i = tid; i < len; i += GS_THREADS) ta[b + c0 + i] = s[i];
      __syncthreads();
    }
    uint64_t *src = ta + b, *dst = tb + b;
    for (uint32_t w = GS_CHUNK; w < m; w <<= 1) {
      for (uint32_t i = tid; i < m; i += GS_THREADS) {
        const uint32_t r0 = i / (2 * w) * (2 * w);
        const uint32_t a1 = min(r0 + w, m), b1 = min(r0 + 2 * w, m);
        const bool in_a = i < a1;
        const uint64_t x = src[i];
        uint32_t lo = in_a ? a1 : r0, hi = in_a ? b1 : a1;
        const uint32_t base = lo;
        while (lo < hi) {
          const uint32_t mid = (lo + hi) >> 1;
          if (src[mid] < x)
            lo = mid + 1;
          else
            hi = mid;
        }
        dst[r0 + (i - (in_a ? r0 : a1)) + (lo - base)] = x;
      }
      __syncthreads();
      uint64_t *t = src;
      src = dst;
      dst = t;
    }
    for (uint32_t i = tid; i < m; i += GS_THREADS) order[b + i] = (uint32_t)src[i];
    __syncthreads();
  }
}

// The sub-columns of the dense cells of 3D grids (SubCols), one wavefront
// per cell of the large-cell list.  The S - 1 inner boundaries in x are
// quantiles of the cell's x (k m / S seeds below the k-th, from a histogram
// of SUB_BINS bins over the cell), likewise in y; the outer ones are -inf and
// +inf.  A seed's sub-column is (slab of x, slab of y), the slab of v being
// the number of inner boundaries <= v, so every seed lies in its
// sub-column's box exactly.  For a seed cluster whose density is a product
// of its x and y marginals (a Gaussian) the sub-columns hold m / S^2 seeds
// each wherever the cluster is.  The slots are then placed 64 at a time in
// the cell's order (a lane's place is its sub-column's cursor plus the
// lanes before it with the same sub-column), so that every sub-column keeps
// the cell's zkey order.  Tables come from two pools; a cell whose table
// does not fit keeps base = ~0u (its gathers read the whole cell).
__global__ void __launch_bounds__(64)
    k_grid_subcols(const uint32_t *__restrict__ cell_start, const uint32_t *__restrict__ big_list,
                   const unsigned int *__restrict__ big_cnt, const double4 *__restrict__ spos,
                   const GridParams *__restrict__ grids, const uint32_t *__restrict__ coff,
                   const uint32_t *__restrict__ prob_of_slot, uint32_t *__restrict__ sub_base,
                   uint32_t *__restrict__ pool, uint32_t pool_cap,
                   unsigned int *__restrict__ pool_cnt, double *__restrict__ xy,
                   uint32_t xy_cap, unsigned int *__restrict__ xy_cnt,
                   uint32_t *__restrict__ sub_slot, uint32_t *__restrict__ sub_key) {
  __shared__ uint32_t hist[SUB_BINS];
  __shared__ uint32_t cur[SUB_SMAX * SUB_SMAX];
  __shared__ double bd[2][SUB_SMAX + 1];
  const uint32_t lane = threadIdx.x;
  const uint32_t nbig = *big_cnt;
  for (uint32_t k = blockIdx.x; k < nbig; k += gridDim.x) {
    const uint32_t c = big_list[k];
    const uint32_t b = cell_start[c], m = cell_start[c + 1] - b;
    if (m <= SUB_MIN) continue;
    const uint32_t pr = prob_of_slot ? prob_of_slot[b] : 0u;
    const GridParams &g = grids[pr];
    if (g.inv_h[0] == 0. || g.inv_h[1] == 0. || g.inv_h[2] == 0.) continue;
    const uint32_t local = c - coff[pr];
    const uint32_t ca = local / ((uint32_t)g.G[1] * (uint32_t)g.G[2]);
    const uint32_t cb = (local / (uint32_t)g.G[2]) % (uint32_t)g.G[1];
    const uint32_t S = sub_side(m), ns = S * S;
    uint32_t base = 0, xo = 0;
    if (lane == 0) {
      base = atomicAdd(pool_cnt, ns + SUB_HEAD + 1u);
      xo = atomicAdd(xy_cnt, 2u * (S + 1u));
    }
    base = (uint32_t)__shfl((int)base, 0, 64);
    xo = (uint32_t)__shfl((int)xo, 0, 64);
    if ((uint64_t)base + ns + SUB_HEAD + 1u > pool_cap || (uint64_t)xo + 2u * (S + 1u) > xy_cap)
      continue;
    // the quantile boundaries in x (d = 0) and y (d = 1)
    for (int d = 0; d < 2; d++) {
      const double c0 = g.lo[d] + (double)(d == 0 ? ca : cb) * g.h[d];
      const double fb = (double)SUB_BINS * g.inv_h[d];
      for (uint32_t i = lane; i < SUB_BINS; i += 64) hist[i] = 0u;
      __syncthreads();
      for (uint32_t i = lane; i < m; i += 64) {
        const double4 p = spos[b + i];
        const double u = floor(((d == 0 ? p.x : p.y) - c0) * fb);
        const uint32_t e = u < 0. ? 0u : (uint32_t)fmin(u, (double)(SUB_BINS - 1));
        atomicAdd(&hist[e], 1u);
      }
      __syncthreads();
      // inclusive scan, 64 bins at a time
      uint32_t carry = 0;
      for (uint32_t s0 = 0; s0 < SUB_BINS; s0 += 64) {
        const uint32_t v = hist[s0 + lane];
        uint32_t tot;
        const uint32_t ex = wave_excl_scan(v, &tot);
        hist[s0 + lane] = carry + ex + v;
        carry += tot;
      }
      __syncthreads();
      if (lane >= 1 && lane < S) {
        // the first bin whose cumulative count reaches ceil(lane m / S);
        // the boundary is its upper edge
        const uint32_t target = (uint32_t)(((uint64_t)lane * m + S - 1u) / S);
        uint32_t lo = 0, hi = SUB_BINS - 1u;
        while (lo < hi) {
          const uint32_t mid = (lo + hi) >> 1;
          if (hist[mid] >= target)
            hi = mid;
          else
            lo = mid + 1;
        }
        bd[d][lane] = c0 + (double)(lo + 1u) / fb;
      }
      if (lane == 0) {
        bd[d][0] = -INFINITY;
        bd[d][S] = INFINITY;
      }
      __syncthreads();
    }
    // slab of v: inner boundaries <= v
    auto slab = [&](double v, const double *bv) -> uint32_t {
      uint32_t lo = 1, hi = S;
      while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (bv[mid] <= v)
          lo = mid + 1;
        else
          hi = mid;
      }
      return lo - 1u;
    };
    auto sub_of = [&](const double4 &p) -> uint32_t {
      return slab(p.x, bd[0]) * S + slab(p.y, bd[1]);
    };
    for (uint32_t i = lane; i < ns; i += 64) cur[i] = 0u;
    __syncthreads();
    for (uint32_t i = lane; i < m; i += 64) atomicAdd(&cur[sub_of(spos[b + i])], 1u);
    __syncthreads();
    // exclusive scan of the sub-column counts: their first positions
    uint32_t carry = 0;
    for (uint32_t s0 = 0; s0 < ns; s0 += 64) {
      const uint32_t i = s0 + lane;
      const uint32_t v = i < ns ? cur[i] : 0u;
      uint32_t tot;
      const uint32_t ex = wave_excl_scan(v, &tot);
      if (i < ns) {
        cur[i] = b + carry + ex;
        pool[base + SUB_HEAD + i] = b + carry + ex;
      }
      carry += tot;
    }
    if (lane == 0) {
      pool[base] = S;
      pool[base + 1] = xo;
      pool[base + SUB_HEAD + ns] = b + m;
    }
    if (lane <= S) {
      xy[xo + lane] = bd[0][lane];
      xy[xo + S + 1u + lane] = bd[1][lane];
    }
    __syncthreads();
    for (uint32_t i0 = 0; i0 < m; i0 += 64) {
      const uint32_t i = i0 + lane;
      const bool valid = i < m;
      double4 p = make_double4(0., 0., 0., 0.);
      if (valid) p = spos[b + i];
      const uint32_t sid = valid ? sub_of(p) : ~0u;
      uint32_t rank = 0;
      bool last = true;
      for (int j = 0; j < 64; j++) {
        const uint32_t sj = (uint32_t)__builtin_amdgcn_readlane((int)sid, j);
        if (sj == sid) {
          if (j < (int)lane) rank++;
          if (j > (int)lane) last = false;
        }
      }
      if (valid) {
        const uint32_t pos = cur[sid] + rank;
        sub_slot[pos] = b + i;
        sub_key[pos] = zkey(p.z);
      }
      __syncthreads();
      if (valid && last) cur[sid] += rank + 1u;
      __syncthreads();
    }
    if (lane == 0) {
      sub_base[c] = base;
      atomicAdd(xy_cnt + 1, 1u);  // cells with sub-columns
    }
    __syncthreads();
  }
}

// sorted positions; w carries the generator index bits
__global__ void k_grid_gather(const double4 *__restrict__ gpos,
                              const uint32_t *__restrict__ order, size_t n,
                              const uint32_t *__restrict__ prob,
                              double4 *__restrict__ spos,
                              uint32_t *__restrict__ slot_of,
                              uint32_t *__restrict__ prob_of_slot) {
  size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  uint32_t i = order[s];
  double4 p = gpos[i];
  p.w = __longlong_as_double((long long)i);
  spos[s] = p;
  slot_of[i] = (uint32_t)s;
  if (prob) prob_of_slot[s] = prob[i];
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

// total of an exclusive scan (last output + last input), written where the
// host reads it (pinned host memory)
template <typename TIn, typename TOut>
__global__ void k_scan_total(const TIn *__restrict__ last_in, const TOut *__restrict__ last_out,
                             unsigned long long *__restrict__ total) {
  *total = (unsigned long long)(*last_out + (TOut)*last_in);
}

// two device words, written where the host reads them
__global__ void k_word_u32(const uint32_t *__restrict__ v, unsigned long long *__restrict__ out) {
  out[0] = v[0];
  out[1] = v[1];
}

// largest of n values (one block), written where the host reads it
__global__ void __launch_bounds__(1024) k_max_u32(const uint32_t *__restrict__ v, uint32_t n,
                                                  unsigned long long *__restrict__ out) {
  __shared__ uint32_t part[32];
  uint32_t m = 0;
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) m = max(m, v[i]);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, o, 64));
  if (__lane_id() == 0) part[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t r = 0;
    for (uint32_t w = 0; w < (blockDim.x >> 6); w++) r = max(r, part[w]);
    *out = r;
  }
}

// the pass counters (staging, error word, exact tests, overflow) and the
// capacity flag, written where the host reads them
__global__ void k_pass_words(const unsigned long long *__restrict__ counters,
                             const unsigned int *__restrict__ cap_err,
                             const unsigned int *__restrict__ handed,
                             unsigned long long *__restrict__ out) {
  const int i = threadIdx.x;
  if (i < 4) out[i] = counters[i];
  if (i == 4) out[4] = *cap_err;
  if (i == 5) out[5] = *handed;
}

// ---------------------------------------------------------------------
// k nearest stream entries (list length L) per work item, one wavefront
// per item.
//
// Order: ascending (key, idx, image), key = wrapping_distance_2 of
// src/rtree_nn.rs:138-146: dx = (q + s) - g, (dx0^2 + dx1^2) + dx2^2 (for
// non-periodic builds the reference's distance_squared, src/rtree_nn.rs:
// 124-134, gives the same bits).
//
// The wavefront gathers every entry with key <= R^2: for each image s the
// grid rows (a, b, c0..c1) crossing the cube around q + s are contiguous
// ranges of the sorted positions, read 64 at a time (coalesced); entries
// inside the ball are appended to an LDS buffer with a ballot.  When the
// ball holds at least L entries (or the whole stream) the L smallest are
// exactly the first L of the stream; each lane ranks its entries by
// counting the smaller ones in the buffer and writes those of rank < L.
// Otherwise R grows (or shrinks when the buffer overflows) and the gather
// is repeated.  Lists are item-major: entry j of item w at kofs[w] + j
// (w * L without offsets).
// ---------------------------------------------------------------------
__device__ __forceinline__ bool entry_less(double ka, uint64_t pa, double kb,
                                           uint64_t pb) {
  return ka < kb || (ka == kb && pa < pb);
}

// Distance from q to a stream entry, computed exactly as the clipper does
// for its stop test (dx = q - (g + shift), sqrt of the dot product): the
// first entry left out of a truncated list is the one the clipper would
// test next, so `sr < entry_dist(...)` is the reference's stop decision
// for it (convex_cell.rs:331-333) without holding the entry itself.
__device__ __forceinline__ double entry_dist(const double4 *__restrict__ spos,
                                            const double *width, const double q[3],
                                            uint64_t pk) {
  d3 ngb = xyz(spos[pk_slot(pk)]);
  d3 s;
  if (image_shift(pk_img(pk), width, &s)) ngb = add3(ngb, s);
  const d3 dx = sub3(mk3(q[0], q[1], q[2]), ngb);
  return sqrt(dot3(dx, dx));
}

// Gather every stream entry with key <= R^2 into (key, pk) buffers (LDS or
// global).  Entries beyond `cap` are counted but not stored.  The count is
// wave-uniform.
//
// Latency-parallel walk: for each periodic image the grid rows (a, b) that
// reach the ball are spread over the lanes (one row per lane, 64 at a time);
// each lane trims its row's c-range to the ball's chord over that column
// and reads the row's slot range with two loads, all rows at once.  The
// candidates of all rows are then numbered consecutively (wave prefix sum)
// and read 64 at a time, each lane finding its row by a binary search over
// the rows' offsets held in the other lanes' registers.  A gather thus waits
// on one dependent load per 64 rows plus one per 64 candidates, instead of
// two per row.
// BLK: the wavefronts of a block share the rows (wavefront wid of nw) and
// reserve their output slots through the LDS counter *blk_cnt (zeroed by
// the caller), which ends as the count; the return value is then unused.
// SUB: the grid has dense cells with sub-columns (sc.base set); without
// them the gather is compiled without that path.
template <bool BLK = false, bool SUB = false>
__device__ __forceinline__ uint32_t knn_gather(
    const double4 *__restrict__ spos, const uint32_t *__restrict__ cell_start,
    const SubCols &sc, const GridParams &g, const double *width, const double q[3], double R,
    double *bkey, uint64_t *bpk, uint32_t cap, uint32_t *blk_cnt = nullptr, int wid = 0,
    int nw = 1, uint32_t *reads = nullptr) {
  const uint32_t lane = __lane_id();
  const double R2 = R * R;
  const double rr = R + g.margin;
  const double rr2 = rr * rr;
  uint32_t cnt = 0;
  // images per axis, nested: an image that misses the grid on an outer axis
  // skips its inner loops (most balls cross no periodic face)
  auto axis = [&](int c, int sv, double &qs, int &lo, int &hi) -> bool {
    qs = q[c] + (double)sv * width[c];
    if (g.inv_h[c] == 0.) {
      lo = 0;
      hi = 0;
      return true;
    }
    const double a = floor((qs - rr - g.lo[c]) * g.inv_h[c]);
    const double b = floor((qs + rr - g.lo[c]) * g.inv_h[c]);
    if (b < 0. || a > (double)(g.G[c] - 1)) return false;
    lo = a < 0. ? 0 : (int)a;
    hi = b > (double)(g.G[c] - 1) ? g.G[c] - 1 : (int)b;
    return true;
  };
  // [t0, t0 + len) cut to the seeds whose zkey lies in [klo, khi] (the
  // range is in zkey order; keys[t] is the key of position t, or null for
  // the key of spos[t])
  auto zcut = [&](uint32_t &t0, uint32_t &len, uint32_t klo, uint32_t khi,
                  const uint32_t *keys) {
    auto key = [&](uint32_t t) { return keys ? keys[t] : zkey(spos[t].z); };
    uint32_t lo = t0, hi = t0 + len;
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (key(mid) < klo)
        lo = mid + 1;
      else
        hi = mid;
    }
    uint32_t lo2 = lo, hi2 = t0 + len;
    while (lo2 < hi2) {
      const uint32_t mid = (lo2 + hi2) >> 1;
      if (key(mid) <= khi)
        lo2 = mid + 1;
      else
        hi2 = mid;
    }
    t0 = lo;
    len = lo2 - lo;
  };
  // the candidates of the lanes' ranges (positions t0 + j, or ind[t0 + j]
  // with an index), numbered consecutively and read 64 at a time
  auto read = [&](uint32_t t0, uint32_t len, const uint32_t *ind, const double *qs,
                  uint32_t img) {
    uint32_t total;
    const uint32_t off = wave_excl_scan(len, &total);
    if (reads) *reads += total;
    for (uint32_t cb = 0; cb < total; cb += 64) {
      const uint32_t gi = cb + lane;
      // largest row r with off[r] <= gi (rows past the end hold
      // off = total, empty rows share the next row's offset)
      uint32_t rsel = 0;
#pragma unroll
      for (uint32_t st = 32; st > 0; st >>= 1) {
        const uint32_t o = (uint32_t)__shfl((int)off, (int)(rsel + st), 64);
        if (rsel + st < 64 && o <= gi) rsel += st;
      }
      const uint32_t orow = (uint32_t)__shfl((int)off, (int)rsel, 64);
      const uint32_t trow = (uint32_t)__shfl((int)t0, (int)rsel, 64);
      bool keep = false;
      double key = 0.;
      uint64_t pk = 0;
      if (gi < total) {
        const uint32_t tp = trow + (gi - orow);
        const uint32_t t = ind ? ind[tp] : tp;
        const double4 gp = spos[t];
        const double d0 = qs[0] - gp.x, d1 = qs[1] - gp.y, dz = qs[2] - gp.z;
        key = (d0 * d0 + d1 * d1) + dz * dz;
        keep = key <= R2;
        const uint64_t gidx = (uint64_t)__double_as_longlong(gp.w);
        pk = (gidx << 32) | ((uint64_t)t << 5) | img;
      }
      const uint64_t m = wave_ballot(keep);
      uint32_t base = cnt;
      if (BLK && m) {
        uint32_t b0 = 0;
        if (lane == 0) b0 = atomicAdd(blk_cnt, (uint32_t)__popcll(m));
        base = (uint32_t)__shfl((int)b0, 0, 64);
      }
      const uint32_t pos = base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
      if (keep && pos < cap) {
        bkey[pos] = key;
        bpk[pos] = pk;
      }
      cnt += (uint32_t)__popcll(m);
    }
  };
  for (int i = g.img_lo[0]; i <= g.img_hi[0]; i++) {
    double qs[3];
    int lo[3], hi[3];
    if (!axis(0, i, qs[0], lo[0], hi[0])) continue;
    for (int j = g.img_lo[1]; j <= g.img_hi[1]; j++) {
      if (!axis(1, j, qs[1], lo[1], hi[1])) continue;
      for (int k = g.img_lo[2]; k <= g.img_hi[2]; k++) {
        if (!axis(2, k, qs[2], lo[2], hi[2])) continue;
        const uint32_t img = (uint32_t)((i + 1) * 9 + (j + 1) * 3 + (k + 1));
        const int nb_ = hi[1] - lo[1] + 1;
        const int nrows = (hi[0] - lo[0] + 1) * nb_;
        for (int r0 = BLK ? 64 * wid : 0; r0 < nrows; r0 += BLK ? 64 * nw : 64) {
          const int r = r0 + (int)lane;
          uint32_t t0 = 0, len = 0;
          // dense cells with sub-columns at the row's ends: read after the
          // rows, sub-column by sub-column
          uint32_t dfirst = ~0u, dlast = ~0u;
          if (r < nrows) {
            const int a = lo[0] + r / nb_, b = lo[1] + r % nb_;
            // squared distance from qs to the column of cells (a, b)
            double d2 = 0.;
            if (g.inv_h[0] != 0.) {
              const double x0 = g.lo[0] + (double)a * g.h[0], x1 = x0 + g.h[0];
              const double dx = fmax(0., fmax(x0 - qs[0], qs[0] - x1));
              d2 += dx * dx;
            }
            if (g.inv_h[1] != 0.) {
              const double y0 = g.lo[1] + (double)b * g.h[1], y1 = y0 + g.h[1];
              const double dy = fmax(0., fmax(y0 - qs[1], qs[1] - y1));
              d2 += dy * dy;
            }
            if (d2 <= rr2) {
              int c0 = lo[2], c1 = hi[2];
              if (g.inv_h[2] != 0.) {
                const double hc = sqrt(rr2 - d2);
                const double za = floor((qs[2] - hc - g.lo[2]) * g.inv_h[2]);
                const double zb = floor((qs[2] + hc - g.lo[2]) * g.inv_h[2]);
                c0 = max(c0, za < 0. ? 0 : (int)fmin(za, (double)g.G[2]));
                c1 = min(c1, zb < 0. ? -1 : (int)fmin(zb, (double)g.G[2]));
              }
              if (c0 <= c1) {
                const size_t row = ((size_t)a * g.G[1] + b) * g.G[2];
                t0 = cell_start[row + c0];
                len = cell_start[row + c1 + 1] - t0;
                if (SUB && len > SUB_MIN) {
                  // the row's first and last cells, when dense, leave the
                  // range (the rest stays in zkey order)
                  const uint32_t te = t0 + len;
                  uint32_t s1 = t0, s2 = te;
                  const uint32_t e0 = c0 == c1 ? te : cell_start[row + c0 + 1];
                  if (e0 - t0 > SUB_MIN) {
                    dfirst = sc.base[row + c0];
                    if (dfirst != ~0u) s1 = e0;
                  }
                  if (c1 > c0) {
                    const uint32_t b1 = cell_start[row + c1];
                    if (te - b1 > SUB_MIN) {
                      dlast = sc.base[row + c1];
                      if (dlast != ~0u) s2 = b1;
                    }
                  }
                  t0 = s1;
                  len = s2 > s1 ? s2 - s1 : 0u;
                }
                if (len > 32u && g.inv_h[2] != 0.) {
                  // a long row (dense cells): its seeds are in zkey order;
                  // keep those whose z can lie in the chord
                  const double hc = sqrt(rr2 - d2);
                  zcut(t0, len, zkey(qs[2] - hc), zkey(qs[2] + hc), nullptr);
                }
              }
            }
          }
          read(t0, len, nullptr, qs, img);
          // the dense cells: the sub-columns that reach the ball, each cut
          // to its chord, the lanes over the sub-columns
          for (int side = 0; side < (SUB ? 2 : 0); side++) {
            uint64_t dmask = wave_ballot((side ? dlast : dfirst) != ~0u);
            while (dmask) {
              const int l = (int)__builtin_ctzll(dmask);
              dmask &= dmask - 1ull;
              const uint32_t sb =
                  (uint32_t)__builtin_amdgcn_readlane((int)(side ? dlast : dfirst), l);
              const uint32_t S = sc.start[sb], xo = sc.start[sb + 1];
              const uint32_t *st = sc.start + sb + SUB_HEAD;
              // boundary `lane` of x and of y
              const double Xl = lane <= S ? sc.xy[xo + lane] : INFINITY;
              const double Yl = lane <= S ? sc.xy[xo + S + 1u + lane] : INFINITY;
              const bool inner = lane >= 1u && lane < S;
              // the slabs that meet [qs - rr, qs + rr] in x and in y
              const uint32_t i0 = (uint32_t)__popcll(wave_ballot(inner && Xl <= qs[0] - rr));
              const uint32_t i1 = (uint32_t)__popcll(wave_ballot(inner && Xl <= qs[0] + rr));
              const uint32_t j0 = (uint32_t)__popcll(wave_ballot(inner && Yl <= qs[1] - rr));
              const uint32_t j1 = (uint32_t)__popcll(wave_ballot(inner && Yl <= qs[1] + rr));
              const uint32_t ny = j1 - j0 + 1u, nxy = (i1 - i0 + 1u) * ny;
              for (uint32_t s0 = 0; s0 < nxy; s0 += 64) {
                const uint32_t e = min(s0 + lane, nxy - 1u);
                const uint32_t ex = e / ny;
                const uint32_t si = i0 + ex, sj = j0 + (e - ex * ny);
                const double xa = __shfl(Xl, (int)si, 64), xb = __shfl(Xl, (int)si + 1, 64);
                const double ya = __shfl(Yl, (int)sj, 64), yb = __shfl(Yl, (int)sj + 1, 64);
                uint32_t u0 = 0, ulen = 0;
                if (s0 + lane < nxy) {
                  const double dx = fmax(0., fmax(xa - qs[0], qs[0] - xb));
                  const double dy = fmax(0., fmax(ya - qs[1], qs[1] - yb));
                  const double d2 = dx * dx + dy * dy;
                  if (d2 <= rr2) {
                    const uint32_t sidx = si * S + sj;
                    u0 = st[sidx];
                    ulen = st[sidx + 1] - u0;
                    if (ulen > 16u) {
                      const double hc = sqrt(rr2 - d2);
                      zcut(u0, ulen, zkey(qs[2] - hc), zkey(qs[2] + hc), sc.key);
                    }
                  }
                }
                read(u0, ulen, sc.slot, qs, img);
              }
            }
          }
        }
      }
    }
  }
  return cnt;
}

// Radius search: the ball must hold at least L entries (or the whole
// stream) and at most `cap`.  Growth/shrink steps from the density, then
// bisection between a too-small and a too-large radius.  Returns the count
// (> cap when no radius fits).
template <bool SUB>
__device__ __forceinline__ uint32_t knn_fill(
    const double4 *__restrict__ spos, const uint32_t *__restrict__ cell_start,
    const SubCols &sc, const GridParams &g, const double *width, const double q[3], double &R,
    uint32_t Lw, double *bkey, uint64_t *bpk, uint32_t cap, uint32_t *iters, uint32_t *reads) {
  double r_lo = 0., r_hi = 0.;
  uint32_t cnt = 0;
  for (int it = 0; it < 96; it++) {
    *iters = it + 1;
    cnt = knn_gather<false, SUB>(spos, cell_start, sc, g, width, q, R, bkey, bpk, cap, nullptr, 0,
                                 1, reads);
    bool all = (uint64_t)cnt >= g.total_stream;
    if (cnt <= cap && (cnt >= Lw || all)) return cnt;
    if (cnt > cap) {
      r_hi = R;
      if (r_lo > 0.) {
        R = 0.5 * (r_lo + r_hi);
      } else {
        double f = cbrt(0.6 * (double)cap / (double)cnt);
        R *= fmax(0.5, fmin(0.9, f));
      }
    } else {
      r_lo = R;
      if (r_hi > 0.) {
        R = 0.5 * (r_lo + r_hi);
      } else {
        double f = cnt ? cbrt(1.4 * (double)Lw / (double)cnt) : 2.0;
        R *= fmax(1.2, fmin(4.0, f));
      }
    }
    if (r_hi > 0. && r_lo > 0. && r_hi - r_lo <= 1e-15 * r_hi) break;
  }
  return cnt > cap ? cnt : cap + 1;
}

// knn_fill with every wavefront of a KB_THREADS block gathering (all threads
// call it; the count comes back through the LDS word *blk_cnt).
// *iters: the gathers made.  take_ball: the first ball is kept whatever its
// count (up to cap; the list ends at the ball, certified by its radius).
template <bool SUB>
__device__ __forceinline__ uint32_t knn_fill_blk(
    const double4 *__restrict__ spos, const uint32_t *__restrict__ cell_start,
    const SubCols &sc, const GridParams &g, const double *width, const double q[3], double &R,
    uint32_t Lw, double *bkey, uint64_t *bpk, uint32_t cap, uint32_t *blk_cnt, int wid,
    int nw, uint32_t *iters, bool take_ball = false) {
  double r_lo = 0., r_hi = 0.;
  uint32_t cnt = 0;
  for (int it = 0; it < 96; it++) {
    *iters = it + 1;
    if (threadIdx.x == 0) *blk_cnt = 0u;
    __syncthreads();
    knn_gather<true, SUB>(spos, cell_start, sc, g, width, q, R, bkey, bpk, cap, blk_cnt, wid, nw);
    __syncthreads();
    cnt = *blk_cnt;
    __syncthreads();
    bool all = (uint64_t)cnt >= g.total_stream;
    if (cnt <= cap && (cnt >= Lw || all || (take_ball && it == 0))) return cnt;
    if (cnt > cap) {
      r_hi = R;
      if (r_lo > 0.) {
        R = 0.5 * (r_lo + r_hi);
      } else {
        double f = cbrt(0.6 * (double)cap / (double)cnt);
        R *= fmax(0.5, fmin(0.9, f));
      }
    } else {
      r_lo = R;
      if (r_hi > 0.) {
        R = 0.5 * (r_lo + r_hi);
      } else {
        double f = cnt ? cbrt(1.4 * (double)Lw / (double)cnt) : 2.0;
        R *= fmax(1.2, fmin(4.0, f));
      }
    }
    if (r_hi > 0. && r_lo > 0. && r_hi - r_lo <= 1e-15 * r_hi) break;
  }
  return cnt > cap ? cnt : cap + 1;
}

// Lower bound on the distance (as the clipper computes it) of every stream
// entry outside a gathered ball of radius R around q: the ball test uses the
// key ((q + s) - g)^2, the clipper |q - (g + s')|, which differ by rounding
// on the scale of the coordinates (and the periodic width).
__device__ __forceinline__ double ball_certificate(const double q[3], double R,
                                                  const double *width) {
  const double M = fmax(fabs(q[0]), fmax(fabs(q[1]), fabs(q[2]))) +
                   fmax(width[0], fmax(width[1], width[2])) + R;
  return R - 1e-13 * M;
}

// Start radius of a list of Lw entries for a seed q of a dense cell with
// sub-columns: the density around q from the z-span of the (up to) 17
// seeds of q's sub-column nearest to q in z order, over the sub-column's
// box; 0 when q's cell has no sub-columns, q's sub-column is unbounded (an
// outer one) or holds fewer than 8 seeds.  Inside a seed cluster it is far
// closer to the radius that holds Lw seeds than one from the density of a
// block of grid cells.
__device__ __forceinline__ double dense_start_radius(
    const double4 *__restrict__ spos, const SubCols &sc, const GridParams &g,
    const double q[3], uint32_t Lw) {
  if (!sc.base || g.inv_h[0] == 0. || g.inv_h[1] == 0. || g.inv_h[2] == 0.) return 0.;
  int C[3];
  for (int c = 0; c < 3; c++) {
    const double f = floor((q[c] - g.lo[c]) * g.inv_h[c]);
    C[c] = (int)fmax(0., fmin(f, (double)(g.G[c] - 1)));
  }
  const uint32_t sb = sc.base[((size_t)C[0] * g.G[1] + C[1]) * g.G[2] + C[2]];
  if (sb == ~0u) return 0.;
  const uint32_t S = sc.start[sb], xo = sc.start[sb + 1];
  auto slab = [&](double v, const double *bv) -> uint32_t {
    uint32_t lo = 1, hi = S;
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (bv[mid] <= v)
        lo = mid + 1;
      else
        hi = mid;
    }
    return lo - 1u;
  };
  const double *X = sc.xy + xo, *Y = sc.xy + xo + S + 1u;
  const uint32_t i = slab(q[0], X), j = slab(q[1], Y);
  const double area = (X[i + 1] - X[i]) * (Y[j + 1] - Y[j]);
  if (!(area > 0.) || !isfinite(area)) return 0.;
  const uint32_t *st = sc.start + sb + SUB_HEAD;
  const uint32_t u0 = st[i * S + j], u1 = st[i * S + j + 1];
  if (u1 - u0 < 8u) return 0.;
  const uint32_t kz = zkey(q[2]);
  uint32_t lo = u0, hi = u1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (sc.key[mid] < kz)
      lo = mid + 1;
    else
      hi = mid;
  }
  const uint32_t K = min(8u, (u1 - u0) / 2u);
  uint32_t a = lo >= u0 + K ? lo - K : u0;
  const uint32_t b = min(u1 - 1u, a + 2u * K);
  a = b >= u0 + 2u * K ? b - 2u * K : u0;
  const double dz = spos[sc.slot[b]].z - spos[sc.slot[a]].z;
  if (!(dz > 0.)) return 0.;
  const double rho = (double)(b - a) / (area * dz);
  return cbrt((1.3 * (double)Lw + 8.) / rho / 4.18879020478639);
}

__device__ __forceinline__ double knn_start_radius(
    const uint32_t *__restrict__ cell_start, const GridParams &g,
    const double q[3], uint32_t Lw) {
  int nd = 0;
  for (int c = 0; c < 3; c++) nd += g.inv_h[c] != 0.;
  int C[3];
  for (int c = 0; c < 3; c++) {
    double f = floor((q[c] - g.lo[c]) * g.inv_h[c]);
    C[c] = g.inv_h[c] == 0. ? 0 : (int)fmax(0., fmin(f, (double)(g.G[c] - 1)));
  }
  uint32_t n27 = 0;
  int cells = 0;
  for (int a = max(0, C[0] - 1); a <= min(g.G[0] - 1, C[0] + 1); a++)
    for (int b = max(0, C[1] - 1); b <= min(g.G[1] - 1, C[1] + 1); b++) {
      size_t row = ((size_t)a * g.G[1] + b) * g.G[2];
      int c0 = max(0, C[2] - 1), c1 = min(g.G[2] - 1, C[2] + 1);
      n27 += cell_start[row + c1 + 1] - cell_start[row + c0];
      cells += c1 - c0 + 1;
    }
  double hv = 1.;
  for (int c = 0; c < 3; c++)
    if (g.inv_h[c] != 0.) hv *= g.h[c];
  double rho = (double)max(n27, 1u) / ((double)cells * hv);
  double vol = (1.3 * (double)Lw + 8.) / rho;
  if (nd == 3) return cbrt(vol / 4.18879020478639);
  if (nd == 2) return sqrt(vol / 3.14159265358979);
  return 0.5 * vol;
}

// LDS variant: rank sort (each lane counts the smaller entries)
template <int CAP>
struct KnnLds {
  double key[CAP];
  uint64_t pk[CAP];
};

// Bitonic sort of cnt (key, pk) entries held in LDS by one wavefront,
// padded with +inf keys to a power of two (the buffer must hold it).
__device__ __forceinline__ void wave_bitonic_lds(double *key, uint64_t *pk, uint32_t cnt) {
  const uint32_t lane = __lane_id();
  uint32_t n2 = 64;
  while (n2 < cnt) n2 <<= 1;
  for (uint32_t e = cnt + lane; e < n2; e += 64) {
    key[e] = INFINITY;
    pk[e] = ~0ull;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  for (uint32_t k = 2; k <= n2; k <<= 1)
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      // lane handles the pairs (i, i ^ j) with i ^ j > i: n2 / 2 pairs
      for (uint32_t p = lane; p < (n2 >> 1); p += 64) {
        const uint32_t i = ((p & ~(j - 1)) << 1) | (p & (j - 1));
        const uint32_t l = i | j;
        const double ki = key[i], kl = key[l];
        const uint64_t pi = pk[i], pl = pk[l];
        const bool up = (i & k) == 0;
        const bool sw = up ? entry_less(kl, pl, ki, pi) : entry_less(ki, pi, kl, pl);
        if (sw) {
          key[i] = kl;
          key[l] = ki;
          pk[i] = pl;
          pk[l] = pi;
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// Lane exchange with lane ^ J: DPP quad permutes for J = 1, 2, a swizzle
// within 32 lanes for J = 4, 8, 16, a backward permute for J = 32.
template <int J>
__device__ __forceinline__ int xchg_i(int v) {
  if constexpr (J == 1)
    return __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, false);  // quad_perm 1,0,3,2
  else if constexpr (J == 2)
    return __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, false);  // quad_perm 2,3,0,1
  else if constexpr (J < 32)
    return __builtin_amdgcn_ds_swizzle(v, (J << 10) | 0x1F);  // and 31, xor J
  else
    return __shfl_xor(v, J, 64);
}
template <int J>
__device__ __forceinline__ void xchg_entry(double &k, uint64_t &p, double &ko, uint64_t &po) {
  const long long kb = __double_as_longlong(k);
  const int klo = xchg_i<J>((int)(kb & 0xffffffffll)), khi = xchg_i<J>((int)(kb >> 32));
  const int plo = xchg_i<J>((int)(p & 0xffffffffull)), phi = xchg_i<J>((int)(p >> 32));
  ko = __longlong_as_double(((long long)khi << 32) | (long long)(unsigned int)klo);
  po = ((uint64_t)(unsigned int)phi << 32) | (uint64_t)(unsigned int)plo;
}
// one compare-exchange stage (k, j < 64) on the H entries a lane holds
// (entry lane + 64 h of the sequence)
template <int K, int J, int H>
__device__ __forceinline__ void bitonic_stage(double (&key)[H], uint64_t (&pk)[H]) {
  const uint32_t lane = __lane_id();
#pragma unroll
  for (int h = 0; h < H; h++) {
    double ko;
    uint64_t po;
    xchg_entry<J>(key[h], pk[h], ko, po);
    const uint32_t i = lane + 64u * h;
    // the lower index of a pair keeps the smaller entry in an ascending
    // block, the larger in a descending one
    const bool keep_min = ((i & J) == 0) == ((i & K) == 0);
    if (keep_min != entry_less(key[h], pk[h], ko, po)) {
      key[h] = ko;
      pk[h] = po;
    }
  }
}
template <int K, int J, int H>
__device__ __forceinline__ void bitonic_merge(double (&key)[H], uint64_t (&pk)[H]) {
  if constexpr (J == 64) {
    // pairs (lane, lane + 64) in registers (K = 128: ascending)
    if (entry_less(key[1], pk[1], key[0], pk[0])) {
      const double tk = key[0];
      const uint64_t tp = pk[0];
      key[0] = key[1];
      pk[0] = pk[1];
      key[1] = tk;
      pk[1] = tp;
    }
  } else {
    bitonic_stage<K, J, H>(key, pk);
  }
  if constexpr (J > 1) bitonic_merge<K, J / 2, H>(key, pk);
}
template <int K, int N, int H>
__device__ __forceinline__ void bitonic_sort_reg(double (&key)[H], uint64_t (&pk)[H]) {
  bitonic_merge<K, K / 2, H>(key, pk);
  if constexpr (K < N) bitonic_sort_reg<2 * K, N, H>(key, pk);
}

// The same network on one 64-bit word per lane (the ascending order of
// the words): a squared distance's bits with its low 7 bits replaced by the
// entry's buffer slot.  Keys are >= +0, so their bit patterns order like the
// keys; the words are distinct, and two entries whose keys agree above the
// low 7 bits come out in slot order, which the caller checks against the
// exact (key, word) order afterwards.
template <int K, int J>
__device__ __forceinline__ void bitonic_stage_u64(uint64_t &v) {
  const uint32_t lane = __lane_id();
  const int lo = xchg_i<J>((int)(v & 0xffffffffull)), hi = xchg_i<J>((int)(v >> 32));
  const uint64_t o = ((uint64_t)(unsigned int)hi << 32) | (uint64_t)(unsigned int)lo;
  const bool keep_min = ((lane & J) == 0) == ((lane & K) == 0);
  if (keep_min != (v < o)) v = o;
}
template <int K, int J>
__device__ __forceinline__ void bitonic_merge_u64(uint64_t &v) {
  bitonic_stage_u64<K, J>(v);
  if constexpr (J > 1) bitonic_merge_u64<K, J / 2>(v);
}
template <int K>
__device__ __forceinline__ void bitonic_sort_u64(uint64_t &v) {
  bitonic_merge_u64<K, K / 2>(v);
  if constexpr (K < 64) bitonic_sort_u64<2 * K>(v);
}
__device__ __forceinline__ uint64_t packed_entry(const double *key, uint32_t s) {
  return ((uint64_t)__double_as_longlong(key[s]) & ~0x7Full) | (uint64_t)s;
}

// SUB: the grid has dense cells with sub-columns (a separate instantiation:
// that path's registers would cost the other one a wavefront per SIMD)
template <int CAP, int WPB, bool SUB>
__global__ void __launch_bounds__(256)
    k_knn(const double4 *__restrict__ spos,
          const uint32_t *__restrict__ cell_start_all, Problems PR,
          const uint32_t *__restrict__ work, uint32_t nwork,
          uint32_t L, const uint32_t *__restrict__ Lper,
          const float *__restrict__ Rper, const uint64_t *__restrict__ kofs,
          uint64_t *__restrict__ kpk, uint32_t *__restrict__ kcnt,
          double *__restrict__ kdist, const uint32_t *__restrict__ dn,
          unsigned long long *__restrict__ kstat, uint32_t proposal) {
  __shared__ KnnLds<CAP> lds[WPB];
  // dn: the item count on the device (nwork bounds it)
  if (dn) nwork = min(*dn, nwork);
  const uint32_t lane = __lane_id();
  const uint32_t wib = threadIdx.x >> 6;
  KnnLds<CAP> &buf = lds[wib];
  const uint32_t nwaves = gridDim.x * (blockDim.x >> 6);
  for (uint32_t w = blockIdx.x * (blockDim.x >> 6) + wib; w < nwork; w += nwaves) {
    // wave-uniform: keep them in scalar registers
    const uint32_t slot = __builtin_amdgcn_readfirstlane(work[w]);
    const uint32_t pr = __builtin_amdgcn_readfirstlane(problem_of(PR, slot));
    // the problem's parameters in registers (wave-uniform scalar loads)
    const GridParams g = PR.grids[pr];
    const BoxParams &boxr = PR.boxes[pr];
    struct {
      double width[3];
    } box = {{boxr.width[0], boxr.width[1], boxr.width[2]}};
    const uint32_t *cell_start = cell_start_all + PR.coff[pr];
    const double4 qq = spos[slot];
    const double q[3] = {qq.x, qq.y, qq.z};
    const uint32_t Lw = Lper ? Lper[w] : L;
    const SubCols sc = sub_cols(PR, pr);
    // radii of redone lists come from the previous ball; first radii
    // (proposed, or from the density of a block of grid cells) inside dense
    // cells come from the density in the seed's sub-column
    const bool given = Rper && Rper[w] > 0.f;
    double R = given ? (double)Rper[w] : knn_start_radius(cell_start, g, q, Lw);
    if (SUB && (!given || proposal)) {
      const double Rd = dense_start_radius(spos, sc, g, q, Lw);
      if (Rd > 0.) R = Rd;
    }
    // a ball with three quarters of the requested entries is kept (its list
    // ends at the ball, certified by its radius) rather than gathered again:
    // a cell that needs more is redone with a longer list anyway
    const uint32_t Lacc = max(1u, (Lw * 3u) / 4u);
    uint32_t iters = 0, reads = 0;
    uint32_t cnt = knn_fill<SUB>(spos, cell_start, sc, g, box.width, q, R, Lacc, buf.key,
                                 buf.pk, CAP, &iters, &reads);
    // ball gathers and candidates read, on 64 counters each (by block) the
    // host adds up
    if (kstat && lane == 0) {
      atomicAdd(kstat + (blockIdx.x & 63u), (unsigned long long)iters);
      atomicAdd(kstat + 64 + (blockIdx.x & 63u), (unsigned long long)reads);
    }
    __builtin_amdgcn_wave_barrier();
    if (cnt > (uint32_t)CAP) {
      if (lane == 0) kcnt[w] = 0xFFFFFFFFu;  // does not fit this variant
      continue;
    }
    const uint32_t outn = min(cnt, Lw);
    uint64_t *out = kpk + (kofs ? kofs[w] : (uint64_t)w * L);
    // up to 128 entries: the packed network, then the exact order of
    // consecutive entries checked (order[] after the buffer's entries); on
    // a violation (keys equal above their low 7 bits, lattices) the exact
    // networks below sort the same buffer again
    bool done = false;
    if (cnt <= 128) {
      uint64_t va = lane < cnt ? packed_entry(buf.key, lane) : ~0ull;
      uint64_t vb = lane + 64 < cnt ? packed_entry(buf.key, lane + 64) : ~0ull;
      bitonic_sort_u64<2>(va);
      uint32_t posa = lane, posb = 64u + lane;
      if (cnt > 64) {
        bitonic_sort_u64<2>(vb);
        // final place: own rank plus the smaller words of the other half
        uint32_t ra = 0, rb = 0;
#pragma unroll
        for (uint32_t st = 64; st > 0; st >>= 1) {
          const uint32_t ia = min(ra + st - 1, 63u), ib = min(rb + st - 1, 63u);
          const uint64_t vbi = (uint64_t)__shfl((unsigned long long)vb, (int)ia, 64);
          const uint64_t vai = (uint64_t)__shfl((unsigned long long)va, (int)ib, 64);
          if (ra + st - 1 < 64 && vbi < va) ra += st;
          if (rb + st - 1 < 64 && vai < vb) rb += st;
        }
        posa = lane + ra;
        posb = lane + rb;
      }
      uint32_t *order = reinterpret_cast<uint32_t *>(buf.pk + 128);
      if (posa < cnt) order[posa] = (uint32_t)(va & 0x7F);
      if (posb < cnt) order[posb] = (uint32_t)(vb & 0x7F);
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      bool bad = false;
      for (uint32_t p = lane; p + 1 < cnt; p += 64) {
        const uint32_t s0 = order[p], s1 = order[p + 1];
        bad |= entry_less(buf.key[s1], buf.pk[s1], buf.key[s0], buf.pk[s0]);
      }
      if (!wave_ballot(bad)) {
        done = true;
        for (uint32_t p = lane; p < outn; p += 64) out[p] = buf.pk[order[p]];
        if (lane == 0 && outn < cnt) kdist[w] = entry_dist(spos, box.width, q, buf.pk[order[outn]]);
      }
    }
    if (done) {
    } else if (cnt <= 64) {
      // bitonic network across the lanes, one entry per lane
      double key[1] = {lane < cnt ? buf.key[lane] : INFINITY};
      uint64_t pk[1] = {lane < cnt ? buf.pk[lane] : ~0ull};
      bitonic_sort_reg<2, 64, 1>(key, pk);
      if (lane < outn) out[lane] = pk[0];
      if (lane == outn && outn < cnt) kdist[w] = entry_dist(spos, box.width, q, pk[0]);
    } else if (cnt <= 128) {
      // entries 0..63 and 64..127 sorted by the network one after the
      // other, then merged: an entry's final rank is its rank in its half
      // plus the number of smaller entries in the other half (a binary
      // search over the other half's lanes)
      double ka[1] = {buf.key[lane]}, kb[1] = {lane + 64 < cnt ? buf.key[lane + 64] : INFINITY};
      uint64_t pa[1] = {buf.pk[lane]}, pb[1] = {lane + 64 < cnt ? buf.pk[lane + 64] : ~0ull};
      bitonic_sort_reg<2, 64, 1>(ka, pa);
      bitonic_sort_reg<2, 64, 1>(kb, pb);
      uint32_t ra = 0, rb = 0;  // smaller entries of the other half
#pragma unroll
      for (uint32_t st = 64; st > 0; st >>= 1) {
        const uint32_t ia = min(ra + st - 1, 63u), ib = min(rb + st - 1, 63u);
        const double kbi = __shfl(kb[0], (int)ia, 64), kai = __shfl(ka[0], (int)ib, 64);
        const uint64_t pbi = (uint64_t)__shfl((unsigned long long)pb[0], (int)ia, 64);
        const uint64_t pai = (uint64_t)__shfl((unsigned long long)pa[0], (int)ib, 64);
        if (ra + st - 1 < 64 && entry_less(kbi, pbi, ka[0], pa[0])) ra += st;
        if (rb + st - 1 < 64 && entry_less(kai, pai, kb[0], pb[0])) rb += st;
      }
      const uint32_t posa = lane + ra, posb = lane + rb;
      if (posa < outn) out[posa] = pa[0];
      if (posb < outn) out[posb] = pb[0];
      if (outn < cnt) {
        if (posa == outn) kdist[w] = entry_dist(spos, box.width, q, pa[0]);
        if (posb == outn) kdist[w] = entry_dist(spos, box.width, q, pb[0]);
      }
    } else {
      wave_bitonic_lds(buf.key, buf.pk, cnt);
      for (uint32_t e = lane; e < outn; e += 64) out[e] = buf.pk[e];
      if (lane == 0 && cnt > outn) kdist[w] = entry_dist(spos, box.width, q, buf.pk[outn]);
    }
    if (lane == 0) {
      kcnt[w] = outn;
      // nothing left out: every entry outside lies beyond the ball
      if (cnt == outn) kdist[w] = ball_certificate(q, R, box.width);
    }
    __builtin_amdgcn_wave_barrier();
  }
}

// Global-memory variant for long lists: one block of 1024 threads per item
// owns two buffers of `cap` entries.  Its first wavefront gathers the ball
// into the first; the block bucket-sorts it: a histogram of the keys over
// KB_BINS equal key ranges (in LDS), a scatter into bin order in the second
// buffer, then windows of consecutive bins of at most KB_WIN entries sorted
// in LDS (a bin larger than that alone, in global memory) back into the
// first; the first L are written out.  Equal keys share a bin, so sorting
// each window by (key, index, image) orders the whole ball.
constexpr int KB_THREADS = 1024;
constexpr uint32_t KB_BINS = 8192;  // histogram bins (32 KiB of LDS)
constexpr uint32_t KB_WIN = 2000;   // entries per LDS window (32,000 B)

// Bitonic sort of m entries, all comparators ascending (the first stage of
// each merge pairs i with its mirror i ^ (k - 1)), so entries past m act as
// +inf and are never touched: no padding.  One block of KB_THREADS.
template <class K, class P>
__device__ __forceinline__ void flip_bitonic(K *key, P *pk, uint32_t m, uint32_t tid) {
  uint32_t n2 = 1;
  while (n2 < m) n2 <<= 1;
  for (uint32_t k = 2; k <= n2; k <<= 1) {
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      for (uint32_t p = tid; p < (n2 >> 1); p += KB_THREADS) {
        const uint32_t i = ((p & ~(j - 1)) << 1) | (p & (j - 1));
        const uint32_t l = j == (k >> 1) ? i ^ (k - 1) : i | j;
        if (l < m) {
          const double ka = key[i], kb = key[l];
          const uint64_t pa = pk[i], pb = pk[l];
          if (entry_less(kb, pb, ka, pa)) {
            key[i] = kb;
            pk[i] = pb;
            key[l] = ka;
            pk[l] = pa;
          }
        }
      }
      __syncthreads();
    }
  }
}

// The items of a long-list launch, longest lists first (64 length classes,
// two per octave; any order within a class), and the launch's work counter
// set to zero.  One block of 1024 threads; dn: the item count on the device
// (n bounds it).
constexpr uint32_t LEN_CLASSES = 64;
__device__ __forceinline__ uint32_t len_class(uint32_t L) {
  if (L < 2u) return LEN_CLASSES - 1u;
  const uint32_t b = 31u - (uint32_t)__clz((int)L);  // floor(log2 L) >= 1
  const uint32_t half = (L >> (b - 1u)) & 1u;         // the next bit
  return LEN_CLASSES - 1u - min(2u * b + half, LEN_CLASSES - 1u);
}
__global__ void __launch_bounds__(1024)
    k_order_by_len(const uint32_t *__restrict__ Lper, uint32_t n, uint32_t *__restrict__ perm,
                   uint32_t *__restrict__ queue, const uint32_t *__restrict__ dn) {
  __shared__ uint32_t cnt[LEN_CLASSES];
  if (dn) n = min(*dn, n);
  const uint32_t tid = threadIdx.x;
  if (tid < LEN_CLASSES) cnt[tid] = 0u;
  if (tid == 0) *queue = 0u;
  __syncthreads();
  for (uint32_t i = tid; i < n; i += blockDim.x) atomicAdd(&cnt[len_class(Lper[i])], 1u);
  __syncthreads();
  if (tid == 0) {
    uint32_t s = 0;
    for (uint32_t c = 0; c < LEN_CLASSES; c++) {
      const uint32_t v = cnt[c];
      cnt[c] = s;
      s += v;
    }
  }
  __syncthreads();
  for (uint32_t i = tid; i < n; i += blockDim.x)
    perm[atomicAdd(&cnt[len_class(Lper[i])], 1u)] = i;
}

template <bool SUB>
__global__ void __launch_bounds__(KB_THREADS)
    k_knn_big(const double4 *__restrict__ spos,
              const uint32_t *__restrict__ cell_start_all, Problems PR,
              const uint32_t *__restrict__ work,
              uint32_t nwork, uint32_t L, const uint32_t *__restrict__ Lper,
              const float *__restrict__ Rper,
              const uint64_t *__restrict__ kofs, uint64_t *__restrict__ kpk,
              uint32_t *__restrict__ kcnt, double *__restrict__ kdist,
              double *__restrict__ gkey, uint64_t *__restrict__ gpk,
              double *__restrict__ gkey2, uint64_t *__restrict__ gpk2,
              uint32_t cap, const uint32_t *__restrict__ dn,
              unsigned long long *__restrict__ gstat, const uint32_t *__restrict__ perm,
              uint32_t *__restrict__ queue) {
  if (dn) nwork = min(*dn, nwork);
  __shared__ uint32_t hist[KB_BINS];
  __shared__ double wkey[KB_WIN];
  __shared__ uint64_t wpk[KB_WIN];
  __shared__ uint32_t s_wsum[KB_THREADS / 64];
  __shared__ uint32_t s_gather, s_next;
  const uint32_t tid = threadIdx.x;
  const uint32_t lane = __lane_id();
  const uint32_t wid = tid >> 6;
  double *bkey = gkey + (size_t)blockIdx.x * cap;
  uint64_t *bpk = gpk + (size_t)blockIdx.x * cap;
  double *tkey = gkey2 + (size_t)blockIdx.x * cap;
  uint64_t *tpk = gpk2 + (size_t)blockIdx.x * cap;
  // items in `perm` order, claimed one at a time from *queue (list costs
  // differ by orders of magnitude), or blockIdx.x, + gridDim.x, ...
  auto claim = [&]() -> uint32_t {
    __syncthreads();
    if (tid == 0) s_next = atomicAdd(queue, 1u);
    __syncthreads();
    return s_next;
  };
  for (uint32_t j = queue ? claim() : blockIdx.x; j < nwork;
       j = queue ? claim() : j + gridDim.x) {
    const uint32_t w = perm ? perm[j] : j;
    const uint32_t slot = work[w];
    const uint32_t pr = problem_of(PR, slot);
    const BoxParams &boxr = PR.boxes[pr];
    const double width[3] = {boxr.width[0], boxr.width[1], boxr.width[2]};
    const double4 qq = spos[slot];
    const double q[3] = {qq.x, qq.y, qq.z};
    const uint32_t Lw = Lper ? Lper[w] : L;
    // block-uniform: keep them in scalar registers
    const uint32_t pru = __builtin_amdgcn_readfirstlane(pr);
    const GridParams g = PR.grids[pru];
    const uint32_t *cell_start = cell_start_all + PR.coff[pru];
    // a negative radius: take the ball of that radius whole
    const float rp = Rper ? Rper[w] : 0.f;
    double R = rp > 0.f ? (double)rp : rp < 0.f ? -(double)rp
                                                : knn_start_radius(cell_start, g, q, Lw);
    uint32_t iters = 0;
    const SubCols sc = sub_cols(PR, pru);
    const uint32_t cnt = knn_fill_blk<SUB>(spos, cell_start, sc, g, width, q, R, Lw, bkey, bpk,
                                           cap, &s_gather, (int)wid, KB_THREADS / 64, &iters,
                                           rp < 0.f);
    if (gstat && tid == 0) {
      // items, ball gathers, entries gathered by the last one
      atomicAdd(gstat, 1ull);
      atomicAdd(gstat + 1, (unsigned long long)iters);
      atomicAdd(gstat + 2, (unsigned long long)cnt);
    }
    if (cnt > cap) {
      if (tid == 0) kcnt[w] = 0xFFFFFFFFu;
      __syncthreads();
      continue;
    }
    // histogram of the keys (all kept keys are <= R^2, the gather's bound)
    const double inv = (double)KB_BINS / (R * R);
    auto bin_of = [&](double k) { return min(KB_BINS - 1u, (uint32_t)(k * inv)); };
    for (uint32_t b = tid; b < KB_BINS; b += KB_THREADS) hist[b] = 0u;
    __syncthreads();
    for (uint32_t e = tid; e < cnt; e += KB_THREADS) atomicAdd(&hist[bin_of(bkey[e])], 1u);
    __syncthreads();
    {
      // exclusive scan of the bins: eight per thread, then across the block
      constexpr uint32_t PER = KB_BINS / KB_THREADS;
      uint32_t v[PER], sum = 0;
#pragma unroll
      for (uint32_t i = 0; i < PER; i++) {
        v[i] = hist[PER * tid + i];
        sum += v[i];
      }
      uint32_t x = sum;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)x, o, 64);
        if ((int)lane >= o) x += y;
      }
      if (lane == 63) s_wsum[wid] = x;
      __syncthreads();
      if (wid == 0) {
        uint32_t wv = lane < KB_THREADS / 64 ? s_wsum[lane] : 0u;
#pragma unroll
        for (int o = 1; o < KB_THREADS / 64; o <<= 1) {
          const uint32_t y = (uint32_t)__shfl_up((int)wv, o, 64);
          if ((int)lane >= o) wv += y;
        }
        if (lane < KB_THREADS / 64) s_wsum[lane] = wv;
      }
      __syncthreads();
      uint32_t base = (wid ? s_wsum[wid - 1] : 0u) + x - sum;
#pragma unroll
      for (uint32_t i = 0; i < PER; i++) {
        hist[PER * tid + i] = base;
        base += v[i];
      }
      __syncthreads();
    }
    // scatter into bin order; hist[b] ends as the end of bin b
    for (uint32_t e = tid; e < cnt; e += KB_THREADS) {
      const double k = bkey[e];
      const uint32_t pos = atomicAdd(&hist[bin_of(k)], 1u);
      tkey[pos] = k;
      tpk[pos] = bpk[e];
    }
    __syncthreads();
    // windows of whole bins, sorted back into the first buffer, up to the
    // one that holds entry outn (the first left out: its distance bounds
    // the list); the bins past it stay unsorted
    const uint32_t outn = min(cnt, Lw);
    for (uint32_t s0 = 0; s0 < cnt && s0 <= outn;) {
      // the bins that end at most KB_WIN past s0 (bin ends do not decrease)
      uint32_t lo = 0, hi = KB_BINS;
      const uint32_t lim = s0 + KB_WIN;
      while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (hist[mid] <= lim)
          lo = mid + 1;
        else
          hi = mid;
      }
      const uint32_t ew = lo ? hist[lo - 1] : 0u;
      if (ew > s0) {
        const uint32_t m = ew - s0;
        for (uint32_t i = tid; i < m; i += KB_THREADS) {
          wkey[i] = tkey[s0 + i];
          wpk[i] = tpk[s0 + i];
        }
        __syncthreads();
        flip_bitonic(wkey, wpk, m, tid);
        for (uint32_t i = tid; i < m; i += KB_THREADS) {
          bkey[s0 + i] = wkey[i];
          bpk[s0 + i] = wpk[i];
        }
        __syncthreads();
        s0 = ew;
      } else {
        // bin `lo` starts at s0 and holds more than KB_WIN entries
        const uint32_t e1 = hist[lo];
        flip_bitonic(tkey + s0, tpk + s0, e1 - s0, tid);
        for (uint32_t i = s0 + tid; i < e1; i += KB_THREADS) {
          bkey[i] = tkey[i];
          bpk[i] = tpk[i];
        }
        __syncthreads();
        s0 = e1;
      }
    }
    uint64_t *out = kpk + (kofs ? kofs[w] : (uint64_t)w * L);
    for (uint32_t e = tid; e < outn; e += KB_THREADS) out[e] = bpk[e];
    if (tid == 0) {
      kcnt[w] = outn;
      kdist[w] = cnt > outn ? entry_dist(spos, width, q, bpk[outn])
                            : ball_certificate(q, R, width);
    }
    __syncthreads();
  }
}

// (ub, when given: the ball bounds of k_redo_count.)
// Cells to redo get a longer list, sized from the cell's safety radius sr
// when its list ran out and the distance kd of the first entry left out
// (sr >= kd, else the cell would have stopped): the entries within sr are
// about L (sr / kd)^3 for a locally uniform density; the length grows by that
// estimate (+10 %), at least 2x and at most 4x per pass (16x from 4,096
// entries on: the few cells with such lists hold up each pass), and the next gather
// starts at the radius expected to hold it.
// An upper bound on the stream entries within r of q: the seeds of the grid
// cells the ball meets (periodic images included).  One wavefront, the rows
// of cells over the lanes; wave-uniform result.  The count stops once it
// reaches `stop` (the result is then >= stop).
__device__ __forceinline__ uint64_t ball_cells_count(const uint32_t *__restrict__ cell_start,
                                                     const GridParams &g, const double *width,
                                                     const double q[3], double r,
                                                     uint64_t stop) {
  const uint32_t lane = __lane_id();
  const double rr = r + g.margin, rr2 = rr * rr;
  uint64_t tot = 0;
  for (int i = g.img_lo[0]; i <= g.img_hi[0]; i++)
    for (int j = g.img_lo[1]; j <= g.img_hi[1]; j++)
      for (int k = g.img_lo[2]; k <= g.img_hi[2]; k++) {
        const int sv[3] = {i, j, k};
        double qs[3];
        int lo[3], hi[3];
        bool miss = false;
        for (int c = 0; c < 3; c++) {
          qs[c] = q[c] + (double)sv[c] * width[c];
          if (g.inv_h[c] == 0.) {
            lo[c] = hi[c] = 0;
            continue;
          }
          const double a = floor((qs[c] - rr - g.lo[c]) * g.inv_h[c]);
          const double b = floor((qs[c] + rr - g.lo[c]) * g.inv_h[c]);
          if (b < 0. || a > (double)(g.G[c] - 1)) miss = true;
          lo[c] = a < 0. ? 0 : (int)fmin(a, (double)(g.G[c] - 1));
          hi[c] = b > (double)(g.G[c] - 1) ? g.G[c] - 1 : (int)fmax(b, 0.);
        }
        if (miss) continue;
        const int nb_ = hi[1] - lo[1] + 1;
        const int nrows = (hi[0] - lo[0] + 1) * nb_;
        for (int r0 = 0; r0 < nrows; r0 += 64) {
          const int rw = r0 + (int)lane;
          uint32_t len = 0;
          if (rw < nrows) {
            const int a = lo[0] + rw / nb_, b = lo[1] + rw % nb_;
            double d2 = 0.;
            if (g.inv_h[0] != 0.) {
              const double x0 = g.lo[0] + (double)a * g.h[0];
              const double dx = fmax(0., fmax(x0 - qs[0], qs[0] - x0 - g.h[0]));
              d2 += dx * dx;
            }
            if (g.inv_h[1] != 0.) {
              const double y0 = g.lo[1] + (double)b * g.h[1];
              const double dy = fmax(0., fmax(y0 - qs[1], qs[1] - y0 - g.h[1]));
              d2 += dy * dy;
            }
            if (d2 <= rr2) {
              int c0 = lo[2], c1 = hi[2];
              if (g.inv_h[2] != 0.) {
                const double hc = sqrt(rr2 - d2);
                const double za = floor((qs[2] - hc - g.lo[2]) * g.inv_h[2]);
                const double zb = floor((qs[2] + hc - g.lo[2]) * g.inv_h[2]);
                c0 = max(c0, za < 0. ? 0 : (int)fmin(za, (double)g.G[2]));
                c1 = min(c1, zb < 0. ? -1 : (int)fmin(zb, (double)g.G[2]));
              }
              if (c0 <= c1) {
                const size_t row = ((size_t)a * g.G[1] + b) * g.G[2];
                len = cell_start[row + c1 + 1] - cell_start[row + c0];
              }
            }
          }
          uint32_t t;
          (void)wave_excl_scan(len, &t);
          tot += t;
          if (tot >= stop) return tot;
        }
      }
  return tot;
}

// redone lists from this length on may grow 16x per pass (4x below)
constexpr uint32_t GROW16_MIN = 4096;

// The length k_redo_size gives a redone list from its last length Lp, the
// cell's safety radius sr when the list ran out and the distance kd of the
// first entry left out.
__device__ __forceinline__ double redo_length(uint32_t Lp, double sr, double kd) {
  const double ratio = (kd > 0. && sr > kd && isfinite(sr)) ? sr / kd : 1.;
  const double est = 1.1 * (double)Lp * ratio * ratio * ratio;
  // long lists (few cells, each pass of them mostly latency) grow up to 16x
  const double lo = 2. * (double)Lp, hi = (Lp >= GROW16_MIN ? 16. : 4.) * (double)Lp;
  return fmin(fmax(est, lo), hi);
}

// Redone lists of at least `lmin` entries: ub[w] = the seeds of the grid
// cells within the safety radius sr[w] of the cell's generator (an upper
// bound on the entries the cell can need, since its safety radius only
// shrinks) when that is below the length redo_length gives, ~0u otherwise
// (and for the other items, and balls over more than 16,384 rows of cells).
// One wavefront per item; dn: the item count on the device (nwork bounds
// it).
__global__ void __launch_bounds__(256)
    k_redo_count(uint32_t nwork, const uint32_t *__restrict__ work,
                 const uint32_t *__restrict__ Lp, const double *__restrict__ sr,
                 const double *__restrict__ kd, const double4 *__restrict__ spos,
                 const uint32_t *__restrict__ cell_start_all, Problems PR, uint32_t lmin,
                 uint32_t *__restrict__ ub, const uint32_t *__restrict__ dn) {
  const uint32_t w = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (w >= nwork || (dn && w >= *dn)) return;
  const double r = sr[w];
  const uint32_t slot = __builtin_amdgcn_readfirstlane(work[w]);
  const uint32_t pr = __builtin_amdgcn_readfirstlane(problem_of(PR, slot));
  const GridParams g = PR.grids[pr];
  double rows = 1.;
  for (int c = 0; c < 2; c++)
    if (g.inv_h[c] != 0.) rows *= fmin(2. * r * g.inv_h[c] + 2., (double)g.G[c]);
  if (Lp[w] < lmin || !(r > 0.) || !isfinite(r) || rows > 16384.) {
    if (__lane_id() == 0) ub[w] = ~0u;
    return;
  }
  const uint64_t Ln = (uint64_t)redo_length(Lp[w], r, kd[w]);
  const BoxParams &box = PR.boxes[pr];
  const double width[3] = {box.width[0], box.width[1], box.width[2]};
  const double4 qq = spos[slot];
  const double q[3] = {qq.x, qq.y, qq.z};
  // (a hair over r: the gather takes the ball of radius r (1 + 1e-6))
  const uint64_t c =
      ball_cells_count(cell_start_all + PR.coff[pr], g, width, q, r * (1. + 2e-6), Ln);
  if (__lane_id() == 0) ub[w] = c < Ln ? (uint32_t)c : ~0u;
}

__global__ void k_redo_size(uint32_t nwork, uint32_t *__restrict__ Lio,
                            const double *__restrict__ sr, const double *__restrict__ kd,
                            float *__restrict__ Rout, uint32_t kmax,
                            unsigned int *__restrict__ cap_err,
                            const uint32_t *__restrict__ dn,
                            const uint32_t *__restrict__ ub = nullptr) {
  const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nwork || (dn && w >= *dn)) return;
  const uint32_t Lp = Lio[w];
  if (Lp >= kmax) atomicOr(cap_err, 1u);
  const double k = kd[w], r = sr[w];
  const uint32_t Lnew = (uint32_t)fmin(redo_length(Lp, r, k), (double)kmax);
  if (ub && ub[w] != ~0u && ub[w] > Lp && ub[w] < Lnew) {
    // fewer seeds within the safety radius than that: the ball of that
    // radius holds every entry the cell can need; the gather takes it whole
    // (a negative radius)
    Lio[w] = ub[w];
    Rout[w] = -(float)(r * (1. + 1e-6));
    return;
  }
  Lio[w] = Lnew;
  Rout[w] = (k > 0. && isfinite(k)) ? (float)(k * cbrt((double)Lnew / (double)Lp)) : 0.f;
}

// Redone main lists longer than `lmax` go to the end of the list, so that
// the wavefront list kernels gather the others and the long-list kernel
// only these.  flag[i] = L[i] > lmax; k_move_long sends flagged items to
// position pos[i] of (lw, lL, lR) and the others, in order, to (sw, sL, sR).
__global__ void k_flag_long(const uint32_t *__restrict__ L, uint32_t n, uint32_t lmax,
                            uint32_t *__restrict__ flag) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flag[i] = L[i] > lmax ? 1u : 0u;
}
__global__ void k_move_long(const uint32_t *__restrict__ w, const uint32_t *__restrict__ L,
                            const float *__restrict__ R, const uint32_t *__restrict__ flag,
                            const uint32_t *__restrict__ pos, uint32_t n,
                            uint32_t *__restrict__ lw, uint32_t *__restrict__ lL,
                            float *__restrict__ lR, uint32_t *__restrict__ sw,
                            uint32_t *__restrict__ sL, float *__restrict__ sR) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (flag[i]) {
    const uint32_t p = pos[i];
    lw[p] = w[i];
    lL[p] = L[i];
    lR[p] = R[i];
  } else {
    const uint32_t p = i - pos[i];
    sw[p] = w[i];
    sL[p] = L[i];
    sR[p] = R[i];
  }
}

__global__ void k_compact_redo(const uint32_t *__restrict__ in_w,
                               const uint32_t *__restrict__ in_L,
                               const double *__restrict__ in_r,
                               const double *__restrict__ in_kd,
                               const uint32_t *__restrict__ flag,
                               const uint32_t *__restrict__ pos, uint32_t n,
                               uint32_t *__restrict__ out_w, uint32_t *__restrict__ out_L,
                               double *__restrict__ out_r, double *__restrict__ out_kd) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && flag[i]) {
    out_w[pos[i]] = in_w[i];
    out_L[pos[i]] = in_L[i];
    out_r[pos[i]] = in_r[i];
    out_kd[pos[i]] = in_kd[i];
  }
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

// work lists carry (slot, list length) pairs
__global__ void k_fill_u32(uint32_t *__restrict__ a, uint32_t n, uint32_t v) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}
__global__ void k_compact_pairs(const uint32_t *__restrict__ in_w,
                                const uint32_t *__restrict__ in_L,
                                const uint32_t *__restrict__ flag,
                                const uint32_t *__restrict__ pos, uint32_t n,
                                uint32_t mult, uint32_t cap,
                                uint32_t *__restrict__ out_w,
                                uint32_t *__restrict__ out_L) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && flag[i]) {
    out_w[pos[i]] = in_w[i];
    uint64_t L = (uint64_t)in_L[i] * mult;
    out_L[pos[i]] = (uint32_t)(L > cap ? cap : L);
  }
}
// counting sort of a work list by list length, longest first (the
// persistent kernel then starts the expensive cells early); bucket = L / 8.
// Each block counts LEN_PER_BLOCK items in LDS and touches the global
// counters once per bucket, so the global atomics do not serialise.
constexpr uint32_t LEN_PER_BLOCK = 4096;
constexpr uint32_t LEN_MAX_BUCKETS = 1024;
__global__ void __launch_bounds__(256)
    k_len_hist(const uint32_t *__restrict__ L, uint32_t n, uint32_t nb,
               uint32_t *__restrict__ hist) {
  __shared__ uint32_t h[LEN_MAX_BUCKETS];
  for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x) h[b] = 0;
  __syncthreads();
  const uint32_t i0 = blockIdx.x * LEN_PER_BLOCK;
  const uint32_t i1 = min(n, i0 + LEN_PER_BLOCK);
  for (uint32_t i = i0 + threadIdx.x; i < i1; i += blockDim.x)
    atomicAdd(&h[nb - 1 - min(L[i] / 8, nb - 1)], 1u);
  __syncthreads();
  for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x)
    if (h[b]) atomicAdd(&hist[b], h[b]);
}
__global__ void __launch_bounds__(256)
    k_len_scatter(const uint32_t *__restrict__ w, const uint32_t *__restrict__ L,
                  const float *__restrict__ R, uint32_t n, uint32_t nb,
                  const uint32_t *__restrict__ start, uint32_t *__restrict__ fill,
                  uint32_t *__restrict__ ow, uint32_t *__restrict__ oL,
                  float *__restrict__ oR) {
  __shared__ uint32_t h[LEN_MAX_BUCKETS];
  constexpr uint32_t PER_THREAD = LEN_PER_BLOCK / 256;
  for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x) h[b] = 0;
  __syncthreads();
  const uint32_t i0 = blockIdx.x * LEN_PER_BLOCK;
  uint32_t bucket[PER_THREAD], rank[PER_THREAD];
#pragma unroll
  for (uint32_t k = 0; k < PER_THREAD; k++) {
    const uint32_t i = i0 + k * 256 + threadIdx.x;
    bucket[k] = 0xFFFFFFFFu;
    if (i < n) {
      bucket[k] = nb - 1 - min(L[i] / 8, nb - 1);
      rank[k] = atomicAdd(&h[bucket[k]], 1u);
    }
  }
  __syncthreads();
  // reserve this block's share of each bucket
  for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x)
    h[b] = h[b] ? start[b] + atomicAdd(&fill[b], h[b]) : 0u;
  __syncthreads();
#pragma unroll
  for (uint32_t k = 0; k < PER_THREAD; k++) {
    if (bucket[k] == 0xFFFFFFFFu) continue;
    const uint32_t i = i0 + k * 256 + threadIdx.x;
    const uint32_t p = h[bucket[k]] + rank[k];
    ow[p] = w[i];
    oL[p] = L[i];
    if (R) oR[p] = R[i];
  }
}
// values by generator index from values by slot
__global__ void k_unpermute(const uint32_t *__restrict__ order, size_t n,
                            int width, const float *__restrict__ in,
                            float *__restrict__ out) {
  size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  size_t i = order[s];
  for (int c = 0; c < width; c++) out[i * width + c] = in[s * width + c];
}

}  // namespace mv
