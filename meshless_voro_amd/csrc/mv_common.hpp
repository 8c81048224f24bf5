// mv_common.hpp -- parameter blocks shared by the kernels of libmvoro.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mv_geom.hpp"

namespace mv {

// Per-cell outcome of one construction pass.
enum CellState : uint8_t {
  CS_DONE = 0,
  CS_NEED_MORE = 1,  // the security radius was not reached within the list
  CS_CAPACITY = 2,   // planes or vertices exceeded this kernel's capacity
  CS_ERROR = 3,      // reference would panic; code in the error word
  CS_EXACT = 4       // met a near-degenerate test; this kernel has no exact
                     // predicate (another one rebuilds the cell)
};

// The simulation box of SimulationBoundary::cuboid (src/voronoi/boundary.rs:
// 14-47): six inward half-spaces and the 52-bit snapping frame.
struct BoxParams {
  double pn[6][3];  // plane normals
  double pp[6][3];  // plane points
  double pd[6];     // n . p
  double A[3];      // snap anchor
  double iw[3];     // 1 / (3 width')
  double width[3];  // period of the neighbour images (user width)
  // ConvexCell::init's eight vertices (the box corners, one per dual of the
  // initial cell), computed once on the host by intersect_planes; every
  // cell of the problem starts from them.  corner_ok: no determinant zero.
  double corner[8][3];
  int corner_ok;
  int dim;
  int periodic;
};

// Uniform grid over the bounding box of the generators.
struct GridParams {
  double lo[3];
  double inv_h[3];  // cells per unit length (0 for a flat dimension)
  double h[3];      // cell extent
  int G[3];
  int img_lo[3], img_hi[3];  // neighbour-image range per axis
  uint64_t total_stream;     // N * number of images
  double margin;             // absolute slack for the grid distance bounds
};

// Neighbour-list entry packing: (idx << 32) | (slot << 5) | image.
// Ordering the packed value orders by (idx, image) because a generator idx
// has one slot.  Image code = (i+1)*9 + (j+1)*3 + (k+1) for the shift
// (i, j, k) * width of the query; 13 = no shift.
// Lanes of the wavefront whose predicate holds.  The builtin takes the
// condition itself, so the compare's lane mask is used as it is (HIP's
// __ballot(int) first turns the condition into an integer and back).
__device__ __forceinline__ uint64_t wave_ballot(bool p) {
  return __builtin_amdgcn_ballot_w64(p);
}

__device__ __forceinline__ uint32_t pk_slot(uint64_t pk) {
  return (uint32_t)(pk >> 5) & 0x7FFFFFFu;
}
__device__ __forceinline__ uint32_t pk_img(uint64_t pk) {
  return (uint32_t)pk & 31u;
}
__device__ __forceinline__ uint32_t pk_idx(uint64_t pk) {
  return (uint32_t)(pk >> 32);
}

// Shift reported for an image code (src/rtree_nn.rs:27-33): None for the
// central image, else -(image * width) per component.
__device__ __forceinline__ bool image_shift(uint32_t img, const double *w,
                                            d3 *s) {
  int i = (int)(img / 9) - 1, j = (int)((img / 3) % 3) - 1,
      k = (int)(img % 3) - 1;
  if (i == 0 && j == 0 && k == 0) {
    *s = mk3(0, 0, 0);
    return false;
  }
  *s = mk3(-((double)i * w[0]), -((double)j * w[1]), -((double)k * w[2]));
  return true;
}

// SimulationBoundary::iloc (src/voronoi/boundary.rs:49-64)
__device__ __forceinline__ void iloc(const BoxParams &b, d3 p, int64_t o[3]) {
  double t0 = 1. + (p.x - b.A[0]) * b.iw[0];
  double t1 = 1. + (p.y - b.A[1]) * b.iw[1];
  double t2 = 1. + (p.z - b.A[2]) * b.iw[2];
  const uint64_t m = 0xFFFFFFFFFFFFFull;
  o[0] = (int64_t)((uint64_t)__double_as_longlong(t0) & m);
  o[1] = (int64_t)((uint64_t)__double_as_longlong(t1) & m);
  o[2] = (int64_t)((uint64_t)__double_as_longlong(t2) & m);
}

__device__ __forceinline__ uint64_t lane_id() {
  return (uint64_t)__lane_id();
}

// exclusive prefix sum over the 64 lanes of a converged wavefront
__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t v,
                                                   uint32_t *total) {
  uint32_t lane = __lane_id();
  uint32_t x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint32_t y = __shfl_up(x, d, 64);
    if (lane >= (uint32_t)d) x += y;
  }
  *total = __shfl(x, 63, 64);
  return x - v;
}

// Per-problem tables of a build over many independent tessellations (one
// entry for a single build).  Slots and generator indices are global over
// the concatenated problems; prob_of_slot == nullptr means one problem.
struct Problems {
  const BoxParams *boxes;
  const GridParams *grids;
  const uint32_t *coff;          // first grid cell of each problem
  const uint32_t *prob_of_slot;  // problem of each sorted slot (or nullptr)
  // dense cells' sub-columns (SubCols; all null when there are none)
  const uint32_t *sub_base, *sub_start, *sub_slot, *sub_key;
  const double *sub_xy;
};
__device__ __forceinline__ uint32_t problem_of(const Problems &P, uint32_t slot) {
  return P.prob_of_slot ? P.prob_of_slot[slot] : 0u;
}

// Dense grid cells (more than SUB_MIN seeds, 3D grids) split into S x S
// sub-columns in x and y, S = sub_side(seeds), at boundaries that are
// quantiles of the cell's x and y: base[cell] is the first word of the
// cell's table in `start` (~0u for the other cells): S, the first of its
// boundaries in `xy` (S + 1 in x from -inf to +inf, then S + 1 in y), an
// unused word, then S*S + 1 positions into `slot` and `key`: the cell's
// slots by sub-column (x-major), each sub-column in the cell's own (zkey)
// order, and their zkeys.  A gather reads only the sub-columns that reach
// its ball, each cut to the ball's chord in z.
constexpr uint32_t SUB_MIN = 256;
constexpr uint32_t SUB_SMAX = 63;   // S + 1 boundaries fit one wavefront
constexpr uint32_t SUB_BINS = 4096; // histogram bins of the quantiles
constexpr uint32_t SUB_HEAD = 3;    // words before the positions
struct SubCols {
  const uint32_t *base, *start, *slot, *key;
  const double *xy;
};
__device__ __forceinline__ SubCols sub_cols(const Problems &P, uint32_t pr) {
  return SubCols{P.sub_base ? P.sub_base + P.coff[pr] : nullptr, P.sub_start, P.sub_slot,
                 P.sub_key, P.sub_xy};
}
// S for a cell of m seeds: the smallest S >= 2 with 16 S^2 >= m (about 16
// seeds per sub-column), at most SUB_SMAX
__device__ __forceinline__ uint32_t sub_side(uint32_t m) {
  uint32_t S = 2;
  while (S < SUB_SMAX && 16u * S * S < m) S++;
  return S;
}

// Kept cells (VoronoiIntegrator): every completed cell's half-spaces and
// vertices, appended to two arenas by atomic cursors and indexed by
// generator.  pl_a == nullptr: nothing is kept.
struct CellStore {
  double4 *pl_a;   // n.xyz, d
  double4 *pl_b;   // p.xyz, reference word
  double4 *vx;     // vertex loc.xyz, packed dual (bits) in w
  uint64_t cap_pl, cap_vx;
  unsigned long long *cur_pl, *cur_vx;
  uint64_t *pbase, *vbase;  // by generator index
  uint32_t *np, *nv;        // by generator index
};

}  // namespace mv
