// mv_clip.hpp -- per-cell construction: convex-cell clipping by the
// neighbour half-spaces with the reference's security-radius stop, then the
// decomposition into oriented tets with the volume/centroid and per-face
// area/centroid integrals.
//
// One cell per lane.  A lane runs the reference's sequential algorithm
// (src/voronoi/convex_cell.rs:294-453, 131-187; voronoi_cell.rs:40-100)
// step for step, so its vertex order, face order and every float
// operation match the reference's; the cell's plane, vertex and cycle
// arrays live in device memory interleaved across lanes (element j of lane
// l at [j * lanes + l]), which keeps the accesses of a wavefront coalesced
// while its lanes walk their arrays in step.
#pragma once
#include "mv_common.hpp"

namespace mv {

struct ClipArgs {
  const double4 *spos;   // sorted generators, w = generator index bits
  const uint8_t *mask;   // by generator index, nullptr = all active
  Problems PR;           // boxes, grids (total stream) per problem
  const uint32_t *work;  // slots to build
  uint32_t nwork;
  const unsigned int *nsub;  // the item count on the device (nwork bounds it)
  const uint64_t *kpk;   // neighbour lists, item-major
  const uint64_t *kofs;  // list offset per item (nullptr: w * L)
  const uint32_t *kcnt;
  const double *kdist;   // distance of the first entry left out (< 0: none)
  double *need_r;        // safety radius of cells that need more entries
  uint32_t L;
  // lane state
  double4 *plA;   // plane: n.xyz, d
  double4 *plB;   // plane: p.xyz, reference word (packed entry, ~0 = box)
  double4 *vtx;   // vertex: loc.xyz, radius2
  uint32_t *vdual;
  uint8_t *cyc;   // SimpleCycle pointers, reused as face flags
  double4 *facc;  // face accumulators: area, centroid sums
  uint32_t lanes;
  // outputs
  uint8_t *wstate;       // CellState per work item
  double *cell_vol;      // by generator index
  double *cell_cen;
  double *cell_sr;
  uint32_t *cell_nf;
  uint64_t *cell_fbase;
  uint32_t *cell_npop;
  uint32_t *link_in;     // faces linked to each cell from a lower one
  // face staging
  uint32_t *st_cell;
  int64_t *st_right;
  uint8_t *st_img;
  uint32_t *st_lslot;    // a linked face's slot among its right cell's links
  double *st_area;
  double *st_cen;
  double *st_norm;
  uint64_t st_cap;
  unsigned long long *st_counter;
  unsigned long long *err_word;  // min over (idx << 8 | code)
  unsigned long long *exact_counter;
  unsigned long long *overflow;  // bit 0: staging area full, bit 1: cell store full
  CellStore cs;                  // kept cells (cs.pl_a == nullptr: none)
  // capacities of the large-cell kernel below its arrays' (0: theirs)
  uint32_t coop_maxp, coop_maxv;
  // the wide kernel: one slab per block of wmaxp planes and wmaxv vertices
  double4 *wplA, *wplB, *wfacc, *wvtx;
  uint64_t *wvdual;
  uint32_t *wcyc;
  uint8_t *wflag;
  uint32_t wmaxp, wmaxv;
};

constexpr uint64_t BOX_REF = ~0ull;

enum ErrCode : uint32_t {
  E_DEGENERATE = 2,
  E_SELF_NOT_FIRST = 3,
  E_DET_ZERO = 4,
  E_NAN = 5,
  E_WALK = 6
};

// A cell's arrays.  CycT holds SimpleCycle pointers (plane indices) and
// DualT a vertex's three plane indices: 8 bits each (up to 255 planes) for
// the kernels with fixed capacity, 21 bits each (up to 2^21 - 1 planes) for
// the wide kernel, whose capacities are set at launch (maxp, maxv) and
// whose per-vertex clip decisions go to `flag` instead of a register mask.
template <typename CycT, typename DualT>
struct LaneT {
  using Cyc = CycT;
  using Dual = DualT;
  static constexpr bool WIDE = sizeof(DualT) == 8;
  double4 *plA, *plB, *vtx, *facc;
  DualT *vdual;
  CycT *cyc;
  uint8_t *flag;
  size_t lane, lanes;
  int maxp, maxv;
  __device__ __forceinline__ size_t at(int j) const {
    return (size_t)j * lanes + lane;
  }
};
using Lane = LaneT<uint8_t, uint32_t>;
using LaneWide = LaneT<uint32_t, uint64_t>;

__device__ __forceinline__ uint32_t pack_dual(int a, int b, int c) {
  return (uint32_t)a | ((uint32_t)b << 8) | ((uint32_t)c << 16);
}
__device__ __forceinline__ int dual_at(uint32_t d, int k) {
  return (int)((d >> (8 * k)) & 0xFFu);
}
__device__ __forceinline__ uint64_t pack_dual_wide(int a, int b, int c) {
  return (uint64_t)a | ((uint64_t)b << 21) | ((uint64_t)c << 42);
}
__device__ __forceinline__ int dual_at(uint64_t d, int k) {
  return (int)((d >> (21 * k)) & 0x1FFFFFull);
}
template <class LS>
__device__ __forceinline__ typename LS::Dual pack_dual_t(int a, int b, int c) {
  if constexpr (LS::WIDE)
    return pack_dual_wide(a, b, c);
  else
    return pack_dual(a, b, c);
}

struct CellRegs {
  const BoxParams *B;    // the cell's problem
  uint64_t total_stream;
  d3 q;
  uint64_t idx;
  int np, nv;
  int cstart, clen;  // SimpleCycle start and len
  double sr;         // safety radius
  uint32_t nexact;   // exact predicate evaluations
};

// Vertex::from_dual (src/voronoi/convex_cell.rs:28-48)
template <class LS>
__device__ __forceinline__ bool make_vertex(const LS &S, const CellRegs &c,
                                            int dim, int i, int j, int k,
                                            double4 *out) {
  double4 a = S.plA[S.at(i)], b = S.plA[S.at(j)], e = S.plA[S.at(k)];
  d3 loc;
  if (!intersect_planes(xyz(a), a.w, xyz(b), b.w, xyz(e), e.w, &loc))
    return false;
  d3 dl = loc;
  if (dim == 1) {
    dl.y = 0.;
    dl.z = 0.;
  } else if (dim == 2) {
    dl.z = 0.;
  }
  d3 df = sub3(c.q, dl);
  *out = make_double4(loc.x, loc.y, loc.z, dot3(df, df));
  return true;
}

// update_safety_radius (src/voronoi/convex_cell.rs:443-453)
template <class LS>
__device__ __forceinline__ bool update_safety_radius(const LS &S, CellRegs &c) {
  double m = -INFINITY;
  for (int v = 0; v < c.nv; v++) {
    double r = S.vtx[S.at(v)].w;
    if (isnan(r)) return false;
    m = fmax(m, r);
  }
  c.sr = 2. * sqrt(m);
  return true;
}

// HalfSpace::right_loc (src/voronoi/half_space.rs:53-65) of stored plane j
template <class LS>
__device__ __forceinline__ d3 right_loc(const LS &S, const CellRegs &c,
                                        const ClipArgs &A, int j) {
  double4 b = S.plB[S.at(j)];
  uint64_t ref = (uint64_t)__double_as_longlong(b.w);
  if (ref == BOX_REF) {
    double4 a = S.plA[S.at(j)];
    d3 pr = plane_project_onto(xyz(a), xyz(b), c.q);
    return sub3(scl3(2., pr), c.q);
  }
  double4 g = A.spos[pk_slot(ref)];
  d3 l = xyz(g);
  d3 s;
  if (image_shift(pk_img(ref), c.B->width, &s)) l = add3(l, s);
  return l;
}

// SimpleCycle (src/simple_cycle.rs:1-101) on the lane's pointer array
template <class LS>
__device__ __forceinline__ void cyc_init(const LS &S, CellRegs &c, int a, int b, int e) {
  using T = typename LS::Cyc;
  int cur = c.cstart;
  for (int i = 0; i < c.clen; i++) {
    int nx = S.cyc[S.at(cur)];
    S.cyc[S.at(cur)] = (T)cur;
    cur = nx;
  }
  c.clen = 3;
  c.cstart = a;
  S.cyc[S.at(a)] = (T)b;
  S.cyc[S.at(b)] = (T)e;
  S.cyc[S.at(e)] = (T)a;
}

template <class LS>
__device__ __forceinline__ bool cyc_try_extend(const LS &S, CellRegs &c, int a, int b, int e) {
  using T = typename LS::Cyc;
  int tri[3] = {a, b, e};
  int pa = S.cyc[S.at(a)], pb = S.cyc[S.at(b)], pe = S.cyc[S.at(e)];
  int ptr[3] = {pa, pb, pe};
  bool con[3] = {pa != a, pb != b, pe != e};
  for (int i = 0; i < 3; i++) {
    int j = i == 2 ? 0 : i + 1, k = i == 0 ? 2 : i - 1;
    if (!con[i] && con[j] && con[k] && ptr[k] == tri[j]) {
      S.cyc[S.at(tri[k])] = (T)tri[i];
      S.cyc[S.at(tri[i])] = (T)tri[j];
      c.clen += 1;
      return true;
    } else if (con[i] && con[j] && con[k] && ptr[k] == tri[j] &&
               ptr[j] == tri[i]) {
      S.cyc[S.at(tri[k])] = (T)tri[i];
      S.cyc[S.at(tri[j])] = (T)tri[j];
      if (c.cstart == tri[j]) c.cstart = tri[i];
      c.clen -= 1;
      return true;
    }
  }
  return false;
}

// Maximum over the wavefront (every lane gets it)
__device__ __forceinline__ double wave_max_all(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// ConvexCell::clip_by_plane (src/voronoi/convex_cell.rs:358-419).
// Returns 0, E_* for a reference panic, or 255 for capacity overflow.
// COOP: the wavefront builds one cell; its lanes classify the vertices 64 at
// a time and every lane then runs the sequential steps on the same data.
template <int MAXP, int MAXV, bool COOP = false, class LS = Lane>
__device__ uint32_t clip_by_plane(const LS &S, CellRegs &c,
                                  const ClipArgs &A, d3 n, d3 p, double d,
                                  double errb, uint64_t ref, d3 ngb) {
  // Every vertex's decision depends on that vertex alone, so all of them
  // are classified first (independent loads, no stores); the reference's
  // swap-with-last scan is then replayed on the resulting bit mask (the
  // wide kernel: on per-vertex flags in memory), which moves the same
  // vertices to the same positions.
  constexpr bool WIDE = LS::WIDE;
  static_assert(WIDE || MAXV <= 1024, "mask width");
  constexpr int MW = WIDE ? 1 : (MAXV + 63) / 64;
  uint64_t m[MW];
#pragma unroll
  for (int k = 0; k < MW; k++) m[k] = 0;
  uint64_t anyw = 0;  // WIDE: any vertex clipped
  double keepmax = -INFINITY;
  bool nan_keep = false;
  if constexpr (COOP) {
    const int lane = (int)__lane_id();
    for (int i0 = 0; i0 < c.nv; i0 += 64) {
      const int i = i0 + lane;
      bool clipped = false, kept = false;
      double r2 = 0.;
      if (i < c.nv) {
        const double4 v = S.vtx[S.at(i)];
        const double cv = dot3(n, xyz(v)) - d;
        if (fabs(cv) < errb) {
          const auto du = S.vdual[S.at(i)];
          int64_t ia[3], ib[3], ic[3], id[3], iv[3];
          iloc(*c.B, c.q, ia);
          iloc(*c.B, right_loc(S, c, A, dual_at(du, 0)), ib);
          iloc(*c.B, right_loc(S, c, A, dual_at(du, 1)), ic);
          iloc(*c.B, right_loc(S, c, A, dual_at(du, 2)), id);
          iloc(*c.B, ngb, iv);
          clipped = in_sphere_exact(ia, ib, ic, id, iv) < 0;
          c.nexact++;
        } else {
          clipped = cv < 0.;
        }
        kept = !clipped;
        r2 = v.w;
        if constexpr (WIDE) S.flag[S.at(i)] = clipped ? 1 : 0;
      }
      if constexpr (WIDE)
        anyw |= wave_ballot(clipped);
      else
        m[i0 >> 6] = wave_ballot(clipped);
      nan_keep |= wave_ballot(kept && isnan(r2)) != 0ull;
      // kept radii are >= 0; fmax drops NaN as the sequential loop does
      keepmax = fmax(keepmax, wave_max_all(kept && !isnan(r2) ? r2 : -INFINITY));
    }
  }
  for (int i = 0; !COOP && i < c.nv; i++) {
    double4 v = S.vtx[S.at(i)];
    double cv = dot3(n, xyz(v)) - d;
    bool clipped;
    if (fabs(cv) < errb) {
      // in-sphere test on snapped coordinates (convex_cell.rs:370-382)
      const auto du = S.vdual[S.at(i)];
      int64_t ia[3], ib[3], ic[3], id[3], iv[3];
      iloc(*c.B, c.q, ia);
      iloc(*c.B, right_loc(S, c, A, dual_at(du, 0)), ib);
      iloc(*c.B, right_loc(S, c, A, dual_at(du, 1)), ic);
      iloc(*c.B, right_loc(S, c, A, dual_at(du, 2)), id);
      iloc(*c.B, ngb, iv);
      clipped = in_sphere_exact(ia, ib, ic, id, iv) < 0;
      c.nexact++;
    } else {
      clipped = cv < 0.;  // NaN keeps the vertex, as in the reference
    }
    if (clipped) {
      if constexpr (WIDE) {
        S.flag[S.at(i)] = 1;
        anyw = 1;
      } else {
        m[i >> 6] |= 1ull << (i & 63);
      }
    } else {
      if constexpr (WIDE) S.flag[S.at(i)] = 0;
      nan_keep |= isnan(v.w);
      keepmax = fmax(keepmax, v.w);
    }
  }
  if constexpr (WIDE && COOP) {
    // every lane reads the flags the other lanes wrote (one workgroup of
    // one wavefront: a workgroup fence makes them visible)
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  }
  uint64_t any = anyw;
#pragma unroll
  for (int k = 0; k < MW; k++) any |= m[k];
  if (!any) return 0;
  if (nan_keep) return E_NAN;
  // replay of the swap-with-last scan (convex_cell.rs:365-390)
  auto isclip = [&](int j) -> bool {
    if constexpr (WIDE)
      return S.flag[S.at(j)] != 0;
    else
      return (m[j >> 6] >> (j & 63)) & 1ull;
  };
  auto setclip = [&](int j, bool v) {
    if constexpr (WIDE) {
      S.flag[S.at(j)] = v ? 1 : 0;
    } else {
      if (v)
        m[j >> 6] |= 1ull << (j & 63);
      else
        m[j >> 6] &= ~(1ull << (j & 63));
    }
  };
  int i = 0, num_v = c.nv;
  while (i < num_v) {
    bool bi = isclip(i);
    if (bi) {
      num_v -= 1;
      bool bn = isclip(num_v);
      if (num_v != i) {
        // position i takes the last element; the clipped vertex's data is
        // dead apart from its dual (used by the boundary walk)
        S.vtx[S.at(i)] = S.vtx[S.at(num_v)];
        const auto di = S.vdual[S.at(i)];
        S.vdual[S.at(i)] = S.vdual[S.at(num_v)];
        S.vdual[S.at(num_v)] = di;
        if (!bn) {
          setclip(i, false);
          setclip(num_v, true);
        }
      }
    } else {
      i += 1;
    }
  }
  // push the plane and grow the cycle
  if (c.np >= S.maxp) return 255;
  const int p_idx = c.np;
  S.plA[S.at(p_idx)] = make_double4(n.x, n.y, n.z, d);
  S.plB[S.at(p_idx)] = make_double4(p.x, p.y, p.z, __longlong_as_double((long long)ref));
  S.cyc[S.at(p_idx)] = (typename LS::Cyc)p_idx;
  c.np += 1;
  // compute_boundary (convex_cell.rs:421-441) over the removed vertices
  {
    const auto d0 = S.vdual[S.at(num_v)];
    cyc_init(S, c, dual_at(d0, 0), dual_at(d0, 1), dual_at(d0, 2));
    for (int r = num_v + 1; r < c.nv; r++) {
      int idx = r;
      for (;;) {
        if (idx >= c.nv) return E_WALK;
        const auto du = S.vdual[S.at(idx)];
        if (cyc_try_extend(S, c, dual_at(du, 0), dual_at(du, 1),
                           dual_at(du, 2))) {
          if (idx > r) {
            S.vdual[S.at(idx)] = S.vdual[S.at(r)];
            S.vdual[S.at(r)] = du;
          }
          break;
        }
        idx++;
      }
    }
  }
  c.nv = num_v;
  if (c.nv + c.clen > S.maxv) return 255;
  int cur = c.cstart;
  double newmax = keepmax;
  for (int t = 0; t < c.clen; t++) {
    int next = S.cyc[S.at(cur)];
    double4 nvtx;
    if (!make_vertex(S, c, c.B->dim, cur, next, p_idx, &nvtx))
      return E_DET_ZERO;
    if (isnan(nvtx.w)) return E_NAN;
    newmax = fmax(newmax, nvtx.w);
    S.vtx[S.at(c.nv)] = nvtx;
    S.vdual[S.at(c.nv)] = pack_dual_t<LS>(cur, next, p_idx);
    c.nv++;
    cur = next;
  }
  // update_safety_radius (convex_cell.rs:443-453): the max over the kept
  // and the new vertices, tracked above
  c.sr = 2. * sqrt(newmax);
  return 0;
}

// should a face be built for plane j (voronoi_cell.rs:53-73)
template <class LS>
__device__ __forceinline__ bool should_construct(const LS &S, const CellRegs &c,
                                                 const ClipArgs &A, int j) {
  double4 a = S.plA[S.at(j)];
  int dim = c.B->dim;
  if (dim == 1 && !(a.y == 0. && a.z == 0.)) return false;
  if (dim == 2 && !(a.z == 0.)) return false;
  uint64_t ref = (uint64_t)__double_as_longlong(S.plB[S.at(j)].w);
  if (ref != BOX_REF && pk_img(ref) == 13) {
    uint32_t right = pk_idx(ref);
    return (uint64_t)right > c.idx || (A.mask != nullptr && !A.mask[right]);
  }
  return true;
}

// COOP: the first entry from j on that the sequential loop of build_cell
// must run -- one whose distance is not finite and positive, that stops
// the cell (sr < dist), or whose plane has a vertex near it or inside
// (the only planes clip_by_plane acts on).  The entries before it leave the
// cell as it is, so skipping them is exact.  Lane l tests entry j + l
// against every vertex (the plane as build_cell computes it, the vertex
// test as clip_by_plane's), 64 entries at a time.  Returns cnt if none.
template <class LS>
__device__ __forceinline__ uint32_t coop_scan(const LS &S, const CellRegs &c,
                                              const ClipArgs &A, const uint64_t *list,
                                              uint32_t j, uint32_t cnt) {
  const BoxParams &B = *c.B;
  const uint32_t lane = __lane_id();
  for (; j < cnt; j += 64) {
    const uint32_t je = j + lane;
    bool hit = false;
    if (je < cnt) {
      const uint64_t pk = list[je];
      d3 ngb = xyz(A.spos[pk_slot(pk)]);
      d3 s;
      if (image_shift(pk_img(pk), B.width, &s)) ngb = add3(ngb, s);
      const d3 dx = sub3(c.q, ngb);
      const double dist = sqrt(dot3(dx, dx));
      if (!(isfinite(dist) && dist > 0.0) || c.sr < dist) {
        hit = true;
      } else {
        const d3 n = div3(dx, dist);
        const d3 p = scl3(0.5, add3(c.q, ngb));
        const double errb = 1e-13 * (1. + dot3(abs3(n), abs3(p)));
        const double d = dot3(n, p);
        for (int v = 0; v < c.nv; v++) {
          const double cv = dot3(n, xyz(S.vtx[S.at(v)])) - d;
          if (fabs(cv) < errb || cv < 0.) {
            hit = true;
            break;
          }
        }
      }
    }
    const uint64_t m = wave_ballot(hit);
    if (m) return j + (uint32_t)__builtin_ctzll(m);
  }
  return cnt;
}

// Runs one cell.  Returns its CellState; on CS_DONE, *nf = faces to emit.
template <int MAXP, int MAXV, bool COOP = false, class LS = Lane>
__device__ uint8_t build_cell(const LS &S, CellRegs &c, const ClipArgs &A,
                              uint32_t w, uint32_t *nf, uint32_t *npop,
                              uint32_t *err) {
  using T = typename LS::Cyc;
  const BoxParams &B = *c.B;
  *nf = 0;
  *npop = 0;
  // ConvexCell::init (convex_cell.rs:294-312)
  for (int k = 0; k < 6; k++) {
    S.plA[S.at(k)] = make_double4(B.pn[k][0], B.pn[k][1], B.pn[k][2], B.pd[k]);
    S.plB[S.at(k)] = make_double4(B.pp[k][0], B.pp[k][1], B.pp[k][2],
                                  __longlong_as_double((long long)BOX_REF));
    S.cyc[S.at(k)] = (T)k;
  }
  c.np = 6;
  c.cstart = 0;
  c.clen = 0;
  c.nexact = 0;
  const uint8_t D[8][3] = {{2, 5, 0}, {5, 3, 0}, {1, 5, 2}, {5, 1, 3},
                           {4, 2, 0}, {4, 0, 3}, {2, 4, 1}, {4, 3, 1}};
  c.nv = 8;
  for (int v = 0; v < 8; v++) {
    double4 nv;
    if (!make_vertex(S, c, B.dim, D[v][0], D[v][1], D[v][2], &nv)) {
      *err = E_DET_ZERO;
      return CS_ERROR;
    }
    S.vtx[S.at(v)] = nv;
    S.vdual[S.at(v)] = pack_dual_t<LS>(D[v][0], D[v][1], D[v][2]);
  }
  if (!update_safety_radius(S, c)) {
    *err = E_NAN;
    return CS_ERROR;
  }
  // ConvexCell::build (convex_cell.rs:316-356)
  const uint32_t cnt = A.kcnt[w];
  if (cnt == 0xFFFFFFFFu) return CS_CAPACITY;  // list did not fit the gather
  const uint64_t *list = A.kpk + (A.kofs ? A.kofs[w] : (uint64_t)w * A.L);
  uint64_t first = list[0];
  if (cnt == 0 || (uint64_t)pk_idx(first) != c.idx) {
    *err = E_SELF_NOT_FIRST;
    return CS_ERROR;
  }
  bool stopped = false;
  uint32_t j = 1;
  for (; j < cnt; j++) {
    // past the first entries (which clip often), a large cell looks for
    // its next event 64 entries at a time
    if (COOP && j >= 32) {
      j = coop_scan(S, c, A, list, j, cnt);
      if (j >= cnt) break;
    }
    uint64_t pk = list[j];
    double4 g = A.spos[pk_slot(pk)];
    d3 ngb = xyz(g);
    d3 s;
    if (image_shift(pk_img(pk), B.width, &s)) ngb = add3(ngb, s);
    d3 dx = sub3(c.q, ngb);
    double dist = sqrt(dot3(dx, dx));
    if (!(isfinite(dist) && dist > 0.0)) {
      *err = E_DEGENERATE;
      return CS_ERROR;
    }
    if (c.sr < dist) {
      stopped = true;
      break;
    }
    d3 n = div3(dx, dist);
    d3 p = scl3(0.5, add3(c.q, ngb));
    // HalfSpace::new (half_space.rs:24-34)
    double errb = 1e-13 * (1. + dot3(abs3(n), abs3(p)));
    double d = dot3(n, p);
    uint32_t r = clip_by_plane<MAXP, MAXV, COOP, LS>(S, c, A, n, p, d, errb, pk, ngb);
    if (r == 255) return CS_CAPACITY;
    if (r) {
      *err = r;
      return CS_ERROR;
    }
  }
  if (!stopped && (uint64_t)cnt < c.total_stream) {
    // list exhausted: the stop test on the first entry left out
    if (!(c.sr < A.kdist[w])) return CS_NEED_MORE;
    stopped = true;
    j = cnt;
  }
  *npop = stopped ? j + 1 : cnt;

  // decomposition and integrals (convex_cell.rs:131-187,
  // voronoi_cell.rs:77-99, integrals.rs:91-105, voronoi_face.rs:30-43)
  for (int k = 0; k < c.np; k++) {
    S.cyc[S.at(k)] = should_construct(S, c, A, k) ? 1 : 0;
    S.facc[S.at(k)] = make_double4(0., 0., 0., 0.);
  }
  double vol = 0.;
  d3 cen = mk3(0, 0, 0);
  const d3 q = c.q;
  for (int vi = 0; vi < c.nv; vi++) {
    const auto du = S.vdual[S.at(vi)];
    const d3 vl = xyz(S.vtx[S.at(vi)]);
    // projections P_t of DecompositionWithoutFaces::load_vertex, streamed:
    // P_{2k} = q projected on plane dual[k], P_{2k+1} = q projected on the
    // line dual[k+1] ^ dual[k]; tet t = (P_t, P_{t-1 mod 6}, vertex).
    d3 prev;  // P_5 first
    {
      double4 a0 = S.plA[S.at(dual_at(du, 0))];
      double4 a2 = S.plA[S.at(dual_at(du, 2))];
      if (!project_onto_intersection(xyz(a0), a0.w, xyz(a2), a2.w, q, &prev)) {
        *err = E_DET_ZERO;
        return CS_ERROR;
      }
    }
    for (int k = 0; k < 3; k++) {
      const int pi = dual_at(du, k);
      const int pn = dual_at(du, k == 2 ? 0 : k + 1);
      const double4 ak = S.plA[S.at(pi)];
      const double4 bk = S.plB[S.at(pi)];
      const d3 pface = plane_project_onto(xyz(ak), xyz(bk), q);
      const double4 an = S.plA[S.at(pn)];
      d3 pedge;
      if (!project_onto_intersection(xyz(an), an.w, xyz(ak), ak.w, q, &pedge)) {
        *err = E_DET_ZERO;
        return CS_ERROR;
      }
      const bool face = S.cyc[S.at(pi)] & 1;
      double4 fa = make_double4(0., 0., 0., 0.);
      if (face) fa = S.facc[S.at(pi)];
      // tet 2k: (P_2k, P_2k-1, v); tet 2k+1: (P_2k+1, P_2k, v)
      for (int h = 0; h < 2; h++) {
        const d3 v0 = h == 0 ? pface : pedge;
        const d3 v1 = h == 0 ? prev : pface;
        double tv = signed_volume_tet(v0, v1, vl, q);
        vol += tv;
        cen = add3(cen, scl3(tv, add3(add3(add3(v0, v1), vl), q)));
        if (face) {
          double ar = signed_area_tri(v0, v1, vl, q);
          d3 sc = scl3(ar, add3(add3(v0, v1), vl));
          fa.x += ar;
          fa.y += sc.x;
          fa.z += sc.y;
          fa.w += sc.z;
        }
      }
      if (face) {
        S.facc[S.at(pi)] = fa;
        S.cyc[S.at(pi)] = 3;
      }
      prev = pedge;
    }
  }
  const size_t gi = c.idx;
  const double norm = vol > 0. ? 0.25 / vol : 0.;
  A.cell_vol[gi] = vol;
  A.cell_cen[3 * gi] = cen.x * norm;
  A.cell_cen[3 * gi + 1] = cen.y * norm;
  A.cell_cen[3 * gi + 2] = cen.z * norm;
  A.cell_sr[gi] = c.sr;
  uint32_t count = 0;
  for (int k = 0; k < c.np; k++) count += S.cyc[S.at(k)] == 3;
  *nf = count;
  return CS_DONE;
}

// Append a completed cell's half-spaces and vertices to the cell store
// (wave-aggregated cursors; every lane of the wavefront calls this).
template <class LS>
__device__ __forceinline__ void keep_cell(const LS &S, const ClipArgs &A,
                                          bool emit, const CellRegs &c) {
  static_assert(!LS::WIDE, "kept cells store 8-bit plane indices");
  const uint32_t lane = __lane_id();
  const uint32_t np = emit ? (uint32_t)c.np : 0u, nv = emit ? (uint32_t)c.nv : 0u;
  uint32_t tp, tv;
  const uint32_t ep = wave_excl_scan(np, &tp);
  const uint32_t ev = wave_excl_scan(nv, &tv);
  unsigned long long bp = 0, bv = 0;
  if (lane == 0 && tp) {
    bp = atomicAdd(A.cs.cur_pl, (unsigned long long)tp);
    bv = atomicAdd(A.cs.cur_vx, (unsigned long long)tv);
  }
  bp = __shfl(bp, 0, 64) + ep;
  bv = __shfl(bv, 0, 64) + ev;
  if (!emit) return;
  if (bp + np > A.cs.cap_pl || bv + nv > A.cs.cap_vx) {
    atomicOr(A.overflow, 2ull);
    return;
  }
  for (uint32_t k = 0; k < np; k++) {
    A.cs.pl_a[bp + k] = S.plA[S.at((int)k)];
    A.cs.pl_b[bp + k] = S.plB[S.at((int)k)];
  }
  for (uint32_t v = 0; v < nv; v++) {
    const double4 x = S.vtx[S.at((int)v)];
    A.cs.vx[bv + v] = make_double4(x.x, x.y, x.z,
                                   __longlong_as_double((long long)S.vdual[S.at((int)v)]));
  }
  A.cs.pbase[c.idx] = bp;
  A.cs.vbase[c.idx] = bv;
  A.cs.np[c.idx] = np;
  A.cs.nv[c.idx] = nv;
}

template <int MAXP, int MAXV, bool KEEP>
__global__ void __launch_bounds__(256) k_clip(ClipArgs A) {
  const size_t gl = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t lane = __lane_id();
  Lane S;
  S.plA = A.plA;
  S.plB = A.plB;
  S.vtx = A.vtx;
  S.facc = A.facc;
  S.vdual = A.vdual;
  S.cyc = A.cyc;
  S.flag = nullptr;
  S.maxp = MAXP;
  S.maxv = MAXV;
  S.lane = gl;
  S.lanes = A.lanes;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  unsigned long long my_exact = 0;
  // wave-uniform trip count: every lane of a wavefront runs the same
  // iterations, so the face-offset scan below sees a converged wavefront
  for (size_t w0 = gl - lane; w0 < A.nwork; w0 += stride) {
    const size_t w = w0 + lane;
    uint32_t nf = 0;
    bool emit = false;
    CellRegs c;
    if (w < A.nwork) {
      const uint32_t slot = A.work[w];
      const uint32_t pr = problem_of(A.PR, slot);
      c.B = &A.PR.boxes[pr];
      c.total_stream = A.PR.grids[pr].total_stream;
      const double4 qq = A.spos[slot];
      c.q = xyz(qq);
      c.idx = (uint64_t)__double_as_longlong(qq.w);
      uint32_t err = 0, npop = 0;
      uint8_t st = build_cell<MAXP, MAXV>(S, c, A, (uint32_t)w, &nf, &npop, &err);
      my_exact += c.nexact;
      A.wstate[w] = st;
      if (st == CS_NEED_MORE) A.need_r[w] = c.sr;
      if (st == CS_ERROR) {
        atomicMin(A.err_word, (unsigned long long)((c.idx << 8) | err));
      }
      if (st == CS_DONE) {
        emit = true;
        A.cell_npop[c.idx] = npop;
      } else {
        nf = 0;
      }
    }
    if (KEEP) keep_cell(S, A, emit, c);
    uint32_t total;
    uint32_t excl = wave_excl_scan(nf, &total);
    unsigned long long base = 0;
    if (lane == 0 && total) base = atomicAdd(A.st_counter, (unsigned long long)total);
    base = __shfl(base, 0, 64);
    if (emit) {
      const uint64_t fb = base + excl;
      A.cell_nf[c.idx] = nf;
      A.cell_fbase[c.idx] = fb;
      if (fb + nf > A.st_cap) {
        atomicOr(A.overflow, 1ull);
      } else {
        uint64_t o = fb;
        for (int k = 0; k < c.np; k++) {
          if (S.cyc[S.at(k)] != 3) continue;
          double4 a = S.plA[S.at(k)];
          uint64_t ref = (uint64_t)__double_as_longlong(S.plB[S.at(k)].w);
          double4 fa = S.facc[S.at(k)];
          double fnrm = fa.x > 0. ? 1. / (3. * fa.x) : 0.;
          A.st_cell[o] = (uint32_t)c.idx;
          A.st_right[o] = ref == BOX_REF ? -1 : (int64_t)pk_idx(ref);
          // a face without periodic shift is also a link of its right cell
          // (the count's old value is the link's slot in that cell's range)
          A.st_lslot[o] = ref != BOX_REF && pk_img(ref) == 13
                              ? atomicAdd(&A.link_in[pk_idx(ref)], 1u)
                              : 0u;
          A.st_img[o] = ref == BOX_REF ? 13 : (uint8_t)pk_img(ref);
          A.st_area[o] = fa.x;
          A.st_cen[3 * o] = fa.y * fnrm;
          A.st_cen[3 * o + 1] = fa.z * fnrm;
          A.st_cen[3 * o + 2] = fa.w * fnrm;
          A.st_norm[3 * o] = a.x;
          A.st_norm[3 * o + 1] = a.y;
          A.st_norm[3 * o + 2] = a.z;
          o++;
        }
      }
    }
  }
  if (my_exact) atomicAdd(A.exact_counter, my_exact);
}

// The arrays of one cell of k_clip_coop, in LDS (~60 KiB at 255 planes
// and 1024 vertices).
template <int MAXP, int MAXV>
struct CoopLds {
  double4 plA[MAXP], plB[MAXP], facc[MAXP], vtx[MAXV];
  uint32_t vdual[MAXV];
  uint8_t cyc[MAXP];
};

// One cell per wavefront (a block of one wavefront), for the cells over
// the wavefront kernel's 64 vertices or planes (long lists, large cells).
// Every lane runs the same sequential steps on the cell's arrays in LDS
// (the lanes store the same values); the vertex classification of each
// plane is split over the lanes.  Results, faces and counters are written
// by lane 0.
// One cell w on the wavefront (every lane runs the same steps; lane 0
// writes the results).
template <int MAXP, int MAXV, bool KEEP, class LS>
__device__ __forceinline__ void coop_cell(const LS &S, const ClipArgs &A, uint32_t w,
                                          unsigned long long &my_exact) {
  const uint32_t lane = __lane_id();
  uint32_t nf = 0;
  CellRegs c;
  const uint32_t slot = A.work[w];
  const uint32_t pr = problem_of(A.PR, slot);
  c.B = &A.PR.boxes[pr];
  c.total_stream = A.PR.grids[pr].total_stream;
  const double4 qq = A.spos[slot];
  c.q = xyz(qq);
  c.idx = (uint64_t)__double_as_longlong(qq.w);
  uint32_t err = 0, npop = 0;
  const uint8_t st = build_cell<MAXP, MAXV, true, LS>(S, c, A, w, &nf, &npop, &err);
  my_exact += c.nexact;
  const bool emit = st == CS_DONE;
  if (lane == 0) {
    A.wstate[w] = st;
    if (st == CS_NEED_MORE) A.need_r[w] = c.sr;
    if (st == CS_ERROR) atomicMin(A.err_word, (unsigned long long)((c.idx << 8) | err));
    if (emit) A.cell_npop[c.idx] = npop;
  }
  if constexpr (KEEP) keep_cell(S, A, emit && lane == 0, c);
  if (emit && lane == 0) {
    const unsigned long long fb = nf ? atomicAdd(A.st_counter, (unsigned long long)nf) : 0ull;
    A.cell_nf[c.idx] = nf;
    A.cell_fbase[c.idx] = fb;
    if (fb + nf > A.st_cap) {
      atomicOr(A.overflow, 1ull);
    } else {
      uint64_t o = fb;
      for (int k = 0; k < c.np; k++) {
        if (S.cyc[S.at(k)] != 3) continue;
        const double4 a = S.plA[S.at(k)];
        const uint64_t ref = (uint64_t)__double_as_longlong(S.plB[S.at(k)].w);
        const double4 fa = S.facc[S.at(k)];
        const double fnrm = fa.x > 0. ? 1. / (3. * fa.x) : 0.;
        A.st_cell[o] = (uint32_t)c.idx;
        A.st_right[o] = ref == BOX_REF ? -1 : (int64_t)pk_idx(ref);
        // a face without periodic shift is also a link of its right cell
        // (the count's old value is the link's slot in that cell's range)
        A.st_lslot[o] = ref != BOX_REF && pk_img(ref) == 13
                            ? atomicAdd(&A.link_in[pk_idx(ref)], 1u)
                            : 0u;
        A.st_img[o] = ref == BOX_REF ? 13 : (uint8_t)pk_img(ref);
        A.st_area[o] = fa.x;
        A.st_cen[3 * o] = fa.y * fnrm;
        A.st_cen[3 * o + 1] = fa.z * fnrm;
        A.st_cen[3 * o + 2] = fa.w * fnrm;
        A.st_norm[3 * o] = a.x;
        A.st_norm[3 * o + 1] = a.y;
        A.st_norm[3 * o + 2] = a.z;
        o++;
      }
    }
  }
}

// One cell per wavefront (a block of one wavefront), for the cells over
// the wavefront kernel's 64 vertices or planes (long lists, large cells).
// Every lane runs the same sequential steps on the cell's arrays in LDS
// (the lanes store the same values); the vertex classification of each
// plane is split over the lanes.  Results, faces and counters are written
// by lane 0.  Cells over A.coop_maxp planes or A.coop_maxv vertices (at
// most the arrays' MAXP, MAXV) report CS_CAPACITY for the wide kernel.
template <int MAXP, int MAXV, bool KEEP>
__global__ void __launch_bounds__(64) k_clip_coop(ClipArgs A) {
  __shared__ CoopLds<MAXP, MAXV> L;
  const uint32_t lane = __lane_id();
  Lane S;
  S.plA = L.plA;
  S.plB = L.plB;
  S.vtx = L.vtx;
  S.facc = L.facc;
  S.vdual = L.vdual;
  S.cyc = L.cyc;
  S.flag = nullptr;
  S.maxp = A.coop_maxp && A.coop_maxp < (uint32_t)MAXP ? (int)A.coop_maxp : MAXP;
  S.maxv = A.coop_maxv && A.coop_maxv < (uint32_t)MAXV ? (int)A.coop_maxv : MAXV;
  S.lane = 0;
  S.lanes = 1;
  unsigned long long my_exact = 0;
  const uint32_t nw = A.nsub ? min(*A.nsub, A.nwork) : A.nwork;
  for (uint32_t w = blockIdx.x; w < nw; w += gridDim.x)
    coop_cell<MAXP, MAXV, KEEP>(S, A, w, my_exact);
  // every lane counted every exact test of its vertices
  for (int o = 32; o > 0; o >>= 1) my_exact += __shfl_xor(my_exact, o, 64);
  if (lane == 0 && my_exact) atomicAdd(A.exact_counter, my_exact);
}

// The wide kernel: the large-cell kernel's steps on arrays in device
// memory, one slab per block (A.wmaxp planes, A.wmaxv vertices, set by the
// host and doubled when a cell reports CS_CAPACITY), with plane indices of
// 21 bits.  It takes the cells over the large-cell kernel's capacity, which
// the reference grows without bound (convex_cell.rs:394-418).
__global__ void __launch_bounds__(64) k_clip_wide(ClipArgs A) {
  const uint32_t lane = __lane_id();
  const size_t b = blockIdx.x;
  LaneWide S;
  S.plA = A.wplA + b * A.wmaxp;
  S.plB = A.wplB + b * A.wmaxp;
  S.facc = A.wfacc + b * A.wmaxp;
  S.cyc = A.wcyc + b * A.wmaxp;
  S.vtx = A.wvtx + b * A.wmaxv;
  S.vdual = A.wvdual + b * A.wmaxv;
  S.flag = A.wflag + b * A.wmaxv;
  S.maxp = (int)A.wmaxp;
  S.maxv = (int)A.wmaxv;
  S.lane = 0;
  S.lanes = 1;
  unsigned long long my_exact = 0;
  for (uint32_t w = blockIdx.x; w < A.nwork; w += gridDim.x)
    coop_cell<0, 0, false>(S, A, w, my_exact);
  for (int o = 32; o > 0; o >>= 1) my_exact += __shfl_xor(my_exact, o, 64);
  if (lane == 0 && my_exact) atomicAdd(A.exact_counter, my_exact);
}

}  // namespace mv
