// mv_clipw.hpp -- per-cell construction with one cell per wavefront.
//
// Same algorithm as mv_clip.hpp (the reference's ConvexCell::build,
// clip_by_plane, compute_boundary, update_safety_radius and decomposition:
// src/voronoi/convex_cell.rs:294-453, 131-187), laid out for CDNA:
//   - vertex slot s of the cell's vertex array lives in lane s (x, y, z,
//     radius2, dual in registers), so one plane test is one VALU pass;
//   - the cell's half-spaces live in LDS (n, d | p, reference word);
//   - the SimpleCycle pointer array lives in one register (lane p holds
//     ptrs[p]); the reference's sequential steps -- the swap-with-last scan,
//     the boundary walk, the new-vertex order -- run as wave-uniform scalar
//     code over ballots and lane reads, so the vertex array, the plane list
//     and every clip decision are the reference's;
//   - the neighbour planes are prepared 64 at a time, one per lane, and
//     consumed in stream order;
//   - waves pull cells from a device work queue (persistent kernel).
// Cells that need more than 64 vertices or 64 planes report CS_CAPACITY and
// are rebuilt by the thread-per-cell kernel with larger capacity.
//
// Integrals: each lane evaluates the two tets of one (vertex, plane) pair
// and stages them in LDS; every face and the cell then sum their tets in the
// reference's order (vertex order, then tet order), so volumes, centroids
// and face areas are the reference's bits.
#pragma once
#include "mv_clip.hpp"
#include "mv_common.hpp"

namespace mv {

constexpr int WCAP = 64;  // vertices and planes per cell on this path
// decomposition: lane 3 i + k takes plane k of the chunk's vertex i, so a
// chunk covers 14 vertices (42 lanes)
constexpr int WCH = 14;

// 8 KiB per wavefront: five blocks of four wavefronts fill a CU's 160 KiB.
// A plane's point p (the midpoint of the generator and its neighbour, or
// the box face's point) is not kept: the few steps that read it rebuild it
// from the reference word, bit for bit (w_plane_point).
struct WaveLds {
  double4 pa[WCAP];      // plane: n.xyz, d
  uint64_t pref[WCAP];   // plane: reference word
  union {
    // the 64 stream entries being consumed (lane t wrote entry t): the
    // reference word, plane n and d, neighbour distance and the plane's
    // rounding bound
    // and, while a clip's boundary is found, the directed edges of the
    // removed vertices' dual triangles (bit y of emat[x]: edge x -> y)
    struct {
      uint64_t ref[64];
      double4 nd[64];
      double2 de[64];
      uint64_t emat[64];
      uint64_t rmat[64];  // bit x of rmat[y]: edge x -> y
    } e;
    // decomposition: per (tet, vertex) of a chunk of WCH vertices
    struct {
      double4 fcf[6][WCH];  // area, area * (v0 + v1 + v2)
      double4 fcv[6][WCH];  // volume, volume * (v0 + v1 + v2 + gen)
    } t;
  } u;
};

// A cell handed from the two-cells-per-wavefront kernel (k_clip_pair) to
// this one: its half-spaces, its vertices (position, 2 |v - q|, packed
// dual) and the next list entry to read, exactly as they stood before that
// entry.  32 half-spaces and vertices at most.
constexpr int PCAP = 32;
struct Handover {
  uint32_t np, nv, jnext, pad;
  double4 pa[PCAP];     // n.xyz, d
  uint64_t pref[PCAP];  // reference word
  double4 v[PCAP];      // x, y, z, 2 |v - q|
  int32_t vd[PCAP];     // packed dual
};

struct WaveArgs {
  const double4 *spos;
  const uint8_t *mask;
  Problems PR;            // boxes, grids (total stream) per problem
  const uint32_t *work;
  uint32_t nwork;
  uint32_t w0;               // first item (items w0 .. w0 + nwork - 1 without sub)
  const uint32_t *sub;       // items to build (nullptr: all nwork) ...
  const unsigned int *nsub;  // ... and their count, read on the device
  uint32_t *exact_list;      // items that met a near-degenerate test ...
  unsigned int *exact_cnt;   // ... and their count (kernels without the
                             // exact predicate)
  const uint64_t *kpk;    // neighbour lists
  const uint64_t *kofs;   // list offset per work item (nullptr: w * L)
  const uint32_t *kcnt;
  const double *kdist;    // distance of the first entry left out (< 0: none)
  double *need_r;         // safety radius of cells that need more entries
  uint32_t L;
  unsigned int *queue;    // work-queue head
  uint8_t *wstate;
  double *cell_vol, *cell_cen, *cell_sr;
  uint32_t *cell_nf;
  uint64_t *cell_fbase;
  uint32_t *cell_npop;
  uint32_t *link_in;      // faces linked to each cell from a lower one
  uint32_t *st_cell;
  int64_t *st_right;
  uint8_t *st_img;
  uint32_t *st_lslot;     // a linked face's slot among its right cell's links
  double *st_area, *st_cen, *st_norm;
  uint64_t st_cap;
  unsigned long long *st_counter;
  unsigned long long *err_word;
  unsigned long long *exact_counter;
  unsigned long long *overflow;  // bit 0: staging area full, bit 1: cell store full
  CellStore cs;                  // kept cells (cs.pl_a == nullptr: none)
  // cells the pair kernel hands over (k_clip_pair writes them; a launch of
  // k_clip_wave with resume != nullptr continues item i of sub from
  // resume[i]); ho_cap bounds both
  Handover *ho;
  uint32_t *ho_list;
  unsigned int *ho_cnt;
  uint32_t ho_cap;
  const Handover *resume;
};

__device__ __forceinline__ int rl_i(int v, int l) {
  return __builtin_amdgcn_readlane(v, l);
}
__device__ __forceinline__ double rl_d(double v, int l) {
  long long b = __double_as_longlong(v);
  int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), l);
  int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned int)lo);
}
__device__ __forceinline__ uint64_t rl_u64(uint64_t v, int l) {
  int lo = __builtin_amdgcn_readlane((int)(v & 0xffffffffull), l);
  int hi = __builtin_amdgcn_readlane((int)(v >> 32), l);
  return ((uint64_t)(unsigned int)hi << 32) | (uint64_t)(unsigned int)lo;
}
// forward permutation: lane `dst` receives this lane's value (ds_permute)
__device__ __forceinline__ int perm_i(int v, int dst) {
  return __builtin_amdgcn_ds_permute(dst << 2, v);
}
__device__ __forceinline__ double perm_d(double v, int dst) {
  long long b = __double_as_longlong(v);
  int lo = __builtin_amdgcn_ds_permute(dst << 2, (int)(b & 0xffffffffll));
  int hi = __builtin_amdgcn_ds_permute(dst << 2, (int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned int)lo);
}
// DPP move of a double (both halves with the same control)
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
  long long b = __double_as_longlong(v);
  int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL, 0xf, 0xf, true);
  int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, true);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned int)lo);
}
// Maximum over the wavefront of non-negative values (0 for inactive
// lanes): row_shr 1/2/4/8 within rows of 16, then row_bcast 15/31 across
// rows; lane 63 holds the result.  DPP keeps the reduction in the VALU.
__device__ __forceinline__ double wave_max_nonneg(double v) {
  v = fmax(v, dpp_d<0x111>(v));
  v = fmax(v, dpp_d<0x112>(v));
  v = fmax(v, dpp_d<0x114>(v));
  v = fmax(v, dpp_d<0x118>(v));
  v = fmax(v, dpp_d<0x142>(v));
  v = fmax(v, dpp_d<0x143>(v));
  return rl_d(v, 63);
}
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ void lds_fence() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// The point p of stored plane j as HalfSpace::new received it
// (half_space.rs:24-34): a box face's point (the face whose normal the
// plane has), else 0.5 (q + neighbour) with the neighbour's image shift,
// the same operations as when the plane was made.
__device__ __forceinline__ d3 w_plane_point(const WaveLds &S, const WaveArgs &A,
                                            const BoxParams &B, d3 q, int j) {
  const uint64_t ref = S.pref[j];
  if (ref == BOX_REF) {
    const double4 a = S.pa[j];
    int k = 0;
    for (int c = 1; c < 6; c++)
      if (a.x == B.pn[c][0] && a.y == B.pn[c][1] && a.z == B.pn[c][2]) k = c;
    return mk3(B.pp[k][0], B.pp[k][1], B.pp[k][2]);
  }
  d3 ngb = xyz(A.spos[pk_slot(ref)]);
  d3 s;
  if (image_shift(pk_img(ref), B.width, &s)) ngb = add3(ngb, s);
  return scl3(0.5, add3(q, ngb));
}

// HalfSpace::right_loc of a stored plane (half_space.rs:53-65)
__device__ __forceinline__ d3 w_right_loc(const WaveLds &S, const WaveArgs &A,
                                          const BoxParams &B, d3 q, int j) {
  const uint64_t ref = S.pref[j];
  if (ref == BOX_REF) {
    const double4 a = S.pa[j];
    d3 pr = plane_project_onto(xyz(a), w_plane_point(S, A, B, q, j), q);
    return sub3(scl3(2., pr), q);
  }
  d3 l = xyz(A.spos[pk_slot(ref)]);
  d3 s;
  if (image_shift(pk_img(ref), B.width, &s)) l = add3(l, s);
  return l;
}

// The position a lane's new vertex takes: with copies kept (at most 32
// vertices after the clip) both halves hold position lane & 31, else the
// lane's own index.
__device__ __forceinline__ int new_pos(bool dup, int nv_after, int lane) {
  return dup && nv_after <= 32 ? (lane & 31) : lane;
}

// wave-uniform SimpleCycle ops on the pointer register `cyc`
__device__ __forceinline__ void wc_set(int &cyc, int lane, int at, int val) {
  cyc = lane == at ? val : cyc;
}

// RESUME: item i of A.sub continues the cell A.resume[i] that the pair kernel
// handed over (its own instantiation, so that the others keep their
// registers).  KEEP: also append every completed cell to the kept-cell
// store.  EXACT:
// decide near-degenerate tests with the exact in-sphere predicate; without
// it (its call holds ~30 VGPRs) the kernel fits 96 VGPRs, i.e. five
// wavefronts per SIMD (as many as the 8 KiB of LDS per wavefront allow), and
// hands such cells (rare outside lattices) to the EXACT instantiation
// through A.exact_list.
template <bool KEEP, bool EXACT, bool RESUME = false>
__global__ void __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(EXACT || KEEP ? 3 : 5)))
k_clip_wave(WaveArgs A) {
  __shared__ WaveLds lds_all[4];  // 4 x 8 KiB
  WaveLds &S = lds_all[threadIdx.x >> 6];
  const int lane = (int)__lane_id();
  unsigned long long my_exact = 0;

  // nsub: the item count on the device, bounded by nwork (or by the
  // hand-over capacity when resuming)
  const uint32_t nwork = A.nsub ? min(*A.nsub, RESUME ? A.ho_cap : A.nwork) : A.nwork;
  uint32_t next = 0;
  if (lane == 0) next = atomicAdd(A.queue, 1u);
  next = (uint32_t)rl_i((int)next, 0);
  for (;;) {
    if (next >= nwork) break;
    const uint32_t w = A.sub ? A.sub[next] : A.w0 + next;
    // claim the following item now; its latency overlaps this cell
    uint32_t nx = 0;
    if (lane == 0) nx = atomicAdd(A.queue, 1u);

    const uint32_t wslot = __builtin_amdgcn_readfirstlane(A.work[w]);
    const uint32_t pr = __builtin_amdgcn_readfirstlane(problem_of(A.PR, wslot));
    const BoxParams &B = A.PR.boxes[pr];
    const int dim = B.dim;
    const uint64_t total_stream = A.PR.grids[pr].total_stream;
    const double4 qq = A.spos[wslot];
    const d3 q = xyz(qq);
    const uint64_t idx = (uint64_t)__double_as_longlong(qq.w);
    const uint32_t cnt = A.kcnt[w];
    // read now, used when the list runs out: its latency hides behind the
    // cell's clipping instead of stalling the wave at the end
    const double kdist_w = A.kdist[w];
    const uint64_t* list = A.kpk + (A.kofs ? A.kofs[w] : (uint64_t)w * A.L);
    uint8_t state = CS_DONE;
    uint32_t err = 0, npop = 0, nexact = 0;
    // a cell handed over by the pair kernel continues from its saved state
    const Handover *R = RESUME ? A.resume + next : nullptr;

    // ---- ConvexCell::init (convex_cell.rs:294-312)
    if (R) {
      if (lane < (int)R->np) {
        S.pa[lane] = R->pa[lane];
        S.pref[lane] = R->pref[lane];
      }
    } else {
      // (the lane through an empty asm: the address is recomputed per cell
      // rather than held across the cell loop)
      int il = lane;
      asm volatile("" : "+v"(il));
      if (il < 6) {
        S.pa[il] = make_double4(B.pn[il][0], B.pn[il][1], B.pn[il][2], B.pd[il]);
        S.pref[il] = BOX_REF;
      }
    }
    lds_fence();
    int np = R ? (int)R->np : 6, nv = R ? (int)R->nv : 8, cstart = 0, clen = 0;
    // dup: while the cell has at most 32 vertices, lanes 32..63 hold copies
    // of lanes 0..31 (vertex position p on lanes p and p + 32), so a scan
    // pass tests two list entries at once.  Every step that moves or adds
    // vertices does the same on both halves until a clip leaves more than
    // 32; from then on the cell uses one lane per position up to 64.
    bool dup = R == nullptr;
    int cyc = lane;  // ptrs[p] of SimpleCycle::new
    // vs = 2 |v - q| as the reference rounds it (2 sqrt(r2)): the safety
    // radius 2 sqrt(max r2) is the maximum of vs over the vertices, sqrt
    // being correctly rounded and so monotone
    double vx = 0., vy = 0., vz = 0., vs = -INFINITY;
    int vd = 0;
    if (R) {
      if (lane < nv) {
        const double4 v = R->v[lane];
        vx = v.x;
        vy = v.y;
        vz = v.z;
        vs = v.w;
        vd = R->vd[lane];
      }
    } else {
      // init vertex duals (2,5,0) (5,3,0) (1,5,2) (5,1,3) (4,2,0) (4,0,3)
      // (2,4,1) (4,3,1), packed 8 bits each; the three components of vertex
      // v are nibble v of three constants (no table in memory)
      int dl0 = lane;
      asm volatile("" : "+v"(dl0));
      const uint32_t sh = 4u * (uint32_t)(dl0 & 7);
      const uint32_t du = ((0x42445152u >> sh) & 15u) | (((0x34021535u >> sh) & 15u) << 8) |
                          (((0x11303200u >> sh) & 15u) << 16);
      // the positions are the box's (Vertex::from_dual of the box planes,
      // computed once per problem with the same operations); only the
      // distance to q is the cell's own
      if ((lane & 31) < 8) {
        const double *cp = B.corner[lane & 7];
        vx = cp[0];
        vy = cp[1];
        vz = cp[2];
        d3 dl = mk3(vx, vy, vz);
        if (dim == 1) {
          dl.y = 0.;
          dl.z = 0.;
        } else if (dim == 2) {
          dl.z = 0.;
        }
        const d3 df = sub3(q, dl);
        vs = 2. * sqrt(dot3(df, df));
        vd = (int)du;
      }
      if (!B.corner_ok) {
        state = CS_ERROR;
        err = E_DET_ZERO;
      }
    }
    if (state == CS_DONE && wave_ballot(lane < nv && isnan(vs))) {
      state = CS_ERROR;
      err = E_NAN;
    }
    if (state == CS_DONE && cnt == 0xFFFFFFFFu) state = CS_CAPACITY;
    if (!R && state == CS_DONE && (cnt == 0 || (uint64_t)pk_idx(list[0]) != idx)) {
      state = CS_ERROR;
      err = E_SELF_NOT_FIRST;
    }

    // ---- ConvexCell::build (convex_cell.rs:316-356)
    bool stopped = false;
    for (uint32_t j0 = R ? R->jnext : 1u; state == CS_DONE && !stopped && j0 < cnt; j0 += 64) {
      // prepare the planes of entries j0 .. j0+63, one per lane, and stage
      // them in LDS
      const uint32_t je = j0 + (uint32_t)lane;
      d3 pn = mk3(0, 0, 0);
      double pdist = 1., pd = 0., perrb = 0.;
      lds_fence();
      if (je < cnt) {
        const uint64_t ppk = list[je];
        d3 ngb = xyz(A.spos[pk_slot(ppk)]);
        d3 s;
        if (image_shift(pk_img(ppk), B.width, &s)) ngb = add3(ngb, s);
        d3 dx = sub3(q, ngb);
        pdist = sqrt(dot3(dx, dx));
        pn = div3(dx, pdist);
        const d3 pp = scl3(0.5, add3(q, ngb));
        // HalfSpace::new (half_space.rs:24-34)
        perrb = 1e-13 * (1. + dot3(abs3(pn), abs3(pp)));
        pd = dot3(pn, pp);
        S.u.e.ref[lane] = ppk;
        S.u.e.nd[lane] = make_double4(pn.x, pn.y, pn.z, pd);
        S.u.e.de[lane] = make_double2(pdist, perrb);
      }
      S.u.e.emat[lane] = 0ull;
      S.u.e.rmat[lane] = 0ull;
      // the first entry whose distance is not finite and positive
      const uint64_t bad_m = wave_ballot(je < cnt && !(isfinite(pdist) && pdist > 0.0));
      const int tbad = bad_m ? (int)__builtin_ctzll(bad_m) : 64;
      lds_fence();
      const int nb = (int)min(64u, cnt - j0);
      int t = 0;
      if (j0 > 128) {
        // a long list past its first 128 entries rarely clips: lane l tests
        // entry l of the batch against every vertex (broadcast from its
        // lane) and the scan below starts at the first entry that is bad,
        // stops the cell (sr < dist, sr the largest vertex term) or touches
        // it; the entries before it change nothing
        const bool act = lane < nv;
        const double srm = wave_max_nonneg(act ? vs : 0.);
        const bool in = lane < nb;
        const double4 e = S.u.e.nd[in ? lane : 0];
        const double2 de = S.u.e.de[in ? lane : 0];
        bool hit = lane == tbad || srm < de.x;
        for (int v = 0; v < nv; v++) {
          const double c = dot3(xyz(e), mk3(rl_d(vx, v), rl_d(vy, v), rl_d(vz, v))) - e.w;
          hit |= (fabs(c) < de.y) | (c < 0.);
        }
        const uint64_t hm = wave_ballot(in && hit);
        t = hm ? (int)__builtin_ctzll(hm) : nb;
      }
      while (t < nb) {
        // scan the entries that leave the cell unchanged (no stop, no
        // vertex near or inside their plane).  Unrolled by two with the
        // next entry's plane read one ahead into the other buffer, so the
        // loop carries only t and one buffer.
        const uint64_t act_m = nv >= 64 ? ~0ull : ((1ull << nv) - 1ull);
        const bool act = lane < nv;
        const d3 vl = mk3(vx, vy, vz);
        // sr < dist: every vertex has vs < dist; an event: a vertex near
        // or inside the plane (clip_by_plane, convex_cell.rs:368-390)
        auto stops = [&](double dist) { return !(wave_ballot(!(vs < dist)) & act_m); };
        auto touches = [&](double4 e, double errb) {
          const double c = dot3(xyz(e), vl) - e.w;
          // two ballots ORed on the scalar unit: a ballot of the two
          // compares ORed per lane costs two more VALU instructions
          return ((wave_ballot(fabs(c) < errb) | wave_ballot(c < 0.)) & act_m) != 0ull;
        };
        bool stop_hit = false;
        if (dup) {
          // the low half tests entry t, the high half entry t + 1 (the
          // reference's order: entry t's stop and clip tests decide before
          // entry t + 1 is looked at); the next pair is read ahead
          const int hi = lane >> 5;
          double4 e = S.u.e.nd[min(t + hi, 63)];
          double2 de = S.u.e.de[min(t + hi, 63)];
          for (;;) {
            if (t == tbad) break;
            const bool two = t + 1 < nb && t + 1 != tbad;
            const uint64_t ns = wave_ballot(!(vs < de.x));
            const double c = dot3(xyz(e), vl) - e.w;
            const uint64_t tm = wave_ballot(fabs(c) < de.y) | wave_ballot(c < 0.);
            // the next pair, read into the same registers now that they are
            // used: the loads overlap the decisions below, and no register
            // copy is needed at the end of the pass
            e = S.u.e.nd[min(t + 2 + hi, 63)];
            de = S.u.e.de[min(t + 2 + hi, 63)];
            if (!(ns & act_m)) {
              stop_hit = true;
              break;
            }
            if (tm & act_m) break;
            if (++t >= nb || !two) break;
            if (!((ns >> 32) & act_m)) {
              stop_hit = true;
              break;
            }
            if ((tm >> 32) & act_m) break;
            if (++t >= nb) break;
          }
          // a pair cut short by the bad entry ends on it: t == tbad
        } else {
        double4 ndA = S.u.e.nd[t];
        double2 deA = S.u.e.de[t];
        for (;;) {
          if (t == tbad) break;
          const double4 ndB = S.u.e.nd[min(t + 1, 63)];
          const double2 deB = S.u.e.de[min(t + 1, 63)];
          if (stops(deA.x)) {
            stop_hit = true;
            break;
          }
          if (touches(ndA, deA.y)) break;
          if (++t >= nb || t == tbad) break;
          ndA = S.u.e.nd[min(t + 1, 63)];
          deA = S.u.e.de[min(t + 1, 63)];
          if (stops(deB.x)) {
            stop_hit = true;
            break;
          }
          if (touches(ndB, deB.y)) break;
          if (++t >= nb) break;
        }
        }
        if (t >= nb) break;
        if (t == tbad) {
          state = CS_ERROR;
          err = E_DEGENERATE;
          break;
        }
        if (stop_hit) {
          stopped = true;
          npop = j0 + (uint32_t)t + 1;
          break;
        }
        // the event's entry, read again
        const double4 nd = S.u.e.nd[t];
        const double errb = S.u.e.de[t].y;
        const double cv = dot3(xyz(nd), vl) - nd.w;
        const d3 n = xyz(nd);
        const double d = nd.w;
        const bool near = act && fabs(cv) < errb;
        bool clipped = act && !near && cv < 0.;
        const uint64_t ref = S.u.e.ref[t];
        if (!EXACT && wave_ballot(near)) {
          state = CS_EXACT;
          break;
        }
        if (EXACT && wave_ballot(near)) {
          d3 ngb = xyz(A.spos[pk_slot(ref)]);
          d3 sh;
          if (image_shift(pk_img(ref), B.width, &sh)) ngb = add3(ngb, sh);
          if (near) {
            int64_t ia[3], ib[3], ic[3], id[3], iv[3];
            iloc(B, q, ia);
            iloc(B, w_right_loc(S, A, B, q, vd & 0xFF), ib);
            iloc(B, w_right_loc(S, A, B, q, (vd >> 8) & 0xFF), ic);
            iloc(B, w_right_loc(S, A, B, q, (vd >> 16) & 0xFF), id);
            iloc(B, ngb, iv);
            clipped = in_sphere_exact(ia, ib, ic, id, iv) < 0;
            nexact++;
          }
        }
        uint64_t M = wave_ballot(clipped);
        if (!M) {
          t++;
          continue;
        }
        // the swap-with-last scan of clip_by_plane (convex_cell.rs:368-390)
        // in closed form.  With K kept vertices, holes h_1 < h_2 < ... (the
        // clipped slots below K) and b_1 > b_2 > ... (the kept slots from K
        // up): h_j takes the vertex of b_j, the vertex of h_j goes to
        // b_{j-1} - 1 (b_0 = nv), a clipped vertex at K goes to b_m - 1
        // (the lowest b, else nv), and every other clipped vertex at p > K
        // moves to p - 1; kept vertices below K stay.
        const uint64_t kept_m = act_m & ~M;
        const int num_v = __popcll(kept_m);  // < nv: M is not empty
        const uint64_t low_m = (1ull << num_v) - 1ull;
        const uint64_t holes = M & low_m, backk = kept_m & ~low_m;
        // pos: the vertex position this lane holds; with copies, the high
        // half does the same moves 32 lanes up
        const int pos = dup ? (lane & 31) : lane;
        const int hbase = lane - pos;
        // every lane's rule but the pairs', then the pairs (h_j, b_j) one by
        // one from the scalar masks (a compare and a select per lane each)
        int dst = pos;
        if ((M >> pos) & 1ull)
          dst = pos == num_v ? (backk ? (int)__builtin_ctzll(backk) : nv) - 1 : pos - 1;
        {
          uint64_t hr = holes, br = backk;
          int prev_b = nv;
          while (hr) {
            const int h = (int)__builtin_ctzll(hr);
            const int b = 63 - (int)__builtin_clzll(br);
            dst = pos == b ? h : dst;
            dst = pos == h ? prev_b - 1 : dst;
            prev_b = b;
            hr &= hr - 1ull;
            br &= ~(1ull << b);
          }
        }
        dst += hbase;
        vx = perm_d(vx, dst);
        vy = perm_d(vy, dst);
        vz = perm_d(vz, dst);
        vs = perm_d(vs, dst);
        vd = perm_i(vd, dst);
        // push the plane (index np); SimpleCycle::grow is the identity
        // pointer already held by lane np
        if (np >= WCAP) {
          state = EXACT || KEEP ? CS_CAPACITY : CS_EXACT;
          break;
        }
        const int p_idx = np;
        if (lane == 0) {
          S.pa[p_idx] = nd;
          S.pref[p_idx] = ref;
        }
        np += 1;
        // compute_boundary (convex_cell.rs:421-441) over positions
        // num_v .. nv-1.  The cycle it ends with is the boundary of the union
        // of the removed vertices' dual triangles, and it starts at the
        // first plane a0 of the triangle at num_v unless the walk enclosed
        // that plane.  When the triangles form a disk (no directed edge
        // twice, Euler characteristic 1) whose boundary is one cycle through
        // a0, the lanes find the boundary edges at once; otherwise the
        // reference's walk runs step by step.  Either way the new vertices
        // (cur, next, p_idx) follow in cycle order from the start.
        // the new vertex at position num_v + r: planes (cur, next) of the
        // boundary's step r, packed cur | next << 8
        int vcn = 0;
        bool fast = false;
        const int nrem = nv - num_v;
        if (nrem <= 2) {
          // one or two removed vertices (44 % of clips on uniform seeds):
          // the walk in closed form.  One triangle (a, b, c) is the cycle
          // a -> b -> c from a; a second one extends it by try_extend's
          // first rule when exactly one of its planes, x, is off the cycle
          // and the other two follow each other on it (g -> f): x goes
          // between them and the start stays a.  Anything else is left to
          // the step-by-step walk below.
          const int d0 = rl_i(vd, num_v);
          const int a = d0 & 0xFF, b = (d0 >> 8) & 0xFF, c = (d0 >> 16) & 0xFF;
          int p0 = a, p1 = b, p2 = c, p3 = 0;
          if (nrem == 1) {
            fast = true;
            clen = 3;
          } else {
            const int d1 = rl_i(vd, num_v + 1);
            const int t0 = d1 & 0xFF, t1 = (d1 >> 8) & 0xFF, t2 = (d1 >> 16) & 0xFF;
            const bool o0 = t0 != a && t0 != b && t0 != c;
            const bool o1 = t1 != a && t1 != b && t1 != c;
            const bool o2 = t2 != a && t2 != b && t2 != c;
            if ((int)o0 + (int)o1 + (int)o2 == 1) {
              // e: the plane off the cycle; f = e + 1, g = e - 1 (mod 3)
              const int x = o0 ? t0 : o1 ? t1 : t2;
              const int tf = o0 ? t1 : o1 ? t2 : t0;
              const int tg = o0 ? t2 : o1 ? t0 : t1;
              // the successor of g on a -> b -> c -> a
              const int sg = tg == a ? b : tg == b ? c : a;
              if (sg == tf) {
                fast = true;
                clen = 4;
                if (tg == a) {
                  p1 = x; p2 = b; p3 = c;
                } else if (tg == b) {
                  p2 = x; p3 = c;
                } else {
                  p3 = x;
                }
              }
            }
          }
          if (fast) {
            const int k = new_pos(dup, num_v + clen, lane) - num_v;
            const int q4 = clen == 4 ? p3 : p0;
            const int vc = k == 0 ? p0 : k == 1 ? p1 : k == 2 ? p2 : p3;
            const int vn = k == 0 ? p1 : k == 1 ? p2 : k == 2 ? q4 : p0;
            vcn = vc | (vn << 8);
          }
        }
        if (!fast) {
          const bool rem = lane >= num_v && lane < nv;
          const int ta = vd & 0xFF, tb = (vd >> 8) & 0xFF, tc = (vd >> 16) & 0xFF;
          // a directed edge set twice (its bit already there) means the
          // triangles do not form a disk
          bool dupe = false;
          if (rem) {
            dupe = ((atomicOr(&S.u.e.emat[ta], 1ull << tb) >> tb) & 1ull) |
                   ((atomicOr(&S.u.e.emat[tb], 1ull << tc) >> tc) & 1ull) |
                   ((atomicOr(&S.u.e.emat[tc], 1ull << ta) >> ta) & 1ull);
            atomicOr(&S.u.e.rmat[tb], 1ull << ta);
            atomicOr(&S.u.e.rmat[tc], 1ull << tb);
            atomicOr(&S.u.e.rmat[ta], 1ull << tc);
          }
          lds_fence();
          // plane `lane`: its edges out and in; an edge is on the boundary
          // when its reverse is in no removed triangle
          const uint64_t eo = S.u.e.emat[lane], ei = S.u.e.rmat[lane];
          S.u.e.emat[lane] = 0ull;
          S.u.e.rmat[lane] = 0ull;
          const uint64_t bo = eo & ~ei;
          const int nbo = __popcll(bo);
          const int nxt = nbo ? (int)__builtin_ctzll(bo) : 0xFF;
          const int n_edge = __popcll(wave_ballot(nbo != 0));
          const int n_node = __popcll(wave_ballot(eo != 0ull));
          const int nf_t = nv - num_v;
          // one boundary edge out of each boundary plane, no directed edge
          // twice, V - E + F = 1 with E = (3 F + n_edge) / 2
          fast = !wave_ballot(nbo > 1 || dupe) && 2 * n_node == nf_t + n_edge + 2;
          if (fast) {
            const int a0 = rl_i(vd, num_v) & 0xFF;
            const int vp = new_pos(dup, num_v + n_edge, lane);
            int cur = a0;
            // a successor at every step, back at a0 after exactly n_edge
            bool closed = false;
            for (int r = 0; r < n_edge; r++) {
              const int nx = rl_i(nxt, cur);
              if (nx == 0xFF) break;
              const bool at = vp == num_v + r;
              vcn = at ? cur | (nx << 8) : vcn;
              if (nx == a0) {
                closed = r + 1 == n_edge;
                break;
              }
              cur = nx;
            }
            fast = closed;
            clen = n_edge;
          }
        }
        if (!fast) {
          // the reference's walk; every pointer of `cyc` is the identity
          // here
          const int d0 = rl_i(vd, num_v);
          const int a0 = d0 & 0xFF, b0 = (d0 >> 8) & 0xFF, c0 = (d0 >> 16) & 0xFF;
          clen = 3;
          cstart = a0;
          wc_set(cyc, lane, a0, b0);
          wc_set(cyc, lane, b0, c0);
          wc_set(cyc, lane, c0, a0);
          // SimpleCycle's `ptrs[x] != x` (x on the cycle) as a scalar mask:
          // no cycle of the walk is shorter than two, so no node on it
          // points to itself
          uint64_t incyc = (1ull << a0) | (1ull << b0) | (1ull << c0);
          for (int r = num_v + 1; r < nv && state == CS_DONE; r++) {
            int ix = r;
            for (;;) {
              if (ix >= nv) {
                state = CS_ERROR;
                err = E_WALK;
                break;
              }
              const int du = rl_i(vd, ix);
              const int tri[3] = {du & 0xFF, (du >> 8) & 0xFF, (du >> 16) & 0xFF};
              const int cm = (int)((incyc >> tri[0]) & 1ull) |
                             ((int)((incyc >> tri[1]) & 1ull) << 1) |
                             ((int)((incyc >> tri[2]) & 1ull) << 2);
              bool ok = false;
              if (cm == 7) {
                // try_extend's second rule: the path g -> f -> e closes
                // around f, for the first e that has it
                const int ptr[3] = {rl_i(cyc, tri[0]), rl_i(cyc, tri[1]), rl_i(cyc, tri[2])};
                for (int e = 0; e < 3 && !ok; e++) {
                  const int f = e == 2 ? 0 : e + 1, g = e == 0 ? 2 : e - 1;
                  if (ptr[g] == tri[f] && ptr[f] == tri[e]) {
                    wc_set(cyc, lane, tri[g], tri[e]);
                    wc_set(cyc, lane, tri[f], tri[f]);
                    incyc &= ~(1ull << tri[f]);
                    if (cstart == tri[f]) cstart = tri[e];
                    clen -= 1;
                    ok = true;
                  }
                }
              } else if (cm == 6 || cm == 5 || cm == 3) {
                // the first rule: tri[e] is the one plane off the cycle and
                // goes between g and f when the cycle has g -> f
                const int e = cm == 6 ? 0 : cm == 5 ? 1 : 2;
                const int f = e == 2 ? 0 : e + 1, g = e == 0 ? 2 : e - 1;
                if (rl_i(cyc, tri[g]) == tri[f]) {
                  wc_set(cyc, lane, tri[g], tri[e]);
                  wc_set(cyc, lane, tri[e], tri[f]);
                  incyc |= 1ull << tri[e];
                  clen += 1;
                  ok = true;
                }
              }
              if (ok) {
                if (ix > r) {
                  const int dr = rl_i(vd, r);
                  vd = pos == r ? du : (pos == ix ? dr : vd);
                }
                break;
              }
              ix++;
            }
          }
          if (state != CS_DONE) break;
          int cur = cstart;
          const int vp = new_pos(dup, num_v + clen, lane);
          for (int r = 0; r < clen; r++) {
            const int nx = rl_i(cyc, cur);
            vcn = vp == num_v + r ? cur | (nx << 8) : vcn;
            cur = nx;
          }
          cyc = lane;  // SimpleCycle starts the next clip from the identity
        }
        nv = num_v;
        if (nv + clen > WCAP) {
          state = CS_CAPACITY;
          break;
        }
        // the copies survive the clip when the cell keeps at most 32
        // vertices; otherwise positions 32 and up take the high lanes
        const bool dup2 = dup && nv + clen <= 32;
        const int vp = new_pos(dup, nv + clen, lane);
        const bool mk = vp >= nv && vp < nv + clen;
        bool ok = true;
        double newr = 0.;
        for (int q0 = 0; q0 < clen; q0 += 16) {
          // Vertex::from_dual on lane quads: quad i (lanes 4i..4i+3) builds
          // new vertex i, lane 4i+c its coordinate c with the operations
          // intersect_planes applies to that coordinate (cross3's component
          // c is a[c+1] b[c+2] - b[c+1] a[c+2]), every lane the determinant;
          // lane 4i then sums |q - v|^2 in dot3's order.  The vertex lanes
          // collect the results.
          // the lane index through an empty asm, so that the lane constants
          // below are recomputed here rather than kept live (and spilled)
          // across the clip loop
          int ql = lane;
          asm volatile("" : "+v"(ql));
          const int qi = q0 + (ql >> 2), qc = ql & 3;
          const int qsrc = min(nv + qi, 63);
          // planes of lanes outside the new range are stale but < 64
          const int pab = __shfl(vcn, qsrc, 64);
          const int ia = pab & 63, ib = (pab >> 8) & 63;
          const bool live = qi < clen && qc < 3;
          // components c+1, c+2 (lane 4i+3 reads valid, unused ones)
          const int c1 = qc >= 2 ? qc - 2 : qc + 1, c2 = qc == 0 ? 2 : qc - 1;
          const double *ra = &S.pa[ia].x, *rb = &S.pa[ib].x, *rp = &S.u.e.nd[t].x;
          const double a1 = ra[c1], a2 = ra[c2], b1 = rb[c1], b2 = rb[c2];
          const double p1 = rp[c1], p2 = rp[c2];
          const double x12 = b1 * p2 - p1 * b2;  // (n_b x n_p)_c
          const double x20 = p1 * a2 - a1 * p2;  // (n_p x n_a)_c
          const double x01 = a1 * b2 - b1 * a2;  // (n_a x n_b)_c
          // det = n_p . (n_a x n_b), the three components from the quad
          const double det = (n.x * dpp_d<0x00>(x01) + n.y * dpp_d<0x55>(x01)) +
                             n.z * dpp_d<0xAA>(x01);
          const double lc = ((ra[3] * x12 + rb[3] * x20) + d * x01) / det;
          const bool zc = (dim == 1 && qc > 0) || (dim == 2 && qc == 2);
          const double df = (qc == 0 ? q.x : qc == 1 ? q.y : q.z) - (zc ? 0. : lc);
          const double sq = df * df;
          const double r2 = (sq + dpp_d<0x55>(sq)) + dpp_d<0xAA>(sq);
          const double sv = 2. * sqrt(r2);
          ok &= !(live && det == 0.);
          if (live && qc == 0 && isnan(r2)) newr = r2;
          // vertex lanes of this group of 16 take their results
          const int k = vp - nv - q0;
          const bool mine = mk && k >= 0 && k < 16;
          const int g = 4 * min(max(k, 0), 15);
          const double nx = __shfl(lc, g, 64), ny = __shfl(lc, g + 1, 64),
                       nz = __shfl(lc, g + 2, 64), ns = __shfl(sv, g, 64);
          if (mine) {
            vx = nx;
            vy = ny;
            vz = nz;
            vs = ns;
          }
        }
        if (mk) vd = vcn | (p_idx << 16);
        if (wave_ballot(!ok)) {
          state = CS_ERROR;
          err = E_DET_ZERO;
          break;
        }
        if (wave_ballot(isnan(newr))) {
          state = CS_ERROR;
          err = E_NAN;
          break;
        }
        nv += clen;
        dup = dup2;
        if (EXACT && !KEEP && np >= WCAP) {
          // the plane list is full: drop the planes no vertex refers to
          // (they build no face and enter no later vertex), keeping the
          // order of the others -- faces follow plane order, so the outputs
          // are unchanged.  The kept-cell store keeps the reference's plane
          // indices, so KEEP builds hand such cells to the larger kernel;
          // the kernel without the exact predicate hands them to the one
          // with it (this code would cost it a wavefront per SIMD).
          const bool act = lane < nv;
          const int t0 = vd & 0xFF, t1 = (vd >> 8) & 0xFF, t2 = (vd >> 16) & 0xFF;
          // the planes the vertices refer to, ORed in an edge-mask word
          // (zero between boundary steps)
          if (act) atomicOr(&S.u.e.emat[0], (1ull << t0) | (1ull << t1) | (1ull << t2));
          lds_fence();
          const uint64_t live = S.u.e.emat[0];
          lds_fence();
          if (lane == 0) S.u.e.emat[0] = 0ull;
          const int np_new = __popcll(live);
          if (np_new < np) {
            auto renum = [&](int x) { return __popcll(live & ((1ull << x) - 1ull)); };
            if ((dup ? (lane & 31) : lane) < nv)
              vd = renum(t0) | (renum(t1) << 8) | (renum(t2) << 16);
            // one double of every plane at a time (few registers live)
            const bool mv_pl = lane < np && ((live >> lane) & 1ull);
            const int to = renum(lane);
#pragma unroll 1
            for (int c = 0; c < 5; c++) {
              double *src = c < 4 ? &S.pa[lane].x + c : reinterpret_cast<double *>(&S.pref[lane]);
              double *dst = c < 4 ? &S.pa[to].x + c : reinterpret_cast<double *>(&S.pref[to]);
              const double v = mv_pl ? *src : 0.;
              lds_fence();
              if (mv_pl) *dst = v;
              lds_fence();
            }
            np = np_new;
          }
        }
        // update_safety_radius (convex_cell.rs:443-453) is the maximum of
        // vs over every vertex (kept ones were checked for NaN when
        // created); the stop test above reads the vs directly
        t++;
      }
    }
    const double sr = wave_max_nonneg(lane < nv ? vs : 0.);
    if (state == CS_DONE && !stopped) {
      // list exhausted: the first entry left out decides (its distance
      // comes with the list), else the whole stream was used
      if ((uint64_t)cnt >= total_stream)
        npop = cnt;
      else if (sr < kdist_w)
        npop = cnt + 1;
      else
        state = CS_NEED_MORE;
    }

    // ---- decomposition and integrals (convex_cell.rs:131-187,
    // voronoi_cell.rs:40-100, integrals.rs:91-105, voronoi_face.rs:30-43)
    lds_fence();  // the entry staging area becomes the tet staging area
    uint32_t nf = 0;
    uint64_t facemask = 0;
    double fa = 0., fx = 0., fy = 0., fz = 0.;
    if (state == CS_DONE) {
      // which planes build a face (voronoi_cell.rs:53-73)
      bool cons = false;
      if (lane < np) {
        const double4 a = S.pa[lane];
        const uint64_t ref = S.pref[lane];
        cons = true;
        if (dim == 1 && !(a.y == 0. && a.z == 0.)) cons = false;
        if (dim == 2 && !(a.z == 0.)) cons = false;
        if (cons && ref != BOX_REF && pk_img(ref) == 13) {
          const uint32_t right = pk_idx(ref);
          cons = (uint64_t)right > idx || (A.mask != nullptr && !A.mask[right]);
        }
      }
      // vertices in chunks of WCH: lane 3 i + k evaluates the two tets on
      // plane k of the chunk's vertex i and stages them in LDS; the cell
      // (lane c sums component c) and the faces (lane p = plane p) then
      // accumulate them in the reference's order: vertex order, tet order
      // within a vertex
      // the projection of q onto plane `lane` (the face's first tet point),
      // handed to the tet lanes by lane permutes
      d3 pfl = mk3(0., 0., 0.);
      if (lane < np)
        pfl = plane_project_onto(xyz(S.pa[lane]), w_plane_point(S, A, B, q, lane), q);
      double cacc = 0.;
      const double *cterm =
          reinterpret_cast<const double *>(&S.u.t.fcv[0][0]) + (lane & 3);
      bool ok = true, has = false;
      for (int c0 = 0; c0 < nv && ok; c0 += WCH) {
        // the lane index through an empty asm: its constants are recomputed
        // in each chunk instead of held (and spilled) across the cell
        int dl_ = lane;
        asm volatile("" : "+v"(dl_));
        const int sl = dl_ / 3, kk = dl_ - 3 * sl;
        const int c1 = min(nv, c0 + WCH);
        const bool mine = c0 + sl < c1;
        // the vertex of this lane, from the lane that holds it
        const int vsrc = min(c0 + sl, 63);
        const d3 vl = mk3(__shfl(vx, vsrc, 64), __shfl(vy, vsrc, 64), __shfl(vz, vsrc, 64));
        const int du = __shfl(vd, vsrc, 64);
        const int pk = (du >> (8 * kk)) & 0xFF;
        const int pn2 = (du >> (kk == 2 ? 0 : 8 * kk + 8)) & 0xFF;
        // edge point of planes (k+1, k); the one of (k, k-1) comes from the
        // lane of plane k-1 of the same vertex (the reference evaluates
        // both with the same arguments)
        d3 pedge = mk3(0., 0., 0.);
        bool lok = true;
        if (mine) {
          const double4 ak = S.pa[pk], an = S.pa[pn2];
          lok = project_onto_intersection(xyz(an), an.w, xyz(ak), ak.w, q, &pedge);
        }
        if (wave_ballot(!lok)) {
          ok = false;
          break;
        }
        const int psrc = min(kk == 0 ? lane + 2 : lane - 1, 63);
        const d3 prev = mk3(__shfl(pedge.x, psrc, 64), __shfl(pedge.y, psrc, 64),
                            __shfl(pedge.z, psrc, 64));
        const d3 pface = mk3(__shfl(pfl.x, pk, 64), __shfl(pfl.y, pk, 64), __shfl(pfl.z, pk, 64));
        if (mine) {
#pragma unroll
          for (int h = 0; h < 2; h++) {
            const d3 v0 = h == 0 ? pface : pedge;
            const d3 v1 = h == 0 ? prev : pface;
            const double tv = signed_volume_tet(v0, v1, vl, q);
            const d3 tc = scl3(tv, add3(add3(add3(v0, v1), vl), q));
            const double a = signed_area_tri(v0, v1, vl, q);
            const d3 sc = scl3(a, add3(add3(v0, v1), vl));
            S.u.t.fcv[2 * kk + h][sl] = make_double4(tv, tc.x, tc.y, tc.z);
            S.u.t.fcf[2 * kk + h][sl] = make_double4(a, sc.x, sc.y, sc.z);
          }
        }
        lds_fence();
        for (int s2 = c0; s2 < c1; s2++) {
          const int sv = s2 - c0;
          // VolumeCentroidIntegral::collect (integrals.rs:91-95), 6 tets
#pragma unroll
          for (int t = 0; t < 6; t++) cacc += cterm[(t * WCH + sv) * 4];
          // VoronoiFaceIntegral::collect for the two tets on plane `lane`
          const int dv = rl_i(vd, s2);
          const int k = (dv & 0xFF) == lane ? 0
                        : ((dv >> 8) & 0xFF) == lane ? 1
                        : ((dv >> 16) & 0xFF) == lane ? 2 : -1;
          if (k >= 0) {
            const double4 f0 = S.u.t.fcf[2 * k][sv], f1 = S.u.t.fcf[2 * k + 1][sv];
            fa += f0.x;
            fx += f0.y;
            fy += f0.z;
            fz += f0.w;
            fa += f1.x;
            fx += f1.y;
            fy += f1.z;
            fz += f1.w;
            has = true;
          }
        }
        lds_fence();
      }
      const double vol = rl_d(cacc, 0), cx = rl_d(cacc, 1), cy = rl_d(cacc, 2),
                   cz = rl_d(cacc, 3);
      if (!ok) {
        state = CS_ERROR;
        err = E_DET_ZERO;
      } else {
        facemask = wave_ballot(lane < np && cons && has);
        nf = (uint32_t)__popcll(facemask);
        if (lane == 0) {
          // VolumeCentroidIntegral::finalize (integrals.rs:97-105)
          const double norm = vol > 0. ? 0.25 / vol : 0.;
          A.cell_vol[idx] = vol;
          A.cell_cen[3 * idx] = cx * norm;
          A.cell_cen[3 * idx + 1] = cy * norm;
          A.cell_cen[3 * idx + 2] = cz * norm;
          A.cell_sr[idx] = sr;
          A.cell_npop[idx] = npop;
        }
      }
    }
    if (lane == 0) {
      A.wstate[w] = state;
      if (state == CS_NEED_MORE) A.need_r[w] = sr;
      if (!EXACT && state == CS_EXACT) A.exact_list[atomicAdd(A.exact_cnt, 1u)] = w;
      if (state == CS_ERROR)
        atomicMin(A.err_word, (unsigned long long)((idx << 8) | err));
    }
    my_exact += nexact;
    if (state == CS_DONE) {
      unsigned long long base = 0;
      if (lane == 0) {
        base = nf ? atomicAdd(A.st_counter, (unsigned long long)nf) : 0ull;
        A.cell_nf[idx] = nf;
        A.cell_fbase[idx] = base;
      }
      base = rl_u64(base, 0);
      if (base + nf > A.st_cap) {
        if (lane == 0) atomicOr(A.overflow, 1ull);
      } else if ((facemask >> lane) & 1ull) {
        // the faces of lower lanes (mbcnt: no per-lane mask register)
        const uint64_t o =
            base + (uint64_t)__builtin_amdgcn_mbcnt_hi(
                       (uint32_t)(facemask >> 32),
                       __builtin_amdgcn_mbcnt_lo((uint32_t)facemask, 0u));
        const double4 a = S.pa[lane];
        const uint64_t ref = S.pref[lane];
        const double fnrm = fa > 0. ? 1. / (3. * fa) : 0.;
        A.st_cell[o] = (uint32_t)idx;
        A.st_right[o] = ref == BOX_REF ? -1 : (int64_t)pk_idx(ref);
        // a face without periodic shift is also a link of its right cell
        // (the count's old value is the link's slot in that cell's range)
        A.st_lslot[o] = ref != BOX_REF && pk_img(ref) == 13
                            ? atomicAdd(&A.link_in[pk_idx(ref)], 1u)
                            : 0u;
        A.st_img[o] = ref == BOX_REF ? 13 : (uint8_t)pk_img(ref);
        A.st_area[o] = fa;
        A.st_cen[3 * o] = fx * fnrm;
        A.st_cen[3 * o + 1] = fy * fnrm;
        A.st_cen[3 * o + 2] = fz * fnrm;
        A.st_norm[3 * o] = a.x;
        A.st_norm[3 * o + 1] = a.y;
        A.st_norm[3 * o + 2] = a.z;
      }
    }
    if (KEEP && state == CS_DONE) {
      // VoronoiIntegrator: keep the cell's half-spaces and vertices
      unsigned long long bp = 0, bv = 0;
      if (lane == 0) {
        bp = atomicAdd(A.cs.cur_pl, (unsigned long long)np);
        bv = atomicAdd(A.cs.cur_vx, (unsigned long long)nv);
      }
      bp = rl_u64(bp, 0);
      bv = rl_u64(bv, 0);
      if (bp + np > A.cs.cap_pl || bv + nv > A.cs.cap_vx) {
        if (lane == 0) atomicOr(A.overflow, 2ull);
      } else {
        if (lane < np) {
          A.cs.pl_a[bp + lane] = S.pa[lane];
          const d3 pp = w_plane_point(S, A, B, q, lane);
          A.cs.pl_b[bp + lane] = make_double4(pp.x, pp.y, pp.z, __longlong_as_double((long long)S.pref[lane]));
        }
        if (lane < nv)
          A.cs.vx[bv + lane] = make_double4(vx, vy, vz, __longlong_as_double((long long)vd));
        if (lane == 0) {
          A.cs.pbase[idx] = bp;
          A.cs.vbase[idx] = bv;
          A.cs.np[idx] = (uint32_t)np;
          A.cs.nv[idx] = (uint32_t)nv;
        }
      }
    }
    next = (uint32_t)rl_i((int)nx, 0);
    lds_fence();
  }
  my_exact = (unsigned long long)wave_sum((double)my_exact);
  if (lane == 0 && my_exact) atomicAdd(A.exact_counter, my_exact);
}

}  // namespace mv
