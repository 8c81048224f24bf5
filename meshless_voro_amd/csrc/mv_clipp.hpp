// mv_clipp.hpp -- per-cell construction with two cells per wavefront.
//
// The construction of k_clip_wave (the reference's ConvexCell::build,
// clip_by_plane, compute_boundary, update_safety_radius and decomposition:
// src/voronoi/convex_cell.rs:294-453, 131-187) with the wavefront split into
// two halves of 32 lanes, each holding one cell:
//   - vertex s of the cell of half h lives in lane 32 h + s, its half-spaces
//     in the half's LDS area, its SimpleCycle pointers in the half's lanes;
//   - the halves scan their neighbour lists together, one entry of each per
//     pass, and stop at their next clip event; the events of both cells are
//     then processed in one pass, so the lane-parallel steps (vertex
//     classification, the swap-with-last moves, the boundary's edge masks,
//     the new vertices on lane quads, the decomposition) serve two cells per
//     instruction, and the sequential steps (the swap pairs, the boundary
//     walks) run per half on the scalar unit;
//   - a cell that would outgrow 32 vertices or 32 half-spaces, or that reads
//     past entry 128 of its list, is handed to k_clip_wave with its state
//     (half-spaces, vertices, next entry): that kernel continues it exactly
//     where it stopped, so every decision, the vertex array and the plane
//     list are the reference's either way.
// Like k_clip_wave<false, false>, the kernel has no exact predicate: a cell
// that meets a near-degenerate test goes to the exact instantiation.
#pragma once
#include "mv_clipw.hpp"

namespace mv {

constexpr int PCH = 10;               // decomposition chunk: vertices per half
constexpr uint32_t PAIR_JMAX = 128;   // list entries read on this path
constexpr uint8_t CS_HANDOVER = 5;    // continued by k_clip_wave

struct PairLds {
  double4 pa[64];     // half h, plane p at [32 h + p]: n.xyz, d
  uint64_t pref[64];  // reference word
  union {
    struct {
      // staged entries, a ring of 32 per half: entry j of half h at
      // [32 h + (j & 31)] (reference word, n and d, distance and rounding
      // bound).  An entry whose distance is not finite and positive is
      // staged as distance -1 with an infinite bound: the scan stops at it
      // and the event step reports it.
      uint64_t ref[64];
      double4 nd[64];
      double2 de[64];
      // boundary edges: half h, plane x at [32 h + x]; bit y of emat: edge
      // x -> y; bit x of rmat[32 h + y]: edge x -> y
      uint32_t emat[64];
      uint32_t rmat[64];
      // capacity check: [2 h] planes of removed vertices, [2 h + 1] of kept
      uint32_t xm[4];
    } e;
    // decomposition: per (tet, vertex) of a chunk of PCH vertices per half
    struct {
      double4 fcf[6][2 * PCH];  // area, area * (v0 + v1 + v2)
      double4 fcv[6][2 * PCH];  // volume, volume * (v0 + v1 + v2 + gen)
    } t;
  } u;
};
static_assert(sizeof(PairLds) <= 10240, "four blocks of four wavefronts per CU");

// Maximum of non-negative values over each half (lane 31: lanes 0-31, lane
// 63: lanes 32-63): row_shr within rows of 16, then row 1 (3) takes row 0's
// (2's) lane 15.
__device__ __forceinline__ double half_max_nonneg(double v) {
  v = fmax(v, dpp_d<0x111>(v));
  v = fmax(v, dpp_d<0x112>(v));
  v = fmax(v, dpp_d<0x114>(v));
  v = fmax(v, dpp_d<0x118>(v));
  long long b = __double_as_longlong(v);
  int lo = __builtin_amdgcn_update_dpp((int)(b & 0xffffffffll), (int)(b & 0xffffffffll), 0x142,
                                       0xA, 0xf, false);
  int hi = __builtin_amdgcn_update_dpp((int)(b >> 32), (int)(b >> 32), 0x142, 0xA, 0xf, false);
  const double u = __longlong_as_double(((long long)hi << 32) | (long long)(unsigned int)lo);
  return fmax(v, u);
}

// The point p of plane j of a half (see w_plane_point)
__device__ __forceinline__ d3 p_plane_point(const PairLds &S, const WaveArgs &A,
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

// Wave-uniform copies (readfirstlane): the per-half state is the same on
// every lane; saying so keeps it in scalar registers and its branches on the
// scalar unit.
__device__ __forceinline__ uint32_t uni(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_readfirstlane((int)x);
}
__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ bool uni(bool b) { return __builtin_amdgcn_readfirstlane((int)b) != 0; }
__device__ __forceinline__ uint64_t uni(uint64_t x) {
  const uint32_t lo = uni((uint32_t)x), hi = uni((uint32_t)(x >> 32));
  return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ double uni(double x) {
  return __longlong_as_double((long long)uni((uint64_t)__double_as_longlong(x)));
}

__device__ __forceinline__ uint32_t low_bits(int n) {
  return n >= 32 ? 0xFFFFFFFFu : ((1u << n) - 1u);
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
k_clip_pair(WaveArgs A) {
  __shared__ PairLds lds_all[4];  // 4 x 10 KiB
  PairLds &S = lds_all[threadIdx.x >> 6];
  const int lane = (int)__lane_id();
  const bool hi = lane >= 32;  // the lane's half
  const int hb = lane & 32;
  const int sl = lane & 31;    // position within the half

  // nsub: the item count on the device (nwork bounds it)
  const uint32_t nwork = A.nsub ? min(*A.nsub, A.nwork) : A.nwork;
  uint32_t next = 0;
  if (lane == 0) next = atomicAdd(A.queue, 2u);
  next = (uint32_t)rl_i((int)next, 0);
  for (;;) {
    if (next >= nwork) break;
    // claim the following pair now; its latency overlaps these cells
    uint32_t nxq = 0;
    if (lane == 0) nxq = atomicAdd(A.queue, 2u);

    // ---- per half: the cell, its list and its state (wave-uniform)
    bool live[2], run[2], pend[2], exh[2];
    uint32_t w[2], cnt[2], jc[2], jend[2], npop[2], err[2], wsl[2], pr[2];
    uint8_t state[2];
    int np[2], nv[2], dimh[2];
    uint64_t idx[2], tstream[2];
    const uint64_t *list[2];
    double kd[2];
#pragma unroll
    for (int h = 0; h < 2; h++) {
      live[h] = next + (uint32_t)h < nwork;
      w[h] = live[h] ? next + (uint32_t)h : next;
      wsl[h] = __builtin_amdgcn_readfirstlane(A.work[w[h]]);
      pr[h] = __builtin_amdgcn_readfirstlane(problem_of(A.PR, wsl[h]));
      dimh[h] = uni(A.PR.boxes[pr[h]].dim);
      tstream[h] = uni(A.PR.grids[pr[h]].total_stream);
      cnt[h] = uni(A.kcnt[w[h]]);
      kd[h] = uni(A.kdist[w[h]]);
      list[h] = A.kpk + uni(A.kofs ? A.kofs[w[h]] : (uint64_t)w[h] * A.L);
      state[h] = CS_DONE;
      err[h] = 0;
      npop[h] = 0;
      np[h] = 6;
      nv[h] = 8;
      jc[h] = 1;
      jend[h] = 1;
      pend[h] = false;
      exh[h] = false;
    }
    const BoxParams &B = A.PR.boxes[hi ? pr[1] : pr[0]];
    const int dim = hi ? dimh[1] : dimh[0];
    const double4 qq = A.spos[hi ? wsl[1] : wsl[0]];
    const d3 q = xyz(qq);
#pragma unroll
    for (int h = 0; h < 2; h++)
      idx[h] = (uint64_t)__double_as_longlong(rl_d(qq.w, 32 * h));

    // ---- ConvexCell::init (convex_cell.rs:294-312), both halves
    if (sl < 6) {
      S.pa[lane] = make_double4(B.pn[sl][0], B.pn[sl][1], B.pn[sl][2], B.pd[sl]);
      S.pref[lane] = BOX_REF;
    }
    // the edge and plane masks start empty (the previous cell's tets were
    // staged over them)
    S.u.e.emat[lane] = 0u;
    S.u.e.rmat[lane] = 0u;
    if (lane < 4) S.u.e.xm[lane] = 0u;
    int cyc = sl;  // ptrs[p] of SimpleCycle::new, plane p of the half on lane 32 h + p
    double vx = 0., vy = 0., vz = 0., vs = -INFINITY;
    int vd = 0;
    if (sl < 8) {
      // duals (2,5,0) (5,3,0) (1,5,2) (5,1,3) (4,2,0) (4,0,3) (2,4,1) (4,3,1)
      const uint32_t sh = 4u * (uint32_t)sl;
      vd = (int)(((0x42445152u >> sh) & 15u) | (((0x34021535u >> sh) & 15u) << 8) |
                 (((0x11303200u >> sh) & 15u) << 16));
      const double *cp = B.corner[sl];
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
    }
    {
      const uint64_t nanm = wave_ballot(sl < 8 && isnan(vs));
#pragma unroll
      for (int h = 0; h < 2; h++) {
        if (!A.PR.boxes[pr[h]].corner_ok) {
          state[h] = CS_ERROR;
          err[h] = E_DET_ZERO;
        } else if ((nanm >> (32 * h)) & 0xFFFFFFFFull) {
          state[h] = CS_ERROR;
          err[h] = E_NAN;
        } else if (cnt[h] == 0xFFFFFFFFu) {
          state[h] = CS_CAPACITY;
        } else if (live[h] && (cnt[h] == 0 || (uint64_t)pk_idx(list[h][0]) != idx[h])) {
          state[h] = CS_ERROR;
          err[h] = E_SELF_NOT_FIRST;
        }
        run[h] = live[h] && state[h] == CS_DONE;
      }
    }
    lds_fence();

    // hand half h to k_clip_wave: its state before entry jc[h]
    auto handover = [&](int h) {
      unsigned int k = 0;
      if (lane == 32 * h) k = atomicAdd(A.ho_cnt, 1u);
      k = (unsigned int)rl_i((int)k, 32 * h);
      if (k < A.ho_cap) {
        Handover &Hd = A.ho[k];
        const bool mine = hi == (h == 1);
        if (lane == 32 * h) {
          A.ho_list[k] = w[h];
          Hd.np = (uint32_t)np[h];
          Hd.nv = (uint32_t)nv[h];
          Hd.jnext = jc[h];
        }
        if (mine && sl < np[h]) {
          Hd.pa[sl] = S.pa[lane];
          Hd.pref[sl] = S.pref[lane];
        }
        if (mine && sl < nv[h]) {
          Hd.v[sl] = make_double4(vx, vy, vz, vs);
          Hd.vd[sl] = vd;
        }
        state[h] = CS_HANDOVER;
      } else {
        // no room: the large-cell kernel builds it again from the start
        state[h] = CS_CAPACITY;
      }
      run[h] = false;
      pend[h] = false;
    };

    // ---- ConvexCell::build (convex_cell.rs:316-356) for both cells
    for (;;) {
#pragma unroll
      for (int h = 0; h < 2; h++) {
        run[h] = uni(run[h]);
        pend[h] = uni(pend[h]);
        exh[h] = uni(exh[h]);
        jc[h] = uni(jc[h]);
        jend[h] = uni(jend[h]);
        np[h] = uni(np[h]);
        nv[h] = uni(nv[h]);
        state[h] = (uint8_t)uni((uint32_t)state[h]);
        err[h] = uni(err[h]);
        npop[h] = uni(npop[h]);
      }
      // a running half whose staged entries are used up: a list used up
      // finishes below, a list past PAIR_JMAX goes to k_clip_wave, others
      // are staged again (every running half's ring is refilled)
      bool refill = false;
#pragma unroll
      for (int h = 0; h < 2; h++) {
        if (run[h] && !pend[h] && jc[h] == jend[h]) {
          if (jc[h] >= cnt[h]) {
            exh[h] = true;
            run[h] = false;
          } else if (jc[h] >= PAIR_JMAX) {
            handover(h);
          } else {
            refill = true;
          }
        }
      }
      if (refill) {
        // slot sl of half h holds entry jc + ((sl - jc) & 31); the entries
        // from jend on are prepared, one per lane (HalfSpace::new,
        // half_space.rs:24-34, and the distance of convex_cell.rs:340-342)
        const bool rn = hi ? run[1] : run[0];
        const uint32_t jcl = hi ? jc[1] : jc[0], jel = hi ? jend[1] : jend[0];
        const uint32_t cl = hi ? cnt[1] : cnt[0];
        const uint32_t je = jcl + (uint32_t)((sl - (int)jcl) & 31);
        lds_fence();
        if (rn && je >= jel && je < cl) {
          const uint64_t ppk = (hi ? list[1] : list[0])[je];
          d3 ngb = xyz(A.spos[pk_slot(ppk)]);
          d3 s;
          if (image_shift(pk_img(ppk), B.width, &s)) ngb = add3(ngb, s);
          const d3 dx = sub3(q, ngb);
          const double pdist = sqrt(dot3(dx, dx));
          const d3 pn = div3(dx, pdist);
          const d3 pp = scl3(0.5, add3(q, ngb));
          const double perrb = 1e-13 * (1. + dot3(abs3(pn), abs3(pp)));
          const double pd = dot3(pn, pp);
          const bool ok = isfinite(pdist) && pdist > 0.0;
          S.u.e.ref[lane] = ppk;
          S.u.e.nd[lane] = ok ? make_double4(pn.x, pn.y, pn.z, pd) : make_double4(0., 0., 0., 0.);
          S.u.e.de[lane] = ok ? make_double2(pdist, perrb) : make_double2(-1., INFINITY);
        }
#pragma unroll
        for (int h = 0; h < 2; h++)
          if (run[h]) jend[h] = min(jc[h] + 32u, cnt[h]);
        lds_fence();
      }

      // scan: each running half tests its entry jc against its vertices,
      // one entry per half per pass, in list order: the stop test
      // (safety_radius < dist: every vertex's 2 |v - q| below dist), then
      // whether a vertex is near or inside the plane (clip_by_plane,
      // convex_cell.rs:368-390).  An entry that does neither leaves the cell
      // as it is.
      for (;;) {
#pragma unroll
        for (int h = 0; h < 2; h++) {
          run[h] = uni(run[h]);
          pend[h] = uni(pend[h]);
          jc[h] = uni(jc[h]);
          npop[h] = uni(npop[h]);
        }
        const bool sc0 = run[0] && !pend[0] && jc[0] < jend[0];
        const bool sc1 = run[1] && !pend[1] && jc[1] < jend[1];
        if (!sc0 && !sc1) break;
        const int slot = hb + (int)((hi ? jc[1] : jc[0]) & 31u);
        const double4 e = S.u.e.nd[slot];
        const double2 de = S.u.e.de[slot];
        const uint64_t ns = wave_ballot(!(vs < de.x));
        const double c = dot3(xyz(e), mk3(vx, vy, vz)) - e.w;
        const uint64_t tm = wave_ballot(fabs(c) < de.y) | wave_ballot(c < 0.);
#pragma unroll
        for (int h = 0; h < 2; h++) {
          if (!(h ? sc1 : sc0)) continue;
          const uint64_t am = (uint64_t)low_bits(nv[h]) << (32 * h);
          if (!(ns & am)) {
            run[h] = false;
            npop[h] = jc[h] + 1u;
          } else if (tm & am) {
            pend[h] = true;
          } else {
            jc[h]++;
          }
        }
      }
      // a half that still has entries to stage goes back up first, so the
      // events of both cells are processed together
      if ((run[0] && !pend[0]) || (run[1] && !pend[1])) continue;
      if (!run[0] && !run[1]) break;

      // ---- clip_by_plane (convex_cell.rs:358-419) for every running half:
      // each has an event at entry jc
      const bool evl = hi ? run[1] : run[0];
      const int slot = hb + (int)((hi ? jc[1] : jc[0]) & 31u);
      const double4 nd = S.u.e.nd[slot];
      const double2 de = S.u.e.de[slot];
      const int nvl = hi ? nv[1] : nv[0];
      const bool act = evl && sl < nvl;
      const double cv = dot3(xyz(nd), mk3(vx, vy, vz)) - nd.w;
      const bool near = act && fabs(cv) < de.y;
      const bool clipped = act && !near && cv < 0.;
      const uint64_t badm = wave_ballot(evl && sl == 0 && de.x < 0.);
      const uint64_t nearm = wave_ballot(near);
      const uint64_t M = wave_ballot(clipped);
      bool clip[2];
#pragma unroll
      for (int h = 0; h < 2; h++) {
        clip[h] = false;
        if (!run[h]) continue;
        pend[h] = false;
        const uint64_t hm = 0xFFFFFFFFull << (32 * h);
        if (badm & hm) {
          // the first entry whose distance is not finite and positive
          state[h] = CS_ERROR;
          err[h] = E_DEGENERATE;
          run[h] = false;
        } else if (nearm & hm) {
          state[h] = CS_EXACT;
          run[h] = false;
        } else if (!(M & hm)) {
          jc[h]++;
        } else {
          clip[h] = true;
        }
      }
      // capacity: the clip adds a half-space and leaves nv + 2 - 2 I
      // vertices, I the half-spaces every vertex of which is removed (the
      // removed vertices' dual triangles form a disk; otherwise the walk
      // fails either way)
      if ((clip[0] && (np[0] >= PCAP || nv[0] >= PCAP - 1)) ||
          (clip[1] && (np[1] >= PCAP || nv[1] >= PCAP - 1))) {
        bool chk[2];
#pragma unroll
        for (int h = 0; h < 2; h++) {
          chk[h] = false;
          if (!clip[h]) continue;
          if (np[h] >= PCAP) {
            clip[h] = false;
            handover(h);
          } else if (nv[h] >= PCAP - 1) {
            chk[h] = true;
          }
        }
        if (chk[0] || chk[1]) {
          const bool ck = hi ? chk[1] : chk[0];
          const uint32_t bits = (1u << (vd & 0xFF)) | (1u << ((vd >> 8) & 0xFF)) |
                                (1u << ((vd >> 16) & 0xFF));
          if (ck && act) atomicOr(&S.u.e.xm[2 * (hi ? 1 : 0) + (clipped ? 0 : 1)], bits);
          lds_fence();
          uint32_t rk[4];
#pragma unroll
          for (int i = 0; i < 4; i++) rk[i] = S.u.e.xm[i];
          lds_fence();
          if (lane < 4) S.u.e.xm[lane] = 0u;
#pragma unroll
          for (int h = 0; h < 2; h++) {
            if (!chk[h]) continue;
            const int I = __popc(rk[2 * h] & ~rk[2 * h + 1]);
            if (nv[h] + 2 - 2 * I > PCAP) {
              clip[h] = false;
              handover(h);
            }
          }
        }
      }
      if (!clip[0] && !clip[1]) continue;
      const bool cll = hi ? clip[1] : clip[0];

      // the swap-with-last scan of clip_by_plane in closed form (see
      // k_clip_wave): with K kept vertices, holes h_1 < h_2 < ... below K
      // and kept slots b_1 > b_2 > ... from K up, h_j takes the vertex of
      // b_j, the vertex of h_j goes to b_{j-1} - 1 (b_0 = nv), a clipped
      // vertex at K to b_m - 1 (else nv - 1), every other clipped one at
      // p > K to p - 1
      int numv[2];
      uint32_t holes[2], backk[2];
#pragma unroll
      for (int h = 0; h < 2; h++) {
        numv[h] = nv[h];
        holes[h] = backk[h] = 0u;
        if (!clip[h]) continue;
        const uint32_t Mh = (uint32_t)(M >> (32 * h));
        const uint32_t kept = low_bits(nv[h]) & ~Mh;
        numv[h] = uni(__popc(kept));
        const uint32_t low = low_bits(numv[h]);
        holes[h] = Mh & low;
        backk[h] = kept & ~low;
      }
      const int numvl = hi ? numv[1] : numv[0];
      int dst = sl;
      if (cll && clipped) {
        const uint32_t bk = hi ? backk[1] : backk[0];
        dst = sl == numvl ? (bk ? (int)__builtin_ctz(bk) : nvl) - 1 : sl - 1;
      }
      {
        uint32_t hr0 = holes[0], br0 = backk[0], hr1 = holes[1], br1 = backk[1];
        int pb0 = nv[0], pb1 = nv[1];
        while (hr0 | hr1) {
          int h0 = -1, b0 = -1, h1 = -1, b1 = -1;
          if (hr0) {
            h0 = (int)__builtin_ctz(hr0);
            b0 = 31 - (int)__builtin_clz(br0);
          }
          if (hr1) {
            h1 = (int)__builtin_ctz(hr1);
            b1 = 31 - (int)__builtin_clz(br1);
          }
          const int hh = hi ? h1 : h0, bb = hi ? b1 : b0, pb = hi ? pb1 : pb0;
          if (cll) {
            dst = sl == bb ? hh : dst;
            dst = sl == hh ? pb - 1 : dst;
          }
          if (hr0) {
            pb0 = b0;
            hr0 &= hr0 - 1u;
            br0 &= ~(1u << b0);
          }
          if (hr1) {
            pb1 = b1;
            hr1 &= hr1 - 1u;
            br1 &= ~(1u << b1);
          }
        }
      }
      dst += hb;
      vx = perm_d(vx, dst);
      vy = perm_d(vy, dst);
      vz = perm_d(vz, dst);
      vs = perm_d(vs, dst);
      vd = perm_i(vd, dst);
      // push the plane (SimpleCycle::grow: the identity pointer is there)
      int pidx[2];
#pragma unroll
      for (int h = 0; h < 2; h++) {
        pidx[h] = np[h];
        if (clip[h]) np[h]++;
      }
      if (cll && sl == 0) {
        const int p = hi ? pidx[1] : pidx[0];
        S.pa[hb + p] = nd;
        S.pref[hb + p] = S.u.e.ref[slot];
      }

      // compute_boundary (convex_cell.rs:421-441) over positions num_v ..
      // nv - 1 of each half (see k_clip_wave): one or two removed vertices
      // in closed form; otherwise the edge masks of the removed triangles
      // give the cycle when they form a disk whose boundary passes through
      // the walk's start; otherwise the reference's walk step by step.  The
      // lane at position num_v + r of a half receives the planes
      // cur | next << 8 of the cycle's step r.
      int vcn = 0;
      bool fast[2];
      int clen[2];
#pragma unroll
      for (int h = 0; h < 2; h++) {
        fast[h] = false;
        clen[h] = 0;
        if (!clip[h]) continue;
        const int nrem = nv[h] - numv[h];
        if (nrem > 2) continue;
        const int d0 = rl_i(vd, 32 * h + numv[h]);
        const int a = d0 & 0xFF, b = (d0 >> 8) & 0xFF, c = (d0 >> 16) & 0xFF;
        int p0 = a, p1 = b, p2 = c, p3 = 0;
        if (nrem == 1) {
          fast[h] = true;
          clen[h] = 3;
        } else {
          const int d1 = rl_i(vd, 32 * h + numv[h] + 1);
          const int t0 = d1 & 0xFF, t1 = (d1 >> 8) & 0xFF, t2 = (d1 >> 16) & 0xFF;
          const bool o0 = t0 != a && t0 != b && t0 != c;
          const bool o1 = t1 != a && t1 != b && t1 != c;
          const bool o2 = t2 != a && t2 != b && t2 != c;
          if ((int)o0 + (int)o1 + (int)o2 == 1) {
            const int x = o0 ? t0 : o1 ? t1 : t2;
            const int tf = o0 ? t1 : o1 ? t2 : t0;
            const int tg = o0 ? t2 : o1 ? t0 : t1;
            const int sg = tg == a ? b : tg == b ? c : a;
            if (sg == tf) {
              fast[h] = true;
              clen[h] = 4;
              if (tg == a) {
                p1 = x;
                p2 = b;
                p3 = c;
              } else if (tg == b) {
                p2 = x;
                p3 = c;
              } else {
                p3 = x;
              }
            }
          }
        }
        if (fast[h] && hi == (h == 1)) {
          const int k = sl - numv[h];
          const int q4 = clen[h] == 4 ? p3 : p0;
          const int vc = k == 0 ? p0 : k == 1 ? p1 : k == 2 ? p2 : p3;
          const int vn = k == 0 ? p1 : k == 1 ? p2 : k == 2 ? q4 : p0;
          vcn = vc | (vn << 8);
        }
      }
      const bool mp0 = clip[0] && !fast[0], mp1 = clip[1] && !fast[1];
      if (mp0 || mp1) {
        const bool mpl = hi ? mp1 : mp0;
        const bool rem = mpl && sl >= numvl && sl < nvl;
        const int ta = vd & 0xFF, tb = (vd >> 8) & 0xFF, tc = (vd >> 16) & 0xFF;
        bool dupe = false;
        if (rem) {
          uint32_t *em = &S.u.e.emat[hb], *rm = &S.u.e.rmat[hb];
          dupe = ((atomicOr(&em[ta], 1u << tb) >> tb) & 1u) |
                 ((atomicOr(&em[tb], 1u << tc) >> tc) & 1u) |
                 ((atomicOr(&em[tc], 1u << ta) >> ta) & 1u);
          atomicOr(&rm[tb], 1u << ta);
          atomicOr(&rm[tc], 1u << tb);
          atomicOr(&rm[ta], 1u << tc);
        }
        lds_fence();
        const uint32_t eo = S.u.e.emat[lane], ei = S.u.e.rmat[lane];
        S.u.e.emat[lane] = 0u;
        S.u.e.rmat[lane] = 0u;
        const uint32_t bo = eo & ~ei;
        const int nbo = __popc(bo);
        const int nxt = nbo ? (int)__builtin_ctz(bo) : 0xFF;
        const uint64_t edge_m = wave_ballot(nbo != 0);
        const uint64_t node_m = wave_ballot(eo != 0u);
        const uint64_t bad_m = wave_ballot(nbo > 1 || dupe);
#pragma unroll
        for (int h = 0; h < 2; h++) {
          if (!(h ? mp1 : mp0)) continue;
          const int n_edge = __popc((uint32_t)(edge_m >> (32 * h)));
          const int n_node = __popc((uint32_t)(node_m >> (32 * h)));
          const int nf_t = nv[h] - numv[h];
          if ((uint32_t)(bad_m >> (32 * h)) || 2 * n_node != nf_t + n_edge + 2) continue;
          const int a0 = rl_i(vd, 32 * h + numv[h]) & 0xFF;
          int cur = a0;
          bool closed = false;
          for (int r = 0; r < n_edge; r++) {
            const int nx = rl_i(nxt, 32 * h + cur);
            if (nx == 0xFF) break;
            if (lane == 32 * h + numv[h] + r) vcn = cur | (nx << 8);
            if (nx == a0) {
              closed = r + 1 == n_edge;
              break;
            }
            cur = nx;
          }
          fast[h] = closed;
          clen[h] = n_edge;
        }
      }
#pragma unroll
      for (int h = 0; h < 2; h++) {
        if (!clip[h] || fast[h]) continue;
        // the reference's walk; the half's pointers are the identity here
        const int base = 32 * h;
        const int d0 = rl_i(vd, base + numv[h]);
        const int a0 = d0 & 0xFF, b0 = (d0 >> 8) & 0xFF, c0 = (d0 >> 16) & 0xFF;
        int cl = 3, cstart = a0;
        wc_set(cyc, lane, base + a0, b0);
        wc_set(cyc, lane, base + b0, c0);
        wc_set(cyc, lane, base + c0, a0);
        uint32_t incyc = (1u << a0) | (1u << b0) | (1u << c0);
        bool failed = false;
        for (int r = numv[h] + 1; r < nv[h] && !failed; r++) {
          int ix = r;
          for (;;) {
            if (ix >= nv[h]) {
              failed = true;
              break;
            }
            const int du = rl_i(vd, base + ix);
            const int tri[3] = {du & 0xFF, (du >> 8) & 0xFF, (du >> 16) & 0xFF};
            const int cm = (int)((incyc >> tri[0]) & 1u) | ((int)((incyc >> tri[1]) & 1u) << 1) |
                           ((int)((incyc >> tri[2]) & 1u) << 2);
            bool ok = false;
            if (cm == 7) {
              const int ptr[3] = {rl_i(cyc, base + tri[0]), rl_i(cyc, base + tri[1]),
                                  rl_i(cyc, base + tri[2])};
              for (int e = 0; e < 3 && !ok; e++) {
                const int f = e == 2 ? 0 : e + 1, g = e == 0 ? 2 : e - 1;
                if (ptr[g] == tri[f] && ptr[f] == tri[e]) {
                  wc_set(cyc, lane, base + tri[g], tri[e]);
                  wc_set(cyc, lane, base + tri[f], tri[f]);
                  incyc &= ~(1u << tri[f]);
                  if (cstart == tri[f]) cstart = tri[e];
                  cl -= 1;
                  ok = true;
                }
              }
            } else if (cm == 6 || cm == 5 || cm == 3) {
              const int e = cm == 6 ? 0 : cm == 5 ? 1 : 2;
              const int f = e == 2 ? 0 : e + 1, g = e == 0 ? 2 : e - 1;
              if (rl_i(cyc, base + tri[g]) == tri[f]) {
                wc_set(cyc, lane, base + tri[g], tri[e]);
                wc_set(cyc, lane, base + tri[e], tri[f]);
                incyc |= 1u << tri[e];
                cl += 1;
                ok = true;
              }
            }
            if (ok) {
              if (ix > r) {
                const int dr = rl_i(vd, base + r);
                vd = lane == base + r ? du : (lane == base + ix ? dr : vd);
              }
              break;
            }
            ix++;
          }
        }
        if (failed) {
          state[h] = CS_ERROR;
          err[h] = E_WALK;
          run[h] = false;
          clip[h] = false;
        } else {
          int cur = cstart;
          for (int r = 0; r < cl; r++) {
            const int nx = rl_i(cyc, base + cur);
            if (lane == base + numv[h] + r) vcn = cur | (nx << 8);
            cur = nx;
          }
          clen[h] = cl;
        }
        // SimpleCycle starts the next clip from the identity
        if (hi == (h == 1)) cyc = sl;
      }
#pragma unroll
      for (int h = 0; h < 2; h++) {
        if (clip[h] && numv[h] + clen[h] > PCAP) {
          // cannot happen for a disk (checked above); built again elsewhere
          state[h] = CS_CAPACITY;
          run[h] = false;
          clip[h] = false;
        }
      }
      if (!clip[0] && !clip[1]) continue;

      // Vertex::from_dual on lane quads (see k_clip_wave): quad i of a half
      // builds new vertex i of its cell, 8 per half per pass
      {
        const bool cl2 = hi ? clip[1] : clip[0];
        const int clenl = cl2 ? (hi ? clen[1] : clen[0]) : 0;
        const bool mk = cl2 && sl >= numvl && sl < numvl + clenl;
        const int qmax = max(clip[0] ? clen[0] : 0, clip[1] ? clen[1] : 0);
        bool ok = true;
        double newr = 0.;
        for (int q0 = 0; q0 < qmax; q0 += 8) {
          int ql = sl;
          asm volatile("" : "+v"(ql));
          const int qi = q0 + (ql >> 2), qc = ql & 3;
          const int qsrc = hb + min(numvl + qi, 31);
          const int pab = __shfl(vcn, qsrc, 64);
          const int ia = pab & 31, ib = (pab >> 8) & 31;
          const bool live_q = qi < clenl && qc < 3;
          const int c1 = qc >= 2 ? qc - 2 : qc + 1, c2 = qc == 0 ? 2 : qc - 1;
          const double *ra = &S.pa[hb + ia].x, *rb = &S.pa[hb + ib].x;
          const double *rp = &S.u.e.nd[slot].x;
          const double a1 = ra[c1], a2 = ra[c2], b1 = rb[c1], b2 = rb[c2];
          const double p1 = rp[c1], p2 = rp[c2];
          const double x12 = b1 * p2 - p1 * b2;
          const double x20 = p1 * a2 - a1 * p2;
          const double x01 = a1 * b2 - b1 * a2;
          const double det = (nd.x * dpp_d<0x00>(x01) + nd.y * dpp_d<0x55>(x01)) +
                             nd.z * dpp_d<0xAA>(x01);
          const double lc = ((ra[3] * x12 + rb[3] * x20) + nd.w * x01) / det;
          const bool zc = (dim == 1 && qc > 0) || (dim == 2 && qc == 2);
          const double df = (qc == 0 ? q.x : qc == 1 ? q.y : q.z) - (zc ? 0. : lc);
          const double sq = df * df;
          const double r2 = (sq + dpp_d<0x55>(sq)) + dpp_d<0xAA>(sq);
          const double sv = 2. * sqrt(r2);
          ok &= !(live_q && det == 0.);
          if (live_q && qc == 0 && isnan(r2)) newr = r2;
          const int k = sl - numvl - q0;
          const bool mine = mk && k >= 0 && k < 8;
          const int g = hb + 4 * min(max(k, 0), 7);
          const double nx = __shfl(lc, g, 64), ny = __shfl(lc, g + 1, 64),
                       nz = __shfl(lc, g + 2, 64), nsv = __shfl(sv, g, 64);
          if (mine) {
            vx = nx;
            vy = ny;
            vz = nz;
            vs = nsv;
          }
        }
        if (mk) vd = vcn | ((hi ? pidx[1] : pidx[0]) << 16);
        const uint64_t detm = wave_ballot(!ok), nanm = wave_ballot(isnan(newr));
#pragma unroll
        for (int h = 0; h < 2; h++) {
          if (!clip[h]) continue;
          const uint64_t hm = 0xFFFFFFFFull << (32 * h);
          if (detm & hm) {
            state[h] = CS_ERROR;
            err[h] = E_DET_ZERO;
            run[h] = false;
          } else if (nanm & hm) {
            state[h] = CS_ERROR;
            err[h] = E_NAN;
            run[h] = false;
          } else {
            nv[h] = numv[h] + clen[h];
            jc[h]++;
          }
        }
      }
    }

    // ---- safety radii (update_safety_radius, convex_cell.rs:443-453: the
    // largest 2 |v - q|); a list used up: the first entry left out decides
    // (its distance comes with the list), else the whole stream was used
    double sr[2];
    {
      const int nvl = hi ? nv[1] : nv[0];
      const double hm = half_max_nonneg(sl < nvl ? vs : 0.);
      sr[0] = rl_d(hm, 31);
      sr[1] = rl_d(hm, 63);
    }
#pragma unroll
    for (int h = 0; h < 2; h++) {
      if (!exh[h] || state[h] != CS_DONE) continue;
      if ((uint64_t)cnt[h] >= tstream[h])
        npop[h] = cnt[h];
      else if (sr[h] < kd[h])
        npop[h] = cnt[h] + 1u;
      else
        state[h] = CS_NEED_MORE;
    }

    // ---- decomposition and integrals (convex_cell.rs:131-187,
    // voronoi_cell.rs:40-100, integrals.rs:91-105, voronoi_face.rs:30-43)
    lds_fence();  // the entry staging area becomes the tet staging area
    bool dec[2];
#pragma unroll
    for (int h = 0; h < 2; h++) dec[h] = live[h] && state[h] == CS_DONE;
    uint64_t facemask = 0;
    double fa = 0., fx = 0., fy = 0., fz = 0.;
    double vol[2] = {0., 0.}, cen[2][3] = {{0., 0., 0.}, {0., 0., 0.}};
    if (dec[0] || dec[1]) {
      const bool dl = hi ? dec[1] : dec[0];
      const int npl = hi ? np[1] : np[0], nvl = hi ? nv[1] : nv[0];
      const uint64_t idxl = hi ? idx[1] : idx[0];
      // which planes build a face (voronoi_cell.rs:53-73)
      bool cons = false;
      d3 pfl = mk3(0., 0., 0.);
      if (dl && sl < npl) {
        const double4 a = S.pa[lane];
        const uint64_t ref = S.pref[lane];
        cons = true;
        if (dim == 1 && !(a.y == 0. && a.z == 0.)) cons = false;
        if (dim == 2 && !(a.z == 0.)) cons = false;
        if (cons && ref != BOX_REF && pk_img(ref) == 13) {
          const uint32_t right = pk_idx(ref);
          cons = (uint64_t)right > idxl || (A.mask != nullptr && !A.mask[right]);
        }
        // the projection of q onto the plane (the face's first tet point)
        pfl = plane_project_onto(xyz(a), p_plane_point(S, A, B, q, lane), q);
      }
      // chunks of PCH vertices per half: lane 32 h + 3 i + k evaluates the
      // two tets on plane k of the chunk's vertex i and stages them; then
      // lane 32 h + c sums component c of the cell and lane 32 h + p the
      // face on plane p, in the reference's order (vertex order, tet order)
      double cacc = 0.;
      const double *cterm =
          reinterpret_cast<const double *>(&S.u.t.fcv[0][hi ? PCH : 0]) + (sl & 3);
      bool has = false, lok_all = true;
      const int nvmax = max(dec[0] ? nv[0] : 0, dec[1] ? nv[1] : 0);
      for (int c0 = 0; c0 < nvmax; c0 += PCH) {
        int dl_ = sl;
        asm volatile("" : "+v"(dl_));
        const int si = dl_ / 3, kk = dl_ - 3 * si;
        const int c1 = min(nvl, c0 + PCH);
        const bool mine = dl && si < PCH && c0 + si < c1;
        const int vsrc = hb + min(c0 + si, 31);
        const d3 vl = mk3(__shfl(vx, vsrc, 64), __shfl(vy, vsrc, 64), __shfl(vz, vsrc, 64));
        const int du = __shfl(vd, vsrc, 64);
        const int pk = (du >> (8 * kk)) & 0xFF;
        const int pn2 = (du >> (kk == 2 ? 0 : 8 * kk + 8)) & 0xFF;
        d3 pedge = mk3(0., 0., 0.);
        if (mine) {
          const double4 ak = S.pa[hb + pk], an = S.pa[hb + pn2];
          lok_all &= project_onto_intersection(xyz(an), an.w, xyz(ak), ak.w, q, &pedge);
        }
        const int psrc = min(kk == 0 ? lane + 2 : lane - 1, 63);
        const d3 prev = mk3(__shfl(pedge.x, psrc, 64), __shfl(pedge.y, psrc, 64),
                            __shfl(pedge.z, psrc, 64));
        const int fsrc = hb + (pk & 31);
        const d3 pface =
            mk3(__shfl(pfl.x, fsrc, 64), __shfl(pfl.y, fsrc, 64), __shfl(pfl.z, fsrc, 64));
        if (mine) {
          const int col = (hi ? PCH : 0) + si;
#pragma unroll
          for (int h2 = 0; h2 < 2; h2++) {
            const d3 v0 = h2 == 0 ? pface : pedge;
            const d3 v1 = h2 == 0 ? prev : pface;
            const double tv = signed_volume_tet(v0, v1, vl, q);
            const d3 tc = scl3(tv, add3(add3(add3(v0, v1), vl), q));
            const double ar = signed_area_tri(v0, v1, vl, q);
            const d3 sc = scl3(ar, add3(add3(v0, v1), vl));
            S.u.t.fcv[2 * kk + h2][col] = make_double4(tv, tc.x, tc.y, tc.z);
            S.u.t.fcf[2 * kk + h2][col] = make_double4(ar, sc.x, sc.y, sc.z);
          }
        }
        lds_fence();
        const int cend = min(nvmax, c0 + PCH);
        for (int s2 = c0; s2 < cend; s2++) {
          const int sv = s2 - c0;
          const bool inr = dl && s2 < nvl;
          if (inr) {
            // VolumeCentroidIntegral::collect (integrals.rs:91-95), 6 tets
#pragma unroll
            for (int t = 0; t < 6; t++) cacc += cterm[(t * 2 * PCH + sv) * 4];
          }
          // VoronoiFaceIntegral::collect for the two tets on plane sl
          const int dv0 = rl_i(vd, s2), dv1 = rl_i(vd, 32 + s2);
          const int dv = hi ? dv1 : dv0;
          const int k = (dv & 0xFF) == sl ? 0
                        : ((dv >> 8) & 0xFF) == sl ? 1
                        : ((dv >> 16) & 0xFF) == sl ? 2 : -1;
          if (inr && k >= 0) {
            const int col = (hi ? PCH : 0) + sv;
            const double4 f0 = S.u.t.fcf[2 * k][col], f1 = S.u.t.fcf[2 * k + 1][col];
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
      const uint64_t badm = wave_ballot(!lok_all);
      facemask = wave_ballot(dl && sl < npl && cons && has);
#pragma unroll
      for (int h = 0; h < 2; h++) {
        if (!dec[h]) continue;
        if ((badm >> (32 * h)) & 0xFFFFFFFFull) {
          state[h] = CS_ERROR;
          err[h] = E_DET_ZERO;
          dec[h] = false;
          facemask &= ~(0xFFFFFFFFull << (32 * h));
          continue;
        }
        vol[h] = rl_d(cacc, 32 * h);
        cen[h][0] = rl_d(cacc, 32 * h + 1);
        cen[h][1] = rl_d(cacc, 32 * h + 2);
        cen[h][2] = rl_d(cacc, 32 * h + 3);
      }
    }

    // ---- outcomes, cell records and staged faces
    unsigned long long base[2] = {0ull, 0ull};
    uint32_t nf[2] = {0u, 0u};
#pragma unroll
    for (int h = 0; h < 2; h++) {
      if (!live[h]) continue;
      nf[h] = (uint32_t)__popcll((facemask >> (32 * h)) & 0xFFFFFFFFull);
      if (lane == 32 * h) {
        A.wstate[w[h]] = state[h];
        if (state[h] == CS_NEED_MORE) A.need_r[w[h]] = sr[h];
        if (state[h] == CS_EXACT) A.exact_list[atomicAdd(A.exact_cnt, 1u)] = w[h];
        if (state[h] == CS_ERROR)
          atomicMin(A.err_word, (unsigned long long)((idx[h] << 8) | err[h]));
        if (state[h] == CS_DONE) {
          // VolumeCentroidIntegral::finalize (integrals.rs:97-105)
          const double norm = vol[h] > 0. ? 0.25 / vol[h] : 0.;
          const uint64_t id = idx[h];
          A.cell_vol[id] = vol[h];
          A.cell_cen[3 * id] = cen[h][0] * norm;
          A.cell_cen[3 * id + 1] = cen[h][1] * norm;
          A.cell_cen[3 * id + 2] = cen[h][2] * norm;
          A.cell_sr[id] = sr[h];
          A.cell_npop[id] = npop[h];
          base[h] = nf[h] ? atomicAdd(A.st_counter, (unsigned long long)nf[h]) : 0ull;
          A.cell_nf[id] = nf[h];
          A.cell_fbase[id] = base[h];
        }
      }
      if (dec[h]) {
        base[h] = rl_u64(base[h], 32 * h);
        if (base[h] + nf[h] > A.st_cap) {
          if (lane == 32 * h) atomicOr(A.overflow, 1ull);
          facemask &= ~(0xFFFFFFFFull << (32 * h));
        }
      }
    }
    if ((facemask >> lane) & 1ull) {
      // the faces of lower lanes of the same half
      const uint64_t hmask = facemask & (hi ? 0xFFFFFFFF00000000ull : 0xFFFFFFFFull);
      const uint64_t o = (hi ? base[1] : base[0]) +
                         (uint64_t)__builtin_amdgcn_mbcnt_hi(
                             (uint32_t)(hmask >> 32),
                             __builtin_amdgcn_mbcnt_lo((uint32_t)hmask, 0u));
      const double4 a = S.pa[lane];
      const uint64_t ref = S.pref[lane];
      const double fnrm = fa > 0. ? 1. / (3. * fa) : 0.;
      A.st_cell[o] = (uint32_t)(hi ? idx[1] : idx[0]);
      A.st_right[o] = ref == BOX_REF ? -1 : (int64_t)pk_idx(ref);
      // a face without periodic shift is also a link of its right cell (the
      // count's old value is the link's slot in that cell's range)
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
    next = (uint32_t)rl_i((int)nxq, 0);
    lds_fence();
  }
}

}  // namespace mv
