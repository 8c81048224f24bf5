// mv_learn.hpp -- learned neighbour-count proposal.
//
// A 16 -> 64 -> 64 -> 2 MLP (ReLU) reads local-density features of each
// seed and proposes, per cell, how many neighbour-stream entries the
// construction will consume (log k) and the final safety radius (log of
// r / local spacing).  Weights are BF16 and every layer runs on MFMA
// (v_mfma_f32_16x16x16_bf16, FP32 accumulation), in training and inference.
// The proposal only sizes the first neighbour list of a cell: the security
// test still decides every cell, and a cell whose list was too short is
// rebuilt with a doubled list, so the output does not depend on the model.
#pragma once
#include <type_traits>

#include "mv_common.hpp"
#include "mv_grid.hpp"  // zkey: the order of a grid row

namespace mv {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int NF = 16;   // features
constexpr int NH = 64;   // hidden width
constexpr int NO = 2;    // outputs
constexpr int NOP = 16;  // outputs padded to one MFMA tile
constexpr uint32_t FEAT_WIN = 64;  // own-cell window that bounds the 8th distance

__device__ __forceinline__ short f2bf(float x) {
  uint32_t b = __float_as_uint(x);
  if ((b & 0x7fffffffu) > 0x7f800000u) return (short)0x7fc0;  // NaN
  b += 0x7fffu + ((b >> 16) & 1u);  // round to nearest even
  return (short)(b >> 16);
}
__device__ __forceinline__ float bf2f(short h) {
  return __uint_as_float(((uint32_t)(uint16_t)h) << 16);
}

// ---------------------------------------------------------------------
// features
// ---------------------------------------------------------------------
struct FeatParams {
  double box_lo[3], box_hi[3];  // simulation box (before periodic widening)
  double h_global;              // (box volume / N)^(1/nd)
};

// 16 features of the seed in sorted slot work[w]:
//  0-7  distances to the 8 nearest seeds found in the 3^d block of grid
//       cells around it (periodic images included), over the local spacing
//  8    log(1 + seeds in its own grid cell)
//  9    log(1 + seeds in the 3^d block)
//  10   log(1 + seeds in the 5^d shell outside the block)
//  11   log(local spacing / global spacing)
//  12   |mean offset of the 8 nearest| / their mean distance
//  13   distance to the box boundary over the local spacing (4 if periodic)
//  14   1 if periodic
//  15   spread of the 8 distances (max - min) / mean
// hloc[w] = local spacing (block volume / block count)^(1/nd).
//
// Two instantiations share the work: DENSE = true takes the waves with a
// seed whose grid cell holds more than FEAT_WIN seeds (dense clusters),
// DENSE = false every other wave; both evaluate the same wave-uniform test
// and return at once on the other's waves (one launch each: in one kernel
// the dense path's registers would cost the common path its occupancy).
//
// A dense wave whose 64 seeds all lie in one grid cell computes the
// features of its first seed, the representative, and gives them to all 64
// (they share the cell, so their counts, spacing and neighbourhood are
// alike): the lanes scan the representative's candidates 64 at a time,
// each keeping its own 8 nearest, and the 64 lists are merged by a
// butterfly.  The features only size the first neighbour lists, so the
// output of a build does not depend on them.
template <bool DENSE>
__global__ void __launch_bounds__(256)
    k_features(const double4 *__restrict__ spos,
               const uint32_t *__restrict__ cell_start_all, Problems PR,
               const FeatParams *__restrict__ fps, const uint32_t *__restrict__ work,
               uint32_t nwork, float *__restrict__ F,
               float *__restrict__ hloc) {
  const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nwork) return;
  uint32_t slot = work[w];
  const uint32_t pr = problem_of(PR, slot);
  const GridParams &g = PR.grids[pr];
  const BoxParams &box = PR.boxes[pr];
  const FeatParams &fp = fps[pr];
  const uint32_t *cell_start = cell_start_all + PR.coff[pr];
  const double4 qq = spos[slot];
  double q[3] = {qq.x, qq.y, qq.z};
  int C[3], nd = 0;
  double cellvol = 1.;
  for (int c = 0; c < 3; c++) {
    const bool act = g.inv_h[c] != 0.;
    nd += act;
    if (act) cellvol *= g.h[c];
    double f = floor((q[c] - g.lo[c]) * g.inv_h[c]);
    C[c] = act ? (int)fmax(0., fmin(f, (double)(g.G[c] - 1))) : 0;
  }
  if (nd == 0) nd = 1;
  double best[8];
  double off[3] = {0., 0., 0.};
  double boff[8][3];
#pragma unroll
  for (int i = 0; i < 8; i++) {
    best[i] = INFINITY;
    boff[i][0] = boff[i][1] = boff[i][2] = 0.;
  }
  uint32_t n1 = 0, n27 = 0, n125 = 0;
  // Dense cells (clusters put thousands of seeds in one grid cell): a bound
  // B on the final 8th squared distance cuts the scan of each row to the z
  // range |dz| < sqrt(B) (rows are in zkey order).  Only seeds whose squared
  // distance exceeds the final 8th are skipped, in the scan's own order, so
  // the result is that of the full scan bit for bit.  B is the 8th smallest
  // squared distance to FEAT_WIN seeds of the own cell next to it in z
  // order, then tightened to the 8th smallest over the own cell's seeds
  // within that first range (both are distances the scan itself computes).
  double zr = INFINITY;
  uint32_t s0 = 0, s1 = 0;
  if (g.inv_h[2] != 0.) {
    const size_t orow = ((size_t)C[0] * g.G[1] + C[1]) * g.G[2] + C[2];
    s0 = cell_start[orow];
    s1 = cell_start[orow + 1];
  }
  const bool dense_lane = g.inv_h[2] != 0. && s1 - s0 > FEAT_WIN;
  if (__any(dense_lane) != DENSE) return;
  // A one-cell wave (its 64 lanes share one grid cell of one problem) scans
  // the same rows on every lane; the union of their z ranges keeps every
  // lane on the same entry at each step (one coalesced load per 64 entries,
  // each broadcast with readlane).  Widening a lane's range only adds
  // entries it would have tested anyway.
  bool uni = false;
  if constexpr (DENSE) {
    const bool full = __popcll(wave_ballot(1)) == 64;
    const uint32_t fp0 = __builtin_amdgcn_readfirstlane(pr);
    const int f0 = __builtin_amdgcn_readfirstlane(C[0]);
    const int f1 = __builtin_amdgcn_readfirstlane(C[1]);
    const int f2 = __builtin_amdgcn_readfirstlane(C[2]);
    uni = __all(full && pr == fp0 && C[0] == f0 && C[1] == f1 && C[2] == f2);
    if (uni) {
      // the representative: lane 0's seed (same cell on every lane)
      slot = (uint32_t)__builtin_amdgcn_readfirstlane((int)slot);
      for (int c = 0; c < 3; c++) q[c] = __shfl(q[c], 0, 64);
    }
  }
  // the 8 smallest of this lane's and every other lane's sorted 8 (with
  // their offsets when O), by a butterfly, then lane 0's for all lanes so
  // that equal distances end in one order everywhere
  auto merge8 = [&](double(&b)[8], double(&bo)[8][3], auto with_off) {
    constexpr bool O = decltype(with_off)::value;
    for (int m = 1; m < 64; m <<= 1) {
      double pb[8], po[8][3];
#pragma unroll
      for (int i = 0; i < 8; i++) {
        pb[i] = __shfl_xor(b[i], m, 64);
        if constexpr (O)
          for (int c = 0; c < 3; c++) po[i][c] = __shfl_xor(bo[i][c], m, 64);
      }
#pragma unroll
      for (int j = 0; j < 8; j++) {
        double v = pb[j];
        double vo[3] = {0., 0., 0.};
        if constexpr (O)
          for (int c = 0; c < 3; c++) vo[c] = po[j][c];
#pragma unroll
        for (int i = 0; i < 8; i++) {
          if (v < b[i]) {
            const double tv = b[i];
            b[i] = v;
            v = tv;
            if constexpr (O)
              for (int c = 0; c < 3; c++) {
                const double to = bo[i][c];
                bo[i][c] = vo[c];
                vo[c] = to;
              }
          }
        }
      }
    }
#pragma unroll
    for (int i = 0; i < 8; i++) {
      b[i] = __shfl(b[i], 0, 64);
      if constexpr (O)
        for (int c = 0; c < 3; c++) bo[i][c] = __shfl(bo[i][c], 0, 64);
    }
  };
  double zlo = -INFINITY, zhi = INFINITY;
  auto widen = [&]() {
    zlo = q[2] - zr;
    zhi = q[2] + zr;
    if (uni)
      for (int o = 32; o > 0; o >>= 1) {
        zlo = fmin(zlo, __shfl_xor(zlo, o, 64));
        zhi = fmax(zhi, __shfl_xor(zhi, o, 64));
      }
  };
  // [t0, t1) of a zkey-ordered range cut to z in [zlo, zhi] (unshifted z)
  auto zcut = [&](uint32_t &t0, uint32_t &t1) {
    if (!(zr < INFINITY) || t1 - t0 <= 32u) return;
    const uint32_t klo = zkey(zlo), khi = zkey(zhi);
    uint32_t lo = t0, hi = t1;
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (zkey(spos[mid].z) < klo)
        lo = mid + 1;
      else
        hi = mid;
    }
    uint32_t lo2 = lo, hi2 = t1;
    while (lo2 < hi2) {
      const uint32_t mid = (lo2 + hi2) >> 1;
      if (zkey(spos[mid].z) <= khi)
        lo2 = mid + 1;
      else
        hi2 = mid;
    }
    t0 = lo;
    t1 = lo2;
  };
  // fn(p, t) for the entries t of [t0, t1): in a one-cell wave each lane
  // takes every 64th entry (coalesced) for the representative; otherwise
  // one lane per seed, in index order
  auto each = [&](uint32_t t0, uint32_t t1, auto &&fn) {
    if (DENSE && uni) {
      for (uint32_t t = t0 + __lane_id(); t < t1; t += 64) fn(spos[t], t);
    } else {
      for (uint32_t t = t0; t < t1; t++) fn(spos[t], t);
    }
  };
  double nooff[8][3];
  if (DENSE && dense_lane) {
    double b8[8];
#pragma unroll
    for (int i = 0; i < 8; i++) b8[i] = INFINITY;
    auto keep8 = [&](const double4 p, uint32_t t) {
      if (t == slot) return;
      const double o[3] = {p.x - q[0], p.y - q[1], p.z - q[2]};
      double v = (o[0] * o[0] + o[1] * o[1]) + o[2] * o[2];
      if (!(v < b8[7])) return;
#pragma unroll
      for (int i = 0; i < 8; i++) {
        if (v < b8[i]) {
          const double tv = b8[i];
          b8[i] = v;
          v = tv;
        }
      }
    };
    uint32_t a0 = slot >= s0 + FEAT_WIN / 2 ? slot - FEAT_WIN / 2 : s0;
    if (a0 + FEAT_WIN > s1) a0 = s1 - FEAT_WIN;
    each(a0, a0 + FEAT_WIN, keep8);
    if (uni) merge8(b8, nooff, std::false_type());
    // |z difference| >= zr gives a squared distance > b8[7] after rounding
    zr = sqrt(b8[7]) * (1. + 1e-6);
    widen();
    // the range holds the window's 8 nearest: start afresh so that none of
    // them is counted twice
#pragma unroll
    for (int i = 0; i < 8; i++) b8[i] = INFINITY;
    uint32_t u0 = s0, u1 = s1;
    zcut(u0, u1);
    each(u0, u1, keep8);
    if (uni) merge8(b8, nooff, std::false_type());
    zr = sqrt(b8[7]) * (1. + 1e-6);
    widen();
  } else if (g.inv_h[2] != 0. && s1 - s0 > 8u) {
    // a lane of a sparse cell with at least 8 other seeds: the 8th nearest
    // of those bounds the final 8th, and the rows' scan of the dense cells
    // next to it (rows of more than 32 seeds) is cut to |dz| below it, as
    // above; the result is that of the full scan
    double b8[8];
#pragma unroll
    for (int i = 0; i < 8; i++) b8[i] = INFINITY;
    for (uint32_t t = s0; t < s1; t++) {
      if (t == slot) continue;
      const double4 p = spos[t];
      const double o[3] = {p.x - q[0], p.y - q[1], p.z - q[2]};
      double v = (o[0] * o[0] + o[1] * o[1]) + o[2] * o[2];
      if (!(v < b8[7])) continue;
#pragma unroll
      for (int i = 0; i < 8; i++) {
        if (v < b8[i]) {
          const double tv = b8[i];
          b8[i] = v;
          v = tv;
        }
      }
    }
    zr = sqrt(b8[7]) * (1. + 1e-6);
    zlo = q[2] - zr;
    zhi = q[2] + zr;
  }
  // entries [t0, t1) with image shift (sx, sy, sz) into the sorted 8
  // (ZS = false: no shift along z)
  auto take = [&](const double4 p, double sx, double sy, double sz, auto zs) {
    const double o[3] = {p.x + sx - q[0], p.y + sy - q[1],
                         decltype(zs)::value ? p.z + sz - q[2] : p.z - q[2]};
    const double d2 = (o[0] * o[0] + o[1] * o[1]) + o[2] * o[2];
    if (!(d2 < best[7])) return;
    // insertion into the sorted 8 (unrolled, registers only)
    double v = d2;
    double vo[3] = {o[0], o[1], o[2]};
#pragma unroll
    for (int i = 0; i < 8; i++) {
      if (v < best[i]) {
        double tv = best[i];
        best[i] = v;
        v = tv;
        for (int c = 0; c < 3; c++) {
          double to = boff[i][c];
          boff[i][c] = vo[c];
          vo[c] = to;
        }
      }
    }
  };
  auto scan = [&](uint32_t t0, uint32_t t1, double sx, double sy, double sz, auto zs) {
    each(t0, t1, [&](const double4 p, uint32_t t) {
      if (t != slot) take(p, sx, sy, sz, zs);
    });
  };
  // cell index along axis c of offset dd from C, with the periodic wrap of
  // 5+ cell axes (shift into *sh); false when the cell does not exist
  auto axis_cell = [&](int c, int dd, int *cc, double *sh) -> bool {
    *sh = 0.;
    if (g.inv_h[c] == 0.) {
      *cc = 0;
      return dd == 0;
    }
    int x = C[c] + dd;
    if (x < 0 || x >= g.G[c]) {
      if (!box.periodic || g.G[c] < 5) return false;
      if (x < 0) {
        x += g.G[c];
        *sh = -box.width[c];
      } else {
        x -= g.G[c];
        *sh = box.width[c];
      }
    }
    *cc = x;
    return true;
  };
  for (int da = -2; da <= 2; da++) {
    int ca;
    double sa;
    if (!axis_cell(0, da, &ca, &sa)) continue;
    for (int db = -2; db <= 2; db++) {
      int cb;
      double sb;
      if (!axis_cell(1, db, &cb, &sb)) continue;
      const size_t row = ((size_t)ca * g.G[1] + cb) * g.G[2];
      const bool inner_row = abs(da) <= 1 && abs(db) <= 1;
      // cells C2-2 .. C2+2 of the row: one range of the sorted seeds when
      // none wraps (cell_start is a prefix count along the row)
      const bool flat = g.inv_h[2] != 0. && C[2] - 2 >= 0 && C[2] + 2 < g.G[2];
      if (flat) {
        const uint32_t e0 = cell_start[row + C[2] - 2], e1 = cell_start[row + C[2] - 1];
        const uint32_t e4 = cell_start[row + C[2] + 2], e5 = cell_start[row + C[2] + 3];
        const uint32_t self = (da == 0 && db == 0) ? 1u : 0u;
        n125 += e5 - e0 - self;
        if (!inner_row) continue;
        n27 += e4 - e1 - self;
        if (da == 0 && db == 0)
          n1 += cell_start[row + C[2] + 1] - cell_start[row + C[2]] - 1u;
        // the three inner cells in order: one range, no shift along c
        uint32_t t0 = e1, t1 = e4;
        zcut(t0, t1);
        scan(t0, t1, sa, sb, 0., std::false_type());
        continue;
      }
      for (int dc = -2; dc <= 2; dc++) {
        int cc2;
        double sc;
        if (!axis_cell(2, dc, &cc2, &sc)) continue;
        const uint32_t t0 = cell_start[row + cc2], t1 = cell_start[row + cc2 + 1];
        const bool center = da == 0 && db == 0 && dc == 0;
        const uint32_t cntc = t1 - t0 - (center ? 1u : 0u);
        n125 += cntc;
        if (!inner_row || abs(dc) > 1) continue;
        n27 += cntc;
        if (center) n1 += cntc;
        uint32_t u0 = t0, u1 = t1;
        if (sc == 0.) zcut(u0, u1);
        scan(u0, u1, sa, sb, sc, std::true_type());
      }
    }
  }
  if constexpr (DENSE)
    if (uni) merge8(best, boff, std::true_type());
  double blocks = 1.;
  for (int c = 0; c < nd; c++) blocks *= 3.;
  const double h = (n27 > 0) ? pow(blocks * cellvol / (double)n27, 1.0 / nd)
                             : fp.h_global * 3.;
  float *f = F + (size_t)w * NF;
  double dsum = 0., dmin = INFINITY, dmax = 0.;
  int nfound = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    double di = best[i] < INFINITY ? sqrt(best[i]) / h : 3.;
    f[i] = (float)fmin(di, 8.);
    if (best[i] < INFINITY) {
      nfound++;
      dsum += di;
      dmin = fmin(dmin, di);
      dmax = fmax(dmax, di);
      for (int c = 0; c < 3; c++) off[c] += boff[i][c];
    }
  }
  const double dmean = nfound ? dsum / nfound : 1.;
  f[8] = (float)log1p((double)n1);
  f[9] = (float)log1p((double)n27);
  f[10] = (float)log1p((double)(n125 - n27));
  f[11] = (float)log(h / fp.h_global);
  const double om = sqrt(off[0] * off[0] + off[1] * off[1] + off[2] * off[2]);
  f[12] = nfound ? (float)(om / (nfound * h) / fmax(dmean, 1e-30)) : 0.f;
  double bd = 4.;
  if (!box.periodic) {
    for (int c = 0; c < 3; c++) {
      if (g.inv_h[c] == 0.) continue;
      bd = fmin(bd, (q[c] - fp.box_lo[c]) / h);
      bd = fmin(bd, (fp.box_hi[c] - q[c]) / h);
    }
  }
  f[13] = (float)fmax(0., bd);
  f[14] = box.periodic ? 1.f : 0.f;
  f[15] = nfound ? (float)((dmax - dmin) / fmax(dmean, 1e-30)) : 0.f;
  hloc[w] = (float)h;
}

// ---------------------------------------------------------------------
// model (device copy of the checkpoint)
// ---------------------------------------------------------------------
struct DevModel {
  float mu[NF], sigma[NF];
  short w1[NF * NH];    // bf16 [in][out]
  short w2[NH * NH];
  short w3[NH * NOP];   // outputs padded with zeros
  float b1[NH], b2[NH], b3[NOP];
};

// Fused inference: each wavefront runs 16 rows at a time through the three
// layers; the accumulator tile (lane l holds rows 4*(l/16)+v of column
// l%16) is transposed through LDS into the next layer's A operand (lane l
// holds row l%16, k = 4*(l/16)+v).
__global__ void __launch_bounds__(256)
    k_mlp_infer(const DevModel *__restrict__ M, const float *__restrict__ F,
                uint32_t n, float *__restrict__ Y) {
  __shared__ short sw1[NF * NH], sw2[NH * NH], sw3[NH * NOP];
  __shared__ float sb1[NH], sb2[NH], sb3[NOP], smu[NF], ssg[NF];
  __shared__ float tbuf[4][16][NH + 1];
  for (int i = threadIdx.x; i < NF * NH; i += blockDim.x) sw1[i] = M->w1[i];
  for (int i = threadIdx.x; i < NH * NH; i += blockDim.x) sw2[i] = M->w2[i];
  for (int i = threadIdx.x; i < NH * NOP; i += blockDim.x) sw3[i] = M->w3[i];
  for (int i = threadIdx.x; i < NH; i += blockDim.x) {
    sb1[i] = M->b1[i];
    sb2[i] = M->b2[i];
  }
  for (int i = threadIdx.x; i < NOP; i += blockDim.x) sb3[i] = M->b3[i];
  for (int i = threadIdx.x; i < NF; i += blockDim.x) {
    smu[i] = M->mu[i];
    ssg[i] = M->sigma[i];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int r = lane & 15, kq = (lane >> 4) * 4;
  float(*T)[NH + 1] = tbuf[wv];
  const uint32_t tiles = (n + 15) / 16;
  const uint32_t nw = gridDim.x * 4;
  for (uint32_t tile = blockIdx.x * 4 + wv; tile < tiles; tile += nw) {
    const uint32_t row = tile * 16 + r;
    bf16x4 a;
#pragma unroll
    for (int v = 0; v < 4; v++) {
      float x = row < n ? (F[(size_t)row * NF + kq + v] - smu[kq + v]) / ssg[kq + v] : 0.f;
      a[v] = f2bf(x);
    }
    // layer 1: 16 x 16 @ 16 x 64
#pragma unroll
    for (int nt = 0; nt < 4; nt++) {
      bf16x4 b;
#pragma unroll
      for (int v = 0; v < 4; v++) b[v] = sw1[(kq + v) * NH + nt * 16 + r];
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
#pragma unroll
      for (int v = 0; v < 4; v++) T[kq + v][nt * 16 + r] = fmaxf(acc[v] + sb1[nt * 16 + r], 0.f);
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    // layer 2: 16 x 64 @ 64 x 64
    bf16x4 a2[4];
#pragma unroll
    for (int kk = 0; kk < 4; kk++)
#pragma unroll
      for (int v = 0; v < 4; v++) a2[kk][v] = f2bf(T[r][kk * 16 + kq + v]);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int nt = 0; nt < 4; nt++) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kk = 0; kk < 4; kk++) {
        bf16x4 b;
#pragma unroll
        for (int v = 0; v < 4; v++) b[v] = sw2[(kk * 16 + kq + v) * NH + nt * 16 + r];
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a2[kk], b, acc, 0, 0, 0);
      }
#pragma unroll
      for (int v = 0; v < 4; v++) T[kq + v][nt * 16 + r] = fmaxf(acc[v] + sb2[nt * 16 + r], 0.f);
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    // layer 3: 16 x 64 @ 64 x 16 (2 used)
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kk = 0; kk < 4; kk++) {
      bf16x4 a3, b;
#pragma unroll
      for (int v = 0; v < 4; v++) {
        a3[v] = f2bf(T[r][kk * 16 + kq + v]);
        b[v] = sw3[(kk * 16 + kq + v) * NOP + r];
      }
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a3, b, acc, 0, 0, 0);
    }
    if (r < NO) {
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const uint32_t orow = tile * 16 + kq + v;
        if (orow < n) Y[(size_t)orow * NO + r] = acc[v] + sb3[r];
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
}

// ---------------------------------------------------------------------
// training kernels
// ---------------------------------------------------------------------
// C[M x N] (+)= A[M x K] B[K x N]; A(i,k) = A[i*sa0 + k*sa1], B(k,j) =
// B[k*sb0 + j*sb1]; FP32 in memory, rounded to BF16 for the MFMA, FP32
// accumulation.  M, N, K multiples of 16; one wavefront per 16 x 16 tile.
__global__ void __launch_bounds__(256)
    k_gemm_bf16(const float *__restrict__ A, long sa0, long sa1,
                const float *__restrict__ B, long sb0, long sb1,
                float *__restrict__ C, int ldc, int M, int N, int K,
                int accumulate) {
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  const int tn = N / 16, tiles = (M / 16) * tn;
  if (wave >= tiles) return;
  const int ti = wave / tn, tj = wave % tn;
  const int r = lane & 15, kq = (lane >> 4) * 4;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += 16) {
    bf16x4 a, b;
#pragma unroll
    for (int v = 0; v < 4; v++) {
      a[v] = f2bf(A[(long)(ti * 16 + r) * sa0 + (long)(k0 + kq + v) * sa1]);
      b[v] = f2bf(B[(long)(k0 + kq + v) * sb0 + (long)(tj * 16 + r) * sb1]);
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  }
#pragma unroll
  for (int v = 0; v < 4; v++) {
    const long idx = (long)(ti * 16 + kq + v) * ldc + tj * 16 + r;
    C[idx] = (accumulate ? C[idx] : 0.f) + acc[v];
  }
}

// Z += bias (per column); A = relu(Z)
__global__ void k_bias_relu(float *__restrict__ Z, float *__restrict__ Aout,
                            const float *__restrict__ b, int rows, int cols,
                            int relu) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows * cols) return;
  const float z = Z[i] + b[i % cols];
  Z[i] = z;
  if (Aout) Aout[i] = relu ? fmaxf(z, 0.f) : z;
}

// pinball loss at quantile tau on the first NO columns of Y (ld NOP);
// writes dY (zero in the padding) and per-row loss.
__global__ void k_pinball(const float *__restrict__ Y,
                          const float *__restrict__ Tg, int rows, float tau,
                          float *__restrict__ dY, float *__restrict__ lrow) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows) return;
  float l = 0.f;
  for (int c = 0; c < NOP; c++) {
    float g = 0.f;
    if (c < NO) {
      const float e = Tg[i * NO + c] - Y[i * NOP + c];
      l += e > 0.f ? tau * e : (tau - 1.f) * e;
      g = (e > 0.f ? -tau : (1.f - tau)) / (float)rows;
    }
    dY[i * NOP + c] = g;
  }
  lrow[i] = l;
}

// dZ = dA * (Z > 0)
__global__ void k_relu_grad(const float *__restrict__ Z,
                            float *__restrict__ dA, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && !(Z[i] > 0.f)) dA[i] = 0.f;
}

// column sums of a rows x cols matrix
__global__ void k_colsum(const float *__restrict__ X, int rows, int cols,
                         float *__restrict__ out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cols) return;
  float s = 0.f;
  for (int r = 0; r < rows; r++) s += X[r * cols + c];
  out[c] = s;
}

// Adam on FP32 master weights
__global__ void k_adam(float *__restrict__ w, const float *__restrict__ g,
                       float *__restrict__ m, float *__restrict__ v, int n,
                       float lr, float b1, float b2, float eps, float c1,
                       float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// gather training rows: X[b] = (F[idx[b]] - mu) / sigma, T[b] = targets
__global__ void k_gather_rows(const float *__restrict__ F,
                              const float *__restrict__ Tg,
                              const uint32_t *__restrict__ idx, int rows,
                              const float *__restrict__ mu,
                              const float *__restrict__ sg,
                              float *__restrict__ X, float *__restrict__ T) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows * NF) return;
  const int r = i / NF, c = i % NF;
  const uint32_t s = idx[r];
  X[i] = (F[(size_t)s * NF + c] - mu[c]) / sg[c];
  if (c < NO) T[r * NO + c] = Tg[(size_t)s * NO + c];
}

// ---------------------------------------------------------------------
// proposal -> per-cell list length and start radius
// ---------------------------------------------------------------------
__global__ void k_proposal(const float *__restrict__ Y,
                           const float *__restrict__ hloc, uint32_t n,
                           uint32_t lmin, uint32_t lmax, float kscale,
                           float rscale, uint32_t *__restrict__ Lper,
                           float *__restrict__ Rper) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float k = expf(fminf(Y[(size_t)i * NO], 12.f)) * kscale + 2.f;
  uint32_t L = (uint32_t)ceilf(k);
  L = (L + 7u) & ~7u;  // 8-wide buckets
  L = L < lmin ? lmin : (L > lmax ? lmax : L);
  Lper[i] = L;
  // the proposed radius, kept within a factor 3 of the radius of a ball of
  // about 1.3 L + 8 seeds at the block density (inside dense clusters the
  // proposal alone misses more often, and a gather that misses is redone)
  const float Rm = expf(fminf(Y[(size_t)i * NO + 1], 8.f)) * hloc[i] * rscale;
  const float Rd = hloc[i] * cbrtf(0.2387324f * (1.3f * (float)L + 8.f));
  Rper[i] = fminf(fmaxf(Rm, 0.333f * Rd), 3.f * Rd);
}

}  // namespace mv
