// mvoro.hip -- host side of libmvoro: the C ABI of include/mvoro.h over the
// gfx950 kernels of mv_grid.hpp, mv_clip.hpp and mv_finalize.hpp.
//
// Pipeline of one build (all device-resident):
//   prepare + bounding box -> counting-sort grid -> [neighbour lists of
//   length L -> per-cell clip + integrals] repeated with L doubled for the
//   cells whose security radius was not reached -> face placement -> links.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "mvoro.h"
#include "mv_clip.hpp"
#include "mv_clipw.hpp"
#include "mv_common.hpp"
#include "mv_finalize.hpp"
#include "mv_grid.hpp"

using namespace mv;

namespace {

// ---------------------------------------------------------------------
// caching device allocator (device memory is reused across builds)
// ---------------------------------------------------------------------
struct Pool {
  std::mutex mu;
  std::multimap<size_t, void *> free_;
  std::unordered_map<void *, size_t> live_;
};
Pool &pool() {
  static Pool *p = new Pool();
  return *p;
}
size_t round_bytes(size_t b) {
  if (b < 512) return 512;
  if (b < (1u << 20)) return (b + 4095) & ~size_t(4095);
  return (b + (1u << 20) - 1) & ~size_t((1u << 20) - 1);
}
void *dalloc(size_t bytes) {
  Pool &P = pool();
  size_t sz = round_bytes(bytes);
  std::lock_guard<std::mutex> g(P.mu);
  auto it = P.free_.lower_bound(sz);
  if (it != P.free_.end() && it->first <= 2 * sz + (4u << 20)) {
    void *p = it->second;
    size_t s = it->first;
    P.free_.erase(it);
    P.live_[p] = s;
    return p;
  }
  void *p = nullptr;
  if (hipMalloc(&p, sz) != hipSuccess) {
    (void)hipGetLastError();
    for (auto &kv : P.free_) (void)hipFree(kv.second);
    P.free_.clear();
    if (hipMalloc(&p, sz) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
  }
  P.live_[p] = sz;
  return p;
}
void dfree(void *p) {
  if (!p) return;
  Pool &P = pool();
  std::lock_guard<std::mutex> g(P.mu);
  auto it = P.live_.find(p);
  if (it == P.live_.end()) return;
  P.free_.emplace(it->second, p);
  P.live_.erase(it);
}

template <typename T>
T *dnew(size_t n) {
  return (T *)dalloc(sizeof(T) * (n ? n : 1));
}

struct Temps {
  std::vector<void *> ptrs;
  template <typename T>
  T *get(size_t n) {
    T *p = dnew<T>(n);
    if (p) ptrs.push_back(p);
    return p;
  }
  ~Temps() {
    for (void *p : ptrs) dfree(p);
  }
};

struct BuildError {
  mv_status st;
};

#define HIPCHK(x)                                   \
  do {                                              \
    hipError_t e_ = (x);                            \
    if (e_ != hipSuccess) throw BuildError{MV_E_DEVICE}; \
  } while (0)

template <typename T>
T *need(T *p) {
  if (!p) throw BuildError{MV_E_OOM};
  return p;
}

inline unsigned nblk(size_t n, unsigned t = 256) {
  return (unsigned)((n + t - 1) / t);
}

// exclusive scan; returns the total
template <typename TIn, typename TOut>
TOut scan_excl(const TIn *in, TOut *out, size_t n, hipStream_t s) {
  if (n == 0) return 0;
  size_t tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
  Temps T;
  TOut *sums = need(T.get<TOut>(tiles));
  hipLaunchKernelGGL((k_scan_tile<TIn, TOut>), dim3((unsigned)tiles),
                     dim3(SCAN_THREADS), 0, s, in, n, out, sums);
  HIPCHK(hipGetLastError());
  if (tiles > 1) {
    TOut *offs = need(T.get<TOut>(tiles));
    scan_excl<TOut, TOut>(sums, offs, tiles, s);
    hipLaunchKernelGGL((k_scan_add<TOut>), dim3((unsigned)tiles),
                       dim3(SCAN_THREADS), 0, s, out, n, offs);
    HIPCHK(hipGetLastError());
  }
  TIn last_in;
  TOut last_out;
  HIPCHK(hipMemcpyAsync(&last_in, in + n - 1, sizeof(TIn), hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(&last_out, out + n - 1, sizeof(TOut), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return last_out + (TOut)last_in;
}

// compaction of a work list by per-item state
uint32_t compact(const uint32_t *work, const uint8_t *st, uint32_t n,
                 uint8_t want, uint32_t *out, hipStream_t s) {
  if (n == 0) return 0;
  Temps T;
  uint32_t *flag = need(T.get<uint32_t>(n));
  uint32_t *pos = need(T.get<uint32_t>(n));
  hipLaunchKernelGGL(k_flag_state, dim3(nblk(n)), dim3(256), 0, s, st, n, want, flag);
  uint32_t cnt = scan_excl<uint32_t, uint32_t>(flag, pos, n, s);
  if (cnt)
    hipLaunchKernelGGL(k_compact, dim3(nblk(n)), dim3(256), 0, s, work, flag, pos, n, out);
  HIPCHK(hipGetLastError());
  return cnt;
}

struct Timer {
  hipEvent_t a, b;
  hipStream_t s;
  explicit Timer(hipStream_t st) : s(st) {
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    (void)hipEventRecord(a, s);
  }
  double stop() {
    (void)hipEventRecord(b, s);
    (void)hipEventSynchronize(b);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return ms;
  }
};

int device_cus() {
  static int cus = 0;
  if (!cus) {
    int dev = 0;
    hipDeviceProp_t p;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipGetDeviceProperties(&p, dev) == hipSuccess)
      cus = p.multiProcessorCount;
    if (cus <= 0) cus = 256;
  }
  return cus;
}

constexpr int MAXP_MAIN = 64, MAXV_MAIN = 128;
constexpr int MAXP_BIG = 255, MAXV_BIG = 1024;

}  // namespace

// ---------------------------------------------------------------------
// the handle
// ---------------------------------------------------------------------
struct mv_tess {
  size_t n = 0, nf = 0, nl = 0;
  double anchor[3], width[3];
  uint32_t dim = 3, periodic = 0;
  std::vector<void *> dev;  // device allocations owned by the handle
  double *d_cell_loc = nullptr, *d_cell_cen = nullptr, *d_cell_vol = nullptr,
         *d_cell_sr = nullptr;
  uint64_t *d_cell_foff = nullptr, *d_cell_fcnt = nullptr,
           *d_cell_idx = nullptr;
  uint32_t *d_npop = nullptr;
  uint64_t *d_face_left = nullptr;
  int64_t *d_face_right = nullptr;
  double *d_face_area = nullptr, *d_face_cen = nullptr, *d_face_norm = nullptr,
         *d_face_shift = nullptr;
  uint8_t *d_face_has_shift = nullptr;
  uint64_t *d_links = nullptr;
  // host copies
  std::mutex mu;
  bool have_host = false;
  std::vector<double> h_cell_loc, h_cell_cen, h_cell_vol, h_cell_sr;
  std::vector<uint64_t> h_cell_foff, h_cell_fcnt, h_cell_idx;
  std::vector<uint32_t> h_npop;
  std::vector<uint64_t> h_face_left;
  std::vector<int64_t> h_face_right;
  std::vector<double> h_face_area, h_face_cen, h_face_norm, h_face_shift;
  std::vector<uint8_t> h_face_has_shift;
  std::vector<uint64_t> h_links;
  mv_stats stats{};

  template <typename T>
  T *own(size_t n) {
    T *p = dnew<T>(n);
    if (!p) throw BuildError{MV_E_OOM};
    dev.push_back(p);
    return p;
  }
  ~mv_tess() {
    for (void *p : dev) dfree(p);
  }
};

namespace {

template <typename T>
void dl(std::vector<T> &h, const T *d, size_t n) {
  h.resize(n);
  if (n) HIPCHK(hipMemcpy(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost));
}

void download(mv_tess *t) {
  std::lock_guard<std::mutex> g(t->mu);
  if (t->have_host) return;
  size_t n = t->n, F = t->nf, L = t->nl;
  dl(t->h_cell_loc, t->d_cell_loc, 3 * n);
  dl(t->h_cell_cen, t->d_cell_cen, 3 * n);
  dl(t->h_cell_vol, t->d_cell_vol, n);
  dl(t->h_cell_sr, t->d_cell_sr, n);
  dl(t->h_cell_foff, t->d_cell_foff, n);
  dl(t->h_cell_fcnt, t->d_cell_fcnt, n);
  dl(t->h_cell_idx, t->d_cell_idx, n);
  dl(t->h_npop, t->d_npop, n);
  dl(t->h_face_left, t->d_face_left, F);
  dl(t->h_face_right, t->d_face_right, F);
  dl(t->h_face_area, t->d_face_area, F);
  dl(t->h_face_cen, t->d_face_cen, 3 * F);
  dl(t->h_face_norm, t->d_face_norm, 3 * F);
  dl(t->h_face_shift, t->d_face_shift, 3 * F);
  dl(t->h_face_has_shift, t->d_face_has_shift, F);
  dl(t->h_links, t->d_links, L);
  t->have_host = true;
}

// SimulationBoundary::cuboid (src/voronoi/boundary.rs:14-47)
BoxParams make_box(const double anchor_in[3], const double width_in[3],
                   int periodic, int dim) {
  BoxParams b;
  memset(&b, 0, sizeof b);
  double a[3] = {anchor_in[0], anchor_in[1], anchor_in[2]};
  double w[3] = {width_in[0], width_in[1], width_in[2]};
  for (int c = 0; c < 3; c++) b.width[c] = w[c];
  if (periodic) {
    a[0] -= w[0];
    w[0] *= 3.;
    if (dim >= 2) {
      a[1] -= w[1];
      w[1] *= 3.;
    }
    if (dim == 3) {
      a[2] -= w[2];
      w[2] *= 3.;
    }
  }
  double aw[3] = {a[0] + w[0], a[1] + w[1], a[2] + w[2]};
  for (int k = 0; k < 6; k++) {
    int ax = k / 2;
    double sgn = (k % 2 == 0) ? 1.0 : -1.0;
    double n[3] = {0., 0., 0.};
    n[ax] = sgn;
    const double *p = (k % 2 == 0) ? a : aw;
    for (int c = 0; c < 3; c++) {
      b.pn[k][c] = n[c];
      b.pp[k][c] = p[c];
    }
    // d = n . p in glam's order
    b.pd[k] = (n[0] * p[0] + n[1] * p[1]) + n[2] * p[2];
  }
  for (int c = 0; c < 3; c++) {
    b.A[c] = a[c] - w[c];
    b.iw[c] = 1. / (3. * w[c]);
  }
  b.dim = dim;
  b.periodic = periodic;
  return b;
}

GridParams make_grid(const double bb[6], size_t n, const double width[3],
                     int dim, int periodic, double occupancy) {
  GridParams g;
  memset(&g, 0, sizeof g);
  double ext[3], vol = 1.;
  int nd = 0;
  double scale = 0.;
  for (int c = 0; c < 3; c++) {
    g.lo[c] = bb[c];
    ext[c] = bb[3 + c] - bb[c];
    if (ext[c] > 0) {
      vol *= ext[c];
      nd++;
    }
    scale = std::max(scale, std::max(std::fabs(bb[c]), std::fabs(bb[3 + c])));
    scale = std::max(scale, std::fabs(width[c]));
  }
  double h = nd ? std::pow(vol * occupancy / (double)n, 1.0 / nd) : 0.;
  size_t ncell = 1;
  for (int c = 0; c < 3; c++) {
    int G = 1;
    if (ext[c] > 0 && h > 0) {
      double gg = std::ceil(ext[c] / h);
      G = gg < 1 ? 1 : (gg > 1024 ? 1024 : (int)gg);
    }
    g.G[c] = G;
    g.inv_h[c] = ext[c] > 0 ? (double)G / ext[c] : 0.;
    g.h[c] = ext[c] > 0 ? ext[c] / G : 0.;
    ncell *= (size_t)G;
  }
  (void)ncell;
  for (int c = 0; c < 3; c++) {
    bool act = periodic && (c == 0 || (c == 1 && dim >= 2) || (c == 2 && dim == 3));
    g.img_lo[c] = act ? -1 : 0;
    g.img_hi[c] = act ? 1 : 0;
  }
  uint64_t nimg = (uint64_t)(g.img_hi[0] - g.img_lo[0] + 1) *
                  (g.img_hi[1] - g.img_lo[1] + 1) *
                  (g.img_hi[2] - g.img_lo[2] + 1);
  g.total_stream = (uint64_t)n * nimg;
  g.margin = 1e-10 * (scale + 1e-300) + 1e-300;
  return g;
}

// neighbour lists: the LDS gather buffer must hold more than L entries;
// long lists (and cells that did not fit) use the global-memory variant
constexpr uint32_t BIG_KNN_CAP = 65536;
constexpr uint32_t BIG_KNN_WAVES = 1024;

void launch_knn(const double4 *spos, const uint32_t *cell_start,
                const GridParams &grid, const BoxParams &box,
                const uint32_t *wl, uint32_t cnt, uint32_t L,
                const uint32_t *Lper, const float *Rper, const uint64_t *kofs,
                uint64_t *kpk, uint32_t *kcnt, bool big, Temps &TT, hipStream_t s) {
  if (cnt == 0) return;
  if (!big && L <= 256) {
    hipLaunchKernelGGL((k_knn<512, 4>), dim3((cnt + 3) / 4), dim3(256), 0, s, spos,
                       cell_start, grid, box, wl, cnt, L, Lper, Rper, kofs, kpk, kcnt);
  } else if (!big && L <= 1024) {
    hipLaunchKernelGGL((k_knn<2048, 1>), dim3(cnt), dim3(64), 0, s, spos,
                       cell_start, grid, box, wl, cnt, L, Lper, Rper, kofs, kpk, kcnt);
  } else {
    uint32_t waves = std::min<uint32_t>(cnt, BIG_KNN_WAVES);
    waves = (waves + 3) / 4 * 4;
    double *gkey = need(TT.get<double>((size_t)waves * BIG_KNN_CAP));
    uint64_t *gpk = need(TT.get<uint64_t>((size_t)waves * BIG_KNN_CAP));
    hipLaunchKernelGGL(k_knn_big, dim3(waves / 4), dim3(256), 0, s, spos, cell_start, grid,
                       box, wl, cnt, L, Lper, kofs, kpk, kcnt, gkey, gpk, BIG_KNN_CAP);
  }
  HIPCHK(hipGetLastError());
}

void build_impl(mv_tess *T, const double *d_xyz, size_t n,
                const uint8_t *d_mask, const double anchor_in[3],
                const double width_in[3], uint32_t dim, uint32_t periodic,
                const mv_options &opt) {
  auto t_start = std::chrono::steady_clock::now();
  hipStream_t s = 0;
  // voronoi.rs:200-210
  double anchor[3] = {anchor_in[0], anchor_in[1], anchor_in[2]};
  double width[3] = {width_in[0], width_in[1], width_in[2]};
  if (dim == 1) {
    anchor[1] = -0.5;
    width[1] = 1.;
  }
  if (dim == 1 || dim == 2) {
    anchor[2] = -0.5;
    width[2] = 1.;
  }
  memcpy(T->anchor, anchor, sizeof anchor);
  memcpy(T->width, width, sizeof width);
  T->dim = dim;
  T->periodic = periodic;
  T->n = n;

  size_t nn = n ? n : 1;
  T->d_cell_loc = T->own<double>(3 * nn);
  T->d_cell_cen = T->own<double>(3 * nn);
  T->d_cell_vol = T->own<double>(nn);
  T->d_cell_sr = T->own<double>(nn);
  T->d_cell_foff = T->own<uint64_t>(nn);
  T->d_cell_fcnt = T->own<uint64_t>(nn);
  T->d_cell_idx = T->own<uint64_t>(nn);
  T->d_npop = T->own<uint32_t>(nn);
  HIPCHK(hipMemsetAsync(T->d_cell_cen, 0, 3 * nn * sizeof(double), s));
  HIPCHK(hipMemsetAsync(T->d_cell_vol, 0, nn * sizeof(double), s));
  HIPCHK(hipMemsetAsync(T->d_cell_sr, 0, nn * sizeof(double), s));
  HIPCHK(hipMemsetAsync(T->d_npop, 0, nn * sizeof(uint32_t), s));
  if (n == 0) {
    T->d_face_left = T->own<uint64_t>(1);
    T->d_face_right = T->own<int64_t>(1);
    T->d_face_area = T->own<double>(1);
    T->d_face_cen = T->own<double>(3);
    T->d_face_norm = T->own<double>(3);
    T->d_face_shift = T->own<double>(3);
    T->d_face_has_shift = T->own<uint8_t>(1);
    T->d_links = T->own<uint64_t>(1);
    HIPCHK(hipStreamSynchronize(s));
    return;
  }

  Temps TT;
  // ---- prepare + bounding box
  Timer tg(s);
  double4 *gpos = need(TT.get<double4>(n));
  unsigned nb = nblk(n);
  double *bbp = need(TT.get<double>(6 * (size_t)nb));
  double *bbd = need(TT.get<double>(6));
  hipLaunchKernelGGL(k_prepare, dim3(nb), dim3(256), 0, s, d_xyz, n, (int)dim, gpos, bbp);
  hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(256), 0, s, bbp, (int)nb, bbd);
  HIPCHK(hipGetLastError());
  double bb[6];
  HIPCHK(hipMemcpyAsync(bb, bbd, sizeof bb, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  for (int c = 0; c < 6; c++)
    if (!std::isfinite(bb[c])) throw BuildError{MV_E_DEGENERATE_POINTS};

  BoxParams box = make_box(anchor, width, (int)periodic, (int)dim);
  double occ = opt.grid_occupancy16 ? (double)opt.grid_occupancy16 / 16.0 : 2.0;
  GridParams grid = make_grid(bb, n, width, (int)dim, (int)periodic, occ);
  size_t ncell = (size_t)grid.G[0] * grid.G[1] * grid.G[2];

  // ---- counting-sort grid
  uint32_t *cellid = need(TT.get<uint32_t>(n));
  uint32_t *counts = need(TT.get<uint32_t>(ncell + 1));
  uint32_t *cell_start = need(TT.get<uint32_t>(ncell + 1));
  uint32_t *fill = need(TT.get<uint32_t>(ncell));
  uint32_t *order = need(TT.get<uint32_t>(n));
  double4 *spos = need(TT.get<double4>(n));
  uint32_t *slot_of = need(TT.get<uint32_t>(n));
  HIPCHK(hipMemsetAsync(counts, 0, (ncell + 1) * sizeof(uint32_t), s));
  HIPCHK(hipMemsetAsync(fill, 0, ncell * sizeof(uint32_t), s));
  hipLaunchKernelGGL(k_grid_count, dim3(nb), dim3(256), 0, s, gpos, n, grid, cellid, counts);
  scan_excl<uint32_t, uint32_t>(counts, cell_start, ncell + 1, s);
  hipLaunchKernelGGL(k_grid_scatter, dim3(nb), dim3(256), 0, s, cellid, n, cell_start, fill, order);
  hipLaunchKernelGGL(k_grid_sort_cells, dim3(nblk(ncell)), dim3(256), 0, s, cell_start, ncell, order);
  hipLaunchKernelGGL(k_grid_gather, dim3(nb), dim3(256), 0, s, gpos, order, n, spos, slot_of);
  HIPCHK(hipGetLastError());

  // ---- initial work list: active slots in grid order
  uint32_t *flag = need(TT.get<uint32_t>(n));
  uint32_t *ident = need(TT.get<uint32_t>(n));
  uint32_t *pos = need(TT.get<uint32_t>(n));
  uint32_t *work = need(TT.get<uint32_t>(n));
  uint32_t *work2 = need(TT.get<uint32_t>(n));
  uint32_t *bigw = need(TT.get<uint32_t>(n));
  uint32_t *bigw2 = need(TT.get<uint32_t>(n));
  hipLaunchKernelGGL(k_flag_active, dim3(nb), dim3(256), 0, s, spos, n, d_mask, flag, ident);
  uint32_t nwork = scan_excl<uint32_t, uint32_t>(flag, pos, n, s);
  if (nwork)
    hipLaunchKernelGGL(k_compact, dim3(nb), dim3(256), 0, s, ident, flag, pos, (uint32_t)n, work);
  HIPCHK(hipGetLastError());
  T->stats.ms_grid = tg.stop();

  // ---- per-cell outputs and face staging
  uint32_t *cell_nf = need(TT.get<uint32_t>(n));
  uint64_t *cell_fbase = need(TT.get<uint64_t>(n));
  HIPCHK(hipMemsetAsync(cell_nf, 0, n * sizeof(uint32_t), s));
  HIPCHK(hipMemsetAsync(cell_fbase, 0, n * sizeof(uint64_t), s));
  uint64_t st_cap = std::max<uint64_t>(4096, 16ull * nwork);
  uint32_t *st_cell = nullptr;
  int64_t *st_right = nullptr;
  uint8_t *st_img = nullptr;
  double *st_area = nullptr, *st_cen = nullptr, *st_norm = nullptr;
  unsigned long long *counters = need(TT.get<unsigned long long>(4));
  // counters: [0] staging cursor, [1] error word, [2] exact tests, [3] overflow
  uint8_t *wstate = need(TT.get<uint8_t>(n));
  unsigned int *queue = need(TT.get<unsigned int>(1));
  const bool exact_order = (opt.flags & MV_FLAG_EXACT_ORDER) != 0;
  uint8_t *bstate = need(TT.get<uint8_t>(n));

  // lane state
  const uint32_t lanes_main = opt.lanes ? opt.lanes : (uint32_t)device_cus() * 512u;
  const uint32_t lanes_big = 8192;

  struct LaneBufs {
    double4 *plA, *plB, *vtx, *facc;
    uint32_t *vdual;
    uint8_t *cyc;
    uint32_t lanes;
  };
  auto alloc_lanes = [&](uint32_t lanes, int maxp, int maxv) {
    LaneBufs L;
    L.lanes = lanes;
    L.plA = need(TT.get<double4>((size_t)maxp * lanes));
    L.plB = need(TT.get<double4>((size_t)maxp * lanes));
    L.facc = need(TT.get<double4>((size_t)maxp * lanes));
    L.cyc = need(TT.get<uint8_t>((size_t)maxp * lanes));
    L.vtx = need(TT.get<double4>((size_t)maxv * lanes));
    L.vdual = need(TT.get<uint32_t>((size_t)maxv * lanes));
    return L;
  };

  uint32_t L = opt.k0 ? opt.k0 : 27;
  const uint32_t kmax = std::min<uint32_t>(opt.k_max ? opt.k_max : 32768, 32768);
  uint64_t redone = 0, bigcells = 0;
  uint32_t passes = 0;
  double ms_knn = 0, ms_clip = 0;

  for (int attempt = 0;; attempt++) {
    // (re)allocate staging
    st_cell = need(TT.get<uint32_t>(st_cap));
    st_right = need(TT.get<int64_t>(st_cap));
    st_img = need(TT.get<uint8_t>(st_cap));
    st_area = need(TT.get<double>(st_cap));
    st_cen = need(TT.get<double>(3 * st_cap));
    st_norm = need(TT.get<double>(3 * st_cap));
    HIPCHK(hipMemsetAsync(counters, 0, 4 * sizeof(unsigned long long), s));
    unsigned long long err_init = ~0ull;
    HIPCHK(hipMemcpyAsync(counters + 1, &err_init, sizeof err_init, hipMemcpyHostToDevice, s));

    uint32_t nw = nwork, nbig = 0;
    uint32_t *cw = work, *cw2 = work2, *cb = bigw, *cb2 = bigw2;
    uint32_t Lp = L;
    LaneBufs LM{}, LB{};
    bool have_main = false, have_big = false;
    bool overflowed = false;
    passes = 0;
    redone = 0;
    bigcells = 0;
    while (nw + nbig > 0) {
      if (Lp > kmax) throw BuildError{MV_E_CAPACITY};
      passes++;
      auto run = [&](uint32_t *wl, uint32_t cnt, uint8_t *state, bool big) {
        Timer tk(s);
        uint64_t *kpk = need(TT.get<uint64_t>((size_t)Lp * cnt));
        uint32_t *kcnt = need(TT.get<uint32_t>(cnt));
        launch_knn(spos, cell_start, grid, box, wl, cnt, Lp, nullptr, nullptr, nullptr, kpk, kcnt,
                   big, TT, s);
        ms_knn += tk.stop();
        Timer tc(s);
        if (!big && !exact_order) {
          WaveArgs W;
          memset(&W, 0, sizeof W);
          W.spos = spos;
          W.mask = d_mask;
          W.box = box;
          W.total_stream = grid.total_stream;
          W.work = wl;
          W.nwork = cnt;
          W.kpk = kpk;
          W.kofs = nullptr;
          W.kcnt = kcnt;
          W.L = Lp;
          W.queue = queue;
          W.wstate = state;
          W.cell_vol = T->d_cell_vol;
          W.cell_cen = T->d_cell_cen;
          W.cell_sr = T->d_cell_sr;
          W.cell_nf = cell_nf;
          W.cell_fbase = cell_fbase;
          W.cell_npop = T->d_npop;
          W.st_cell = st_cell;
          W.st_right = st_right;
          W.st_img = st_img;
          W.st_area = st_area;
          W.st_cen = st_cen;
          W.st_norm = st_norm;
          W.st_cap = st_cap;
          W.st_counter = counters;
          W.err_word = counters + 1;
          W.exact_counter = counters + 2;
          W.overflow = counters + 3;
          HIPCHK(hipMemsetAsync(queue, 0, sizeof(unsigned int), s));
          uint32_t blocks = std::min<uint32_t>((cnt + 3) / 4, (uint32_t)device_cus() * 4u);
          hipLaunchKernelGGL(k_clip_wave, dim3(blocks), dim3(256), 0, s, W);
          HIPCHK(hipGetLastError());
          ms_clip += tc.stop();
          return;
        }
        ClipArgs A;
        memset(&A, 0, sizeof A);
        A.spos = spos;
        A.mask = d_mask;
        A.box = box;
        A.total_stream = grid.total_stream;
        A.work = wl;
        A.nwork = cnt;
        A.kpk = kpk;
        A.kofs = nullptr;
        A.kcnt = kcnt;
        A.L = Lp;
        uint32_t lanes;
        if (!big) {
          lanes = std::min<uint32_t>(lanes_main, (cnt + 255) / 256 * 256);
          if (!have_main) {
            LM = alloc_lanes(lanes_main, MAXP_MAIN, MAXV_MAIN);
            have_main = true;
          }
          A.plA = LM.plA; A.plB = LM.plB; A.vtx = LM.vtx; A.facc = LM.facc;
          A.vdual = LM.vdual; A.cyc = LM.cyc;
          A.lanes = LM.lanes;
        } else {
          lanes = std::min<uint32_t>(lanes_big, (cnt + 255) / 256 * 256);
          if (!have_big) {
            LB = alloc_lanes(lanes_big, MAXP_BIG, MAXV_BIG);
            have_big = true;
          }
          A.plA = LB.plA; A.plB = LB.plB; A.vtx = LB.vtx; A.facc = LB.facc;
          A.vdual = LB.vdual; A.cyc = LB.cyc;
          A.lanes = LB.lanes;
        }
        A.wstate = state;
        A.cell_vol = T->d_cell_vol;
        A.cell_cen = T->d_cell_cen;
        A.cell_sr = T->d_cell_sr;
        A.cell_nf = cell_nf;
        A.cell_fbase = cell_fbase;
        A.cell_npop = T->d_npop;
        A.st_cell = st_cell;
        A.st_right = st_right;
        A.st_img = st_img;
        A.st_area = st_area;
        A.st_cen = st_cen;
        A.st_norm = st_norm;
        A.st_cap = st_cap;
        A.st_counter = counters;
        A.err_word = counters + 1;
        A.exact_counter = counters + 2;
        A.overflow = counters + 3;
        if (!big)
          hipLaunchKernelGGL((k_clip<MAXP_MAIN, MAXV_MAIN>), dim3(lanes / 256), dim3(256), 0, s, A);
        else
          hipLaunchKernelGGL((k_clip<MAXP_BIG, MAXV_BIG>), dim3(lanes / 256), dim3(256), 0, s, A);
        HIPCHK(hipGetLastError());
        ms_clip += tc.stop();
      };
      if (nw) run(cw, nw, wstate, false);
      if (nbig) run(cb, nbig, bstate, true);
      unsigned long long cnts[4];
      HIPCHK(hipMemcpyAsync(cnts, counters, sizeof cnts, hipMemcpyDeviceToHost, s));
      HIPCHK(hipStreamSynchronize(s));
      if (cnts[1] != ~0ull) {
        static const mv_status map[] = {MV_OK, MV_E_ARG, MV_E_DEGENERATE_POINTS,
                                        MV_E_SELF_NOT_FIRST, MV_E_PLANE_DET_ZERO,
                                        MV_E_NAN, MV_E_BOUNDARY_WALK};
        uint32_t code = (uint32_t)(cnts[1] & 0xFF);
        throw BuildError{code < 7 ? map[code] : MV_E_DEVICE};
      }
      if (cnts[3]) {
        overflowed = true;
        break;
      }
      // next lists: main NEED_MORE; big NEED_MORE + main CAPACITY
      uint32_t nw_next = compact(cw, wstate, nw, CS_NEED_MORE, cw2, s);
      uint32_t nb_next = compact(cb, bstate, nbig, CS_NEED_MORE, cb2, s);
      uint32_t ncap = compact(cw, wstate, nw, CS_CAPACITY, cb2 + nb_next, s);
      uint32_t nbcap = compact(cb, bstate, nbig, CS_CAPACITY, cw2 + nw_next, s);
      if (nbcap) throw BuildError{MV_E_CAPACITY};
      bigcells += ncap;
      redone += nw_next + nb_next;
      nw = nw_next;
      nbig = nb_next + ncap;
      std::swap(cw, cw2);
      std::swap(cb, cb2);
      if (nw + nbig) Lp *= 2;
    }
    if (!overflowed) {
      unsigned long long cnts[4];
      HIPCHK(hipMemcpy(cnts, counters, sizeof cnts, hipMemcpyDeviceToHost));
      T->stats.exact_tests = cnts[2];
      T->stats.k_last = Lp;
      break;
    }
    if (attempt > 4) throw BuildError{MV_E_CAPACITY};
    st_cap *= 2;
    HIPCHK(hipMemsetAsync(cell_nf, 0, n * sizeof(uint32_t), s));
  }
  T->stats.passes = passes;
  T->stats.cells_redone = redone;
  T->stats.big_cells = bigcells;
  T->stats.ms_knn = ms_knn;
  T->stats.ms_clip = ms_clip;

  // ---- faces and links
  Timer tf(s);
  uint64_t *face_off = need(TT.get<uint64_t>(n));
  uint64_t F = scan_excl<uint32_t, uint64_t>(cell_nf, face_off, n, s);
  T->nf = F;
  size_t FF = F ? F : 1;
  T->d_face_left = T->own<uint64_t>(FF);
  T->d_face_right = T->own<int64_t>(FF);
  T->d_face_area = T->own<double>(FF);
  T->d_face_cen = T->own<double>(3 * FF);
  T->d_face_norm = T->own<double>(3 * FF);
  T->d_face_shift = T->own<double>(3 * FF);
  T->d_face_has_shift = T->own<uint8_t>(FF);
  uint32_t *link_in = need(TT.get<uint32_t>(n));
  HIPCHK(hipMemsetAsync(link_in, 0, n * sizeof(uint32_t), s));
  if (F) {
    FinalArgs FA;
    FA.st_cell = st_cell;
    FA.st_right = st_right;
    FA.st_img = st_img;
    FA.st_area = st_area;
    FA.st_cen = st_cen;
    FA.st_norm = st_norm;
    FA.n_staged = F;
    FA.cell_fbase = cell_fbase;
    FA.face_off = face_off;
    for (int c = 0; c < 3; c++) FA.width[c] = width[c];
    FA.face_left = T->d_face_left;
    FA.face_right = T->d_face_right;
    FA.face_area = T->d_face_area;
    FA.face_centroid = T->d_face_cen;
    FA.face_normal = T->d_face_norm;
    FA.face_shift = T->d_face_shift;
    FA.face_has_shift = T->d_face_has_shift;
    FA.link_in = link_in;
    hipLaunchKernelGGL(k_place_faces, dim3(nblk(F)), dim3(256), 0, s, FA);
    HIPCHK(hipGetLastError());
  }
  hipLaunchKernelGGL(k_link_count, dim3(nb), dim3(256), 0, s, cell_nf, link_in, n, T->d_cell_fcnt);
  uint64_t NL = scan_excl<uint64_t, uint64_t>(T->d_cell_fcnt, T->d_cell_foff, n, s);
  T->nl = NL;
  T->d_links = T->own<uint64_t>(NL ? NL : 1);
  if (F) {
    uint32_t *lfill = need(TT.get<uint32_t>(n));
    HIPCHK(hipMemsetAsync(lfill, 0, n * sizeof(uint32_t), s));
    hipLaunchKernelGGL(k_fill_links, dim3(nblk(F)), dim3(256), 0, s, T->d_face_left,
                       T->d_face_right, T->d_face_has_shift, (size_t)F, T->d_cell_foff,
                       lfill, T->d_links);
    hipLaunchKernelGGL(k_sort_links, dim3(nb), dim3(256), 0, s, T->d_cell_foff,
                       T->d_cell_fcnt, n, T->d_links);
  }
  hipLaunchKernelGGL(k_cell_records, dim3(nb), dim3(256), 0, s, gpos, n, d_mask,
                     T->d_cell_loc, T->d_cell_idx);
  HIPCHK(hipGetLastError());
  T->stats.ms_finalize = tf.stop();
  HIPCHK(hipStreamSynchronize(s));
  auto t_end = std::chrono::steady_clock::now();
  T->stats.ms_total = std::chrono::duration<double, std::milli>(t_end - t_start).count();
}

bool valid_args(size_t n, uint32_t dim, const double *anchor,
                const double *width) {
  if (dim < 1 || dim > 3) return false;
  if (!anchor || !width) return false;
  if (n >= (1ull << 27)) return false;
  return true;
}

mv_status run_build(const double *d_xyz, size_t n, const uint8_t *d_mask,
                    const double anchor[3], const double width[3],
                    uint32_t dim, uint32_t periodic, const mv_options *opt,
                    mv_tess **out) {
  mv_options o;
  mv_default_options(&o);
  if (opt) o = *opt;
  if (o.mode != MV_MODE_PLAIN) return MV_E_ARG;
  mv_tess *T = new (std::nothrow) mv_tess();
  if (!T) return MV_E_OOM;
  try {
    build_impl(T, d_xyz, n, d_mask, anchor, width, dim, periodic ? 1 : 0, o);
  } catch (const BuildError &e) {
    (void)hipDeviceSynchronize();
    (void)hipGetLastError();
    delete T;
    return e.st;
  }
  *out = T;
  return MV_OK;
}

}  // namespace

// ---------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------
extern "C" {

void mv_default_options(mv_options *opt) {
  if (!opt) return;
  memset(opt, 0, sizeof *opt);
  opt->mode = MV_MODE_PLAIN;
  opt->k0 = 27;
  opt->k_max = 32768;
  opt->lanes = 0;
}

mv_status mv_build_device(const double *d_xyz, size_t n,
                          const uint8_t *d_mask, const double anchor[3],
                          const double width[3], uint32_t dimensionality,
                          uint32_t periodic, const mv_options *opt,
                          mv_tess **out) {
  if (!out) return MV_E_ARG;
  *out = nullptr;
  if (!valid_args(n, dimensionality, anchor, width) || (n && !d_xyz))
    return MV_E_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    (void)hipGetLastError();
    return MV_E_DEVICE;
  }
  return run_build(d_xyz, n, d_mask, anchor, width, dimensionality, periodic,
                   opt, out);
}

mv_status mv_build(const double *xyz, size_t n, const uint8_t *mask,
                   const double anchor[3], const double width[3],
                   uint32_t dimensionality, uint32_t periodic,
                   const mv_options *opt, mv_tess **out) {
  if (!out) return MV_E_ARG;
  *out = nullptr;
  if (!valid_args(n, dimensionality, anchor, width) || (n && !xyz))
    return MV_E_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    (void)hipGetLastError();
    return MV_E_DEVICE;
  }
  double *d_xyz = nullptr;
  uint8_t *d_mask = nullptr;
  if (n) {
    d_xyz = (double *)dalloc(3 * n * sizeof(double));
    if (!d_xyz) return MV_E_OOM;
    if (hipMemcpy(d_xyz, xyz, 3 * n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
      dfree(d_xyz);
      return MV_E_DEVICE;
    }
    if (mask) {
      d_mask = (uint8_t *)dalloc(n);
      if (!d_mask) {
        dfree(d_xyz);
        return MV_E_OOM;
      }
      if (hipMemcpy(d_mask, mask, n, hipMemcpyHostToDevice) != hipSuccess) {
        dfree(d_xyz);
        dfree(d_mask);
        return MV_E_DEVICE;
      }
    }
  }
  mv_status st = run_build(d_xyz, n, d_mask, anchor, width, dimensionality,
                           periodic, opt, out);
  dfree(d_xyz);
  dfree(d_mask);
  if (st == MV_OK) {
    try {
      download(*out);
    } catch (const BuildError &e) {
      delete *out;
      *out = nullptr;
      return e.st;
    }
  }
  return st;
}

void mv_free(mv_tess *t) { delete t; }

void mv_anchor(const mv_tess *t, double out[3]) {
  memcpy(out, t->anchor, 3 * sizeof(double));
}
void mv_width(const mv_tess *t, double out[3]) {
  memcpy(out, t->width, 3 * sizeof(double));
}
uint32_t mv_dimensionality(const mv_tess *t) { return t->dim; }
uint32_t mv_periodic(const mv_tess *t) { return t->periodic; }
size_t mv_n_cells(const mv_tess *t) { return t->n; }
size_t mv_n_faces(const mv_tess *t) { return t->nf; }
size_t mv_n_links(const mv_tess *t) { return t->nl; }

#define HOST_ACCESSOR(ret, name, field)          \
  ret mv_##name(mv_tess *t) {                    \
    try {                                        \
      download(t);                               \
    } catch (const BuildError &) {               \
      return nullptr;                            \
    }                                            \
    return t->field.data();                      \
  }
HOST_ACCESSOR(const double *, cell_loc, h_cell_loc)
HOST_ACCESSOR(const double *, cell_centroid, h_cell_cen)
HOST_ACCESSOR(const double *, cell_volume, h_cell_vol)
HOST_ACCESSOR(const double *, cell_safety_radius, h_cell_sr)
HOST_ACCESSOR(const uint64_t *, cell_face_offset, h_cell_foff)
HOST_ACCESSOR(const uint64_t *, cell_face_count, h_cell_fcnt)
HOST_ACCESSOR(const uint64_t *, cell_idx, h_cell_idx)
HOST_ACCESSOR(const uint32_t *, cell_n_popped, h_npop)
HOST_ACCESSOR(const uint64_t *, face_left, h_face_left)
HOST_ACCESSOR(const int64_t *, face_right, h_face_right)
HOST_ACCESSOR(const double *, face_area, h_face_area)
HOST_ACCESSOR(const double *, face_centroid, h_face_cen)
HOST_ACCESSOR(const double *, face_normal, h_face_norm)
HOST_ACCESSOR(const uint8_t *, face_has_shift, h_face_has_shift)
HOST_ACCESSOR(const double *, face_shift, h_face_shift)
HOST_ACCESSOR(const uint64_t *, cell_face_connections, h_links)

void mv_get_device_view(const mv_tess *t, mv_device_view *v) {
  v->cell_loc = t->d_cell_loc;
  v->cell_centroid = t->d_cell_cen;
  v->cell_volume = t->d_cell_vol;
  v->cell_safety_radius = t->d_cell_sr;
  v->cell_face_offset = t->d_cell_foff;
  v->cell_face_count = t->d_cell_fcnt;
  v->cell_idx = t->d_cell_idx;
  v->face_left = t->d_face_left;
  v->face_right = t->d_face_right;
  v->face_area = t->d_face_area;
  v->face_centroid = t->d_face_cen;
  v->face_normal = t->d_face_norm;
  v->face_has_shift = t->d_face_has_shift;
  v->face_shift = t->d_face_shift;
  v->cell_face_connections = t->d_links;
}

void mv_get_stats(const mv_tess *t, mv_stats *s) { *s = t->stats; }

const char *mv_status_string(mv_status s) {
  switch (s) {
    case MV_OK: return "ok";
    case MV_E_ARG: return "invalid argument";
    case MV_E_DEGENERATE_POINTS: return "degenerate point set (zero or non-finite distance)";
    case MV_E_SELF_NOT_FIRST: return "first nearest neighbour is not the generator itself (duplicate generators)";
    case MV_E_PLANE_DET_ZERO: return "degenerate 3-plane intersection";
    case MV_E_NAN: return "NaN distance encountered";
    case MV_E_BOUNDARY_WALK: return "no suitable vertex found to extend boundary";
    case MV_E_DEVICE: return "HIP device error or no GPU";
    case MV_E_OOM: return "out of memory";
    case MV_E_CAPACITY: return "per-cell capacity exceeded";
  }
  return "unknown status";
}

mv_status mv_in_sphere_exact(const int64_t *abcdv, size_t count,
                             int8_t *out_sign) {
  if (count == 0) return MV_OK;
  if (!abcdv || !out_sign) return MV_E_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    (void)hipGetLastError();
    return MV_E_DEVICE;
  }
  int64_t *d_in = (int64_t *)dalloc(15 * count * sizeof(int64_t));
  int8_t *d_out = (int8_t *)dalloc(count);
  if (!d_in || !d_out) {
    dfree(d_in);
    dfree(d_out);
    return MV_E_OOM;
  }
  mv_status st = MV_OK;
  if (hipMemcpy(d_in, abcdv, 15 * count * sizeof(int64_t), hipMemcpyHostToDevice) != hipSuccess)
    st = MV_E_DEVICE;
  if (st == MV_OK) {
    hipLaunchKernelGGL(k_exact_batch, dim3(nblk(count)), dim3(256), 0, 0, d_in, count, d_out);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpy(out_sign, d_out, count, hipMemcpyDeviceToHost) != hipSuccess)
      st = MV_E_DEVICE;
  }
  dfree(d_in);
  dfree(d_out);
  return st;
}

}  // extern "C"
