// mvoro.hip -- host side of libmvoro: the C ABI of include/mvoro.h over the
// gfx950 kernels of mv_grid.hpp, mv_clip.hpp and mv_finalize.hpp.
//
// Pipeline of one build (all device-resident):
//   prepare + bounding box -> counting-sort grid -> [neighbour lists of
//   length L -> per-cell clip + integrals] repeated with L doubled for the
//   cells whose security radius was not reached -> face placement -> links.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <system_error>
#include <thread>
#include <unordered_map>
#include <vector>

#include "mvoro.h"
#include "mv_clip.hpp"
#include "mv_clipw.hpp"
#include "mv_clipp.hpp"
#include "mv_common.hpp"
#include "mv_faces.hpp"
#include "mv_finalize.hpp"
#include "mv_graph.hpp"
#include "mv_grid.hpp"
#include "mv_learn.hpp"
#include "mvoro_integrals.hpp"

// Device memory of the allocation-free build (mv_build_device_noalloc):
// every scratch buffer of such a build is carved from this array.
static __device__ __align__(256) uint8_t scratch[MV_NOALLOC_SCRATCH_BYTES];

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
// While a thread runs the allocation-free build, its device allocations are
// carved from one region (the static scratch) by a bump pointer and never
// released; nothing reaches hipMalloc.
struct Arena {
  uint8_t *base = nullptr;
  size_t size = 0, used = 0;
  bool exhausted = false;
  unsigned long long *words = nullptr;  // device words the host reads totals from
};
thread_local Arena *t_arena = nullptr;
// set while a captured build (mv_graph) is recorded
thread_local bool t_graph = false;
// the list kernels' counters of the current build, or null: the long-list
// kernel's items, ball gathers and entries gathered, then 64 counters of
// the wavefront list kernels' ball gathers and 64 of their candidates read
thread_local unsigned long long *t_knn_stat = nullptr;
unsigned long long *knn_stripes() { return t_knn_stat && !t_graph ? t_knn_stat + 3 : nullptr; }
bool arena_active() { return t_arena != nullptr; }
size_t arena_free() {
  return t_arena && t_arena->used < t_arena->size ? t_arena->size - t_arena->used : 0;
}
void *arena_alloc(Arena &A, size_t bytes) {
  const size_t off = (A.used + 255) & ~size_t(255);
  const size_t b = bytes ? bytes : 1;
  if (off > A.size || b > A.size - off) {
    A.exhausted = true;
    return nullptr;
  }
  A.used = off + b;
  return A.base + off;
}
// hipMalloc calls made by this process (test hook)
std::atomic<unsigned long long> g_device_mallocs{0};

// The allocation-free build's scratch: one copy per GPU, held until the
// process ends.  In host code `scratch` names a host-side stand-in, not a
// device address; the device address comes from hipGetSymbolAddress (once
// per device, with that device current) and only that pointer is used.
constexpr int NOALLOC_MAX_DEVICES = 64;
struct NoAllocDevice {
  std::mutex mu;     // serialises this process's calls on the device
  void *ptr = nullptr;
};
NoAllocDevice &noalloc_device(int dev) {
  static NoAllocDevice d[NOALLOC_MAX_DEVICES];
  return d[dev];
}

size_t round_bytes(size_t b) {
  if (b < 512) return 512;
  if (b < (1u << 20)) return (b + 4095) & ~size_t(4095);
  return (b + (1u << 20) - 1) & ~size_t((1u << 20) - 1);
}
void *dalloc(size_t bytes) {
  if (t_arena) return arena_alloc(*t_arena, bytes);
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
  g_device_mallocs++;
  if (hipMalloc(&p, sz) != hipSuccess) {
    (void)hipGetLastError();
    for (auto &kv : P.free_) (void)hipFree(kv.second);
    P.free_.clear();
    g_device_mallocs++;
    if (hipMalloc(&p, sz) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
  }
  P.live_[p] = sz;
  return p;
}
void dfree(void *p) {
  // arena memory is never returned (nor found among the pool's blocks)
  if (!p || t_arena) return;
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

// The status of the exception being handled (call inside a catch block):
// nothing thrown by the library may cross the C ABI.
mv_status status_of_exception() noexcept {
  try {
    throw;
  } catch (const BuildError &e) {
    return e.st;
  } catch (const std::bad_alloc &) {
    return MV_E_OOM;
  } catch (...) {
    return MV_E_DEVICE;
  }
}

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

// exclusive scan on the device (tile scans, then the tile offsets)
template <typename TIn, typename TOut>
void scan_excl_dev(const TIn *in, TOut *out, size_t n, hipStream_t s) {
  size_t tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
  Temps T;
  TOut *sums = need(T.get<TOut>(tiles));
  hipLaunchKernelGGL((k_scan_tile<TIn, TOut>), dim3((unsigned)tiles),
                     dim3(SCAN_THREADS), 0, s, in, n, out, sums);
  HIPCHK(hipGetLastError());
  if (tiles > 1) {
    TOut *offs = need(T.get<TOut>(tiles));
    scan_excl_dev<TOut, TOut>(sums, offs, tiles, s);
    hipLaunchKernelGGL((k_scan_add<TOut>), dim3((unsigned)tiles),
                       dim3(SCAN_THREADS), 0, s, out, n, offs);
    HIPCHK(hipGetLastError());
  }
}

// Eight words per host thread that kernels write totals to: pinned host
// memory (one small kernel and a stream synchronisation per read, no
// copy), or, in the allocation-free build, device words of its scratch
// read back with a copy.
unsigned long long *host_word() {
  if (t_arena) return t_arena->words;
  thread_local unsigned long long *w = nullptr;
  if (!w && hipHostMalloc((void **)&w, 64, hipHostMallocDefault) != hipSuccess) {
    w = nullptr;
    throw BuildError{MV_E_OOM};
  }
  return w;
}
// words [0, k) of host_word(), after the work queued on s has finished
void read_words(hipStream_t s, unsigned long long *dst, int k) {
  if (t_arena) {
    HIPCHK(hipMemcpyAsync(dst, t_arena->words, (size_t)k * sizeof *dst, hipMemcpyDeviceToHost,
                          s));
    HIPCHK(hipStreamSynchronize(s));
    return;
  }
  HIPCHK(hipStreamSynchronize(s));
  const volatile unsigned long long *w = host_word();
  for (int i = 0; i < k; i++) dst[i] = w[i];
}

// exclusive scan; returns the total
template <typename TIn, typename TOut>
TOut scan_excl(const TIn *in, TOut *out, size_t n, hipStream_t s) {
  if (n == 0) return 0;
  scan_excl_dev<TIn, TOut>(in, out, n, s);
  unsigned long long *hw = host_word();
  hipLaunchKernelGGL((k_scan_total<TIn, TOut>), dim3(1), dim3(1), 0, s, in + n - 1,
                     (const TOut *)(out + n - 1), hw);
  HIPCHK(hipGetLastError());
  unsigned long long tot;
  read_words(s, &tot, 1);
  return (TOut)tot;
}

struct Timer {
  hipEvent_t a, b;
  hipStream_t s;
  bool on;
  // no events in the allocation-free build (it creates nothing)
  explicit Timer(hipStream_t st, bool enabled = true) : s(st), on(enabled && !t_arena) {
    if (!on) return;
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    (void)hipEventRecord(a, s);
  }
  double stop() {
    if (!on) return 0.;
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
// Device allocations of a build; shared by the handles of a batched build.
struct DevStore {
  std::vector<void *> dev;
  ~DevStore() {
    for (void *p : dev) dfree(p);
  }
};

// std::vector storage whose resize leaves the new elements uninitialised
// (the host copies are overwritten by the download at once).
template <typename T>
struct NoInitAlloc : std::allocator<T> {
  template <typename U>
  struct rebind {
    using other = NoInitAlloc<U>;
  };
  NoInitAlloc() = default;
  template <typename U>
  NoInitAlloc(const NoInitAlloc<U> &) noexcept {}
  template <typename U>
  void construct(U *p) noexcept {
    ::new ((void *)p) U;
  }
  template <typename U, typename... A>
  void construct(U *p, A &&...a) {
    ::new ((void *)p) U(std::forward<A>(a)...);
  }
};
template <typename T>
using hvec = std::vector<T, NoInitAlloc<T>>;

struct mv_tess {
  size_t n = 0, nf = 0, nl = 0;
  double anchor[3], width[3];
  uint32_t dim = 3, periodic = 0;
  std::shared_ptr<DevStore> store = std::make_shared<DevStore>();
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
  // kept cells (MV_FLAG_KEEP_CELLS)
  bool keep = false;
  CellStore cs{};
  uint64_t used_pl = 0, used_vx = 0, n_active = 0;
  uint64_t idx_base = 0;           // first generator of this problem in the store
  uint8_t *d_mask = nullptr;       // copy of the build's mask (symmetric faces)
  uint64_t *d_active = nullptr;    // active cells in generator order
  // face lists of the kept cells (mv_with_faces); wf_on selects the fan
  // decomposition, wf_ready says the lists exist
  bool wf_on = false, wf_ready = false;
  uint64_t wf_nfaces = 0, wf_nentries = 0;
  uint64_t *d_wf_fbase = nullptr, *d_wf_voff = nullptr, *d_wf_cell_foff = nullptr;
  uint32_t *d_wf_nf = nullptr, *d_wf_nvert = nullptr, *d_wf_vtx = nullptr;
  int32_t *d_wf_plane = nullptr;
  // a Voronoi made from an integrator reads arrays of the integrator's build
  std::shared_ptr<DevStore> parent;
  // host copies
  std::mutex mu;
  bool have_host = false;
  hvec<double> h_cell_loc, h_cell_cen, h_cell_vol, h_cell_sr;
  hvec<uint64_t> h_cell_foff, h_cell_fcnt, h_cell_idx;
  hvec<uint32_t> h_npop;
  hvec<uint64_t> h_face_left;
  hvec<int64_t> h_face_right;
  hvec<double> h_face_area, h_face_cen, h_face_norm, h_face_shift;
  hvec<uint8_t> h_face_has_shift;
  hvec<uint64_t> h_links;
  mv_stats stats{};

  template <typename T>
  T *own(size_t n) {
    T *p = dnew<T>(n);
    if (!p) throw BuildError{MV_E_OOM};
    store->dev.push_back(p);
    return p;
  }
  void release(void *p) {
    auto it = std::find(store->dev.begin(), store->dev.end(), p);
    if (it != store->dev.end()) {
      store->dev.erase(it);
      dfree(p);
    }
  }
};

namespace {

// Device -> host copy of a handle's arrays.  Into pageable memory the
// runtime stages every byte through its own pinned buffer on the calling
// thread; here the bytes are split over up to DL_THREADS threads, each
// with 2 x DL_HALF of pinned staging and a stream of its own: the DMA into
// one half overlaps the host copy out of the other, and the destination
// pages are first touched by many cores at once.
struct Piece {
  void *dst;
  const void *src;
  size_t bytes;
};

constexpr int DL_THREADS = 16;
constexpr size_t DL_HALF = size_t(4) << 20;
constexpr size_t DL_MIN_THREADED = size_t(32) << 20;  // below: one hipMemcpy per array

struct DlStaging {
  uint8_t *buf[DL_THREADS] = {};
  hipStream_t st[DL_THREADS] = {};
  hipEvent_t ev[DL_THREADS][2] = {};
};

// One download at a time uses a device's staging (held under dl_mu).
std::mutex dl_mu;

// The staging of device dev, made on its first download and kept until the
// process ends (8 MiB of pinned host memory per thread); nullptr if it
// cannot be made.  Called with dl_mu held.
DlStaging *dl_staging(int dev) {
  static std::map<int, DlStaging *> made;
  auto it = made.find(dev);
  if (it != made.end()) return it->second;
  DlStaging *S = new DlStaging;
  bool ok = hipSetDevice(dev) == hipSuccess;
  for (int w = 0; ok && w < DL_THREADS; w++) {
    ok = hipHostMalloc((void **)&S->buf[w], 2 * DL_HALF, hipHostMallocDefault) == hipSuccess &&
         hipStreamCreateWithFlags(&S->st[w], hipStreamNonBlocking) == hipSuccess &&
         hipEventCreateWithFlags(&S->ev[w][0], hipEventDisableTiming) == hipSuccess &&
         hipEventCreateWithFlags(&S->ev[w][1], hipEventDisableTiming) == hipSuccess;
  }
  if (!ok) {
    (void)hipGetLastError();
    S = nullptr;  // what was made stays allocated; the plain copies are used
  }
  made[dev] = S;
  return S;
}

// Worker w: bytes [lo, hi) of the pieces' concatenation, in chunks of at
// most DL_HALF that do not cross a piece.  Returns with its stream idle.
void dl_share(DlStaging *S, int w, int dev, const std::vector<Piece> *P, size_t lo,
              size_t hi, std::atomic<int> *bad) {
  if (hipSetDevice(dev) != hipSuccess) {
    bad->store(1);
    return;
  }
  uint8_t *pend_dst = nullptr;
  size_t pend_n = 0;
  int h = 0;
  bool ok = true;
  size_t base = 0;
  for (const Piece &p : *P) {
    const size_t a = std::max(lo, base), b = std::min(hi, base + p.bytes);
    for (size_t o = a; ok && o < b;) {
      const size_t c = std::min(DL_HALF, b - o);
      uint8_t *stg = S->buf[w] + h * DL_HALF;
      ok = hipMemcpyAsync(stg, (const uint8_t *)p.src + (o - base), c, hipMemcpyDeviceToHost,
                          S->st[w]) == hipSuccess &&
           hipEventRecord(S->ev[w][h], S->st[w]) == hipSuccess;
      if (ok && pend_n) {
        ok = hipEventSynchronize(S->ev[w][h ^ 1]) == hipSuccess;
        if (ok) memcpy(pend_dst, S->buf[w] + (h ^ 1) * DL_HALF, pend_n);
      }
      pend_dst = (uint8_t *)p.dst + (o - base);
      pend_n = c;
      h ^= 1;
      o += c;
    }
    base += p.bytes;
  }
  if (ok && pend_n) {
    ok = hipEventSynchronize(S->ev[w][h ^ 1]) == hipSuccess;
    if (ok) memcpy(pend_dst, S->buf[w] + (h ^ 1) * DL_HALF, pend_n);
  }
  if (!ok) {
    (void)hipStreamSynchronize(S->st[w]);
    (void)hipGetLastError();
    bad->store(1);
  }
}

// Runs fn(0) .. fn(T - 1), one thread each.  A thread that cannot be
// started (std::system_error under a thread limit, std::bad_alloc) leaves
// its share, and every later one, to the calling thread; the threads that
// did start are joined in every case, so no joinable std::thread is ever
// destroyed.  fail_after >= 0 makes the start of thread fail_after (and of
// every later one) throw, for the test of this path.
template <typename F>
void run_shares(int T, int fail_after, F &&fn) {
  std::vector<std::thread> th;
  int started = 0;
  try {
    th.reserve((size_t)T);
    for (; started < T; started++) {
      if (fail_after >= 0 && started >= fail_after)
        throw std::system_error(std::make_error_code(std::errc::resource_unavailable_try_again));
      th.emplace_back([&fn, started] { fn(started); });
    }
  } catch (...) {
    // fewer threads: the rest of the shares run here
  }
  for (int w = started; w < T; w++) fn(w);
  for (auto &x : th) x.join();
}

void copy_to_host(const std::vector<Piece> &P) {
  size_t total = 0;
  for (const Piece &p : P) total += p.bytes;
  DlStaging *S = nullptr;
  int dev = 0;
  std::unique_lock<std::mutex> lk(dl_mu, std::defer_lock);
  if (total >= DL_MIN_THREADED) {
    hipPointerAttribute_t at;
    if (hipPointerGetAttributes(&at, P[0].src) == hipSuccess) {
      dev = at.device;
      lk.lock();
      S = dl_staging(dev);
    } else {
      (void)hipGetLastError();
    }
  }
  if (!S) {
    for (const Piece &p : P) HIPCHK(hipMemcpy(p.dst, p.src, p.bytes, hipMemcpyDeviceToHost));
    return;
  }
  const int T = (int)std::min<size_t>(DL_THREADS, (total + 4 * DL_HALF - 1) / (4 * DL_HALF));
  const size_t share = ((total + T - 1) / T + 4095) & ~size_t(4095);
  std::atomic<int> bad{0};
  run_shares(T, -1, [&](int w) {
    const size_t lo = std::min(total, w * share), hi = std::min(total, lo + share);
    dl_share(S, w, dev, &P, lo, hi, &bad);
  });
  if (bad.load()) throw BuildError{MV_E_DEVICE};
}

template <typename T>
void dl(std::vector<Piece> &P, hvec<T> &h, const T *d, size_t n) {
  h.resize(n);
  if (n) P.push_back({h.data(), d, n * sizeof(T)});
}

void download(mv_tess *t) {
  std::lock_guard<std::mutex> g(t->mu);
  if (t->have_host) return;
  size_t n = t->n, F = t->nf, L = t->nl;
  std::vector<Piece> P;
  dl(P, t->h_cell_loc, t->d_cell_loc, 3 * n);
  dl(P, t->h_cell_cen, t->d_cell_cen, 3 * n);
  dl(P, t->h_cell_vol, t->d_cell_vol, n);
  dl(P, t->h_cell_sr, t->d_cell_sr, n);
  dl(P, t->h_cell_foff, t->d_cell_foff, n);
  dl(P, t->h_cell_fcnt, t->d_cell_fcnt, n);
  dl(P, t->h_cell_idx, t->d_cell_idx, n);
  dl(P, t->h_npop, t->d_npop, n);
  dl(P, t->h_face_left, t->d_face_left, F);
  dl(P, t->h_face_right, t->d_face_right, F);
  dl(P, t->h_face_area, t->d_face_area, F);
  dl(P, t->h_face_cen, t->d_face_cen, 3 * F);
  dl(P, t->h_face_norm, t->d_face_norm, 3 * F);
  dl(P, t->h_face_shift, t->d_face_shift, 3 * F);
  dl(P, t->h_face_has_shift, t->d_face_has_shift, F);
  dl(P, t->h_links, t->d_links, L);
  if (!P.empty()) copy_to_host(P);
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
  // the initial cell's vertices: duals (2,5,0) (5,3,0) (1,5,2) (5,1,3)
  // (4,2,0) (4,0,3) (2,4,1) (4,3,1) (convex_cell.rs:294-312), with the same
  // operations as on the device (this file is compiled without contraction)
  static const int DU[8][3] = {{2, 5, 0}, {5, 3, 0}, {1, 5, 2}, {5, 1, 3},
                               {4, 2, 0}, {4, 0, 3}, {2, 4, 1}, {4, 3, 1}};
  b.corner_ok = 1;
  for (int v = 0; v < 8; v++) {
    const int i = DU[v][0], j = DU[v][1], k = DU[v][2];
    d3 loc = mk3(0., 0., 0.);
    if (!intersect_planes(mk3(b.pn[i][0], b.pn[i][1], b.pn[i][2]), b.pd[i],
                          mk3(b.pn[j][0], b.pn[j][1], b.pn[j][2]), b.pd[j],
                          mk3(b.pn[k][0], b.pn[k][1], b.pn[k][2]), b.pd[k], &loc))
      b.corner_ok = 0;
    b.corner[v][0] = loc.x;
    b.corner[v][1] = loc.y;
    b.corner[v][2] = loc.z;
  }
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
// long lists (and cells that did not fit) use the global-memory variant,
// one block per item, whose two buffers hold at least 2 L entries each (a
// power of two, at least 65,536); at most 8 GiB of such buffers are in
// flight
constexpr uint32_t BIG_KNN_CAP = 65536;
// words of a build's pass counters
constexpr int NCOUNTERS = 7 + 128;
// wide kernel: starting capacities per cell (planes, vertices), their
// limit (plane indices have 21 bits) and the blocks of one launch
constexpr uint32_t WIDE_P0 = 1024, WIDE_V0 = 4096;
constexpr uint32_t WIDE_P_MAX = 1u << 20;
constexpr uint32_t WIDE_BLOCKS = 256;
constexpr uint32_t BIG_KNN_BLOCKS = 1024;
constexpr uint64_t BIG_KNN_BYTES = 8ull << 30;
// the longest list the wavefront list kernels build (2,048-entry LDS tile)
constexpr uint32_t MAIN_LIST_MAX = 1024;
// redone lists from this length on are bounded by the seeds within the
// cell's safety radius (k_redo_count)
constexpr uint32_t REDO_COUNT_MIN = 1024;

// The list kernel for lists of at most L entries: the smallest LDS tile
// (512, 1024, 2048 gathered entries) that holds twice L and at least
// min_tile, else the long-list kernel.
void launch_knn(const double4 *spos, const uint32_t *cell_start, const Problems &PR,
                const uint32_t *wl, uint32_t cnt, uint32_t L,
                const uint32_t *Lper, const float *Rper, const uint64_t *kofs,
                uint64_t *kpk, uint32_t *kcnt, double *kdist, bool big, Temps &TT,
                hipStream_t s, uint32_t min_tile = 0, const uint32_t *dn = nullptr,
                uint32_t big_blocks = 0, uint32_t max_blocks_knn = 0, bool proposal = false) {
  // proposal: Rper are first radii (proposed), not radii of redone lists
  // max_blocks_knn: a smaller grid (the kernels loop over their items)
  auto grid = [&](uint32_t b) {
    return dim3(max_blocks_knn ? std::min(b, max_blocks_knn) : b);
  };
  if (cnt == 0) return;
  if (!big && min_tile > 512 && L <= 512) L = min_tile >= 2048 ? 1024u : 512u;
  // the instantiation with the dense-cell path only when the grid has some
  const bool sub = PR.sub_base != nullptr;
  auto wave_knn = [&](auto kern, uint32_t wpb) {
    hipLaunchKernelGGL(kern, grid((cnt + wpb - 1) / wpb), dim3(64 * wpb), 0, s, spos, cell_start,
                       PR, wl, cnt, L, Lper, Rper, kofs, kpk, kcnt, kdist, dn, knn_stripes(),
                       proposal ? 1u : 0u);
  };
  if (!big && L <= 256) {
    if (sub)
      wave_knn(k_knn<512, 4, true>, 4);
    else
      wave_knn(k_knn<512, 4, false>, 4);
  } else if (!big && L <= 512) {
    if (sub)
      wave_knn(k_knn<1024, 2, true>, 2);
    else
      wave_knn(k_knn<1024, 2, false>, 2);
  } else if (!big && L <= 1024) {
    if (sub)
      wave_knn(k_knn<2048, 1, true>, 1);
    else
      wave_knn(k_knn<2048, 1, false>, 1);
  } else {    uint32_t cap = BIG_KNN_CAP;
    while (cap < 2 * (uint64_t)L && cap < (1u << 24)) cap <<= 1;
    const uint64_t per_block = 2ull * cap * (sizeof(double) + sizeof(uint64_t));
    // in the allocation-free build, half of what is left of its scratch
    const uint64_t budget = arena_active() && !t_graph ? arena_free() / 2 : BIG_KNN_BYTES;
    const uint32_t max_blocks = (uint32_t)std::max<uint64_t>(
        1, std::min<uint64_t>(BIG_KNN_BLOCKS, budget / per_block));
    const uint32_t blocks = std::min<uint32_t>(big_blocks ? big_blocks : cnt, max_blocks);
    double *gkey = need(TT.get<double>((size_t)blocks * cap));
    uint64_t *gpk = need(TT.get<uint64_t>((size_t)blocks * cap));
    double *gkey2 = need(TT.get<double>((size_t)blocks * cap));
    uint64_t *gpk2 = need(TT.get<uint64_t>((size_t)blocks * cap));
    // lists of different lengths: longest first, claimed one at a time
    uint32_t *perm = nullptr, *queue = nullptr;
    if (Lper && cnt > 1) {
      perm = need(TT.get<uint32_t>(cnt));
      queue = need(TT.get<uint32_t>(1));
      hipLaunchKernelGGL(k_order_by_len, dim3(1), dim3(1024), 0, s, Lper, cnt, perm, queue, dn);
    }
    if (sub)
      hipLaunchKernelGGL(k_knn_big<true>, dim3(blocks), dim3(KB_THREADS), 0, s, spos, cell_start,
                         PR, wl, cnt, L, Lper, Rper, kofs, kpk, kcnt, kdist, gkey, gpk, gkey2,
                         gpk2, cap, dn, t_graph ? nullptr : t_knn_stat, (const uint32_t *)perm,
                         queue);
    else
      hipLaunchKernelGGL(k_knn_big<false>, dim3(blocks), dim3(KB_THREADS), 0, s, spos,
                         cell_start, PR, wl, cnt, L, Lper, Rper, kofs, kpk, kcnt, kdist, gkey, gpk,
                         gkey2, gpk2, cap, dn, t_graph ? nullptr : t_knn_stat,
                         (const uint32_t *)perm, queue);
  }
  HIPCHK(hipGetLastError());
}

// One problem of a build: its box after the 1D/2D normalisation
// (src/voronoi.rs:200-210) and its generators [g0, g1) of the concatenation.
struct ProblemSpec {
  double anchor[3], width[3];
  uint32_t periodic;
  uint64_t g0, g1;
};

void normalise_box(const double a_in[3], const double w_in[3], uint32_t dim, double a[3],
                   double w[3]) {
  for (int c = 0; c < 3; c++) {
    a[c] = a_in[c];
    w[c] = w_in[c];
  }
  if (dim == 1) {
    a[1] = -0.5;
    w[1] = 1.;
  }
  if (dim == 1 || dim == 2) {
    a[2] = -0.5;
    w[2] = 1.;
  }
}

// The counting-sort grids of all problems of a build (device buffers owned
// by `TT`): problem p owns grid cells [coff[p], coff[p+1]) and therefore
// sorted slots [g0, g1).
struct GridData {
  uint32_t P = 1;
  std::vector<BoxParams> boxes;
  std::vector<GridParams> grids;
  std::vector<uint32_t> coff;
  size_t ncell = 0;
  double4 *gpos = nullptr, *spos = nullptr;
  uint32_t *cell_start = nullptr, *order = nullptr, *slot_of = nullptr;
  uint32_t *prob = nullptr, *prob_of_slot = nullptr;  // nullptr when P == 1
  uint64_t *d_goff = nullptr;
  BoxParams *d_boxes = nullptr;
  GridParams *d_grids = nullptr;
  uint32_t *d_coff = nullptr;
  // the dense cells' sub-columns (SubCols; null for 1D and 2D builds)
  uint32_t *sub_base = nullptr, *sub_start = nullptr, *sub_slot = nullptr, *sub_key = nullptr;
  double *sub_xy = nullptr;
  uint32_t dense_cells = 0;
  Problems PR() const {
    return Problems{d_boxes, d_grids, d_coff, prob_of_slot, sub_base, sub_start, sub_slot,
                    sub_key, sub_xy};
  }
};

// The sub-column tables of the dense cells (of the large-cell list the
// grid sort made), allocated from TT and built on s; cnt: words used in the
// two pools, then the cells with sub-columns.  A cell of m > 256 seeds has
// S^2 <= 3 m / 32 + 1 sub-columns (S = ceil(sqrt(m / 16)) <= m / 64 + 1):
// tables of S^2 + 4 words, at most n / 8 + n / 64 in all, and 2 (S + 1) <=
// 128 boundaries.
struct SubTables {
  uint32_t *base, *start, *slot, *key;
  double *xy;
  unsigned int *cnt;
};
SubTables launch_subcols(size_t n, size_t ncell, const uint32_t *cell_start,
                         const uint32_t *big_list, const unsigned int *big_cnt,
                         const double4 *spos, const GridParams *d_grids, const uint32_t *d_coff,
                         const uint32_t *prob_of_slot, Temps &TT, hipStream_t s) {
  const uint64_t pool_cap = (uint64_t)n / 8 + (uint64_t)n / 64 + 64;
  const uint64_t xy_cap = (uint64_t)n / 2 + 256;
  SubTables t;
  t.base = need(TT.get<uint32_t>(ncell));
  t.start = need(TT.get<uint32_t>(pool_cap));
  t.xy = need(TT.get<double>(xy_cap));
  t.slot = need(TT.get<uint32_t>(n));
  t.key = need(TT.get<uint32_t>(n));
  t.cnt = need(TT.get<unsigned int>(3));
  hipLaunchKernelGGL(k_fill_u32, dim3(nblk(ncell)), dim3(256), 0, s, t.base, (uint32_t)ncell,
                     ~0u);
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, (uint32_t *)t.cnt, 3u, 0u);
  hipLaunchKernelGGL(k_grid_subcols, dim3(2048), dim3(64), 0, s, cell_start, big_list, big_cnt,
                     spos, d_grids, d_coff, prob_of_slot, t.base, t.start, (uint32_t)pool_cap,
                     t.cnt, t.xy, (uint32_t)xy_cap, t.cnt + 1, t.slot, t.key);
  HIPCHK(hipGetLastError());
  return t;
}

void build_grid(GridData &D, const double *d_xyz, size_t n,
                const std::vector<ProblemSpec> &probs, uint32_t dim, double occ, Temps &TT,
                hipStream_t s) {
  const uint32_t P = (uint32_t)probs.size();
  D.P = P;
  std::vector<uint64_t> goff(P + 1);
  uint64_t maxn = 1;
  for (uint32_t p = 0; p < P; p++) {
    goff[p] = probs[p].g0;
    maxn = std::max<uint64_t>(maxn, probs[p].g1 - probs[p].g0);
  }
  goff[P] = n;
  D.d_goff = need(TT.get<uint64_t>(P + 1));
  HIPCHK(hipMemcpyAsync(D.d_goff, goff.data(), (P + 1) * sizeof(uint64_t),
                        hipMemcpyHostToDevice, s));
  D.gpos = need(TT.get<double4>(n));
  const unsigned nb = nblk(n);
  hipLaunchKernelGGL(k_prepare, dim3(nb), dim3(256), 0, s, d_xyz, n, (int)dim, D.gpos);
  if (P > 1) {
    D.prob = need(TT.get<uint32_t>(n));
    D.prob_of_slot = need(TT.get<uint32_t>(n));
    const uint32_t gx = (uint32_t)std::min<uint64_t>((maxn + 255) / 256, 64);
    hipLaunchKernelGGL(k_problem_of, dim3(gx, P), dim3(256), 0, s, (const uint64_t *)D.d_goff,
                       P, D.prob);
  }
  // bounding boxes
  const uint32_t S = (uint32_t)((maxn + BBOX_CHUNK - 1) / BBOX_CHUNK);
  double *bbp = need(TT.get<double>(6 * (size_t)S * P));
  double *bbd = need(TT.get<double>(6 * (size_t)P));
  hipLaunchKernelGGL(k_bbox_part, dim3(S, P), dim3(256), 0, s, (const double4 *)D.gpos,
                     (const uint64_t *)D.d_goff, S, bbp);
  hipLaunchKernelGGL(k_bbox_final, dim3((P + 63) / 64), dim3(64), 0, s, (const double *)bbp, S,
                     P, bbd);
  HIPCHK(hipGetLastError());
  std::vector<double> bb(6 * (size_t)P);
  HIPCHK(hipMemcpyAsync(bb.data(), bbd, bb.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  D.boxes.resize(P);
  D.grids.resize(P);
  D.coff.resize(P + 1);
  size_t ncell = 0;
  for (uint32_t p = 0; p < P; p++) {
    const uint64_t np_ = probs[p].g1 - probs[p].g0;
    D.boxes[p] = make_box(probs[p].anchor, probs[p].width, (int)probs[p].periodic, (int)dim);
    double b6[6];
    for (int c = 0; c < 6; c++) {
      b6[c] = bb[6 * p + c];
      if (np_ && !std::isfinite(b6[c])) throw BuildError{MV_E_DEGENERATE_POINTS};
      if (!np_) b6[c] = 0.;
    }
    D.grids[p] = make_grid(b6, np_ ? np_ : 1, probs[p].width, (int)dim, (int)probs[p].periodic,
                           occ);
    // total stream of this problem only
    uint64_t nimg = (uint64_t)(D.grids[p].img_hi[0] - D.grids[p].img_lo[0] + 1) *
                    (D.grids[p].img_hi[1] - D.grids[p].img_lo[1] + 1) *
                    (D.grids[p].img_hi[2] - D.grids[p].img_lo[2] + 1);
    D.grids[p].total_stream = np_ * nimg;
    D.coff[p] = (uint32_t)ncell;
    ncell += (size_t)D.grids[p].G[0] * D.grids[p].G[1] * D.grids[p].G[2];
    if (ncell >= (1ull << 32) - 1) throw BuildError{MV_E_CAPACITY};
  }
  D.coff[P] = (uint32_t)ncell;
  D.ncell = ncell;
  D.d_boxes = need(TT.get<BoxParams>(P));
  D.d_grids = need(TT.get<GridParams>(P));
  D.d_coff = need(TT.get<uint32_t>(P + 1));
  HIPCHK(hipMemcpyAsync(D.d_boxes, D.boxes.data(), P * sizeof(BoxParams), hipMemcpyHostToDevice,
                        s));
  HIPCHK(hipMemcpyAsync(D.d_grids, D.grids.data(), P * sizeof(GridParams),
                        hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(D.d_coff, D.coff.data(), (P + 1) * sizeof(uint32_t),
                        hipMemcpyHostToDevice, s));
  uint32_t *cellid = need(TT.get<uint32_t>(n));
  uint32_t *counts = need(TT.get<uint32_t>(ncell + 1));
  uint32_t *fill = need(TT.get<uint32_t>(ncell));
  D.cell_start = need(TT.get<uint32_t>(ncell + 1));
  D.order = need(TT.get<uint32_t>(n));
  D.spos = need(TT.get<double4>(n));
  D.slot_of = need(TT.get<uint32_t>(n));
  HIPCHK(hipMemsetAsync(counts, 0, (ncell + 1) * sizeof(uint32_t), s));
  HIPCHK(hipMemsetAsync(fill, 0, ncell * sizeof(uint32_t), s));
  hipLaunchKernelGGL(k_grid_count, dim3(nb), dim3(256), 0, s, (const double4 *)D.gpos, n,
                     (const GridParams *)D.d_grids, (const uint32_t *)D.d_coff,
                     (const uint32_t *)D.prob, cellid, counts);
  scan_excl<uint32_t, uint32_t>(counts, D.cell_start, ncell + 1, s);
  hipLaunchKernelGGL(k_grid_scatter, dim3(nb), dim3(256), 0, s, cellid, n, D.cell_start, fill,
                     D.order);
  {
    // after the scatter, `fill` holds the list of large cells and `counts`
    // their count; spos, written only by the gather below, serves as the
    // two sort buffers (8 of its 32 bytes per seed each)
    uint32_t *big_list = fill;
    unsigned int *big_cnt = counts;
    uint64_t *ta = reinterpret_cast<uint64_t *>(D.spos);
    uint64_t *tb = ta + n;
    HIPCHK(hipMemsetAsync(big_cnt, 0, sizeof(unsigned int), s));
    hipLaunchKernelGGL(k_grid_sort_cells, dim3(nblk(ncell)), dim3(256), 0, s, D.cell_start,
                       ncell, (const double4 *)D.gpos, D.order, big_list, big_cnt);
    // the cells over GRID_SMALL seeds (their count stays on the device)
    hipLaunchKernelGGL(k_grid_sort_big, dim3(1024), dim3(GS_THREADS), 0, s,
                       (const uint32_t *)D.cell_start, (const uint32_t *)big_list,
                       (const unsigned int *)big_cnt, (const double4 *)D.gpos, D.order, ta,
                       tb);
  }
  hipLaunchKernelGGL(k_grid_gather, dim3(nb), dim3(256), 0, s, (const double4 *)D.gpos,
                     (const uint32_t *)D.order, n, (const uint32_t *)D.prob, D.spos, D.slot_of,
                     D.prob_of_slot);
  if (dim == 3) {
    const SubTables t = launch_subcols(n, ncell, D.cell_start, fill, counts, D.spos, D.d_grids,
                                       D.d_coff, D.prob_of_slot, TT, s);
    // no dense cell: the gathers run without the sub-column path
    hipLaunchKernelGGL(k_word_u32, dim3(1), dim3(1), 0, s, (const uint32_t *)(t.cnt + 1),
                       host_word());
    unsigned long long w2[2];
    read_words(s, w2, 2);
    D.dense_cells = (uint32_t)w2[1];
    if (D.dense_cells) {
      D.sub_base = t.base;
      D.sub_start = t.start;
      D.sub_slot = t.slot;
      D.sub_key = t.key;
      D.sub_xy = t.xy;
    }
  }
  HIPCHK(hipGetLastError());
}

// feature normalisation of each problem: its box and mean spacing
std::vector<FeatParams> feat_params(const std::vector<ProblemSpec> &probs, int dim) {
  std::vector<FeatParams> out(probs.size());
  for (size_t p = 0; p < probs.size(); p++) {
    FeatParams &fp = out[p];
    double vol = 1.;
    int nd = 0;
    for (int c = 0; c < 3; c++) {
      fp.box_lo[c] = probs[p].anchor[c];
      fp.box_hi[c] = probs[p].anchor[c] + probs[p].width[c];
      if (c < dim) {
        vol *= probs[p].width[c];
        nd++;
      }
    }
    const double np_ = (double)std::max<uint64_t>(probs[p].g1 - probs[p].g0, 1);
    fp.h_global = std::pow(vol / np_, 1.0 / std::max(nd, 1));
  }
  return out;
}

void upload_model(const mv_model *m, DevModel *d, hipStream_t s) {
  DevModel h;
  memset(&h, 0, sizeof h);
  for (int i = 0; i < NF; i++) {
    h.mu[i] = m->mu[i];
    h.sigma[i] = m->sigma[i] > 0.f ? m->sigma[i] : 1.f;
  }
  memcpy(h.w1, m->w1, sizeof h.w1);
  memcpy(h.w2, m->w2, sizeof h.w2);
  for (int i = 0; i < NH; i++)
    for (int o = 0; o < NO; o++) h.w3[i * NOP + o] = (short)m->w3[i * NO + o];
  memcpy(h.b1, m->b1, sizeof h.b1);
  memcpy(h.b2, m->b2, sizeof h.b2);
  for (int o = 0; o < NO; o++) h.b3[o] = m->b3[o];
  HIPCHK(hipMemcpyAsync(d, &h, sizeof h, hipMemcpyHostToDevice, s));
  HIPCHK(hipStreamSynchronize(s));
}

// What a build did, pass by pass (recorded for mv_graph_create's plan):
// per tier (0 main, 1 large cells, 2 wide) the cells, the list-length bound
// the list kernel was chosen for, and the list entries.
struct PassRec {
  uint32_t cnt[3] = {0, 0, 0};
  uint32_t lbnd[3] = {0, 0, 0};
  uint64_t total[3] = {0, 0, 0};
};
struct BuildTrace {
  std::vector<PassRec> passes;
  std::vector<BoxParams> boxes;
  std::vector<GridParams> grids;
  std::vector<uint32_t> coff;
  size_t ncell = 0;
  uint32_t lbound0 = 0;
  uint64_t st_cap = 0, F = 0, NL = 0;
  uint32_t dense_cells = 0;
};
thread_local BuildTrace *t_trace = nullptr;

// One build over P independent problems (P = 1: Voronoi::build /
// build_partial).  Ts[p] receives problem p; Ts[0] owns the device memory,
// which the other handles share.  All kernels run once over the
// concatenation: slots, generator indices and faces are global during the
// build and made problem-relative at the end.
void build_multi(const std::vector<mv_tess *> &Ts, const double *d_xyz, size_t n,
                 const uint8_t *d_mask, const std::vector<ProblemSpec> &probs, uint32_t dim,
                 const mv_options &opt) {
  auto t_start = std::chrono::steady_clock::now();
  hipStream_t s = 0;
  mv_tess *T = Ts[0];
  const uint32_t P = (uint32_t)probs.size();
  for (uint32_t p = 1; p < P; p++) Ts[p]->store = T->store;
  T->dim = dim;
  T->periodic = probs[0].periodic;
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
  for (uint32_t p = 0; p < P; p++) {
    memcpy(Ts[p]->anchor, probs[p].anchor, sizeof T->anchor);
    memcpy(Ts[p]->width, probs[p].width, sizeof T->width);
  }
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
    for (uint32_t p = 1; p < P; p++) {
      mv_tess *U = Ts[p];
      U->dim = dim;
      U->periodic = probs[p].periodic;
      U->d_cell_loc = T->d_cell_loc; U->d_cell_cen = T->d_cell_cen;
      U->d_cell_vol = T->d_cell_vol; U->d_cell_sr = T->d_cell_sr;
      U->d_cell_foff = T->d_cell_foff; U->d_cell_fcnt = T->d_cell_fcnt;
      U->d_cell_idx = T->d_cell_idx; U->d_npop = T->d_npop;
      U->d_face_left = T->d_face_left; U->d_face_right = T->d_face_right;
      U->d_face_area = T->d_face_area; U->d_face_cen = T->d_face_cen;
      U->d_face_norm = T->d_face_norm; U->d_face_shift = T->d_face_shift;
      U->d_face_has_shift = T->d_face_has_shift; U->d_links = T->d_links;
    }
    return;
  }

  Temps TT;
  struct Stream2 {
    hipStream_t h = nullptr;
    hipEvent_t ev = nullptr;
    ~Stream2() {
      if (ev) (void)hipEventDestroy(ev);
      if (h) (void)hipStreamDestroy(h);
    }
  } st2;
  // the allocation-free build creates no stream: its big lists run on s
  if (!arena_active()) {
    HIPCHK(hipStreamCreateWithFlags(&st2.h, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&st2.ev, hipEventDisableTiming));
  }
  hipStream_t s2 = arena_active() ? s : st2.h;
  // ---- grid
  Timer tg(s);
  GridData D;
  const double occ = opt.grid_occupancy16 ? (double)opt.grid_occupancy16 / 16.0 : 2.0;
  build_grid(D, d_xyz, n, probs, dim, occ, TT, s);
  if (t_trace) {
    t_trace->boxes = D.boxes;
    t_trace->grids = D.grids;
    t_trace->coff = D.coff;
    t_trace->ncell = D.ncell;
    t_trace->dense_cells = D.dense_cells;
  }
  const Problems PR = D.PR();
  const double4 *spos = D.spos;
  const uint32_t *cell_start = D.cell_start;
  const unsigned nb = nblk(n);

  // ---- initial work list: active slots in grid order
  uint32_t *flag = need(TT.get<uint32_t>(n));
  uint32_t *ident = need(TT.get<uint32_t>(n));
  uint32_t *pos = need(TT.get<uint32_t>(n));
  // [0]: first-pass list, [1]/[2]: main lists of later passes, [3]/[4]:
  // lists of the large-cell kernel, [5]/[6]: lists of the wide kernel
  uint32_t *lists[7], *lens[7];
  for (int k = 0; k < 7; k++) {
    lists[k] = need(TT.get<uint32_t>(n));
    lens[k] = need(TT.get<uint32_t>(n));
  }
  hipLaunchKernelGGL(k_flag_active, dim3(nb), dim3(256), 0, s, spos, n, d_mask, flag, ident);
  uint32_t nwork = scan_excl<uint32_t, uint32_t>(flag, pos, n, s);
  if (nwork)
    hipLaunchKernelGGL(k_compact, dim3(nb), dim3(256), 0, s, ident, flag, pos, (uint32_t)n,
                       lists[0]);
  HIPCHK(hipGetLastError());
  T->stats.ms_grid = tg.stop();

  // ---- list lengths of the first pass
  const uint32_t kmax =
      std::max<uint32_t>(std::min<uint32_t>(opt.k_max ? opt.k_max : MV_K_MAX, MV_K_MAX), 2);
  const uint32_t k0 = std::min(std::max<uint32_t>(opt.k0 ? opt.k0 : 27, 2), kmax);
  float *Rfirst = nullptr;
  uint32_t lbound = k0;  // upper bound of the current main list's lengths
  double ms_model = 0;
  if (opt.mode == MV_MODE_LEARNED && nwork) {
    Timer tm(s);
    float *F = need(TT.get<float>((size_t)nwork * NF));
    float *hl = need(TT.get<float>(nwork));
    float *Y = need(TT.get<float>((size_t)nwork * NO));
    DevModel *dm = need(TT.get<DevModel>(1));
    upload_model(opt.model, dm, s);
    const std::vector<FeatParams> fph = feat_params(probs, (int)dim);
    FeatParams *fpd = need(TT.get<FeatParams>(P));
    HIPCHK(hipMemcpyAsync(fpd, fph.data(), P * sizeof(FeatParams), hipMemcpyHostToDevice, s));
    // waves with a seed in a dense grid cell, then every other wave
    hipLaunchKernelGGL(k_features<true>, dim3(nblk(nwork)), dim3(256), 0, s, spos, cell_start,
                       PR, (const FeatParams *)fpd, lists[0], nwork, F, hl);
    hipLaunchKernelGGL(k_features<false>, dim3(nblk(nwork)), dim3(256), 0, s, spos, cell_start,
                       PR, (const FeatParams *)fpd, lists[0], nwork, F, hl);
    const uint32_t tiles = (nwork + 15) / 16;
    hipLaunchKernelGGL(k_mlp_infer, dim3(std::min<uint32_t>((tiles + 3) / 4, 4096)), dim3(256),
                       0, s, dm, F, nwork, Y);
    float *R0 = need(TT.get<float>(nwork));
    const float kscale = opt.model->kscale > 0.f ? opt.model->kscale : 1.f;
    // first-pass lists are cut to first_list_cap (a multiple of 8 in
    // [16, 1024]; 0 = 256)
    const uint32_t cap0 = opt.first_list_cap ? opt.first_list_cap : 256u;
    lbound = std::min<uint32_t>(std::min<uint32_t>(std::max<uint32_t>(cap0, 16u), 1024u) & ~7u,
                                kmax);
    const float rscale = opt.model->rscale > 0.f ? opt.model->rscale : 0.f;
    hipLaunchKernelGGL(k_proposal, dim3(nblk(nwork)), dim3(256), 0, s, Y, hl, nwork,
                       std::min<uint32_t>(16, kmax), lbound, kscale, rscale, lens[0], R0);
    // longest lists first (or grid order on request)
    const uint32_t NB = std::min<uint32_t>(lbound / 8 + 1, LEN_MAX_BUCKETS);
    uint32_t *hist = need(TT.get<uint32_t>(NB + 1));
    uint32_t *bst = need(TT.get<uint32_t>(NB + 1));
    uint32_t *bfill = need(TT.get<uint32_t>(NB));
    HIPCHK(hipMemsetAsync(hist, 0, (NB + 1) * sizeof(uint32_t), s));
    HIPCHK(hipMemsetAsync(bfill, 0, NB * sizeof(uint32_t), s));
    if (opt.flags & MV_FLAG_NO_LENGTH_ORDER) {
      Rfirst = R0;
    } else {
      const uint32_t lblocks = (nwork + LEN_PER_BLOCK - 1) / LEN_PER_BLOCK;
      hipLaunchKernelGGL(k_len_hist, dim3(lblocks), dim3(256), 0, s, lens[0], nwork, NB, hist);
      scan_excl<uint32_t, uint32_t>(hist, bst, NB + 1, s);
      Rfirst = need(TT.get<float>(nwork));
      hipLaunchKernelGGL(k_len_scatter, dim3(lblocks), dim3(256), 0, s, lists[0], lens[0], R0,
                         nwork, NB, bst, bfill, lists[1], lens[1], Rfirst);
      HIPCHK(hipGetLastError());
      std::swap(lists[0], lists[1]);
      std::swap(lens[0], lens[1]);
    }
    ms_model = tm.stop();
  } else if (nwork) {
    hipLaunchKernelGGL(k_fill_u32, dim3(nblk(nwork)), dim3(256), 0, s, lens[0], nwork, k0);
  }

  // ---- per-cell outputs and face staging
  uint32_t *cell_nf = need(TT.get<uint32_t>(n));
  uint64_t *cell_fbase = need(TT.get<uint64_t>(n));
  HIPCHK(hipMemsetAsync(cell_nf, 0, n * sizeof(uint32_t), s));
  HIPCHK(hipMemsetAsync(cell_fbase, 0, n * sizeof(uint64_t), s));
  // links from lower cells' faces, counted as the faces are emitted
  uint32_t *link_in = need(TT.get<uint32_t>(n));
  HIPCHK(hipMemsetAsync(link_in, 0, n * sizeof(uint32_t), s));
  uint64_t st_cap = std::max<uint64_t>(4096, 16ull * nwork);
  unsigned long long *counters = need(TT.get<unsigned long long>(NCOUNTERS));
  // counters: [0] staging cursor, [1] error word, [2] exact tests, [3] overflow,
  // [4..6] the long-list kernel's items, ball gathers and entries gathered,
  // [7..71) the wavefront list kernels' ball gathers, [71..135) their
  // candidates read (64 counters each)
  struct KnnStatScope {
    explicit KnnStatScope(unsigned long long *p) { t_knn_stat = p; }
    ~KnnStatScope() { t_knn_stat = nullptr; }
  } knn_stat_scope(counters + 4);
  // kept cells: arenas sized for ~48 half-spaces and ~64 vertices per cell,
  // doubled when a pass runs out
  T->keep = (opt.flags & MV_FLAG_KEEP_CELLS) != 0;
  unsigned long long *cs_cur = need(TT.get<unsigned long long>(2));
  uint64_t cap_pl = 0, cap_vx = 0;
  auto alloc_store = [&]() {
    if (T->cs.pl_a) {
      T->release(T->cs.pl_a);
      T->release(T->cs.pl_b);
      T->release(T->cs.vx);
    }
    T->cs.pl_a = T->own<double4>(cap_pl);
    T->cs.pl_b = T->own<double4>(cap_pl);
    T->cs.vx = T->own<double4>(cap_vx);
    T->cs.cap_pl = cap_pl;
    T->cs.cap_vx = cap_vx;
  };
  if (T->keep) {
    T->cs.pbase = T->own<uint64_t>(n);
    T->cs.vbase = T->own<uint64_t>(n);
    T->cs.np = T->own<uint32_t>(n);
    T->cs.nv = T->own<uint32_t>(n);
    HIPCHK(hipMemsetAsync(T->cs.np, 0, n * sizeof(uint32_t), s));
    HIPCHK(hipMemsetAsync(T->cs.nv, 0, n * sizeof(uint32_t), s));
    T->cs.cur_pl = cs_cur;
    T->cs.cur_vx = cs_cur + 1;
    cap_pl = 48ull * nwork + 4096;
    cap_vx = 64ull * nwork + 4096;
    alloc_store();
    if (d_mask) {
      T->d_mask = T->own<uint8_t>(n);
      HIPCHK(hipMemcpyAsync(T->d_mask, d_mask, n, hipMemcpyDeviceToDevice, s));
    }
  }
  uint8_t *wstate = need(TT.get<uint8_t>(n));
  uint8_t *bstate = need(TT.get<uint8_t>(n));
  uint8_t *xstate = need(TT.get<uint8_t>(n));
  // wide-kernel capacities (planes, vertices per cell), doubled when a
  // cell runs out
  uint32_t wide_p = WIDE_P0, wide_v = WIDE_V0;
  // work-queue heads of the clip launches: [0] main, [1] cells handed to the
  // exact kernel, [2] its queue, [3] the resume launch's queue, [4] cells the
  // pair kernel handed over
  unsigned int *queue = need(TT.get<unsigned int>(5));
  HIPCHK(hipMemsetAsync(queue, 0, 5 * sizeof(unsigned int), s));
  // the reference's summation order, or 64 cells per wavefront: the
  // thread-per-cell kernel for the main lists
  const bool exact_order =
      (opt.flags & MV_FLAG_EXACT_ORDER) != 0 || opt.cells_per_wave == 64;
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
  LaneBufs LM{};
  bool have_main = false;
  uint64_t redone = 0, bigcells = 0, widecells = 0, handovers = 0;
  uint32_t passes = 0;
  double ms_knn = 0, ms_clip = 0;
  // per-kernel figures (reset with each attempt)
  double ms_knn_k = 0, ms_cw = 0;
  uint32_t n_knn = 0, n_cw = 0;
  uint64_t q_knn = 0, cells_cw = 0, e_knn = 0;
  const uint32_t nwork0 = nwork;

  uint32_t *st_cell = nullptr;
  int64_t *st_right = nullptr;
  uint8_t *st_img = nullptr;
  uint32_t *st_lslot = nullptr;
  double *st_area = nullptr, *st_cen = nullptr, *st_norm = nullptr;
  unsigned int *cap_err = need(TT.get<unsigned int>(1));
  for (int attempt = 0;; attempt++) {
    st_cell = need(TT.get<uint32_t>(st_cap));
    st_right = need(TT.get<int64_t>(st_cap));
    st_img = need(TT.get<uint8_t>(st_cap));
    st_lslot = need(TT.get<uint32_t>(st_cap));
    st_area = need(TT.get<double>(st_cap));
    st_cen = need(TT.get<double>(3 * st_cap));
    st_norm = need(TT.get<double>(3 * st_cap));
    HIPCHK(hipMemsetAsync(counters, 0, NCOUNTERS * sizeof(unsigned long long), s));
    HIPCHK(hipMemsetAsync(cs_cur, 0, 2 * sizeof(unsigned long long), s));
    unsigned long long err_init = ~0ull;
    HIPCHK(hipMemcpyAsync(counters + 1, &err_init, sizeof err_init, hipMemcpyHostToDevice, s));

    // the first pass reads lists[0], which is never written, so that an
    // attempt can start over; later passes alternate between two buffers
    uint32_t nw = nwork0, nbig = 0, nwide = 0;
    uint32_t *cw = lists[0], *cl = lens[0], *cb = lists[3], *cbl = lens[3];
    uint32_t *cx = lists[5], *cxl = lens[5];
    int ping_m = 1, ping_b = 4, ping_x = 6;
    // bounds on the list lengths of the main and big lists (all clamped to
    // kmax; a list whose lower bound is kmax cannot grow any more)
    uint32_t lb_main = std::min(lbound, kmax), lb_big = lb_main, lb_wide = lb_main;
    uint32_t nw_tail = 0;  // long lists at the end of the main list
    const float *Rbig = nullptr, *Rwide = nullptr;
    HIPCHK(hipMemsetAsync(cap_err, 0, sizeof(unsigned int), s));
    const float *Rmain = Rfirst;
    bool overflowed = false;
    unsigned long long ovf_bits = 0;
    passes = 0;
    redone = 0;
    handovers = 0;
    if (t_trace) t_trace->passes.clear();
    bigcells = 0;
    widecells = 0;
    ms_knn_k = ms_cw = 0;
    n_knn = n_cw = 0;
    q_knn = cells_cw = e_knn = 0;
    while (nw + nbig + nwide > 0) {
      passes++;
      if (t_trace) t_trace->passes.emplace_back();
      double *nr_main = need(TT.get<double>(std::max<uint32_t>(nw, 1)));
      double *nr_big = need(TT.get<double>(std::max<uint32_t>(nbig, 1)));
      double *nr_wide = need(TT.get<double>(std::max<uint32_t>(nwide, 1)));
      // this pass's kdist per list
      const double *kd_main = nullptr, *kd_big = nullptr, *kd_wide = nullptr;
      // events of a pipelined first pass (destroyed once the pass is done)
      struct PipeEvents {
        std::vector<hipEvent_t> ev;
        ~PipeEvents() {
          for (hipEvent_t e : ev) (void)hipEventDestroy(e);
        }
      } pev;
      // one list on one stream; untimed runs do not wait for their kernels
      // tier 0: the main list (wavefront kernel), 1: the large-cell kernel,
      // 2: the wide kernel; tiers 1 and 2 gather with the long-list kernel
      // ntail: the last ntail lists of the main list are the long ones
      auto run = [&](uint32_t *wl, uint32_t *ll, uint32_t cnt, uint32_t lbnd, const float *R0,
                     uint8_t *state, int tier, hipStream_t s, bool timed, uint32_t ntail = 0) {
        const bool big = tier > 0;
        double *need_r = tier == 2 ? nr_wide : big ? nr_big : nr_main;
        Timer tk(s, timed);
        uint64_t *kofs = need(TT.get<uint64_t>(cnt));
        // redone lists: the longest one picks the list kernel (the growth
        // bound is loose); read with the scan's total
        const uint32_t nhead = cnt - ntail;
        // the first pass's radii (when given) are proposals
        const bool first_prop = passes == 1 && R0 != nullptr;
        if (passes > 1) {
          hipLaunchKernelGGL(k_max_u32, dim3(1), dim3(1024), 0, s, (const uint32_t *)ll, nhead,
                             host_word() + 1);
          if (ntail)
            hipLaunchKernelGGL(k_max_u32, dim3(1), dim3(1024), 0, s,
                               (const uint32_t *)(ll + nhead), ntail, host_word() + 2);
        }
        const uint64_t total = scan_excl<uint32_t, uint64_t>(ll, kofs, cnt, s);
        uint32_t ltail = 0;
        if (passes > 1) {
          unsigned long long w2[3];
          read_words(s, w2, 3);
          ltail = ntail ? std::min<uint32_t>(lbnd, (uint32_t)w2[2]) : 0u;
          lbnd = std::min<uint32_t>(lbnd, (uint32_t)w2[1]);
        }
        uint64_t *kpk = need(TT.get<uint64_t>(total));
        uint32_t *kcnt = need(TT.get<uint32_t>(cnt));
        double *kdist = need(TT.get<double>(cnt));
        (tier == 2 ? kd_wide : big ? kd_big : kd_main) = kdist;
        if (t_trace) {
          PassRec &r = t_trace->passes.back();
          r.cnt[tier] = cnt;
          r.lbnd[tier] = lbnd;
          r.total[tier] = total;
        }
        // pipelined first pass (main list, one cell per wavefront): the lists
        // of chunk k + 1 are built on the second stream, with a few blocks
        // per CU, while chunk k is clipped (the clip kernel's blocks leave
        // room for them and the list kernel's work fills the VALU issue slots
        // the clipper leaves)
        const uint32_t pchunks = opt.pipeline & 0xFFu;
        const uint32_t pblocks = std::max<uint32_t>((opt.pipeline >> 8) & 0xFFu, 1u);
        const bool pipe = pchunks > 1 && passes == 1 && tier == 0 && !exact_order && !T->keep &&
                          opt.cells_per_wave != 2 && s2 != s && cnt >= 64u * pchunks;
        std::vector<uint32_t> cbnd;
        Timer tkk(s, timed);
        if (!pipe && ntail && nhead) {
          // short lists on the wavefront list kernels, the long tail on the
          // long-list kernel
          launch_knn(spos, cell_start, PR, wl, nhead, lbnd, ll, R0, kofs, kpk, kcnt, kdist, false,
                     TT, s, opt.knn_min_tile);
          launch_knn(spos, cell_start, PR, wl + nhead, ntail, ltail, ll + nhead,
                     R0 ? R0 + nhead : nullptr, kofs + nhead, kpk, kcnt + nhead, kdist + nhead,
                     true, TT, s);
        } else if (!pipe) {
          launch_knn(spos, cell_start, PR, wl, cnt, std::max(lbnd, ltail), ll, R0, kofs, kpk,
                     kcnt, kdist, big, TT, s, opt.knn_min_tile, nullptr, 0, 0, first_prop);
        } else {
          cbnd.resize(pchunks + 1);
          for (uint32_t k = 0; k <= pchunks; k++) cbnd[k] = (uint32_t)((uint64_t)cnt * k / pchunks);
          pev.ev.assign(pchunks, nullptr);
          for (hipEvent_t &e : pev.ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
          // the second stream starts after the offsets
          HIPCHK(hipEventRecord(pev.ev[0], s));
          HIPCHK(hipStreamWaitEvent(s2, pev.ev[0], 0));
          launch_knn(spos, cell_start, PR, wl, cbnd[1], lbnd, ll, R0, kofs, kpk, kcnt, kdist, false,
                     TT, s, opt.knn_min_tile, nullptr, 0, 0, first_prop);
          for (uint32_t k = 1; k < pchunks; k++) {
            const uint32_t c0 = cbnd[k];
            launch_knn(spos, cell_start, PR, wl + c0, cbnd[k + 1] - c0, lbnd, ll + c0,
                       R0 ? R0 + c0 : nullptr, kofs + c0, kpk, kcnt + c0, kdist + c0, false, TT,
                       s2, opt.knn_min_tile, nullptr, 0, (uint32_t)device_cus() * pblocks,
                       first_prop);
            HIPCHK(hipEventRecord(pev.ev[k], s2));
          }
        }
        ms_knn_k += tkk.stop();
        n_knn++;
        q_knn += cnt;
        e_knn += total;
        ms_knn += tk.stop();
        Timer tc(s, timed);
        if (!big && !exact_order) {
          WaveArgs W;
          memset(&W, 0, sizeof W);
          W.spos = spos;
          W.mask = d_mask;
          W.PR = PR;
          W.work = wl;
          W.nwork = cnt;
          W.kpk = kpk;
          W.kofs = kofs;
          W.kcnt = kcnt;
          W.kdist = kdist;
          W.need_r = need_r;
          W.L = 0;
          W.queue = queue;
          W.wstate = state;
          W.cell_vol = T->d_cell_vol;
          W.cell_cen = T->d_cell_cen;
          W.cell_sr = T->d_cell_sr;
          W.cell_nf = cell_nf;
          W.link_in = link_in;
          W.cell_fbase = cell_fbase;
          W.cell_npop = T->d_npop;
          W.st_cell = st_cell;
          W.st_right = st_right;
          W.st_img = st_img;
          W.st_lslot = st_lslot;
          W.st_area = st_area;
          W.st_cen = st_cen;
          W.st_norm = st_norm;
          W.st_cap = st_cap;
          W.st_counter = counters;
          W.err_word = counters + 1;
          W.exact_counter = counters + 2;
          W.overflow = counters + 3;
          if (T->keep) W.cs = T->cs;
          HIPCHK(hipMemsetAsync(queue, 0, 5 * sizeof(unsigned int), s));
          const uint32_t bpc = opt.clip_blocks_per_cu ? opt.clip_blocks_per_cu : 4u;
          uint32_t blocks = std::min<uint32_t>((cnt + 3) / 4, (uint32_t)device_cus() * bpc);
          if (!T->keep) {
            W.exact_list = need(TT.get<uint32_t>(cnt));
            W.exact_cnt = queue + 1;
          }
          // two cells per wavefront (cells that outgrow it continue on the
          // wavefront kernel from their saved state)
          const bool pair = !T->keep && opt.cells_per_wave == 2;
          if (pair) {
            const uint32_t hcap = std::min<uint32_t>(cnt, cnt / 4 + 256);
            W.ho = need(TT.get<Handover>(hcap));
            W.ho_list = need(TT.get<uint32_t>(hcap));
            W.ho_cnt = queue + 4;
            W.ho_cap = hcap;
          }
          Timer tcw(s, timed);
          if (T->keep) {
            hipLaunchKernelGGL((k_clip_wave<true, true>), dim3(blocks), dim3(256), 0, s, W);
          } else if (pipe) {
            // chunk k after its lists
            unsigned int *qk = need(TT.get<unsigned int>(pchunks));
            HIPCHK(hipMemsetAsync(qk, 0, pchunks * sizeof(unsigned int), s));
            for (uint32_t k = 0; k < pchunks; k++) {
              if (k) HIPCHK(hipStreamWaitEvent(s, pev.ev[k], 0));
              WaveArgs Wk = W;
              Wk.w0 = cbnd[k];
              Wk.nwork = cbnd[k + 1] - cbnd[k];
              Wk.queue = qk + k;
              const uint32_t bk =
                  std::min<uint32_t>((Wk.nwork + 3) / 4, (uint32_t)device_cus() * bpc);
              hipLaunchKernelGGL((k_clip_wave<false, false>), dim3(bk), dim3(256), 0, s, Wk);
            }
          } else if (pair) {
            // four blocks of four wavefronts fill a CU's LDS
            const uint32_t pblocks =
                std::min<uint32_t>((cnt + 7) / 8, (uint32_t)device_cus() * std::min(bpc, 4u));
            hipLaunchKernelGGL(k_clip_pair, dim3(pblocks), dim3(256), 0, s, W);
            HIPCHK(hipGetLastError());
            WaveArgs R = W;
            R.sub = W.ho_list;
            R.nsub = W.ho_cnt;
            R.queue = queue + 3;
            R.resume = W.ho;
            hipLaunchKernelGGL((k_clip_wave<false, false, true>), dim3(blocks), dim3(256), 0, s,
                               R);
          } else {
            hipLaunchKernelGGL((k_clip_wave<false, false>), dim3(blocks), dim3(256), 0, s, W);
          }
          HIPCHK(hipGetLastError());
          ms_cw += tcw.stop();
          n_cw++;
          if (!T->keep) {
            // the cells that met a near-degenerate test, with the exact
            // predicate (the count stays on the device: an empty list costs
            // one queue claim per wavefront)
            WaveArgs X = W;
            X.sub = W.exact_list;
            X.nsub = W.exact_cnt;
            X.queue = queue + 2;
            hipLaunchKernelGGL((k_clip_wave<false, true>), dim3(blocks), dim3(256), 0, s, X);
            HIPCHK(hipGetLastError());
          }
          ms_clip += tc.stop();
          return;
        }
        ClipArgs A;
        memset(&A, 0, sizeof A);
        A.spos = spos;
        A.mask = d_mask;
        A.PR = PR;
        A.work = wl;
        A.nwork = cnt;
        A.kpk = kpk;
        A.kofs = kofs;
        A.kcnt = kcnt;
        A.kdist = kdist;
        A.need_r = need_r;
        A.L = 0;
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
          // k_clip_coop keeps each cell's arrays in LDS
          lanes = std::min<uint32_t>(lanes_big, cnt);
          A.lanes = lanes;
        }
        A.wstate = state;
        A.cell_vol = T->d_cell_vol;
        A.cell_cen = T->d_cell_cen;
        A.cell_sr = T->d_cell_sr;
        A.cell_nf = cell_nf;
        A.link_in = link_in;
        A.cell_fbase = cell_fbase;
        A.cell_npop = T->d_npop;
        A.st_cell = st_cell;
        A.st_right = st_right;
        A.st_img = st_img;
        A.st_lslot = st_lslot;
        A.st_area = st_area;
        A.st_cen = st_cen;
        A.st_norm = st_norm;
        A.st_cap = st_cap;
        A.st_counter = counters;
        A.err_word = counters + 1;
        A.exact_counter = counters + 2;
        A.overflow = counters + 3;
        if (T->keep) A.cs = T->cs;
        A.coop_maxp = opt.coop_max_planes;
        A.coop_maxv = opt.coop_max_vertices;
        if (tier == 2) {
          // one slab per block, its planes and vertices sized for the cells
          const uint32_t blocks = std::min<uint32_t>(cnt, WIDE_BLOCKS);
          A.wmaxp = wide_p;
          A.wmaxv = wide_v;
          A.wplA = need(TT.get<double4>((size_t)blocks * wide_p));
          A.wplB = need(TT.get<double4>((size_t)blocks * wide_p));
          A.wfacc = need(TT.get<double4>((size_t)blocks * wide_p));
          A.wcyc = need(TT.get<uint32_t>((size_t)blocks * wide_p));
          A.wvtx = need(TT.get<double4>((size_t)blocks * wide_v));
          A.wvdual = need(TT.get<uint64_t>((size_t)blocks * wide_v));
          A.wflag = need(TT.get<uint8_t>((size_t)blocks * wide_v));
          hipLaunchKernelGGL(k_clip_wide, dim3(blocks), dim3(64), 0, s, A);
          HIPCHK(hipGetLastError());
          ms_clip += tc.stop();
          return;
        }
        const dim3 grid(lanes / 256);
        // large cells: one cell per wavefront, a block each
        const dim3 grid_coop(std::max<uint32_t>(lanes, 1));
        if (!big && !T->keep)
          hipLaunchKernelGGL((k_clip<MAXP_MAIN, MAXV_MAIN, false>), grid, dim3(256), 0, s, A);
        else if (!big)
          hipLaunchKernelGGL((k_clip<MAXP_MAIN, MAXV_MAIN, true>), grid, dim3(256), 0, s, A);
        else if (!T->keep)
          hipLaunchKernelGGL((k_clip_coop<MAXP_BIG, MAXV_BIG, false>), grid_coop, dim3(64), 0,
                             s, A);
        else
          hipLaunchKernelGGL((k_clip_coop<MAXP_BIG, MAXV_BIG, true>), grid_coop, dim3(64), 0,
                             s, A);
        HIPCHK(hipGetLastError());
        ms_clip += tc.stop();
      };
      // the big list (long lists, large cells: few items, long latency) runs
      // on its own stream beside the main list
      if (nbig && s2 != s) {
        // s2 starts after everything queued on s (the lists it reads)
        HIPCHK(hipEventRecord(st2.ev, s));
        HIPCHK(hipStreamWaitEvent(s2, st2.ev, 0));
      }
      if (nwide && !nbig && s2 != s) {
        HIPCHK(hipEventRecord(st2.ev, s));
        HIPCHK(hipStreamWaitEvent(s2, st2.ev, 0));
      }
      Timer tbig(s2, nbig + nwide > 0);
      if (nbig) run(cb, cbl, nbig, lb_big, Rbig, bstate, 1, s2, false);
      if (nwide) run(cx, cxl, nwide, lb_wide, Rwide, xstate, 2, s2, false);
      if (nw) run(cw, cl, nw, lb_main, Rmain, wstate, 0, s, true, nw_tail);
      if (nbig + nwide) ms_clip += tbig.stop();
      HIPCHK(hipStreamSynchronize(s2));
      unsigned long long cnts[6];
      hipLaunchKernelGGL(k_pass_words, dim3(1), dim3(64), 0, s, counters, cap_err,
                         (const unsigned int *)(queue + 4), host_word());
      HIPCHK(hipGetLastError());
      read_words(s, cnts, 6);
      handovers += cnts[5];
      const unsigned long long cap_err_h = cnts[4];
      // a cell that reached kmax entries without its security radius
      if (cap_err_h) throw BuildError{MV_E_CAPACITY};
      if (cnts[1] != ~0ull) {
        static const mv_status map[] = {MV_OK, MV_E_ARG, MV_E_DEGENERATE_POINTS,
                                        MV_E_SELF_NOT_FIRST, MV_E_PLANE_DET_ZERO,
                                        MV_E_NAN, MV_E_BOUNDARY_WALK};
        uint32_t code = (uint32_t)(cnts[1] & 0xFF);
        throw BuildError{code < 7 ? map[code] : MV_E_DEVICE};
      }
      if (cnts[3]) {
        ovf_bits = cnts[3];
        overflowed = true;
        break;
      }
      // next lists: cells over the on-chip capacity move to the big list
      // with their list; cells whose security radius was not reached get a
      // longer list (k_redo_size: doubled, or quadrupled once long)
      auto take = [&](const uint32_t *w, const uint32_t *L, const uint8_t *st, uint32_t cnt,
                      uint8_t want, uint32_t *ow, uint32_t *oL, const double *nr,
                      double *onr, const double *kd, double *okd) -> uint32_t {
        if (cnt == 0) return 0;
        Temps T2;
        uint32_t *fl = need(T2.get<uint32_t>(cnt));
        uint32_t *ps = need(T2.get<uint32_t>(cnt));
        hipLaunchKernelGGL(k_flag_state, dim3(nblk(cnt)), dim3(256), 0, s, st, cnt, want, fl);
        uint32_t c = scan_excl<uint32_t, uint32_t>(fl, ps, cnt, s);
        if (c && nr)
          hipLaunchKernelGGL(k_compact_redo, dim3(nblk(cnt)), dim3(256), 0, s, w, L, nr, kd, fl,
                             ps, cnt, ow, oL, onr, okd);
        else if (c)
          hipLaunchKernelGGL(k_compact_pairs, dim3(nblk(cnt)), dim3(256), 0, s, w, L, fl, ps,
                             cnt, 1u, kmax, ow, oL);
        HIPCHK(hipGetLastError());
        return c;
      };
      uint32_t *cw2 = lists[ping_m], *cl2 = lens[ping_m];
      uint32_t *cb2 = lists[ping_b], *cbl2 = lens[ping_b];
      uint32_t *cx2 = lists[ping_x], *cxl2 = lens[ping_x];
      double *rr_m = need(TT.get<double>(std::max<uint32_t>(nw, 1)));
      double *rr_b = need(TT.get<double>(std::max<uint32_t>(nw + nbig, 1)));
      double *rr_x = need(TT.get<double>(std::max<uint32_t>(nwide, 1)));
      double *kk_m = need(TT.get<double>(std::max<uint32_t>(nw, 1)));
      double *kk_b = need(TT.get<double>(std::max<uint32_t>(nw + nbig, 1)));
      double *kk_x = need(TT.get<double>(std::max<uint32_t>(nwide, 1)));
      const uint32_t nw_next =
          take(cw, cl, wstate, nw, CS_NEED_MORE, cw2, cl2, nr_main, rr_m, kd_main, kk_m);
      const uint32_t nb_next =
          take(cb, cbl, bstate, nbig, CS_NEED_MORE, cb2, cbl2, nr_big, rr_b, kd_big, kk_b);
      const uint32_t nx_next =
          take(cx, cxl, xstate, nwide, CS_NEED_MORE, cx2, cxl2, nr_wide, rr_x, kd_wide, kk_x);
      const uint32_t ncap = take(cw, cl, wstate, nw, CS_CAPACITY, cb2 + nb_next, cbl2 + nb_next,
                                 nullptr, nullptr, nullptr, nullptr);
      // cells over the large-cell kernel's capacity go to the wide kernel
      // (kept-cell stores hold 8-bit plane indices: not for them)
      const uint32_t nbcap = take(cb, cbl, bstate, nbig, CS_CAPACITY, cx2 + nx_next,
                                  cxl2 + nx_next, nullptr, nullptr, nullptr, nullptr);
      if (nbcap && T->keep) throw BuildError{MV_E_CAPACITY};
      // cells over the wide kernel's capacity run again with twice as much
      const uint32_t nxcap = take(cx, cxl, xstate, nwide, CS_CAPACITY, cx2 + nx_next + nbcap,
                                  cxl2 + nx_next + nbcap, nullptr, nullptr, nullptr, nullptr);
      if (nxcap) {
        if (wide_p >= WIDE_P_MAX) throw BuildError{MV_E_CAPACITY};
        wide_p *= 2;
        wide_v *= 2;
      }
      float *Rm_next = need(TT.get<float>(std::max<uint32_t>(nw_next, 1)));
      float *Rb_next = need(TT.get<float>(std::max<uint32_t>(nb_next + ncap, 1)));
      float *Rx_next = need(TT.get<float>(std::max<uint32_t>(nx_next + nbcap + nxcap, 1)));
      // 2x to 4x per pass (to 16x past 4,096 entries), from the safety
      // radius at exhaustion
      constexpr uint32_t GROW = 16;
      // lists of 1,024 entries and more: bounded by the seeds of the grid
      // cells within the safety radius
      auto ball_bound = [&](uint32_t cnt, const uint32_t *wl, const uint32_t *ll,
                            const double *rr, const double *kk) -> const uint32_t * {
        if (!cnt) return nullptr;
        uint32_t *ub = need(TT.get<uint32_t>(cnt));
        hipLaunchKernelGGL(k_redo_count, dim3((cnt + 3) / 4), dim3(256), 0, s, cnt, wl, ll, rr,
                           kk, spos, (const uint32_t *)cell_start, PR, REDO_COUNT_MIN, ub,
                           (const uint32_t *)nullptr);
        HIPCHK(hipGetLastError());
        return ub;
      };
      if (nw_next)
        hipLaunchKernelGGL(k_redo_size, dim3(nblk(nw_next)), dim3(256), 0, s, nw_next, cl2,
                           (const double *)rr_m, (const double *)kk_m, Rm_next, kmax, cap_err,
                           (const uint32_t *)nullptr,
                           ball_bound(nw_next, cw2, cl2, (const double *)rr_m, (const double *)kk_m));
      if (nb_next)
        hipLaunchKernelGGL(k_redo_size, dim3(nblk(nb_next)), dim3(256), 0, s, nb_next, cbl2,
                           (const double *)rr_b, (const double *)kk_b, Rb_next, kmax, cap_err,
                           (const uint32_t *)nullptr,
                           ball_bound(nb_next, cb2, cbl2, (const double *)rr_b, (const double *)kk_b));
      if (nx_next)
        hipLaunchKernelGGL(k_redo_size, dim3(nblk(nx_next)), dim3(256), 0, s, nx_next, cxl2,
                           (const double *)rr_x, (const double *)kk_x, Rx_next, kmax, cap_err,
                           (const uint32_t *)nullptr,
                           ball_bound(nx_next, cx2, cxl2, (const double *)rr_x, (const double *)kk_x));
      if (ncap)  // moved cells keep their list; their gather starts from the density
        hipLaunchKernelGGL(k_fill_u32, dim3(nblk(ncap)), dim3(256), 0, s,
                           (uint32_t *)(Rb_next + nb_next), ncap, 0u);
      if (nbcap + nxcap)
        hipLaunchKernelGGL(k_fill_u32, dim3(nblk(nbcap + nxcap)), dim3(256), 0, s,
                           (uint32_t *)(Rx_next + nx_next), nbcap + nxcap, 0u);
      HIPCHK(hipGetLastError());
      bigcells += ncap;
      widecells += nbcap;
      redone += nw_next + nb_next + nx_next;
      if (nw && !exact_order) cells_cw += nw - nw_next - ncap;
      // the main lists longer than the wavefront list kernels take go to
      // its end (gathered by the long-list kernel); not in a build that
      // plans a captured one, whose passes gather a main list with one
      // kernel (the two kernels accept different balls, and so leave
      // different cells for the next pass)
      nw_tail = 0;
      if (nw_next && (uint64_t)lb_main * GROW > MAIN_LIST_MAX && !exact_order && !t_trace) {
        Temps T3;
        uint32_t *fl = need(T3.get<uint32_t>(nw_next));
        uint32_t *ps = need(T3.get<uint32_t>(nw_next));
        hipLaunchKernelGGL(k_flag_long, dim3(nblk(nw_next)), dim3(256), 0, s,
                           (const uint32_t *)cl2, nw_next, MAIN_LIST_MAX, fl);
        const uint32_t nl = scan_excl<uint32_t, uint32_t>(fl, ps, nw_next, s);
        if (nl && nl < nw_next) {
          uint32_t *sw = need(T3.get<uint32_t>(nw_next));
          uint32_t *sL = need(T3.get<uint32_t>(nw_next));
          float *sR = need(T3.get<float>(nw_next));
          const uint32_t o = nw_next - nl;
          hipLaunchKernelGGL(k_move_long, dim3(nblk(nw_next)), dim3(256), 0, s,
                             (const uint32_t *)cw2, (const uint32_t *)cl2, (const float *)Rm_next,
                             (const uint32_t *)fl, (const uint32_t *)ps, nw_next, sw + o, sL + o,
                             sR + o, sw, sL, sR);
          HIPCHK(hipGetLastError());
          HIPCHK(hipMemcpyAsync(cw2, sw, nw_next * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
          HIPCHK(hipMemcpyAsync(cl2, sL, nw_next * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
          HIPCHK(hipMemcpyAsync(Rm_next, sR, nw_next * sizeof(float), hipMemcpyDeviceToDevice,
                                s));
          // (the temporaries are freed when T3 ends: the copies finish first)
          HIPCHK(hipStreamSynchronize(s));
          nw_tail = nl;
        }
      }
      const uint32_t lb_main_old = lb_main;
      // bounds on the new lengths (k_redo_size grows them at most 4x, 16x
      // from 4,096 entries)
      auto grown = [&](uint32_t lb) {
        return (uint32_t)std::min<uint64_t>((uint64_t)GROW * lb, kmax);
      };
      const uint32_t lb_big_old = lb_big;
      lb_main = grown(lb_main);
      lb_big = std::max(grown(lb_big), std::min(lb_main_old, kmax));
      lb_wide = std::max(grown(lb_wide), std::min(lb_big_old, kmax));
      Rmain = Rm_next;
      Rbig = Rb_next;
      Rwide = Rx_next;
      nw = nw_next;
      nbig = nb_next + ncap;
      nwide = nx_next + nbcap + nxcap;
      cw = cw2;
      cl = cl2;
      cb = cb2;
      cbl = cbl2;
      cx = cx2;
      cxl = cxl2;
      ping_m = 3 - ping_m;
      ping_b = 7 - ping_b;
      ping_x = 11 - ping_x;
    }
    if (!overflowed) {
      unsigned long long cnts[NCOUNTERS];
      HIPCHK(hipMemcpy(cnts, counters, sizeof cnts, hipMemcpyDeviceToHost));
      T->stats.knn_gathers = T->stats.knn_reads = 0;
      for (int k = 7; k < 71; k++) T->stats.knn_gathers += cnts[k];
      for (int k = 71; k < NCOUNTERS; k++) T->stats.knn_reads += cnts[k];
      T->stats.exact_tests = cnts[2];
      T->stats.knn_big_items = cnts[4];
      T->stats.knn_big_gathers = cnts[5];
      T->stats.knn_big_entries = cnts[6];
      if (T->keep) {
        unsigned long long used[2];
        HIPCHK(hipMemcpy(used, cs_cur, sizeof used, hipMemcpyDeviceToHost));
        T->used_pl = used[0];
        T->used_vx = used[1];
      }
      T->stats.k_last = lb_main;
      break;
    }
    if (attempt > 4) throw BuildError{MV_E_CAPACITY};
    // grow what ran out (bit 0: face staging, bit 1: kept-cell store)
    if (ovf_bits & 1) st_cap *= 2;
    if (ovf_bits & 2) {
      cap_pl *= 2;
      cap_vx *= 2;
      alloc_store();
    }
    HIPCHK(hipMemsetAsync(cell_nf, 0, n * sizeof(uint32_t), s));
    HIPCHK(hipMemsetAsync(link_in, 0, n * sizeof(uint32_t), s));
  }
  if (t_trace) {
    t_trace->lbound0 = lbound;
    t_trace->st_cap = st_cap;
  }
  T->stats.passes = passes;
  T->stats.cells_redone = redone;
  T->stats.big_cells = bigcells;
  T->stats.wide_cells = widecells;
  T->stats.pair_handovers = handovers;
  T->stats.dense_cells = D.dense_cells;
  T->stats.ms_knn = ms_knn + ms_model;
  T->stats.ms_clip = ms_clip;
  T->stats.clip_wave_launches = n_cw;
  T->stats.knn_launches = n_knn;
  T->stats.clip_wave_cells = cells_cw;
  T->stats.knn_queries = q_knn;
  T->stats.knn_entries = e_knn;
  T->stats.ms_clip_wave = ms_cw;
  T->stats.ms_knn_kernels = ms_knn_k;
  T->stats.ms_model = ms_model;

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
  hipLaunchKernelGGL(k_link_count, dim3(nb), dim3(256), 0, s, cell_nf, link_in, n, T->d_cell_fcnt);
  uint64_t NL = scan_excl<uint64_t, uint64_t>(T->d_cell_fcnt, T->d_cell_foff, n, s);
  T->nl = NL;
  if (t_trace) {
    t_trace->F = F;
    t_trace->NL = NL;
  }
  T->d_links = T->own<uint64_t>(NL ? NL : 1);
  if (F) {
    FinalArgs FA;
    FA.n_staged_dev = nullptr;
    FA.st_cell = st_cell;
    FA.st_right = st_right;
    FA.st_img = st_img;
    FA.st_lslot = st_lslot;
    FA.st_area = st_area;
    FA.st_cen = st_cen;
    FA.st_norm = st_norm;
    FA.n_staged = F;
    FA.cell_fbase = cell_fbase;
    FA.face_off = face_off;
    FA.boxes = D.d_boxes;
    FA.prob = D.prob;
    FA.face_left = T->d_face_left;
    FA.face_right = T->d_face_right;
    FA.face_area = T->d_face_area;
    FA.face_centroid = T->d_face_cen;
    FA.face_normal = T->d_face_norm;
    FA.face_shift = T->d_face_shift;
    FA.face_has_shift = T->d_face_has_shift;
    FA.link_in = link_in;
    FA.link_off = T->d_cell_foff;
    FA.links = T->d_links;
    hipLaunchKernelGGL(k_place_faces, dim3(nblk(F)), dim3(256), 0, s, FA);
    hipLaunchKernelGGL(k_sort_links, dim3(nb), dim3(256), 0, s, T->d_cell_foff, link_in, n,
                       T->d_links, (const unsigned long long *)nullptr);
    HIPCHK(hipGetLastError());
  }
  hipLaunchKernelGGL(k_cell_records, dim3(nb), dim3(256), 0, s, (const double4 *)D.gpos, n,
                     d_mask, (const uint64_t *)D.d_goff, (const uint32_t *)D.prob,
                     T->d_cell_loc, T->d_cell_idx);
  HIPCHK(hipGetLastError());
  // many problems: per-problem face, link and generator numbering
  std::vector<uint64_t> foff(P + 1, 0), loff(P + 1, 0);
  if (P > 1) {
    uint64_t *d_foff = need(TT.get<uint64_t>(P + 1));
    uint64_t *d_loff = need(TT.get<uint64_t>(P + 1));
    hipLaunchKernelGGL(k_problem_bases, dim3((P + 1 + 63) / 64), dim3(64), 0, s,
                       (const uint64_t *)D.d_goff, P, (const uint64_t *)face_off, F,
                       (const uint64_t *)T->d_cell_foff, NL, n, d_foff, d_loff,
                       (const unsigned long long *)nullptr, (const unsigned long long *)nullptr);
    HIPCHK(hipMemcpyAsync(foff.data(), d_foff, (P + 1) * sizeof(uint64_t),
                          hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(loff.data(), d_loff, (P + 1) * sizeof(uint64_t),
                          hipMemcpyDeviceToHost, s));
    if (F)
      hipLaunchKernelGGL(k_localize_faces, dim3(nblk(F)), dim3(256), 0, s, T->d_face_left,
                         T->d_face_right, F, (const uint64_t *)D.d_goff,
                         (const uint32_t *)D.prob, (const unsigned long long *)nullptr);
    hipLaunchKernelGGL(k_localize_links, dim3(nb), dim3(256), 0, s, T->d_cell_foff,
                       (const uint64_t *)T->d_cell_fcnt, n, (const uint32_t *)D.prob,
                       (const uint64_t *)d_foff, (const uint64_t *)d_loff, T->d_links,
                       (const unsigned long long *)nullptr);
    HIPCHK(hipGetLastError());
  } else {
    foff[1] = F;
    loff[1] = NL;
  }
  T->stats.ms_finalize = tf.stop();
  HIPCHK(hipStreamSynchronize(s));
  auto t_end = std::chrono::steady_clock::now();
  T->stats.ms_total = std::chrono::duration<double, std::milli>(t_end - t_start).count();
  // the handles: problem p sees its slices
  double *cl = T->d_cell_loc, *cc = T->d_cell_cen, *cv = T->d_cell_vol, *csr = T->d_cell_sr;
  uint64_t *cfo = T->d_cell_foff, *cfc = T->d_cell_fcnt, *cid = T->d_cell_idx;
  uint32_t *npp = T->d_npop;
  uint64_t *fl = T->d_face_left;
  int64_t *fr = T->d_face_right;
  double *fa = T->d_face_area, *fc = T->d_face_cen, *fn = T->d_face_norm, *fs = T->d_face_shift;
  uint8_t *fh = T->d_face_has_shift;
  uint64_t *lk = T->d_links;
  const CellStore cs = T->cs;
  uint8_t *km = T->d_mask;
  for (uint32_t p = 0; p < P; p++) {
    mv_tess *U = Ts[p];
    const uint64_t g0 = probs[p].g0;
    U->dim = dim;
    U->periodic = probs[p].periodic;
    U->n = probs[p].g1 - g0;
    U->nf = foff[p + 1] - foff[p];
    U->nl = loff[p + 1] - loff[p];
    U->d_cell_loc = cl + 3 * g0;
    U->d_cell_cen = cc + 3 * g0;
    U->d_cell_vol = cv + g0;
    U->d_cell_sr = csr + g0;
    U->d_cell_foff = cfo + g0;
    U->d_cell_fcnt = cfc + g0;
    U->d_cell_idx = cid + g0;
    U->d_npop = npp + g0;
    U->d_face_left = fl + foff[p];
    U->d_face_right = fr + foff[p];
    U->d_face_area = fa + foff[p];
    U->d_face_cen = fc + 3 * foff[p];
    U->d_face_norm = fn + 3 * foff[p];
    U->d_face_shift = fs + 3 * foff[p];
    U->d_face_has_shift = fh + foff[p];
    U->d_links = lk + loff[p];
    U->stats = T->stats;
    U->keep = T->keep;
    if (T->keep) {
      U->cs = cs;
      U->cs.pbase = cs.pbase + g0;
      U->cs.vbase = cs.vbase + g0;
      U->cs.np = cs.np + g0;
      U->cs.nv = cs.nv + g0;
      U->used_pl = T->used_pl;
      U->used_vx = T->used_vx;
      U->idx_base = g0;
      U->d_mask = km ? km + g0 : nullptr;
    }
  }
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
  if (o.mode != MV_MODE_PLAIN && o.mode != MV_MODE_LEARNED) return MV_E_ARG;
  if (o.mode == MV_MODE_LEARNED &&
      (!o.model || o.model->magic != MV_MODEL_MAGIC || o.model->version != 1))
    return MV_E_ARG;
  mv_tess *T = new (std::nothrow) mv_tess();
  if (!T) return MV_E_OOM;
  try {
    std::vector<ProblemSpec> probs(1);
    normalise_box(anchor, width, dim, probs[0].anchor, probs[0].width);
    probs[0].periodic = periodic ? 1 : 0;
    probs[0].g0 = 0;
    probs[0].g1 = n;
    build_multi({T}, d_xyz, n, d_mask, probs, dim, o);
  } catch (...) {
    const mv_status st = status_of_exception();
    (void)hipDeviceSynchronize();
    (void)hipGetLastError();
    delete T;
    return st;
  }
  *out = T;
  return MV_OK;
}

// ---------------------------------------------------------------------
// learned proposal: host side of the training and inference kernels
// ---------------------------------------------------------------------
// FP32 parameter block of the trainer: w1 16x64 | b1 | w2 64x64 | b2 |
// w3 64x16 (2 used) | b3 16 (2 used)
constexpr int P_W1 = 0, P_B1 = P_W1 + NF * NH, P_W2 = P_B1 + NH, P_B2 = P_W2 + NH * NH,
              P_W3 = P_B2 + NH, P_B3 = P_W3 + NH * NOP, P_TOTAL = P_B3 + NOP;

uint16_t host_f2bf(float x) {
  uint32_t b;
  memcpy(&b, &x, 4);
  if ((b & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
  b += 0x7fffu + ((b >> 16) & 1u);
  return (uint16_t)(b >> 16);
}
float host_bf2f(uint16_t h) {
  uint32_t b = (uint32_t)h << 16;
  float x;
  memcpy(&x, &b, 4);
  return x;
}

void model_to_params(const mv_model &m, std::vector<float> &P) {
  P.assign(P_TOTAL, 0.f);
  for (int i = 0; i < NF * NH; i++) P[P_W1 + i] = host_bf2f(m.w1[i]);
  for (int i = 0; i < NH * NH; i++) P[P_W2 + i] = host_bf2f(m.w2[i]);
  for (int i = 0; i < NH; i++)
    for (int o = 0; o < NO; o++) P[P_W3 + i * NOP + o] = host_bf2f(m.w3[i * NO + o]);
  for (int i = 0; i < NH; i++) {
    P[P_B1 + i] = m.b1[i];
    P[P_B2 + i] = m.b2[i];
  }
  for (int o = 0; o < NO; o++) P[P_B3 + o] = m.b3[o];
}

void params_to_model(const std::vector<float> &P, mv_model &m) {
  for (int i = 0; i < NF * NH; i++) m.w1[i] = host_f2bf(P[P_W1 + i]);
  for (int i = 0; i < NH * NH; i++) m.w2[i] = host_f2bf(P[P_W2 + i]);
  for (int i = 0; i < NH; i++)
    for (int o = 0; o < NO; o++) m.w3[i * NO + o] = host_f2bf(P[P_W3 + i * NOP + o]);
  for (int i = 0; i < NH; i++) {
    m.b1[i] = P[P_B1 + i];
    m.b2[i] = P[P_B2 + i];
  }
  for (int o = 0; o < NO; o++) m.b3[o] = P[P_B3 + o];
}

// Device buffers of one forward + backward pass over `B` rows.
struct TrainBufs {
  int B;
  float *X, *Tg, *Z1, *A1, *Z2, *A2, *Y, *dY, *dA2, *dA1, *lrow;
};

TrainBufs train_bufs(int B, Temps &TT) {
  TrainBufs t;
  t.B = B;
  t.X = need(TT.get<float>((size_t)B * NF));
  t.Tg = need(TT.get<float>((size_t)B * NO));
  t.Z1 = need(TT.get<float>((size_t)B * NH));
  t.A1 = need(TT.get<float>((size_t)B * NH));
  t.Z2 = need(TT.get<float>((size_t)B * NH));
  t.A2 = need(TT.get<float>((size_t)B * NH));
  t.Y = need(TT.get<float>((size_t)B * NOP));
  t.dY = need(TT.get<float>((size_t)B * NOP));
  t.dA2 = need(TT.get<float>((size_t)B * NH));
  t.dA1 = need(TT.get<float>((size_t)B * NH));
  t.lrow = need(TT.get<float>((size_t)B));
  return t;
}

void gemm(const float *A, long sa0, long sa1, const float *Bm, long sb0, long sb1, float *C,
          int ldc, int M, int N, int K, hipStream_t s) {
  const int tiles = (M / 16) * (N / 16);
  hipLaunchKernelGGL(k_gemm_bf16, dim3((tiles + 3) / 4), dim3(256), 0, s, A, sa0, sa1, Bm, sb0,
                     sb1, C, ldc, M, N, K, 0);
}

// forward and backward of the pinball loss over the rows in t.X / t.Tg;
// P: parameters, G: gradient (same layout)
void forward_backward(const TrainBufs &t, const float *P, float *G, float tau, hipStream_t s) {
  const int B = t.B;
  auto g1 = [](long n) { return dim3((unsigned)((n + 255) / 256)); };
  gemm(t.X, NF, 1, P + P_W1, NH, 1, t.Z1, NH, B, NH, NF, s);
  hipLaunchKernelGGL(k_bias_relu, g1((long)B * NH), dim3(256), 0, s, t.Z1, t.A1, P + P_B1, B,
                     NH, 1);
  gemm(t.A1, NH, 1, P + P_W2, NH, 1, t.Z2, NH, B, NH, NH, s);
  hipLaunchKernelGGL(k_bias_relu, g1((long)B * NH), dim3(256), 0, s, t.Z2, t.A2, P + P_B2, B,
                     NH, 1);
  gemm(t.A2, NH, 1, P + P_W3, NOP, 1, t.Y, NOP, B, NOP, NH, s);
  hipLaunchKernelGGL(k_bias_relu, g1((long)B * NOP), dim3(256), 0, s, t.Y, (float *)nullptr,
                     P + P_B3, B, NOP, 0);
  hipLaunchKernelGGL(k_pinball, g1(B), dim3(256), 0, s, t.Y, t.Tg, B, tau, t.dY, t.lrow);
  // layer 3
  gemm(t.A2, 1, NH, t.dY, NOP, 1, G + P_W3, NOP, NH, NOP, B, s);
  hipLaunchKernelGGL(k_colsum, dim3(1), dim3(64), 0, s, t.dY, B, NOP, G + P_B3);
  gemm(t.dY, NOP, 1, P + P_W3, 1, NOP, t.dA2, NH, B, NH, NOP, s);
  hipLaunchKernelGGL(k_relu_grad, g1((long)B * NH), dim3(256), 0, s, t.Z2, t.dA2, B * NH);
  // layer 2
  gemm(t.A1, 1, NH, t.dA2, NH, 1, G + P_W2, NH, NH, NH, B, s);
  hipLaunchKernelGGL(k_colsum, dim3(1), dim3(64), 0, s, t.dA2, B, NH, G + P_B2);
  gemm(t.dA2, NH, 1, P + P_W2, 1, NH, t.dA1, NH, B, NH, NH, s);
  hipLaunchKernelGGL(k_relu_grad, g1((long)B * NH), dim3(256), 0, s, t.Z1, t.dA1, B * NH);
  // layer 1
  gemm(t.X, 1, NF, t.dA1, NH, 1, G + P_W1, NH, NF, NH, B, s);
  hipLaunchKernelGGL(k_colsum, dim3(1), dim3(64), 0, s, t.dA1, B, NH, G + P_B1);
  HIPCHK(hipGetLastError());
}

uint64_t splitmix(uint64_t &st) {
  uint64_t z = (st += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
double unif(uint64_t &st) { return ((splitmix(st) >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
double gauss(uint64_t &st) {
  const double u = unif(st), v = unif(st);
  return std::sqrt(-2.0 * std::log(u)) * std::cos(6.283185307179586 * v);
}

mv_status device_ready() {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    (void)hipGetLastError();
    return MV_E_DEVICE;
  }
  return MV_OK;
}

template <typename F>
mv_status guarded(F &&f) {
  try {
    f();
  } catch (...) {
    const mv_status st = status_of_exception();
    (void)hipDeviceSynchronize();
    (void)hipGetLastError();
    return st;
  }
  return MV_OK;
}

// many problems in one build
mv_status run_build_batch(const double *d_xyz, const uint64_t *offsets, uint32_t P,
                          const uint8_t *d_mask, const double *anchors, const double *widths,
                          const uint32_t *periodic, uint32_t dim, const mv_options *opt,
                          mv_tess **out) {
  mv_options o;
  mv_default_options(&o);
  if (opt) o = *opt;
  if (o.mode != MV_MODE_PLAIN && o.mode != MV_MODE_LEARNED) return MV_E_ARG;
  if (o.mode == MV_MODE_LEARNED &&
      (!o.model || o.model->magic != MV_MODEL_MAGIC || o.model->version != 1))
    return MV_E_ARG;
  std::vector<mv_tess *> Ts(P, nullptr);
  for (uint32_t p = 0; p < P; p++) {
    Ts[p] = new (std::nothrow) mv_tess();
    if (!Ts[p]) {
      for (mv_tess *t : Ts) delete t;
      return MV_E_OOM;
    }
  }
  try {
    std::vector<ProblemSpec> probs(P);
    for (uint32_t p = 0; p < P; p++) {
      normalise_box(anchors + 3 * p, widths + 3 * p, dim, probs[p].anchor, probs[p].width);
      probs[p].periodic = periodic && periodic[p] ? 1u : 0u;
      probs[p].g0 = offsets[p];
      probs[p].g1 = offsets[p + 1];
    }
    build_multi(Ts, d_xyz, offsets[P], d_mask, probs, dim, o);
  } catch (...) {
    const mv_status st = status_of_exception();
    (void)hipDeviceSynchronize();
    (void)hipGetLastError();
    for (mv_tess *t : Ts) delete t;
    return st;
  }
  for (uint32_t p = 0; p < P; p++) out[p] = Ts[p];
  return MV_OK;
}

bool valid_batch(const uint64_t *offsets, uint32_t P, uint32_t dim, const double *anchors,
                 const double *widths) {
  if (!offsets || P == 0 || P > (1u << 24) || offsets[0] != 0) return false;
  if (!anchors || !widths || dim < 1 || dim > 3) return false;
  for (uint32_t p = 0; p < P; p++)
    if (offsets[p + 1] < offsets[p]) return false;
  return offsets[P] < (1ull << 27);
}

// ---------------------------------------------------------------------
// kept cells: device view and the active-cell list
// ---------------------------------------------------------------------
CellGeom geom_of(const mv_tess *t) {
  CellGeom g;
  g.pl_a = t->cs.pl_a;
  g.pl_b = t->cs.pl_b;
  g.vx = t->cs.vx;
  g.pbase = t->cs.pbase;
  g.vbase = t->cs.vbase;
  g.np = t->cs.np;
  g.nv = t->cs.nv;
  g.cell_loc = t->d_cell_loc;
  for (int c = 0; c < 3; c++) g.width[c] = t->width[c];
  g.dim = (int)t->dim;
  g.idx_base = t->idx_base;
  const bool wf = t->wf_on && t->wf_ready;
  g.fbase = wf ? t->d_wf_fbase : nullptr;
  g.nf = wf ? t->d_wf_nf : nullptr;
  g.face_plane = wf ? t->d_wf_plane : nullptr;
  g.face_nvert = wf ? t->d_wf_nvert : nullptr;
  g.face_voff = wf ? t->d_wf_voff : nullptr;
  g.face_vtx = wf ? t->d_wf_vtx : nullptr;
  return g;
}

__global__ void k_flag_built(const uint32_t *__restrict__ np, size_t n,
                             uint32_t *__restrict__ flag) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flag[i] = np[i] ? 1u : 0u;
}
__global__ void k_compact_index(const uint32_t *__restrict__ flag,
                                const uint64_t *__restrict__ pos, size_t n,
                                uint64_t *__restrict__ out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && flag[i]) out[pos[i]] = i;
}
// tets per active cell: six per vertex, or with faces m - 2 per face of m
// vertices, i.e. 3 nv - 2 nf
__global__ void k_gather_nv(const uint64_t *__restrict__ cells, uint64_t n,
                            const uint32_t *__restrict__ nv, const uint32_t *__restrict__ nf,
                            uint64_t *__restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t c = cells[i];
  out[i] = nf ? 3ull * nv[c] - 2ull * nf[c] : 6ull * nv[c];
}
struct TetOut {
  double *v;
  int32_t *plane;
  uint64_t k;
  __device__ void operator()(d3 v0, d3 v1, d3 v2, int p) {
    double *o = v + 9 * k;
    o[0] = v0.x; o[1] = v0.y; o[2] = v0.z;
    o[3] = v1.x; o[4] = v1.y; o[5] = v1.z;
    o[6] = v2.x; o[7] = v2.y; o[8] = v2.z;
    plane[k] = p;
    k++;
  }
};
__global__ void __launch_bounds__(256)
    k_export_tets(CellGeom G, const uint64_t *__restrict__ cells, uint64_t n,
                  const uint64_t *__restrict__ off, double *__restrict__ v,
                  int32_t *__restrict__ plane, unsigned int *err) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  TetOut f{v, plane, off[i]};
  if (!for_each_tet(G, cells[i], f)) atomicOr(err, 1u);
}

// active cells (those the build kept) in generator order
void ensure_active(mv_tess *t) {
  if (t->d_active || t->n == 0) return;
  hipStream_t s = 0;
  Temps TT;
  const size_t n = t->n;
  uint32_t *flag = need(TT.get<uint32_t>(n));
  uint64_t *pos = need(TT.get<uint64_t>(n));
  hipLaunchKernelGGL(k_flag_built, dim3(nblk(n)), dim3(256), 0, s, t->cs.np, n, flag);
  const uint64_t na = scan_excl<uint32_t, uint64_t>(flag, pos, n, s);
  t->d_active = t->own<uint64_t>(na ? na : 1);
  if (na)
    hipLaunchKernelGGL(k_compact_index, dim3(nblk(n)), dim3(256), 0, s, flag, pos, n,
                       t->d_active);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s));
  t->n_active = na;
}

template <class I>
void run_cell_integrals(mv_tess *t, double *out, int width) {
  hipStream_t s = 0;
  Temps TT;
  ensure_active(t);
  const uint64_t na = t->n_active;
  if (!na) return;
  I *d = need(TT.get<I>(na));
  unsigned int *err = need(TT.get<unsigned int>(1));
  HIPCHK(hipMemsetAsync(err, 0, sizeof(unsigned int), s));
  hipLaunchKernelGGL((cell_integrals<I>), dim3(nblk(na)), dim3(256), 0, s, geom_of(t),
                     (const uint64_t *)t->d_active, na, (const void *)nullptr, d, err);
  HIPCHK(hipGetLastError());
  std::vector<I> h(na);
  unsigned int herr = 0;
  HIPCHK(hipMemcpyAsync(h.data(), d, na * sizeof(I), hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(&herr, err, sizeof herr, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (herr) throw BuildError{MV_E_PLANE_DET_ZERO};
  static_assert(sizeof(I) % sizeof(double) == 0, "integral of doubles");
  for (uint64_t i = 0; i < na; i++)
    memcpy(out + (size_t)i * width, &h[i], width * sizeof(double));
}

template <class I>
void run_face_integrals(mv_tess *t, int symmetric, size_t *count, uint64_t *left,
                        int64_t *right, uint8_t *has_shift, double *shift, int32_t *plane,
                        double *values, int width) {
  hipStream_t s = 0;
  Temps TT;
  ensure_active(t);
  const uint64_t na = t->n_active;
  if (!na) {
    *count = 0;
    return;
  }
  int32_t *face_local = need(TT.get<int32_t>(t->used_pl ? t->used_pl : 1));
  uint64_t *cnt = need(TT.get<uint64_t>(na));
  uint64_t *off = need(TT.get<uint64_t>(na));
  const CellGeom G = geom_of(t);
  hipLaunchKernelGGL(face_counts, dim3(nblk(na)), dim3(256), 0, s, G,
                     (const uint64_t *)t->d_active, na, symmetric,
                     (const uint8_t *)t->d_mask, face_local, cnt);
  HIPCHK(hipGetLastError());
  const uint64_t F = scan_excl<uint64_t, uint64_t>(cnt, off, na, s);
  if (!left) {
    *count = F;
    return;
  }
  if (*count < F) throw BuildError{MV_E_ARG};
  *count = F;
  if (!F) return;
  FaceIntegrator<I> *d = need(TT.get<FaceIntegrator<I>>(F));
  unsigned int *err = need(TT.get<unsigned int>(1));
  HIPCHK(hipMemsetAsync(err, 0, sizeof(unsigned int), s));
  hipLaunchKernelGGL((face_integrals<I>), dim3(nblk(na)), dim3(256), 0, s, G,
                     (const uint64_t *)t->d_active, na, (const int32_t *)face_local,
                     (const uint64_t *)off, (const void *)nullptr, d, err);
  HIPCHK(hipGetLastError());
  std::vector<FaceIntegrator<I>> h(F);
  unsigned int herr = 0;
  HIPCHK(hipMemcpyAsync(h.data(), d, F * sizeof(FaceIntegrator<I>), hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(&herr, err, sizeof herr, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (herr) throw BuildError{MV_E_PLANE_DET_ZERO};
  for (uint64_t f = 0; f < F; f++) {
    left[f] = h[f].left;
    if (right) right[f] = h[f].right;
    if (has_shift) has_shift[f] = (uint8_t)h[f].has_shift;
    if (shift)
      for (int c = 0; c < 3; c++) shift[3 * f + c] = h[f].shift[c];
    if (plane) plane[f] = h[f].plane;
    if (values) memcpy(values + (size_t)f * width, &h[f].integral, width * sizeof(double));
  }
}

// ConvexCell::with_faces for every kept cell (mv_faces.hpp); the lists are
// built once per handle and kept until mv_free.
void run_with_faces(mv_tess *t) {
  if (t->dim != 3) throw BuildError{MV_E_NOT_3D};
  if (t->wf_ready) {
    t->wf_on = true;
    return;
  }
  hipStream_t s = 0;
  Temps TT;
  ensure_active(t);
  const uint64_t na = t->n_active;
  const size_t n = t->n ? t->n : 1;
  t->d_wf_fbase = t->own<uint64_t>(n);
  t->d_wf_nf = t->own<uint32_t>(n);
  t->d_wf_cell_foff = t->own<uint64_t>(na + 1);
  HIPCHK(hipMemsetAsync(t->d_wf_fbase, 0, n * sizeof(uint64_t), s));
  HIPCHK(hipMemsetAsync(t->d_wf_nf, 0, n * sizeof(uint32_t), s));
  HIPCHK(hipMemsetAsync(t->d_wf_cell_foff, 0, (na + 1) * sizeof(uint64_t), s));
  uint64_t NF = 0, NE = 0;
  if (na) {
    CellGeom G = geom_of(t);
    uint32_t *pcnt = need(TT.get<uint32_t>(t->used_pl ? t->used_pl : 1));
    uint64_t *nfc = need(TT.get<uint64_t>(na));
    uint64_t *nec = need(TT.get<uint64_t>(na));
    uint64_t *eoff = need(TT.get<uint64_t>(na));
    hipLaunchKernelGGL(k_wf_count, dim3(nblk(na)), dim3(256), 0, s, G,
                       (const uint64_t *)t->d_active, na, pcnt, nfc, nec);
    HIPCHK(hipGetLastError());
    NF = scan_excl<uint64_t, uint64_t>(nfc, t->d_wf_cell_foff, na, s);
    NE = scan_excl<uint64_t, uint64_t>(nec, eoff, na, s);
    HIPCHK(hipMemcpyAsync(t->d_wf_cell_foff + na, &NF, sizeof NF, hipMemcpyHostToDevice, s));
    t->d_wf_plane = t->own<int32_t>(NF ? NF : 1);
    t->d_wf_nvert = t->own<uint32_t>(NF ? NF : 1);
    t->d_wf_voff = t->own<uint64_t>(NF ? NF : 1);
    t->d_wf_vtx = t->own<uint32_t>(NE ? NE : 1);
    unsigned int *err = need(TT.get<unsigned int>(1));
    HIPCHK(hipMemsetAsync(err, 0, sizeof(unsigned int), s));
    hipLaunchKernelGGL(k_wf_fill, dim3(nblk(na)), dim3(256), 0, s, G,
                       (const uint64_t *)t->d_active, na, (const uint64_t *)t->d_wf_cell_foff,
                       (const uint64_t *)eoff, pcnt, t->d_wf_fbase, t->d_wf_nf, t->d_wf_plane,
                       t->d_wf_nvert, t->d_wf_voff, t->d_wf_vtx, err);
    HIPCHK(hipGetLastError());
    unsigned int herr = 0;
    HIPCHK(hipMemcpyAsync(&herr, err, sizeof herr, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (herr) throw BuildError{MV_E_FACE_WALK};
  } else {
    t->d_wf_plane = t->own<int32_t>(1);
    t->d_wf_nvert = t->own<uint32_t>(1);
    t->d_wf_voff = t->own<uint64_t>(1);
    t->d_wf_vtx = t->own<uint32_t>(1);
    HIPCHK(hipStreamSynchronize(s));
  }
  t->wf_nfaces = NF;
  t->wf_nentries = NE;
  t->wf_ready = true;
  t->wf_on = true;
}

// From<&VoronoiIntegrator<M>> for Voronoi: the build's connectivity with
// cell and face integrals of the kept cells' current decomposition.
mv_tess *voronoi_from_kept(mv_tess *t) {
  hipStream_t s = 0;
  Temps TT;
  ensure_active(t);
  std::unique_ptr<mv_tess> U(new mv_tess);
  U->parent = t->store;
  U->n = t->n;
  U->nf = t->nf;
  U->nl = t->nl;
  for (int c = 0; c < 3; c++) {
    U->anchor[c] = t->anchor[c];
    U->width[c] = t->width[c];
  }
  U->dim = t->dim;
  U->periodic = t->periodic;
  U->stats = t->stats;
  U->d_cell_loc = t->d_cell_loc;
  U->d_cell_sr = t->d_cell_sr;
  U->d_cell_foff = t->d_cell_foff;
  U->d_cell_fcnt = t->d_cell_fcnt;
  U->d_cell_idx = t->d_cell_idx;
  U->d_npop = t->d_npop;
  U->d_face_left = t->d_face_left;
  U->d_face_right = t->d_face_right;
  U->d_face_norm = t->d_face_norm;
  U->d_face_shift = t->d_face_shift;
  U->d_face_has_shift = t->d_face_has_shift;
  U->d_links = t->d_links;
  const size_t n = t->n ? t->n : 1, F = t->nf ? t->nf : 1;
  U->d_cell_vol = U->own<double>(n);
  U->d_cell_cen = U->own<double>(3 * n);
  U->d_face_area = U->own<double>(F);
  U->d_face_cen = U->own<double>(3 * F);
  // VoronoiCell::default for the generators without a cell
  HIPCHK(hipMemsetAsync(U->d_cell_vol, 0, n * sizeof(double), s));
  HIPCHK(hipMemsetAsync(U->d_cell_cen, 0, 3 * n * sizeof(double), s));
  const uint64_t na = t->n_active;
  if (na) {
    const CellGeom G = geom_of(t);
    unsigned int *err = need(TT.get<unsigned int>(1));
    HIPCHK(hipMemsetAsync(err, 0, sizeof(unsigned int), s));
    VolumeCentroidIntegral *ci = need(TT.get<VolumeCentroidIntegral>(na));
    hipLaunchKernelGGL((cell_integrals<VolumeCentroidIntegral>), dim3(nblk(na)), dim3(256), 0,
                       s, G, (const uint64_t *)t->d_active, na, (const void *)nullptr, ci, err);
    hipLaunchKernelGGL(k_scatter_cell_integrals, dim3(nblk(na)), dim3(256), 0, s,
                       (const uint64_t *)t->d_active, na, (const VolumeCentroidIntegral *)ci,
                       U->d_cell_vol, U->d_cell_cen);
    // the faces VoronoiCell::from_convex_cell creates (voronoi_cell.rs:53-73)
    // are those of compute_face_integrals_sym, in the same order
    int32_t *face_local = need(TT.get<int32_t>(t->used_pl ? t->used_pl : 1));
    uint64_t *cnt = need(TT.get<uint64_t>(na));
    uint64_t *off = need(TT.get<uint64_t>(na));
    hipLaunchKernelGGL(face_counts, dim3(nblk(na)), dim3(256), 0, s, G,
                       (const uint64_t *)t->d_active, na, 1, (const uint8_t *)t->d_mask,
                       face_local, cnt);
    HIPCHK(hipGetLastError());
    const uint64_t NF = scan_excl<uint64_t, uint64_t>(cnt, off, na, s);
    if (NF != t->nf) throw BuildError{MV_E_DEVICE};
    if (NF) {
      FaceIntegrator<AreaCentroidIntegral> *fi =
          need(TT.get<FaceIntegrator<AreaCentroidIntegral>>(NF));
      hipLaunchKernelGGL((face_integrals<AreaCentroidIntegral>), dim3(nblk(na)), dim3(256), 0,
                         s, G, (const uint64_t *)t->d_active, na, (const int32_t *)face_local,
                         (const uint64_t *)off, (const void *)nullptr, fi, err);
      hipLaunchKernelGGL(k_scatter_face_integrals, dim3(nblk(NF)), dim3(256), 0, s, NF,
                         (const FaceIntegrator<AreaCentroidIntegral> *)fi, U->d_face_area,
                         U->d_face_cen);
    }
    HIPCHK(hipGetLastError());
    unsigned int herr = 0;
    HIPCHK(hipMemcpyAsync(&herr, err, sizeof herr, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (herr) throw BuildError{MV_E_PLANE_DET_ZERO};
  }
  HIPCHK(hipStreamSynchronize(s));
  return U.release();
}

}  // namespace

// ---------------------------------------------------------------------
// captured builds (mv_graph_create / mv_graph_launch)
// ---------------------------------------------------------------------
// The build of a fixed set of problems recorded once as a hipGraph and
// replayed on the current contents of the caller's seed array.  Its pass
// structure and capacities come from a planning build on the seeds given at
// creation, with headroom; every count a pass needs stays on the device, so
// the recorded sequence never waits for the host.  A replay that needs more
// than was recorded (see mv_graph.hpp) runs the ordinary build instead.
struct mv_graph {
  const double *d_xyz = nullptr;
  const uint8_t *d_mask = nullptr;
  size_t n = 0;
  uint32_t P = 0, dim = 3;
  std::vector<uint64_t> offsets;
  std::vector<double> anchors, widths;
  std::vector<uint32_t> periodic;
  std::vector<ProblemSpec> probs;
  mv_options opt{};
  mv_model model{};
  // the plan: passes, and per pass the cells, list-length bound and list
  // entries each tier is recorded for
  uint32_t npass = 0;
  std::vector<uint32_t> capM, capB, LbM, LbB;
  std::vector<uint64_t> EM, EB;
  uint64_t st_cap = 0;
  size_t ncell = 0;
  uint32_t lbound0 = 0;
  // the planning build had dense cells: the recorded build splits them into
  // sub-columns too (so that its gathers start from the same radii)
  bool sub = false;
  // device memory: constant tables, the status words and the arena the
  // recorded build carves every other buffer from
  std::shared_ptr<DevStore> store = std::make_shared<DevStore>();
  uint64_t *d_goff = nullptr;
  BoxParams *d_boxes = nullptr;
  GridParams *d_grids = nullptr;
  uint32_t *d_coff = nullptr;
  DevModel *d_model = nullptr;
  FeatParams *d_fpd = nullptr;
  unsigned long long *d_status = nullptr;
  size_t status_words = 0, arena_bytes = 0;
  // outputs (inside the arena)
  double *cell_loc = nullptr, *cell_cen = nullptr, *cell_vol = nullptr, *cell_sr = nullptr;
  uint64_t *cell_foff = nullptr, *cell_fcnt = nullptr, *cell_idx = nullptr;
  uint32_t *npop = nullptr;
  uint64_t *face_left = nullptr;
  int64_t *face_right = nullptr;
  double *face_area = nullptr, *face_cen = nullptr, *face_norm = nullptr, *face_shift = nullptr;
  uint8_t *face_has_shift = nullptr;
  uint64_t *links = nullptr;
  unsigned long long *h_status = nullptr;  // pinned
  // the recorded stream, and a second one the large-cell tier forks onto
  hipStream_t s = nullptr, s2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // a launch starts after the work queued before it on the null stream
  // (the caller's writes to the seeds), as the ordinary build, which runs
  // on that stream
  hipEvent_t ev_in = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  mv_stats plan_stats{};
  uint32_t last_violation = 0;  // violation bits of the last replay
  ~mv_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    if (ev_in) (void)hipEventDestroy(ev_in);
    if (s2) (void)hipStreamDestroy(s2);
    if (s) (void)hipStreamDestroy(s);
    if (h_status) (void)hipHostFree(h_status);
  }
  template <typename T>
  T *own(size_t k) {
    T *p = dnew<T>(k);
    if (!p) throw BuildError{MV_E_OOM};
    store->dev.push_back(p);
    return p;
  }
};

namespace {

// The recorded build: the ordinary build's kernels in its order, with the
// counts on the device and the buffers sized by the plan.  Called while the
// stream is captured, with the arena active (nothing is freed).
void graph_body(mv_graph &G) {
  hipStream_t s = G.s;
  const size_t n = G.n;
  const uint32_t P = G.P, dim = G.dim, NPG = G.npass;
  const mv_options &opt = G.opt;
  Temps TT;
  auto fill32 = [&](void *p, size_t k, uint32_t v) {
    if (k)
      hipLaunchKernelGGL(k_fill_u32, dim3(nblk(k)), dim3(256), 0, s, (uint32_t *)p, (uint32_t)k,
                         v);
  };
  auto fill8 = [&](void *p, size_t k, uint8_t v, hipStream_t st) {
    if (k) hipLaunchKernelGGL(k_fill_u8, dim3(nblk(k)), dim3(256), 0, st, (uint8_t *)p, k, v);
  };
  auto fill64 = [&](void *p, size_t k, unsigned long long v) {
    if (k)
      hipLaunchKernelGGL(k_fill_u64, dim3(nblk(k)), dim3(256), 0, s, (unsigned long long *)p, k,
                         v);
  };
  unsigned int *viol = need(TT.get<unsigned int>(1));
  fill32(viol, 1, 0u);

  // ---- grid (build_grid with the recorded grid parameters)
  uint64_t maxn = 1;
  for (uint32_t p = 0; p < P; p++) maxn = std::max<uint64_t>(maxn, G.offsets[p + 1] - G.offsets[p]);
  const unsigned nb = nblk(n);
  double4 *gpos = need(TT.get<double4>(n));
  hipLaunchKernelGGL(k_prepare, dim3(nb), dim3(256), 0, s, G.d_xyz, n, (int)dim, gpos);
  uint32_t *prob = nullptr, *prob_of_slot = nullptr;
  if (P > 1) {
    prob = need(TT.get<uint32_t>(n));
    prob_of_slot = need(TT.get<uint32_t>(n));
    const uint32_t gx = (uint32_t)std::min<uint64_t>((maxn + 255) / 256, 64);
    hipLaunchKernelGGL(k_problem_of, dim3(gx, P), dim3(256), 0, s, (const uint64_t *)G.d_goff,
                       P, prob);
  }
  const uint32_t S = (uint32_t)((maxn + BBOX_CHUNK - 1) / BBOX_CHUNK);
  double *bbp = need(TT.get<double>(6 * (size_t)S * P));
  double *bbd = need(TT.get<double>(6 * (size_t)P));
  hipLaunchKernelGGL(k_bbox_part, dim3(S, P), dim3(256), 0, s, (const double4 *)gpos,
                     (const uint64_t *)G.d_goff, S, bbp);
  hipLaunchKernelGGL(k_bbox_final, dim3((P + 63) / 64), dim3(64), 0, s, (const double *)bbp, S,
                     P, bbd);
  hipLaunchKernelGGL(k_graph_grid_check, dim3((P + 63) / 64), dim3(64), 0, s,
                     (const double *)bbd, (const GridParams *)G.d_grids, P,
                     (const uint64_t *)G.d_goff, viol);
  const size_t ncell = G.ncell;
  uint32_t *cellid = need(TT.get<uint32_t>(n));
  uint32_t *counts = need(TT.get<uint32_t>(ncell + 1));
  uint32_t *gfill = need(TT.get<uint32_t>(ncell));
  uint32_t *cell_start = need(TT.get<uint32_t>(ncell + 1));
  uint32_t *order = need(TT.get<uint32_t>(n));
  double4 *spos = need(TT.get<double4>(n));
  uint32_t *slot_of = need(TT.get<uint32_t>(n));
  fill32(counts, ncell + 1, 0u);
  fill32(gfill, ncell, 0u);
  hipLaunchKernelGGL(k_grid_count, dim3(nb), dim3(256), 0, s, (const double4 *)gpos, n,
                     (const GridParams *)G.d_grids, (const uint32_t *)G.d_coff,
                     (const uint32_t *)prob, cellid, counts);
  scan_excl_dev<uint32_t, uint32_t>(counts, cell_start, ncell + 1, s);
  hipLaunchKernelGGL(k_grid_scatter, dim3(nb), dim3(256), 0, s, cellid, n, cell_start, gfill,
                     order);
  {
    uint32_t *big_list = gfill;
    unsigned int *big_cnt = counts;
    uint64_t *ta = reinterpret_cast<uint64_t *>(spos);
    uint64_t *tb = ta + n;
    fill32(big_cnt, 1, 0u);
    hipLaunchKernelGGL(k_grid_sort_cells, dim3(nblk(ncell)), dim3(256), 0, s, cell_start, ncell,
                       (const double4 *)gpos, order, big_list, big_cnt);
    hipLaunchKernelGGL(k_grid_sort_big, dim3(1024), dim3(GS_THREADS), 0, s,
                       (const uint32_t *)cell_start, (const uint32_t *)big_list,
                       (const unsigned int *)big_cnt, (const double4 *)gpos, order, ta, tb);
  }
  hipLaunchKernelGGL(k_grid_gather, dim3(nb), dim3(256), 0, s, (const double4 *)gpos,
                     (const uint32_t *)order, n, (const uint32_t *)prob, spos, slot_of,
                     prob_of_slot);
  Problems PR{G.d_boxes, G.d_grids, G.d_coff, prob_of_slot};
  if (G.sub && G.dim == 3) {
    const SubTables t = launch_subcols(n, ncell, cell_start, gfill, counts, spos, G.d_grids,
                                       G.d_coff, prob_of_slot, TT, s);
    PR.sub_base = t.base;
    PR.sub_start = t.start;
    PR.sub_slot = t.slot;
    PR.sub_key = t.key;
    PR.sub_xy = t.xy;
  }

  // ---- the first pass's cells (active slots in grid order) and lengths
  uint32_t *cm = need(TT.get<uint32_t>(NPG + 1));   // main-tier cells per pass
  uint32_t *cb = need(TT.get<uint32_t>(NPG + 1));   // large-cell tier
  uint32_t *cbn = need(TT.get<uint32_t>(NPG + 1));  // ... of those, the redone ones
  fill32(cm, NPG + 1, 0u);
  fill32(cb, NPG + 1, 0u);
  fill32(cbn, NPG + 1, 0u);
  uint32_t *flag = need(TT.get<uint32_t>(n));
  uint32_t *ident = need(TT.get<uint32_t>(n));
  uint32_t *pos = need(TT.get<uint32_t>(n));
  uint32_t *lists[5], *lens[5];  // [0]/[1] main, [2]/[3] large cells, [4] sort buffer
  for (int k = 0; k < 5; k++) {
    lists[k] = need(TT.get<uint32_t>(n));
    lens[k] = need(TT.get<uint32_t>(n));
  }
  hipLaunchKernelGGL(k_flag_active, dim3(nb), dim3(256), 0, s, spos, n, G.d_mask, flag, ident);
  scan_excl_dev<uint32_t, uint32_t>(flag, pos, n, s);
  hipLaunchKernelGGL(k_compact, dim3(nb), dim3(256), 0, s, ident, flag, pos, (uint32_t)n,
                     lists[0]);
  // the first pass's capacity: the planning build's active cells and more
  const uint32_t cap0 = G.capM[0];
  const unsigned nb0 = nblk(cap0);
  hipLaunchKernelGGL(k_graph_count, dim3(1), dim3(1), 0, s, (const uint32_t *)flag,
                     (const uint32_t *)pos, (uint32_t)n, (const uint32_t *)nullptr, cm, cap0,
                     GV_ITEMS, viol);
  hipLaunchKernelGGL(k_tail_fill_u32, dim3(nb0), dim3(256), 0, s, lists[0], (const uint32_t *)cm,
                     cap0, 0u);
  const uint32_t kmax =
      std::max<uint32_t>(std::min<uint32_t>(opt.k_max ? opt.k_max : MV_K_MAX, MV_K_MAX), 2);
  const uint32_t k0 = std::min(std::max<uint32_t>(opt.k0 ? opt.k0 : 27, 2), kmax);
  const float *Rm = nullptr;
  if (opt.mode == MV_MODE_LEARNED) {
    float *F = need(TT.get<float>((size_t)cap0 * NF));
    float *hl = need(TT.get<float>(cap0));
    float *Y = need(TT.get<float>((size_t)cap0 * NO));
    float *R0 = need(TT.get<float>(cap0));
    hipLaunchKernelGGL(k_features<true>, dim3(nb0), dim3(256), 0, s, spos, cell_start, PR,
                       (const FeatParams *)G.d_fpd, lists[0], cap0, F, hl);
    hipLaunchKernelGGL(k_features<false>, dim3(nb0), dim3(256), 0, s, spos, cell_start, PR,
                       (const FeatParams *)G.d_fpd, lists[0], cap0, F, hl);
    const uint32_t tiles = (cap0 + 15) / 16;
    hipLaunchKernelGGL(k_mlp_infer, dim3(std::min<uint32_t>((tiles + 3) / 4, 4096)), dim3(256),
                       0, s, (const DevModel *)G.d_model, (const float *)F, cap0, Y);
    const float kscale = G.model.kscale > 0.f ? G.model.kscale : 1.f;
    const float rscale = G.model.rscale > 0.f ? G.model.rscale : 0.f;
    hipLaunchKernelGGL(k_proposal, dim3(nb0), dim3(256), 0, s, (const float *)Y,
                       (const float *)hl, cap0, std::min<uint32_t>(16, kmax), G.lbound0,
                       kscale, rscale, lens[0], R0);
    hipLaunchKernelGGL(k_tail_fill_u32, dim3(nb0), dim3(256), 0, s, lens[0],
                       (const uint32_t *)cm, cap0, 0u);
    if (opt.flags & MV_FLAG_NO_LENGTH_ORDER) {
      Rm = R0;
    } else {
      // longest first; the unused tail (length 0) sorts last
      const uint32_t NB = std::min<uint32_t>(G.lbound0 / 8 + 1, LEN_MAX_BUCKETS);
      uint32_t *hist = need(TT.get<uint32_t>(NB + 1));
      uint32_t *bst = need(TT.get<uint32_t>(NB + 1));
      uint32_t *bfill = need(TT.get<uint32_t>(NB));
      fill32(hist, NB + 1, 0u);
      fill32(bfill, NB, 0u);
      const uint32_t lblocks = (cap0 + LEN_PER_BLOCK - 1) / LEN_PER_BLOCK;
      hipLaunchKernelGGL(k_len_hist, dim3(lblocks), dim3(256), 0, s, lens[0], cap0, NB, hist);
      scan_excl_dev<uint32_t, uint32_t>(hist, bst, NB + 1, s);
      float *Rs = need(TT.get<float>(cap0));
      hipLaunchKernelGGL(k_len_scatter, dim3(lblocks), dim3(256), 0, s, lists[0], lens[0], R0,
                         cap0, NB, bst, bfill, lists[4], lens[4], Rs);
      std::swap(lists[0], lists[4]);
      std::swap(lens[0], lens[4]);
      Rm = Rs;
    }
  } else {
    fill32(lens[0], cap0, k0);
    hipLaunchKernelGGL(k_tail_fill_u32, dim3(nb0), dim3(256), 0, s, lens[0],
                       (const uint32_t *)cm, cap0, 0u);
  }

  // ---- per-cell outputs, face staging, pass counters
  G.cell_loc = need(TT.get<double>(3 * n));
  G.cell_cen = need(TT.get<double>(3 * n));
  G.cell_vol = need(TT.get<double>(n));
  G.cell_sr = need(TT.get<double>(n));
  G.cell_foff = need(TT.get<uint64_t>(n));
  G.cell_fcnt = need(TT.get<uint64_t>(n));
  G.cell_idx = need(TT.get<uint64_t>(n));
  G.npop = need(TT.get<uint32_t>(n));
  fill64(G.cell_cen, 3 * n, 0ull);
  fill64(G.cell_vol, n, 0ull);
  fill64(G.cell_sr, n, 0ull);
  fill32(G.npop, n, 0u);
  uint32_t *cell_nf = need(TT.get<uint32_t>(n));
  uint64_t *cell_fbase = need(TT.get<uint64_t>(n));
  uint32_t *link_in = need(TT.get<uint32_t>(n));
  fill32(cell_nf, n, 0u);
  fill64(cell_fbase, n, 0ull);
  fill32(link_in, n, 0u);
  const uint64_t st_cap = G.st_cap;
  uint32_t *st_cell = need(TT.get<uint32_t>(st_cap));
  int64_t *st_right = need(TT.get<int64_t>(st_cap));
  uint8_t *st_img = need(TT.get<uint8_t>(st_cap));
  uint32_t *st_lslot = need(TT.get<uint32_t>(st_cap));
  double *st_area = need(TT.get<double>(st_cap));
  double *st_cen = need(TT.get<double>(3 * st_cap));
  double *st_norm = need(TT.get<double>(3 * st_cap));
  unsigned long long *counters = need(TT.get<unsigned long long>(4));
  fill64(counters, 4, 0ull);
  fill64(counters + 1, 1, ~0ull);
  unsigned int *cap_err = need(TT.get<unsigned int>(1));
  fill32(cap_err, 1, 0u);
  uint8_t *wstate = need(TT.get<uint8_t>(n));
  uint8_t *bstate = need(TT.get<uint8_t>(n));
  unsigned int *queue = need(TT.get<unsigned int>(5));
  uint32_t *exact_list = need(TT.get<uint32_t>(n));
  const uint32_t bpc = opt.clip_blocks_per_cu ? opt.clip_blocks_per_cu : 4u;
  const uint32_t cus = (uint32_t)device_cus();

  // ---- passes: the main tier (one cell per wavefront), the large-cell
  // tier; cells not certified go to the next pass with longer lists, cells
  // over the wavefront kernel's capacity to the large-cell tier
  int im = 0, ib = 2;
  const float *Rb = nullptr;
  for (uint32_t p = 0; p < NPG; p++) {
    const uint32_t capM = G.capM[p], capB = G.capB[p];
    double *need_m = need(TT.get<double>(std::max<uint32_t>(capM, 1)));
    double *need_b = need(TT.get<double>(std::max<uint32_t>(capB, 1)));
    double *kd_m = nullptr, *kd_b = nullptr;
    // the large-cell tier (few cells, long lists and latency) on a stream of
    // its own beside the main tier, as in the ordinary build
    const bool fork = capB && capM;
    hipStream_t sb = fork ? G.s2 : s;
    if (fork) {
      HIPCHK(hipEventRecord(G.ev_fork, s));
      HIPCHK(hipStreamWaitEvent(G.s2, G.ev_fork, 0));
    }
    if (capB) {
      fill8(bstate, capB, (uint8_t)CS_DONE, sb);
      uint64_t *kofs = need(TT.get<uint64_t>(capB));
      scan_excl_dev<uint32_t, uint64_t>(lens[ib], kofs, capB, sb);
      hipLaunchKernelGGL(k_graph_entries, dim3(1), dim3(1), 0, sb, (const uint32_t *)lens[ib],
                         (const uint64_t *)kofs, capB, (uint64_t)G.EB[p], cb + p, viol);
      uint64_t *kpk = need(TT.get<uint64_t>(G.EB[p]));
      uint32_t *kcnt = need(TT.get<uint32_t>(capB));
      double *kdist = need(TT.get<double>(capB));
      kd_b = kdist;
      launch_knn(spos, cell_start, PR, lists[ib], capB, G.LbB[p], lens[ib], Rb, kofs, kpk, kcnt,
                 kdist, true, TT, sb, 0, cb + p, std::min<uint32_t>(capB, BIG_KNN_BLOCKS));
      ClipArgs A;
      memset(&A, 0, sizeof A);
      A.spos = spos;
      A.mask = G.d_mask;
      A.PR = PR;
      A.work = lists[ib];
      A.nwork = capB;
      A.nsub = cb + p;
      A.kpk = kpk;
      A.kofs = kofs;
      A.kcnt = kcnt;
      A.kdist = kdist;
      A.need_r = need_b;
      A.lanes = std::min<uint32_t>(8192, capB);
      A.wstate = bstate;
      A.cell_vol = G.cell_vol;
      A.cell_cen = G.cell_cen;
      A.cell_sr = G.cell_sr;
      A.cell_nf = cell_nf;
      A.link_in = link_in;
      A.cell_fbase = cell_fbase;
      A.cell_npop = G.npop;
      A.st_cell = st_cell;
      A.st_right = st_right;
      A.st_img = st_img;
      A.st_lslot = st_lslot;
      A.st_area = st_area;
      A.st_cen = st_cen;
      A.st_norm = st_norm;
      A.st_cap = st_cap;
      A.st_counter = counters;
      A.err_word = counters + 1;
      A.exact_counter = counters + 2;
      A.overflow = counters + 3;
      A.coop_maxp = opt.coop_max_planes;
      A.coop_maxv = opt.coop_max_vertices;
      hipLaunchKernelGGL((k_clip_coop<MAXP_BIG, MAXV_BIG, false>),
                         dim3(std::max<uint32_t>(A.lanes, 1)), dim3(64), 0, sb, A);
    }
    if (capM) {
      fill8(wstate, capM, (uint8_t)CS_DONE, s);
      fill32(queue, 5, 0u);
      uint64_t *kofs = need(TT.get<uint64_t>(capM));
      scan_excl_dev<uint32_t, uint64_t>(lens[im], kofs, capM, s);
      hipLaunchKernelGGL(k_graph_entries, dim3(1), dim3(1), 0, s, (const uint32_t *)lens[im],
                         (const uint64_t *)kofs, capM, (uint64_t)G.EM[p], cm + p, viol);
      uint64_t *kpk = need(TT.get<uint64_t>(G.EM[p]));
      uint32_t *kcnt = need(TT.get<uint32_t>(capM));
      double *kdist = need(TT.get<double>(capM));
      kd_m = kdist;
      // (the first pass's radii, when given, are proposals)
      launch_knn(spos, cell_start, PR, lists[im], capM, G.LbM[p], lens[im], Rm, kofs, kpk, kcnt,
                 kdist, false, TT, s, opt.knn_min_tile, cm + p, 0, 0, p == 0 && Rm != nullptr);
      WaveArgs W;
      memset(&W, 0, sizeof W);
      W.spos = spos;
      W.mask = G.d_mask;
      W.PR = PR;
      W.work = lists[im];
      W.nwork = capM;
      W.nsub = cm + p;
      W.kpk = kpk;
      W.kofs = kofs;
      W.kcnt = kcnt;
      W.kdist = kdist;
      W.need_r = need_m;
      W.queue = queue;
      W.wstate = wstate;
      W.cell_vol = G.cell_vol;
      W.cell_cen = G.cell_cen;
      W.cell_sr = G.cell_sr;
      W.cell_nf = cell_nf;
      W.link_in = link_in;
      W.cell_fbase = cell_fbase;
      W.cell_npop = G.npop;
      W.st_cell = st_cell;
      W.st_right = st_right;
      W.st_img = st_img;
      W.st_lslot = st_lslot;
      W.st_area = st_area;
      W.st_cen = st_cen;
      W.st_norm = st_norm;
      W.st_cap = st_cap;
      W.st_counter = counters;
      W.err_word = counters + 1;
      W.exact_counter = counters + 2;
      W.overflow = counters + 3;
      W.exact_list = exact_list;
      W.exact_cnt = queue + 1;
      const uint32_t blocks = std::min<uint32_t>((capM + 3) / 4, cus * bpc);
      hipLaunchKernelGGL((k_clip_wave<false, false>), dim3(blocks), dim3(256), 0, s, W);
      WaveArgs X = W;
      X.sub = exact_list;
      X.nsub = queue + 1;
      X.queue = queue + 2;
      hipLaunchKernelGGL((k_clip_wave<false, true>), dim3(blocks), dim3(256), 0, s, X);
    }
    if (fork) {
      HIPCHK(hipEventRecord(G.ev_join, G.s2));
      HIPCHK(hipStreamWaitEvent(s, G.ev_join, 0));
    }
    hipLaunchKernelGGL(k_graph_pass_check, dim3(1), dim3(1), 0, s,
                       (const unsigned long long *)counters, (const unsigned int *)cap_err, viol);
    if (p + 1 < NPG) {
      const uint32_t capM2 = G.capM[p + 1], capB2 = G.capB[p + 1];
      const int im2 = 1 - im, ib2 = 5 - ib;
      double *rr_m = need(TT.get<double>(std::max<uint32_t>(capM2, 1)));
      double *kk_m = need(TT.get<double>(std::max<uint32_t>(capM2, 1)));
      double *rr_b = need(TT.get<double>(std::max<uint32_t>(capB2, 1)));
      double *kk_b = need(TT.get<double>(std::max<uint32_t>(capB2, 1)));
      float *Rm2 = need(TT.get<float>(std::max<uint32_t>(capM2, 1)));
      float *Rb2 = need(TT.get<float>(std::max<uint32_t>(capB2, 1)));
      auto take = [&](const uint32_t *w, const uint32_t *L, const uint8_t *st, uint32_t cnt,
                      uint8_t want, const double *nr, const double *kd, const uint32_t *base,
                      uint32_t *ow, uint32_t *oL, double *onr, double *okd, uint32_t ocap,
                      uint32_t *dst) {
        hipLaunchKernelGGL(k_flag_state, dim3(nblk(cnt)), dim3(256), 0, s, st, cnt, want, flag);
        scan_excl_dev<uint32_t, uint32_t>(flag, pos, cnt, s);
        hipLaunchKernelGGL(k_take_at, dim3(nblk(cnt)), dim3(256), 0, s, w, L, nr, kd,
                           (const uint32_t *)flag, (const uint32_t *)pos, cnt, base, ocap, ow, oL,
                           onr, okd);
        hipLaunchKernelGGL(k_graph_count, dim3(1), dim3(1), 0, s, (const uint32_t *)flag,
                           (const uint32_t *)pos, cnt, base, dst, ocap, GV_ITEMS, viol);
      };
      // not certified: longer lists in the same tier; over the wavefront
      // kernel's capacity: the large-cell tier with the same list
      if (capM)
        take(lists[im], lens[im], wstate, capM, (uint8_t)CS_NEED_MORE, need_m, kd_m, nullptr,
             lists[im2], lens[im2], rr_m, kk_m, capM2, cm + p + 1);
      if (capB)
        take(lists[ib], lens[ib], bstate, capB, (uint8_t)CS_NEED_MORE, need_b, kd_b, nullptr,
             lists[ib2], lens[ib2], rr_b, kk_b, capB2, cbn + p + 1);
      if (capM)
        take(lists[im], lens[im], wstate, capM, (uint8_t)CS_CAPACITY, nullptr, nullptr,
             cbn + p + 1, lists[ib2], lens[ib2], nullptr, nullptr, capB2, cb + p + 1);
      else
        hipLaunchKernelGGL(k_graph_count, dim3(1), dim3(1), 0, s, (const uint32_t *)flag,
                           (const uint32_t *)pos, 0u, (const uint32_t *)(cbn + p + 1),
                           cb + p + 1, capB2, GV_ITEMS, viol);
      if (capB)
        hipLaunchKernelGGL(k_graph_any_state, dim3(nblk(capB)), dim3(256), 0, s,
                           (const uint8_t *)bstate, capB, (uint8_t)CS_CAPACITY, GV_WIDE, viol);
      // (long lists bounded by the seeds within the safety radius, as in
      // the ordinary build)
      auto ball_bound = [&](uint32_t cap, const uint32_t *wl, const uint32_t *ll,
                            const double *rr, const double *kk,
                            const uint32_t *dcnt) -> const uint32_t * {
        uint32_t *ub = need(TT.get<uint32_t>(cap));
        hipLaunchKernelGGL(k_redo_count, dim3((cap + 3) / 4), dim3(256), 0, s, cap, wl, ll, rr,
                           kk, (const double4 *)spos, (const uint32_t *)cell_start, PR, REDO_COUNT_MIN,
                           ub, dcnt);
        return ub;
      };
      if (capM2) {
        const uint32_t *ubm = ball_bound(capM2, lists[im2], lens[im2], (const double *)rr_m,
                                         (const double *)kk_m, (const uint32_t *)(cm + p + 1));
        hipLaunchKernelGGL(k_redo_size, dim3(nblk(capM2)), dim3(256), 0, s, capM2, lens[im2],
                           (const double *)rr_m, (const double *)kk_m, Rm2, kmax, cap_err,
                           (const uint32_t *)(cm + p + 1), ubm);
        hipLaunchKernelGGL(k_tail_fill_u32, dim3(nblk(capM2)), dim3(256), 0, s, lens[im2],
                           (const uint32_t *)(cm + p + 1), capM2, 0u);
        hipLaunchKernelGGL(k_tail_fill_u32, dim3(nblk(capM2)), dim3(256), 0, s, lists[im2],
                           (const uint32_t *)(cm + p + 1), capM2, 0u);
      }
      if (capB2) {
        const uint32_t *ubb = ball_bound(capB2, lists[ib2], lens[ib2], (const double *)rr_b,
                                         (const double *)kk_b, (const uint32_t *)(cbn + p + 1));
        hipLaunchKernelGGL(k_redo_size, dim3(nblk(capB2)), dim3(256), 0, s, capB2, lens[ib2],
                           (const double *)rr_b, (const double *)kk_b, Rb2, kmax, cap_err,
                           (const uint32_t *)(cbn + p + 1), ubb);
        // moved cells keep their list; their gather starts from the density
        hipLaunchKernelGGL(k_range_fill_f32, dim3(nblk(capB2)), dim3(256), 0, s, Rb2,
                           (const uint32_t *)(cbn + p + 1), (const uint32_t *)(cb + p + 1),
                           capB2, 0.f);
        hipLaunchKernelGGL(k_tail_fill_u32, dim3(nblk(capB2)), dim3(256), 0, s, lens[ib2],
                           (const uint32_t *)(cb + p + 1), capB2, 0u);
        hipLaunchKernelGGL(k_tail_fill_u32, dim3(nblk(capB2)), dim3(256), 0, s, lists[ib2],
                           (const uint32_t *)(cb + p + 1), capB2, 0u);
      }
      im = im2;
      ib = ib2;
      Rm = Rm2;
      Rb = Rb2;
    } else {
      // the last recorded pass: every cell must be done
      if (capM) {
        hipLaunchKernelGGL(k_graph_any_state, dim3(nblk(capM)), dim3(256), 0, s,
                           (const uint8_t *)wstate, capM, (uint8_t)CS_NEED_MORE, GV_LEFT, viol);
        hipLaunchKernelGGL(k_graph_any_state, dim3(nblk(capM)), dim3(256), 0, s,
                           (const uint8_t *)wstate, capM, (uint8_t)CS_CAPACITY, GV_LEFT, viol);
      }
      if (capB) {
        hipLaunchKernelGGL(k_graph_any_state, dim3(nblk(capB)), dim3(256), 0, s,
                           (const uint8_t *)bstate, capB, (uint8_t)CS_NEED_MORE, GV_LEFT, viol);
        hipLaunchKernelGGL(k_graph_any_state, dim3(nblk(capB)), dim3(256), 0, s,
                           (const uint8_t *)bstate, capB, (uint8_t)CS_CAPACITY, GV_WIDE, viol);
      }
    }
  }

  // ---- faces and links (flatten! + finalize), totals on the device
  uint64_t *face_off = need(TT.get<uint64_t>(n));
  scan_excl_dev<uint32_t, uint64_t>(cell_nf, face_off, n, s);
  hipLaunchKernelGGL(k_link_count, dim3(nb), dim3(256), 0, s, cell_nf, link_in, n, G.cell_fcnt);
  scan_excl_dev<uint64_t, uint64_t>(G.cell_fcnt, G.cell_foff, n, s);
  unsigned long long *dF = need(TT.get<unsigned long long>(1));
  unsigned long long *dNL = need(TT.get<unsigned long long>(1));
  hipLaunchKernelGGL(k_graph_totals, dim3(1), dim3(1), 0, s, (const uint32_t *)cell_nf,
                     (const uint64_t *)face_off, (const uint64_t *)G.cell_fcnt,
                     (const uint64_t *)G.cell_foff, n, st_cap, viol, dF, dNL);
  G.face_left = need(TT.get<uint64_t>(st_cap));
  G.face_right = need(TT.get<int64_t>(st_cap));
  G.face_area = need(TT.get<double>(st_cap));
  G.face_cen = need(TT.get<double>(3 * st_cap));
  G.face_norm = need(TT.get<double>(3 * st_cap));
  G.face_shift = need(TT.get<double>(3 * st_cap));
  G.face_has_shift = need(TT.get<uint8_t>(st_cap));
  G.links = need(TT.get<uint64_t>(2 * st_cap));
  FinalArgs FA;
  memset(&FA, 0, sizeof FA);
  FA.st_cell = st_cell;
  FA.st_right = st_right;
  FA.st_img = st_img;
  FA.st_lslot = st_lslot;
  FA.st_area = st_area;
  FA.st_cen = st_cen;
  FA.st_norm = st_norm;
  FA.n_staged = st_cap;
  FA.n_staged_dev = dF;
  FA.cell_fbase = cell_fbase;
  FA.face_off = face_off;
  FA.boxes = G.d_boxes;
  FA.prob = prob;
  FA.face_left = G.face_left;
  FA.face_right = G.face_right;
  FA.face_area = G.face_area;
  FA.face_centroid = G.face_cen;
  FA.face_normal = G.face_norm;
  FA.face_shift = G.face_shift;
  FA.face_has_shift = G.face_has_shift;
  FA.link_in = link_in;
  FA.link_off = G.cell_foff;
  FA.links = G.links;
  hipLaunchKernelGGL(k_place_faces, dim3(nblk(st_cap)), dim3(256), 0, s, FA);
  hipLaunchKernelGGL(k_sort_links, dim3(nb), dim3(256), 0, s, (const uint64_t *)G.cell_foff,
                     (const uint32_t *)link_in, n, G.links, (const unsigned long long *)dNL);
  hipLaunchKernelGGL(k_cell_records, dim3(nb), dim3(256), 0, s, (const double4 *)gpos, n,
                     G.d_mask, (const uint64_t *)G.d_goff, (const uint32_t *)prob, G.cell_loc,
                     G.cell_idx);
  uint64_t *d_foff = nullptr, *d_loff = nullptr;
  if (P > 1) {
    d_foff = need(TT.get<uint64_t>(P + 1));
    d_loff = need(TT.get<uint64_t>(P + 1));
    hipLaunchKernelGGL(k_problem_bases, dim3((P + 1 + 63) / 64), dim3(64), 0, s,
                       (const uint64_t *)G.d_goff, P, (const uint64_t *)face_off, st_cap,
                       (const uint64_t *)G.cell_foff, 2 * st_cap, n, d_foff, d_loff,
                       (const unsigned long long *)dF, (const unsigned long long *)dNL);
    hipLaunchKernelGGL(k_localize_faces, dim3(nblk(st_cap)), dim3(256), 0, s, G.face_left,
                       G.face_right, st_cap, (const uint64_t *)G.d_goff, (const uint32_t *)prob,
                       (const unsigned long long *)dF);
    hipLaunchKernelGGL(k_localize_links, dim3(nb), dim3(256), 0, s, G.cell_foff,
                       (const uint64_t *)G.cell_fcnt, n, (const uint32_t *)prob,
                       (const uint64_t *)d_foff, (const uint64_t *)d_loff, G.links,
                       (const unsigned long long *)dNL);
  }
  hipLaunchKernelGGL(k_graph_status, dim3(nblk(P + 1)), dim3(256), 0, s,
                     (const unsigned int *)viol, (const unsigned long long *)counters,
                     (const unsigned long long *)dF, (const unsigned long long *)dNL,
                     (const uint64_t *)d_foff, (const uint64_t *)d_loff, P, G.d_status);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(G.h_status, G.d_status, G.status_words * sizeof(unsigned long long),
                        hipMemcpyDeviceToHost, s));
}

// Record graph_body on G.s with the arena `A` active.
hipGraph_t record_graph(mv_graph &G, Arena &A) {
  t_arena = &A;
  t_graph = true;
  hipGraph_t g = nullptr;
  if (hipStreamBeginCapture(G.s, hipStreamCaptureModeRelaxed) != hipSuccess) {
    t_arena = nullptr;
    t_graph = false;
    throw BuildError{MV_E_DEVICE};
  }
  try {
    graph_body(G);
  } catch (...) {
    (void)hipStreamEndCapture(G.s, &g);
    if (g) (void)hipGraphDestroy(g);
    t_arena = nullptr;
    t_graph = false;
    throw;
  }
  const hipError_t e = hipStreamEndCapture(G.s, &g);
  t_arena = nullptr;
  t_graph = false;
  if (e != hipSuccess || !g) throw BuildError{MV_E_DEVICE};
  return g;
}

// capacities from the planning build: headroom on every count a replay with
// moved seeds may change, one spare pass
void plan_graph(mv_graph &G, const BuildTrace &tr) {
  const size_t n = G.n;
  const uint32_t np = (uint32_t)tr.passes.size();
  G.npass = std::max<uint32_t>(np + 1, 2);
  G.capM.assign(G.npass, 0);
  G.capB.assign(G.npass, 0);
  G.LbM.assign(G.npass, 0);
  G.LbB.assign(G.npass, 0);
  G.EM.assign(G.npass, 0);
  G.EB.assign(G.npass, 0);
  auto grow = [&](uint64_t c, double f, uint64_t add) {
    return (uint32_t)std::min<uint64_t>(n, (uint64_t)((double)c * f) + add);
  };
  const PassRec last = tr.passes.back();
  for (uint32_t p = 0; p < G.npass; p++) {
    if (p == 0) {
      const PassRec &r = tr.passes[0];
      // the active cells (all, or the masked ones) and a tenth more
      G.capM[0] = grow(r.cnt[0], 1.1, 1024);
      G.LbM[0] = std::max<uint32_t>(r.lbnd[0], 2);
      G.EM[0] = (uint64_t)((double)r.total[0] * 1.1) + 4ull * G.capM[0] + 4096;
      continue;
    }
    if (p < np) {
      // the cells this pass took in the planning build, and half again
      const PassRec &r = tr.passes[p];
      G.capM[p] = grow(r.cnt[0], 1.5, 1024);
      G.capB[p] = grow(r.cnt[1], 1.5, 256);
      G.LbM[p] = std::max<uint32_t>(r.lbnd[0], 2);
      G.LbB[p] = std::max<uint32_t>(r.lbnd[1], 2048);
      G.EM[p] = (uint64_t)((double)r.total[0] * 1.5) +
                (uint64_t)G.capM[p] * std::min<uint32_t>(G.LbM[p], 64) + 4096;
      G.EB[p] = (uint64_t)((double)r.total[1] * 1.5) + (uint64_t)G.capB[p] * 1024 + 65536;
    } else {
      // the spare pass: a few cells with lists up to the growth of the last
      // recorded pass's (at most fourfold from one pass to the next, or
      // sixteenfold from 4,096 entries)
      auto gf = [](uint32_t lb) -> uint32_t { return lb >= GROW16_MIN ? 16u : 4u; };
      G.capM[p] = grow(0, 1.0, 1024);
      G.capB[p] = grow(0, 1.0, 256);
      G.LbM[p] = (uint32_t)std::min<uint64_t>(
          n, std::max<uint64_t>((uint64_t)gf(last.lbnd[0]) * std::max<uint32_t>(last.lbnd[0], 2),
                                64));
      G.LbB[p] = (uint32_t)std::min<uint64_t>(
          n, std::max<uint64_t>((uint64_t)gf(last.lbnd[1]) * last.lbnd[1], 2048));
      G.EM[p] = std::max<uint64_t>((uint64_t)G.capM[p] * G.LbM[p],
                                   (uint64_t)gf(last.lbnd[0]) * last.total[0]) + 4096;
      G.EB[p] = std::max<uint64_t>((uint64_t)G.capB[p] * 1024,
                                   (uint64_t)gf(last.lbnd[1]) * last.total[1]) + 65536;
    }
  }
  G.st_cap = std::max<uint64_t>(tr.F + tr.F / 4, 4096);
  G.ncell = tr.ncell;
  G.lbound0 = tr.lbound0;
  G.sub = tr.dense_cells > 0;
}

mv_status graph_create(const double *d_xyz, const uint64_t *offsets, uint32_t P,
                       const uint8_t *d_mask, const double *anchors, const double *widths,
                       const uint32_t *periodic, uint32_t dim, const mv_options *opt,
                       mv_graph **out) {
  mv_options o;
  mv_default_options(&o);
  if (opt) o = *opt;
  if (o.mode != MV_MODE_PLAIN && o.mode != MV_MODE_LEARNED) return MV_E_ARG;
  if (o.mode == MV_MODE_LEARNED &&
      (!o.model || o.model->magic != MV_MODEL_MAGIC || o.model->version != 1))
    return MV_E_ARG;
  if ((o.flags & (MV_FLAG_KEEP_CELLS | MV_FLAG_EXACT_ORDER)) || o.cells_per_wave > 1)
    return MV_E_ARG;
  if (offsets[P] == 0) return MV_E_ARG;
  std::unique_ptr<mv_graph> G(new (std::nothrow) mv_graph());
  if (!G) return MV_E_OOM;
  G->d_xyz = d_xyz;
  G->d_mask = d_mask;
  G->n = offsets[P];
  G->P = P;
  G->dim = dim;
  G->offsets.assign(offsets, offsets + P + 1);
  G->anchors.assign(anchors, anchors + 3 * P);
  G->widths.assign(widths, widths + 3 * P);
  G->periodic.assign(P, 0u);
  for (uint32_t p = 0; p < P; p++) G->periodic[p] = periodic && periodic[p] ? 1u : 0u;
  if (o.mode == MV_MODE_LEARNED) G->model = *o.model;
  o.model = o.mode == MV_MODE_LEARNED ? &G->model : nullptr;
  G->opt = o;
  // the planning build
  BuildTrace tr;
  {
    std::vector<mv_tess *> hs(P, nullptr);
    t_trace = &tr;
    const mv_status st = run_build_batch(d_xyz, offsets, P, d_mask, anchors, widths, periodic,
                                         dim, &G->opt, hs.data());
    t_trace = nullptr;
    if (st != MV_OK) return st;
    G->plan_stats = hs[0]->stats;
    for (mv_tess *h : hs) delete h;
  }
  if (tr.passes.empty()) return MV_E_ARG;
  try {
    plan_graph(*G, tr);
    G->probs.resize(P);
    for (uint32_t p = 0; p < P; p++) {
      normalise_box(anchors + 3 * p, widths + 3 * p, dim, G->probs[p].anchor,
                    G->probs[p].width);
      G->probs[p].periodic = G->periodic[p];
      G->probs[p].g0 = offsets[p];
      G->probs[p].g1 = offsets[p + 1];
    }
    // the planning build's grids, widened to hold each problem's box (a
    // seed anywhere in it stays inside); the cell counts do not change
    std::vector<GridParams> grids = tr.grids;
    for (uint32_t p = 0; p < P; p++) {
      GridParams &g = grids[p];
      double scale = 0.;
      for (int c = 0; c < 3; c++) {
        if (g.h[c] == 0.) continue;
        const double lo = std::min(g.lo[c], G->probs[p].anchor[c]);
        const double hi = std::max(g.lo[c] + (double)g.G[c] * g.h[c],
                                   G->probs[p].anchor[c] + G->probs[p].width[c]);
        g.lo[c] = lo;
        g.h[c] = (hi - lo) / g.G[c];
        g.inv_h[c] = (double)g.G[c] / (hi - lo);
        scale = std::max(scale, std::max(std::fabs(lo), std::fabs(hi)));
        scale = std::max(scale, std::fabs(G->probs[p].width[c]));
      }
      g.margin = std::max(g.margin, 1e-10 * (scale + 1e-300) + 1e-300);
    }
    hipStream_t s0 = 0;
    G->d_goff = G->own<uint64_t>(P + 1);
    G->d_boxes = G->own<BoxParams>(P);
    G->d_grids = G->own<GridParams>(P);
    G->d_coff = G->own<uint32_t>(P + 1);
    HIPCHK(hipMemcpyAsync(G->d_goff, G->offsets.data(), (P + 1) * sizeof(uint64_t),
                          hipMemcpyHostToDevice, s0));
    HIPCHK(hipMemcpyAsync(G->d_boxes, tr.boxes.data(), P * sizeof(BoxParams),
                          hipMemcpyHostToDevice, s0));
    HIPCHK(hipMemcpyAsync(G->d_grids, grids.data(), P * sizeof(GridParams),
                          hipMemcpyHostToDevice, s0));
    HIPCHK(hipMemcpyAsync(G->d_coff, tr.coff.data(), (P + 1) * sizeof(uint32_t),
                          hipMemcpyHostToDevice, s0));
    if (o.mode == MV_MODE_LEARNED) {
      G->d_model = G->own<DevModel>(1);
      upload_model(&G->model, G->d_model, s0);
      const std::vector<FeatParams> fph = feat_params(G->probs, (int)dim);
      G->d_fpd = G->own<FeatParams>(P);
      HIPCHK(hipMemcpyAsync(G->d_fpd, fph.data(), P * sizeof(FeatParams), hipMemcpyHostToDevice,
                            s0));
    }
    HIPCHK(hipStreamSynchronize(s0));
    G->status_words = 4 + 2 * ((size_t)P + 1);
    G->d_status = G->own<unsigned long long>(G->status_words);
    if (hipHostMalloc((void **)&G->h_status, G->status_words * sizeof(unsigned long long),
                      hipHostMallocDefault) != hipSuccess) {
      G->h_status = nullptr;
      throw BuildError{MV_E_OOM};
    }
    (void)device_cus();
    HIPCHK(hipStreamCreateWithFlags(&G->s, hipStreamNonBlocking));
    HIPCHK(hipStreamCreateWithFlags(&G->s2, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&G->ev_fork, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&G->ev_join, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&G->ev_in, hipEventDisableTiming));
    // record once against a counting arena (its addresses are only kernel
    // arguments of a graph that never runs) to size the arena, then for real
    {
      Arena A;
      A.base = reinterpret_cast<uint8_t *>(uintptr_t(1) << 44);
      A.size = size_t(1) << 44;
      hipGraph_t g0 = record_graph(*G, A);
      (void)hipGraphDestroy(g0);
      G->arena_bytes = A.used + 4096;
    }
    Arena A;
    A.base = G->own<uint8_t>(G->arena_bytes);
    A.size = G->arena_bytes;
    G->graph = record_graph(*G, A);
    if (A.exhausted) throw BuildError{MV_E_OOM};
    HIPCHK(hipGraphInstantiate(&G->exec, G->graph, nullptr, nullptr, 0));
  } catch (...) {
    const mv_status st = status_of_exception();
    (void)hipDeviceSynchronize();
    (void)hipGetLastError();
    return st;
  }
  *out = G.release();
  return MV_OK;
}

mv_status graph_launch(mv_graph *G, mv_tess **out, uint32_t *replayed) {
  const uint32_t P = G->P;
  auto t0 = std::chrono::steady_clock::now();
  try {
    HIPCHK(hipEventRecord(G->ev_in, 0));
    HIPCHK(hipStreamWaitEvent(G->s, G->ev_in, 0));
    HIPCHK(hipGraphLaunch(G->exec, G->s));
    HIPCHK(hipStreamSynchronize(G->s));
  } catch (...) {
    const mv_status st = status_of_exception();
    (void)hipGetLastError();
    return st;
  }
  const unsigned long long *st = G->h_status;
  G->last_violation = (uint32_t)st[0] | (st[1] != ~0ull ? (uint32_t)GV_PASS : 0u);
  if (st[0] != 0ull || st[1] != ~0ull) {
    // the replay needed more than was recorded (or met an error): the
    // ordinary build
    if (replayed) *replayed = 0;
    return run_build_batch(G->d_xyz, G->offsets.data(), P, G->d_mask, G->anchors.data(),
                           G->widths.data(), G->periodic.data(), G->dim, &G->opt, out);
  }
  const double ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  std::vector<mv_tess *> Ts(P, nullptr);
  for (uint32_t p = 0; p < P; p++) {
    Ts[p] = new (std::nothrow) mv_tess();
    if (!Ts[p]) {
      for (mv_tess *t : Ts) delete t;
      return MV_E_OOM;
    }
  }
  const uint64_t F = st[2], NL = st[3];
  for (uint32_t p = 0; p < P; p++) {
    mv_tess *U = Ts[p];
    const uint64_t g0 = G->offsets[p];
    const uint64_t f0 = P > 1 ? st[4 + p] : 0, f1 = P > 1 ? st[4 + p + 1] : F;
    const uint64_t l0 = P > 1 ? st[5 + P + p] : 0, l1 = P > 1 ? st[5 + P + p + 1] : NL;
    U->parent = G->store;
    memcpy(U->anchor, G->probs[p].anchor, sizeof U->anchor);
    memcpy(U->width, G->probs[p].width, sizeof U->width);
    U->dim = G->dim;
    U->periodic = G->probs[p].periodic;
    U->n = G->offsets[p + 1] - g0;
    U->nf = f1 - f0;
    U->nl = l1 - l0;
    U->d_cell_loc = G->cell_loc + 3 * g0;
    U->d_cell_cen = G->cell_cen + 3 * g0;
    U->d_cell_vol = G->cell_vol + g0;
    U->d_cell_sr = G->cell_sr + g0;
    U->d_cell_foff = G->cell_foff + g0;
    U->d_cell_fcnt = G->cell_fcnt + g0;
    U->d_cell_idx = G->cell_idx + g0;
    U->d_npop = G->npop + g0;
    U->d_face_left = G->face_left + f0;
    U->d_face_right = G->face_right + f0;
    U->d_face_area = G->face_area + f0;
    U->d_face_cen = G->face_cen + 3 * f0;
    U->d_face_norm = G->face_norm + 3 * f0;
    U->d_face_shift = G->face_shift + 3 * f0;
    U->d_face_has_shift = G->face_has_shift + f0;
    U->d_links = G->links + l0;
    U->stats = G->plan_stats;
    U->stats.passes = G->npass;
    U->stats.ms_total = ms;
    out[p] = U;
  }
  if (replayed) *replayed = 1;
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
  opt->k_max = MV_K_MAX;
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

mv_status mv_build_batch_device(const double *d_xyz, const uint64_t *offsets, uint32_t P,
                                const uint8_t *d_mask, const double *anchors,
                                const double *widths, const uint32_t *periodic,
                                uint32_t dimensionality, const mv_options *opt, mv_tess **out) {
  if (!out || !valid_batch(offsets, P, dimensionality, anchors, widths)) return MV_E_ARG;
  for (uint32_t p = 0; p < P; p++) out[p] = nullptr;
  if (offsets[P] && !d_xyz) return MV_E_ARG;
  if (mv_status st = device_ready()) return st;
  return run_build_batch(d_xyz, offsets, P, d_mask, anchors, widths, periodic, dimensionality,
                         opt, out);
}

mv_status mv_graph_create(const double *d_xyz, const uint64_t *offsets, uint32_t P,
                          const uint8_t *d_mask, const double *anchors, const double *widths,
                          const uint32_t *periodic, uint32_t dimensionality,
                          const mv_options *opt, mv_graph **out) {
  if (!out || !valid_batch(offsets, P, dimensionality, anchors, widths)) return MV_E_ARG;
  *out = nullptr;
  if (!d_xyz) return MV_E_ARG;
  if (mv_status st = device_ready()) return st;
  return graph_create(d_xyz, offsets, P, d_mask, anchors, widths, periodic, dimensionality, opt,
                      out);
}

mv_status mv_graph_launch(mv_graph *g, mv_tess **out, uint32_t *replayed) {
  if (!g || !out) return MV_E_ARG;
  for (uint32_t p = 0; p < g->P; p++) out[p] = nullptr;
  if (replayed) *replayed = 0;
  return graph_launch(g, out, replayed);
}

void mv_graph_free(mv_graph *g) { delete g; }

uint32_t mv_graph_last_violation(const mv_graph *g) { return g ? g->last_violation : 0u; }

mv_status mv_build_batch(const double *xyz, const uint64_t *offsets, uint32_t P,
                         const uint8_t *mask, const double *anchors, const double *widths,
                         const uint32_t *periodic, uint32_t dimensionality,
                         const mv_options *opt, mv_tess **out) {
  if (!out || !valid_batch(offsets, P, dimensionality, anchors, widths)) return MV_E_ARG;
  for (uint32_t p = 0; p < P; p++) out[p] = nullptr;
  const size_t n = offsets[P];
  if (n && !xyz) return MV_E_ARG;
  if (mv_status st = device_ready()) return st;
  double *d_xyz = nullptr;
  uint8_t *d_mask = nullptr;
  mv_status st = MV_OK;
  if (n) {
    d_xyz = (double *)dalloc(3 * n * sizeof(double));
    if (!d_xyz) return MV_E_OOM;
    if (hipMemcpy(d_xyz, xyz, 3 * n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
      st = MV_E_DEVICE;
    if (st == MV_OK && mask) {
      d_mask = (uint8_t *)dalloc(n);
      if (!d_mask)
        st = MV_E_OOM;
      else if (hipMemcpy(d_mask, mask, n, hipMemcpyHostToDevice) != hipSuccess)
        st = MV_E_DEVICE;
    }
  }
  if (st == MV_OK)
    st = run_build_batch(d_xyz, offsets, P, d_mask, anchors, widths, periodic, dimensionality,
                         opt, out);
  dfree(d_xyz);
  dfree(d_mask);
  if (st == MV_OK) {
    st = guarded([&] {
      for (uint32_t p = 0; p < P; p++) download(out[p]);
    });
    if (st != MV_OK)
      for (uint32_t p = 0; p < P; p++) {
        delete out[p];
        out[p] = nullptr;
      }
  }
  return st;
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
    } catch (...) {
      const mv_status st2 = status_of_exception();
      delete *out;
      *out = nullptr;
      return st2;
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
    } catch (...) {                              \
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
    case MV_E_FACE_WALK: return "no next vertex connected to the current one around a face";
    case MV_E_NOT_3D: return "faces can only be extracted in 3D";
    case MV_E_SCRATCH:
      return "input too large for the allocation-free build's static scratch";
    case MV_E_OUT_SMALL: return "an output array is too short (counts returned)";
  }
  return "unknown status";
}

mv_status mv_model_features(const double *xyz, size_t n, const double anchor[3],
                            const double width[3], uint32_t dimensionality,
                            uint32_t periodic, float *features, float *hloc) {
  if (!valid_args(n, dimensionality, anchor, width) || (n && (!xyz || !features || !hloc)))
    return MV_E_ARG;
  if (n == 0) return MV_OK;
  if (mv_status st = device_ready()) return st;
  return guarded([&] {
    hipStream_t s = 0;
    Temps TT;
    double *d_xyz = need(TT.get<double>(3 * n));
    HIPCHK(hipMemcpyAsync(d_xyz, xyz, 3 * n * sizeof(double), hipMemcpyHostToDevice, s));
    GridData D;
    std::vector<ProblemSpec> probs(1);
    normalise_box(anchor, width, dimensionality, probs[0].anchor, probs[0].width);
    probs[0].periodic = periodic ? 1u : 0u;
    probs[0].g0 = 0;
    probs[0].g1 = n;
    build_grid(D, d_xyz, n, probs, dimensionality, 2.0, TT, s);
    const std::vector<FeatParams> fph = feat_params(probs, (int)dimensionality);
    FeatParams *fpd = need(TT.get<FeatParams>(1));
    HIPCHK(hipMemcpyAsync(fpd, fph.data(), sizeof(FeatParams), hipMemcpyHostToDevice, s));
    uint32_t *flag = need(TT.get<uint32_t>(n));
    uint32_t *ident = need(TT.get<uint32_t>(n));
    hipLaunchKernelGGL(k_flag_active, dim3(nblk(n)), dim3(256), 0, s, D.spos, n,
                       (const uint8_t *)nullptr, flag, ident);
    float *F = need(TT.get<float>(n * NF));
    float *hl = need(TT.get<float>(n));
    float *Fo = need(TT.get<float>(n * NF));
    float *ho = need(TT.get<float>(n));
    hipLaunchKernelGGL(k_features<true>, dim3(nblk(n)), dim3(256), 0, s, D.spos, D.cell_start,
                       D.PR(), (const FeatParams *)fpd, ident, (uint32_t)n, F, hl);
    hipLaunchKernelGGL(k_features<false>, dim3(nblk(n)), dim3(256), 0, s, D.spos, D.cell_start,
                       D.PR(), (const FeatParams *)fpd, ident, (uint32_t)n, F, hl);
    hipLaunchKernelGGL(k_unpermute, dim3(nblk(n)), dim3(256), 0, s, D.order, n, NF, F, Fo);
    hipLaunchKernelGGL(k_unpermute, dim3(nblk(n)), dim3(256), 0, s, D.order, n, 1, hl, ho);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(features, Fo, n * NF * sizeof(float), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(hloc, ho, n * sizeof(float), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
  });
}

mv_status mv_model_predict(const mv_model *model, const float *features, size_t n,
                           float *out) {
  if (!model || model->magic != MV_MODEL_MAGIC || model->version != 1) return MV_E_ARG;
  if (n && (!features || !out)) return MV_E_ARG;
  if (n == 0) return MV_OK;
  if (n >= (1ull << 31)) return MV_E_ARG;
  if (mv_status st = device_ready()) return st;
  return guarded([&] {
    hipStream_t s = 0;
    Temps TT;
    DevModel *dm = need(TT.get<DevModel>(1));
    upload_model(model, dm, s);
    float *F = need(TT.get<float>(n * NF));
    float *Y = need(TT.get<float>(n * NO));
    HIPCHK(hipMemcpyAsync(F, features, n * NF * sizeof(float), hipMemcpyHostToDevice, s));
    const uint32_t tiles = (uint32_t)((n + 15) / 16);
    hipLaunchKernelGGL(k_mlp_infer, dim3(std::min<uint32_t>((tiles + 3) / 4, 4096)), dim3(256),
                       0, s, dm, F, (uint32_t)n, Y);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, Y, n * NO * sizeof(float), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
  });
}

mv_status mv_model_loss_grad(const mv_model *model, const float *features,
                             const float *targets, size_t n, float tau, float *loss,
                             float *grad) {
  if (!model || model->magic != MV_MODEL_MAGIC || model->version != 1) return MV_E_ARG;
  if (n == 0 || n % 16 != 0 || n > (1u << 24) || !features || !targets || !loss || !grad)
    return MV_E_ARG;
  if (mv_status st = device_ready()) return st;
  return guarded([&] {
    hipStream_t s = 0;
    Temps TT;
    const int B = (int)n;
    std::vector<float> P;
    model_to_params(*model, P);
    float *dP = need(TT.get<float>(P_TOTAL));
    float *dG = need(TT.get<float>(P_TOTAL));
    HIPCHK(hipMemcpyAsync(dP, P.data(), P_TOTAL * sizeof(float), hipMemcpyHostToDevice, s));
    TrainBufs t = train_bufs(B, TT);
    // normalised inputs, as the inference kernel forms them
    std::vector<float> X((size_t)B * NF);
    for (int r = 0; r < B; r++)
      for (int c = 0; c < NF; c++) {
        const float sg = model->sigma[c] > 0.f ? model->sigma[c] : 1.f;
        X[(size_t)r * NF + c] = (features[(size_t)r * NF + c] - model->mu[c]) / sg;
      }
    HIPCHK(hipMemcpyAsync(t.X, X.data(), X.size() * sizeof(float), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(t.Tg, targets, (size_t)B * NO * sizeof(float), hipMemcpyHostToDevice,
                          s));
    forward_backward(t, dP, dG, tau, s);
    std::vector<float> G(P_TOTAL), lr(B);
    HIPCHK(hipMemcpyAsync(G.data(), dG, P_TOTAL * sizeof(float), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(lr.data(), t.lrow, B * sizeof(float), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    double L = 0;
    for (float v : lr) L += v;
    *loss = (float)(L / B);
    float *g = grad;
    for (int i = 0; i < NF * NH; i++) *g++ = G[P_W1 + i];
    for (int i = 0; i < NH; i++) *g++ = G[P_B1 + i];
    for (int i = 0; i < NH * NH; i++) *g++ = G[P_W2 + i];
    for (int i = 0; i < NH; i++) *g++ = G[P_B2 + i];
    for (int i = 0; i < NH; i++)
      for (int o = 0; o < NO; o++) *g++ = G[P_W3 + i * NOP + o];
    for (int o = 0; o < NO; o++) *g++ = G[P_B3 + o];
  });
}

mv_status mv_model_train(const float *features, const float *targets, size_t n,
                         const mv_train_config *cfg, mv_model *model, float *final_loss) {
  if (!features || !targets || !cfg || !model || n < 16 || n >= (1ull << 31)) return MV_E_ARG;
  if (cfg->batch == 0 || cfg->batch % 16 != 0 || cfg->batch > (1u << 20) || cfg->steps == 0 ||
      !(cfg->lr > 0.f) || !(cfg->tau > 0.f && cfg->tau < 1.f))
    return MV_E_ARG;
  if (mv_status st = device_ready()) return st;
  return guarded([&] {
    hipStream_t s = 0;
    Temps TT;
    // normalisation constants
    mv_model m;
    memset(&m, 0, sizeof m);
    m.magic = MV_MODEL_MAGIC;
    m.version = 1;
    m.tau = cfg->tau;
    m.kscale = 1.f;
    m.rscale = 0.f;
    for (int c = 0; c < NF; c++) {
      double a = 0, b = 0;
      for (size_t r = 0; r < n; r++) {
        const double v = features[r * NF + c];
        a += v;
        b += v * v;
      }
      const double mu = a / n, var = std::max(b / n - mu * mu, 0.0);
      m.mu[c] = (float)mu;
      m.sigma[c] = var > 1e-12 ? (float)std::sqrt(var) : 1.f;
    }
    // He initialisation; output bias at the target quantile
    uint64_t rs = 0x6d76ull ^ ((uint64_t)cfg->seed << 20);
    std::vector<float> P(P_TOTAL, 0.f);
    for (int i = 0; i < NF * NH; i++) P[P_W1 + i] = (float)(gauss(rs) * std::sqrt(2.0 / NF));
    for (int i = 0; i < NH * NH; i++) P[P_W2 + i] = (float)(gauss(rs) * std::sqrt(2.0 / NH));
    for (int i = 0; i < NH; i++)
      for (int o = 0; o < NO; o++)
        P[P_W3 + i * NOP + o] = (float)(gauss(rs) * std::sqrt(1.0 / NH) * 0.1);
    for (int o = 0; o < NO; o++) {
      std::vector<float> col(n);
      for (size_t r = 0; r < n; r++) col[r] = targets[r * NO + o];
      const size_t q = std::min(n - 1, (size_t)(cfg->tau * (double)n));
      std::nth_element(col.begin(), col.begin() + q, col.end());
      P[P_B3 + o] = col[q];
    }
    float *dF = need(TT.get<float>(n * NF));
    float *dT = need(TT.get<float>(n * NO));
    float *dmu = need(TT.get<float>(NF));
    float *dsg = need(TT.get<float>(NF));
    float *dP = need(TT.get<float>(P_TOTAL));
    float *dG = need(TT.get<float>(P_TOTAL));
    float *dm1 = need(TT.get<float>(P_TOTAL));
    float *dm2 = need(TT.get<float>(P_TOTAL));
    HIPCHK(hipMemcpyAsync(dF, features, n * NF * sizeof(float), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dT, targets, n * NO * sizeof(float), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dmu, m.mu, sizeof m.mu, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dsg, m.sigma, sizeof m.sigma, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dP, P.data(), P_TOTAL * sizeof(float), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemsetAsync(dm1, 0, P_TOTAL * sizeof(float), s));
    HIPCHK(hipMemsetAsync(dm2, 0, P_TOTAL * sizeof(float), s));
    const int B = (int)cfg->batch;
    TrainBufs t = train_bufs(B, TT);
    // the row indices of every step, drawn up front
    const uint32_t chunk = 256;
    std::vector<uint32_t> idx((size_t)B * chunk);
    uint32_t *didx = need(TT.get<uint32_t>((size_t)B * chunk));
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    double c1 = 1, c2 = 1;
    for (uint32_t step = 0; step < cfg->steps; step++) {
      const uint32_t j = step % chunk;
      if (j == 0) {
        HIPCHK(hipStreamSynchronize(s));
        for (auto &v : idx) v = (uint32_t)(splitmix(rs) % n);
        HIPCHK(hipMemcpyAsync(didx, idx.data(), idx.size() * sizeof(uint32_t),
                              hipMemcpyHostToDevice, s));
      }
      hipLaunchKernelGGL(k_gather_rows, dim3((B * NF + 255) / 256), dim3(256), 0, s, dF, dT,
                         didx + (size_t)j * B, B, dmu, dsg, t.X, t.Tg);
      forward_backward(t, dP, dG, cfg->tau, s);
      c1 *= b1;
      c2 *= b2;
      // cosine decay of the step size over the run
      const float lr = (float)(cfg->lr * 0.5 *
                               (1.0 + std::cos(3.141592653589793 * step / cfg->steps)));
      hipLaunchKernelGGL(k_adam, dim3((P_TOTAL + 255) / 256), dim3(256), 0, s, dP, dG, dm1, dm2,
                         P_TOTAL, lr, b1, b2, eps, (float)(1 - c1), (float)(1 - c2));
      HIPCHK(hipGetLastError());
    }
    HIPCHK(hipMemcpyAsync(P.data(), dP, P_TOTAL * sizeof(float), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    params_to_model(P, m);
    *model = m;
    if (final_loss) {
      // loss of the BF16 checkpoint over every row, through the inference kernel
      DevModel *dmod = need(TT.get<DevModel>(1));
      upload_model(&m, dmod, s);
      float *Y = need(TT.get<float>(n * NO));
      const uint32_t tiles = (uint32_t)((n + 15) / 16);
      hipLaunchKernelGGL(k_mlp_infer, dim3(std::min<uint32_t>((tiles + 3) / 4, 4096)),
                         dim3(256), 0, s, dmod, dF, (uint32_t)n, Y);
      HIPCHK(hipGetLastError());
      std::vector<float> Yh(n * NO);
      HIPCHK(hipMemcpyAsync(Yh.data(), Y, n * NO * sizeof(float), hipMemcpyDeviceToHost, s));
      HIPCHK(hipStreamSynchronize(s));
      double L = 0;
      for (size_t i = 0; i < n * NO; i++) {
        const double e = (double)targets[i] - Yh[i];
        L += e > 0 ? cfg->tau * e : (cfg->tau - 1.0) * e;
      }
      *final_loss = (float)(L / n);
    }
  });
}

size_t mv_n_active(const mv_tess *t) {
  if (!t) return 0;
  if (!t->keep) return 0;
  (void)guarded([&] { ensure_active(const_cast<mv_tess *>(t)); });
  return t->n_active;
}

mv_status mv_cell_integrals(mv_tess *t, uint32_t kind, double *out) {
  if (!t || !t->keep || kind > MV_CELL_VOLUME_CENTROID) return MV_E_ARG;
  std::lock_guard<std::mutex> g(t->mu);
  if (t->n && !out) return MV_E_ARG;
  return guarded([&] {
    if (kind == MV_CELL_VOLUME)
      run_cell_integrals<VolumeIntegral>(t, out, 1);
    else
      run_cell_integrals<VolumeCentroidIntegral>(t, out, 4);
  });
}

mv_status mv_face_integrals(mv_tess *t, uint32_t kind, uint32_t symmetric, size_t *count,
                            uint64_t *left, int64_t *right, uint8_t *has_shift, double *shift,
                            int32_t *plane, double *values) {
  if (!t || !t->keep || !count || kind > MV_FACE_AREA_CENTROID) return MV_E_ARG;
  std::lock_guard<std::mutex> g(t->mu);
  const int sym = symmetric ? 1 : 0;
  return guarded([&] {
    if (kind == MV_FACE_AREA)
      run_face_integrals<AreaIntegral>(t, sym, count, left, right, has_shift, shift, plane,
                                       values, 1);
    else
      run_face_integrals<AreaCentroidIntegral>(t, sym, count, left, right, has_shift, shift,
                                               plane, values, 4);
  });
}

mv_status mv_tets(mv_tess *t, size_t *count, double *v012, int32_t *plane,
                  uint64_t *cell_offset) {
  if (!t || !t->keep || !count) return MV_E_ARG;
  std::lock_guard<std::mutex> g(t->mu);
  return guarded([&] {
    hipStream_t s = 0;
    Temps TT;
    ensure_active(t);
    const uint64_t na = t->n_active;
    if (!na) {
      *count = 0;
      if (v012 && cell_offset) cell_offset[0] = 0;
      return;
    }
    uint64_t *cnt = need(TT.get<uint64_t>(na));
    uint64_t *off = need(TT.get<uint64_t>(na));
    const CellGeom G = geom_of(t);
    hipLaunchKernelGGL(k_gather_nv, dim3(nblk(na)), dim3(256), 0, s,
                       (const uint64_t *)t->d_active, na, (const uint32_t *)t->cs.nv, G.nf,
                       cnt);
    const uint64_t K = scan_excl<uint64_t, uint64_t>(cnt, off, na, s);
    if (!v012) {
      *count = K;
      return;
    }
    if (*count < K || !plane || !cell_offset) throw BuildError{MV_E_ARG};
    *count = K;
    double *dv = need(TT.get<double>(9 * K));
    int32_t *dp = need(TT.get<int32_t>(K));
    unsigned int *err = need(TT.get<unsigned int>(1));
    HIPCHK(hipMemsetAsync(err, 0, sizeof(unsigned int), s));
    hipLaunchKernelGGL(k_export_tets, dim3(nblk(na)), dim3(256), 0, s, G,
                       (const uint64_t *)t->d_active, na, (const uint64_t *)off, dv, dp, err);
    HIPCHK(hipGetLastError());
    unsigned int herr = 0;
    HIPCHK(hipMemcpyAsync(v012, dv, 9 * K * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(plane, dp, K * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(cell_offset, off, na * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(&herr, err, sizeof herr, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    cell_offset[na] = K;
    if (herr) throw BuildError{MV_E_PLANE_DET_ZERO};
  });
}

mv_status mv_kept_cells(mv_tess *t, size_t *n_planes, size_t *n_vertices,
                        uint64_t *plane_offset, uint64_t *vertex_offset, double *plane_n,
                        double *plane_p, int64_t *plane_right, uint8_t *plane_has_shift,
                        double *plane_shift, double *vertex_loc, int32_t *vertex_dual) {
  if (!t || !t->keep || !n_planes || !n_vertices) return MV_E_ARG;
  std::lock_guard<std::mutex> g(t->mu);
  return guarded([&] {
    hipStream_t s = 0;
    Temps TT;
    ensure_active(t);
    const uint64_t na = t->n_active;
    if (!na) {
      *n_planes = *n_vertices = 0;
      if (plane_n && plane_offset && vertex_offset) plane_offset[0] = vertex_offset[0] = 0;
      return;
    }
    const CellGeom G = geom_of(t);
    uint64_t *npc = need(TT.get<uint64_t>(na));
    uint64_t *nvc = need(TT.get<uint64_t>(na));
    uint64_t *poff = need(TT.get<uint64_t>(na));
    uint64_t *voff = need(TT.get<uint64_t>(na));
    hipLaunchKernelGGL(k_kept_counts, dim3(nblk(na)), dim3(256), 0, s, G,
                       (const uint64_t *)t->d_active, na, npc, nvc);
    HIPCHK(hipGetLastError());
    const uint64_t P = scan_excl<uint64_t, uint64_t>(npc, poff, na, s);
    const uint64_t V = scan_excl<uint64_t, uint64_t>(nvc, voff, na, s);
    if (!plane_n) {
      *n_planes = P;
      *n_vertices = V;
      return;
    }
    if (*n_planes < P || *n_vertices < V || !plane_p || !plane_right || !plane_has_shift ||
        !plane_shift || !vertex_loc || !vertex_dual || !plane_offset || !vertex_offset)
      throw BuildError{MV_E_ARG};
    *n_planes = P;
    *n_vertices = V;
    double *dpn = need(TT.get<double>(3 * P)), *dpp = need(TT.get<double>(3 * P));
    double *dsh = need(TT.get<double>(3 * P)), *dvl = need(TT.get<double>(3 * V));
    int64_t *drt = need(TT.get<int64_t>(P));
    uint8_t *dhs = need(TT.get<uint8_t>(P));
    int32_t *dvd = need(TT.get<int32_t>(3 * V));
    hipLaunchKernelGGL(k_kept_export, dim3(nblk(na)), dim3(256), 0, s, G,
                       (const uint64_t *)t->d_active, na, (const uint64_t *)poff,
                       (const uint64_t *)voff, dpn, dpp, drt, dhs, dsh, dvl, dvd);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(plane_n, dpn, 3 * P * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(plane_p, dpp, 3 * P * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(plane_shift, dsh, 3 * P * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(plane_right, drt, P * sizeof(int64_t), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(plane_has_shift, dhs, P, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(vertex_loc, dvl, 3 * V * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(vertex_dual, dvd, 3 * V * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(plane_offset, poff, na * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(vertex_offset, voff, na * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    plane_offset[na] = P;
    vertex_offset[na] = V;
  });
}

mv_status mv_with_faces(mv_tess *t) {
  if (!t || !t->keep) return MV_E_ARG;
  std::lock_guard<std::mutex> g(t->mu);
  return guarded([&] { run_with_faces(t); });
}

mv_status mv_discard_faces(mv_tess *t) {
  if (!t || !t->keep) return MV_E_ARG;
  std::lock_guard<std::mutex> g(t->mu);
  t->wf_on = false;
  return MV_OK;
}

mv_status mv_face_vertices(mv_tess *t, size_t *n_faces, size_t *n_entries,
                           uint64_t *cell_face_offset, int32_t *face_plane,
                           uint32_t *face_vertex_count, uint64_t *face_vertex_offset,
                           uint32_t *face_vertices) {
  if (!t || !t->keep || !n_faces || !n_entries) return MV_E_ARG;
  std::lock_guard<std::mutex> g(t->mu);
  if (!t->wf_on || !t->wf_ready) return MV_E_ARG;
  const uint64_t NF = t->wf_nfaces, NE = t->wf_nentries, na = t->n_active;
  if (!face_plane) {
    *n_faces = NF;
    *n_entries = NE;
    return MV_OK;
  }
  if (*n_faces < NF || *n_entries < NE || !cell_face_offset || !face_vertex_count ||
      !face_vertex_offset || !face_vertices)
    return MV_E_ARG;
  *n_faces = NF;
  *n_entries = NE;
  return guarded([&] {
    hipStream_t s = 0;
    HIPCHK(hipMemcpyAsync(cell_face_offset, t->d_wf_cell_foff, (na + 1) * sizeof(uint64_t),
                          hipMemcpyDeviceToHost, s));
    if (NF) {
      HIPCHK(hipMemcpyAsync(face_plane, t->d_wf_plane, NF * sizeof(int32_t),
                            hipMemcpyDeviceToHost, s));
      HIPCHK(hipMemcpyAsync(face_vertex_count, t->d_wf_nvert, NF * sizeof(uint32_t),
                            hipMemcpyDeviceToHost, s));
      HIPCHK(hipMemcpyAsync(face_vertex_offset, t->d_wf_voff, NF * sizeof(uint64_t),
                            hipMemcpyDeviceToHost, s));
    }
    if (NE)
      HIPCHK(hipMemcpyAsync(face_vertices, t->d_wf_vtx, NE * sizeof(uint32_t),
                            hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
  });
}

mv_status mv_voronoi_from_integrator(const mv_tess *tc, mv_tess **out) {
  if (!tc || !tc->keep || !out) return MV_E_ARG;
  *out = nullptr;
  mv_tess *t = const_cast<mv_tess *>(tc);
  std::lock_guard<std::mutex> g(t->mu);
  return guarded([&] { *out = voronoi_from_kept(t); });
}

mv_status mv_get_cell_geometry(const mv_tess *t, mv_cell_geometry *g) {
  if (!t || !g || !t->keep) return MV_E_ARG;
  const CellGeom G = geom_of(t);
  static_assert(sizeof(CellGeom) == sizeof(mv_cell_geometry), "mv_cell_geometry layout");
  memcpy(g, &G, sizeof G);
  return MV_OK;
}

uint64_t mv_test_device_mallocs(void) { return g_device_mallocs.load(); }

mv_status mv_build_device_noalloc(const double *d_xyz, size_t n, const uint8_t *d_mask,
                                  const double anchor[3], const double width[3],
                                  uint32_t dimensionality, uint32_t periodic,
                                  const mv_options *opt, mv_noalloc_out *out) {
  if (!out) return MV_E_ARG;
  out->n_faces = out->n_links = out->scratch_used = 0;
  if (!valid_args(n, dimensionality, anchor, width) || (n && !d_xyz)) return MV_E_ARG;
  if (n > MV_NOALLOC_MAX_SEEDS) return MV_E_SCRATCH;
  mv_options o;
  mv_default_options(&o);
  if (opt) o = *opt;
  if (o.flags & MV_FLAG_KEEP_CELLS) return MV_E_ARG;
  if (o.mode != MV_MODE_PLAIN && o.mode != MV_MODE_LEARNED) return MV_E_ARG;
  if (o.mode == MV_MODE_LEARNED &&
      (!o.model || o.model->magic != MV_MODEL_MAGIC || o.model->version != 1))
    return MV_E_ARG;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NOALLOC_MAX_DEVICES) {
    (void)hipGetLastError();
    return MV_E_DEVICE;
  }
  NoAllocDevice &ND = noalloc_device(dev);
  std::lock_guard<std::mutex> g(ND.mu);
  if (!ND.ptr) {
    // first call on this device (current): the device address of its copy
    // of `scratch`; host code uses only this pointer, never the symbol
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
      (void)hipGetLastError();
      return MV_E_DEVICE;
    }
    ND.ptr = p;
  }
  Arena A;
  A.base = (uint8_t *)ND.ptr;
  A.size = MV_NOALLOC_SCRATCH_BYTES;
  A.words = (unsigned long long *)arena_alloc(A, 8 * sizeof(unsigned long long));
  t_arena = &A;
  struct Leave {
    ~Leave() { t_arena = nullptr; }
  } leave;
  mv_status st = MV_OK;
  try {
    mv_tess T;
    std::vector<ProblemSpec> probs(1);
    normalise_box(anchor, width, dimensionality, probs[0].anchor, probs[0].width);
    probs[0].periodic = periodic ? 1 : 0;
    probs[0].g0 = 0;
    probs[0].g1 = n;
    build_multi({&T}, d_xyz, n, d_mask, probs, dimensionality, o);
    out->n_faces = T.nf;
    out->n_links = T.nl;
    if (T.nf > out->face_cap || T.nl > out->link_cap) {
      st = MV_E_OUT_SMALL;
    } else {
      const hipStream_t s = 0;
      auto put = [&](void *dst, const void *src, size_t bytes) {
        if (dst && bytes) HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s));
      };
      const size_t F = T.nf, NL = T.nl;
      put(out->cell_loc, T.d_cell_loc, 3 * n * sizeof(double));
      put(out->cell_centroid, T.d_cell_cen, 3 * n * sizeof(double));
      put(out->cell_volume, T.d_cell_vol, n * sizeof(double));
      put(out->cell_safety_radius, T.d_cell_sr, n * sizeof(double));
      put(out->cell_face_offset, T.d_cell_foff, n * sizeof(uint64_t));
      put(out->cell_face_count, T.d_cell_fcnt, n * sizeof(uint64_t));
      put(out->cell_idx, T.d_cell_idx, n * sizeof(uint64_t));
      put(out->face_left, T.d_face_left, F * sizeof(uint64_t));
      put(out->face_right, T.d_face_right, F * sizeof(int64_t));
      put(out->face_area, T.d_face_area, F * sizeof(double));
      put(out->face_centroid, T.d_face_cen, 3 * F * sizeof(double));
      put(out->face_normal, T.d_face_norm, 3 * F * sizeof(double));
      put(out->face_has_shift, T.d_face_has_shift, F);
      put(out->face_shift, T.d_face_shift, 3 * F * sizeof(double));
      put(out->cell_face_connections, T.d_links, NL * sizeof(uint64_t));
      HIPCHK(hipStreamSynchronize(s));
    }
  } catch (...) {
    st = A.exhausted ? MV_E_SCRATCH : status_of_exception();
    (void)hipDeviceSynchronize();
    (void)hipGetLastError();
  }
  out->scratch_used = A.used;
  return st;
}

mv_status mv_test_run_shares(int32_t threads, int32_t fail_after, int32_t *out) {
  if (threads < 0 || threads > 4096 || (threads && !out)) return MV_E_ARG;
  return guarded([&] { run_shares(threads, fail_after, [&](int w) { out[w] = w + 1; }); });
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
