// mv_graph.hpp -- kernels of the captured build (mv_graph_create /
// mv_graph_launch): every count a pass needs stays on the device, so the
// captured sequence never waits for the host.  Capacities come from a
// planning build; when a replay needs more than was captured (more cells to
// redo, longer lists, more faces, a seed outside the captured grid, an
// error), a violation bit is set, the counts that would overflow are cut to
// zero, and the host runs the ordinary build instead.
#pragma once
#include "mv_common.hpp"

namespace mv {

// violation bits (status word 0)
enum : unsigned int {
  GV_GRID = 1u,      // a seed outside the captured grid
  GV_ITEMS = 2u,     // more cells for a pass than its capacity
  GV_ENTRIES = 4u,   // more list entries than captured
  GV_PASS = 8u,      // an error status or full face staging in a pass
  GV_LEFT = 16u,     // cells still unfinished after the last captured pass
  GV_WIDE = 32u,     // a cell over the large-cell kernel's capacity
};

__global__ void k_fill_u64(unsigned long long *__restrict__ a, size_t n, unsigned long long v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}
__global__ void k_fill_u8(uint8_t *__restrict__ a, size_t n, uint8_t v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}
// a[i] = v for *from <= i < cap (the unused tail of a work list)
__global__ void k_tail_fill_u32(uint32_t *__restrict__ a, const uint32_t *__restrict__ from,
                                uint32_t cap, uint32_t v) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < cap && i >= *from) a[i] = v;
}
// a[i] = v for *from <= i < *to
__global__ void k_range_fill_f32(float *__restrict__ a, const uint32_t *__restrict__ from,
                                 const uint32_t *__restrict__ to, uint32_t cap, float v) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < cap && i >= *from && i < *to) a[i] = v;
}

// Every problem's bounding box lies inside its captured grid (the grid of
// the planning build, widened to hold the problem's box).
__global__ void k_graph_grid_check(const double *__restrict__ bbd,
                                   const GridParams *__restrict__ grids, uint32_t P,
                                   const uint64_t *__restrict__ goff,
                                   unsigned int *__restrict__ viol) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P || goff[p + 1] == goff[p]) return;
  const GridParams &g = grids[p];
  bool bad = false;
  for (int c = 0; c < 3; c++) {
    const double lo = bbd[6 * p + c], hi = bbd[6 * p + 3 + c];
    // (the grid's distance bounds carry the margin)
    const double glo = g.lo[c], ghi = g.lo[c] + (double)g.G[c] * g.h[c];
    if (!(lo >= glo - g.margin) || !(hi <= ghi + g.margin)) bad = true;
    // a flat dimension stays flat
    if (g.h[c] == 0. && !(hi == lo && lo == glo)) bad = true;
  }
  if (bad) atomicOr(viol, GV_GRID);
}

// The list entries of a pass fit what was captured: kofs is the exclusive
// scan of the lengths over `cap` items (zero past the count); otherwise the
// pass's count becomes zero and the violation is recorded.
__global__ void k_graph_entries(const uint32_t *__restrict__ L, const uint64_t *__restrict__ kofs,
                                uint32_t cap, uint64_t ecap, uint32_t *__restrict__ count,
                                unsigned int *__restrict__ viol) {
  const uint64_t total = kofs[cap - 1] + L[cap - 1];
  if (total > ecap) {
    *count = 0u;
    atomicOr(viol, GV_ENTRIES);
  }
}

// The error word, the overflow bits and the kmax flag of a pass.
__global__ void k_graph_pass_check(const unsigned long long *__restrict__ counters,
                                   const unsigned int *__restrict__ cap_err,
                                   unsigned int *__restrict__ viol) {
  if (counters[1] != ~0ull || counters[3] != 0ull || *cap_err != 0u) atomicOr(viol, GV_PASS);
}

// Count of a compaction: *dst = base + the flagged items (the exclusive scan
// pos of flag over cap items); over `limit` (or after any violation) the
// count is zero and, over the limit, `bit` is recorded.
__global__ void k_graph_count(const uint32_t *__restrict__ flag, const uint32_t *__restrict__ pos,
                              uint32_t cap, const uint32_t *__restrict__ base,
                              uint32_t *__restrict__ dst, uint32_t limit, unsigned int bit,
                              unsigned int *__restrict__ viol) {
  const uint64_t t = (uint64_t)(base ? *base : 0u) + (cap ? pos[cap - 1] + flag[cap - 1] : 0u);
  if (t > limit) atomicOr(viol, bit);
  *dst = t > limit || *viol ? 0u : (uint32_t)t;
}

// Compaction of the flagged items to position *base + pos (bounded by
// out_cap); in_r == nullptr: no radius and distance columns.
__global__ void k_take_at(const uint32_t *__restrict__ in_w, const uint32_t *__restrict__ in_L,
                          const double *__restrict__ in_r, const double *__restrict__ in_kd,
                          const uint32_t *__restrict__ flag, const uint32_t *__restrict__ pos,
                          uint32_t n, const uint32_t *__restrict__ base, uint32_t out_cap,
                          uint32_t *__restrict__ out_w, uint32_t *__restrict__ out_L,
                          double *__restrict__ out_r, double *__restrict__ out_kd) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !flag[i]) return;
  const uint64_t o = (uint64_t)(base ? *base : 0u) + pos[i];
  if (o >= out_cap) return;
  out_w[o] = in_w[i];
  out_L[o] = in_L[i];
  if (in_r) {
    out_r[o] = in_r[i];
    out_kd[o] = in_kd[i];
  }
}

// Any item of n in state `want`: violation `bit`.
__global__ void k_graph_any_state(const uint8_t *__restrict__ st, uint32_t n, uint8_t want,
                                  unsigned int bit, unsigned int *__restrict__ viol) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && st[i] == want) atomicOr(viol, bit);
}

// Totals of the final scans (faces, links), zero after any violation so
// that the face placement does nothing with inconsistent counts.
__global__ void k_graph_totals(const uint32_t *__restrict__ cell_nf,
                               const uint64_t *__restrict__ face_off,
                               const uint64_t *__restrict__ cell_fcnt,
                               const uint64_t *__restrict__ cell_foff, size_t n,
                               uint64_t fcap, unsigned int *__restrict__ viol,
                               unsigned long long *__restrict__ dF,
                               unsigned long long *__restrict__ dNL) {
  const uint64_t F = face_off[n - 1] + cell_nf[n - 1];
  const uint64_t NL = cell_foff[n - 1] + cell_fcnt[n - 1];
  if (F > fcap || NL > 2 * fcap) atomicOr(viol, GV_PASS);
  const bool ok = *viol == 0u && F <= fcap && NL <= 2 * fcap;
  *dF = ok ? F : 0ull;
  *dNL = ok ? NL : 0ull;
}

// The status words the host reads after a replay: violation bits, error
// word, faces, links, then the per-problem face and link offsets.
__global__ void k_graph_status(const unsigned int *__restrict__ viol,
                               const unsigned long long *__restrict__ counters,
                               const unsigned long long *__restrict__ dF,
                               const unsigned long long *__restrict__ dNL,
                               const uint64_t *__restrict__ foff,
                               const uint64_t *__restrict__ loff, uint32_t P,
                               unsigned long long *__restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    out[0] = *viol;
    out[1] = counters[1];
    out[2] = *dF;
    out[3] = *dNL;
  }
  if (foff && i <= P) {
    out[4 + i] = foff[i];
    out[5 + P + i] = loff[i];
  }
}

}  // namespace mv
