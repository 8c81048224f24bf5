/*
 * mvoro_oracle.c -- CPU oracle: a plain-C restatement of the per-cell Voronoi
 * construction of the meshless_voronoi crate (v0.7.2, yuyttenhove/meshless_voro).
 *
 * TEST INFRASTRUCTURE ONLY.  The tests use it as the checker and bench.py
 * times it as the CPU baseline ("port").  The product library never links,
 * loads or calls this file.
 *
 * Compile with -ffp-contract=off: the reference is Rust, which never fuses a
 * multiply and an add, and every float formula below is written in the
 * operand order of glam 0.27 (the reference's vector crate):
 *   dot(a,b)   = (a.x*b.x + a.y*b.y) + a.z*b.z
 *   cross(a,b) = (a.y*b.z - b.y*a.z, a.z*b.x - b.z*a.x, a.x*b.y - b.x*a.y)
 *   det(c0,c1,c2) = c2 . (c0 x c1)
 *   project_onto(self, rhs) = (rhs * self.dot(rhs)) * (1 / rhs.dot(rhs))
 * glam itself is not vendored with the reference, so the ulp-level bits of
 * these formulas are "parity unpinned"; the closed forms in the tests pin
 * their values.
 *
 * What each part restates:
 *   geometry primitives ............ src/geometry.rs:29-100
 *   exact in-sphere sign ........... src/geometry.rs:232-314
 *   HalfSpace ...................... src/voronoi/half_space.rs:21-65
 *   SimulationBoundary (+ iloc) .... src/voronoi/boundary.rs:14-64
 *   Generator ...................... src/voronoi/generator.rs:13-24
 *   SimpleCycle .................... src/simple_cycle.rs:1-101
 *   ConvexCell init/build/clip ..... src/voronoi/convex_cell.rs:294-453
 *   decomposition .................. src/voronoi/convex_cell.rs:131-187
 *   VoronoiCell/VoronoiFace ........ src/voronoi/voronoi_cell.rs:40-100,
 *                                    src/voronoi/voronoi_face.rs:11-45,
 *                                    src/voronoi/integrals.rs:86-106
 *   build orchestration + finalize . src/voronoi.rs:192-315
 *   neighbour stream ordering ...... src/rtree_nn.rs:10-36, 124-164
 * The R-tree is replaced by a uniform grid that yields the same ascending
 * key order; among exactly equal keys the order is (idx, image code), since
 * rstar's order among ties is internal to that crate.
 */
#define _GNU_SOURCE
#include "mvoro_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#ifdef ORC_COUNT_FLOPS
static __thread uint64_t t_flops;
#define FL(n) (t_flops += (uint64_t)(n))
#else
#define FL(n) ((void)0)
#endif

/* ------------------------------------------------------------------ */
/* vectors (glam DVec3 semantics)                                      */
/* ------------------------------------------------------------------ */
typedef struct {
  double x, y, z;
} v3;

static inline v3 mk(double x, double y, double z) {
  v3 r;
  r.x = x;
  r.y = y;
  r.z = z;
  return r;
}
static inline v3 vadd(v3 a, v3 b) {
  FL(3);
  return mk(a.x + b.x, a.y + b.y, a.z + b.z);
}
static inline v3 vsub(v3 a, v3 b) {
  FL(3);
  return mk(a.x - b.x, a.y - b.y, a.z - b.z);
}
/* f64 * DVec3 and DVec3 * f64 give the same bits */
static inline v3 vscale(double s, v3 a) {
  FL(3);
  return mk(s * a.x, s * a.y, s * a.z);
}
static inline v3 vdivs(v3 a, double s) {
  FL(3);
  return mk(a.x / s, a.y / s, a.z / s);
}
static inline double vdot(v3 a, v3 b) {
  FL(5);
  return (a.x * b.x + a.y * b.y) + a.z * b.z;
}
static inline v3 vcross(v3 a, v3 b) {
  FL(9);
  return mk(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x,
            a.x * b.y - b.x * a.y);
}
static inline v3 vabs(v3 a) { return mk(fabs(a.x), fabs(a.y), fabs(a.z)); }
static inline double vlen(v3 a) {
  FL(1);
  return sqrt(vdot(a, a));
}
/* Rust f64::signum: +1 for +0, -1 for -0 */
static inline double rsignum(double x) {
  if (isnan(x)) return x;
  return copysign(1.0, x);
}
static inline v3 ld3(const double *p) { return mk(p[0], p[1], p[2]); }
static inline void st3(double *p, v3 a) {
  p[0] = a.x;
  p[1] = a.y;
  p[2] = a.z;
}

/* ------------------------------------------------------------------ */
/* geometry.rs                                                         */
/* ------------------------------------------------------------------ */

/* Plane::project_onto, geometry.rs:36-38 (glam DVec3::project_onto) */
static inline v3 plane_project_onto(v3 n, v3 p, v3 point) {
  v3 diff = vsub(p, point);
  FL(1);
  double rcp = 1.0 / vdot(n, n);
  double s = vdot(diff, n);
  v3 proj = vscale(rcp, vscale(s, n));
  return vadd(point, proj);
}

/* intersect_planes, geometry.rs:68-76.  Returns 0 if det == 0. */
static inline int intersect_planes(v3 n0, double d0, v3 n1, double d1, v3 n2,
                                   double d2, v3 *out) {
  v3 c01 = vcross(n0, n1);
  double det = vdot(n2, c01);
  if (det == 0.0) return 0;
  v3 a = vscale(d0, vcross(n1, n2));
  v3 b = vscale(d1, vcross(n2, n0));
  v3 c = vscale(d2, c01);
  *out = vdivs(vadd(vadd(a, b), c), det);
  return 1;
}

/* Plane::project_onto_intersection, geometry.rs:41-47: self = `a`,
 * other = `b`; the third plane passes through `point` with normal a x b */
static inline int project_onto_intersection(v3 na, double da, v3 nb,
                                            double db, v3 point, v3 *out) {
  v3 nperp = vcross(na, nb);
  double dperp = vdot(point, nperp);
  return intersect_planes(na, da, nb, db, nperp, dperp, out);
}

/* signed_volume_tet, geometry.rs:82-88 */
static inline double signed_volume_tet(v3 v0, v3 v1, v3 v2, v3 v3_) {
  v3 v01 = vsub(v1, v0), v02 = vsub(v2, v0), v03 = vsub(v3_, v0);
  FL(1);
  return vdot(v03, vcross(v01, v02)) / 6.0;
}

/* signed_area_tri, geometry.rs:95-100 */
static inline double signed_area_tri(v3 v0, v3 v1, v3 v2, v3 t) {
  v3 n = vscale(0.5, vcross(vsub(v1, v0), vsub(v2, v0)));
  double sign = rsignum(vdot(vsub(t, v0), n));
  FL(1);
  return vlen(n) * sign;
}

/* ---- exact in-sphere sign, geometry.rs:232-314 --------------------- */
/* 384-bit two's complement integers with 32-bit limbs; the determinant
 * of snapped 52-bit coordinates is below 2^267 in magnitude, so every
 * intermediate is exact. */
#define BL 12
typedef struct {
  uint32_t w[BL];
} bi;

static bi bi_i64(int64_t v) {
  bi r;
  uint64_t u = (uint64_t)v;
  r.w[0] = (uint32_t)u;
  r.w[1] = (uint32_t)(u >> 32);
  uint32_t ext = v < 0 ? 0xffffffffu : 0u;
  for (int i = 2; i < BL; i++) r.w[i] = ext;
  return r;
}
static bi bi_add(bi a, bi b) {
  bi r;
  uint64_t c = 0;
  for (int i = 0; i < BL; i++) {
    uint64_t s = (uint64_t)a.w[i] + b.w[i] + c;
    r.w[i] = (uint32_t)s;
    c = s >> 32;
  }
  return r;
}
static bi bi_neg(bi a) {
  bi one = bi_i64(1);
  for (int i = 0; i < BL; i++) a.w[i] = ~a.w[i];
  return bi_add(a, one);
}
static bi bi_sub(bi a, bi b) { return bi_add(a, bi_neg(b)); }
static bi bi_mul(bi a, bi b) {
  bi r;
  memset(&r, 0, sizeof r);
  for (int i = 0; i < BL; i++) {
    uint64_t carry = 0;
    for (int j = 0; i + j < BL; j++) {
      uint64_t t = (uint64_t)a.w[i] * b.w[j] + r.w[i + j] + carry;
      r.w[i + j] = (uint32_t)t;
      carry = t >> 32;
    }
  }
  return r;
}
static int bi_sign(bi a) {
  if (a.w[BL - 1] & 0x80000000u) return -1;
  for (int i = 0; i < BL; i++)
    if (a.w[i]) return 1;
  return 0;
}
/* big_int_det2x2!: a*d - b*c */
static bi bi_det2(bi a, bi b, bi c, bi d) {
  return bi_sub(bi_mul(a, d), bi_mul(b, c));
}
/* big_int_det3x3! (geometry.rs:257-267) */
static bi bi_det3(bi a0, bi a1, bi a2, bi b0, bi b1, bi b2, bi c0, bi c1,
                  bi c2) {
  bi det = bi_mul(a0, bi_det2(b1, b2, c1, c2));
  det = bi_sub(det, bi_mul(a1, bi_det2(b0, b2, c0, c2)));
  det = bi_add(det, bi_mul(a2, bi_det2(b0, b1, c0, c1)));
  return det;
}
/* big_int!: (x - a) per component, plus the squared norm */
static void bi_lift(const int64_t *x, const int64_t *a, bi out[4]) {
  for (int i = 0; i < 3; i++) out[i] = bi_i64(x[i] - a[i]);
  bi n2 = bi_mul(out[0], out[0]);
  n2 = bi_add(n2, bi_mul(out[1], out[1]));
  n2 = bi_add(n2, bi_mul(out[2], out[2]));
  out[3] = n2;
}

int orc_in_sphere_exact(const int64_t a[3], const int64_t b[3],
                        const int64_t c[3], const int64_t d[3],
                        const int64_t v[3]) {
  bi B[4], C[4], D[4], V[4];
  bi_lift(b, a, B);
  bi_lift(c, a, C);
  bi_lift(d, a, D);
  bi_lift(v, a, V);
  bi det;
  memset(&det, 0, sizeof det);
  bi t;
  t = bi_det3(B[1], C[1], D[1], B[2], C[2], D[2], B[3], C[3], D[3]);
  det = bi_sub(det, bi_mul(V[0], t));
  t = bi_det3(B[0], C[0], D[0], B[2], C[2], D[2], B[3], C[3], D[3]);
  det = bi_add(det, bi_mul(V[1], t));
  t = bi_det3(B[0], C[0], D[0], B[1], C[1], D[1], B[3], C[3], D[3]);
  det = bi_sub(det, bi_mul(V[2], t));
  t = bi_det3(B[0], C[0], D[0], B[1], C[1], D[1], B[2], C[2], D[2]);
  det = bi_add(det, bi_mul(V[3], t));
  return bi_sign(det);
}

/* ------------------------------------------------------------------ */
/* boundary.rs, half_space.rs, generator.rs                            */
/* ------------------------------------------------------------------ */
typedef struct {
  v3 n, p;
  double d, errb;
  int64_t right; /* -1: None (boundary plane) */
  int has_shift;
  v3 shift;
} hspace;

/* HalfSpace::new, half_space.rs:24-34 */
static inline hspace hs_new(v3 n, v3 p, int64_t right, int has_shift,
                            v3 shift) {
  hspace h;
  FL(2);
  h.errb = 1e-13 * (1. + vdot(vabs(n), vabs(p)));
  h.n = n;
  h.p = p;
  h.d = vdot(n, p);
  h.right = right;
  h.has_shift = has_shift;
  h.shift = shift;
  return h;
}

/* HalfSpace::clip, half_space.rs:40-47 */
static inline double hs_clip(const hspace *h, v3 v) {
  FL(1);
  double c = vdot(h->n, v) - h->d;
  if (fabs(c) < h->errb) return 0.;
  return rsignum(c);
}

typedef struct {
  v3 anchor, inv_width; /* snap frame (boundary.rs:41-46) */
  int dim;
  hspace planes[6];
} sim_boundary;

/* SimulationBoundary::cuboid, boundary.rs:14-47 */
static void boundary_cuboid(sim_boundary *sb, v3 anchor, v3 width,
                            int periodic, int dim) {
  if (periodic) {
    anchor.x -= width.x;
    width.x *= 3.;
    if (dim >= 2) {
      anchor.y -= width.y;
      width.y *= 3.;
    }
    if (dim == 3) {
      anchor.z -= width.z;
      width.z *= 3.;
    }
  }
  v3 aw = vadd(anchor, width);
  v3 zero = mk(0, 0, 0);
  sb->planes[0] = hs_new(mk(1, 0, 0), anchor, -1, 0, zero);
  sb->planes[1] = hs_new(mk(-1, 0, 0), aw, -1, 0, zero);
  sb->planes[2] = hs_new(mk(0, 1, 0), anchor, -1, 0, zero);
  sb->planes[3] = hs_new(mk(0, -1, 0), aw, -1, 0, zero);
  sb->planes[4] = hs_new(mk(0, 0, 1), anchor, -1, 0, zero);
  sb->planes[5] = hs_new(mk(0, 0, -1), aw, -1, 0, zero);
  sb->anchor = vsub(anchor, width);
  sb->inv_width = mk(1. / (3. * width.x), 1. / (3. * width.y),
                     1. / (3. * width.z));
  sb->dim = dim;
}

/* SimulationBoundary::iloc, boundary.rs:49-64 */
static inline void boundary_iloc(const sim_boundary *sb, v3 loc,
                                 int64_t out[3]) {
  double t[3];
  t[0] = 1. + (loc.x - sb->anchor.x) * sb->inv_width.x;
  t[1] = 1. + (loc.y - sb->anchor.y) * sb->inv_width.y;
  t[2] = 1. + (loc.z - sb->anchor.z) * sb->inv_width.z;
  for (int i = 0; i < 3; i++) {
    uint64_t bits;
    memcpy(&bits, &t[i], 8);
    out[i] = (int64_t)(bits & 0xFFFFFFFFFFFFFull);
  }
}

/* Generator::new, generator.rs:13-24 */
static inline v3 generator_loc(const double *xyz, size_t i, int dim) {
  v3 g = ld3(xyz + 3 * i);
  if (dim == 1) {
    g.y = 0.;
    g.z = 0.;
  } else if (dim == 2) {
    g.z = 0.;
  }
  return g;
}

/* ------------------------------------------------------------------ */
/* neighbour stream: uniform grid, ascending (key, idx, image)          */
/* ------------------------------------------------------------------ */
typedef struct {
  double key;
  uint64_t idx;
  int32_t img; /* (i+1)*9 + (j+1)*3 + (k+1), 13 = no shift */
} cand;

static int cand_cmp(const void *pa, const void *pb) {
  const cand *a = (const cand *)pa, *b = (const cand *)pb;
  if (a->key < b->key) return -1;
  if (a->key > b->key) return 1;
  if (a->idx < b->idx) return -1;
  if (a->idx > b->idx) return 1;
  return (a->img > b->img) - (a->img < b->img);
}
static inline int cand_gt(const cand *a, const cand *b) {
  return cand_cmp(a, b) > 0;
}

typedef struct {
  const double *gen; /* 3N, dimension-zeroed generators */
  size_t n;
  double lo[3], inv_h[3];
  int G[3];
  uint32_t *cell_start; /* G0*G1*G2 + 1 */
  uint32_t *items;      /* N, generator indices sorted by cell */
  double h_min;         /* smallest non-degenerate cell extent */
  int nimg_lo[3], nimg_hi[3];
  double width[3];
  int periodic;
  size_t total_stream; /* N * number of images */
} grid_t;

static int grid_coord(const grid_t *g, int c, double x) {
  double f = (x - g->lo[c]) * g->inv_h[c];
  if (!(f >= 0.)) return 0; /* also NaN */
  if (f >= (double)g->G[c]) return g->G[c] - 1;
  return (int)f;
}

static int grid_build(grid_t *g, const double *gen, size_t n,
                      const double width[3], int dim, int periodic) {
  memset(g, 0, sizeof *g);
  g->gen = gen;
  g->n = n;
  double hi[3];
  for (int c = 0; c < 3; c++) {
    g->lo[c] = INFINITY;
    hi[c] = -INFINITY;
  }
  for (size_t i = 0; i < n; i++)
    for (int c = 0; c < 3; c++) {
      double x = gen[3 * i + c];
      if (x < g->lo[c]) g->lo[c] = x;
      if (x > hi[c]) hi[c] = x;
    }
  double ext[3], vol = 1.;
  int nd = 0;
  for (int c = 0; c < 3; c++) {
    ext[c] = hi[c] - g->lo[c];
    if (ext[c] > 0) {
      vol *= ext[c];
      nd++;
    }
  }
  double h = 0.;
  if (nd > 0) h = pow(vol * 2.0 / (double)n, 1.0 / nd);
  g->h_min = INFINITY;
  for (int c = 0; c < 3; c++) {
    int G = 1;
    if (ext[c] > 0 && h > 0) {
      double gg = ceil(ext[c] / h);
      G = gg < 1 ? 1 : (gg > 512 ? 512 : (int)gg);
    }
    g->G[c] = G;
    g->inv_h[c] = ext[c] > 0 ? (double)G / ext[c] : 0.;
    if (ext[c] > 0 && ext[c] / G < g->h_min) g->h_min = ext[c] / G;
    g->width[c] = width[c];
  }
  if (!(g->h_min < INFINITY)) g->h_min = 1.;
  size_t ncell = (size_t)g->G[0] * g->G[1] * g->G[2];
  g->cell_start = (uint32_t *)calloc(ncell + 1, sizeof(uint32_t));
  g->items = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
  uint32_t *cid = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
  if (!g->cell_start || !g->items || !cid) {
    free(cid);
    return ORC_E_OOM;
  }
  for (size_t i = 0; i < n; i++) {
    int a = grid_coord(g, 0, gen[3 * i]);
    int b = grid_coord(g, 1, gen[3 * i + 1]);
    int c = grid_coord(g, 2, gen[3 * i + 2]);
    cid[i] = (uint32_t)(((size_t)a * g->G[1] + b) * g->G[2] + c);
    g->cell_start[cid[i] + 1]++;
  }
  for (size_t c = 0; c < ncell; c++) g->cell_start[c + 1] += g->cell_start[c];
  uint32_t *fill = (uint32_t *)malloc((ncell ? ncell : 1) * sizeof(uint32_t));
  memcpy(fill, g->cell_start, ncell * sizeof(uint32_t));
  for (size_t i = 0; i < n; i++) g->items[fill[cid[i]]++] = (uint32_t)i;
  free(fill);
  free(cid);
  /* image ranges: rtree_nn.rs:75-90 */
  g->periodic = periodic;
  for (int c = 0; c < 3; c++) {
    int active = periodic && (c == 0 || (c == 1 && dim >= 2) ||
                              (c == 2 && dim == 3));
    g->nimg_lo[c] = active ? -1 : 0;
    g->nimg_hi[c] = active ? 1 : 0;
  }
  size_t nimg = (size_t)(g->nimg_hi[0] - g->nimg_lo[0] + 1) *
                (g->nimg_hi[1] - g->nimg_lo[1] + 1) *
                (g->nimg_hi[2] - g->nimg_lo[2] + 1);
  g->total_stream = n * nimg;
  return ORC_OK;
}

static void grid_free(grid_t *g) {
  free(g->cell_start);
  free(g->items);
}

typedef struct {
  cand *buf;
  size_t cap, len;
} cand_vec;

static int cv_push(cand_vec *v, cand c) {
  if (v->len == v->cap) {
    size_t nc = v->cap ? 2 * v->cap : 256;
    cand *nb = (cand *)realloc(v->buf, nc * sizeof(cand));
    if (!nb) return 0;
    v->buf = nb;
    v->cap = nc;
  }
  v->buf[v->len++] = c;
  return 1;
}

/* All stream entries with key <= R2.  The key is the reference's
 * wrapping_distance_2 (rtree_nn.rs:138-146): dx = (q + s) - g summed as
 * (dx0^2 + dx1^2) + dx2^2, s = image * width; for non-periodic builds the
 * reference's distance_squared (rtree_nn.rs:124-134) gives the same bits. */
static int gather(const grid_t *g, v3 q, double R, cand_vec *out) {
  double R2 = R * R;
  out->len = 0;
  for (int i = g->nimg_lo[0]; i <= g->nimg_hi[0]; i++)
    for (int j = g->nimg_lo[1]; j <= g->nimg_hi[1]; j++)
      for (int k = g->nimg_lo[2]; k <= g->nimg_hi[2]; k++) {
        double s[3] = {(double)i * g->width[0], (double)j * g->width[1],
                       (double)k * g->width[2]};
        double qs[3] = {q.x + s[0], q.y + s[1], q.z + s[2]};
        int lo[3], hi[3], skip = 0;
        for (int c = 0; c < 3; c++) {
          double rr = R * (1. + 1e-9) + 1e-300;
          double a = (qs[c] - rr - g->lo[c]) * g->inv_h[c];
          double b = (qs[c] + rr - g->lo[c]) * g->inv_h[c];
          if (g->inv_h[c] == 0.) {
            a = 0.;
            b = 0.;
          }
          if (b < -1. || a > (double)g->G[c] + 1.) {
            skip = 1;
            break;
          }
          double fa = floor(a) - 1., fb = floor(b) + 1.;
          lo[c] = fa < 0. ? 0 : (fa > g->G[c] - 1 ? g->G[c] - 1 : (int)fa);
          hi[c] = fb < 0. ? 0 : (fb > g->G[c] - 1 ? g->G[c] - 1 : (int)fb);
        }
        if (skip) continue;
        int32_t img = (i + 1) * 9 + (j + 1) * 3 + (k + 1);
        for (int a = lo[0]; a <= hi[0]; a++)
          for (int b = lo[1]; b <= hi[1]; b++)
            for (int c = lo[2]; c <= hi[2]; c++) {
              size_t cell = ((size_t)a * g->G[1] + b) * g->G[2] + c;
              for (uint32_t t = g->cell_start[cell];
                   t < g->cell_start[cell + 1]; t++) {
                uint32_t gi = g->items[t];
                const double *gp = g->gen + 3 * (size_t)gi;
                double d0 = qs[0] - gp[0], d1 = qs[1] - gp[1],
                       d2 = qs[2] - gp[2];
                double key = (d0 * d0 + d1 * d1) + d2 * d2;
                if (key <= R2) {
                  cand cc;
                  cc.key = key;
                  cc.idx = gi;
                  cc.img = img;
                  if (!cv_push(out, cc)) return ORC_E_OOM;
                }
              }
            }
      }
  qsort(out->buf, out->len, sizeof(cand), cand_cmp);
  return ORC_OK;
}

/* decode image code into the reported shift (rtree_nn.rs:27-33):
 * None if the image is (0,0,0), else Some(-(image * width)) */
static inline int image_shift(int32_t img, const double width[3], v3 *shift) {
  int i = img / 9 - 1, j = (img / 3) % 3 - 1, k = img % 3 - 1;
  if (i == 0 && j == 0 && k == 0) {
    *shift = mk(0, 0, 0);
    return 0;
  }
  *shift = mk(-((double)i * width[0]), -((double)j * width[1]),
              -((double)k * width[2]));
  return 1;
}

/* ------------------------------------------------------------------ */
/* SimpleCycle, simple_cycle.rs:1-101                                  */
/* ------------------------------------------------------------------ */
typedef struct {
  int32_t *ptrs;
  int cap, n;
  int start, len;
} scycle;

static int sc_reserve(scycle *s, int n) {
  if (n > s->cap) {
    int nc = s->cap ? s->cap : 64;
    while (nc < n) nc *= 2;
    int32_t *np = (int32_t *)realloc(s->ptrs, nc * sizeof(int32_t));
    if (!np) return 0;
    s->ptrs = np;
    s->cap = nc;
  }
  return 1;
}
/* SimpleCycle::new */
static int sc_new(scycle *s, int capacity) {
  if (!sc_reserve(s, capacity)) return 0;
  for (int i = 0; i < capacity; i++) s->ptrs[i] = i;
  s->n = capacity;
  s->start = 0;
  s->len = 0;
  return 1;
}
/* SimpleCycle::grow */
static int sc_grow(scycle *s) {
  if (!sc_reserve(s, s->n + 1)) return 0;
  s->ptrs[s->n] = s->n;
  s->n++;
  return 1;
}
/* SimpleCycle::init */
static void sc_init(scycle *s, int a, int b, int c) {
  int cur = s->start;
  for (int i = 0; i < s->len; i++) {
    int nx = s->ptrs[cur];
    s->ptrs[cur] = cur;
    cur = nx;
  }
  s->len = 3;
  s->start = a;
  s->ptrs[a] = b;
  s->ptrs[b] = c;
  s->ptrs[c] = a;
}
/* SimpleCycle::try_extend: 1 = Ok, 0 = Err */
static int sc_try_extend(scycle *s, int a, int b, int c) {
  int tri[3] = {a, b, c};
  int con[3] = {s->ptrs[a] != a, s->ptrs[b] != b, s->ptrs[c] != c};
  for (int i = 0; i < 3; i++) {
    int j = (i + 1) % 3, k = (i + 2) % 3;
    if (!con[i] && con[j] && con[k] && s->ptrs[tri[k]] == tri[j]) {
      s->ptrs[tri[k]] = tri[i];
      s->ptrs[tri[i]] = tri[j];
      s->len += 1;
      return 1;
    } else if (con[i] && con[j] && con[k] && s->ptrs[tri[k]] == tri[j] &&
               s->ptrs[tri[j]] == tri[i]) {
      s->ptrs[tri[k]] = tri[i];
      s->ptrs[tri[j]] = tri[j];
      if (s->start == tri[j]) s->start = tri[i];
      s->len -= 1;
      return 1;
    }
  }
  return 0;
}

int orc_simple_cycle(const int32_t *tris_in, int n_tris, int capacity,
                     int32_t *out) {
  scycle s;
  memset(&s, 0, sizeof s);
  int32_t *tris = (int32_t *)malloc(3 * (size_t)n_tris * sizeof(int32_t));
  memcpy(tris, tris_in, 3 * (size_t)n_tris * sizeof(int32_t));
  sc_new(&s, capacity);
  sc_init(&s, tris[0], tris[1], tris[2]);
  int ok = 1;
  for (int i = 1; i < n_tris && ok; i++) {
    int idx = i;
    for (;;) {
      if (idx >= n_tris) {
        ok = 0;
        break;
      }
      if (sc_try_extend(&s, tris[3 * idx], tris[3 * idx + 1],
                        tris[3 * idx + 2])) {
        if (idx > i)
          for (int c = 0; c < 3; c++) {
            int32_t t = tris[3 * i + c];
            tris[3 * i + c] = tris[3 * idx + c];
            tris[3 * idx + c] = t;
          }
        break;
      }
      idx++;
    }
  }
  int len = -1;
  if (ok) {
    len = s.len;
    int cur = s.start;
    for (int i = 0; i < s.len; i++) {
      out[i] = cur;
      cur = s.ptrs[cur];
    }
  }
  free(tris);
  free(s.ptrs);
  return len;
}

/* ------------------------------------------------------------------ */
/* ConvexCell, convex_cell.rs:14-453                                   */
/* ------------------------------------------------------------------ */
typedef struct {
  v3 loc;
  int32_t dual[3];
  double r2;
} vert;

typedef struct {
  int64_t idx;
  v3 loc;
  hspace *planes;
  int np, cap_p;
  vert *verts;
  int nv, cap_v;
  scycle cyc;
  double safety_radius;
  int dim;
} ccell;

static int cc_push_plane(ccell *c, hspace h) {
  if (c->np == c->cap_p) {
    int nc = c->cap_p ? 2 * c->cap_p : 64;
    hspace *nb = (hspace *)realloc(c->planes, nc * sizeof(hspace));
    if (!nb) return 0;
    c->planes = nb;
    c->cap_p = nc;
  }
  c->planes[c->np++] = h;
  return 1;
}
static int cc_reserve_verts(ccell *c, int n) {
  if (n > c->cap_v) {
    int nc = c->cap_v ? c->cap_v : 128;
    while (nc < n) nc *= 2;
    vert *nb = (vert *)realloc(c->verts, nc * sizeof(vert));
    if (!nb) return 0;
    c->verts = nb;
    c->cap_v = nc;
  }
  return 1;
}

/* Vertex::from_dual, convex_cell.rs:28-48 */
static int vertex_from_dual(const ccell *c, int i, int j, int k, vert *out) {
  const hspace *a = &c->planes[i], *b = &c->planes[j], *d = &c->planes[k];
  v3 loc;
  if (!intersect_planes(a->n, a->d, b->n, b->d, d->n, d->d, &loc)) return 0;
  v3 dl = loc;
  if (c->dim == 1) {
    dl.y = 0.;
    dl.z = 0.;
  } else if (c->dim == 2) {
    dl.z = 0.;
  }
  v3 df = vsub(c->loc, dl);
  out->loc = loc;
  out->dual[0] = i;
  out->dual[1] = j;
  out->dual[2] = k;
  out->r2 = vdot(df, df);
  return 1;
}

/* update_safety_radius, convex_cell.rs:443-453 */
static int cc_update_safety_radius(ccell *c) {
  double m = c->verts[0].r2;
  if (isnan(m)) return ORC_E_NAN;
  for (int i = 1; i < c->nv; i++) {
    double r = c->verts[i].r2;
    if (isnan(r)) return ORC_E_NAN;
    if (r >= m) m = r;
  }
  FL(2);
  c->safety_radius = 2. * sqrt(m);
  return ORC_OK;
}

/* ConvexCell::init, convex_cell.rs:294-312 */
static int cc_init(ccell *c, v3 loc, int64_t idx, const sim_boundary *sb) {
  static const int duals[8][3] = {{2, 5, 0}, {5, 3, 0}, {1, 5, 2},
                                  {5, 1, 3}, {4, 2, 0}, {4, 0, 3},
                                  {2, 4, 1}, {4, 3, 1}};
  c->idx = idx;
  c->loc = loc;
  c->dim = sb->dim;
  c->np = 0;
  for (int i = 0; i < 6; i++)
    if (!cc_push_plane(c, sb->planes[i])) return ORC_E_OOM;
  if (!cc_reserve_verts(c, 8)) return ORC_E_OOM;
  c->nv = 8;
  for (int i = 0; i < 8; i++)
    if (!vertex_from_dual(c, duals[i][0], duals[i][1], duals[i][2],
                          &c->verts[i]))
      return ORC_E_PLANE_DET_ZERO;
  if (!sc_new(&c->cyc, 6)) return ORC_E_OOM;
  c->safety_radius = 0.;
  return cc_update_safety_radius(c);
}

/* HalfSpace::right_loc, half_space.rs:53-65 */
static inline v3 hs_right_loc(const hspace *h, int64_t left,
                              const double *gen) {
  if (h->right >= 0) {
    v3 l = ld3(gen + 3 * (size_t)h->right);
    if (h->has_shift) l = vadd(l, h->shift);
    return l;
  }
  v3 left_loc = ld3(gen + 3 * (size_t)left);
  v3 pr = plane_project_onto(h->n, h->p, left_loc);
  return vsub(vscale(2., pr), left_loc);
}

/* compute_boundary, convex_cell.rs:421-441 */
static int cc_compute_boundary(scycle *s, vert *v, int n) {
  sc_init(s, v[0].dual[0], v[0].dual[1], v[0].dual[2]);
  for (int i = 1; i < n; i++) {
    int idx = i;
    for (;;) {
      if (idx >= n) return ORC_E_BOUNDARY_WALK;
      if (sc_try_extend(s, v[idx].dual[0], v[idx].dual[1], v[idx].dual[2])) {
        if (idx > i) {
          vert t = v[i];
          v[i] = v[idx];
          v[idx] = t;
        }
        break;
      }
      idx++;
    }
  }
  return ORC_OK;
}

typedef struct {
  const double *gen;
  const sim_boundary *sb;
  uint64_t exact_tests;
} clip_ctx;

/* clip_by_plane, convex_cell.rs:358-419 */
static int cc_clip_by_plane(ccell *c, hspace p, clip_ctx *ctx) {
  int i = 0, num_v = c->nv, num_r = 0;
  while (i < num_v) {
    double clip = hs_clip(&p, c->verts[i].loc);
    if (clip == 0.) {
      const int32_t *dual = c->verts[i].dual;
      int64_t a[3], b[3], cc[3], d[3], v[3];
      boundary_iloc(ctx->sb, c->loc, a);
      boundary_iloc(ctx->sb, hs_right_loc(&c->planes[dual[0]], c->idx, ctx->gen), b);
      boundary_iloc(ctx->sb, hs_right_loc(&c->planes[dual[1]], c->idx, ctx->gen), cc);
      boundary_iloc(ctx->sb, hs_right_loc(&c->planes[dual[2]], c->idx, ctx->gen), d);
      boundary_iloc(ctx->sb, hs_right_loc(&p, c->idx, ctx->gen), v);
      clip = (double)orc_in_sphere_exact(a, b, cc, d, v);
      ctx->exact_tests++;
    }
    if (clip < 0.) {
      num_v -= 1;
      num_r += 1;
      vert t = c->verts[i];
      c->verts[i] = c->verts[num_v];
      c->verts[num_v] = t;
    } else {
      i += 1;
    }
  }
  if (num_r > 0) {
    int p_idx = c->np;
    if (!cc_push_plane(c, p)) return ORC_E_OOM;
    if (!sc_grow(&c->cyc)) return ORC_E_OOM;
    int st = cc_compute_boundary(&c->cyc, c->verts + num_v, c->nv - num_v);
    if (st) return st;
    c->nv = num_v;
    int len = c->cyc.len;
    if (!cc_reserve_verts(c, num_v + len)) return ORC_E_OOM;
    int cur = c->cyc.start;
    for (int t = 0; t < len; t++) {
      int next = c->cyc.ptrs[cur];
      if (!vertex_from_dual(c, cur, next, p_idx, &c->verts[c->nv]))
        return ORC_E_PLANE_DET_ZERO;
      c->nv++;
      cur = next;
    }
    return cc_update_safety_radius(c);
  }
  return ORC_OK;
}

/* per-thread workspace */
typedef struct {
  ccell cell;
  cand_vec cands;
  /* faces produced by this thread */
  size_t fcap, flen;
  double *f_area, *f_cen, *f_norm, *f_shift;
  int64_t *f_right;
  uint8_t *f_has_shift;
  /* per-face plane index scratch */
  int32_t *maybe; /* per plane: face slot or -1 */
  int maybe_cap;
  double *facc; /* per plane accumulators: area, cx, cy, cz */
  uint64_t exact_tests;
} workspace;

typedef struct {
  int status;
  uint32_t n_popped;
  int exhausted;
} build_info;

/* ConvexCell::build, convex_cell.rs:316-356, with the neighbour stream of
 * voronoi.rs:290-294 */
static int cc_build(workspace *ws, const grid_t *g, const sim_boundary *sb,
                    const double *gen, size_t idx, build_info *info) {
  ccell *c = &ws->cell;
  v3 loc = ld3(gen + 3 * idx);
  int st = cc_init(c, loc, (int64_t)idx, sb);
  info->n_popped = 0;
  info->exhausted = 0;
  if (st) return st;
  clip_ctx ctx = {gen, sb, 0};
  double R = 2.0 * g->h_min;
  if (!(R > 0.)) R = 1.;
  cand last = {0., 0, 0};
  int have_last = 0;
  for (;;) {
    st = gather(g, loc, R, &ws->cands);
    if (st) return st;
    size_t start = 0;
    if (have_last)
      while (start < ws->cands.len && !cand_gt(&ws->cands.buf[start], &last))
        start++;
    for (size_t t = start; t < ws->cands.len; t++) {
      cand e = ws->cands.buf[t];
      last = e;
      have_last = 1;
      info->n_popped++;
      if (info->n_popped == 1) {
        if (e.idx != idx) return ORC_E_SELF_NOT_FIRST;
        continue;
      }
      v3 shift;
      int has_shift = image_shift(e.img, g->width, &shift);
      v3 ngb = ld3(gen + 3 * e.idx);
      if (has_shift) ngb = vadd(ngb, shift);
      v3 dx = vsub(c->loc, ngb);
      double dist = vlen(dx);
      if (!(isfinite(dist) && dist > 0.0)) return ORC_E_DEGENERATE_POINTS;
      if (c->safety_radius < dist) {
        ws->exact_tests += ctx.exact_tests;
        return ORC_OK;
      }
      v3 n = vdivs(dx, dist);
      FL(1);
      v3 p = vscale(0.5, vadd(c->loc, ngb));
      st = cc_clip_by_plane(c, hs_new(n, p, (int64_t)e.idx, has_shift, shift),
                            &ctx);
      if (st) return st;
    }
    if (ws->cands.len >= g->total_stream) {
      info->exhausted = 1;
      ws->exact_tests += ctx.exact_tests;
      return ORC_OK;
    }
    R *= 2.;
  }
}

/* Plane index -> should a face be built (voronoi_cell.rs:53-73) */
static inline int should_construct(const ccell *c, int pi,
                                   const uint8_t *mask) {
  const hspace *h = &c->planes[pi];
  int valid = 1;
  if (c->dim == 1)
    valid = h->n.y == 0. && h->n.z == 0.;
  else if (c->dim == 2)
    valid = h->n.z == 0.;
  if (!valid) return 0;
  if (h->right >= 0 && !h->has_shift)
    return h->right > c->idx || (mask && !mask[h->right]);
  return 1;
}

/* DecompositionWithoutFaces::load_vertex, convex_cell.rs:150-159 */
static int load_projections(const ccell *c, const vert *v, v3 proj[6]) {
  for (int i = 0; i < 3; i++) {
    const hspace *cur = &c->planes[v->dual[i]];
    const hspace *next = &c->planes[v->dual[(i + 1) % 3]];
    proj[2 * i] = plane_project_onto(cur->n, cur->p, c->loc);
    if (!project_onto_intersection(next->n, next->d, cur->n, cur->d, c->loc,
                                   &proj[2 * i + 1]))
      return 0;
  }
  return 1;
}

typedef struct {
  v3 centroid;
  double volume;
} cell_out;

/* VoronoiCell::from_convex_cell, voronoi_cell.rs:40-100.  Appends the
 * faces to the workspace; returns the number of faces. */
static int cell_integrate(workspace *ws, const uint8_t *mask, cell_out *co,
                          int *nfaces) {
  ccell *c = &ws->cell;
  if (c->np > ws->maybe_cap) {
    int nc = ws->maybe_cap ? ws->maybe_cap : 64;
    while (nc < c->np) nc *= 2;
    ws->maybe = (int32_t *)realloc(ws->maybe, nc * sizeof(int32_t));
    ws->facc = (double *)realloc(ws->facc, 4 * (size_t)nc * sizeof(double));
    if (!ws->maybe || !ws->facc) return ORC_E_OOM;
    ws->maybe_cap = nc;
  }
  for (int i = 0; i < c->np; i++) ws->maybe[i] = -1;
  double vol = 0.;
  v3 cen = mk(0, 0, 0);
  v3 loc = c->loc;
  for (int vi = 0; vi < c->nv; vi++) {
    const vert *v = &c->verts[vi];
    v3 proj[6];
    if (!load_projections(c, v, proj)) return ORC_E_PLANE_DET_ZERO;
    for (int t = 0; t < 6; t++) {
      v3 v0 = proj[t], v1 = proj[(t + 5) % 6], v2 = v->loc;
      int pi = v->dual[t / 2];
      /* VolumeCentroidIntegral::collect, integrals.rs:91-95 */
      double tv = signed_volume_tet(v0, v1, v2, loc);
      FL(1);
      vol += tv;
      cen = vadd(cen, vscale(tv, vadd(vadd(vadd(v0, v1), v2), loc)));
      if (ws->maybe[pi] == -1)
        ws->maybe[pi] = should_construct(c, pi, mask) ? 1 : -2;
      if (ws->maybe[pi] == -2) continue;
      if (ws->maybe[pi] == 1) {
        ws->maybe[pi] = 0;
        ws->facc[4 * pi] = 0.;
        ws->facc[4 * pi + 1] = 0.;
        ws->facc[4 * pi + 2] = 0.;
        ws->facc[4 * pi + 3] = 0.;
      }
      /* VoronoiFaceIntegral::collect, voronoi_face.rs:30-34 */
      double a = signed_area_tri(v0, v1, v2, loc);
      v3 s = vscale(a, vadd(vadd(v0, v1), v2));
      FL(4);
      ws->facc[4 * pi] += a;
      ws->facc[4 * pi + 1] += s.x;
      ws->facc[4 * pi + 2] += s.y;
      ws->facc[4 * pi + 3] += s.z;
    }
  }
  /* VolumeCentroidIntegral::finalize, integrals.rs:97-105 */
  double norm = vol > 0. ? 0.25 / vol : 0.;
  co->volume = vol;
  co->centroid = vscale(norm, cen);
  int nf = 0;
  for (int pi = 0; pi < c->np; pi++) {
    if (ws->maybe[pi] != 0) continue;
    if (ws->flen == ws->fcap) {
      size_t nc = ws->fcap ? 2 * ws->fcap : 4096;
      ws->f_area = (double *)realloc(ws->f_area, nc * sizeof(double));
      ws->f_cen = (double *)realloc(ws->f_cen, 3 * nc * sizeof(double));
      ws->f_norm = (double *)realloc(ws->f_norm, 3 * nc * sizeof(double));
      ws->f_shift = (double *)realloc(ws->f_shift, 3 * nc * sizeof(double));
      ws->f_right = (int64_t *)realloc(ws->f_right, nc * sizeof(int64_t));
      ws->f_has_shift = (uint8_t *)realloc(ws->f_has_shift, nc);
      if (!ws->f_area || !ws->f_cen || !ws->f_norm || !ws->f_shift ||
          !ws->f_right || !ws->f_has_shift)
        return ORC_E_OOM;
      ws->fcap = nc;
    }
    size_t f = ws->flen++;
    const hspace *h = &c->planes[pi];
    double area = ws->facc[4 * pi];
    /* VoronoiFaceIntegral::finalize, voronoi_face.rs:36-43 */
    FL(2);
    double fn = area > 0. ? 1. / (3. * area) : 0.;
    ws->f_area[f] = area;
    st3(ws->f_cen + 3 * f,
        vscale(fn, mk(ws->facc[4 * pi + 1], ws->facc[4 * pi + 2],
                      ws->facc[4 * pi + 3])));
    st3(ws->f_norm + 3 * f, h->n);
    st3(ws->f_shift + 3 * f, h->has_shift ? h->shift : mk(0, 0, 0));
    ws->f_right[f] = h->right;
    ws->f_has_shift[f] = (uint8_t)h->has_shift;
    nf++;
  }
  *nfaces = nf;
  return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* orchestration: voronoi.rs:192-315                                   */
/* ------------------------------------------------------------------ */
typedef struct {
  size_t n;
  const double *gen;
  const uint8_t *mask;
  const grid_t *grid;
  const sim_boundary *sb;
  atomic_size_t next;
  int nthreads;
  workspace *ws;
  /* per cell results */
  double *cen, *vol, *sr;
  int32_t *cell_thread;
  uint64_t *cell_foff; /* offset into the thread's face buffer */
  uint32_t *cell_nf;
  uint32_t *n_popped, *n_planes, *n_vertices;
  uint8_t *exhausted;
  int *status;
  /* integrator export */
  int want_tets;
  uint64_t flops;
  pthread_mutex_t mu;
} job_t;

typedef struct {
  job_t *job;
  int tid;
} worker_arg;

static void *worker(void *pa) {
  worker_arg *wa = (worker_arg *)pa;
  job_t *J = wa->job;
  workspace *ws = &J->ws[wa->tid];
#ifdef ORC_COUNT_FLOPS
  t_flops = 0;
#endif
  for (;;) {
    size_t b = atomic_fetch_add(&J->next, 64);
    if (b >= J->n) break;
    size_t e = b + 64 < J->n ? b + 64 : J->n;
    for (size_t i = b; i < e; i++) {
      J->cell_thread[i] = wa->tid;
      J->cell_foff[i] = ws->flen;
      J->cell_nf[i] = 0;
      if (J->mask && !J->mask[i]) {
        J->status[i] = ORC_OK;
        continue;
      }
      build_info info;
      int st = cc_build(ws, J->grid, J->sb, J->gen, i, &info);
      J->n_popped[i] = info.n_popped;
      J->exhausted[i] = (uint8_t)info.exhausted;
      if (st) {
        J->status[i] = st;
        continue;
      }
      J->n_planes[i] = (uint32_t)ws->cell.np;
      J->n_vertices[i] = (uint32_t)ws->cell.nv;
      cell_out co;
      int nf = 0;
      st = cell_integrate(ws, J->mask, &co, &nf);
      J->status[i] = st;
      if (st) continue;
      J->cen[3 * i] = co.centroid.x;
      J->cen[3 * i + 1] = co.centroid.y;
      J->cen[3 * i + 2] = co.centroid.z;
      J->vol[i] = co.volume;
      J->sr[i] = ws->cell.safety_radius;
      J->cell_nf[i] = (uint32_t)nf;
    }
  }
#ifdef ORC_COUNT_FLOPS
  pthread_mutex_lock(&J->mu);
  J->flops += t_flops;
  pthread_mutex_unlock(&J->mu);
#endif
  return NULL;
}

static void ws_free(workspace *ws) {
  free(ws->cell.planes);
  free(ws->cell.verts);
  free(ws->cell.cyc.ptrs);
  free(ws->cands.buf);
  free(ws->f_area);
  free(ws->f_cen);
  free(ws->f_norm);
  free(ws->f_shift);
  free(ws->f_right);
  free(ws->f_has_shift);
  free(ws->maybe);
  free(ws->facc);
}

static int resolve_threads(int nthreads) {
  if (nthreads > 0) return nthreads;
  long c = sysconf(_SC_NPROCESSORS_ONLN);
  return c > 0 ? (int)c : 1;
}

/* Normalise the box for 1D/2D (voronoi.rs:200-210) and zero the unused
 * generator components (generator.rs:13-24). */
static double *prepare(const double *xyz, size_t n, const double anchor_in[3],
                       const double width_in[3], int dim, double anchor[3],
                       double width[3]) {
  for (int c = 0; c < 3; c++) {
    anchor[c] = anchor_in[c];
    width[c] = width_in[c];
  }
  if (dim == 1) {
    anchor[1] = -0.5;
    width[1] = 1.;
  }
  if (dim == 1 || dim == 2) {
    anchor[2] = -0.5;
    width[2] = 1.;
  }
  double *gen = (double *)malloc(3 * (n ? n : 1) * sizeof(double));
  if (!gen) return NULL;
  for (size_t i = 0; i < n; i++) {
    v3 g = generator_loc(xyz, i, dim);
    st3(gen + 3 * i, g);
  }
  return gen;
}

int orc_build(const double *xyz, size_t n, const uint8_t *mask,
              const double anchor_in[3], const double width_in[3], int dim,
              int periodic, int nthreads, orc_voronoi **out) {
  *out = NULL;
  if (dim < 1 || dim > 3 || (n > 0 && !xyz)) return ORC_E_ARG;
  orc_voronoi *V = (orc_voronoi *)calloc(1, sizeof(orc_voronoi));
  if (!V) return ORC_E_OOM;
  *out = V;
  double anchor[3], width[3];
  double *gen = prepare(xyz, n, anchor_in, width_in, dim, anchor, width);
  if (!gen) return V->status = ORC_E_OOM;
  memcpy(V->anchor, anchor, sizeof anchor);
  memcpy(V->width, width, sizeof width);
  V->dim = dim;
  V->periodic = periodic;
  V->n_cells = n;
  V->status_cell = -1;
  size_t nn = n ? n : 1;
  V->cell_loc = (double *)calloc(3 * nn, sizeof(double));
  V->cell_centroid = (double *)calloc(3 * nn, sizeof(double));
  V->cell_volume = (double *)calloc(nn, sizeof(double));
  V->cell_safety_radius = (double *)calloc(nn, sizeof(double));
  V->cell_face_offset = (uint64_t *)calloc(nn, sizeof(uint64_t));
  V->cell_face_count = (uint64_t *)calloc(nn, sizeof(uint64_t));
  V->cell_idx = (uint64_t *)calloc(nn, sizeof(uint64_t));
  V->cell_n_popped = (uint32_t *)calloc(nn, sizeof(uint32_t));
  V->cell_n_planes = (uint32_t *)calloc(nn, sizeof(uint32_t));
  V->cell_n_vertices = (uint32_t *)calloc(nn, sizeof(uint32_t));
  V->cell_exhausted = (uint8_t *)calloc(nn, 1);
  if (n == 0) {
    free(gen);
    return ORC_OK;
  }

  grid_t grid;
  int st = grid_build(&grid, gen, n, width, dim, periodic);
  if (st) {
    free(gen);
    return V->status = st;
  }
  sim_boundary sb;
  boundary_cuboid(&sb, ld3(anchor), ld3(width), periodic, dim);

  job_t J;
  memset(&J, 0, sizeof J);
  J.n = n;
  J.gen = gen;
  J.mask = mask;
  J.grid = &grid;
  J.sb = &sb;
  atomic_init(&J.next, 0);
  J.nthreads = resolve_threads(nthreads);
  J.ws = (workspace *)calloc(J.nthreads, sizeof(workspace));
  J.cen = V->cell_centroid;
  J.vol = V->cell_volume;
  J.sr = V->cell_safety_radius;
  J.cell_thread = (int32_t *)malloc(n * sizeof(int32_t));
  J.cell_foff = (uint64_t *)malloc(n * sizeof(uint64_t));
  J.cell_nf = (uint32_t *)malloc(n * sizeof(uint32_t));
  J.n_popped = V->cell_n_popped;
  J.n_planes = V->cell_n_planes;
  J.n_vertices = V->cell_n_vertices;
  J.exhausted = V->cell_exhausted;
  J.status = (int *)calloc(n, sizeof(int));
  pthread_mutex_init(&J.mu, NULL);

  pthread_t *th = (pthread_t *)malloc(J.nthreads * sizeof(pthread_t));
  worker_arg *args = (worker_arg *)malloc(J.nthreads * sizeof(worker_arg));
  for (int t = 0; t < J.nthreads; t++) {
    args[t].job = &J;
    args[t].tid = t;
    pthread_create(&th[t], NULL, worker, &args[t]);
  }
  for (int t = 0; t < J.nthreads; t++) pthread_join(th[t], NULL);
  V->flops = J.flops;
  for (int t = 0; t < J.nthreads; t++) V->exact_tests += J.ws[t].exact_tests;

  /* first failing cell decides the status (the reference would panic) */
  for (size_t i = 0; i < n; i++)
    if (J.status[i]) {
      V->status = J.status[i];
      V->status_cell = (int64_t)i;
      break;
    }

  if (V->status == ORC_OK) {
    /* cells (VoronoiCell::init / Default) */
    for (size_t i = 0; i < n; i++) {
      if (mask && !mask[i]) continue;
      V->cell_loc[3 * i] = gen[3 * i];
      V->cell_loc[3 * i + 1] = gen[3 * i + 1];
      V->cell_loc[3 * i + 2] = gen[3 * i + 2];
      V->cell_idx[i] = i;
    }
    /* flatten! (voronoi.rs:225-226): faces in cell order */
    size_t F = 0;
    for (size_t i = 0; i < n; i++) F += J.cell_nf[i];
    size_t FF = F ? F : 1;
    V->n_faces = F;
    V->face_left = (uint64_t *)malloc(FF * sizeof(uint64_t));
    V->face_right = (int64_t *)malloc(FF * sizeof(int64_t));
    V->face_area = (double *)malloc(FF * sizeof(double));
    V->face_centroid = (double *)malloc(3 * FF * sizeof(double));
    V->face_normal = (double *)malloc(3 * FF * sizeof(double));
    V->face_has_shift = (uint8_t *)malloc(FF);
    V->face_shift = (double *)malloc(3 * FF * sizeof(double));
    size_t f = 0;
    for (size_t i = 0; i < n; i++) {
      workspace *ws = &J.ws[J.cell_thread[i]];
      size_t o = J.cell_foff[i];
      for (uint32_t k = 0; k < J.cell_nf[i]; k++, f++, o++) {
        V->face_left[f] = i;
        V->face_right[f] = ws->f_right[o];
        V->face_area[f] = ws->f_area[o];
        memcpy(V->face_centroid + 3 * f, ws->f_cen + 3 * o, 24);
        memcpy(V->face_normal + 3 * f, ws->f_norm + 3 * o, 24);
        memcpy(V->face_shift + 3 * f, ws->f_shift + 3 * o, 24);
        V->face_has_shift[f] = ws->f_has_shift[o];
      }
    }
    /* finalize (voronoi.rs:241-263) */
    uint64_t *cnt = V->cell_face_count;
    for (size_t k = 0; k < F; k++) {
      cnt[V->face_left[k]]++;
      if (V->face_right[k] >= 0 && !V->face_has_shift[k])
        cnt[V->face_right[k]]++;
    }
    uint64_t off = 0;
    for (size_t i = 0; i < n; i++) {
      V->cell_face_offset[i] = off;
      off += cnt[i];
    }
    V->n_links = off;
    V->links = (uint64_t *)malloc((off ? off : 1) * sizeof(uint64_t));
    uint64_t *fill = (uint64_t *)malloc(n * sizeof(uint64_t));
    memcpy(fill, V->cell_face_offset, n * sizeof(uint64_t));
    for (size_t k = 0; k < F; k++) {
      V->links[fill[V->face_left[k]]++] = k;
      if (V->face_right[k] >= 0 && !V->face_has_shift[k])
        V->links[fill[V->face_right[k]]++] = k;
    }
    free(fill);
  }

  for (int t = 0; t < J.nthreads; t++) ws_free(&J.ws[t]);
  free(J.ws);
  free(J.cell_thread);
  free(J.cell_foff);
  free(J.cell_nf);
  free(J.status);
  free(th);
  free(args);
  pthread_mutex_destroy(&J.mu);
  grid_free(&grid);
  free(gen);
  return V->status;
}

void orc_free(orc_voronoi *V) {
  if (!V) return;
  free(V->cell_loc);
  free(V->cell_centroid);
  free(V->cell_volume);
  free(V->cell_safety_radius);
  free(V->cell_face_offset);
  free(V->cell_face_count);
  free(V->cell_idx);
  free(V->face_left);
  free(V->face_right);
  free(V->face_area);
  free(V->face_centroid);
  free(V->face_normal);
  free(V->face_has_shift);
  free(V->face_shift);
  free(V->links);
  free(V->cell_n_popped);
  free(V->cell_n_planes);
  free(V->cell_n_vertices);
  free(V->cell_exhausted);
  free(V);
}

/* ------------------------------------------------------------------ */
/* decomposition export (single-threaded; small inputs)                */
/* ------------------------------------------------------------------ */
int orc_build_tets(const double *xyz, size_t n, const uint8_t *mask,
                   const double anchor_in[3], const double width_in[3],
                   int dim, int periodic, int nthreads, orc_tets **out) {
  (void)nthreads;
  *out = NULL;
  if (dim < 1 || dim > 3 || n == 0 || !xyz) return ORC_E_ARG;
  orc_tets *T = (orc_tets *)calloc(1, sizeof(orc_tets));
  *out = T;
  double anchor[3], width[3];
  double *gen = prepare(xyz, n, anchor_in, width_in, dim, anchor, width);
  grid_t grid;
  int st = grid_build(&grid, gen, n, width, dim, periodic);
  if (st) {
    free(gen);
    return T->status = st;
  }
  sim_boundary sb;
  boundary_cuboid(&sb, ld3(anchor), ld3(width), periodic, dim);
  workspace ws;
  memset(&ws, 0, sizeof ws);
  T->n_cells = n;
  T->tet_offset = (uint64_t *)calloc(n + 1, sizeof(uint64_t));
  T->plane_offset = (uint64_t *)calloc(n + 1, sizeof(uint64_t));
  T->vert_offset = (uint64_t *)calloc(n + 1, sizeof(uint64_t));
  size_t tcap = 1024, pcap = 256, vcap = 256;
  T->tet_v = (double *)malloc(9 * tcap * sizeof(double));
  T->tet_plane = (int32_t *)malloc(tcap * sizeof(int32_t));
  T->plane_n = (double *)malloc(3 * pcap * sizeof(double));
  T->plane_p = (double *)malloc(3 * pcap * sizeof(double));
  T->plane_right = (int64_t *)malloc(pcap * sizeof(int64_t));
  T->plane_has_shift = (uint8_t *)malloc(pcap);
  T->plane_shift = (double *)malloc(3 * pcap * sizeof(double));
  T->vert_loc = (double *)malloc(3 * vcap * sizeof(double));
  T->vert_dual = (int32_t *)malloc(3 * vcap * sizeof(int32_t));
  size_t nt = 0, np = 0, nv = 0;
  for (size_t i = 0; i < n && !st; i++) {
    if (!mask || mask[i]) {
      build_info info;
      st = cc_build(&ws, &grid, &sb, gen, i, &info);
      if (st) {
        T->status = st;
        break;
      }
      ccell *c = &ws.cell;
      while (np + c->np > pcap) {
        pcap *= 2;
        T->plane_n = (double *)realloc(T->plane_n, 3 * pcap * sizeof(double));
        T->plane_p = (double *)realloc(T->plane_p, 3 * pcap * sizeof(double));
        T->plane_right =
            (int64_t *)realloc(T->plane_right, pcap * sizeof(int64_t));
        T->plane_has_shift = (uint8_t *)realloc(T->plane_has_shift, pcap);
        T->plane_shift =
            (double *)realloc(T->plane_shift, 3 * pcap * sizeof(double));
      }
      for (int k = 0; k < c->np; k++, np++) {
        st3(T->plane_n + 3 * np, c->planes[k].n);
        st3(T->plane_p + 3 * np, c->planes[k].p);
        T->plane_right[np] = c->planes[k].right;
        T->plane_has_shift[np] = (uint8_t)c->planes[k].has_shift;
        st3(T->plane_shift + 3 * np, c->planes[k].shift);
      }
      while (nv + c->nv > vcap) {
        vcap *= 2;
        T->vert_loc = (double *)realloc(T->vert_loc, 3 * vcap * sizeof(double));
        T->vert_dual =
            (int32_t *)realloc(T->vert_dual, 3 * vcap * sizeof(int32_t));
      }
      while (nt + 6 * (size_t)c->nv > tcap) {
        tcap *= 2;
        T->tet_v = (double *)realloc(T->tet_v, 9 * tcap * sizeof(double));
        T->tet_plane = (int32_t *)realloc(T->tet_plane, tcap * sizeof(int32_t));
      }
      for (int k = 0; k < c->nv; k++, nv++) {
        const vert *v = &c->verts[k];
        st3(T->vert_loc + 3 * nv, v->loc);
        for (int q = 0; q < 3; q++) T->vert_dual[3 * nv + q] = v->dual[q];
        v3 proj[6];
        if (!load_projections(c, v, proj)) {
          T->status = st = ORC_E_PLANE_DET_ZERO;
          break;
        }
        for (int t = 0; t < 6; t++, nt++) {
          st3(T->tet_v + 9 * nt, proj[t]);
          st3(T->tet_v + 9 * nt + 3, proj[(t + 5) % 6]);
          st3(T->tet_v + 9 * nt + 6, v->loc);
          T->tet_plane[nt] = v->dual[t / 2];
        }
      }
    }
    T->tet_offset[i + 1] = nt;
    T->plane_offset[i + 1] = np;
    T->vert_offset[i + 1] = nv;
  }
  ws_free(&ws);
  grid_free(&grid);
  free(gen);
  return T->status;
}

void orc_free_tets(orc_tets *T) {
  if (!T) return;
  free(T->tet_offset);
  free(T->tet_v);
  free(T->tet_plane);
  free(T->plane_offset);
  free(T->plane_n);
  free(T->plane_p);
  free(T->plane_right);
  free(T->plane_has_shift);
  free(T->plane_shift);
  free(T->vert_offset);
  free(T->vert_loc);
  free(T->vert_dual);
  free(T);
}

/* ------------------------------------------------------------------ */
/* with_faces: face vertex lists and the fan decomposition             */
/* ------------------------------------------------------------------ */

/* Vertex::plane_idx (convex_cell.rs:50-59): position of half-space p in a
 * vertex's dual, -1 if absent */
static int dual_pos(const int32_t *d, int32_t p) {
  for (int k = 0; k < 3; k++)
    if (d[k] == p) return k;
  return -1;
}

/* sort_face_vertices (convex_cell.rs:508-543): vi holds the m vertices of
 * half-space `plane` in ascending index; starting from vi[0], repeatedly
 * move the first later vertex that contains the current edge's other plane
 * to the next slot (swap), the next edge plane being the one after `plane`
 * in that vertex's dual. */
static int sort_face_vertices(const int32_t *dual, uint32_t *vi, uint32_t m,
                              int32_t plane) {
  if (m == 0) return ORC_OK;
  const int32_t *d = dual + 3 * (size_t)vi[0];
  int pos = dual_pos(d, plane);
  if (pos < 0) return ORC_E_FACE_WALK;
  int32_t next_plane = d[(pos + 1) % 3];
  uint32_t cur = 1;
  while (cur + 1 < m) {
    uint32_t t = cur;
    for (; t < m; t++) {
      d = dual + 3 * (size_t)vi[t];
      pos = dual_pos(d, plane);
      if (pos < 0) return ORC_E_FACE_WALK;
      if (d[0] == next_plane || d[1] == next_plane || d[2] == next_plane) {
        next_plane = d[(pos + 1) % 3];
        const uint32_t sw = vi[cur];
        vi[cur] = vi[t];
        vi[t] = sw;
        cur++;
        break;
      }
    }
    if (t >= m) return ORC_E_FACE_WALK; /* :538-541 */
  }
  return ORC_OK;
}

/* ConvexCell::with_faces (convex_cell.rs:461-505) per cell, then
 * DecompositionWithFaces (:89-129): for each face in order, the fan
 * (v[0], v[k], v[k+1]), k = 1 .. count-2, tagged with the face's
 * half-space. */
int orc_with_faces(const orc_tets *T, int dim, orc_faces **out) {
  orc_faces *F = (orc_faces *)calloc(1, sizeof(orc_faces));
  *out = F;
  if (!F) return ORC_E_OOM;
  if (dim != 3) return F->status = ORC_E_NOT_3D; /* convex_cell.rs:462-466 */
  const size_t n = T->n_cells;
  F->n_cells = n;
  const uint64_t np_all = T->plane_offset[n], nv_all = T->vert_offset[n];
  F->face_offset = (uint64_t *)calloc(n + 1, sizeof(uint64_t));
  F->conn_offset = (uint64_t *)calloc(n + 1, sizeof(uint64_t));
  F->tet_offset = (uint64_t *)calloc(n + 1, sizeof(uint64_t));
  F->face_plane = (int32_t *)malloc((np_all + 1) * sizeof(int32_t));
  F->face_count = (uint32_t *)malloc((np_all + 1) * sizeof(uint32_t));
  F->face_vertex_offset = (uint64_t *)malloc((np_all + 1) * sizeof(uint64_t));
  F->conn = (uint32_t *)malloc((3 * nv_all + 1) * sizeof(uint32_t));
  /* a face of m vertices gives m - 2 tets: at most 3 per vertex */
  F->tet_v = (double *)malloc((9 * 3 * nv_all + 1) * sizeof(double));
  F->tet_plane = (int32_t *)malloc((3 * nv_all + 1) * sizeof(int32_t));
  if (!F->face_offset || !F->conn_offset || !F->tet_offset || !F->face_plane ||
      !F->face_count || !F->face_vertex_offset || !F->conn || !F->tet_v ||
      !F->tet_plane)
    return F->status = ORC_E_OOM;
  uint32_t *cnt = NULL, *cursor = NULL;
  size_t cap = 0;
  size_t nf = 0, ne = 0, nt = 0;
  int st = ORC_OK;
  for (size_t i = 0; i < n && !st; i++) {
    const size_t np = T->plane_offset[i + 1] - T->plane_offset[i];
    const size_t nv = T->vert_offset[i + 1] - T->vert_offset[i];
    const int32_t *dual = T->vert_dual + 3 * T->vert_offset[i];
    const double *loc = T->vert_loc + 3 * T->vert_offset[i];
    if (np > cap) {
      cap = np;
      cnt = (uint32_t *)realloc(cnt, cap * sizeof(uint32_t));
      cursor = (uint32_t *)realloc(cursor, cap * sizeof(uint32_t));
    }
    /* face_vertex_connections[dual[k]].push(idx), vertices in order */
    for (size_t p = 0; p < np; p++) cnt[p] = 0;
    for (size_t v = 0; v < nv; v++)
      for (int k = 0; k < 3; k++) cnt[dual[3 * v + k]]++;
    uint32_t run = 0;
    for (size_t p = 0; p < np; p++) {
      cursor[p] = run;
      run += cnt[p];
    }
    uint32_t *conn = F->conn + ne;
    for (size_t v = 0; v < nv; v++)
      for (int k = 0; k < 3; k++) conn[cursor[dual[3 * v + k]]++] = (uint32_t)v;
    /* faces in half-space order; the offset is relative to the cell */
    run = 0;
    for (size_t p = 0; p < np && !st; p++) {
      if (!cnt[p]) continue;
      st = sort_face_vertices(dual, conn + run, cnt[p], (int32_t)p);
      if (st) break;
      F->face_plane[nf] = (int32_t)p;
      F->face_count[nf] = cnt[p];
      F->face_vertex_offset[nf] = run;
      nf++;
      if (cnt[p] < 3) { /* the fan would index past the face (:113-118) */
        st = ORC_E_FACE_WALK;
        break;
      }
      const uint32_t *fv = conn + run;
      for (uint32_t k = 1; k + 1 < cnt[p]; k++, nt++) {
        memcpy(F->tet_v + 9 * nt, loc + 3 * (size_t)fv[0], 3 * sizeof(double));
        memcpy(F->tet_v + 9 * nt + 3, loc + 3 * (size_t)fv[k], 3 * sizeof(double));
        memcpy(F->tet_v + 9 * nt + 6, loc + 3 * (size_t)fv[k + 1], 3 * sizeof(double));
        F->tet_plane[nt] = (int32_t)p;
      }
      run += cnt[p];
    }
    ne += 3 * nv;
    F->face_offset[i + 1] = nf;
    F->conn_offset[i + 1] = ne;
    F->tet_offset[i + 1] = nt;
  }
  free(cnt);
  free(cursor);
  return F->status = st;
}

void orc_free_faces(orc_faces *F) {
  if (!F) return;
  free(F->face_offset);
  free(F->face_plane);
  free(F->face_count);
  free(F->face_vertex_offset);
  free(F->conn_offset);
  free(F->conn);
  free(F->tet_offset);
  free(F->tet_v);
  free(F->tet_plane);
  free(F);
}

/* ------------------------------------------------------------------ */
/* exported primitives                                                 */
/* ------------------------------------------------------------------ */
void orc_iloc(const double anchor[3], const double width[3], int periodic,
              int dim, const double loc[3], int64_t out[3]) {
  sim_boundary sb;
  boundary_cuboid(&sb, ld3(anchor), ld3(width), periodic, dim);
  boundary_iloc(&sb, ld3(loc), out);
}

double orc_signed_volume_tet(const double v0[3], const double v1[3],
                             const double v2[3], const double v3_[3]) {
  return signed_volume_tet(ld3(v0), ld3(v1), ld3(v2), ld3(v3_));
}

double orc_signed_area_tri(const double v0[3], const double v1[3],
                           const double v2[3], const double t[3]) {
  return signed_area_tri(ld3(v0), ld3(v1), ld3(v2), ld3(t));
}

int orc_intersect_planes(const double n0[3], const double p0[3],
                         const double n1[3], const double p1[3],
                         const double n2[3], const double p2[3],
                         double out[3]) {
  v3 r;
  int ok = intersect_planes(ld3(n0), vdot(ld3(p0), ld3(n0)), ld3(n1),
                            vdot(ld3(p1), ld3(n1)), ld3(n2),
                            vdot(ld3(p2), ld3(n2)), &r);
  if (ok) st3(out, r);
  return ok;
}

int orc_halfspace_clip(const double n[3], const double p[3],
                       const double v[3]) {
  hspace h = hs_new(ld3(n), ld3(p), -1, 0, mk(0, 0, 0));
  return (int)hs_clip(&h, ld3(v));
}

size_t orc_neighbour_stream(const double *xyz, size_t n, const double width[3],
                            int dim, int periodic, size_t query,
                            size_t max_out, uint64_t *out_idx,
                            int32_t *out_image, double *out_key) {
  double anchor0[3] = {0, 0, 0}, anchor[3], w[3];
  double *gen = prepare(xyz, n, anchor0, width, dim, anchor, w);
  grid_t g;
  if (grid_build(&g, gen, n, w, dim, periodic)) {
    free(gen);
    return 0;
  }
  cand_vec cv;
  memset(&cv, 0, sizeof cv);
  double R = 2.0 * g.h_min;
  size_t cnt = 0;
  for (;;) {
    gather(&g, ld3(gen + 3 * query), R, &cv);
    if (cv.len >= max_out || cv.len >= g.total_stream) break;
    R *= 2.;
  }
  cnt = cv.len < max_out ? cv.len : max_out;
  for (size_t i = 0; i < cnt; i++) {
    out_idx[i] = cv.buf[i].idx;
    out_image[i] = cv.buf[i].img;
    out_key[i] = cv.buf[i].key;
  }
  free(cv.buf);
  grid_free(&g);
  free(gen);
  return cnt;
}
