// mv_geom.hpp -- FP64 geometry on the device, in the exact operand order of
// the reference (meshless_voronoi v0.7.2 over glam 0.27), so that every
// clip decision of a cell is taken on the same bits as the reference's.
//
// This translation unit is compiled with -ffp-contract=off: the reference
// is Rust, which never fuses a multiply with an add.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mv {

struct d3 {
  double x, y, z;
};

__host__ __device__ __forceinline__ d3 mk3(double x, double y, double z) {
  d3 r;
  r.x = x;
  r.y = y;
  r.z = z;
  return r;
}
__host__ __device__ __forceinline__ d3 add3(d3 a, d3 b) {
  return mk3(a.x + b.x, a.y + b.y, a.z + b.z);
}
__host__ __device__ __forceinline__ d3 sub3(d3 a, d3 b) {
  return mk3(a.x - b.x, a.y - b.y, a.z - b.z);
}
__host__ __device__ __forceinline__ d3 scl3(double s, d3 a) {
  return mk3(s * a.x, s * a.y, s * a.z);
}
__host__ __device__ __forceinline__ d3 div3(d3 a, double s) {
  return mk3(a.x / s, a.y / s, a.z / s);
}
// glam DVec3::dot: (x*x' + y*y') + z*z'
__host__ __device__ __forceinline__ double dot3(d3 a, d3 b) {
  return (a.x * b.x + a.y * b.y) + a.z * b.z;
}
// glam DVec3::cross
__host__ __device__ __forceinline__ d3 cross3(d3 a, d3 b) {
  return mk3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x,
             a.x * b.y - b.x * a.y);
}
__device__ __forceinline__ d3 abs3(d3 a) {
  return mk3(fabs(a.x), fabs(a.y), fabs(a.z));
}
__device__ __forceinline__ d3 xyz(double4 v) { return mk3(v.x, v.y, v.z); }

// Plane::project_onto (src/geometry.rs:36-38) over glam's project_onto:
// point + (n * ((p - point).n)) * (1 / n.n)
__device__ __forceinline__ d3 plane_project_onto(d3 n, d3 p, d3 point) {
  d3 diff = sub3(p, point);
  double rcp = 1.0 / dot3(n, n);
  double s = dot3(diff, n);
  return add3(point, scl3(rcp, scl3(s, n)));
}

// intersect_planes (src/geometry.rs:68-76); d_i = p_i . n_i.
// Returns false when the determinant is exactly zero (reference panics).
__host__ __device__ __forceinline__ bool intersect_planes(d3 n0, double d0, d3 n1,
                                                 double d1, d3 n2, double d2,
                                                 d3 *out) {
  d3 c01 = cross3(n0, n1);
  double det = dot3(n2, c01);
  if (det == 0.0) return false;
  d3 a = scl3(d0, cross3(n1, n2));
  d3 b = scl3(d1, cross3(n2, n0));
  d3 c = scl3(d2, c01);
  *out = div3(add3(add3(a, b), c), det);
  return true;
}

// Plane::project_onto_intersection (src/geometry.rs:41-47): the third plane
// passes through `point` with normal na x nb.
__device__ __forceinline__ bool project_onto_intersection(d3 na, double da,
                                                          d3 nb, double db,
                                                          d3 point, d3 *out) {
  d3 np = cross3(na, nb);
  double dp = dot3(point, np);
  return intersect_planes(na, da, nb, db, np, dp, out);
}

// signed_volume_tet (src/geometry.rs:82-88)
__device__ __forceinline__ double signed_volume_tet(d3 v0, d3 v1, d3 v2,
                                                    d3 v3) {
  d3 a = sub3(v1, v0), b = sub3(v2, v0), c = sub3(v3, v0);
  return dot3(c, cross3(a, b)) / 6.0;
}

// Rust f64::signum for a non-NaN argument: +1 for +0, -1 for -0
__device__ __forceinline__ double rsignum(double x) {
  return isnan(x) ? x : copysign(1.0, x);
}

// signed_area_tri (src/geometry.rs:95-100)
__device__ __forceinline__ double signed_area_tri(d3 v0, d3 v1, d3 v2, d3 t) {
  d3 n = scl3(0.5, cross3(sub3(v1, v0), sub3(v2, v0)));
  double sign = rsignum(dot3(sub3(t, v0), n));
  return sqrt(dot3(n, n)) * sign;
}

// ---------------------------------------------------------------------
// Exact in-sphere sign (src/geometry.rs:271-314) on snapped 52-bit integer
// coordinates.  The determinant is that of the 4x4 matrix M whose columns
// are b-a, c-a, d-a, v-a, each extended by its squared norm.  Any exact
// evaluation of det(M) has the reference's sign; here M is expanded by
// Laplace along rows {x, y} against rows {z, |.|^2}:
//   det M = sum_{j<k} (-1)^(1+j+k) m01(j,k) m23(~(j,k))
// with m01 = 2x2 minors of the x/y rows (< 2^106, int128) and m23 = 2x2
// minors of the z/norm rows (< 2^160, 192-bit); |det M| < 2^268 fits the
// 320-bit accumulator.  A floating-point evaluation of the same expansion
// with an error bound decides first; only ties and near-ties reach the
// integer path.
// ---------------------------------------------------------------------
typedef unsigned __int128 u128;
typedef __int128 s128;

struct u192 {
  uint64_t w[3];
};
struct u320 {
  uint64_t w[5];
};

// |z| * w for w >= 0 (w < 2^107): 64 x 128 -> 192 bits
__device__ __forceinline__ u192 mul_u64_u128(uint64_t a, u128 b) {
  u128 lo = (u128)a * (uint64_t)b;
  u128 hi = (u128)a * (uint64_t)(b >> 64);
  u192 r;
  r.w[0] = (uint64_t)lo;
  u128 mid = (lo >> 64) + (u128)(uint64_t)hi;
  r.w[1] = (uint64_t)mid;
  r.w[2] = (uint64_t)(hi >> 64) + (uint64_t)(mid >> 64);
  return r;
}
__device__ __forceinline__ u192 neg192(u192 a) {
  u192 r;
  u128 c = 1;
  for (int i = 0; i < 3; i++) {
    c += (u128)(~a.w[i]);
    r.w[i] = (uint64_t)c;
    c >>= 64;
  }
  return r;
}
__device__ __forceinline__ u192 add192(u192 a, u192 b) {
  u192 r;
  u128 c = 0;
  for (int i = 0; i < 3; i++) {
    c += (u128)a.w[i] + b.w[i];
    r.w[i] = (uint64_t)c;
    c >>= 64;
  }
  return r;
}
// z * w as 192-bit two's complement
__device__ __forceinline__ u192 zmul(int64_t z, u128 w) {
  u192 m = mul_u64_u128((uint64_t)(z < 0 ? -z : z), w);
  return z < 0 ? neg192(m) : m;
}
__device__ __forceinline__ void add320(u320 &acc, const u320 &b) {
  u128 c = 0;
  for (int i = 0; i < 5; i++) {
    c += (u128)acc.w[i] + b.w[i];
    acc.w[i] = (uint64_t)c;
    c >>= 64;
  }
}
// acc += s * m01 * m23 (s = +1 or -1), all in two's complement
__device__ __forceinline__ void acc_term(u320 &acc, int s, s128 m01,
                                         u192 m23) {
  bool neg = m01 < 0;
  u128 a = (u128)(neg ? -m01 : m01);
  if ((int64_t)m23.w[2] < 0) {
    m23 = neg192(m23);
    neg = !neg;
  }
  if (s < 0) neg = !neg;
  // |m01| * |m23|: 2 x 3 limbs
  uint64_t al[2] = {(uint64_t)a, (uint64_t)(a >> 64)};
  u320 p;
  for (int i = 0; i < 5; i++) p.w[i] = 0;
  for (int i = 0; i < 2; i++) {
    u128 carry = 0;
    for (int j = 0; j < 3; j++) {
      u128 t = (u128)al[i] * m23.w[j] + p.w[i + j] + carry;
      p.w[i + j] = (uint64_t)t;
      carry = t >> 64;
    }
    p.w[i + 3] += (uint64_t)carry;
  }
  if (neg) {
    u128 c = 1;
    for (int i = 0; i < 5; i++) {
      c += (u128)(~p.w[i]);
      p.w[i] = (uint64_t)c;
      c >>= 64;
    }
  }
  add320(acc, p);
}

__device__ __noinline__ int in_sphere_exact_int(const int64_t *x,
                                                const int64_t *y,
                                                const int64_t *z) {
  // x, y, z: the 4 columns' coordinate differences
  u128 w[4];
  for (int j = 0; j < 4; j++) {
    w[j] = (u128)((s128)x[j] * x[j]) + (u128)((s128)y[j] * y[j]) +
           (u128)((s128)z[j] * z[j]);
  }
  const int J[6] = {0, 0, 0, 1, 1, 2}, K[6] = {1, 2, 3, 2, 3, 3};
  const int CJ[6] = {2, 1, 1, 0, 0, 0}, CK[6] = {3, 3, 2, 3, 2, 1};
  const int SG[6] = {1, -1, 1, 1, -1, 1};
  u320 acc;
  for (int i = 0; i < 5; i++) acc.w[i] = 0;
  for (int t = 0; t < 6; t++) {
    int j = J[t], k = K[t], cj = CJ[t], ck = CK[t];
    s128 m01 = (s128)x[j] * y[k] - (s128)x[k] * y[j];
    u192 m23 = add192(zmul(z[cj], w[ck]), neg192(zmul(z[ck], w[cj])));
    acc_term(acc, SG[t], m01, m23);
  }
  if ((int64_t)acc.w[4] < 0) return -1;
  return (acc.w[0] | acc.w[1] | acc.w[2] | acc.w[3] | acc.w[4]) ? 1 : 0;
}

// a, b, c, d, v: snapped coordinates (3 each).  Returns -1, 0, +1.
__device__ __noinline__ int in_sphere_exact(const int64_t *a, const int64_t *b,
                                            const int64_t *c, const int64_t *d,
                                            const int64_t *v) {
  int64_t x[4], y[4], z[4];
  const int64_t *pts[4] = {b, c, d, v};
  for (int j = 0; j < 4; j++) {
    x[j] = pts[j][0] - a[0];
    y[j] = pts[j][1] - a[1];
    z[j] = pts[j][2] - a[2];
  }
  // floating-point filter on the same expansion; differences are exact
  // in double, squared norms carry <= 3 ulp, and every path of the
  // expansion adds <= 12 more roundings: |error| <= 16 u * permanent,
  // checked here against 1e-13 * permanent.
  double fx[4], fy[4], fz[4], fw[4];
  for (int j = 0; j < 4; j++) {
    fx[j] = (double)x[j];
    fy[j] = (double)y[j];
    fz[j] = (double)z[j];
    fw[j] = (fx[j] * fx[j] + fy[j] * fy[j]) + fz[j] * fz[j];
  }
  const int J[6] = {0, 0, 0, 1, 1, 2}, K[6] = {1, 2, 3, 2, 3, 3};
  const int CJ[6] = {2, 1, 1, 0, 0, 0}, CK[6] = {3, 3, 2, 3, 2, 1};
  const double SG[6] = {1., -1., 1., 1., -1., 1.};
  double det = 0., per = 0.;
  for (int t = 0; t < 6; t++) {
    int j = J[t], k = K[t], cj = CJ[t], ck = CK[t];
    double m01 = fx[j] * fy[k] - fx[k] * fy[j];
    double p01 = fabs(fx[j] * fy[k]) + fabs(fx[k] * fy[j]);
    double m23 = fz[cj] * fw[ck] - fz[ck] * fw[cj];
    double p23 = fabs(fz[cj] * fw[ck]) + fabs(fz[ck] * fw[cj]);
    det += SG[t] * (m01 * m23);
    per += p01 * p23;
  }
  double bound = 1e-13 * per;
  if (det > bound) return 1;
  if (det < -bound) return -1;
  return in_sphere_exact_int(x, y, z);
}

}  // namespace mv
