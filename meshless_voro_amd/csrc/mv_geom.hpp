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

__device__ __forceinline__ d3 mk3(double x, double y, double z) {
  d3 r;
  r.x = x;
  r.y = y;
  r.z = z;
  return r;
}
__device__ __forceinline__ d3 add3(d3 a, d3 b) {
  return mk3(a.x + b.x, a.y + b.y, a.z + b.z);
}
__device__ __forceinline__ d3 sub3(d3 a, d3 b) {
  return mk3(a.x - b.x, a.y - b.y, a.z - b.z);
}
__device__ __forceinline__ d3 scl3(double s, d3 a) {
  return mk3(s * a.x, s * a.y, s * a.z);
}
__device__ __forceinline__ d3 div3(d3 a, double s) {
  return mk3(a.x / s, a.y / s, a.z / s);
}
// glam DVec3::dot: (x*x' + y*y') + z*z'
__device__ __forceinline__ double dot3(d3 a, d3 b) {
  return (a.x * b.x + a.y * b.y) + a.z * b.z;
}
// glam DVec3::cross
__device__ __forceinline__ d3 cross3(d3 a, d3 b) {
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
__device__ __forceinline__ bool intersect_planes(d3 n0, double d0, d3 n1,
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
// Exact in-sphere sign (src/geometry.rs:271-314) on snapped 52-bit
// integer coordinates.  |det| < 2^267, so 320-bit two's complement
// integers hold every intermediate exactly.  A floating-point evaluation
// of the same determinant with an error bound decides first; only when it
// cannot, the integer expansion runs.  Both give the sign of the same
// integer determinant.
// ---------------------------------------------------------------------
struct i320 {
  uint64_t w[5];
};

__device__ __forceinline__ i320 i320_of(int64_t v) {
  i320 r;
  uint64_t e = v < 0 ? ~0ull : 0ull;
  r.w[0] = (uint64_t)v;
  r.w[1] = e;
  r.w[2] = e;
  r.w[3] = e;
  r.w[4] = e;
  return r;
}
__device__ __forceinline__ i320 i320_add(const i320 &a, const i320 &b) {
  i320 r;
  unsigned __int128 c = 0;
#pragma unroll
  for (int i = 0; i < 5; i++) {
    c += (unsigned __int128)a.w[i] + b.w[i];
    r.w[i] = (uint64_t)c;
    c >>= 64;
  }
  return r;
}
__device__ __forceinline__ i320 i320_neg(const i320 &a) {
  i320 r;
  unsigned __int128 c = 1;
#pragma unroll
  for (int i = 0; i < 5; i++) {
    c += (unsigned __int128)(~a.w[i]);
    r.w[i] = (uint64_t)c;
    c >>= 64;
  }
  return r;
}
__device__ __forceinline__ i320 i320_sub(const i320 &a, const i320 &b) {
  return i320_add(a, i320_neg(b));
}
// product modulo 2^320; exact because every product here fits
__device__ __forceinline__ i320 i320_mul(const i320 &a, const i320 &b) {
  i320 r;
#pragma unroll
  for (int i = 0; i < 5; i++) r.w[i] = 0;
#pragma unroll
  for (int i = 0; i < 5; i++) {
    unsigned __int128 carry = 0;
#pragma unroll
    for (int j = 0; i + j < 5; j++) {
      unsigned __int128 t =
          (unsigned __int128)a.w[i] * b.w[j] + r.w[i + j] + carry;
      r.w[i + j] = (uint64_t)t;
      carry = t >> 64;
    }
  }
  return r;
}
__device__ __forceinline__ int i320_sign(const i320 &a) {
  if ((int64_t)a.w[4] < 0) return -1;
  return (a.w[0] | a.w[1] | a.w[2] | a.w[3] | a.w[4]) ? 1 : 0;
}
__device__ __forceinline__ i320 i320_det2(const i320 &a, const i320 &b,
                                          const i320 &c, const i320 &d) {
  return i320_sub(i320_mul(a, d), i320_mul(b, c));
}
// rows (a0 a1 a2), (b0 b1 b2), (c0 c1 c2) expanded along the first row
__device__ __noinline__ i320 i320_det3(const i320 &a0, const i320 &a1,
                                       const i320 &a2, const i320 &b0,
                                       const i320 &b1, const i320 &b2,
                                       const i320 &c0, const i320 &c1,
                                       const i320 &c2) {
  i320 d = i320_mul(a0, i320_det2(b1, b2, c1, c2));
  d = i320_sub(d, i320_mul(a1, i320_det2(b0, b2, c0, c2)));
  d = i320_add(d, i320_mul(a2, i320_det2(b0, b1, c0, c1)));
  return d;
}

// lifted column: (x - a) per component and its squared norm
__device__ __forceinline__ void i320_lift(const int64_t *x, const int64_t *a,
                                          i320 out[4]) {
  out[0] = i320_of(x[0] - a[0]);
  out[1] = i320_of(x[1] - a[1]);
  out[2] = i320_of(x[2] - a[2]);
  i320 n = i320_mul(out[0], out[0]);
  n = i320_add(n, i320_mul(out[1], out[1]));
  n = i320_add(n, i320_mul(out[2], out[2]));
  out[3] = n;
}

__device__ __noinline__ int in_sphere_exact_int(const int64_t *a,
                                                const int64_t *b,
                                                const int64_t *c,
                                                const int64_t *d,
                                                const int64_t *v) {
  i320 B[4], C[4], D[4], V[4];
  i320_lift(b, a, B);
  i320_lift(c, a, C);
  i320_lift(d, a, D);
  i320_lift(v, a, V);
  i320 det = i320_of(0);
  i320 t;
  t = i320_det3(B[1], C[1], D[1], B[2], C[2], D[2], B[3], C[3], D[3]);
  det = i320_sub(det, i320_mul(V[0], t));
  t = i320_det3(B[0], C[0], D[0], B[2], C[2], D[2], B[3], C[3], D[3]);
  det = i320_add(det, i320_mul(V[1], t));
  t = i320_det3(B[0], C[0], D[0], B[1], C[1], D[1], B[3], C[3], D[3]);
  det = i320_sub(det, i320_mul(V[2], t));
  t = i320_det3(B[0], C[0], D[0], B[1], C[1], D[1], B[2], C[2], D[2]);
  det = i320_add(det, i320_mul(V[3], t));
  return i320_sign(det);
}

// Floating-point evaluation of the same 4x4 determinant, with the
// permanent (same expansion on absolute values) for the error bound.
// Inputs differ by less than 2^52 so the differences are exact; the
// squared norms carry a relative error of at most 3 ulp.  The expansion
// has 3 multiplications and at most 6 additions along every path, so the
// computed value is within 16 u of the permanent (u = 2^-53); the filter
// uses 1e-13, a factor of ~60 above that.
__device__ __forceinline__ double det3f(double a0, double a1, double a2,
                                        double b0, double b1, double b2,
                                        double c0, double c1, double c2) {
  return (a0 * (b1 * c2 - b2 * c1) - a1 * (b0 * c2 - b2 * c0)) +
         a2 * (b0 * c1 - b1 * c0);
}
__device__ __forceinline__ double per3f(double a0, double a1, double a2,
                                        double b0, double b1, double b2,
                                        double c0, double c1, double c2) {
  return (a0 * (b1 * c2 + b2 * c1) + a1 * (b0 * c2 + b2 * c0)) +
         a2 * (b0 * c1 + b1 * c0);
}

__device__ __forceinline__ int in_sphere_exact(const int64_t *a,
                                               const int64_t *b,
                                               const int64_t *c,
                                               const int64_t *d,
                                               const int64_t *v) {
  double B[4], C[4], D[4], V[4];
  double aB[4], aC[4], aD[4], aV[4];
  const int64_t *pts[4] = {b, c, d, v};
  double *outs[4] = {B, C, D, V};
  double *aouts[4] = {aB, aC, aD, aV};
#pragma unroll
  for (int k = 0; k < 4; k++) {
    double x = (double)(pts[k][0] - a[0]);
    double y = (double)(pts[k][1] - a[1]);
    double z = (double)(pts[k][2] - a[2]);
    outs[k][0] = x;
    outs[k][1] = y;
    outs[k][2] = z;
    outs[k][3] = (x * x + y * y) + z * z;
    aouts[k][0] = fabs(x);
    aouts[k][1] = fabs(y);
    aouts[k][2] = fabs(z);
    aouts[k][3] = outs[k][3];
  }
  double det =
      ((-V[0] * det3f(B[1], C[1], D[1], B[2], C[2], D[2], B[3], C[3], D[3]) +
        V[1] * det3f(B[0], C[0], D[0], B[2], C[2], D[2], B[3], C[3], D[3])) -
       V[2] * det3f(B[0], C[0], D[0], B[1], C[1], D[1], B[3], C[3], D[3])) +
      V[3] * det3f(B[0], C[0], D[0], B[1], C[1], D[1], B[2], C[2], D[2]);
  double per =
      ((aV[0] * per3f(aB[1], aC[1], aD[1], aB[2], aC[2], aD[2], aB[3], aC[3],
                      aD[3]) +
        aV[1] * per3f(aB[0], aC[0], aD[0], aB[2], aC[2], aD[2], aB[3], aC[3],
                      aD[3])) +
       aV[2] * per3f(aB[0], aC[0], aD[0], aB[1], aC[1], aD[1], aB[3], aC[3],
                     aD[3])) +
      aV[3] * per3f(aB[0], aC[0], aD[0], aB[1], aC[1], aD[1], aB[2], aC[2],
                    aD[2]);
  double bound = 1e-13 * per;
  if (det > bound) return 1;
  if (det < -bound) return -1;
  return in_sphere_exact_int(a, b, c, d, v);
}

}  // namespace mv
