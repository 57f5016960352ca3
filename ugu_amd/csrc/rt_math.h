// Ray-triangle and ray-box tests of the raytracer, shared by the HIP kernels
// and host-side checks of the hierarchy (compiled by hipcc for the device and
// by g++ for the host; both with -ffp-contract=off). V4 is any type with
// float members x, y, z (float4 on the device).
#pragma once

#include <math.h>

#ifdef __HIPCC__
#define RT_HD __host__ __device__ __forceinline__
#else
#define RT_HD inline
#endif

#pragma clang fp contract(off)

namespace ugu_hip {

// Eigen's dot of two float64 3-vectors. With SSE2 Eigen reduces a 3-vector
// of doubles with one Packet2d for coefficients 0..1 and a scalar tail
// (redux_impl, LinearVectorizedTraversal): (x0 + x1) + x2.
RT_HD double dot_d(double a0, double a1, double a2, double b0, double b1, double b2) {
  return (a0 * b0 + a1 * b1) + a2 * b2;
}

// IntersectImpl (geom_util.cc:60-115) on float vertices promoted to double.
// Returns true with the float casts of (t, u, v) on a hit.
template <typename V4>
RT_HD bool mt_hit(double ox, double oy, double oz, double dx, double dy, double dz, const V4& p0,
                  const V4& p1, const V4& p2, float& t, float& u, float& v) {
  const double eps = static_cast<double>(1e-8f);
  const double v0x = p0.x, v0y = p0.y, v0z = p0.z;
  const double e1x = double(p1.x) - v0x, e1y = double(p1.y) - v0y, e1z = double(p1.z) - v0z;
  const double e2x = double(p2.x) - v0x, e2y = double(p2.y) - v0y, e2z = double(p2.z) - v0z;
  // alpha = ray x e2
  const double ax = dy * e2z - dz * e2y;
  const double ay = dz * e2x - dx * e2z;
  const double az = dx * e2y - dy * e2x;
  const double det = dot_d(e1x, e1y, e1z, ax, ay, az);
  if (-eps < det && det < eps) return false;
  const double inv_det = 1.0 / det;
  const double rx = ox - v0x, ry = oy - v0y, rz = oz - v0z;
  const double uu = dot_d(ax, ay, az, rx, ry, rz) * inv_det;
  if (uu < 0.0 || uu > 1.0) return false;
  // beta = r x e1
  const double bx = ry * e1z - rz * e1y;
  const double by = rz * e1x - rx * e1z;
  const double bz = rx * e1y - ry * e1x;
  const double vv = dot_d(dx, dy, dz, bx, by, bz) * inv_det;
  if (vv < 0.0 || uu + vv > 1.0) return false;
  const double tt = dot_d(e2x, e2y, e2z, bx, by, bz) * inv_det;
  if (tt < 0.0) return false;
  const double w = 1.0 - uu - vv;
  t = static_cast<float>(tt);
  u = static_cast<float>(w);
  v = static_cast<float>(uu);
  return true;
}

// Relative slack of the box test (2^-20): covers the float rounding of the
// slab arithmetic (a few ulp).
constexpr float kBoxSlack = 9.5367431640625e-07f;

// End of the interval a ray still searches once it holds a hit at t: t
// widened by the slack (t + t 2^-20, one rounding).
RT_HD float box_tmax(float t) { return t + t * kBoxSlack; }

// Slack applied to a slab interval [tn, tf] (tn >= 0): the box is entered at
// tn (1 - 2^-20) if that is <= tf (1 + 2^-20), else missed (+inf). Both
// products are single roundings of tn - tn 2^-20 and tf + tf 2^-20 (the
// scaled terms are exact); when tf < 0 both forms miss.
RT_HD float slab_entry(float tn, float tf) {
  const float tn_lo = tn * (1.0f - kBoxSlack);
  const float tf_hi = tf * (1.0f + kBoxSlack);
  return tn_lo <= tf_hi ? tn_lo : __builtin_huge_valf();
}

// Float slab test of a widened box. Returns the interval entry (>= 0) or
// +inf when the box is missed or lies beyond tmax.
template <typename V4>
RT_HD float box_entry(const V4& lo, const V4& hi, float ox, float oy, float oz, float ix,
                      float iy, float iz, float tmax) {
  const float tx0 = (lo.x - ox) * ix, tx1 = (hi.x - ox) * ix;
  const float ty0 = (lo.y - oy) * iy, ty1 = (hi.y - oy) * iy;
  const float tz0 = (lo.z - oz) * iz, tz1 = (hi.z - oz) * iz;
  // fminf/fmaxf return the other operand for NaN ((b - o) * inf with b == o);
  // that happens only where the ray runs along a widened face, which holds no
  // geometry.
  const float tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), 0.0f));
  const float tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), tmax));
  return slab_entry(tn, tf);
}

}  // namespace ugu_hip
