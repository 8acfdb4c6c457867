// swf_weno.h - WENO-JS reconstructions in first-difference form (FP64 and
// FP32-bulk modes).
//
// The reference's weno_right / weno_left (weno5.f90:86-146) act on five values
// a(-2..2).  Every candidate stencil is consistent (its coefficients sum to 1) and
// the smoothness indicators only see differences, so
//     weno(a) = a(0) + weno(a - a(0)),
// and the right-hand side depends on the four first differences
//     e = (a(-1)-a(-2), a(0)-a(-1), a(1)-a(0), a(2)-a(1))
// alone.  The functions below return that correction.  In these variables
//   beta1: x = e1-e0,    y = e1-3e0      (stencil 0..2)
//   beta2: x = e0-em1,   y = em1+e0      (stencil -1..1)
//   beta3: x = em1-em2,  y = 3em1-em2    (stencil -2..0)
//   beta_k = 13/12 x^2 + 1/4 y^2
// and six times the candidate values minus a(0) are
//   right: 4e0-e1, em1+2e0, 5em1-2em2      (um3, um2, um1 of weno_right)
//   left:  2e1-5e0, -(2em1+e0), em2-4em1   (up3, up2, up1 of weno_left).
// Working with the correction keeps the reconstruction exactly consistent in
// any precision: rounding of the nonlinear weights only scales the (small)
// correction, never a(0).
//
// Weights: alpha_k = d_k/(1e-6+beta_k)^2 normalised.  Both forms scale beta by 4
// (4 beta_k = 13/3 x^2 + y^2, with 4e-6 in place of 1e-6), which leaves the
// weights unchanged.
//   FP64: a common denominator, alpha_k ~ d_k * s_j * s_l with s = (4e-6+4beta)^2,
//         so one reciprocal per reconstruction.
//   FP32: ratios t_k = q_k / min(q) >= 1 of q = 4e-6+4beta, capped at 1e9 (a
//         weight below 1e-18 of the largest is negligible in FP32), then
//         alpha_k ~ d_k (t_j t_l)^2 -- no overflow for any finite beta.
#pragma once

#include <float.h>
#include <math.h>

#include "swf_dd.h"

namespace swf {

// FP32 reciprocal: the hardware estimate on the device, a division on the host
SWF_HD float rcp32(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rcpf(x);
#else
  return 1.0f / x;
#endif
}

// FP64 reciprocal for normal, finite x: the hardware estimate (relative error
// below 2^-29) refined by one Newton step (below 2^-52 before the final
// rounding); a division on the host
SWF_HD double rcp64(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  const double r = __builtin_amdgcn_rcp(x);
  return fma(r, fma(-x, r, 1.0), r);
#else
  return 1.0 / x;
#endif
}

// weno_right(a) - a(0) (RIGHT) or weno_left(a) - a(0), FP64
template <bool RIGHT>
SWF_HD double weno_d64(const double e[4]) {
  const double em2 = e[0], em1 = e[1], e0 = e[2], e1 = e[3];
  const double x1 = e1 - e0, y1 = e1 - 3.0 * e0;
  const double x2 = e0 - em1, y2 = em1 + e0;
  const double x3 = em1 - em2, y3 = 3.0 * em1 - em2;
  const double c = 13.0 / 3.0;
  const double q1 = 4e-6 + ((c * x1) * x1 + y1 * y1);
  const double q2 = 4e-6 + ((c * x2) * x2 + y2 * y2);
  const double q3 = 4e-6 + ((c * x3) * x3 + y3 * y3);
  const double d1 = RIGHT ? 0.3 : 0.1, d2 = 0.6, d3 = RIGHT ? 0.1 : 0.3;
  const double c3 = RIGHT ? 4.0 * e0 - e1 : 2.0 * e1 - 5.0 * e0;
  const double c2 = RIGHT ? em1 + 2.0 * e0 : -(2.0 * em1 + e0);
  const double c1 = RIGHT ? 5.0 * em1 - 2.0 * em2 : em2 - 4.0 * em1;
  const double s1 = q1 * q1, s2 = q2 * q2, s3 = q3 * q3;
  const double a1 = d1 * (s2 * s3), a2 = d2 * (s1 * s3), a3 = d3 * (s1 * s2);
  const double den = (a1 + a2) + a3;
  if (den <= DBL_MAX) return ((a1 * c3 + a2 * c2) + a3 * c1) * rcp64(6.0 * den);
  // beta beyond ~1e77: ratios to the smallest q instead of products
  const double m = fmin(q1, fmin(q2, q3));
  const double r1 = m / q1, r2 = m / q2, r3 = m / q3;
  const double b1 = d1 * (r1 * r1), b2 = d2 * (r2 * r2), b3 = d3 * (r3 * r3);
  return ((b1 * c3 + b2 * c2) + b3 * c1) / (6.0 * ((b1 + b2) + b3));
}

// weno_right(a) - a(0) (RIGHT) or weno_left(a) - a(0), FP32 (inputs rounded
// from FP64 differences).  Non-finite only when a difference exceeds FP32 range
// or beta overflows; callers then use weno_d64.
template <bool RIGHT>
SWF_HD float weno_d32(const float e[4]) {
  const float em2 = e[0], em1 = e[1], e0 = e[2], e1 = e[3];
  const float x1 = e1 - e0, y1 = e1 - 3.0f * e0;
  const float x2 = e0 - em1, y2 = em1 + e0;
  const float x3 = em1 - em2, y3 = 3.0f * em1 - em2;
  const float c = 13.0f / 3.0f;
  const float q1 = 4e-6f + ((c * x1) * x1 + y1 * y1);
  const float q2 = 4e-6f + ((c * x2) * x2 + y2 * y2);
  const float q3 = 4e-6f + ((c * x3) * x3 + y3 * y3);
  const float d1 = RIGHT ? 0.3f : 0.1f, d2 = 0.6f, d3 = RIGHT ? 0.1f : 0.3f;
  const float c3 = RIGHT ? 4.0f * e0 - e1 : 2.0f * e1 - 5.0f * e0;
  const float c2 = RIGHT ? em1 + 2.0f * e0 : -(2.0f * em1 + e0);
  const float c1 = RIGHT ? 5.0f * em1 - 2.0f * em2 : em2 - 4.0f * em1;
  const float sm = rcp32(fminf(q1, fminf(q2, q3)));
  const float t1 = fminf(q1 * sm, 1e9f), t2 = fminf(q2 * sm, 1e9f), t3 = fminf(q3 * sm, 1e9f);
  const float p1 = t2 * t3, p2 = t1 * t3, p3 = t1 * t2;
  const float a1 = d1 * (p1 * p1), a2 = d2 * (p2 * p2), a3 = d3 * (p3 * p3);
  const float inv = rcp32((a1 + a2) + a3) * (1.0f / 6.0f);
  return ((a1 * inv) * c3 + (a2 * inv) * c2) + (a3 * inv) * c1;
}

// FP32 form when every |difference| is at most 1e18 (then 4 beta < 1e38 and the
// weight ratios stay in range), the FP64 form otherwise
template <bool RIGHT>
SWF_HD double weno_d32_or_64(const double e[4]) {
  const float ef[4] = {(float)e[0], (float)e[1], (float)e[2], (float)e[3]};
  const float big = fmaxf(fmaxf(fabsf(ef[0]), fabsf(ef[1])), fmaxf(fabsf(ef[2]), fabsf(ef[3])));
  if (big <= 1e18f) return (double)weno_d32<RIGHT>(ef);
  return weno_d64<RIGHT>(e);
}

// the four first differences of a(-2..2)
SWF_HD void first_differences(const double *a, double e[4]) {
  e[0] = a[-1] - a[-2];
  e[1] = a[0] - a[-1];
  e[2] = a[1] - a[0];
  e[3] = a[2] - a[1];
}

}  // namespace swf
