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
// (4 beta_k = 13/3 x^2 + y^2, with q4 = 4e-6 in place of 1e-6), which leaves the
// weights unchanged.  The step diffraction program's weno5.f90 has 1e-12 (q4 = 4e-12).
//   FP64: a common denominator, alpha_k ~ d_k * s_j * s_l with s = (4e-6+4beta)^2,
//         so one reciprocal per reconstruction.
//   FP32: alpha_k = d_k (min(q)/q_k)^2 with q = 4e-6+4beta (no overflow for any
//         finite beta; negligible weights underflow to 0); the candidates and
//         the normalised weighted sum stay FP64.
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

// FP64 square root for x = 0 and x in [2^-900, 2^900]: the hardware reciprocal
// square root y, then one Goldschmidt step on (g, h) = (x y, y / 2) and one
// correction with the residual x - g^2 (within 1 ulp); the library routine on the
// host.  The library's device routine adds range scaling and a second correction.
SWF_HD double sqrt64(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y, h = 0.5 * y;
  const double r = fma(-h, g, 0.5);
  g = fma(g, r, g);
  h = fma(h, r, h);
  const double d = fma(-g, g, x);
  return (x == 0.0) ? 0.0 : fma(d, h, g);
#else
  return sqrt(x);
#endif
}

// weno_right(a) - a(0) (RIGHT) or weno_left(a) - a(0), FP64
template <bool RIGHT>
SWF_HD double weno_d64(const double e[4], double q4) {
  const double em2 = e[0], em1 = e[1], e0 = e[2], e1 = e[3];
  const double x1 = e1 - e0, y1 = e1 - 3.0 * e0;
  const double x2 = e0 - em1, y2 = em1 + e0;
  const double x3 = em1 - em2, y3 = 3.0 * em1 - em2;
  const double c = 13.0 / 3.0;
  const double q1 = q4 + ((c * x1) * x1 + y1 * y1);
  const double q2 = q4 + ((c * x2) * x2 + y2 * y2);
  const double q3 = q4 + ((c * x3) * x3 + y3 * y3);
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

// Unnormalised nonlinear weights in FP32 from the smoothness indicators'
// arguments x_k, y_k in FP32: alpha_k = d_k (min(q)/q_k)^2 with
// q = 4e-6 + 4 beta.  Every alpha_k carries its own relative rounding error only;
// the weight of a stencil across a jump underflows to 0 as its exact value would
// be negligible.  The smallest-q stencil has alpha = d_k >= 0.1, so the sum never
// vanishes.  Requires max |x|, |y| <= 1e18 (then 4 beta < 1e38).
template <bool RIGHT>
SWF_HD void alphas_d32(const float x[3], const float y[3], float a[3], float q4) {
  const float c = 13.0f / 3.0f;
  const float q1 = q4 + ((c * x[0]) * x[0] + y[0] * y[0]);
  const float q2 = q4 + ((c * x[1]) * x[1] + y[1] * y[1]);
  const float q3 = q4 + ((c * x[2]) * x[2] + y[2] * y[2]);
  const float m = fminf(q1, fminf(q2, q3));
  const float r1 = m * rcp32(q1), r2 = m * rcp32(q2), r3 = m * rcp32(q3);
  a[0] = (RIGHT ? 0.3f : 0.1f) * (r1 * r1);
  a[1] = 0.6f * (r2 * r2);
  a[2] = (RIGHT ? 0.1f : 0.3f) * (r3 * r3);
}

// weno_right(a) - a(0) (RIGHT) or weno_left(a) - a(0), FP32-bulk: smoothness
// indicators and nonlinear weights in FP32; the candidate values and the
// normalised weighted sum sum(alpha_k c_k) / sum(alpha_k) in FP64 (the FP32
// weights convert exactly), so the result is a weighted mean of the FP64
// candidates with each weight perturbed by ~1e-7 of itself.  Falls back to
// weno_d64 beyond FP32 range.
template <bool RIGHT>
SWF_HD double weno_d32_or_64(const double e[4], double q4, float q4f) {
  const double em2 = e[0], em1 = e[1], e0 = e[2], e1 = e[3];
  // the smoothness arguments in FP32 from the FP32 differences; |e| <= 2.5e17
  // keeps |x|, |y| <= 1e18
  const float f0 = (float)em2, f1 = (float)em1, f2 = (float)e0, f3 = (float)e1;
  const float big = fmaxf(fmaxf(fabsf(f0), fabsf(f1)), fmaxf(fabsf(f2), fabsf(f3)));
  if (!(big <= 2.5e17f)) return weno_d64<RIGHT>(e, q4);
  const float x[3] = {f3 - f2, f2 - f1, f1 - f0}, y[3] = {f3 - 3.0f * f2, f1 + f2, 3.0f * f1 - f0};
  float al[3];
  alphas_d32<RIGHT>(x, y, al, q4f);
  const double a1 = al[0], a2 = al[1], a3 = al[2];
  const double c3 = RIGHT ? 4.0 * e0 - e1 : 2.0 * e1 - 5.0 * e0;
  const double c2 = RIGHT ? em1 + 2.0 * e0 : -(2.0 * em1 + e0);
  const double c1 = RIGHT ? 5.0 * em1 - 2.0 * em2 : em2 - 4.0 * em1;
  return ((a1 * c3 + a2 * c2) + a3 * c1) * rcp64(6.0 * ((a1 + a2) + a3));
}

// weno_right(a) - a(0) (RIGHT) or weno_left(a) - a(0) with everything in FP32: the
// differences e, the smoothness indicators and weights (as weno_d32_or_64), the
// candidate values and their weighted sum.  Beyond FP32 range of the indicators
// (|e| > 2.5e17) the FP64 form of the same differences.
template <bool RIGHT>
SWF_HD float weno_d32f(const float e[4], double q4, float q4f) {
  const float em2 = e[0], em1 = e[1], e0 = e[2], e1 = e[3];
  const float big = fmaxf(fmaxf(fabsf(em2), fabsf(em1)), fmaxf(fabsf(e0), fabsf(e1)));
  if (!(big <= 2.5e17f)) {
    const double d[4] = {em2, em1, e0, e1};
    return (float)weno_d64<RIGHT>(d, q4);
  }
  const float x[3] = {e1 - e0, e0 - em1, em1 - em2}, y[3] = {e1 - 3.0f * e0, em1 + e0, 3.0f * em1 - em2};
  float al[3];
  alphas_d32<RIGHT>(x, y, al, q4f);
  const float c3 = RIGHT ? 4.0f * e0 - e1 : 2.0f * e1 - 5.0f * e0;
  const float c2 = RIGHT ? em1 + 2.0f * e0 : -(2.0f * em1 + e0);
  const float c1 = RIGHT ? 5.0f * em1 - 2.0f * em2 : em2 - 4.0f * em1;
  return ((al[0] * c3 + al[1] * c2) + al[2] * c1) * rcp32(6.0f * ((al[0] + al[1]) + al[2]));
}

// the four first differences of a(-2..2)
SWF_HD void first_differences(const double *a, double e[4]) {
  e[0] = a[-1] - a[-2];
  e[1] = a[0] - a[-1];
  e[2] = a[1] - a[0];
  e[3] = a[2] - a[1];
}

}  // namespace swf
