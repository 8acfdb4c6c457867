// swf_dd.h - binary128 expression sites of the reference, emulated in double-double.
//
// The reference is built with `-fdefault-real-8`, which makes every `d`-exponent
// literal REAL(16).  Each expression that contains one is evaluated in IEEE
// binary128 and rounded once to binary64 on assignment.  These functions reproduce
// that final binary64 result with double-double (~106-bit) arithmetic built from
// error-free transformations (Knuth two-sum, FMA two-product), then one
// round-to-nearest of (hi + lo).  The two agree unless the exact value lies within
// ~2^-104 (relative) of a binary64 rounding midpoint while the binary128 evaluation
// falls on the other side; for data of the reference's kind that is a ~2^-50 event
// per site.  Exactly representable cases (equal stencil values, beta == 0,
// short expansions) are computed exactly and therefore match.
//
// Sites (file:line in the reference):
//   weno5.f90:100-102,134-136  smoothness indicators beta(1..3)      -> beta_q
//   weno5.f90:104-106,138-140  alpha = d/(1d-6+beta)**2              -> alpha_q
//   ex4_sedov_2d/main.f90:234,455  w(i)/(w(i)+w(i+1)+1d-15)          -> roe_weight_q
//   ex4_sedov_2d/main.f90:240,461  sqrt(...) + 1d-15                  -> cm_q
//   ex4_sedov_2d/main.f90:426  1./(den+1d-15)                         -> yrecip_q
//   ex4_sedov_2d/main.f90:738  0.75d0*u0 + 0.25d0*x                   -> rk2_q
//   ex4_sedov_2d/main.f90:760  (u0 + 2d0*x)/3d0                       -> rk3_q
//   ex4_sedov_2d/main.f90:112  time <= time_terminal - 1d-10          -> time_continue_q
//
// Header-only, usable from HIP device code and from host C++.
#pragma once

#include <math.h>

#if defined(__HIPCC__)
#define SWF_HD __host__ __device__ __forceinline__
#else
#define SWF_HD static inline
#endif

namespace swf {

struct dd {
  double hi, lo;
};

// Knuth two-sum: s + e == a + b exactly.
SWF_HD dd two_sum(double a, double b) {
  double s = a + b;
  double bb = s - a;
  double e = (a - (s - bb)) + (b - bb);
  return {s, e};
}
// |a| >= |b| (or a == 0)
SWF_HD dd quick_two_sum(double a, double b) {
  double s = a + b;
  double e = b - (s - a);
  return {s, e};
}
SWF_HD dd two_prod(double a, double b) {
  double p = a * b;
  double e = fma(a, b, -p);
  return {p, e};
}
// Sums renormalise with two_sum (no magnitude precondition): under cancellation
// the leading part can be smaller than the trailing one.
SWF_HD dd dd_add(dd a, dd b) {
  dd s = two_sum(a.hi, b.hi);
  dd t = two_sum(a.lo, b.lo);
  double c = s.lo + t.hi;
  dd v = two_sum(s.hi, c);
  double w = t.lo + v.lo;
  return two_sum(v.hi, w);
}
SWF_HD dd dd_add_d(dd a, double b) {
  dd s = two_sum(a.hi, b);
  double v = a.lo + s.lo;
  return two_sum(s.hi, v);
}
SWF_HD dd dd_mul(dd a, dd b) {
  dd p = two_prod(a.hi, b.hi);
  double t = fma(a.hi, b.lo, a.lo * b.hi);
  return quick_two_sum(p.hi, p.lo + t);
}
SWF_HD dd dd_mul_d(dd a, double b) {
  dd p = two_prod(a.hi, b);
  return quick_two_sum(p.hi, fma(a.lo, b, p.lo));
}
SWF_HD dd dd_sqr(dd a) {
  dd p = two_prod(a.hi, a.hi);
  double t = 2.0 * a.hi * a.lo;
  return quick_two_sum(p.hi, p.lo + t);
}
// a / b for double a, double-double b (three-term long division)
SWF_HD dd d_div_dd(double a, dd b) {
  double q1 = a / b.hi;
  // r = a - q1*b
  dd p = dd_mul_d(b, q1);
  dd r = two_sum(a, -p.hi);
  r.lo -= p.lo;
  r = two_sum(r.hi, r.lo);
  double q2 = r.hi / b.hi;
  p = dd_mul_d(b, q2);
  dd r2 = dd_add(r, dd{-p.hi, -p.lo});
  double q3 = r2.hi / b.hi;
  dd q = two_sum(q1, q2);
  return dd_add_d(q, q3);
}
SWF_HD dd dd_div_d(dd a, double b) {
  double q1 = a.hi / b;
  dd p = two_prod(q1, b);
  dd r = dd_add(a, dd{-p.hi, -p.lo});
  double q2 = r.hi / b;
  p = two_prod(q2, b);
  dd r2 = dd_add(r, dd{-p.hi, -p.lo});
  double q3 = r2.hi / b;
  dd q = two_sum(q1, q2);
  return dd_add_d(q, q3);
}
// round-to-nearest binary64 of hi + lo
SWF_HD double dd_round(dd a) { return a.hi + a.lo; }

// binary128 constants as double-double (hi, lo); the residual below lo is
// under 2^-105 relative.
struct qconst {
  static constexpr double c13_12_hi = 0x1.1555555555555p+0, c13_12_lo = 0x1.5555555555555p-54;
  static constexpr double e6_hi = 0x1.0c6f7a0b5ed8dp-20, e6_lo = 0x1.b5a63f9a49c2cp-75;
  static constexpr double e12_hi = 0x1.19799812dea11p-40, e12_lo = 0x1.97f27f0f6e886p-96;
  static constexpr double e15_hi = 0x1.203af9ee75616p-50, e15_lo = -0x1.937831647f5ap-104;
  static constexpr double e10_hi = 0x1.b7cdfd9d7bdbbp-34, e10_lo = -0x1.20a5465df8d2cp-88;
  // 1d-8 rounds to 0x1.5798ee2308c3ap-27 with a negative residual: for binary64 x,
  // x < 1d-8 (binary128)  <=>  x < 0x1.5798ee2308c3ap-27.
  static constexpr double e8_hi = 0x1.5798ee2308c3ap-27;
};

// alpha = d/(1d-6 + beta)**2 with d the binary64 value of the REAL(8) literal;
// (ehi, elo) is the binary128 1d-6 (1d-12 in the weno5.f90 of the step program)
SWF_HD double alpha_q(double d, double beta, double ehi, double elo) {
  dd e = dd_add_d(dd{ehi, elo}, beta);
  return dd_round(d_div_dd(d, dd_sqr(e)));
}

// w(i)/(w(i) + w(i+1) + 1d-15): the first sum is binary64, the rest binary128
SWF_HD double roe_weight_q(double wa, double wb) {
  double s = wa + wb;
  dd den = dd_add_d(dd{qconst::e15_hi, qconst::e15_lo}, s);
  return dd_round(d_div_dd(wa, den));
}

// sqrt(...) + 1d-15
SWF_HD double cm_q(double s) { return dd_round(dd_add_d(dd{qconst::e15_hi, qconst::e15_lo}, s)); }

// 1./(den + 1d-15)
SWF_HD double yrecip_q(double den) {
  dd d = dd_add_d(dd{qconst::e15_hi, qconst::e15_lo}, den);
  return dd_round(d_div_dd(1.0, d));
}

// (u0 + 2d0*x)/3d0
SWF_HD double rk3_q(double u0, double x) {
  dd s = two_sum(u0, 2.0 * x);
  return dd_round(dd_div_d(s, 3.0));
}

// time <= time_terminal - 1d-10 (binary128)
SWF_HD int time_continue_q(double t, double tf) {
  dd s = dd_add(two_sum(tf, -qconst::e10_hi), dd{-qconst::e10_lo, 0.0});
  return (t < s.hi) || (t == s.hi && s.lo >= 0.0);
}

}  // namespace swf
