// swf_q128.h - minimal IEEE binary128 arithmetic (add, multiply, conversions)
// with round-to-nearest-even, in 128-bit integer arithmetic.  Host and device.
//
// Used as the exact slow path of the binary128 expression sites (swf_dd.h): when
// a double-double result lies too close to a binary64 rounding boundary to decide
// the final rounding, the site is re-evaluated here operation by operation, each
// result rounded to 113 significant bits exactly as binary128 hardware or
// libquadmath would.  Finite values only; results are assumed to stay in the
// binary64 normal range on conversion back (the reference's data do).
#pragma once

#include <stdint.h>
#include <string.h>

#include "swf_dd.h"

namespace swf {

typedef unsigned __int128 u128;

// value = (-1)^neg * m * 2^(e - 112), m in [2^112, 2^113), or zero
struct q128 {
  u128 m;
  int e;
  int neg;
  int zero;
};

SWF_HD int clz128(u128 x) {
  const uint64_t hi = (uint64_t)(x >> 64), lo = (uint64_t)x;
  if (hi) return __builtin_clzll(hi);
  if (lo) return 64 + __builtin_clzll(lo);
  return 128;
}

SWF_HD q128 q_from_double(double d) {
  q128 r;
  r.neg = 0;
  r.e = 0;
  r.m = 0;
  r.zero = 0;
  uint64_t bits;
  memcpy(&bits, &d, 8);
  r.neg = (int)(bits >> 63);
  const int ex = (int)((bits >> 52) & 0x7ff);
  uint64_t man = bits & ((1ull << 52) - 1);
  if (ex == 0 && man == 0) {
    r.zero = 1;
    return r;
  }
  int e2;
  if (ex == 0) {  // subnormal: value = man * 2^-1074
    const int lz = __builtin_clzll(man) - 11;  // bring the leading 1 to bit 52
    man <<= lz;
    e2 = -1022 - lz;
  } else {
    man |= (1ull << 52);
    e2 = ex - 1023;
  }
  r.m = (u128)man << 60;
  r.e = e2;
  return r;
}

// round a mantissa with the leading one at bit `top` (>= 112) to 113 bits, RNE,
// with an extra sticky flag for bits already shifted out
SWF_HD q128 q_round(int neg, int e, u128 m, int top, int sticky_in) {
  q128 r;
  r.neg = neg;
  r.zero = 0;
  const int sh = top - 112;
  u128 mm = m;
  if (sh > 0) {
    const u128 half = (u128)1 << (sh - 1);
    const u128 rem = m & (((u128)1 << sh) - 1);
    mm = m >> sh;
    if (rem > half || (rem == half && (sticky_in || (mm & 1)))) mm += 1;
    if (mm >> 113) {
      mm >>= 1;
      e += 1;
    }
  }
  r.m = mm;
  r.e = e;
  return r;
}

SWF_HD q128 q_neg(q128 a) {
  a.neg ^= 1;
  return a;
}

// a + b, correctly rounded
SWF_HD q128 q_add(q128 a, q128 b) {
  if (a.zero) return b;
  if (b.zero) return a;
  // |a| >= |b|
  if (b.e > a.e || (b.e == a.e && b.m > a.m)) {
    q128 t = a;
    a = b;
    b = t;
  }
  const int d = a.e - b.e;
  // three guard bits: leading one at bit 115
  u128 ma = a.m << 3, mb = b.m << 3;
  int sticky = 0;
  if (d > 0) {
    if (d >= 120) {
      sticky = (mb != 0);
      mb = 0;
    } else {
      sticky = (mb & (((u128)1 << d) - 1)) != 0;
      mb >>= d;
    }
  }
  u128 m;
  if (a.neg == b.neg) {
    m = ma + mb;
  } else {
    m = ma - mb;
    if (sticky) m -= 1;  // the discarded bits of b make the difference smaller
  }
  if (m == 0 && !sticky) {
    q128 z;
    z.zero = 1;
    z.neg = 0;
    z.e = 0;
    z.m = 0;
    return z;
  }
  const int top = 127 - clz128(m);
  int e = a.e + (top - 115);
  if (top < 115) {
    // cancellation: bring the leading one back to bit 115 (exact when d <= 1)
    m <<= (115 - top);
    return q_round(a.neg, e, m, 115, sticky);
  }
  return q_round(a.neg, e, m, top, sticky);
}

// a * b, correctly rounded
SWF_HD q128 q_mul(q128 a, q128 b) {
  if (a.zero || b.zero) {
    q128 z;
    z.zero = 1;
    z.neg = a.neg ^ b.neg;
    z.e = 0;
    z.m = 0;
    return z;
  }
  const uint64_t ah = (uint64_t)(a.m >> 64), al = (uint64_t)a.m;
  const uint64_t bh = (uint64_t)(b.m >> 64), bl = (uint64_t)b.m;
  const u128 ll = (u128)al * bl, lh = (u128)al * bh, hl = (u128)ah * bl, hh = (u128)ah * bh;
  // 256-bit product (hi:lo)
  u128 lo = ll;
  u128 mid = lh + hl;                       // < 2^114, no overflow
  u128 hi = hh + (mid >> 64);
  const u128 mid_lo = mid << 64;
  lo += mid_lo;
  if (lo < mid_lo) hi += 1;
  // product in [2^224, 2^226): keep the top 116 bits in a 128-bit word
  const int top256 = (hi >> (225 - 128)) ? 225 : 224;
  const int sh = top256 - 115;  // right shift of the 256-bit value
  // m = (hi:lo) >> sh, sticky = any bit below
  const u128 m = (hi << (128 - sh)) | (lo >> sh);
  const int sticky = (lo & (((u128)1 << sh) - 1)) != 0;
  const int e = a.e + b.e + (top256 - 224);
  return q_round(a.neg ^ b.neg, e, m, 115, sticky);
}

// round to binary64 (normal range)
SWF_HD double q_to_double(q128 a) {
  if (a.zero) return a.neg ? -0.0 : 0.0;
  const u128 rem = a.m & (((u128)1 << 60) - 1);
  const u128 half = (u128)1 << 59;
  uint64_t man = (uint64_t)(a.m >> 60);
  int e = a.e;
  if (rem > half || (rem == half && (man & 1))) man += 1;
  if (man >> 53) {
    man >>= 1;
    e += 1;
  }
  double r = ldexp((double)man, e - 52);
  return a.neg ? -r : r;
}

// 13d0/12d0 rounded to binary128: mantissa round(13/12 * 2^112)
SWF_HD q128 q_13_12() {
  q128 r;
  r.zero = 0;
  r.neg = 0;
  r.e = 0;
  r.m = ((u128)0x1155555555555ull << 64) | (u128)0x5555555555555555ull;
  return r;
}

// binary128 evaluation of one smoothness indicator of weno5.f90:100-102
//   which 0: 13/12*(a0-2a1+a2)**2 + 0.25*(3a0-4a1+a2)**2
//   which 1: 13/12*(a-1-2a0+a1)**2 + 0.25*(binary64 (a-1-a1)**2)
//   which 2: 13/12*(a-2-2a-1+a0)**2 + 0.25*(a-2-4a-1+3a0)**2
SWF_HD double beta_q128(const double *a, int which) {
  const q128 c = q_13_12();
  q128 x, y;
  if (which == 0) {
    x = q_add(q_add(q_from_double(a[0]), q_neg(q_from_double(2.0 * a[1]))), q_from_double(a[2]));
    y = q_add(q_add(q_mul(q_from_double(3.0), q_from_double(a[0])), q_neg(q_from_double(4.0 * a[1]))),
              q_from_double(a[2]));
  } else if (which == 1) {
    x = q_add(q_add(q_from_double(a[-1]), q_neg(q_from_double(2.0 * a[0]))), q_from_double(a[1]));
    const double yd = a[-1] - a[1];
    y = q_from_double(yd * yd);  // binary64 square, promoted
  } else {
    x = q_add(q_add(q_from_double(a[-2]), q_neg(q_from_double(2.0 * a[-1]))), q_from_double(a[0]));
    y = q_add(q_add(q_from_double(a[-2]), q_neg(q_from_double(4.0 * a[-1]))),
              q_mul(q_from_double(3.0), q_from_double(a[0])));
  }
  const q128 t1 = q_mul(c, q_mul(x, x));
  q128 t2;
  if (which == 1) {
    t2 = y;
    t2.e -= 2;  // 0.25 * y, exact
  } else {
    t2 = q_mul(y, y);
    t2.e -= 2;
  }
  return q_to_double(q_add(t1, t2));
}

// 0.75d0*u0 + 0.25d0*x in binary128
SWF_HD double rk2_q128(double u0, double x) {
  q128 a = q_mul(q_from_double(0.75), q_from_double(u0));
  q128 b = q_from_double(x);
  b.e -= 2;
  return q_to_double(q_add(a, b));
}

// true when the double-double value v cannot safely be rounded to binary64:
// it lies within 2^-100 |v.hi| of a rounding midpoint of v.hi's neighbourhood
SWF_HD bool dd_near_midpoint(dd v) {
  if (v.hi == 0.0) return false;
  int e;
  frexp(v.hi, &e);
  const double ulp = ldexp(1.0, e - 53);   // spacing above |v.hi|
  const double margin = ldexp(fabs(v.hi), -100);
  const double lo = (v.hi > 0) ? v.lo : -v.lo;  // oriented away from zero
  const double frac = fabs(v.hi);
  // spacing below |v.hi| halves when |v.hi| is a power of two
  const double ulp_dn = (frac == ldexp(1.0, e - 1)) ? 0.5 * ulp : ulp;
  return fabs(lo - 0.5 * ulp) < margin || fabs(lo + 0.5 * ulp_dn) < margin;
}

// beta(1..3) of weno5.f90:100-102 (= :134-136), `a` pointing at average(0):
// double-double evaluation, with the exact binary128 path when the result is
// within 2^-100 of a binary64 rounding boundary.
SWF_HD void beta_q(const double *a, double beta[3]) {
  const dd c1312 = {qconst::c13_12_hi, qconst::c13_12_lo};
  dd t[3];
  {
    dd x = dd_add_d(two_sum(a[0], -2.0 * a[1]), a[2]);
    dd y = dd_add_d(dd_add_d(two_prod(3.0, a[0]), -4.0 * a[1]), a[2]);
    t[0] = dd_add(dd_mul(c1312, dd_sqr(x)), dd_mul_d(dd_sqr(y), 0.25));
  }
  {
    dd x = dd_add_d(two_sum(a[-1], -2.0 * a[0]), a[1]);
    double yd = a[-1] - a[1];
    double y2 = yd * yd;
    t[1] = dd_add_d(dd_mul(c1312, dd_sqr(x)), 0.25 * y2);
  }
  {
    dd x = dd_add_d(two_sum(a[-2], -2.0 * a[-1]), a[0]);
    dd y = dd_add(two_sum(a[-2], -4.0 * a[-1]), two_prod(3.0, a[0]));
    t[2] = dd_add(dd_mul(c1312, dd_sqr(x)), dd_mul_d(dd_sqr(y), 0.25));
  }
#pragma unroll
  for (int m = 0; m < 3; ++m) beta[m] = dd_near_midpoint(t[m]) ? beta_q128(a, m) : dd_round(t[m]);
}

// 0.75d0*u0 + 0.25d0*x (ex4_sedov_2d/main.f90:738); both products are exact in
// binary128, the sum is rounded once to 113 bits then to binary64.
SWF_HD double rk2_q(double u0, double x) {
  const dd r = dd_add_d(two_prod(0.75, u0), 0.25 * x);
  return dd_near_midpoint(r) ? rk2_q128(u0, x) : dd_round(r);
}

}  // namespace swf
