// Host check of the library's double-double emulation of the reference's
// binary128 expression sites (sweeping_fd_weno_amd/csrc/swf_dd.h) against
// __float128 evaluation of the same expressions.  Built and run by
// tests/test_dd_emulation.py; prints "site mismatches samples" per site.
#include <math.h>
#include <quadmath.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "swf_q128.h"

typedef __float128 q;
#define Q(x) ((q)(x))
static double D(q x) { return (double)x; }

static uint64_t s_state = 88172645463325252ull;
static uint64_t rnd() {
  s_state ^= s_state << 13;
  s_state ^= s_state >> 7;
  s_state ^= s_state << 17;
  return s_state;
}
static double uni() { return (rnd() >> 11) * 0x1.0p-53; }
static double gauss() { return sqrt(-2.0 * log(uni() + 1e-300)) * cos(6.283185307179586 * uni()); }
// values of the kind the sweep and WENO see: wide exponents, near-equal groups
static double value(int kind, double base) {
  switch (kind) {
    case 0: return gauss();
    case 1: return base * (1.0 + gauss() * ldexp(1.0, -(int)(rnd() % 50)));
    case 2: return gauss() * pow(10.0, (double)((int)(rnd() % 60) - 30));
    default: return (double)((int64_t)(rnd() % 64) - 32) / 8.0;
  }
}

static void beta_ref(const double *a, double b[3]) {
  const q c = Q(13) / Q(12);
  q x = (Q(a[0]) - Q(2) * Q(a[1])) + Q(a[2]);
  q y = (Q(3) * Q(a[0]) - Q(4) * Q(a[1])) + Q(a[2]);
  b[0] = D(c * (x * x) + Q(0.25) * (y * y));
  x = (Q(a[-1]) - Q(2) * Q(a[0])) + Q(a[1]);
  double yd = a[-1] - a[1];
  b[1] = D(c * (x * x) + Q(0.25) * Q(yd * yd));
  x = (Q(a[-2]) - Q(2) * Q(a[-1])) + Q(a[0]);
  y = (Q(a[-2]) - Q(4) * Q(a[-1])) + Q(3) * Q(a[0]);
  b[2] = D(c * (x * x) + Q(0.25) * (y * y));
}

static int neq(double a, double b) { return memcmp(&a, &b, 8) != 0; }

int main(int argc, char **argv) {
  long n = argc > 1 ? atol(argv[1]) : 1000000;
  long kb[4] = {0, 0, 0, 0}, kr[4] = {0, 0, 0, 0}, bad_q128 = 0;
  long bad_alpha12 = 0;
  long bad_beta = 0, bad_alpha = 0, bad_roe = 0, bad_cm = 0, bad_yr = 0, bad_rk2 = 0, bad_rk3 = 0, bad_t = 0;
  for (long it = 0; it < n; ++it) {
    const int kind = (int)(it % 4);
    const double base = gauss() * pow(10.0, (double)((int)(rnd() % 20) - 10));
    double s[5];
    for (int m = 0; m < 5; ++m) s[m] = value(kind, base);
    double b1[3], b2[3];
    swf::beta_q(&s[2], b1);
    beta_ref(&s[2], b2);
    for (int m = 0; m < 3; ++m) {
      bad_beta += neq(b1[m], b2[m]);
      kb[kind] += neq(b1[m], b2[m]);
      bad_q128 += neq(swf::beta_q128(&s[2], m), b2[m]);  // the exact path on its own
    }
    const double dlin[3] = {0.1, 0.6, 0.3};
    for (int m = 0; m < 3; ++m) {
      const q e = 1e-6Q + Q(b2[m]);
      bad_alpha += neq(swf::alpha_q(dlin[m], b2[m], swf::qconst::e6_hi, swf::qconst::e6_lo), D(Q(dlin[m]) / (e * e)));
      // the step program's weno5.f90: 1d-12 (its betas are near 1d-12 for small data)
      const double bs = b2[m] * 1e-6;
      const q e12 = 1e-12Q + Q(bs);
      bad_alpha12 +=
          neq(swf::alpha_q(dlin[m], bs, swf::qconst::e12_hi, swf::qconst::e12_lo), D(Q(dlin[m]) / (e12 * e12)));
    }
    const double wa = fabs(value(kind, fabs(base) + 1e-300)), wb = fabs(value(kind, fabs(base) + 1e-300));
    bad_roe += neq(swf::roe_weight_q(wa, wb), D(Q(wa) / (Q(wa + wb) + 1e-15Q)));
    bad_cm += neq(swf::cm_q(wa), D(Q(wa) + 1e-15Q));
    bad_yr += neq(swf::yrecip_q(wa), D(Q(1.) / (Q(wa) + 1e-15Q)));
    const double u0 = value(kind, base), x = value(kind, base);
    bad_rk2 += neq(swf::rk2_q(u0, x), D(Q(0.75) * Q(u0) + Q(0.25) * Q(x)));
    kr[kind] += neq(swf::rk2_q(u0, x), D(Q(0.75) * Q(u0) + Q(0.25) * Q(x)));
    bad_q128 += neq(swf::rk2_q128(u0, x), D(Q(0.75) * Q(u0) + Q(0.25) * Q(x)));
    bad_rk3 += neq(swf::rk3_q(u0, x), D((Q(u0) + Q(2) * Q(x)) / Q(3)));
    const double tf = uni() * 2.0, t = tf - 1e-10 * (uni() * 2.0);
    bad_t += (swf::time_continue_q(t, tf) != (Q(t) <= Q(tf) - 1e-10Q));
  }
  printf("alpha12 %ld %ld\n", bad_alpha12, 3 * n);
  printf("beta %ld %ld\nalpha %ld %ld\nroe_weight %ld %ld\ncm %ld %ld\nyrecip %ld %ld\nrk2 %ld %ld\nrk3 %ld %ld\n"
         "time_continue %ld %ld\nq128_path %ld %ld\n",
         bad_beta, 3 * n, bad_alpha, 3 * n, bad_roe, n, bad_cm, n, bad_yr, n, bad_rk2, n, bad_rk3, n, bad_t, n, bad_q128, 4 * n);
  printf("by input kind (gauss, near-equal, wide-exponent, eighths): beta %ld %ld %ld %ld  rk2 %ld %ld %ld %ld\n",
         kb[0], kb[1], kb[2], kb[3], kr[0], kr[1], kr[2], kr[3]);
  return 0;
}
