// Host check of the first-difference WENO forms (FP64, and FP32-bulk with its FP64 guard) (sweeping_fd_weno_amd/csrc/swf_weno.h)
// against weno_right / weno_left of weno5.f90:86-146 evaluated in binary128
// throughout.  Prints one line per (form, side, input kind):
//     <form> <side> <kind> <n> <max err / scale>
// where scale = max|a(k)-a(0)| (the size of the correction), and a last line
//     rl identity all <n> <max err>
// for the closed form of R * L used by the fast modes' centre value.  A consistent form
// keeps this ratio at rounding level of its own precision however large |a(0)|
// is; a wrong coefficient or sign makes it O(1).
#include <math.h>
#include <quadmath.h>
#include <stdio.h>
#include <stdlib.h>

#include <random>

#include "swf_weno.h"

using q = __float128;

// weno5.f90 formulas, every operation in binary128
static q weno_ref(const double *a, bool right) {
  q A[5];
  for (int i = 0; i < 5; ++i) A[i] = a[i - 2];
  const q am2 = A[0], am1 = A[1], a0 = A[2], a1 = A[3], a2 = A[4];
  q b1 = (q)13 / 12 * (a0 - 2 * a1 + a2) * (a0 - 2 * a1 + a2) + (q)0.25Q * (3 * a0 - 4 * a1 + a2) * (3 * a0 - 4 * a1 + a2);
  q b2 = (q)13 / 12 * (am1 - 2 * a0 + a1) * (am1 - 2 * a0 + a1) + (q)0.25Q * (am1 - a1) * (am1 - a1);
  q b3 = (q)13 / 12 * (am2 - 2 * am1 + a0) * (am2 - 2 * am1 + a0) +
         (q)0.25Q * (am2 - 4 * am1 + 3 * a0) * (am2 - 4 * am1 + 3 * a0);
  q d1 = right ? 0.3Q : 0.1Q, d2 = 0.6Q, d3 = right ? 0.1Q : 0.3Q;
  q e = 1e-6Q;
  q al1 = d1 / ((e + b1) * (e + b1)), al2 = d2 / ((e + b2) * (e + b2)), al3 = d3 / ((e + b3) * (e + b3));
  q s = al1 + al2 + al3;
  q c1, c2, c3;  // um1..um3 (right) or up1..up3 (left)
  if (right) {
    c1 = (2 * am2 - 7 * am1 + 11 * a0) / 6;
    c2 = (-am1 + 5 * a0 + 2 * a1) / 6;
    c3 = (2 * a0 + 5 * a1 - a2) / 6;
  } else {
    c1 = (-am2 + 5 * am1 + 2 * a0) / 6;
    c2 = (2 * am1 + 5 * a0 - a1) / 6;
    c3 = (11 * a0 - 7 * a1 + 2 * a2) / 6;
  }
  return (al1 * c3 + al2 * c2 + al3 * c1) / s;
}

// R * L of the reference's eigenvectors (ex4_sedov_2d/main.f90:242-280, x form)
// for a random Roe state, in binary128, against I + dH e_E (L row 1 + L row 4)
// with dH = (hm - qm) - cm^2/gamma1; returns the largest entry difference
// relative to the largest entry of R * L.
static double rl_identity_error(std::mt19937_64 &rng) {
  std::uniform_real_distribution<double> u(-1.0, 1.0);
  const q g1 = 0.4Q;
  const q vx = 3 * u(rng), vy = 3 * u(rng);
  const q qm = 0.5Q * (vx * vx + vy * vy);
  const q c2 = powq(10.0Q, 6 * u(rng));  // sound speed^2 over 12 decades
  const q hm = c2 / g1 + qm + (q)(1e-3 * u(rng)) * c2;  // H off the consistent value
  const q cm = sqrtq(fabsq(g1 * (hm - qm))) + 1e-15Q;
  const q rcm = 1 / cm, b1 = g1 * rcm * rcm, b2 = qm * b1, t0 = vx * rcm, t1 = b1 * vx, t2 = 0.5Q * b1,
          t3 = b1 * vy, tc = vx * cm;
  const q R[4][4] = {{1, 0, 1, 1}, {vx - cm, 0, vx, vx + cm}, {vy, 1, vy, vy}, {hm - tc, vy, qm, hm + tc}};
  const q L[4][4] = {{0.5Q * (b2 + t0), -0.5Q * (t1 + rcm), -0.5Q * t3, t2},
                     {-vy, 0, 1, 0},
                     {1 - b2, t1, t3, -b1},
                     {0.5Q * (b2 - t0), -0.5Q * (t1 - rcm), -0.5Q * t3, t2}};
  const q dH = (hm - qm) - cm * cm / g1;
  const q sumrow[4] = {b2, -t1, -t3, b1};
  q big = 0, err = 0;
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) {
      q rl = 0;
      for (int m = 0; m < 4; ++m) rl += R[r][m] * L[m][c];
      const q model = (r == c ? 1 : 0) + (r == 3 ? dH * sumrow[c] : 0);
      big = fmaxq(big, fabsq(rl));
      err = fmaxq(err, fabsq(rl - model));
    }
  return (double)(err / big);
}

int main(int argc, char **argv) {
  const long n = argc > 1 ? atol(argv[1]) : 200000;
  std::mt19937_64 rng(20240611);
  std::normal_distribution<double> g(0.0, 1.0);
  std::uniform_real_distribution<double> u(-1.0, 1.0);
  const char *kinds[] = {"gaussian", "offset1e8", "smooth", "jump", "wide", "tiny"};
  for (int kind = 0; kind < 6; ++kind) {
    double m64[2] = {0, 0}, m32[2] = {0, 0};
    for (long t = 0; t < n; ++t) {
      double s[5];
      switch (kind) {
        case 0:
          for (double &v : s) v = g(rng);
          break;
        case 1:  // large common offset: the correction is tiny relative to a(0)
          for (double &v : s) v = 1e8 + g(rng);
          break;
        case 2: {  // samples of a smooth function: beta well below eps
          const double x0 = u(rng), h = 1e-3 * (1 + u(rng));
          for (int i = 0; i < 5; ++i) s[i] = sin(3 * (x0 + (i - 2) * h)) + 2;
          break;
        }
        case 3: {  // a jump somewhere in the stencil
          const int at = (int)(rng() % 5);
          const double hi = pow(10.0, 6 * u(rng));
          for (int i = 0; i < 5; ++i) s[i] = (i < at ? 1.0 : hi) + 1e-3 * g(rng);
          break;
        }
        case 4:  // values spread over 40 decades
          for (double &v : s) v = g(rng) * pow(10.0, 20 * u(rng));
          break;
        default:  // near-vacuum magnitudes
          for (double &v : s) v = 1e-13 * (1 + 0.5 * u(rng));
          break;
      }
      const double *a = &s[2];
      double e[4];
      swf::first_differences(a, e);
      double scale = 0;
      for (int i = -2; i <= 2; ++i) scale = fmax(scale, fabs(a[i] - a[0]));
      if (scale == 0) continue;
      for (int side = 0; side < 2; ++side) {
        const bool right = side == 0;
        const q ref = weno_ref(a, right) - (q)a[0];
        const double r64 = right ? swf::weno_d64<true>(e, 4e-6) : swf::weno_d64<false>(e, 4e-6);
        const double r32 =
            right ? swf::weno_d32_or_64<true>(e, 4e-6, 4e-6f) : swf::weno_d32_or_64<false>(e, 4e-6, 4e-6f);
        const double d64 = (double)fabsq((q)r64 - ref) / scale;
        const double d32 = (double)fabsq((q)r32 - ref) / scale;
        m64[side] = fmax(m64[side], d64);
        m32[side] = (d32 == d32) ? fmax(m32[side], d32) : INFINITY;
      }
    }
    for (int side = 0; side < 2; ++side) {
      printf("d64 %s %s %ld %.3e\n", side ? "left" : "right", kinds[kind], n, m64[side]);
      printf("d32 %s %s %ld %.3e\n", side ? "left" : "right", kinds[kind], n, m32[side]);
    }
  }
  std::mt19937_64 rng2(7);
  double rl = 0;
  for (long t = 0; t < n; ++t) rl = fmax(rl, rl_identity_error(rng2));
  printf("rl identity all %ld %.3e\n", n, rl);
  return 0;
}
