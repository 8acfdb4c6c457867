// swf_ex7.h - the detonation program (ex7_detonation/main.f90) on the GPU, FP64-exact:
// the reactive Euler equations with a fifth conserved variable rho*Y (mass fraction of
// the unburnt gas), pressure p = gamma1 (E - |m|^2/(2 rho) - q rho Y) (eos, :168-174),
// a one-step Arrhenius source -k rho Y exp(-t rho / p) in the rho*Y equation
// (chemsource, :752-771), characteristic-wise WENO with the 5x5 eigen-system of that
// system (wenofd, :176-736), per-line alphas (the reference's alpha_local with mg = 0),
// the step's walls written into u inside wenofd (:186-198, :456-471), its bc (:915-1071)
// and its sweep (sweep_lim, scalar_sweep, nod_lim, :1399-1806) with the program's own
// component-5 indexing in the order-II passes.
//
// Every operation follows the reference's order; the binary128 sites (the Roe weight and
// sound speed with 1d-15, the 1d-12 of its weno5.f90, the RK combinations, the loop
// test) use the double-double emulation of swf_dd.h / swf_q128.h, and exp is the host
// C library's algorithm (glibc 2.35, x86-64 with FMA; exp_glibc below).  The grid
// kernels are one thread per item; the sweep is one workgroup that finds each pass's
// flagged positions in parallel and walks the chains they start in the reference's
// order (k7_sweep).  Bit-identical results are the contract.
#pragma once

#include "swf_device.h"

namespace swf {
namespace ex7 {

constexpr int NC = 5;

// Configuration (one domain: cut_x = cut_y = 1, nx = ny = 5 total)
struct K7 {
  int nx, ny, P;
  long long plane;
  int total, step, swp_num, track_cap;
  double dx, dy, gamma, gamma1, eps, eps2, q, tc, kc, cfl, weight;
};

// Device scalars of one context
struct Dev7 {
  double dt, time, tfinal;
  unsigned long long amax_bits;
  int postag, status, active;
  long long steps_done, sweep_count, den_count, pre_count;
  long long fail_q;
  double pbar;
};

struct B7 {
  double *U, *Uin, *U0;        // stage output, stage input, the step's state (after walls)
  double *fpf, *fnf, *gpf, *gnf;  // fp_flux, fn_flux (x), gp_flux, gn_flux (y): 5 planes each
  double *al;                  // alpha(i, j) of findalpha over the frame
  double *arow, *acol;         // its maxima per row (arow[j + 2]) and per column (acol[i + 2])
  Dev7 *g;
  long long *trk;              // 3 per sweep: q, negative pairs, positive neighbours
  unsigned *bm;                // the sweep: flags of one pass, one bit per position
  int *lst;                    // the sweep: flagged positions of one pass, ascending
};

enum { ST7_OK = 0, ST7_PBAR = 1, ST7_SWEEP_FAIL = 2, ST7_PSUM = 3, ST7_ORDER2_EDGE = 4 };

SWF_HD long long c7(const K7 &k, int i, int j) { return (long long)(j + 2) * k.P + (i + XOFF); }

// ---------------------------------------------------------------------------
// exp: the host C library's exp (glibc 2.35 sysdeps/ieee754/dbl-64/e_exp.c, from
// the ARM optimized-routines, as the x86-64 FMA variant evaluates it), restated:
// k = round(x N / ln2), r = x - k ln2 / N (two-part ln2), 2^(k/N) from a table of
// N = 128 entries (H, T) with 2^(j/N) ~ H (1 + T), and a degree-5 polynomial:
//     exp(x) = scale + scale * (T + r + r^2 (C2 + r C3) + r^4 (C4 + r C5)).
// The main path and the large positive x path are evaluated with fused
// multiply-adds, the large negative x path without, as the library build does (checked bit for bit against the host library by
// tests/test_ex7_exp.py).
// ---------------------------------------------------------------------------
// 2^(j/128) = H (1 + T) for j = 0..127: entry 2j holds the bits of T, entry 2j+1 the
// bits of H minus j << 45 (the exponent of 2^(k/128) is added from k); H is 2^(j/128)
// rounded to binary64 and T the remainder 2^(j/128)/H - 1 rounded, both from the exact
// values (tests/test_ex7_exp.py derives them again with 80-digit decimal arithmetic).
constexpr unsigned long long kExpTab[256] = {
    0x0000000000000000ull, 0x3ff0000000000000ull,
    0x3c9b3b4f1a88bf6eull, 0x3feff63da9fb3335ull,
    0xbc7160139cd8dc5dull, 0x3fefec9a3e778061ull,
    0xbc905e7a108766d1ull, 0x3fefe315e86e7f85ull,
    0x3c8cd2523567f613ull, 0x3fefd9b0d3158574ull,
    0xbc8bce8023f98efaull, 0x3fefd06b29ddf6deull,
    0x3c60f74e61e6c861ull, 0x3fefc74518759bc8ull,
    0x3c90a3e45b33d399ull, 0x3fefbe3ecac6f383ull,
    0x3c979aa65d837b6dull, 0x3fefb5586cf9890full,
    0x3c8eb51a92fdeffcull, 0x3fefac922b7247f7ull,
    0x3c3ebe3d702f9cd1ull, 0x3fefa3ec32d3d1a2ull,
    0xbc6a033489906e0bull, 0x3fef9b66affed31bull,
    0xbc9556522a2fbd0eull, 0x3fef9301d0125b51ull,
    0xbc5080ef8c4eea55ull, 0x3fef8abdc06c31ccull,
    0xbc91c923b9d5f416ull, 0x3fef829aaea92de0ull,
    0x3c80d3e3e95c55afull, 0x3fef7a98c8a58e51ull,
    0xbc801b15eaa59348ull, 0x3fef72b83c7d517bull,
    0xbc8f1ff055de323dull, 0x3fef6af9388c8deaull,
    0x3c8b898c3f1353bfull, 0x3fef635beb6fcb75ull,
    0xbc96d99c7611eb26ull, 0x3fef5be084045cd4ull,
    0x3c9aecf73e3a2f60ull, 0x3fef54873168b9aaull,
    0xbc8fe782cb86389dull, 0x3fef4d5022fcd91dull,
    0x3c8a6f4144a6c38dull, 0x3fef463b88628cd6ull,
    0x3c807a05b0e4047dull, 0x3fef3f49917ddc96ull,
    0x3c968efde3a8a894ull, 0x3fef387a6e756238ull,
    0x3c875e18f274487dull, 0x3fef31ce4fb2a63full,
    0x3c80472b981fe7f2ull, 0x3fef2b4565e27cddull,
    0xbc96b87b3f71085eull, 0x3fef24dfe1f56381ull,
    0x3c82f7e16d09ab31ull, 0x3fef1e9df51fdee1ull,
    0xbc3d219b1a6fbffaull, 0x3fef187fd0dad990ull,
    0x3c8b3782720c0ab4ull, 0x3fef1285a6e4030bull,
    0x3c6e149289cecb8full, 0x3fef0cafa93e2f56ull,
    0x3c834d754db0abb6ull, 0x3fef06fe0a31b715ull,
    0x3c864201e2ac744cull, 0x3fef0170fc4cd831ull,
    0x3c8fdd395dd3f84aull, 0x3feefc08b26416ffull,
    0xbc86a3803b8e5b04ull, 0x3feef6c55f929ff1ull,
    0xbc924aedcc4b5068ull, 0x3feef1a7373aa9cbull,
    0xbc9907f81b512d8eull, 0x3feeecae6d05d866ull,
    0xbc71d1e83e9436d2ull, 0x3feee7db34e59ff7ull,
    0xbc991919b3ce1b15ull, 0x3feee32dc313a8e5ull,
    0x3c859f48a72a4c6dull, 0x3feedea64c123422ull,
    0xbc9312607a28698aull, 0x3feeda4504ac801cull,
    0xbc58a78f4817895bull, 0x3feed60a21f72e2aull,
    0xbc7c2c9b67499a1bull, 0x3feed1f5d950a897ull,
    0x3c4363ed60c2ac11ull, 0x3feece086061892dull,
    0x3c9666093b0664efull, 0x3feeca41ed1d0057ull,
    0x3c6ecce1daa10379ull, 0x3feec6a2b5c13cd0ull,
    0x3c93ff8e3f0f1230ull, 0x3feec32af0d7d3deull,
    0x3c7690cebb7aafb0ull, 0x3feebfdad5362a27ull,
    0x3c931dbdeb54e077ull, 0x3feebcb299fddd0dull,
    0xbc8f94340071a38eull, 0x3feeb9b2769d2ca7ull,
    0xbc87deccdc93a349ull, 0x3feeb6daa2cf6642ull,
    0xbc78dec6bd0f385full, 0x3feeb42b569d4f82ull,
    0xbc861246ec7b5cf6ull, 0x3feeb1a4ca5d920full,
    0x3c93350518fdd78eull, 0x3feeaf4736b527daull,
    0x3c7b98b72f8a9b05ull, 0x3feead12d497c7fdull,
    0x3c9063e1e21c5409ull, 0x3feeab07dd485429ull,
    0x3c34c7855019c6eaull, 0x3feea9268a5946b7ull,
    0x3c9432e62b64c035ull, 0x3feea76f15ad2148ull,
    0xbc8ce44a6199769full, 0x3feea5e1b976dc09ull,
    0xbc8c33c53bef4da8ull, 0x3feea47eb03a5585ull,
    0xbc845378892be9aeull, 0x3feea34634ccc320ull,
    0xbc93cedd78565858ull, 0x3feea23882552225ull,
    0x3c5710aa807e1964ull, 0x3feea155d44ca973ull,
    0xbc93b3efbf5e2228ull, 0x3feea09e667f3bcdull,
    0xbc6a12ad8734b982ull, 0x3feea012750bdabfull,
    0xbc6367efb86da9eeull, 0x3fee9fb23c651a2full,
    0xbc80dc3d54e08851ull, 0x3fee9f7df9519484ull,
    0xbc781f647e5a3ecfull, 0x3fee9f75e8ec5f74ull,
    0xbc86ee4ac08b7db0ull, 0x3fee9f9a48a58174ull,
    0xbc8619321e55e68aull, 0x3fee9feb564267c9ull,
    0x3c909ccb5e09d4d3ull, 0x3feea0694fde5d3full,
    0xbc7b32dcb94da51dull, 0x3feea11473eb0187ull,
    0x3c94ecfd5467c06bull, 0x3feea1ed0130c132ull,
    0x3c65ebe1abd66c55ull, 0x3feea2f336cf4e62ull,
    0xbc88a1c52fb3cf42ull, 0x3feea427543e1a12ull,
    0xbc9369b6f13b3734ull, 0x3feea589994cce13ull,
    0xbc805e843a19ff1eull, 0x3feea71a4623c7adull,
    0xbc94d450d872576eull, 0x3feea8d99b4492edull,
    0x3c90ad675b0e8a00ull, 0x3feeaac7d98a6699ull,
    0x3c8db72fc1f0eab4ull, 0x3feeace5422aa0dbull,
    0xbc65b6609cc5e7ffull, 0x3feeaf3216b5448cull,
    0x3c7bf68359f35f44ull, 0x3feeb1ae99157736ull,
    0xbc93091fa71e3d83ull, 0x3feeb45b0b91ffc6ull,
    0xbc5da9b88b6c1e29ull, 0x3feeb737b0cdc5e5ull,
    0xbc6c23f97c90b959ull, 0x3feeba44cbc8520full,
    0xbc92434322f4f9aaull, 0x3feebd829fde4e50ull,
    0xbc85ca6cd7668e4bull, 0x3feec0f170ca07baull,
    0x3c71affc2b91ce27ull, 0x3feec49182a3f090ull,
    0x3c6dd235e10a73bbull, 0x3feec86319e32323ull,
    0xbc87c50422622263ull, 0x3feecc667b5de565ull,
    0x3c8b1c86e3e231d5ull, 0x3feed09bec4a2d33ull,
    0xbc91bbd1d3bcbb15ull, 0x3feed503b23e255dull,
    0x3c90cc319cee31d2ull, 0x3feed99e1330b358ull,
    0x3c8469846e735ab3ull, 0x3feede6b5579fdbfull,
    0xbc82dfcd978e9db4ull, 0x3feee36bbfd3f37aull,
    0x3c8c1a7792cb3387ull, 0x3feee89f995ad3adull,
    0xbc907b8f4ad1d9faull, 0x3feeee07298db666ull,
    0xbc55c3d956dcaebaull, 0x3feef3a2b84f15fbull,
    0xbc90a40e3da6f640ull, 0x3feef9728de5593aull,
    0xbc68d6f438ad9334ull, 0x3feeff76f2fb5e47ull,
    0xbc91eee26b588a35ull, 0x3fef05b030a1064aull,
    0x3c74ffd70a5fddcdull, 0x3fef0c1e904bc1d2ull,
    0xbc91bdfbfa9298acull, 0x3fef12c25bd71e09ull,
    0x3c736eae30af0cb3ull, 0x3fef199bdd85529cull,
    0x3c8ee3325c9ffd94ull, 0x3fef20ab5fffd07aull,
    0x3c84e08fd10959acull, 0x3fef27f12e57d14bull,
    0x3c63cdaf384e1a67ull, 0x3fef2f6d9406e7b5ull,
    0x3c676b2c6c921968ull, 0x3fef3720dcef9069ull,
    0xbc808a1883ccb5d2ull, 0x3fef3f0b555dc3faull,
    0xbc8fad5d3ffffa6full, 0x3fef472d4a07897cull,
    0xbc900dae3875a949ull, 0x3fef4f87080d89f2ull,
    0x3c74a385a63d07a7ull, 0x3fef5818dcfba487ull,
    0xbc82919e2040220full, 0x3fef60e316c98398ull,
    0x3c8e5a50d5c192acull, 0x3fef69e603db3285ull,
    0x3c843a59ac016b4bull, 0x3fef7321f301b460ull,
    0xbc82d52107b43e1full, 0x3fef7c97337b9b5full,
    0xbc892ab93b470dc9ull, 0x3fef864614f5a129ull,
    0x3c74b604603a88d3ull, 0x3fef902ee78b3ff6ull,
    0x3c83c5ec519d7271ull, 0x3fef9a51fbc74c83ull,
    0xbc8ff7128fd391f0ull, 0x3fefa4afa2a490daull,
    0xbc8dae98e223747dull, 0x3fefaf482d8e67f1ull,
    0x3c8ec3bc41aa2008ull, 0x3fefba1bee615a27ull,
    0x3c842b94c3a9eb32ull, 0x3fefc52b376bba97ull,
    0x3c8a64a931d185eeull, 0x3fefd0765b6e4540ull,
    0xbc8e37bae43be3edull, 0x3fefdbfdad9cbe14ull,
    0x3c77893b4d91cd9dull, 0x3fefe7c1819e90d8ull,
    0x3c5305c14160cc89ull, 0x3feff3c22b8f71f1ull,
};

SWF_HD unsigned long long dbits(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return (unsigned long long)__double_as_longlong(x);
#else
  unsigned long long u;
  __builtin_memcpy(&u, &x, 8);
  return u;
#endif
}
SWF_HD double bitsd(unsigned long long u) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __longlong_as_double((long long)u);
#else
  double x;
  __builtin_memcpy(&x, &u, 8);
  return x;
#endif
}

SWF_HD double exp_special(double tmp, unsigned long long sbits, unsigned long long ki) {
#pragma clang fp contract(off)
  double scale, y;
  if ((ki & 0x80000000ull) == 0) {
    // k > 0: the product and sum fused, as the library build evaluates this branch
    sbits -= 1009ull << 52;
    scale = bitsd(sbits);
    return 0x1p1009 * fma(scale, tmp, scale);
  }
  sbits += 1022ull << 52;
  scale = bitsd(sbits);
  y = scale + scale * tmp;
  if (y < 1.0) {
    // rounded to the result's precision before the scaling into the subnormal range
    double lo = scale - y + scale * tmp;
    const double hi = 1.0 + y;
    lo = 1.0 - hi + y + lo;
    y = (hi + lo) - 1.0;
    if (y == 0.0) y = 0.0;
  }
  return 0x1p-1022 * y;
}

SWF_HD double exp_glibc(double x) {
  const unsigned long long *tab = kExpTab;
  const double InvLn2N = 0x1.71547652b82fep0 * 128, Shift = 0x1.8p52;
  const double NegLn2hiN = -0x1.62e42fefa0000p-8, NegLn2loN = -0x1.cf79abc9e3b3ap-47;
  const double C2 = 0x1.ffffffffffdbdp-2, C3 = 0x1.555555555543cp-3, C4 = 0x1.55555cf172b91p-5,
               C5 = 0x1.1111167a4d017p-7;
  const unsigned top_small = 0x3c9;  // top12(2^-54)
  unsigned abstop = (unsigned)(dbits(x) >> 52) & 0x7ffu;
  if (abstop - top_small >= 0x408u - top_small) {  // 0x408 = top12(512)
    if (abstop - top_small >= 0x80000000u) return 1.0 + x;
    if (abstop >= 0x409u) {  // top12(1024)
      if (dbits(x) == 0xfff0000000000000ull) return 0.0;
      if (abstop >= 0x7ffu) return 1.0 + x;
      return (dbits(x) >> 63) ? 0.0 : bitsd(0x7ff0000000000000ull);
    }
    abstop = 0;
  }
  double kd = fma(InvLn2N, x, Shift);
  const unsigned long long ki = dbits(kd);
  kd = kd - Shift;
  const double r = fma(kd, NegLn2loN, fma(kd, NegLn2hiN, x));
  const unsigned long long idx = 2 * (ki % 128), top = ki << 45;
  const double tail = bitsd(tab[idx]);
  const unsigned long long sbits = tab[idx + 1] + top;
  const double r2 = r * r;
  const double tmp = fma(r2 * r2, fma(r, C5, C4), fma(r2, fma(r, C3, C2), tail + r));
  if (abstop == 0) return exp_special(tmp, sbits, ki);
  const double scale = bitsd(sbits);
  return fma(scale, tmp, scale);
}

// ---------------------------------------------------------------------------
// Point physics (the reference's expression order; contraction off in this unit)
// ---------------------------------------------------------------------------
// eos (:168-174) with the velocities of fflux / gflux / chemsource
SWF_HD double eos7(const K7 &k, double rho, double vx, double vy, double e, double y) {
  return k.gamma1 * (((e - (0.5 * rho) * (vx * vx)) - (0.5 * rho) * (vy * vy)) - k.q * y);
}

// fflux (DIR 0, :110-137) / gflux (DIR 1, :139-166)
template <int DIR>
SWF_HD void flux7(const K7 &k, const double u[NC], double f[NC]) {
  const double rho = u[0], mx = u[1], my = u[2], e = u[3], y = u[4];
  double vx = mx / rho, vy = my / rho;
  if (rho < k.eps2) vx = 0.;
  if (rho < k.eps2) vy = 0.;
  const double pres = eos7(k, rho, vx, vy, e, y);
  if (DIR == 0) {
    f[0] = mx;
    f[1] = rho * (vx * vx) + pres;
    f[2] = (rho * vx) * vy;
    f[3] = (e + pres) * vx;
    f[4] = y * vx;
  } else {
    f[0] = my;
    f[1] = (rho * vx) * vy;
    f[2] = rho * (vy * vy) + pres;
    f[3] = (e + pres) * vy;
    f[4] = y * vy;
  }
}

// chemsource (:752-771): the rho*Y equation's source at a cell
SWF_HD double chem7(const K7 &k, const double u[NC]) {
  const double rho = u[0], y = u[4];
  double vx = u[1] / rho, vy = u[2] / rho;
  if (rho < qconst::e8_hi) vx = 0.;  // rho < 1d-8 (binary128 literal)
  if (rho < qconst::e8_hi) vy = 0.;
  const double pres = eos7(k, rho, vx, vy, u[3], y);
  return -((k.kc * y) * exp_glibc(-((k.tc * rho) / pres)));
}

// findalpha's alpha(i, j) (:774-808)
SWF_HD double alpha7(const K7 &k, const double u[NC]) {
  const double den = u[0];
  double vx, vy;
  if (den >= k.eps2) {
    vx = u[1] / den;
    vy = u[2] / den;
  } else {
    vx = 0.;
    vy = 0.;
  }
  const double pres = k.gamma1 * (((u[3] - (0.5 * den) * (vx * vx)) - (0.5 * den) * (vy * vy)) - u[4] * k.q);
  const double sound = sqrt((k.gamma * pres) / den);
  const double a = fabs(u[1] / u[0]) + sound, b = fabs(u[2] / u[0]) + sound;
  return fmax2(a, b);
}

// the primitive values of a point for the Roe average (:201-214, :476-491)
struct Prim7 {
  double w, vx, vy, ratio, h;
};
SWF_HD Prim7 prim7(const K7 &k, const double u[NC]) {
  const double den = u[0], xmt = u[1], ymt = u[2], eng = u[3], yrho = u[4];
  const double t0 = 1. / den;
  const double vex = xmt * t0, vey = ymt * t0;
  const double pre = k.gamma1 * ((eng - 0.5 * (xmt * vex + ymt * vey)) - k.q * yrho);
  Prim7 p;
  p.w = sqrt(fabs(den));
  p.vx = vex;
  p.vy = vey;
  p.ratio = yrho * t0;
  p.h = (pre + eng) * t0;
  return p;
}

// Both halves of centre c of a line: the Roe system of (c, c+1) (:235-318 x,
// :509-589 y), projection of the 5-point stencils of the split fluxes sp (f+) and sn
// (f-), WENO with 1d-12 (ex7's weno5.f90), back-projection, times the line's alpha.
// sp[m][c] is component c of stencil point m (m = 0..4 for c-2..c+2).
template <int DIR>
SWF_HD void centre7(const K7 &k, const Prim7 &pa, const Prim7 &pb, const double sp[5][NC], const double sn[5][NC],
                    double al, double outp[NC], double outn[NC]) {
  const double t0 = roe_weight_q(pa.w, pb.w);
  const double t1 = 1. - t0;
  const double vxm = t0 * pa.vx + t1 * pb.vx;
  const double vym = t0 * pa.vy + t1 * pb.vy;
  const double ym = t0 * pa.ratio + t1 * pb.ratio;
  const double hm = t0 * pa.h + t1 * pb.h;
  const double qm = 0.5 * (vxm * vxm + vym * vym);
  const double cm = cm_q(sqrt(fabs(k.gamma1 * ((hm - qm) - k.q * ym))));
  const double q = k.q;
  double R[NC][NC], L[NC][NC];
  const double rcm = 1. / cm;
  const double b1 = (k.gamma1 * rcm) * rcm;
  const double b2 = qm * b1;
  if (DIR == 0) {
    const double tc = vxm * cm;
    R[0][0] = 1.0;      R[0][1] = 1.0; R[0][2] = 0.0;  R[0][3] = 0.0; R[0][4] = 1.0;
    R[1][0] = vxm - cm; R[1][1] = vxm; R[1][2] = 0.;   R[1][3] = 0.;  R[1][4] = vxm + cm;
    R[2][0] = vym;      R[2][1] = vym; R[2][2] = -1.;  R[2][3] = 0.;  R[2][4] = vym;
    R[3][0] = hm - tc;  R[3][1] = qm;  R[3][2] = -vym; R[3][3] = q;   R[3][4] = hm + tc;
    R[4][0] = ym;       R[4][1] = 0.;  R[4][2] = 0.;   R[4][3] = 1.;  R[4][4] = ym;
    const double s0 = vxm * rcm, s1 = b1 * vxm, s2 = 0.5 * b1, s3 = b1 * vym;
    L[0][0] = 0.5 * (b2 + s0); L[0][1] = -(0.5 * (s1 + rcm)); L[0][2] = -(0.5 * s3); L[0][3] = s2; L[0][4] = -(s2 * q);
    L[1][0] = 1. - b2;         L[1][1] = s1;                  L[1][2] = s3;          L[1][3] = -b1; L[1][4] = b1 * q;
    L[2][0] = vym;             L[2][1] = 0.0;                 L[2][2] = -1.0;        L[2][3] = 0.0; L[2][4] = 0.0;
    L[3][0] = -(b2 * ym);      L[3][1] = s1 * ym;             L[3][2] = s3 * ym;     L[3][3] = -(b1 * ym);
    L[3][4] = 1. + (q * ym) * b1;
    L[4][0] = 0.5 * (b2 - s0); L[4][1] = -(0.5 * (s1 - rcm)); L[4][2] = -(0.5 * s3); L[4][3] = s2; L[4][4] = -(s2 * q);
  } else {
    const double tc = vym * cm;
    R[0][0] = 1.0;      R[0][1] = 1.0; R[0][2] = 0.0; R[0][3] = 0.0; R[0][4] = 1.0;
    R[1][0] = vxm;      R[1][1] = vxm; R[1][2] = 1.0; R[1][3] = 0.;  R[1][4] = vxm;
    R[2][0] = vym - cm; R[2][1] = vym; R[2][2] = 0.;  R[2][3] = 0.;  R[2][4] = vym + cm;
    R[3][0] = hm - tc;  R[3][1] = qm;  R[3][2] = vxm; R[3][3] = q;   R[3][4] = hm + tc;
    R[4][0] = ym;       R[4][1] = 0.;  R[4][2] = 0.;  R[4][3] = 1.;  R[4][4] = ym;
    const double s0 = vym * rcm, s1 = b1 * vym, s2 = 0.5 * b1, s3 = b1 * vxm;
    L[0][0] = 0.5 * (b2 + s0); L[0][1] = -(0.5 * s3); L[0][2] = -(0.5 * (s1 + rcm)); L[0][3] = s2; L[0][4] = -(q * s2);
    L[1][0] = 1. - b2;         L[1][1] = s3;          L[1][2] = s1;                  L[1][3] = -b1; L[1][4] = b1 * q;
    L[2][0] = -vxm;            L[2][1] = 1.0;         L[2][2] = 0.0;                 L[2][3] = 0.0; L[2][4] = 0.0;
    L[3][0] = -(ym * b2);      L[3][1] = ym * s3;     L[3][2] = ym * s1;             L[3][3] = -(ym * b1);
    L[3][4] = 1. + (ym * b1) * q;
    L[4][0] = 0.5 * (b2 - s0); L[4][1] = -(0.5 * s3); L[4][2] = -(0.5 * (s1 - rcm)); L[4][3] = s2; L[4][4] = -(q * s2);
  }
  double ur[NC], ul[NC];
  for (int m = 0; m < NC; ++m) {
    double cp[5], cn[5];
    for (int g = 0; g < 5; ++g) {
      cp[g] = (((L[m][0] * sp[g][0] + L[m][1] * sp[g][1]) + L[m][2] * sp[g][2]) + L[m][3] * sp[g][3]) +
              L[m][4] * sp[g][4];
      cn[g] = (((L[m][0] * sn[g][0] + L[m][1] * sn[g][1]) + L[m][2] * sn[g][2]) + L[m][3] * sn[g][3]) +
              L[m][4] * sn[g][4];
    }
    ur[m] = weno_right<MODE_EXACT>(&cp[2], kWeps12);
    ul[m] = weno_left<MODE_EXACT>(&cn[2], kWeps12);
  }
  for (int c = 0; c < NC; ++c) {
    const double fp = (((R[c][0] * ur[0] + R[c][1] * ur[1]) + R[c][2] * ur[2]) + R[c][3] * ur[3]) + R[c][4] * ur[4];
    const double fn = (((R[c][0] * ul[0] + R[c][1] * ul[1]) + R[c][2] * ul[2]) + R[c][3] * ul[3]) + R[c][4] * ul[4];
    outp[c] = al * fp;
    outn[c] = al * fn;
  }
}

SWF_HD void load7(const K7 &k, const double *U, int i, int j, double v[NC]) {
  const long long o = c7(k, i, j);
  for (int c = 0; c < NC; ++c) v[c] = U[c * k.plane + o];
}
SWF_HD void store7(const K7 &k, double *U, int i, int j, const double v[NC]) {
  const long long o = c7(k, i, j);
  for (int c = 0; c < NC; ++c) U[c * k.plane + o] = v[c];
}

// bc (:915-1071) on one domain, as the value of cell (i, j) after its passes:
// bottom rows fixed gas (1, 0, 0, 55, 1), right columns the same, top rows the row ny
// reflected (m2 negated; at the right columns row ny is the fixed gas, so m2 = -0),
// left columns the inflow (11, 11*6.18, 0, 970, 11), then every cell with
// i < total and j < step*total (ghosts too) reset to the gas at rest.  The rank's own
// exchange (cut 1 x 1) is overwritten by these passes everywhere.  Interior cells
// outside the step keep their value (false).
SWF_HD bool bc7_value(const K7 &k, const double *U, int i, int j, double v[NC]) {
  const bool stepcell = i < k.total && j < k.step * k.total;
  if (stepcell) {
    v[0] = 1.; v[1] = 0.; v[2] = 0.; v[3] = 55.; v[4] = 1.;
  } else if (i <= 0) {
    v[0] = 11.; v[1] = 11. * 6.18; v[2] = 0.; v[3] = 970.; v[4] = 11.;
  } else if (j >= k.ny + 1) {
    if (i >= k.nx + 1) {
      v[0] = 1.; v[1] = 0.; v[2] = -0.; v[3] = 55.; v[4] = 1.;
    } else {
      load7(k, U, i, k.ny, v);
      v[2] = -v[2];
    }
  } else if (i >= k.nx + 1 || j <= 0) {
    v[0] = 1.; v[1] = 0.; v[2] = 0.; v[3] = 55.; v[4] = 1.;
  } else {
    return false;
  }
  return true;
}

// ---------------------------------------------------------------------------
// Kernels (compiled once, in swf_k_exact.hip)
// ---------------------------------------------------------------------------
#ifdef SWF_COMMON_KERNELS
SWF_D bool live7(const B7 &b) { return b.g->active && b.g->status == ST7_OK; }

// findalpha: alpha of every cell of the frame
__global__ __launch_bounds__(256) void k7_alpha(K7 k, B7 b) {
  if (!live7(b)) return;
  const int i = -2 + (int)(blockIdx.x * 64 + threadIdx.x), j = -2 + (int)(blockIdx.y * 4 + threadIdx.y);
  if (i > k.nx + 3 || j > k.ny + 3) return;
  double v[NC];
  load7(k, b.Uin, i, j, v);
  b.al[c7(k, i, j)] = alpha7(k, v);
}

// alpha_local of each row (x-pass) and column (y-pass), one wave per line: lines
// 0..ny+5 the rows, the others the columns.  The maximum of non-negative alphas does
// not depend on the order of the comparisons (a NaN never replaces the running value,
// as in the reference's serial max).
__global__ __launch_bounds__(256) void k7_lines(K7 k, B7 b) {
  if (!live7(b)) return;
  const int line = blockIdx.x * 4 + (int)(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int nr = k.ny + 6, ncl = k.nx + 6;
  if (line >= nr + ncl) return;  // per wave
  double a = 0.;
  if (line < nr) {
    const int j = line - 2;
    for (int i = -2 + lane; i <= k.nx + 3; i += 64) a = fmax2(a, b.al[c7(k, i, j)]);
  } else {
    const int i = line - nr - 2;
    for (int j = -2 + lane; j <= k.ny + 3; j += 64) a = fmax2(a, b.al[c7(k, i, j)]);
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) a = fmax2(a, __shfl_xor(a, off, 64));
  if (lane == 0) {
    if (line < nr)
      b.arow[line] = a;
    else
      b.acol[line - nr] = a;
  }
}

// stage 1: alphamax, dt (rk :816-821) and the time after the step; one wave
__global__ __launch_bounds__(64) void k7_dt(K7 k, B7 b) {
  if (!live7(b)) return;
  Dev7 *g = b.g;
  const int lane = threadIdx.x;
  double a = 0.;
  for (int j = lane; j < k.ny + 6; j += 64) a = fmax2(a, b.arow[j]);
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) a = fmax2(a, __shfl_xor(a, off, 64));
  if (lane == 0) {
    g->amax_bits = dbits(a);
    const double dtx = (k.dx * k.cfl) / a, dty = (k.dy * k.cfl) / a;
    g->dt = fmin2(fmin2(dtx, dty), g->tfinal - g->time);
  }
}

// the walls inside wenofd: before the x-pass (dir 0) the three cells left of x =
// total mirror the three right of it in rows 0..min(ny+1, step*total), m1 negated; before
// the y-pass (dir 1) the rows step*total-2 .. step*total mirror the three above, m2
// negated, in columns 1..total
__global__ __launch_bounds__(256) void k7_walls(K7 k, B7 b, int dir) {
  if (!live7(b)) return;
  const int t = blockIdx.x * 256 + threadIdx.x;
  const int T = k.total, js = k.step * T;
  if (dir == 0) {
    const int nrow = (k.ny + 1 < js ? k.ny + 1 : js) + 1;  // rows 0..min(ny+1, js)
    if (t >= 3 * nrow) return;
    const int j = t / 3, ii = t % 3;
    double v[NC];
    load7(k, b.Uin, T + 1 + ii, j, v);
    v[1] = -v[1];
    store7(k, b.Uin, T - ii, j, v);
  } else {
    if (js < -2 || js > k.ny + 3 || t >= 3 * T) return;
    const int ii = t / T, kk = 1 + t % T;
    double v[NC];
    load7(k, b.Uin, kk, js + 1 + ii, v);
    v[2] = -v[2];
    store7(k, b.Uin, kk, js - ii, v);
  }
}

// wenofd pass DIR for centres c = 0..n+1 of lines 0..m+1 (x: rows, y: columns)
template <int DIR>
__global__ __launch_bounds__(128) void k7_flux(K7 k, B7 b) {
  if (!live7(b)) return;
  const int n = DIR == 0 ? k.nx : k.ny, m = DIR == 0 ? k.ny : k.nx;
  const int t = blockIdx.x * 128 + threadIdx.x;
  if (t >= (n + 2) * (m + 2)) return;
  const int c = t % (n + 2), line = t / (n + 2);
  const double al = DIR == 0 ? b.arow[line + 2] : b.acol[line + 2];
  double sp[5][NC], sn[5][NC];
  Prim7 pa, pb;
  for (int g = 0; g < 5; ++g) {
    const int p = c - 2 + g;
    double u[NC], f[NC];
    if (DIR == 0)
      load7(k, b.Uin, p, line, u);
    else
      load7(k, b.Uin, line, p, u);
    flux7<DIR>(k, u, f);
    for (int q = 0; q < NC; ++q) {
      const double fa = f[q] / al;
      sp[g][q] = 0.5 * (u[q] + fa);
      sn[g][q] = 0.5 * (u[q] - fa);
    }
    if (g == 2) pa = prim7(k, u);
    if (g == 3) pb = prim7(k, u);
  }
  double op[NC], on[NC];
  centre7<DIR>(k, pa, pb, sp, sn, al, op, on);
  const long long o = DIR == 0 ? c7(k, c, line) : c7(k, line, c);
  double *FP = DIR == 0 ? b.fpf : b.gpf, *FN = DIR == 0 ? b.fnf : b.gnf;
  for (int q = 0; q < NC; ++q) {
    FP[q * k.plane + o] = op[q];
    FN[q * k.plane + o] = on[q];
  }
}

// scheme (:738-750), the RK combination of stage s (rk :825, :848, :870) and
// checkpositivity (:890-913) for the interior cells
__global__ __launch_bounds__(256) void k7_stage(K7 k, B7 b, int s) {
  if (!live7(b)) return;
  const int i = 1 + (int)(blockIdx.x * 64 + threadIdx.x), j = 1 + (int)(blockIdx.y * 4 + threadIdx.y);
  if (i > k.nx || j > k.ny) return;
  const long long o = c7(k, i, j), ox = c7(k, i + 1, j), oxm = c7(k, i - 1, j), oy = c7(k, i, j + 1),
                  oym = c7(k, i, j - 1);
  double u[NC], rhs[NC];
  load7(k, b.Uin, i, j, u);
  for (int c = 0; c < NC; ++c) {
    const long long pc = c * k.plane;
    const double X = (b.fpf[pc + o] - b.fnf[pc + ox]) - (b.fpf[pc + oxm] - b.fnf[pc + o]);
    const double Y = (b.gpf[pc + o] - b.gnf[pc + oy]) - (b.gpf[pc + oym] - b.gnf[pc + o]);
    rhs[c] = -(X / k.dx) - Y / k.dy;
  }
  rhs[4] = rhs[4] + chem7(k, u);
  const double dt = b.g->dt;
  double v[NC];
  for (int c = 0; c < NC; ++c) {
    const double x = u[c] + dt * rhs[c];
    if (s == 1)
      v[c] = x;
    else if (s == 2)
      v[c] = rk2_q(b.U0[c * k.plane + o], x);
    else
      v[c] = rk3_q(b.U0[c * k.plane + o], x);
  }
  store7(k, b.U, i, j, v);
  const bool bad = v[0] <= k.eps || (((v[3] - ((v[2] * v[2]) / v[0]) / 2.) - ((v[1] * v[1]) / v[0]) / 2.) -
                                     k.q * v[4]) <= k.eps;
  if (bad) atomicOr(&b.g->postag, 1);
}

// bc of the stage output (every ghost and step cell; interior row ny first read)
__global__ __launch_bounds__(256) void k7_bc(K7 k, B7 b, int force) {
  if (!force && !live7(b)) return;
  const int i = -2 + (int)(blockIdx.x * 64 + threadIdx.x), j = -2 + (int)(blockIdx.y * 4 + threadIdx.y);
  if (i > k.nx + 3 || j > k.ny + 3) return;
  double v[NC];
  if (bc7_value(k, b.U, i, j, v)) store7(k, b.U, i, j, v);
}

// ---------------------------------------------------------------------------
// The sweep (sweep_lim :1433-1737 with scalar_sweep :1740-1806), one thread, in the
// reference's serial order on the interior (its gathered utemp).  Snake orders
// (sweeping_order :1808-1862): order I row by row (posI), order II column by column.
// ---------------------------------------------------------------------------
struct SweepCounts {
  long long pos, neg;
};

// pressure forms of the sweep
SWF_HD double pres_flag7(const K7 &k, const double v[NC]) {
  // gamma1*(v(4) - 0.5*v(2)*v(2)/v(1) - 0.5*v(3)**2/v(1) - q_const*v(5))
  return k.gamma1 * (((v[3] - ((0.5 * v[1]) * v[1]) / v[0]) - (0.5 * (v[2] * v[2])) / v[0]) - k.q * v[4]);
}
SWF_HD double pres_sq7(const K7 &k, const double v[NC]) {
  // gamma1*(v(4) - 0.5*v(2)**2/v(1) - 0.5*v(3)**2/v(1) - q_const*v(5))
  return k.gamma1 * (((v[3] - (0.5 * (v[1] * v[1])) / v[0]) - (0.5 * (v[2] * v[2])) / v[0]) - k.q * v[4]);
}

// nod_lim (:1399-1429)
SWF_HD void nod_lim7(const K7 &k, const double v1[NC], const double v2[NC], const double vb[NC], double out[NC],
                     SweepCounts &cnt) {
  const double p1 = pres_sq7(k, v1), p2 = pres_sq7(k, v2);
  double t;
  if (p2 > k.eps) {
    t = (p1 - k.eps) / (p1 - p2);
    cnt.pos++;
  } else {
    const double pb = pres_sq7(k, vb);
    const double t1 = (p1 - k.eps) / (p1 - pb);
    double s12 = 0., s1b = 0.;
    for (int c = NC - 1; c >= 0; --c) {
      const double a = v2[c] - v1[c], d = vb[c] - v1[c];
      s12 = (c == NC - 1) ? a * a : s12 + a * a;
      s1b = (c == NC - 1) ? d * d : s1b + d * d;
    }
    const double d12 = sqrt(s12), d1b = sqrt(s1b);
    t = fmin2((t1 * d1b) / d12, 0.25);
    cnt.neg++;
  }
  for (int c = 0; c < NC; ++c) out[c] = (1. - t) * v1[c] + t * v2[c];
}

SWF_HD void cell_I(const K7 &k, long long p, int &i, int &j) {
  // order I (row by row): p = 1-based position
  const long long q = p - 1;
  j = (int)(q / k.nx) + 1;
  const int r = (int)(q - (long long)(j - 1) * k.nx);
  i = (j & 1) ? r + 1 : k.nx - r;
}
SWF_HD void cell_II(const K7 &k, long long p, int &i, int &j) {
  const long long q = p - 1;
  i = (int)(q / k.ny) + 1;
  const int r = (int)(q - (long long)(i - 1) * k.ny);
  j = (i & 1) ? r + 1 : k.ny - r;
}

// The sweep in one workgroup of SW7_T threads.  In each pass, the adjustment at
// position p (when the cell there is flagged) changes only the cells of positions p
// and p + 1 (p - 1 in a backward pass); in the order-II passes component 5 is changed
// at the order-I positions p and p - 1, which no later position of the pass reads.
// So a later position's flag can change only through the chain that reaches it from
// its predecessor, and every other position is flagged at its turn exactly when it
// was at the start of the pass.  A pass therefore (1) finds the positions flagged at
// its start, all threads over contiguous words of a bitmap, then an ordered list by
// a block scan, and (2) one thread walks the chains they start, in the reference's
// order, re-testing each chain's next position on the current state.  The minima of
// the density and pressure (the density sweep's test and the convergence test) are
// folded over contiguous segments and combined in order, with the reference's
// comparison (fmin2(x, m): a NaN replaces the running value, the next value replaces
// the NaN), so they equal its serial loops bit for bit for any input.  u_bar is the
// reference's serial sum, staged through LDS in chunks.
constexpr int SW7_T = 1024;  // threads of the sweep workgroup
constexpr int SW7_CH = 512;  // cells per u_bar chunk

struct Sweep7Shared {
  double ubc[NC][SW7_CH];
  double pm[2][SW7_T];        // per-thread minima of a segment
  unsigned char pf[2][SW7_T];  // 1: the segment has values, 2: one of them is a NaN
  int wscan[SW7_T / 64];
  int total;
  int stop;
  double ub[NC];
};

// one value into a segment's running minimum (started at +inf)
SWF_D void min_push(double &m, unsigned char &f, double x) {
  f |= (unsigned char)(1 | (x != x ? 2 : 0));
  m = fmin2(x, m);
}

// the reference's serial minimum from 1e20 over the segments 0..SW7_T-1 in order
SWF_D double min_combine(const double *pm, const unsigned char *pf) {
  double acc = 1e20;
  for (int t = 0; t < SW7_T; ++t) {
    if (!(pf[t] & 1)) continue;
    if ((pf[t] & 2) || acc != acc)
      acc = pm[t];
    else
      acc = fmin2(pm[t], acc);
  }
  return acc;
}

// exclusive prefix sum over the workgroup in thread order; total to sh.total
SWF_D int scan7(int x, Sweep7Shared &sh, int tid) {
  int incl = x;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int v = __shfl_up(incl, off, 64);
    if ((tid & 63) >= off) incl += v;
  }
  if ((tid & 63) == 63) sh.wscan[tid >> 6] = incl;
  __syncthreads();
  if (tid == 0) {
    int run = 0;
    for (int q = 0; q < SW7_T / 64; ++q) {
      const int t = sh.wscan[q];
      sh.wscan[q] = run;
      run += t;
    }
    sh.total = run;
  }
  __syncthreads();
  return sh.wscan[tid >> 6] + incl - x;
}

enum { PASS_DEN = 0, PASS_I = 1, PASS_II = 2 };

// the cell's flag at position p of a pass of this kind, on the current state
SWF_D bool flag7(const K7 &k, const double *U, int kind, long long p) {
  int i, j;
  if (kind == PASS_II)
    cell_II(k, p, i, j);
  else
    cell_I(k, p, i, j);
  const long long o = c7(k, i, j);
  if (kind == PASS_DEN) return U[o] < 1e-12;
  double v[NC];
  for (int c = 0; c < 4; ++c) v[c] = U[c * k.plane + o];
  if (kind == PASS_II) {
    int i5, j5;
    cell_I(k, p, i5, j5);
    v[4] = U[4 * k.plane + c7(k, i5, j5)];
  } else {
    v[4] = U[4 * k.plane + o];
  }
  return pres_flag7(k, v) <= k.eps || v[0] <= k.eps;
}

// Step 1 of a pass: the positions 1..N flagged on the current state, ascending, into
// b.lst; their number in sh.total.
SWF_D void flagged_list(const K7 &k, const B7 &b, int kind, Sweep7Shared &sh, int tid) {
  const long long N = (long long)k.nx * k.ny;
  const int nw = (int)((N + 31) / 32), wpt = (nw + SW7_T - 1) / SW7_T;
  const int w0 = tid * wpt, w1 = min(nw, w0 + wpt);
  int cnt = 0;
  for (int w = w0; w < w1; ++w) {
    unsigned bits = 0;
    for (int bit = 0; bit < 32; ++bit) {
      const long long p = 32LL * w + bit + 1;
      if (p <= N && flag7(k, b.U, kind, p)) bits |= 1u << bit;
    }
    b.bm[w] = bits;
    cnt += __popc(bits);
  }
  int o = scan7(cnt, sh, tid);
  for (int w = w0; w < w1; ++w) {
    unsigned bits = b.bm[w];
    while (bits) {
      const int bit = __ffs(bits) - 1;
      bits &= bits - 1;
      b.lst[o++] = 32 * w + bit + 1;
    }
  }
  __syncthreads();
}

// thread 0's part of the sweep: the adjustments of sweep_lim / scalar_sweep
struct Walk7 {
  const K7 &k;
  const B7 &b;
  double ub[NC];
  SweepCounts cnt;
  int q;

  SWF_D double &at(int c, int i, int j) const { return b.U[c * k.plane + c7(k, i, j)]; }

  // one adjustment of the pair (cell a, neighbour p); component 5 at (a5, p5).
  // 0: the cell is not flagged, 1: adjusted, 2: the pressure-sum check failed
  SWF_D int adjust(int ia, int ja, int ip, int jp, int ia5, int ja5, int ip5, int jp5, bool check) {
    double vi[NC], vp[NC], vit[NC], vpt[NC];
    for (int c = 0; c < 4; ++c) {
      vi[c] = at(c, ia, ja);
      vp[c] = at(c, ip, jp);
    }
    vi[4] = at(4, ia5, ja5);
    const double pvi = pres_flag7(k, vi);
    if (!(pvi <= k.eps || vi[0] <= k.eps)) return 0;
    vp[4] = at(4, ip5, jp5);
    nod_lim7(k, vi, vp, ub, vit, cnt);
    for (int c = 0; c < NC; ++c) {
      const double d = vi[c] - vit[c];
      vpt[c] = vp[c] + (k.weight / k.weight) * d;
    }
    for (int c = 0; c < 4; ++c) at(c, ia, ja) = vit[c];
    at(4, ia5, ja5) = vit[4];
    for (int c = 0; c < 4; ++c) at(c, ip, jp) = vpt[c];
    at(4, ip5, jp5) = vpt[4];
    if (check) {
      const double pvp = pres_flag7(k, vp), pvit = pres_sq7(k, vit), pvpt = pres_sq7(k, vpt);
      if ((pvi + pvp) > (pvit + pvpt) + k.eps) return 2;
    }
    return 1;
  }

  // position p of a pass (fwd: neighbour p + 1, else p - 1): 0 not flagged, 1 adjusted,
  // 2 stopped (status set)
  SWF_D int at_position(int kind, bool fwd, long long p) {
    int i, j, in, jn;
    if (kind == PASS_DEN) {
      cell_I(k, p, i, j);
      cell_I(k, fwd ? p + 1 : p - 1, in, jn);
      const double vi = at(0, i, j), vn = at(0, in, jn);
      if (!(vi < 1e-12)) return 0;
      at(0, in, jn) = vn + (k.weight / k.weight) * (vi - 1e-12);
      at(0, i, j) = 1e-12;
      return 1;
    }
    if (kind == PASS_I) {
      cell_I(k, p, i, j);
      cell_I(k, fwd ? p + 1 : p - 1, in, jn);
      const int r = adjust(i, j, in, jn, i, j, in, jn, true);
      if (r == 2) {
        b.g->status = ST7_PSUM;
        b.g->fail_q = q;
      }
      return r;
    }
    // order II; component 5 at the order-I positions p (cell) and p - 1 (neighbour).  At
    // p = 1 of the forward pass that neighbour is position 0, outside the order arrays:
    // not reproducible, stop.
    int i5, j5, in5 = 0, jn5 = 0;
    cell_II(k, p, i, j);
    cell_II(k, fwd ? p + 1 : p - 1, in, jn);
    cell_I(k, p, i5, j5);
    if (fwd) {
      double v[NC];
      for (int c = 0; c < 4; ++c) v[c] = at(c, i, j);
      v[4] = at(4, i5, j5);
      const bool flagged = pres_flag7(k, v) <= k.eps || v[0] <= k.eps;
      if (!flagged) return 0;
      if (p == 1) {
        b.g->status = ST7_ORDER2_EDGE;
        b.g->fail_q = q;
        return 2;
      }
    }
    cell_I(k, p - 1, in5, jn5);
    return adjust(i, j, in, jn, i5, j5, in5, jn5, false);
  }

  // step 2 of a pass: the chains from the listed positions, in the pass's order;
  // false when the pass stopped the sweep
  SWF_D bool walk(int kind, bool fwd, int L) {
    const long long N = (long long)k.nx * k.ny;
    if (fwd) {
      long long cur = 1;  // positions below cur are done
      for (int idx = 0; idx < L; ++idx) {
        long long p = b.lst[idx];
        if (p < cur) continue;
        if (p > N - 1) break;
        while (true) {
          const int r = at_position(kind, true, p);
          if (r == 2) return false;
          if (r == 0 || p + 1 > N - 1) {
            cur = p + 1;
            break;
          }
          ++p;
        }
      }
    } else {
      long long cur = N;  // positions above cur are done
      for (int idx = L - 1; idx >= 0; --idx) {
        long long p = b.lst[idx];
        if (p > cur) continue;
        if (p < 2) break;
        while (true) {
          const int r = at_position(kind, false, p);
          if (r == 2) return false;
          if (r == 0 || p - 1 < 2) {
            cur = p - 1;
            break;
          }
          --p;
        }
      }
    }
    return true;
  }
};

// one pass of every thread; false when it stopped the sweep
SWF_D bool pass7(const K7 &k, const B7 &b, Walk7 &w, int kind, bool fwd, Sweep7Shared &sh, int tid) {
  flagged_list(k, b, kind, sh, tid);
  if (tid == 0) sh.stop = w.walk(kind, fwd, sh.total) ? 0 : 1;
  __syncthreads();
  return !sh.stop;
}

__global__ __launch_bounds__(SW7_T) void k7_sweep(K7 k, B7 b) {
  __shared__ Sweep7Shared sh;
  if (!live7(b)) return;
  Dev7 *g = b.g;
  if (!g->postag) return;
  const int tid = threadIdx.x;
  const long long N = (long long)k.nx * k.ny;
  const long long seg = (N + SW7_T - 1) / SW7_T, q0 = tid * seg, q1 = (q0 + seg < N) ? q0 + seg : N;
  const double *U = b.U;
  Walk7 w{k, b, {0., 0., 0., 0., 0.}, {0, 0}, 0};
  // scalar_sweep: the density sweep when the smallest density is below delta
  {
    double m = __builtin_inf();
    unsigned char f = 0;
    for (long long q = q0; q < q1; ++q) {
      const int i = (int)(q % k.nx) + 1, j = (int)(q / k.nx) + 1;
      min_push(m, f, U[c7(k, i, j)]);
    }
    sh.pm[0][tid] = m;
    sh.pf[0][tid] = f;
    __syncthreads();
    if (tid == 0) sh.stop = min_combine(sh.pm[0], sh.pf[0]) < 1e-12 ? 1 : 0;
    __syncthreads();
    if (sh.stop) {
      pass7(k, b, w, PASS_DEN, true, sh, tid);
      pass7(k, b, w, PASS_DEN, false, sh, tid);
      if (tid == 0) g->den_count++;
    }
  }
  // u_bar in the reference's order (j outer, i inner), one thread per component
  double ubs = 0.;
  for (long long c0 = 0; c0 < N; c0 += SW7_CH) {
    const int n = (int)((N - c0 < SW7_CH) ? N - c0 : SW7_CH);
    for (int t = tid; t < NC * n; t += SW7_T) {
      const int c = t / n, e = t - c * n;
      const long long q = c0 + e;
      const int i = (int)(q % k.nx) + 1, j = (int)(q / k.nx) + 1;
      sh.ubc[c][e] = U[c * k.plane + c7(k, i, j)];
    }
    __syncthreads();
    if (tid < NC)
      for (int e = 0; e < n; ++e) ubs = ubs + k.weight * sh.ubc[tid][e];
    __syncthreads();
  }
  if (tid < NC) sh.ub[tid] = ubs;
  __syncthreads();
  if (tid == 0) {
    const double *ub = sh.ub;
    const double pbar =
        k.gamma1 * (((ub[3] - (0.5 * (ub[1] * ub[1])) / ub[0]) - (0.5 * (ub[2] * ub[2])) / ub[0]) - k.q * ub[4]);
    sh.stop = 0;
    if (pbar <= 0.) {
      g->status = ST7_PBAR;
      g->pbar = pbar;
      sh.stop = 1;
    } else {
      g->sweep_count++;
    }
  }
  __syncthreads();
  if (sh.stop) return;
  for (int c = 0; c < NC; ++c) w.ub[c] = sh.ub[c];
  for (int q = 1; q <= k.swp_num; ++q) {
    w.q = q;
    const int kind = (q & 1) ? PASS_I : PASS_II;
    if (!pass7(k, b, w, kind, true, sh, tid)) return;
    if (!pass7(k, b, w, kind, false, sh, tid)) return;
    // convergence: the pressure without the reaction term, and the density
    double pm = __builtin_inf(), rm = __builtin_inf();
    unsigned char pf = 0, rf = 0;
    for (long long t = q0; t < q1; ++t) {
      const int i = (int)(t % k.nx) + 1, j = (int)(t / k.nx) + 1;
      const long long o = c7(k, i, j);
      const double r = U[o], m1 = U[k.plane + o], m2 = U[2 * k.plane + o], e = U[3 * k.plane + o];
      min_push(pm, pf, k.gamma1 * (e - (0.5 * (m1 * m1 + m2 * m2)) / r));
      min_push(rm, rf, r);
    }
    sh.pm[0][tid] = pm;
    sh.pf[0][tid] = pf;
    sh.pm[1][tid] = rm;
    sh.pf[1][tid] = rf;
    __syncthreads();
    if (tid == 0) {
      const double pmin = min_combine(sh.pm[0], sh.pf[0]), rmin = min_combine(sh.pm[1], sh.pf[1]);
      g->pre_count++;
      sh.stop = 0;
      if (pmin > k.eps && rmin > k.eps) {
        const long long s = g->sweep_count - 1;
        if (s < k.track_cap) {
          b.trk[3 * s] = q;
          b.trk[3 * s + 1] = w.cnt.neg;
          b.trk[3 * s + 2] = w.cnt.pos;
        }
        g->postag = 0;
        sh.stop = 1;
      } else if (q == k.swp_num) {
        g->status = ST7_SWEEP_FAIL;
        g->fail_q = q;
        sh.stop = 1;
      }
    }
    __syncthreads();
    if (sh.stop) return;
  }
}

// postag cleared at the start of each stage's checkpositivity
__global__ void k7_clear(B7 b) { b.g->postag = 0; }

// u(5, -2:nx+3, -2:ny+3) in the Fortran layout <-> the planes
__global__ void k7_from_fortran(K7 k, const double *__restrict__ src, double *__restrict__ dst) {
  const long long n = (long long)(k.nx + 6) * (k.ny + 6);
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < NC * n;
       t += (long long)gridDim.x * blockDim.x) {
    const int c = (int)(t % NC);
    const long long cell = t / NC;
    const int i = -2 + (int)(cell % (k.nx + 6)), j = -2 + (int)(cell / (k.nx + 6));
    dst[c * k.plane + c7(k, i, j)] = src[t];
  }
}
__global__ void k7_to_fortran(K7 k, const double *__restrict__ src, double *__restrict__ dst) {
  const long long n = (long long)(k.nx + 6) * (k.ny + 6);
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < NC * n;
       t += (long long)gridDim.x * blockDim.x) {
    const int c = (int)(t % NC);
    const long long cell = t / NC;
    const int i = -2 + (int)(cell % (k.nx + 6)), j = -2 + (int)(cell / (k.nx + 6));
    dst[t] = src[c * k.plane + c7(k, i, j)];
  }
}

#endif  // SWF_COMMON_KERNELS

}  // namespace ex7
}  // namespace swf
