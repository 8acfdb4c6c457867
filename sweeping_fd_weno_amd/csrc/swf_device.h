// swf_device.h - shared device-side types and the per-cell physics of the
// characteristic-wise FD WENO-JS scheme with Lax-Friedrichs splitting.
//
// Templated on the precision mode M:
//   M == 0  FP64 exact: the reference's operation order, no contraction (the
//           translation unit is compiled with -ffp-contract=off), binary128 sites
//           emulated in double-double (swf_dd.h).  Bit-identical to the reference.
//   M == 1  FP64: the same formulas in binary64, binary128 sites evaluated in
//           binary64, contraction allowed.
//   M == 2  FP32-bulk: smoothness indicators and nonlinear weights in FP32 with
//           hardware reciprocals; state, fluxes, projections, RK, pressure,
//           positivity tests and the sweep in FP64.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "swf_q128.h"

#define SWF_D __device__ __forceinline__

namespace swf {

enum { MODE_EXACT = 0, MODE_FP64 = 1, MODE_FP32 = 2 };
enum { BC_SEDOV = 0, BC_PERIODIC = 1, BC_JET_HALF = 2, BC_JET_FULL = 3 };
enum { ST_OK = 0, ST_PBAR = 1, ST_SWEEP_FAIL = 2 };

// Kernel-side configuration (plain values, passed by value to every kernel).
struct KCfg {
  int nx, ny, P;     // interior size, row pitch (doubles) of every plane
  int nwords;        // words of one sweep bitmap: ceil(nx*ny/32)
  long long plane;   // doubles per plane: (ny+6)*P
  double dx, dy, gamma, gamma1, cfl, eps, eps2, delta, weight, dt_num;
  double xlo, ylo, jet_rho, jet_mx, jet_e, amb_rho, amb_e, jet_halfwidth;
  double alpha_const;  // alpha of the prescribed boundary states (jet inflow/ambient), else 0
  int bc, use_limiter, swp_num, track_cap;
  int alpha_in_scheme, roe_eps_double, ydens_plain, eps2_quad_literal, findalpha_ex3;
  int dt_rule_dx53, poscheck_le, success_ge, pbar_check, fail_stop;
};

// Device-resident scalars of one solver context.
struct DevScalars {
  double time[2];            // by step parity
  int active[2];             // step gate by parity
  double tfinal, dt;
  unsigned long long alpha_bits[4];  // alpha max: inputs of stage 1,2,3; [3] queries
  int postag, swept, status, rho_any;
  unsigned done_blocks;      // update workgroups finished in this stage
  long long steps_done;
  long long sweep_count, den_count, pre_count, stage_sweeps;
  long long total_sweep, max_sweep, total_neg, max_neg, total_pos, max_pos;
  long long fail_q;          // q at which the last failure happened
  double pbar;               // P(u_bar) of the last pbar failure
};

// Device buffers of one solver context.
struct Bufs {
  double *U, *U0, *Fx, *Fy, *R;
  unsigned *bmR, *bmP, *bmA, *bmB;  // sweep bitmaps (order positions)
  DevScalars *g;
  long long *trk;                   // 3 * track_cap: q, neg, pos per sweep trigger
  double *apart;                    // alpha max of each update workgroup
};

// index of cell (i, j) (Fortran indices, ghosts from -2) inside a plane
SWF_D long long cidx(const KCfg &k, int i, int j) { return (long long)(j + 2) * k.P + (i + 2); }

// ---------------------------------------------------------------------------
// Sweep orders (ex4_sedov_2d/main.f90:1836-1895), closed form, 0-based positions
// ---------------------------------------------------------------------------
SWF_D int posI(const KCfg &k, int i, int j) { return (j - 1) * k.nx + ((j & 1) ? (i - 1) : (k.nx - i)); }
SWF_D int posII(const KCfg &k, int i, int j) { return (i - 1) * k.ny + ((i & 1) ? (j - 1) : (k.ny - j)); }
SWF_D void cellI(const KCfg &k, int p, int &i, int &j) {
  j = p / k.nx + 1;
  int r = p - (j - 1) * k.nx;
  i = (j & 1) ? r + 1 : k.nx - r;
}
SWF_D void cellII(const KCfg &k, int p, int &i, int &j) {
  i = p / k.ny + 1;
  int r = p - (i - 1) * k.ny;
  j = (i & 1) ? r + 1 : k.ny - r;
}

// Fortran MAX / MIN of two non-NaN reals
SWF_HD double fmax2(double a, double b) { return (b > a) ? b : a; }
SWF_HD double fmin2(double a, double b) { return (b < a) ? b : a; }

// ---------------------------------------------------------------------------
// Point physics
// ---------------------------------------------------------------------------

// rho < eps2 test of fflux/gflux (ex4: REAL(8) eps2; ex3/ex5: the binary128 1d-8)
SWF_D bool below_eps2(const KCfg &k, double rho) {
  return k.eps2_quad_literal ? (rho < qconst::e8_hi) : (rho < k.eps2);
}

// Lax-Friedrichs split of the Euler flux: fp = 0.5*(u + f/a), fn = 0.5*(u - f/a)
// (ex4_sedov_2d/main.f90:125-177 fflux/gflux, :223-227 split).  DIR 0 = x, 1 = y.
template <int M, int DIR>
SWF_D void split_point(const KCfg &k, const double u[4], double alpha, double ia, double fp[4], double fn[4]) {
  double rho = u[0], mx = u[1], my = u[2], e = u[3];
  double vx = mx / rho, vy = my / rho;
  if (below_eps2(k, rho)) {
    vx = 0.;
    vy = 0.;
  }
  double pres = k.gamma1 * ((e - (0.5 * rho) * (vx * vx)) - (0.5 * rho) * (vy * vy));
  double f[4];
  if (DIR == 0) {
    f[0] = mx;
    f[1] = rho * (vx * vx) + pres;
    f[2] = (rho * vx) * vy;
    f[3] = (e + pres) * vx;
  } else {
    f[0] = my;
    f[1] = (rho * vx) * vy;
    f[2] = rho * (vy * vy) + pres;
    f[3] = (e + pres) * vy;
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    double fa = (M == MODE_FP32) ? f[c] * ia : f[c] / alpha;
    fp[c] = 0.5 * (u[c] + fa);
    fn[c] = 0.5 * (u[c] - fa);
  }
}

// Primitive data for the Roe average (ex4_sedov_2d/main.f90:201-214 x, :421-434 y)
template <int M, int DIR>
SWF_D void roe_prims(const KCfg &k, const double u[4], double &w, double &vx, double &vy, double &h) {
  double den = u[0], xmt = u[1], ymt = u[2], eng = u[3];
  double t0;
  if (DIR == 1 && !k.ydens_plain) {
    if (M == MODE_EXACT)
      t0 = yrecip_q(den);
    else
      t0 = 1. / (den + 1e-15);
  } else {
    t0 = 1. / den;
  }
  double vex = xmt * t0, vey = ymt * t0;
  double pre = k.gamma1 * (eng - 0.5 * (xmt * vex + ymt * vey));
  w = sqrt(fabs(den));
  vx = vex;
  vy = vey;
  h = (pre + eng) * t0;
}

// ---------------------------------------------------------------------------
// WENO-JS reconstructions (weno5.f90:86-146)
// ---------------------------------------------------------------------------

// c(-1:2,0:2) of weno5.f90:3-14
#define SWF_C_M1_0 1.83333333333333333333333333333333
#define SWF_C_M1_1 (-1.16666666666666666666666666666667)
#define SWF_C_M1_2 0.333333333333333333333333333333333
#define SWF_C_0_0 0.333333333333333333333333333333333
#define SWF_C_0_1 0.833333333333333333333333333333333
#define SWF_C_0_2 (-0.166666666666666666666666666666667)
#define SWF_C_1_0 (-0.166666666666666666666666666666667)
#define SWF_C_1_1 0.833333333333333333333333333333333
#define SWF_C_1_2 0.333333333333333333333333333333333
#define SWF_C_2_0 0.333333333333333333333333333333333
#define SWF_C_2_1 (-1.16666666666666666666666666666667)
#define SWF_C_2_2 1.83333333333333333333333333333333

// Nonlinear weights in FP32 with hardware reciprocals (FP32-bulk mode).  The
// weights are d_k/(1e-6+beta_k)^2 normalised; scaling every (1e-6+beta_k) by the
// smallest one keeps all intermediates in [0, 1] so nothing overflows in FP32.
SWF_D void weights_fp32(const double *a, float d0, float d1, float d2, float w[3]) {
  float a2 = (float)a[-2], a1m = (float)a[-1], a0 = (float)a[0], a1 = (float)a[1], a2p = (float)a[2];
  float x1 = a0 - 2.f * a1 + a2p, y1 = 3.f * a0 - 4.f * a1 + a2p;
  float x2 = a1m - 2.f * a0 + a1, y2 = a1m - a1;
  float x3 = a2 - 2.f * a1m + a0, y3 = a2 - 4.f * a1m + 3.f * a0;
  const float c = 13.f / 12.f;
  float s1 = 1e-6f + (c * x1 * x1 + 0.25f * y1 * y1);
  float s2 = 1e-6f + (c * x2 * x2 + 0.25f * y2 * y2);
  float s3 = 1e-6f + (c * x3 * x3 + 0.25f * y3 * y3);
  float m = fminf(s1, fminf(s2, s3));
  float r1 = m * __builtin_amdgcn_rcpf(s1), r2 = m * __builtin_amdgcn_rcpf(s2), r3 = m * __builtin_amdgcn_rcpf(s3);
  float al1 = d0 * r1 * r1, al2 = d1 * r2 * r2, al3 = d2 * r3 * r3;
  float inv = __builtin_amdgcn_rcpf(al1 + al2 + al3);
  w[0] = al1 * inv;
  w[1] = al2 * inv;
  w[2] = al3 * inv;
}

// weno_right: value at x_{i+1/2}; `a` points at average(0)
template <int M>
SWF_D double weno_right(const double *a) {
  double um1 = ((0.0 + SWF_C_2_0 * a[-2]) + SWF_C_2_1 * a[-1]) + SWF_C_2_2 * a[0];
  double um2 = ((0.0 + SWF_C_1_0 * a[-1]) + SWF_C_1_1 * a[0]) + SWF_C_1_2 * a[1];
  double um3 = ((0.0 + SWF_C_0_0 * a[0]) + SWF_C_0_1 * a[1]) + SWF_C_0_2 * a[2];
  if (M == MODE_FP32) {
    float w[3];
    weights_fp32(a, 0.3f, 0.6f, 0.1f, w);
    return ((double)w[0] * um3 + (double)w[1] * um2) + (double)w[2] * um1;
  }
  double b[3];
  if (M == MODE_EXACT) {
    beta_q(a, b);
  } else {
    b[0] = (13. / 12.) * ((a[0] - 2. * a[1] + a[2]) * (a[0] - 2. * a[1] + a[2])) +
           0.25 * ((3. * a[0] - 4. * a[1] + a[2]) * (3. * a[0] - 4. * a[1] + a[2]));
    b[1] = (13. / 12.) * ((a[-1] - 2. * a[0] + a[1]) * (a[-1] - 2. * a[0] + a[1])) +
           0.25 * ((a[-1] - a[1]) * (a[-1] - a[1]));
    b[2] = (13. / 12.) * ((a[-2] - 2. * a[-1] + a[0]) * (a[-2] - 2. * a[-1] + a[0])) +
           0.25 * ((a[-2] - 4. * a[-1] + 3. * a[0]) * (a[-2] - 4. * a[-1] + 3. * a[0]));
  }
  double al1, al2, al3;
  if (M == MODE_EXACT) {
    al1 = alpha_q(0.3, b[0]);
    al2 = alpha_q(0.6, b[1]);
    al3 = alpha_q(0.1, b[2]);
  } else {
    al1 = 0.3 / ((1e-6 + b[0]) * (1e-6 + b[0]));
    al2 = 0.6 / ((1e-6 + b[1]) * (1e-6 + b[1]));
    al3 = 0.1 / ((1e-6 + b[2]) * (1e-6 + b[2]));
  }
  double s = (al1 + al2) + al3;
  double w1 = al1 / s, w2 = al2 / s, w3 = al3 / s;
  return (w1 * um3 + w2 * um2) + w3 * um1;
}

// weno_left: value at x_{i-1/2}; `a` points at average(0)
template <int M>
SWF_D double weno_left(const double *a) {
  double up1 = ((0.0 + SWF_C_1_0 * a[-2]) + SWF_C_1_1 * a[-1]) + SWF_C_1_2 * a[0];
  double up2 = ((0.0 + SWF_C_0_0 * a[-1]) + SWF_C_0_1 * a[0]) + SWF_C_0_2 * a[1];
  double up3 = ((0.0 + SWF_C_M1_0 * a[0]) + SWF_C_M1_1 * a[1]) + SWF_C_M1_2 * a[2];
  if (M == MODE_FP32) {
    float w[3];
    weights_fp32(a, 0.1f, 0.6f, 0.3f, w);
    return ((double)w[0] * up3 + (double)w[1] * up2) + (double)w[2] * up1;
  }
  double b[3];
  if (M == MODE_EXACT) {
    beta_q(a, b);
  } else {
    b[0] = (13. / 12.) * ((a[0] - 2. * a[1] + a[2]) * (a[0] - 2. * a[1] + a[2])) +
           0.25 * ((3. * a[0] - 4. * a[1] + a[2]) * (3. * a[0] - 4. * a[1] + a[2]));
    b[1] = (13. / 12.) * ((a[-1] - 2. * a[0] + a[1]) * (a[-1] - 2. * a[0] + a[1])) +
           0.25 * ((a[-1] - a[1]) * (a[-1] - a[1]));
    b[2] = (13. / 12.) * ((a[-2] - 2. * a[-1] + a[0]) * (a[-2] - 2. * a[-1] + a[0])) +
           0.25 * ((a[-2] - 4. * a[-1] + 3. * a[0]) * (a[-2] - 4. * a[-1] + 3. * a[0]));
  }
  double al1, al2, al3;
  if (M == MODE_EXACT) {
    al1 = alpha_q(0.1, b[0]);
    al2 = alpha_q(0.6, b[1]);
    al3 = alpha_q(0.3, b[2]);
  } else {
    al1 = 0.1 / ((1e-6 + b[0]) * (1e-6 + b[0]));
    al2 = 0.6 / ((1e-6 + b[1]) * (1e-6 + b[1]));
    al3 = 0.3 / ((1e-6 + b[2]) * (1e-6 + b[2]));
  }
  double s = (al1 + al2) + al3;
  double w1 = al1 / s, w2 = al2 / s, w3 = al3 / s;
  return (w1 * up3 + w2 * up2) + w3 * up1;
}

// ---------------------------------------------------------------------------
// One half of an interface flux: Roe matrices at (a|b), projection of the
// 5-point stencil `st` (stride 4 doubles, pointer at the centre point),
// 4 reconstructions and back-projection (ex4_sedov_2d/main.f90:228-304, :452-526).
// RIGHT selects weno_right (the f+ part) or weno_left (the f- part).
// ---------------------------------------------------------------------------
struct Prim {
  double w, vx, vy, h;
};

template <int M, int DIR, bool RIGHT>
SWF_D void half_flux(const KCfg &k, const Prim &pa, const Prim &pb, const double *st, int sstride,
                     double out[4]) {
  const double g1 = k.gamma1;
  double t0;
  if (M == MODE_EXACT && !k.roe_eps_double)
    t0 = roe_weight_q(pa.w, pb.w);
  else
    t0 = pa.w / ((pa.w + pb.w) + 1e-15);
  double t1 = 1. - t0;
  double vxm = t0 * pa.vx + t1 * pb.vx;
  double vym = t0 * pa.vy + t1 * pb.vy;
  double hm = t0 * pa.h + t1 * pb.h;
  double qm = 0.5 * (vxm * vxm + vym * vym);
  double sq = sqrt(fabs(g1 * (hm - qm)));
  double cm;
  if (M == MODE_EXACT && !k.roe_eps_double)
    cm = cm_q(sq);
  else
    cm = sq + 1e-15;
  // normal (vn) and tangential (vt) velocity for this direction
  const double vn = (DIR == 0) ? vxm : vym;
  const double vt = (DIR == 0) ? vym : vxm;
  const double tc = vn * cm;
  double rcm = 1. / cm;
  double b1 = (g1 * rcm) * rcm;
  double b2 = qm * b1;
  double tr = vn * rcm;
  double tn = b1 * vn;
  double t2 = 0.5 * b1;
  double tt = b1 * vt;
  // Left eigenvectors, rows k = 1..4, written in (normal, tangential) form;
  // component order (rho, m_n, m_t, E) is mapped to (rho, mx, my, E) below.
  double L[4][4];
  L[0][0] = 0.5 * (b2 + tr); L[0][1] = -(0.5 * (tn + rcm)); L[0][2] = -(0.5 * tt); L[0][3] = t2;
  L[1][0] = -vt;             L[1][1] = 0.0;                 L[1][2] = 1.0;           L[1][3] = 0.0;
  L[2][0] = 1. - b2;         L[2][1] = tn;                  L[2][2] = tt;            L[2][3] = -b1;
  L[3][0] = 0.5 * (b2 - tr); L[3][1] = -(0.5 * (tn - rcm)); L[3][2] = -(0.5 * tt); L[3][3] = t2;
  // Right eigenvectors (columns), (rho, m_n, m_t, E) rows
  double R[4][4];
  R[0][0] = 1.0;     R[0][1] = 0.0; R[0][2] = 1.0; R[0][3] = 1.0;
  R[1][0] = vn - cm; R[1][1] = 0.0; R[1][2] = vn;  R[1][3] = vn + cm;
  R[2][0] = vt;      R[2][1] = 1.0; R[2][2] = vt;  R[2][3] = vt;
  R[3][0] = hm - tc; R[3][1] = vt;  R[3][2] = qm;  R[3][3] = hm + tc;
  // physical component index of the (n, t) slot
  const int cn = (DIR == 0) ? 1 : 2, ct = (DIR == 0) ? 2 : 1;
  double r[4];
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    // evl(m, :) in physical component order
    double e[4];
    e[0] = L[m][0];
    e[cn] = L[m][1];
    e[ct] = L[m][2];
    e[3] = L[m][3];
    double cu[5];
#pragma unroll
    for (int ig = -2; ig <= 2; ++ig) {
      const double *f = st + ig * sstride;
      if (M == MODE_EXACT)
        cu[ig + 2] = ((e[0] * f[0] + e[1] * f[1]) + e[2] * f[2]) + e[3] * f[3];
      else if (m == 1)  // structural zeros/ones of row 2 dropped outside the exact mode
        cu[ig + 2] = e[0] * f[0] + f[ct];
      else
        cu[ig + 2] = ((e[0] * f[0] + e[1] * f[1]) + e[2] * f[2]) + e[3] * f[3];
    }
    r[m] = RIGHT ? weno_right<M>(&cu[2]) : weno_left<M>(&cu[2]);
  }
  // back-projection, rows in physical component order
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int rr = (c == 0) ? 0 : (c == 3) ? 3 : (c == cn) ? 1 : 2;
    if (M == MODE_EXACT)
      out[c] = ((R[rr][0] * r[0] + R[rr][1] * r[1]) + R[rr][2] * r[2]) + R[rr][3] * r[3];
    else if (rr == 0)
      out[c] = (r[0] + r[2]) + r[3];
    else
      out[c] = ((R[rr][0] * r[0] + R[rr][1] * r[1]) + R[rr][2] * r[2]) + R[rr][3] * r[3];
  }
}

// ---------------------------------------------------------------------------
// findalpha per cell (ex4_sedov_2d/main.f90:652-683; ex3 :549-557)
// ---------------------------------------------------------------------------
SWF_HD double cell_alpha(const KCfg &k, double u1, double u2, double u3, double u4) {
  double den = u1;
  if (k.findalpha_ex3) {
    double pres = k.gamma1 * ((u4 - ((u2 * u2) / u1) / 2.) - ((u3 * u3) / u1) / 2.);
    double sound = sqrt((k.gamma * pres) / den);
    return fmax2(fabs(u2 / u1) + sound, fabs(u3 / u1) + sound);
  }
  double vx, vy;
  if (den >= k.eps2) {
    vx = u2 / den;
    vy = u3 / den;
  } else {
    vx = 0.;
    vy = 0.;
  }
  double pres = k.gamma1 * ((u4 - (0.5 * den) * (vx * vx)) - (0.5 * den) * (vy * vy));
  double sound = sqrt((k.gamma * pres) / den);
  return fmax2(fabs(u2 / u1) + sound, fabs(u3 / u1) + sound);
}

// sweep flag: P <= eps or rho <= eps (ex4_sedov_2d/main.f90:1465-1466)
SWF_D double sweep_pres(double g1, double v0, double v1, double v2, double v3) {
#pragma clang fp contract(off)
  return g1 * (v3 - (0.5 * (v1 * v1 + v2 * v2)) / v0);
}
SWF_D bool sweep_flag(double g1, double eps, double v0, double v1, double v2, double v3) {
#pragma clang fp contract(off)
  double p = sweep_pres(g1, v0, v1, v2, v3);
  return (p <= eps) || (v0 <= eps);
}

// wave-wide max of a non-negative double, then one atomic per wave
SWF_D void alpha_atomic_max(unsigned long long *slot, double a) {
  unsigned long long bits = __double_as_longlong(a);
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    unsigned long long o = __shfl_xor(bits, off, 64);
    bits = (o > bits) ? o : bits;
  }
  if ((threadIdx.x + threadIdx.y * blockDim.x) % 64 == 0) atomicMax(slot, bits);
}

}  // namespace swf
