// swf_device.h - shared device-side types and the per-cell physics of the
// characteristic-wise FD WENO-JS scheme with Lax-Friedrichs splitting.
//
// Templated on the precision mode M:
//   M == 0  FP64 exact: the reference's operation order, no contraction (the
//           translation unit is compiled with -ffp-contract=off), binary128 sites
//           emulated in double-double (swf_dd.h).  Bit-identical to the reference.
//   M == 1  FP64: the same scheme in binary64 with contraction; reconstructions in
//           first-difference form with one reciprocal each (swf_weno.h).
//   M == 2  FP32-bulk: as M == 1, with the smoothness indicators and nonlinear
//           weights in FP32; state, split fluxes, eigen-projections, candidate
//           values, RK, pressure, positivity tests and the sweep in FP64.
//   M == 3  FP32-full: as M == 2, with the eigen-projections of the flux
//           differences, the candidate values, the reconstruction and the
//           back-projection in FP32 too; the Roe averages, the split fluxes, the
//           centre values, RK, pressure, positivity tests and the sweep in FP64.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "swf_q128.h"
#include "swf_weno.h"

#define SWF_D __device__ __forceinline__
// the per-point physics below is also compiled for the host (host checks of the
// fast modes against the oracle); kernels and wave-level helpers stay device-only

namespace swf {

enum { MODE_EXACT = 0, MODE_FP64 = 1, MODE_FP32 = 2, MODE_FP32F = 3 };
enum { BC_SEDOV = 0, BC_PERIODIC = 1, BC_JET_HALF = 2, BC_JET_FULL = 3, BC_STEP = 4 };
enum { ST_OK = 0, ST_PBAR = 1, ST_SWEEP_FAIL = 2, ST_PSUM = 3 };
enum { YS_RHS_ONLY = 1, YS_NO_SCALARS = 2 };  // flags of the stage kernel

// Kernel-side configuration (plain values, passed by value to every kernel).
struct KCfg {
  int nx, ny, P;     // interior size, row pitch (doubles) of every plane
  int nwords;        // words of one sweep bitmap: ceil(nx*ny/32)
  long long plane;   // doubles per plane: (ny+6)*P
  double dx, dy, gamma, gamma1, cfl, eps, eps2, delta, weight, dt_num;
  double idx, idy, ig1;   // 1/dx, 1/dy, 1/gamma1 (FP64 and FP32-bulk modes)
  double xlo, ylo, jet_rho, jet_mx, jet_e, amb_rho, amb_e, jet_halfwidth;
  double alpha_const;  // alpha of the prescribed boundary states (jet inflow/ambient), else 0
  int bc, use_limiter, swp_num, track_cap;
  int step_total, alpha_lines, weno_eps12, psum_check;  // the step diffraction program
  double in_rho, in_mx, in_e;                           // its post-shock inflow state
  int alpha_in_scheme, roe_eps_double, ydens_plain, eps2_quad_literal, findalpha_ex3;
  int dt_rule_dx53, poscheck_le, success_ge, pbar_check, fail_stop;
  int fast_div;  // the sweep's divisions by refined reciprocals (FP64 / FP32 modes)
};

// the cells the step program's bc sets: the ghost frame (3 full rows below and
// above, 3 columns left and right of the interior rows), then column 1 of rows
// 1 .. 6*step_total-1
SWF_HD long long step_bc_cells(const KCfg &k) {
  const long long W = k.nx + 6;
  const int lim = 6 * k.step_total - 1;
  const int nres = k.ny < lim ? k.ny : lim;
  return 6 * W + 6LL * k.ny + (nres > 0 ? nres : 0);
}

// per-line alpha tiles (the step program): 64 columns x 8 rows of the frame
constexpr int LA_W = 64, LA_H = 8;
SWF_HD int la_gx(const KCfg &k) { return (k.nx + 6 + LA_W - 1) / LA_W; }
SWF_HD int la_gy(const KCfg &k) { return (k.ny + 6 + LA_H - 1) / LA_H; }
// their reduction: 8 lanes per row or column line, 256 threads per block
constexpr int LA_LANES = 8;
SWF_HD int la_reduce_blocks(const KCfg &k) { return ((k.nx + 6 + k.ny + 6) * LA_LANES + 255) / 256; }
// the step's wall images (ex6): 3 cells per row 0 .. min(ny+1, 6T) before the
// x-pass (dir 0), 3 per column 1..T before the y-pass (dir 1)
SWF_HD int step_reflect_cells(const KCfg &k, int dir) {
  return dir == 0 ? 3 * ((k.ny + 1 < 6 * k.step_total ? k.ny + 1 : 6 * k.step_total) + 1) : 3 * k.step_total;
}

// Device-resident scalars of one solver context.
struct DevScalars {
  double time[2];            // by step parity
  int active[2];             // step gate by parity
  double tfinal, dt;
  unsigned long long alpha_bits[4];  // alpha max: inputs of stage 1,2,3; [3] queries
  int postag, swept, status, rho_any;
  int ubar_tiles;            // the stage kernel sums weight*u per workgroup (fast modes,
                             // after a stage that swept)
  unsigned done_blocks;      // update workgroups finished in this stage
  long long steps_done;
  long long sweep_count, den_count, pre_count, stage_sweeps;
  long long total_sweep, max_sweep, total_neg, max_neg, total_pos, max_pos;
  long long fail_q;          // q at which the last failure happened
  double pbar;               // P(u_bar) of the last pbar failure
};

// The sweep's workgroup size, and the steps of a speculative walk it records per
// thread (longer walks are computed again when they write; 64 MB of records).
constexpr int SW_T = 512;
constexpr int SW_REC = 4096;

// Device buffers of one solver context.  A step cycles three state buffers
// S0 -> S1 -> S2 -> S0: stage s reads its stencils from Uin and writes U; the
// RK combinations of stages 2 and 3 read the step's initial state U0 (= S0)
// pointwise, so stage 3 writes S0 in place.  Outside a stage Uin == U == S0.
struct Bufs {
  double *U, *Uin, *U0, *Fx, *R;
  unsigned *bmR, *bmP, *bmA, *bmB;  // sweep bitmaps (order positions)
  DevScalars *g;
  long long *trk;                   // 3 * track_cap: q, neg, pos per sweep trigger
  double *apart;                    // alpha max of each update workgroup
  double *upart;                    // 4 per update workgroup: sum of weight*u over its cells
  double *wrec;                     // sweep walk records: SW_REC rows x SW_T slots x 4
  unsigned *wpos;                   // their branch bits: SW_REC/32 words x SW_T slots
  int *zs;                          // flux limiter: (den, pre) counts per workgroup of the
                                    // x-flux, stage and edge kernels (zs_pairs)
  double *arow, *acol;              // per-line alphas: max of row j at arow[j+2], of column
                                    // i at acol[i+2] (alpha_lines)
  double *lpart;                    // their partials per 64 x 8 cell tile
};

// One instance of a batch: its configuration, its buffers outside a stage (Uin ==
// U == U0 == S0) and the two other state buffers of the stage cycle.
struct BatchInst {
  KCfg k;
  Bufs b;
  double *S1, *S2;
};

// What the host's time loop reads of each instance of a batch.
struct BatchStatus {
  double time;
  long long steps_done;
  int status, active;
};

// Row layout of every plane: cell i (Fortran index, ghosts from -2) at column
// i + XOFF, so interior cell 1 starts a 128-byte line and each 16-cell tile row of
// the stage kernel is exactly one cache line; the pitch P is a multiple of 16.
constexpr int XOFF = 15;
SWF_HD int row_pitch(int nx) { return (nx + 3 + XOFF + 1 + 15) / 16 * 16; }

// index of cell (i, j) (ghost rows from j = -2) inside a plane
SWF_HD long long cidx(const KCfg &k, int i, int j) { return (long long)(j + 2) * k.P + (i + XOFF); }

// Update tiles: UB_W x UB_H interior cells per workgroup of YS_T threads of the
// stage kernel (one thread per y interface, UB_H + 1 interfaces per column); each
// writes one alpha partial, and the sweep marks the tiles it changed.  Taller
// tiles stage fewer points per cell (1.19 at 31 cells against 1.4), but 16 x 31
// cells on 512 threads ran the Sedov 640^2 stage in 28.7 us against 24.5 us.
constexpr int UB_W = 16, UB_H = 15, YS_T = UB_W * (UB_H + 1);
SWF_HD int ub_gx(const KCfg &k) { return (k.nx + UB_W - 1) / UB_W; }
SWF_HD int ub_gy(const KCfg &k) { return (k.ny + UB_H - 1) / UB_H; }
SWF_HD int ub_of(const KCfg &k, int i, int j) { return ((j - 1) / UB_H) * ub_gx(k) + (i - 1) / UB_W; }

// x-flux workgroups: 64 centres x 4 rows, tile origins XT_C bx covering the
// interfaces 0..nx (lanes 0..XT_C-1 of a row give interfaces)
constexpr int XT_C = 63;
SWF_HD int xf_gx(const KCfg &k) { return (k.nx + 1 + XT_C - 1) / XT_C; }
SWF_HD int xf_gy(const KCfg &k) { return (k.ny + 3) / 4; }

// Flux-limiter half fluxes that `scheme` does not read but the reference's
// wenofd limits and counts (use_limiter == 1): x-pass ghost rows 0 and ny+1
// (f+ and f- of cells 0..nx+1), f+ of cell nx+1 and f- of cell 0 in rows 1..ny,
// and the same for the y-pass.
SWF_HD int zs_edge_halves(const KCfg &k) { return 6 * (k.nx + k.ny) + 16; }
// (den, pre) count pairs: x-flux workgroups, stage workgroups, edge workgroups
SWF_HD int zs_pairs(const KCfg &k) {
  return xf_gx(k) * xf_gy(k) + ub_gx(k) * ub_gy(k) + (zs_edge_halves(k) + 255) / 256;
}

// ---------------------------------------------------------------------------
// Sweep orders (ex4_sedov_2d/main.f90:1836-1895), closed form, 0-based positions
// ---------------------------------------------------------------------------
SWF_HD int posI(const KCfg &k, int i, int j) { return (j - 1) * k.nx + ((j & 1) ? (i - 1) : (k.nx - i)); }
SWF_HD int posII(const KCfg &k, int i, int j) { return (i - 1) * k.ny + ((i & 1) ? (j - 1) : (k.ny - j)); }
SWF_HD void cellI(const KCfg &k, int p, int &i, int &j) {
  j = p / k.nx + 1;
  int r = p - (j - 1) * k.nx;
  i = (j & 1) ? r + 1 : k.nx - r;
}
SWF_HD void cellII(const KCfg &k, int p, int &i, int &j) {
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
SWF_HD bool below_eps2(const KCfg &k, double rho) {
  return k.eps2_quad_literal ? (rho < qconst::e8_hi) : (rho < k.eps2);
}

// Roe-average inputs of one stencil point: sqrt(rho), velocities, enthalpy
struct Prim {
  double w, vx, vy, h;
};

// Lax-Friedrichs split of the Euler flux: fp = 0.5*(u + f/a), fn = 0.5*(u - f/a)
// (ex4_sedov_2d/main.f90:125-177 fflux/gflux, :223-227 split).  DIR 0 = x, 1 = y.
template <int M, int DIR>
SWF_HD void split_point(const KCfg &k, const double u[4], double alpha, double ia, double fp[4], double fn[4]) {
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
    double fa = f[c] / alpha;
    fp[c] = 0.5 * (u[c] + fa);
    fn[c] = 0.5 * (u[c] - fa);
  }
}

// Primitive data for the Roe average (ex4_sedov_2d/main.f90:201-214 x, :421-434 y)
template <int M, int DIR>
SWF_HD void roe_prims(const KCfg &k, const double u[4], double &w, double &vx, double &vy, double &h) {
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

// Split fluxes and Roe primitives of one stencil point in the fast modes: the
// formulas of split_point and roe_prims with one reciprocal of rho shared by the
// velocities and a multiplication by 1/alpha.  Reciprocals and the square root
// are the refined hardware estimates (within 1 ulp).
template <int DIR>
SWF_HD void stage_point_fast(const KCfg &k, const double u[4], double ia, double fp[4], double fn[4], Prim &pr) {
  const double rho = u[0], mx = u[1], my = u[2], e = u[3];
  const double r = rcp64(rho);
  const double vx0 = mx * r, vy0 = my * r;
  const bool vac = below_eps2(k, rho);
  const double vx = vac ? 0.0 : vx0, vy = vac ? 0.0 : vy0;
  const double pres = k.gamma1 * ((e - (0.5 * rho) * (vx * vx)) - (0.5 * rho) * (vy * vy));
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
    const double fa = f[c] * ia;
    fp[c] = 0.5 * (u[c] + fa);
    fn[c] = 0.5 * (u[c] - fa);
  }
  double t0 = r, vex = vx0, vey = vy0;
  if (DIR == 1 && !k.ydens_plain) {
    t0 = rcp64(rho + 1e-15);
    vex = mx * t0;
    vey = my * t0;
  }
  const double pre = k.gamma1 * (e - 0.5 * (mx * vex + my * vey));
  pr.w = sqrt64(fabs(rho));
  pr.vx = vex;
  pr.vy = vey;
  pr.h = (pre + e) * t0;
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

// The constant of the WENO weights alpha_k = d_k/(eps + beta_k)**2: eps = 1d-6 as a
// binary128 double-double (exact mode) and 4*eps in FP64 / FP32 (fast modes).
struct WenoEps {
  double hi, lo, q4;
  float q4f;
};
constexpr WenoEps kWeps6 = {qconst::e6_hi, qconst::e6_lo, 4e-6, 4e-6f};
// ex6_shock_diffraction/weno5.f90:104-106,138-140
constexpr WenoEps kWeps12 = {qconst::e12_hi, qconst::e12_lo, 4e-12, 4e-12f};

// weno_right: value at x_{i+1/2}; `a` points at average(0).  The exact mode
// follows weno5.f90:120-146 operation by operation (beta and alpha at binary128);
// the fast modes use the first-difference forms of swf_weno.h.
template <int M>
SWF_HD double weno_right(const double *a, const WenoEps &we) {
  if (M == MODE_FP32F) {
    double e[4];
    first_differences(a, e);
    const float f[4] = {(float)e[0], (float)e[1], (float)e[2], (float)e[3]};
    return a[0] + (double)weno_d32f<true>(f, we.q4, we.q4f);
  }
  if (M != MODE_EXACT) {
    double e[4];
    first_differences(a, e);
    return a[0] + (M == MODE_FP64 ? weno_d64<true>(e, we.q4) : weno_d32_or_64<true>(e, we.q4, we.q4f));
  }
  const double um1 = ((0.0 + SWF_C_2_0 * a[-2]) + SWF_C_2_1 * a[-1]) + SWF_C_2_2 * a[0];
  const double um2 = ((0.0 + SWF_C_1_0 * a[-1]) + SWF_C_1_1 * a[0]) + SWF_C_1_2 * a[1];
  const double um3 = ((0.0 + SWF_C_0_0 * a[0]) + SWF_C_0_1 * a[1]) + SWF_C_0_2 * a[2];
  double b[3];
  beta_q(a, b);
  const double al1 = alpha_q(0.3, b[0], we.hi, we.lo), al2 = alpha_q(0.6, b[1], we.hi, we.lo),
               al3 = alpha_q(0.1, b[2], we.hi, we.lo);
  const double s = (al1 + al2) + al3;
  const double w1 = al1 / s, w2 = al2 / s, w3 = al3 / s;
  return (w1 * um3 + w2 * um2) + w3 * um1;
}

// weno_left: value at x_{i-1/2}; `a` points at average(0) (weno5.f90:86-113)
template <int M>
SWF_HD double weno_left(const double *a, const WenoEps &we) {
  if (M == MODE_FP32F) {
    double e[4];
    first_differences(a, e);
    const float f[4] = {(float)e[0], (float)e[1], (float)e[2], (float)e[3]};
    return a[0] + (double)weno_d32f<false>(f, we.q4, we.q4f);
  }
  if (M != MODE_EXACT) {
    double e[4];
    first_differences(a, e);
    return a[0] + (M == MODE_FP64 ? weno_d64<false>(e, we.q4) : weno_d32_or_64<false>(e, we.q4, we.q4f));
  }
  const double up1 = ((0.0 + SWF_C_1_0 * a[-2]) + SWF_C_1_1 * a[-1]) + SWF_C_1_2 * a[0];
  const double up2 = ((0.0 + SWF_C_0_0 * a[-1]) + SWF_C_0_1 * a[0]) + SWF_C_0_2 * a[1];
  const double up3 = ((0.0 + SWF_C_M1_0 * a[0]) + SWF_C_M1_1 * a[1]) + SWF_C_M1_2 * a[2];
  double b[3];
  beta_q(a, b);
  const double al1 = alpha_q(0.1, b[0], we.hi, we.lo), al2 = alpha_q(0.6, b[1], we.hi, we.lo),
               al3 = alpha_q(0.3, b[2], we.hi, we.lo);
  const double s = (al1 + al2) + al3;
  const double w1 = al1 / s, w2 = al2 / s, w3 = al3 / s;
  return (w1 * up3 + w2 * up2) + w3 * up1;
}

// ---------------------------------------------------------------------------
// One half of an interface flux: Roe matrices at (a|b), projection of the
// 5-point stencil `st` (pointer at component 0 of the centre point; points SS
// doubles apart, components CS doubles apart),
// 4 reconstructions and back-projection (ex4_sedov_2d/main.f90:228-304, :452-526).
// RIGHT selects weno_right (the f+ part) or weno_left (the f- part).
// ---------------------------------------------------------------------------
// FP64-exact: the reference's operations in its order, structural zeros kept
// (they decide signs of zero), binary128 sites emulated.
// The Roe system at (a|b) (ex4_sedov_2d/main.f90:232-276, :452-497): the scalars
// every entry of the eigenvector matrices is formed from, in the reference's
// expressions.  One system serves both halves of a centre (fp_flux(i) and
// fn_flux(i) share Roe(i, i+1)).
struct RoeExact {
  double vn, vt, hm, qm, cm, tc, rcm, b1, b2, tr, tn, t2, tt;
};
template <int DIR>
SWF_HD RoeExact roe_exact(const KCfg &k, const Prim &pa, const Prim &pb) {
  const double g1 = k.gamma1;
  RoeExact r;
  const double t0 = k.roe_eps_double ? pa.w / ((pa.w + pb.w) + 1e-15) : roe_weight_q(pa.w, pb.w);
  const double t1 = 1. - t0;
  const double vxm = t0 * pa.vx + t1 * pb.vx;
  const double vym = t0 * pa.vy + t1 * pb.vy;
  r.hm = t0 * pa.h + t1 * pb.h;
  r.qm = 0.5 * (vxm * vxm + vym * vym);
  const double sq = sqrt(fabs(g1 * (r.hm - r.qm)));
  r.cm = k.roe_eps_double ? sq + 1e-15 : cm_q(sq);
  // normal (vn) and tangential (vt) velocity for this direction
  r.vn = (DIR == 0) ? vxm : vym;
  r.vt = (DIR == 0) ? vym : vxm;
  r.tc = r.vn * r.cm;
  r.rcm = 1. / r.cm;
  r.b1 = (g1 * r.rcm) * r.rcm;
  r.b2 = r.qm * r.b1;
  r.tr = r.vn * r.rcm;
  r.tn = r.b1 * r.vn;
  r.t2 = 0.5 * r.b1;
  r.tt = r.b1 * r.vt;
  return r;
}

// One half of a centre's flux from its Roe system: projection of the 5-point
// stencil `st`, 4 reconstructions, back-projection (ex4_sedov_2d/main.f90:278-304).
// E12: the weights' constant is 1d-12 (the step program's weno5.f90), else 1d-6
template <int DIR, bool RIGHT, int SS, int CS, bool E12 = false>
SWF_HD void half_exact(const RoeExact &q, const double *st, double out[4]) {
  const WenoEps we = E12 ? kWeps12 : kWeps6;
  const double vn = q.vn, vt = q.vt, hm = q.hm, qm = q.qm, cm = q.cm, tc = q.tc, rcm = q.rcm;
  const double b1 = q.b1, b2 = q.b2, tr = q.tr, tn = q.tn, t2 = q.t2, tt = q.tt;
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
    double e[4];
    e[0] = L[m][0];
    e[cn] = L[m][1];
    e[ct] = L[m][2];
    e[3] = L[m][3];
    double cu[5];
#pragma unroll
    for (int ig = -2; ig <= 2; ++ig) {
      const double *f = st + ig * SS;
      cu[ig + 2] = ((e[0] * f[0] + e[1] * f[CS]) + e[2] * f[2 * CS]) + e[3] * f[3 * CS];
    }
    r[m] = RIGHT ? weno_right<MODE_EXACT>(&cu[2], we) : weno_left<MODE_EXACT>(&cu[2], we);
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int rr = (c == 0) ? 0 : (c == 3) ? 3 : (c == cn) ? 1 : 2;
    out[c] = ((R[rr][0] * r[0] + R[rr][1] * r[1]) + R[rr][2] * r[2]) + R[rr][3] * r[3];
  }
}

// One half flux with the Roe system at (a|b) (ex4_sedov_2d/main.f90:228-304, :452-526)
template <int DIR, bool RIGHT, int SS, int CS, bool E12 = false>
SWF_HD void half_flux_exact(const KCfg &k, const Prim &pa, const Prim &pb, const double *st, double out[4]) {
  half_exact<DIR, RIGHT, SS, CS, E12>(roe_exact<DIR>(k, pa, pb), st, out);
}

// FP64 / FP32-bulk: the same Roe eigen-system, with the reconstruction written
// as centre value plus correction (swf_weno.h):
//     out = f(0) + R * weno_d(L * (first differences of f))
// which is R * weno(L * f) in exact arithmetic (R = L^-1, L linear).  Only the
// correction passes through the eigenvectors and the nonlinear weights.
// The acoustic rows 1 and 4 of L share A = (b2 d0 - tn dn - tt dt + b1 d3) / 2 and
// differ by B = (vn d0 - dn) / (2 c): row 1 = A + B, row 4 = A - B, and the entropy
// row 3 is d0 - 2 A (10 operations per projected difference instead of 13).
// FP32-full (M == 3): the projections, the weights and candidates, and the
// back-projection in FP32; the Roe average, the centre value and the R * L energy
// term below in FP64.
// The Roe system of the fast modes at (a|b): what both halves of a centre need.
struct RoeFast {
  double vn, vt, hm, qm, hq, cm, rcm, b1, b2, tn, tt;
};
template <int DIR>
SWF_HD RoeFast roe_fast(const KCfg &k, const Prim &pa, const Prim &pb) {
  const double g1 = k.gamma1;
  RoeFast q;
  const double t0 = pa.w * rcp64((pa.w + pb.w) + 1e-15);
  const double t1 = 1. - t0;
  const double vxm = t0 * pa.vx + t1 * pb.vx;
  const double vym = t0 * pa.vy + t1 * pb.vy;
  q.hm = t0 * pa.h + t1 * pb.h;
  q.qm = 0.5 * (vxm * vxm + vym * vym);
  q.hq = q.hm - q.qm;
  q.cm = sqrt64(fabs(g1 * q.hq)) + 1e-15;
  q.vn = (DIR == 0) ? vxm : vym;
  q.vt = (DIR == 0) ? vym : vxm;
  q.rcm = rcp64(q.cm);
  q.b1 = (g1 * q.rcm) * q.rcm;
  q.b2 = q.qm * q.b1;
  q.tn = q.b1 * q.vn;
  q.tt = q.b1 * q.vt;
  return q;
}

// One half of a centre's flux in the fast modes from its Roe system.
template <int M, int DIR, bool RIGHT, int SS, int CS, bool E12 = false>
SWF_HD void half_fast(const KCfg &k, const RoeFast &rs, const double *st, double out[4]) {
  const WenoEps we = E12 ? kWeps12 : kWeps6;
  const double vn = rs.vn, vt = rs.vt, hm = rs.hm, qm = rs.qm, hq = rs.hq, cm = rs.cm, rcm = rs.rcm;
  const double b1 = rs.b1, b2 = rs.b2, tn = rs.tn, tt = rs.tt;
  const int cn = (DIR == 0) ? 1 : 2, ct = (DIR == 0) ? 2 : 1;
  double corr[4];
  if constexpr (M == MODE_FP32F) {
    const float hb2 = (float)(0.5 * b2), htn = (float)(0.5 * tn), htt = (float)(0.5 * tt), hb1 = (float)(0.5 * b1);
    const float hr = (float)(0.5 * rcm), fvn = (float)vn, fvt = (float)vt;
    float e[4][4];  // e[field][difference]
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const double *f0 = st + (q - 2) * SS, *f1 = st + (q - 1) * SS;
      const float d0 = (float)(f1[0] - f0[0]), dn = (float)(f1[cn * CS] - f0[cn * CS]),
                  dt = (float)(f1[ct * CS] - f0[ct * CS]), d3 = (float)(f1[3 * CS] - f0[3 * CS]);
      const float A = ((hb2 * d0 - htn * dn) - htt * dt) + hb1 * d3;
      const float B = hr * (fvn * d0 - dn);
      e[0][q] = A + B;
      e[1][q] = dt - fvt * d0;
      e[2][q] = d0 - 2.0f * A;
      e[3][q] = A - B;
    }
    float r[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) r[m] = weno_d32f<RIGHT>(e[m], we.q4, we.q4f);
    const float fcm = (float)cm, fhm = (float)hm, fqm = (float)qm;
    const float s = (r[0] + r[2]) + r[3], dr = r[3] - r[0];
    corr[0] = s;
    corr[cn] = fvn * s + fcm * dr;
    corr[ct] = fvt * s + r[1];
    corr[3] = (fhm * (r[0] + r[3]) + (fvn * fcm) * dr) + (fvt * r[1] + fqm * r[2]);
  } else {
    const double hb2 = 0.5 * b2, htn = 0.5 * tn, htt = 0.5 * tt, hb1 = 0.5 * b1, hr = 0.5 * rcm;
    double e[4][4];  // e[field][difference]
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const double *f0 = st + (q - 2) * SS, *f1 = st + (q - 1) * SS;
      const double d0 = f1[0] - f0[0], dn = f1[cn * CS] - f0[cn * CS], dt = f1[ct * CS] - f0[ct * CS],
                   d3 = f1[3 * CS] - f0[3 * CS];
      const double A = ((hb2 * d0 - htn * dn) - htt * dt) + hb1 * d3;
      const double B = hr * (vn * d0 - dn);
      e[0][q] = A + B;
      e[1][q] = dt - vt * d0;
      e[2][q] = d0 - 2.0 * A;
      e[3][q] = A - B;
    }
    double r[4];
#pragma unroll
    for (int m = 0; m < 4; ++m)
      r[m] = (M == MODE_FP64) ? weno_d64<RIGHT>(e[m], we.q4) : weno_d32_or_64<RIGHT>(e[m], we.q4, we.q4f);
    // R * r with the right eigenvector columns (1, vn-c, vt, H-vn c), (0, 0, 1, vt),
    // (1, vn, vt, q), (1, vn+c, vt, H+vn c)
    const double s = (r[0] + r[2]) + r[3], dr = r[3] - r[0];
    corr[0] = s;
    corr[cn] = vn * s + cm * dr;
    corr[ct] = vt * s + r[1];
    corr[3] = (hm * (r[0] + r[3]) + (vn * cm) * dr) + (vt * r[1] + qm * r[2]);
  }
  // The reference's R * L is not the identity: cm carries +1e-15, so
  // H != cm^2/gamma1 + q, and R * L = I + dH * e_E * (L row 1 + L row 4) with
  // dH = (hm - qm) - cm^2/gamma1.  Its centre value R * (L * f(0)) is therefore
  // f(0) plus this energy term (relative size ~1e-15/cm: 1e-12 where cm ~ 1e-3).
  const double dH = hq - (cm * cm) * k.ig1;
  corr[3] += dH * (((b2 * st[0] - tn * st[cn * CS]) - tt * st[ct * CS]) + b1 * st[3 * CS]);
#pragma unroll
  for (int c = 0; c < 4; ++c) out[c] = st[c * CS] + corr[c];
}

// One half flux with the Roe system at (a|b)
template <int M, int DIR, bool RIGHT, int SS, int CS, bool E12 = false>
SWF_HD void half_flux_fast(const KCfg &k, const Prim &pa, const Prim &pb, const double *st, double out[4]) {
  half_fast<M, DIR, RIGHT, SS, CS, E12>(k, roe_fast<DIR>(k, pa, pb), st, out);
}

template <int M, int DIR, bool RIGHT, int SS, int CS, bool E12 = false>
SWF_HD void half_flux(const KCfg &k, const Prim &pa, const Prim &pb, const double *st, double out[4]) {
  if (M == MODE_EXACT)
    half_flux_exact<DIR, RIGHT, SS, CS, E12>(k, pa, pb, st, out);
  else
    half_flux_fast<M, DIR, RIGHT, SS, CS, E12>(k, pa, pb, st, out);
}

// Both halves of centre i with the one Roe system of (i, i+1): fp_flux(i) =
// R weno_right(L fp) from the stencil `stp` and fn_flux(i) = R weno_left(L fn)
// from `stn` (ex4_sedov_2d/main.f90:228-304: the reference forms both in one
// iteration of its i loop).  Each result equals half_flux's for that half.
template <int M, int DIR, int SS, int CS, bool E12 = false>
SWF_HD void centre_fluxes(const KCfg &k, const Prim &pa, const Prim &pb, const double *stp, const double *stn,
                          double outp[4], double outn[4]) {
  if constexpr (M == MODE_EXACT) {
    const RoeExact q = roe_exact<DIR>(k, pa, pb);
    half_exact<DIR, true, SS, CS, E12>(q, stp, outp);
    half_exact<DIR, false, SS, CS, E12>(q, stn, outn);
  } else {
    const RoeFast q = roe_fast<DIR>(k, pa, pb);
    half_fast<M, DIR, true, SS, CS, E12>(k, q, stp, outp);
    half_fast<M, DIR, false, SS, CS, E12>(k, q, stn, outn);
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

// findalpha per cell in the fast modes: the same expression with one
// reciprocal of rho
SWF_HD double cell_alpha_fast(const KCfg &k, double u1, double u2, double u3, double u4) {
  const double r = rcp64(u1);
  const double ax = fabs(u2 * r), ay = fabs(u3 * r);
  double pres;
  if (k.findalpha_ex3) {
    pres = k.gamma1 * ((u4 - 0.5 * (u2 * u2) * r) - 0.5 * (u3 * u3) * r);
  } else {
    const bool vac = !(u1 >= k.eps2);
    const double vx = vac ? 0.0 : u2 * r, vy = vac ? 0.0 : u3 * r;
    pres = k.gamma1 * ((u4 - (0.5 * u1) * (vx * vx)) - (0.5 * u1) * (vy * vy));
  }
  const double sound = sqrt64((k.gamma * pres) * r);
  return fmax2(ax + sound, ay + sound);
}

// Zhang-Shu positivity limiter of one reconstructed split flux F (before the
// alpha scaling) against the split flux `ave` of its centre cell, components
// CS doubles apart (use_limiter == 1; ex4_sedov_2d/main.f90:306-412 x-pass,
// :527-629 y-pass).  Density first: theta from the points F and
// (ave - F/12)/(11/12); then pressure, with the density component of both points
// updated and the others as before.  The ex4 y-pass clamps the pressure theta of
// gp at 0 (:573), which for finite data (theta >= 0) changes nothing.  Returns
// bit 0 / bit 1 when the density / pressure step changed F (den_count /
// pre_count of the reference).
template <int CS>
SWF_HD int zs_limit(double g1, const double *ave, double F[4]) {
#pragma clang fp contract(off)
  int changed = 0;
  const double d_ave = ave[0], m_ave = ave[CS], n_ave = ave[2 * CS], e_ave = ave[3 * CS];
  const double p_ave = g1 * ((e_ave - (0.5 * (m_ave * m_ave)) / d_ave) - (0.5 * (n_ave * n_ave)) / d_ave);
  const double av[4] = {d_ave, m_ave, n_ave, e_ave};
  double pt[2][4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    pt[0][c] = F[c];
    pt[1][c] = (av[c] - F[c] / 12.) / (11. / 12.);
  }
  double theta = 1.;
#pragma unroll
  for (int ig = 0; ig < 2; ++ig) {
    const double d_l = pt[ig][0];
    if (d_l < 0.) theta = fmin2(theta, fabs(d_ave - 0.) / fabs(d_ave - d_l));
  }
  if (theta < 1.) {
    F[0] = theta * F[0] + (1. - theta) * d_ave;
    changed |= 1;
  }
  theta = 1.;
  pt[0][0] = F[0];
  pt[1][0] = (d_ave - F[0] / 12.) / (11. / 12.);
#pragma unroll
  for (int ig = 0; ig < 2; ++ig) {
    const double d_l = pt[ig][0], m_l = pt[ig][1], n_l = pt[ig][2], e_l = pt[ig][3];
    const double p_l = g1 * ((e_l - ((m_l * m_l) / d_l) / 2.) - ((n_l * n_l) / d_l) / 2.);
    if (p_l < 0.) theta = fmin2(theta, fabs(p_ave - 0.) / fabs(p_ave - p_l));
  }
  if (theta < 1.) {
#pragma unroll
    for (int c = 0; c < 4; ++c) F[c] = theta * F[c] + (1. - theta) * av[c];
    changed |= 2;
  }
  return changed;
}

// sweep flag: P <= eps or rho <= eps (ex4_sedov_2d/main.f90:1465-1466)
SWF_HD double sweep_pres(double g1, double v0, double v1, double v2, double v3) {
#pragma clang fp contract(off)
  return g1 * (v3 - (0.5 * (v1 * v1 + v2 * v2)) / v0);
}
SWF_HD bool sweep_flag(double g1, double eps, double v0, double v1, double v2, double v3) {
#pragma clang fp contract(off)
  double p = sweep_pres(g1, v0, v1, v2, v3);
  return (p <= eps) || (v0 <= eps);
}
// The fast modes' form: the division by rho a refined reciprocal (within 1 ulp) and
// the two products fused, g1 * (E - (m1^2 + m2^2) / (2 rho)).
SWF_HD double sweep_pres_fast(double g1, double v0, double v1, double v2, double v3) {
  return g1 * fma(-0.5 * fma(v1, v1, v2 * v2), rcp64(v0), v3);
}
// The same for a context: the fast form in the fast modes (k.fast_div), here and
// wherever the sweep forms a pressure, so that every flag test of a run uses one
// function.
SWF_HD double sweep_pres(const KCfg &k, double v0, double v1, double v2, double v3) {
  if (k.fast_div) return sweep_pres_fast(k.gamma1, v0, v1, v2, v3);
  return sweep_pres(k.gamma1, v0, v1, v2, v3);
}
// The same with the mode fixed at compile time (the sweep's speculative walks).
template <bool FAST>
SWF_HD double sweep_pres_t(double g1, double v0, double v1, double v2, double v3) {
  if (FAST) return sweep_pres_fast(g1, v0, v1, v2, v3);
  return sweep_pres(g1, v0, v1, v2, v3);
}
SWF_HD bool sweep_flag(const KCfg &k, double v0, double v1, double v2, double v3) {
#pragma clang fp contract(off)
  const double p = sweep_pres(k, v0, v1, v2, v3);
  return (p <= k.eps) || (v0 <= k.eps);
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
