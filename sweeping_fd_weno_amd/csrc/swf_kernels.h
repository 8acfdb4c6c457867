// swf_kernels.h - grid kernels of one SSP-RK3 stage, templated on precision mode.
//
// Per stage (ex4_sedov_2d/main.f90:685-777 `rk`):
//   xflux<M>   net x-interface flux  F(k+1/2) = fp_flux(k) - fn_flux(k+1)   (wenofd x-pass)
//   yflux<M>   net y-interface flux  G(j+1/2) = gp_flux(j) - gn_flux(j+1)   (wenofd y-pass)
//   update<M>  RHS (scheme), RK combination, checkpositivity, sweep flags, alpha max
//   sweep      (swf_sweep.h) single-workgroup order-preserving sweep, when flagged
//   alpha_fix  alpha max again when the sweep changed the state
//   bc         ghost cells (bc) and their alpha
// Each interface kernel stages the stencil points of its tile (split fluxes and
// Roe primitives) in LDS once; every thread then computes one interface: two
// Roe eigen-systems, eight WENO reconstructions, two back-projections.
#pragma once

#include "swf_device.h"

namespace swf {

// x tile: 64 interfaces x 4 rows; points k0-2 .. k0+66
constexpr int XT_W = 64, XT_H = 4, XT_PTS = XT_W + 5;
// y tile: 16 columns x 16 interfaces; points jf0-2 .. jf0+18
constexpr int YT_W = 16, YT_H = 16, YT_PTS = YT_H + 5;

SWF_D double alpha_from_bits(unsigned long long b) { return __longlong_as_double((long long)b); }

template <int M>
__global__ __launch_bounds__(256) void k_xflux(KCfg k, Bufs b, int slot_in, int slot_reset, int parity,
                                               int clear_maps) {
  const DevScalars *g = b.g;
  if (!g->active[parity]) return;
  const int tx = threadIdx.x, ty = threadIdx.y, tid = tx + ty * XT_W;
  const int k0 = blockIdx.x * XT_W, j0 = 1 + blockIdx.y * XT_H;
  // housekeeping for the update kernel of this stage
  if (clear_maps) {
    const long long gid = (long long)(blockIdx.x + blockIdx.y * gridDim.x) * 256 + tid;
    if (gid < k.nwords) {
      b.bmR[gid] = 0u;
      b.bmP[gid] = 0u;
    }
    if (gid == 0) {
      b.g->alpha_bits[slot_reset] = 0ull;
      b.g->postag = 0;
      b.g->rho_any = 0;
    }
  }
  const double alpha = alpha_from_bits(g->alpha_bits[slot_in]);
  const double ia = 1.0 / alpha;

  __shared__ double s_fp[XT_H][XT_PTS][4];
  __shared__ double s_fn[XT_H][XT_PTS][4];
  __shared__ Prim s_pr[XT_H][XT_PTS];

  for (int p = tid; p < XT_H * XT_PTS; p += 256) {
    const int r = p / XT_PTS, q = p - r * XT_PTS;
    const int i = k0 - 2 + q, j = j0 + r;
    if (j <= k.ny && i <= k.nx + 3) {
      const long long o = cidx(k, i, j);
      double u[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) u[c] = b.U[c * k.plane + o];
      split_point<M, 0>(k, u, alpha, ia, s_fp[r][q], s_fn[r][q]);
      Prim pr;
      roe_prims<M, 0>(k, u, pr.w, pr.vx, pr.vy, pr.h);
      s_pr[r][q] = pr;
    }
  }
  __syncthreads();
  const int kk = k0 + tx, j = j0 + ty;
  if (kk > k.nx || j > k.ny) return;
  double fp[4], fn[4];
  // f+ at interface kk (Roe(kk, kk+1), stencil kk-2..kk+2)
  half_flux<M, 0, true>(k, s_pr[ty][tx + 2], s_pr[ty][tx + 3], &s_fp[ty][tx + 2][0], 4, fp);
  // f- at interface kk+1 (Roe(kk+1, kk+2), stencil kk-1..kk+3)
  half_flux<M, 0, false>(k, s_pr[ty][tx + 3], s_pr[ty][tx + 4], &s_fn[ty][tx + 3][0], 4, fn);
  const long long o = cidx(k, kk, j);
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    double f = k.alpha_in_scheme ? (fp[c] - fn[c]) : (alpha * fp[c] - alpha * fn[c]);
    b.Fx[c * k.plane + o] = f;
  }
}

template <int M>
__global__ __launch_bounds__(256) void k_yflux(KCfg k, Bufs b, int slot_in, int parity) {
  const DevScalars *g = b.g;
  if (!g->active[parity]) return;
  const int tx = threadIdx.x, ty = threadIdx.y, tid = tx + ty * YT_W;
  const int i0 = 1 + blockIdx.x * YT_W, jf0 = blockIdx.y * YT_H;
  const double alpha = alpha_from_bits(g->alpha_bits[slot_in]);
  const double ia = 1.0 / alpha;

  __shared__ double s_fp[YT_PTS][YT_W][4];
  __shared__ double s_fn[YT_PTS][YT_W][4];
  __shared__ Prim s_pr[YT_PTS][YT_W];

  for (int p = tid; p < YT_PTS * YT_W; p += 256) {
    const int r = p / YT_W, q = p - r * YT_W;
    const int i = i0 + q, j = jf0 - 2 + r;
    if (i <= k.nx && j <= k.ny + 3) {
      const long long o = cidx(k, i, j);
      double u[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) u[c] = b.U[c * k.plane + o];
      split_point<M, 1>(k, u, alpha, ia, s_fp[r][q], s_fn[r][q]);
      Prim pr;
      roe_prims<M, 1>(k, u, pr.w, pr.vx, pr.vy, pr.h);
      s_pr[r][q] = pr;
    }
  }
  __syncthreads();
  const int i = i0 + tx, jf = jf0 + ty;
  if (i > k.nx || jf > k.ny) return;
  double fp[4], fn[4];
  half_flux<M, 1, true>(k, s_pr[ty + 2][tx], s_pr[ty + 3][tx], &s_fp[ty + 2][tx][0], YT_W * 4, fp);
  half_flux<M, 1, false>(k, s_pr[ty + 3][tx], s_pr[ty + 4][tx], &s_fn[ty + 3][tx][0], YT_W * 4, fn);
  const long long o = cidx(k, i, jf);
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    double f = k.alpha_in_scheme ? (fp[c] - fn[c]) : (alpha * fp[c] - alpha * fn[c]);
    b.Fy[c * k.plane + o] = f;
  }
}

// scheme: RHS of one interior cell from the net interface fluxes
SWF_D double cell_rhs(const KCfg &k, const Bufs &b, int c, long long o, double alpha) {
  const double *fx = b.Fx + c * k.plane, *fy = b.Fy + c * k.plane;
  double X = fx[o] - fx[o - 1];
  double Y = fy[o] - fy[o - k.P];
  if (k.alpha_in_scheme) return -((alpha * X) / k.dx) - (alpha * Y) / k.dy;
  return -(X / k.dx) - Y / k.dy;
}

// dt of the step (ex4_sedov_2d/main.f90:696-700; ex3 :570-572)
SWF_D double step_dt(const KCfg &k, double alpha, double time, double tfinal) {
  if (k.dt_rule_dx53) return k.dt_num / alpha;
  double dtx = (k.dx * k.cfl) / alpha;
  double dty = (k.dy * k.cfl) / alpha;
  return fmin2(fmin2(dtx, dty), tfinal - time);
}

// RK stage combination + checkpositivity + sweep flags + alpha of the new state.
// `rhs_only` writes RHS into b.R instead (parity tests of the spatial operator).
template <int M>
__global__ __launch_bounds__(256) void k_update(KCfg k, Bufs b, int stage, int slot_in, int slot_out,
                                                int parity, int rhs_only) {
  DevScalars *g = b.g;
  if (!g->active[parity]) return;
  const int i = 1 + blockIdx.x * 64 + threadIdx.x, j = 1 + blockIdx.y * 4 + threadIdx.y;
  const bool in = (i <= k.nx && j <= k.ny);
  const double alpha = alpha_from_bits(g->alpha_bits[slot_in]);
  double anew = 0.0;
  if (in) {
    const long long o = cidx(k, i, j);
    if (rhs_only) {
#pragma unroll
      for (int c = 0; c < 4; ++c) b.R[c * k.plane + o] = cell_rhs(k, b, c, o, alpha);
      return;
    }
    double dt;
    if (stage == 1) {
      dt = step_dt(k, alpha, g->time[parity], g->tfinal);
      if (i == 1 && j == 1) g->dt = dt;
    } else {
      dt = g->dt;
    }
    double v[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const double r = cell_rhs(k, b, c, o, alpha);
      const double u = b.U[c * k.plane + o];
      double un;
      if (stage == 1) {
        b.U0[c * k.plane + o] = u;
        un = u + dt * r;
      } else if (stage == 2) {
        const double x = u + dt * r;
        const double u0 = b.U0[c * k.plane + o];
        un = (M == MODE_EXACT) ? rk2_q(u0, x) : 0.75 * u0 + 0.25 * x;
      } else {
        const double x = u + dt * r;
        const double u0 = b.U0[c * k.plane + o];
        un = (M == MODE_EXACT) ? rk3_q(u0, x) : (u0 + 2.0 * x) / 3.0;
      }
      v[c] = un;
      b.U[c * k.plane + o] = un;
    }
    // checkpositivity (ex4_sedov_2d/main.f90:779-801)
    const double ke = (v[3] - ((v[2] * v[2]) / v[0]) / 2.) - ((v[1] * v[1]) / v[0]) / 2.;
    const bool bad = k.poscheck_le ? (v[0] <= k.eps || ke <= k.eps) : (v[0] < k.eps || ke < k.eps);
    if (bad) atomicOr(&g->postag, 1);
    if (k.use_limiter == 2) {
      const int p = posI(k, i, j);
      if (sweep_flag(k.gamma1, k.eps, v[0], v[1], v[2], v[3])) atomicOr(&b.bmP[p >> 5], 1u << (p & 31));
      if (v[0] < k.delta) {
        atomicOr(&b.bmR[p >> 5], 1u << (p & 31));
        atomicOr(&g->rho_any, 1);
      }
    }
    anew = cell_alpha(k, v[0], v[1], v[2], v[3]);
  }
  if (!rhs_only) alpha_atomic_max(&g->alpha_bits[slot_out], anew);
}

#ifdef SWF_COMMON_KERNELS
// Mode-independent grid kernels, compiled once (swf_k_exact.hip).

// alpha max over the interior, after the sweep changed the state
__global__ __launch_bounds__(256) void k_alpha_fix(KCfg k, Bufs b, int slot_out, int parity) {
  DevScalars *g = b.g;
  if (!g->active[parity] || !g->swept) return;
  const int i = 1 + blockIdx.x * 64 + threadIdx.x, j = 1 + blockIdx.y * 4 + threadIdx.y;
  double a = 0.0;
  if (i <= k.nx && j <= k.ny) {
    const long long o = cidx(k, i, j);
    a = cell_alpha(k, b.U[o], b.U[k.plane + o], b.U[2 * k.plane + o], b.U[3 * k.plane + o]);
  }
  alpha_atomic_max(&g->alpha_bits[slot_out], a);
}

// alpha max over every cell (ghosts included), for a freshly uploaded state
__global__ __launch_bounds__(256) void k_alpha_all(KCfg k, Bufs b, int slot_out) {
  const int i = -2 + blockIdx.x * 64 + threadIdx.x, j = -2 + blockIdx.y * 4 + threadIdx.y;
  double a = 0.0;
  if (i <= k.nx + 3 && j <= k.ny + 3) {
    const long long o = cidx(k, i, j);
    a = cell_alpha(k, b.U[o], b.U[k.plane + o], b.U[2 * k.plane + o], b.U[3 * k.plane + o]);
  }
  alpha_atomic_max(&b.g->alpha_bits[slot_out], a);
}

// Ghost cells.  Each ghost is a closed-form image of one interior cell (or of a
// prescribed inflow/ambient state), equal to the result of the reference's
// sequential passes (x-low, y-low, x-high, y-high) in that order:
//   Sedov  (ex4_sedov_2d/main.f90:916-994): reflect at x=0 (m1 -> -m1), y=0
//          (m2 -> -m2), zero-gradient at x=xm, y=ym; corners compose the images.
//   periodic (ex3_vortex_accuracy_2d/main.f90:653-669).
//   jet (ex5_mach2000_jet/main.f90:686-741): x=0 inflow for y<=0.05 else ambient,
//          reflect at y=0, zero-gradient elsewhere; full-domain variant: inflow for
//          |y|<0.05, zero-gradient at both y faces.
__global__ __launch_bounds__(256) void k_bc(KCfg k, Bufs b, int slot_out, int parity) {
  DevScalars *g = b.g;
  if (parity >= 0 && !g->active[parity]) return;
  const int W = k.nx + 6;
  const long long nrow = 3LL * W;                 // cells in 3 full ghost rows
  const long long nside = 3LL * k.ny;             // cells in 3 ghost columns of interior rows
  const long long total = 2 * nrow + 2 * nside;
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  double a = 0.0;
  if (t < total) {
    int i, j;
    if (t < nrow) {
      j = -2 + (int)(t / W);
      i = -2 + (int)(t % W);
    } else if (t < 2 * nrow) {
      j = k.ny + 1 + (int)((t - nrow) / W);
      i = -2 + (int)((t - nrow) % W);
    } else if (t < 2 * nrow + nside) {
      long long s = t - 2 * nrow;
      j = 1 + (int)(s / 3);
      i = -2 + (int)(s % 3);
    } else {
      long long s = t - 2 * nrow - nside;
      j = 1 + (int)(s / 3);
      i = k.nx + 1 + (int)(s % 3);
    }
    double v[4];
    if (k.bc == BC_PERIODIC) {
      const int si = (i <= 0) ? i + k.nx : (i > k.nx ? i - k.nx : i);
      const int sj = (j <= 0) ? j + k.ny : (j > k.ny ? j - k.ny : j);
      const long long so = cidx(k, si, sj);
#pragma unroll
      for (int c = 0; c < 4; ++c) v[c] = b.U[c * k.plane + so];
    } else {
      // source row: reflect (Sedov, half jet) or clamp (full jet) below, clamp above
      bool flip_y = false;
      int sj;
      if (j <= 0) {
        if (k.bc == BC_JET_FULL) {
          sj = 1;
        } else {
          sj = 1 - j;
          flip_y = true;
        }
      } else {
        sj = (j > k.ny) ? k.ny : j;
      }
      if (i <= 0 && (k.bc == BC_JET_HALF || k.bc == BC_JET_FULL)) {
        const double ys = k.ylo + sj * k.dy - k.dy / 2.;
        const bool jet = (k.bc == BC_JET_HALF) ? (ys <= k.jet_halfwidth) : (fabs(ys) < k.jet_halfwidth);
        if (jet) {
          v[0] = k.jet_rho; v[1] = k.jet_mx; v[2] = 0.; v[3] = k.jet_e;
        } else {
          v[0] = k.amb_rho; v[1] = 0.; v[2] = 0.; v[3] = k.amb_e;
        }
      } else {
        bool flip_x = false;
        int si;
        if (i <= 0) {
          si = 1 - i;
          flip_x = true;
        } else {
          si = (i > k.nx) ? k.nx : i;
        }
        const long long so = cidx(k, si, sj);
#pragma unroll
        for (int c = 0; c < 4; ++c) v[c] = b.U[c * k.plane + so];
        if (flip_x) v[1] = -v[1];
      }
      if (flip_y) v[2] = -v[2];
    }
    const long long o = cidx(k, i, j);
#pragma unroll
    for (int c = 0; c < 4; ++c) b.U[c * k.plane + o] = v[c];
    a = cell_alpha(k, v[0], v[1], v[2], v[3]);
  }
  if (slot_out >= 0) alpha_atomic_max(&g->alpha_bits[slot_out], a);
}

// Fortran-layout u(4,-2:nx+3,-2:ny+3) <-> SoA planes
__global__ void k_from_fortran(KCfg k, const double *__restrict__ src, double *__restrict__ dst) {
  const long long n = (long long)(k.nx + 6) * (k.ny + 6);
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < 4 * n;
       t += (long long)gridDim.x * blockDim.x) {
    const int c = (int)(t % 4);
    const long long cell = t / 4;
    const int i = -2 + (int)(cell % (k.nx + 6)), j = -2 + (int)(cell / (k.nx + 6));
    dst[c * k.plane + cidx(k, i, j)] = src[t];
  }
}
__global__ void k_to_fortran(KCfg k, const double *__restrict__ src, double *__restrict__ dst) {
  const long long n = (long long)(k.nx + 6) * (k.ny + 6);
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < 4 * n;
       t += (long long)gridDim.x * blockDim.x) {
    const int c = (int)(t % 4);
    const long long cell = t / 4;
    const int i = -2 + (int)(cell % (k.nx + 6)), j = -2 + (int)(cell / (k.nx + 6));
    dst[t] = src[c * k.plane + cidx(k, i, j)];
  }
}

#endif  // SWF_COMMON_KERNELS

// weno_right / weno_left of n 5-point stencils (parity tests of the reconstruction)
template <int M>
__global__ void k_weno_points(const double *__restrict__ a, double *__restrict__ right,
                              double *__restrict__ left, long long n) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  double s[5];
#pragma unroll
  for (int q = 0; q < 5; ++q) s[q] = a[5 * t + q];
  right[t] = weno_right<M>(&s[2]);
  left[t] = weno_left<M>(&s[2]);
}

}  // namespace swf
