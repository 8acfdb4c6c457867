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
#include "swf_sweep.h"

namespace swf {

// x tile: 64 interfaces x 4 rows; points k0-2 .. k0+66
constexpr int XT_W = 64, XT_H = 4, XT_PTS = XT_W + 5;
// y tile: 16 columns x 16 interfaces; points jf0-2 .. jf0+18
constexpr int YT_W = 16, YT_H = 16, YT_PTS = YT_H + 5;

SWF_D double alpha_from_bits(unsigned long long b) { return __longlong_as_double((long long)b); }

// LDS of one flux tile (x and y tiles share the allocation)
union FluxSmem {
  struct {
    double fp[XT_H][XT_PTS][4];
    double fn[XT_H][XT_PTS][4];
    Prim pr[XT_H][XT_PTS];
  } x;
  struct {
    double fp[YT_PTS][YT_W][4];
    double fn[YT_PTS][YT_W][4];
    Prim pr[YT_PTS][YT_W];
  } y;
};

SWF_D int flux_xblocks(const KCfg &k) { return ((k.nx + 1 + XT_W - 1) / XT_W) * ((k.ny + XT_H - 1) / XT_H); }
SWF_D int flux_yblocks(const KCfg &k) { return ((k.nx + YT_W - 1) / YT_W) * ((k.ny + 1 + YT_H - 1) / YT_H); }

// x tile: interfaces k0..k0+63 of rows j0..j0+3
template <int M>
SWF_D void xflux_tile(const KCfg &k, const Bufs &b, int bx, int by, double alpha, double ia, FluxSmem &sm,
                      int tid) {
  const int k0 = bx * XT_W, j0 = 1 + by * XT_H;
  for (int p = tid; p < XT_H * XT_PTS; p += 256) {
    const int r = p / XT_PTS, q = p - r * XT_PTS;
    const int i = k0 - 2 + q, j = j0 + r;
    if (j <= k.ny && i <= k.nx + 3) {
      const long long o = cidx(k, i, j);
      double u[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) u[c] = b.U[c * k.plane + o];
      split_point<M, 0>(k, u, alpha, ia, sm.x.fp[r][q], sm.x.fn[r][q]);
      Prim pr;
      roe_prims<M, 0>(k, u, pr.w, pr.vx, pr.vy, pr.h);
      sm.x.pr[r][q] = pr;
    }
  }
  __syncthreads();
  const int tx = tid & (XT_W - 1), ty = tid / XT_W;
  const int kk = k0 + tx, j = j0 + ty;
  if (kk > k.nx || j > k.ny) return;
  double fp[4], fn[4];
  // f+ at interface kk (Roe(kk, kk+1), stencil kk-2..kk+2)
  half_flux<M, 0, true>(k, sm.x.pr[ty][tx + 2], sm.x.pr[ty][tx + 3], &sm.x.fp[ty][tx + 2][0], 4, fp);
  // f- at interface kk+1 (Roe(kk+1, kk+2), stencil kk-1..kk+3)
  half_flux<M, 0, false>(k, sm.x.pr[ty][tx + 3], sm.x.pr[ty][tx + 4], &sm.x.fn[ty][tx + 3][0], 4, fn);
  const long long o = cidx(k, kk, j);
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const double f = k.alpha_in_scheme ? (fp[c] - fn[c]) : (alpha * fp[c] - alpha * fn[c]);
    b.Fx[c * k.plane + o] = f;
  }
}

// y tile: interfaces jf0..jf0+15 of columns i0..i0+15
template <int M>
SWF_D void yflux_tile(const KCfg &k, const Bufs &b, int bx, int by, double alpha, double ia, FluxSmem &sm,
                      int tid) {
  const int i0 = 1 + bx * YT_W, jf0 = by * YT_H;
  for (int p = tid; p < YT_PTS * YT_W; p += 256) {
    const int r = p / YT_W, q = p - r * YT_W;
    const int i = i0 + q, j = jf0 - 2 + r;
    if (i <= k.nx && j <= k.ny + 3) {
      const long long o = cidx(k, i, j);
      double u[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) u[c] = b.U[c * k.plane + o];
      split_point<M, 1>(k, u, alpha, ia, sm.y.fp[r][q], sm.y.fn[r][q]);
      Prim pr;
      roe_prims<M, 1>(k, u, pr.w, pr.vx, pr.vy, pr.h);
      sm.y.pr[r][q] = pr;
    }
  }
  __syncthreads();
  const int tx = tid & (YT_W - 1), ty = tid / YT_W;
  const int i = i0 + tx, jf = jf0 + ty;
  if (i > k.nx || jf > k.ny) return;
  double fp[4], fn[4];
  half_flux<M, 1, true>(k, sm.y.pr[ty + 2][tx], sm.y.pr[ty + 3][tx], &sm.y.fp[ty + 2][tx][0], YT_W * 4, fp);
  half_flux<M, 1, false>(k, sm.y.pr[ty + 3][tx], sm.y.pr[ty + 4][tx], &sm.y.fn[ty + 3][tx][0], YT_W * 4, fn);
  const long long o = cidx(k, i, jf);
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const double f = k.alpha_in_scheme ? (fp[c] - fn[c]) : (alpha * fp[c] - alpha * fn[c]);
    b.Fy[c * k.plane + o] = f;
  }
}

// wenofd: both directions in one launch; the first flux_xblocks() workgroups
// take x tiles, the rest y tiles.  `clear_maps` also prepares this stage's
// bookkeeping (sweep bitmaps, flags, the alpha slot the update kernel fills).
template <int M>
__global__ __launch_bounds__(256) void k_flux(KCfg k, Bufs b, int slot_in, int slot_reset, int parity,
                                              int clear_maps) {
  __shared__ FluxSmem sm;
  DevScalars *g = b.g;
  if (!g->active[parity]) return;
  const int tid = threadIdx.x;
  if (clear_maps) {
    const long long gid = (long long)blockIdx.x * 256 + tid;
    if (gid < k.nwords) {
      b.bmR[gid] = 0u;
      b.bmP[gid] = 0u;
    }
    if (gid == 0) {
      g->postag = 0;
      g->rho_any = 0;
    }
  }
  const double alpha = alpha_from_bits(g->alpha_bits[slot_in]);
  const double ia = 1.0 / alpha;
  const int nbx = flux_xblocks(k);
  if ((int)blockIdx.x < nbx) {
    const int gx = (k.nx + 1 + XT_W - 1) / XT_W;
    xflux_tile<M>(k, b, blockIdx.x % gx, blockIdx.x / gx, alpha, ia, sm, tid);
  } else {
    const int gx = (k.nx + YT_W - 1) / YT_W;
    const int t = blockIdx.x - nbx;
    yflux_tile<M>(k, b, t % gx, t / gx, alpha, ia, sm, tid);
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

// Inflow / ambient state of the jet's x = 0 ghost at row j (ex5_mach2000_jet/main.f90:688-708)
SWF_D void jet_state(const KCfg &k, int j, double v[4]) {
  const double ys = k.ylo + j * k.dy - k.dy / 2.;
  const bool jet = (k.bc == BC_JET_HALF) ? (ys <= k.jet_halfwidth) : (fabs(ys) < k.jet_halfwidth);
  if (jet) {
    v[0] = k.jet_rho; v[1] = k.jet_mx; v[2] = 0.; v[3] = k.jet_e;
  } else {
    v[0] = k.amb_rho; v[1] = 0.; v[2] = 0.; v[3] = k.amb_e;
  }
}

// Write the edge ghost cells whose value is an image of interior cell (i, j) with
// state v.  These are the ghosts the flux stencils read (edges, not corners);
// the images equal the reference's `bc` passes (see k_bc).
SWF_D void write_ghosts(const KCfg &k, const Bufs &b, int i, int j, const double v[4]) {
  double w[4];
  switch (k.bc) {
    case BC_SEDOV:
      if (i <= 3) {
        w[0] = v[0]; w[1] = -v[1]; w[2] = v[2]; w[3] = v[3];
        store_ij(k, b, 1 - i, j, w);
      }
      if (i == k.nx)
        for (int g = 1; g <= 3; ++g) store_ij(k, b, k.nx + g, j, v);
      if (j <= 3) {
        w[0] = v[0]; w[1] = v[1]; w[2] = -v[2]; w[3] = v[3];
        store_ij(k, b, i, 1 - j, w);
      }
      if (j == k.ny)
        for (int g = 1; g <= 3; ++g) store_ij(k, b, i, k.ny + g, v);
      break;
    case BC_PERIODIC:
      if (i >= k.nx - 2) store_ij(k, b, i - k.nx, j, v);
      if (i <= 3) store_ij(k, b, i + k.nx, j, v);
      if (j >= k.ny - 2) store_ij(k, b, i, j - k.ny, v);
      if (j <= 3) store_ij(k, b, i, j + k.ny, v);
      break;
    default:  // jets
      if (i == 1) {
        jet_state(k, j, w);
        for (int g = 0; g <= 2; ++g) store_ij(k, b, -g, j, w);
      }
      if (i == k.nx)
        for (int g = 1; g <= 3; ++g) store_ij(k, b, k.nx + g, j, v);
      if (k.bc == BC_JET_HALF) {
        if (j <= 3) {
          w[0] = v[0]; w[1] = v[1]; w[2] = -v[2]; w[3] = v[3];
          store_ij(k, b, i, 1 - j, w);
        }
      } else if (j == 1) {
        for (int g = 0; g <= 2; ++g) store_ij(k, b, i, -g, v);
      }
      if (j == k.ny)
        for (int g = 1; g <= 3; ++g) store_ij(k, b, i, k.ny + g, v);
      break;
  }
}

// block-wide max of a non-negative value over NT threads; result valid in thread 0
template <int NT>
SWF_D double block_max(double a, double *red, int tid) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const double o = __shfl_xor(a, off, 64);
    a = (o > a) ? o : a;
  }
  if ((tid & 63) == 0) red[tid >> 6] = a;
  __syncthreads();
  double m = 0.0;
  if (tid == 0)
    for (int q = 0; q < NT / 64; ++q) m = fmax2(m, red[q]);
  __syncthreads();
  return m;
}

// After a sweep: alpha of the update blocks whose cells changed, and the edge
// ghosts of the boundary band.  One workgroup of NT threads (NT >= 256).
template <int NT>
SWF_D void after_sweep(const KCfg &k, const Bufs &b, SweepShared<NT> &sh, double *red, int tid) {
  const int gx = (k.nx + 63) >> 6, gy = (k.ny + 3) >> 2, nblk = gx * gy;
  const int nw = sh.all_dirty ? (nblk + 31) / 32 : min(SW_DIRTY_WORDS, (nblk + 31) / 32);
  constexpr int PER = NT / 256;  // update blocks handled at once
  for (int w = 0; w < nw; ++w) {
    unsigned word = sh.all_dirty ? 0xffffffffu : sh.dirty[w];
    while (word) {
      // up to PER dirty blocks of this word, one 256-thread group each
      int blks[PER];
#pragma unroll
      for (int q = 0; q < PER; ++q) {
        if (word) {
          blks[q] = w * 32 + (__ffs(word) - 1);
          word &= word - 1;
        } else {
          blks[q] = -1;
        }
      }
      const int grp = tid >> 8, lt = tid & 255;
      const int blk = blks[grp];
      double a = 0.0;
      if (blk >= 0 && blk < nblk) {
        const int i = 1 + (blk % gx) * 64 + (lt & 63), j = 1 + (blk / gx) * 4 + (lt >> 6);
        if (i <= k.nx && j <= k.ny) {
          double v[4];
          load_ij(k, b, i, j, v);
          a = cell_alpha(k, v[0], v[1], v[2], v[3]);
        }
      }
      // max within each 256-thread group: waves 4*grp .. 4*grp+3
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) {
        const double o = __shfl_xor(a, off, 64);
        a = (o > a) ? o : a;
      }
      if ((tid & 63) == 0) red[tid >> 6] = a;
      __syncthreads();
      if (lt == 0 && blk >= 0 && blk < nblk)
        b.apart[blk] = fmax2(fmax2(red[4 * grp], red[4 * grp + 1]), fmax2(red[4 * grp + 2], red[4 * grp + 3]));
      __syncthreads();
    }
  }
  // ghosts of the boundary band
  const int nb = 6 * k.ny + 6 * k.nx;
  for (int t = tid; t < nb; t += NT) {
    int i, j;
    if (t < 6 * k.ny) {
      const int c = t % 6;
      j = 1 + t / 6;
      i = (c < 3) ? 1 + c : k.nx - 5 + c;
    } else {
      const int s = t - 6 * k.ny, c = s % 6;
      i = 1 + s / 6;
      j = (c < 3) ? 1 + c : k.ny - 5 + c;
    }
    if (i >= 1 && i <= k.nx && j >= 1 && j <= k.ny) {
      double v[4];
      load_ij(k, b, i, j, v);
      write_ghosts(k, b, i, j, v);
    }
  }
  __syncthreads();
}

// RK stage combination + checkpositivity + sweep flags + ghost images + the
// workgroup's alpha max of the new state (reduced by k_stage_end).  `rhs_only`
// writes RHS into b.R instead (parity tests of the spatial operator).
template <int M>
__global__ __launch_bounds__(256) void k_update(KCfg k, Bufs b, int stage, int slot_in, int slot_out,
                                                int parity, int rhs_only) {
  __shared__ double red4[4];
  DevScalars *g = b.g;
  const int tid = threadIdx.x + 64 * threadIdx.y;
  const int bid = blockIdx.x + blockIdx.y * gridDim.x;
  if (!g->active[parity]) return;
  const int i = 1 + blockIdx.x * 64 + threadIdx.x, j = 1 + blockIdx.y * 4 + threadIdx.y;
  const bool in = (i <= k.nx && j <= k.ny);
  const double alpha = alpha_from_bits(g->alpha_bits[slot_in]);
  if (rhs_only) {
    if (in) {
      const long long o = cidx(k, i, j);
#pragma unroll
      for (int c = 0; c < 4; ++c) b.R[c * k.plane + o] = cell_rhs(k, b, c, o, alpha);
    }
    return;
  }
  double anew = 0.0;
  if (in) {
    const long long o = cidx(k, i, j);
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
    write_ghosts(k, b, i, j, v);
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
  const double am = block_max<256>(anew, red4, tid);
  if (tid == 0) b.apart[bid] = am;
}

#ifdef SWF_COMMON_KERNELS
// Mode-independent grid kernels, compiled once (swf_k_exact.hip).

// End of an RK stage, one workgroup: sweep_lim when checkpositivity flagged the
// stage, the alpha max of the new state (findalpha) from the per-workgroup
// maxima and the prescribed boundary states, and at stage 3 the time update of
// the reference's loop (time = time + dt, DO WHILE test).
__global__ __launch_bounds__(SW_T) void k_stage_end(KCfg k, Bufs b, int stage, int slot_out, int parity,
                                                    int serial_ubar) {
  __shared__ SweepShared<SW_T> sh;
  __shared__ double red[SW_T / 64];
  DevScalars *g = b.g;
  const int tid = threadIdx.x;
  if (!g->active[parity]) {
    // an inactive step leaves the next one inactive too
    if (stage == 3 && tid == 0) g->active[1 - parity] = 0;
    return;
  }
  if (k.use_limiter == 2 && g->postag) {
    sweep_body<SW_T>(k, b, serial_ubar != 0, sh, tid);
    after_sweep<SW_T>(k, b, sh, red, tid);
  }
  const int nblk = ((k.nx + 63) >> 6) * ((k.ny + 3) >> 2);
  double a = 0.0;
  for (int q = tid; q < nblk; q += SW_T) a = fmax2(a, b.apart[q]);
  a = block_max<SW_T>(a, red, tid);
  if (tid == 0) {
    g->alpha_bits[slot_out] = (unsigned long long)__double_as_longlong(fmax2(a, k.alpha_const));
    if (stage == 3) {
      const double tn = g->time[parity] + g->dt;
      g->time[1 - parity] = tn;
      g->steps_done++;
      g->active[1 - parity] = (g->status == ST_OK) && time_continue_q(tn, g->tfinal);
    }
  }
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
