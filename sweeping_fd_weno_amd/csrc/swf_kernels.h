// swf_kernels.h - grid kernels of one SSP-RK3 stage, templated on precision mode.
//
// One stage (ex4_sedov_2d/main.f90:685-777 `rk`) is three launches:
//   k_xflux<M,LIM,STEP> wenofd x-pass: net x-interface flux F(i+1/2) = fp_flux(i) - fn_flux(i+1)
//                (scaled by alpha as :414-415) into Fx; also clears the stage's
//                sweep bitmaps and flags
//   k_ystage<M,LIM,STEP> wenofd y-pass for a tile, then per cell `scheme` (RHS from Fx and
//                the tile's y fluxes), the RK combination, checkpositivity, sweep
//                flags, ghost images (`bc`) and the tile's alpha max (findalpha)
//   k_stage_end  (mode-independent, one workgroup) sweep_lim when flagged, alpha
//                reduction, and at stage 3 the time update
// With the Zhang-Shu flux limiter (use_limiter = 1; LIM = true) the two flux
// kernels limit every half flux and count the changes per workgroup, and two
// launches come before k_stage_end: k_zs_edges<M> (the halves of the ghost rows
// and columns, counted only) and k_bc (the new state's full ghost frame).
// Each flux tile stages its stencil points (split fluxes and Roe primitives) in
// LDS once.  An x-flux lane computes one centre i: the f+ and f- halves of its
// stencil with the one Roe system of (i, i+1), eight WENO reconstructions, and
// the interface flux from the next lane's f-; a stage-kernel thread computes one
// y interface: the f+ half with the Roe system of (j, j+1) and the f- half with
// that of (j+1, j+2).
#pragma once

#include "swf_device.h"
#include "swf_sweep.h"

namespace swf {

// x tile: 64 centres (k0 .. k0+63) of each of 4 rows; lanes 0..62 give the
// interfaces k0+1/2 .. k0+62+1/2, so tile origins step by XT_C = 63.  Stencil points
// k0-2 .. k0+65 of each row.
constexpr int XT_H = 4, XT_PTS = 64 + 4;
// y tile: 16 columns x 16 interfaces (jf0 .. jf0+15) -> 15 cells (jf0+1 .. jf0+15);
// stencil points jf0-2 .. jf0+18
constexpr int YT_W = UB_W, YT_H = UB_H + 1, YT_PTS = YT_H + 5;

SWF_D double alpha_from_bits(unsigned long long b) { return __longlong_as_double((long long)b); }


// LDS of a flux tile, component-major so that consecutive lanes read
// consecutive doubles: split fluxes f+ (fp) and f- (fn) of every stencil point,
// then the Roe-average inputs w, vx, vy, h.
struct XSmem {
  static constexpr int CS = XT_H * XT_PTS;  // component stride
  double fp[4][XT_H][XT_PTS];
  double fn[4][XT_H][XT_PTS];
  double pr[4][XT_H][XT_PTS];
  int zred[8];
};
struct YSmem {
  static constexpr int CS = YT_PTS * YT_W;
  double fp[4][YT_PTS][YT_W];
  double fn[4][YT_PTS][YT_W];
  double pr[4][YT_PTS][YT_W];
  // net y-interface fluxes of each wave's last row (the next wave's first cells
  // read them; within a wave the row below comes by a shuffle)
  double G[YS_T / 64][4][YT_W];
  double red[YS_T / 64];
  double ured[4][YS_T / 64];  // weight*u sums of the waves
  int zred[2 * YS_T / 64];
};

// split fluxes and Roe primitives of stencil point u
template <int M, int DIR>
SWF_D void stage_point(const KCfg &k, const double u[4], double alpha, double ia, double fp[4], double fn[4],
                       double pr[4]) {
  Prim p;
  if (M == MODE_EXACT) {
    split_point<M, DIR>(k, u, alpha, ia, fp, fn);
    roe_prims<M, DIR>(k, u, p.w, p.vx, p.vy, p.h);
  } else {
    stage_point_fast<DIR>(k, u, ia, fp, fn, p);
  }
  pr[0] = p.w;
  pr[1] = p.vx;
  pr[2] = p.vy;
  pr[3] = p.h;
}

// Stage the points p = tid and p = tid + NT of a tile of NT threads (at most two
// per thread):
// both loads are issued before either point is processed, so the tile waits for
// one memory round trip.  point(p, i, j) maps a point to its cell and returns
// false outside the grid; store(p, fp, fn, pr) writes the staged values.
// STEP (the step program, alpha_lines): the split of a point uses the alpha of its
// line, row j for the x-pass (aline[j+2]) and column i for the y-pass (aline[i+2]).
template <int M, int DIR, int NT, int NPTS, bool STEP, class PointF, class StoreF>
SWF_D void stage_tile(const KCfg &k, const Bufs &b, double alpha, double ia, const double *aline, int tid,
                      PointF point, StoreF store) {
  static_assert(NPTS <= 2 * NT, "at most two points per thread");
  double u[2][4];
  bool ok[2];
  int pp[2], line[2];
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    pp[s] = tid + NT * s;
    int i = 0, j = 0;
    ok[s] = pp[s] < NPTS && point(pp[s], i, j);
    line[s] = DIR == 0 ? j : i;
    if (ok[s]) {
      const long long o = cidx(k, i, j);
#pragma unroll
      for (int c = 0; c < 4; ++c) u[s][c] = b.Uin[c * k.plane + o];
    }
  }
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    if (ok[s]) {
      double fp[4], fn[4], pr[4];
      double al = alpha, ial = ia;
      if (STEP) {
        al = aline[line[s] + 2];
        ial = 1.0 / al;
      }
      stage_point<M, DIR>(k, u[s], al, ial, fp, fn, pr);
      store(pp[s], fp, fn, pr);
    }
  }
}

// Roe-average inputs of a staged point (component stride CS)
template <int CS>
SWF_D Prim load_prim(const double *pr) {
  Prim p;
  p.w = pr[0];
  p.vx = pr[CS];
  p.vy = pr[2 * CS];
  p.h = pr[3 * CS];
  return p;
}

// net flux at one interface from the staged points: f+ at the interface (centre
// point p, Roe(p, p+1)) minus f- of the next point (centre p+1, Roe(p+1, p+2)),
// scaled by alpha inside wenofd (ex4) or later in `scheme` (ex3/ex5).  Pointers
// are at component 0 of point p; points SS doubles apart, components CS.
template <int M, int DIR, int SS, int CS, bool LIM, bool E12 = false>
SWF_D int iface_flux(const KCfg &k, const double *pr, const double *fp, const double *fn, double alpha,
                     double F[4]) {
  const Prim p0 = load_prim<CS>(pr), p1 = load_prim<CS>(pr + SS), p2 = load_prim<CS>(pr + 2 * SS);
  double a[4], c[4];
  half_flux<M, DIR, true, SS, CS, E12>(k, p0, p1, fp, a);
  half_flux<M, DIR, false, SS, CS, E12>(k, p1, p2, fn + SS, c);
  int lim = 0;
  if (LIM) lim = zs_limit<CS>(k.gamma1, fp, a) | (zs_limit<CS>(k.gamma1, fn + SS, c) << 2);
#pragma unroll
  for (int q = 0; q < 4; ++q) F[q] = k.alpha_in_scheme ? (a[q] - c[q]) : (alpha * a[q] - alpha * c[q]);
  return lim;
}

// den_count / pre_count increments of the limiter change bits of iface_flux
SWF_D int zs_den(int lim) { return (lim & 1) + ((lim >> 2) & 1); }
SWF_D int zs_pre(int lim) { return ((lim >> 1) & 1) + ((lim >> 3) & 1); }

// sums of two counts over a block of NT threads (4 or 8 waves), stored by thread 0
// as one (den, pre) pair of the limiter count partials
template <int NT>
SWF_D void zs_store(int den, int pre, int *pair, int *red, int tid) {
  static_assert(NT == 256 || NT == 512, "4 or 8 waves");
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    den += __shfl_xor(den, off, 64);
    pre += __shfl_xor(pre, off, 64);
  }
  if ((tid & 63) == 0) {
    red[2 * (tid >> 6)] = den;
    red[2 * (tid >> 6) + 1] = pre;
  }
  __syncthreads();
  if (tid == 0) {
    int d = 0, p = 0;
#pragma unroll
    for (int w = 0; w < NT / 64; ++w) {
      d += red[2 * w];
      p += red[2 * w + 1];
    }
    pair[0] = d;
    pair[1] = p;
  }
}

// Occupancy of the fast-mode flux kernels: 4 waves per SIMD (at most 128 VGPRs);
// the exact mode's double-double and binary128 sites need more registers.
// x-flux workgroups have 256 threads, stage workgroups YS_T.
#define SWF_FLUX_BOUNDS(M) __launch_bounds__(256, (M) == MODE_EXACT ? 1 : 4)
#define SWF_STAGE_BOUNDS(M) __launch_bounds__(YS_T, (M) == MODE_EXACT ? 1 : 4)

// wenofd x-pass -> Fx for the centres k0 = XT_C bx .. k0 + 63 of rows j0 = 1 + 4 by
// .. j0 + 3.  Each lane forms both halves of its centre k with the one Roe system
// of (k, k+1) (centre_fluxes), and lanes 0..XT_C-1 the net interface flux
// F(k+1/2) = fp_flux(k) - fn_flux(k+1), fn_flux(k+1) from the next lane; the
// tile's last centre is the next tile's first.  In an RK stage (stage = 1, 2, 3; 0
// for a stand-alone evaluation) the launch also resets the stage's sweep bitmaps
// and flags (the stage kernel sets them), and at stage 1 of an inactive step keeps
// the next step inactive.  Tile (bx, by) of a gx x gy grid of x-flux tiles.
// STEP: the step diffraction program's wenofd (ex6_shock_diffraction/main.f90:
// 170-647): each row (x-pass) / column (y-pass) split and scaled with its own
// alpha max, and 1d-12 in the WENO weights.
template <int M, bool LIM, bool STEP = false>
SWF_D void xflux_tile(const KCfg &k, const Bufs &b, int stage, int slot_in, int parity, int bx, int by, int gx,
                      int gy, XSmem &sm) {
  constexpr int CS = XSmem::CS;
  DevScalars *g = b.g;
  const int tid = threadIdx.x;
  const bool first = bx == 0 && by == 0 && tid == 0;
  if (!g->active[parity] || g->status != ST_OK) {
    if (stage == 1 && first) g->active[1 - parity] = 0;
    return;
  }
  if (stage > 0) {
    const long long nthreads = (long long)gx * gy * 256;
    for (long long w = (long long)(bx + by * gx) * 256 + tid; w < k.nwords; w += nthreads) {
      b.bmR[w] = 0u;
      b.bmP[w] = 0u;
    }
    if (first) {
      g->postag = 0;
      g->rho_any = 0;
    }
  }
  const double alpha = alpha_from_bits(g->alpha_bits[slot_in]);
  const double ia = 1.0 / alpha;
  const int k0 = bx * XT_C, j0 = 1 + by * XT_H;
  const double *aline = STEP ? b.arow : nullptr;
  stage_tile<M, 0, 256, XT_H * XT_PTS, STEP>(
      k, b, alpha, ia, aline, tid,
      [&](int p, int &i, int &j) {
        const int r = p / XT_PTS;
        i = k0 - 2 + (p - r * XT_PTS);
        j = j0 + r;
        return j <= k.ny && i <= k.nx + 3;
      },
      [&](int p, const double *fp, const double *fn, const double *pr) {
        const int r = p / XT_PTS, q = p - r * XT_PTS;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          sm.fp[c][r][q] = fp[c];
          sm.fn[c][r][q] = fn[c];
          sm.pr[c][r][q] = pr[c];
        }
      });
  __syncthreads();
  const int tx = tid & 63, ty = tid >> 6;
  const int kc = k0 + tx, j = j0 + ty;
  const bool own = kc <= k.nx + 1 && j <= k.ny;
  double a[4] = {0.0, 0.0, 0.0, 0.0}, c[4] = {0.0, 0.0, 0.0, 0.0};
  int lim = 0;
  if (own) {
    centre_fluxes<M, 0, 1, CS, STEP>(k, load_prim<CS>(&sm.pr[0][ty][tx + 2]), load_prim<CS>(&sm.pr[0][ty][tx + 3]),
                                     &sm.fp[0][ty][tx + 2], &sm.fn[0][ty][tx + 2], a, c);
    if (LIM) {
      // the halves fp(0..nx) and fn(1..nx+1) of the row, each counted once (the
      // tile's last centre by the next tile, unless it ends the row)
      if (kc <= k.nx) lim |= zs_limit<CS>(k.gamma1, &sm.fp[0][ty][tx + 2], a);
      if (kc >= 1) lim |= zs_limit<CS>(k.gamma1, &sm.fn[0][ty][tx + 2], c) << 2;
      if (tx == XT_C && kc != k.nx + 1) lim = 0;
    }
  }
  double cn[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) cn[q] = __shfl_down(c[q], 1, 64);
  if (own && tx < XT_C && kc <= k.nx) {
    const double al = STEP ? aline[j + 2] : alpha;
    const long long o = cidx(k, kc, j);
#pragma unroll
    for (int q = 0; q < 4; ++q) b.Fx[q * k.plane + o] = k.alpha_in_scheme ? (a[q] - cn[q]) : (al * a[q] - al * cn[q]);
  }
  if (LIM && stage > 0) zs_store<256>(zs_den(lim), zs_pre(lim), b.zs + 2 * (bx + by * gx), sm.zred, tid);
}

// Workgroups are dealt to the eight XCDs in turn by linear id.  In the fast modes
// each XCD gets a contiguous run of tiles instead, so that neighbouring tiles,
// which share stencil halos, share an L2 (Sedov 640^2 FP64: 0.137 -> 0.131 ms per
// step).  The exact mode keeps the dealt order: its tiles differ in cost (the
// binary128 fallback gathers in near-constant regions), and contiguous runs left
// the XCDs unbalanced (0.89 -> 1.04 ms per step).  Returns the tile of workgroup
// (blockIdx.x, blockIdx.y).
template <int M>
SWF_D void xcd_tile(int gx, int gy, int &bx, int &by) {
  if (M == MODE_EXACT) {
    bx = blockIdx.x;
    by = blockIdx.y;
    return;
  }
  const int n = gx * gy, b = blockIdx.x + blockIdx.y * gx;
  const int q = n >> 3, r = n & 7, x = b & 7, kk = b >> 3;
  const int t = x * q + min(x, r) + kk;
  bx = t % gx;
  by = t / gx;
}

template <int M, bool LIM, bool STEP = false>
__global__ SWF_FLUX_BOUNDS(M) void k_xflux(KCfg k, Bufs b, int stage, int slot_in, int parity) {
  __shared__ XSmem sm;
  int bx, by;
  xcd_tile<M>(gridDim.x, gridDim.y, bx, by);
  xflux_tile<M, LIM, STEP>(k, b, stage, slot_in, parity, bx, by, gridDim.x, gridDim.y, sm);
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
    case BC_STEP:  // the whole frame and column 1 after the stage (step_frame)
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

// sum of n = 2^m values as a balanced tree: ((r0 + r1) + (r2 + r3)) + ...
template <int N>
SWF_D double pair_sum(const double *r) {
  if constexpr (N == 1)
    return r[0];
  else
    return pair_sum<N / 2>(r) + pair_sum<N / 2>(r + N / 2);
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

// After a sweep: alpha partials of the update tiles whose cells changed, and
// the edge ghosts of their cells in the boundary band.  One workgroup of NT
// threads (NT >= 512).  The changed tiles go NT / 32 at a time, 32 threads per
// tile, each thread's cells loaded together; the tile's max by shuffles within
// its 32 lanes.
template <int NT>
SWF_D void after_sweep(const KCfg &k, const Bufs &b, SweepShared<NT> &sh, double *red, int tid) {
  static_assert(NT % 64 == 0 && NT >= 512, "whole waves, 16 tiles per round at least");
  constexpr int RT = NT / 32;                      // tiles per round
  constexpr int CELLS = UB_W * UB_H;
  constexpr int PER = (CELLS + 31) / 32;           // cells per thread
  const int gx = ub_gx(k), nblk = gx * ub_gy(k);
  const int nw = sh.all_dirty ? (nblk + 31) / 32 : min(SW_DIRTY_WORDS, (nblk + 31) / 32);
  const int grp = tid >> 5, lt = tid & 31;
  // the changed tiles as a list in sh.list (free after the sweep): each thread counts
  // the set bits of its bitmap words, one block scan places them
  int cnt = 0;
  for (int w = tid; w < nw; w += NT) {
    unsigned word = sh.all_dirty ? 0xffffffffu : sh.dirty[w];
    if (w * 32 + 32 > nblk) word &= (nblk - w * 32 >= 32) ? 0xffffffffu : ((1u << (nblk - w * 32)) - 1u);
    cnt += __popc(word);
  }
  const int first = block_exclusive_scan<NT>(cnt, sh, &sh.nscan, tid);
  const int ntiles = sh.nscan;
  if (ntiles <= SW_WCH * 32 && cnt) {
    int o = first;
    for (int w = tid; w < nw; w += NT) {
      unsigned word = sh.all_dirty ? 0xffffffffu : sh.dirty[w];
      if (w * 32 + 32 > nblk) word &= (nblk - w * 32 >= 32) ? 0xffffffffu : ((1u << (nblk - w * 32)) - 1u);
      while (word) {
        sh.list[o++] = w * 32 + (__ffs(word) - 1);
        word &= word - 1;
      }
    }
  }
  __syncthreads();
  // beyond the list's capacity (whole-grid sweeps of very large grids): the bitmap in
  // order, every thread walking it alike
  int w = 0;
  unsigned word = (nw > 0) ? (sh.all_dirty ? 0xffffffffu : sh.dirty[0]) : 0u;
  for (int r0 = 0;; r0 += RT) {
    int myblk = -1;
    bool any = false;
    if (ntiles <= SW_WCH * 32) {
      any = r0 < ntiles;
      if (r0 + grp < ntiles) myblk = sh.list[r0 + grp];
    } else {
#pragma unroll
      for (int q = 0; q < RT; ++q) {
        while (!word && w + 1 < nw) {
          ++w;
          word = sh.all_dirty ? 0xffffffffu : sh.dirty[w];
        }
        if (word) {
          const int blk = w * 32 + (__ffs(word) - 1);
          word &= word - 1;
          if (blk < nblk) {
            any = true;
            if (q == grp) myblk = blk;
          }
        }
      }
    }
    if (!any) break;
    double a = 0.0;
    if (myblk >= 0) {
      const int i0 = 1 + (myblk % gx) * UB_W, j0 = 1 + (myblk / gx) * UB_H;
      double v[PER][4];
      bool ok[PER];
#pragma unroll
      for (int r = 0; r < PER; ++r) {
        const int c = lt + 32 * r, i = i0 + c % UB_W, j = j0 + c / UB_W;
        ok[r] = c < CELLS && i <= k.nx && j <= k.ny;
        if (ok[r]) load_ij(k, b, i, j, v[r]);
      }
#pragma unroll
      for (int r = 0; r < PER; ++r)
        if (ok[r])  // the stage kernel's function for the mode (fast_div: every mode but the exact one)
          a = fmax2(a, k.fast_div ? cell_alpha_fast(k, v[r][0], v[r][1], v[r][2], v[r][3])
                                  : cell_alpha(k, v[r][0], v[r][1], v[r][2], v[r][3]));
      // the ghost images of the tile's cells in the boundary band (3 cells deep)
#pragma unroll
      for (int r = 0; r < PER; ++r) {
        const int c = lt + 32 * r, i = i0 + c % UB_W, j = j0 + c / UB_W;
        if (ok[r] && (i <= 3 || i >= k.nx - 2 || j <= 3 || j >= k.ny - 2)) write_ghosts(k, b, i, j, v[r]);
      }
    }
#pragma unroll
    for (int off = 16; off >= 1; off >>= 1) {
      const double o = __shfl_xor(a, off, 64);
      a = (o > a) ? o : a;
    }
    if (lt == 0 && myblk >= 0) b.apart[myblk] = a;
  }
  __syncthreads();
}

// One stage for a tile of UB_W x UB_H cells: the y-pass of wenofd over the
// tile's 16 interfaces, then per cell scheme (ex4_sedov_2d/main.f90:639-650),
// the RK combination (:715, :738, :760), checkpositivity (:779-801), the sweep
// flags, ghost images and the tile's alpha max (findalpha, :652-683) as one
// partial of the stage's alpha.  Stage 1 also fixes dt (:696-700) and with it the
// time after the step and the gate of the next step (the loop test of :112).
// flags: YS_RHS_ONLY stores RHS into b.R instead (parity tests of the spatial
// operator); YS_NO_SCALARS leaves the device scalars alone (kernel timing).
// Stage tile (bx, by) of a grid gx tiles wide.
template <int M, bool LIM, bool STEP = false>
SWF_D void ystage_tile(const KCfg &k, const Bufs &b, int stage, int slot_in, int parity, int flags, int bx, int by,
                       int gx, YSmem &sm) {
  DevScalars *g = b.g;
  if (!g->active[parity] || g->status != ST_OK) return;
  const bool rhs_only = flags & YS_RHS_ONLY, scalars = !(flags & YS_NO_SCALARS);
  // after a stage that swept, the tile's sums of weight*u for the next sweep's u_bar
  const bool usum = M != MODE_EXACT && scalars && !rhs_only && g->ubar_tiles;
  const int tid = threadIdx.x;
  const double alpha = alpha_from_bits(g->alpha_bits[slot_in]);
  const double ia = 1.0 / alpha;
  const int i0 = 1 + bx * UB_W, j0 = 1 + by * UB_H;  // first cell
  const int jf0 = j0 - 1;                                            // first interface
  const int tx = tid & (YT_W - 1), ty = tid / YT_W;
  const int i = i0 + tx, jf = jf0 + ty;
  const int j = jf;
  const bool in = ty >= 1 && i <= k.nx && j <= k.ny;
  const double *aline = STEP ? b.acol : nullptr;
  stage_tile<M, 1, YS_T, YT_PTS * YT_W, STEP>(
      k, b, alpha, ia, aline, tid,
      [&](int p, int &i, int &j) {
        const int r = p / YT_W;
        i = i0 + (p - r * YT_W);
        j = jf0 - 2 + r;
        return i <= k.nx && j <= k.ny + 3;
      },
      [&](int p, const double *fp, const double *fn, const double *pr) {
        const int r = p / YT_W, q = p - r * YT_W;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          sm.fp[c][r][q] = fp[c];
          sm.fn[c][r][q] = fn[c];
          sm.pr[c][r][q] = pr[c];
        }
      });
  __syncthreads();
  double G[4] = {0.0, 0.0, 0.0, 0.0};
  int lim = 0;
  if (i <= k.nx && jf <= k.ny)
    lim = iface_flux<M, 1, YT_W, YSmem::CS, LIM, STEP>(k, &sm.pr[0][ty + 2][tx], &sm.fp[0][ty + 2][tx], &sm.fn[0][ty + 2][tx],
                                            STEP ? aline[i + 2] : alpha, G);
  // a tile's first interface is the previous tile's last: counted there
  if (ty == 0 && by > 0) lim = 0;
  // the cell's loads, all issued here, before the exchange and before any store
  // (at stage 3 the new state is written over U0 itself)
  const long long o = cidx(k, i, j);
  double fxo[4], fxm[4], uin[4], u0v[4];
  if (in) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      fxo[c] = b.Fx[c * k.plane + o];
      fxm[c] = b.Fx[c * k.plane + o - 1];
    }
    if (!rhs_only) {
#pragma unroll
      for (int c = 0; c < 4; ++c) uin[c] = b.Uin[c * k.plane + o];
      if (stage > 1) {
#pragma unroll
        for (int c = 0; c < 4; ++c) u0v[c] = b.U0[c * k.plane + o];
      }
    }
  }
  // G(j-1/2) of cell j = jf: the thread YT_W lanes below, across a wave boundary
  // the previous wave's last row through LDS
  constexpr int RPW = 64 / YT_W;  // rows per wave
  const int wv = tid >> 6;
  double Gb[4];
  if (ty % RPW == RPW - 1) {
#pragma unroll
    for (int c = 0; c < 4; ++c) sm.G[wv][c][tx] = G[c];
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) Gb[c] = __shfl_up(G[c], YT_W, 64);
  __syncthreads();
  if (ty % RPW == 0 && wv > 0) {
#pragma unroll
    for (int c = 0; c < 4; ++c) Gb[c] = sm.G[wv - 1][c][tx];
  }
  // cell j = jf: Y = G(j+1/2) - G(j-1/2), X = F(i+1/2) - F(i-1/2)
  double anew = 0.0;
  double wu[4] = {0.0, 0.0, 0.0, 0.0};
  if (in) {
    double rhs[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const double X = fxo[c] - fxm[c];
      const double Y = G[c] - Gb[c];
      if (M == MODE_EXACT)
        rhs[c] = k.alpha_in_scheme ? -((alpha * X) / k.dx) - (alpha * Y) / k.dy : -(X / k.dx) - Y / k.dy;
      else
        rhs[c] = k.alpha_in_scheme ? -((alpha * X) * k.idx) - (alpha * Y) * k.idy : -(X * k.idx) - Y * k.idy;
    }
    if (rhs_only) {
#pragma unroll
      for (int c = 0; c < 4; ++c) b.R[c * k.plane + o] = rhs[c];
    } else {
      double dt;
      if (stage == 1) {
        dt = step_dt(k, alpha, g->time[parity], g->tfinal);
        if (i == 1 && j == 1 && scalars) {
          // time = time + dt after the step; the DO WHILE test gates the next one
          const double tn = g->time[parity] + dt;
          g->dt = dt;
          g->time[1 - parity] = tn;
          g->steps_done++;
          g->active[1 - parity] = time_continue_q(tn, g->tfinal);
        }
      } else {
        dt = g->dt;
      }
      double v[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const double x = uin[c] + dt * rhs[c];
        double un;
        if (stage == 1) {
          un = x;
        } else if (stage == 2) {
          const double u0 = u0v[c];
          un = (M == MODE_EXACT) ? rk2_q(u0, x) : 0.75 * u0 + 0.25 * x;
        } else {
          const double u0 = u0v[c];
          un = (M == MODE_EXACT) ? rk3_q(u0, x) : (u0 + 2.0 * x) * (1.0 / 3.0);
        }
        v[c] = un;
        b.U[c * k.plane + o] = un;
        wu[c] = k.weight * un;
      }
      write_ghosts(k, b, i, j, v);
      // checkpositivity (ex4_sedov_2d/main.f90:779-801); the fast modes divide by rho
      // through the refined reciprocal that sweep_flag and cell_alpha_fast form as well
      double ke;
      if (M == MODE_EXACT) {
        ke = (v[3] - ((v[2] * v[2]) / v[0]) / 2.) - ((v[1] * v[1]) / v[0]) / 2.;
      } else {
        const double r = rcp64(v[0]);
        ke = (v[3] - (0.5 * (v[2] * v[2])) * r) - (0.5 * (v[1] * v[1])) * r;
      }
      const bool bad = k.poscheck_le ? (v[0] <= k.eps || ke <= k.eps) : (v[0] < k.eps || ke < k.eps);
      if (bad) atomicOr(&g->postag, 1);
      if (k.use_limiter == 2) {
        const int p = posI(k, i, j);
        if (sweep_flag(k, v[0], v[1], v[2], v[3])) atomicOr(&b.bmP[p >> 5], 1u << (p & 31));
        if (v[0] < k.delta) {
          atomicOr(&b.bmR[p >> 5], 1u << (p & 31));
          atomicOr(&g->rho_any, 1);
        }
      }
      anew = (M == MODE_EXACT) ? cell_alpha(k, v[0], v[1], v[2], v[3]) : cell_alpha_fast(k, v[0], v[1], v[2], v[3]);
    }
  }
  if (rhs_only) return;
  const double am = block_max<YS_T>(anew, sm.red, tid);
  if (tid == 0 && scalars) b.apart[bx + by * gx] = am;
  if (usum) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      double v = wu[c];
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) v = v + __shfl_xor(v, off, 64);
      wu[c] = v;
    }
    if ((tid & 63) == 0) {
#pragma unroll
      for (int c = 0; c < 4; ++c) sm.ured[c][tid >> 6] = wu[c];
    }
    __syncthreads();
    if (tid < 4) b.upart[4 * (bx + by * gx) + tid] = pair_sum<YS_T / 64>(sm.ured[tid]);
  }
  if (LIM && scalars)
    zs_store<YS_T>(zs_den(lim), zs_pre(lim), b.zs + 2 * (xf_gx(k) * xf_gy(k) + bx + by * gx), sm.zred, tid);
}

template <int M, bool LIM, bool STEP = false>
__global__ SWF_STAGE_BOUNDS(M) void k_ystage(KCfg k, Bufs b, int stage, int slot_in, int parity, int flags) {
  __shared__ YSmem sm;
  int bx, by;
  xcd_tile<M>(gridDim.x, gridDim.y, bx, by);
  ystage_tile<M, LIM, STEP>(k, b, stage, slot_in, parity, flags, bx, by, gridDim.x, sm);
}

// One limited half flux outside the flux tiles (use_limiter == 1): the half of
// cell `a` on line `line` (row for DIR 0, column for DIR 1) with the Roe system
// of (a, a+1), f+ (RIGHT) or f-, from the stage input.  Returns its limiter
// change bits (density, pressure).
template <int M, int DIR, bool RIGHT>
SWF_D int zs_edge_half(const KCfg &k, const Bufs &b, int line, int a, double alpha, double ia) {
  constexpr int NS = 5;  // stencil points a-2 .. a+2 (Roe uses a, a+1)
  double st[4 * NS], prm[4 * NS];
#pragma unroll
  for (int q = 0; q < NS; ++q) {
    const long long o = DIR == 0 ? cidx(k, a - 2 + q, line) : cidx(k, line, a - 2 + q);
    double u[4], fp[4], fn[4], pr[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) u[c] = b.Uin[c * k.plane + o];
    stage_point<M, DIR>(k, u, alpha, ia, fp, fn, pr);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      st[c * NS + q] = RIGHT ? fp[c] : fn[c];
      prm[c * NS + q] = pr[c];
    }
  }
  double out[4];
  half_flux<M, DIR, RIGHT, 1, NS>(k, load_prim<NS>(&prm[2]), load_prim<NS>(&prm[3]), &st[2], out);
  return zs_limit<NS>(k.gamma1, &st[2], out);
}

// Limiter counts of the halves that wenofd evaluates (ex3_vortex_accuracy_2d/
// main.f90:224-340, :428-540 with use_limiter = 1: every cell 0..n+1 of rows
// 0..ny+1 / columns 0..nx+1, both halves) and the flux tiles do not: the ghost
// rows (x) and columns (y) in full, and in the interior lines f+ of cell n+1 and
// f- of cell 0.  One thread per half, zs_edge_halves in all; the stage input's
// ghost frame must be complete (k_bc after every stage).
template <int M>
__global__ __launch_bounds__(256) void k_zs_edges(KCfg k, Bufs b, int slot_in, int parity) {
  __shared__ int zred[8];
  DevScalars *g = b.g;
  if (!g->active[parity] || g->status != ST_OK) return;
  const int tid = threadIdx.x;
  const double alpha = alpha_from_bits(g->alpha_bits[slot_in]);
  const double ia = 1.0 / alpha;
  const int nx = k.nx, ny = k.ny;
  int t = blockIdx.x * 256 + tid;
  int lim = 0;
  if (t < zs_edge_halves(k)) {
    int dir, line, a;
    bool right;
    if (t < 4 * (nx + 2)) {  // x: rows 0 and ny+1, cells 0..nx+1
      dir = 0;
      line = t < 2 * (nx + 2) ? 0 : ny + 1;
      const int r = t % (2 * (nx + 2));
      a = r >> 1;
      right = (r & 1) == 0;
    } else if ((t -= 4 * (nx + 2)) < 2 * ny) {  // x: rows 1..ny, f+ of cell nx+1, f- of cell 0
      dir = 0;
      line = 1 + (t >> 1);
      right = (t & 1) == 0;
      a = right ? nx + 1 : 0;
    } else if ((t -= 2 * ny) < 4 * (ny + 2)) {  // y: columns 0 and nx+1, cells 0..ny+1
      dir = 1;
      line = t < 2 * (ny + 2) ? 0 : nx + 1;
      const int r = t % (2 * (ny + 2));
      a = r >> 1;
      right = (r & 1) == 0;
    } else {  // y: columns 1..nx, g+ of cell ny+1, g- of cell 0
      t -= 4 * (ny + 2);
      dir = 1;
      line = 1 + (t >> 1);
      right = (t & 1) == 0;
      a = right ? ny + 1 : 0;
    }
    if (dir == 0)
      lim = right ? zs_edge_half<M, 0, true>(k, b, line, a, alpha, ia) : zs_edge_half<M, 0, false>(k, b, line, a, alpha, ia);
    else
      lim = right ? zs_edge_half<M, 1, true>(k, b, line, a, alpha, ia) : zs_edge_half<M, 1, false>(k, b, line, a, alpha, ia);
  }
  zs_store<256>(lim & 1, (lim >> 1) & 1, b.zs + 2 * (xf_gx(k) * xf_gy(k) + ub_gx(k) * ub_gy(k) + blockIdx.x), zred, tid);
}

// ---------------------------------------------------------------------------
// Batches: independent instances (own grid, configuration, buffers, time and
// sweep) advanced together, one launch per kernel of an RK stage for all of
// them.  A block of a batch launch finds its instance and tile in a map built by
// the host (the schedule: the order of the tiles of all instances).
// ---------------------------------------------------------------------------

// buffers of stage s (1, 2, 3): stencils from S0, S1, S2, results into S1, S2, S0
SWF_D Bufs batch_stage_bufs(const BatchInst &I, int s) {
  Bufs b = I.b;
  b.Uin = (s == 1) ? I.b.U0 : (s == 2) ? I.S1 : I.S2;
  b.U = (s == 1) ? I.S1 : (s == 2) ? I.S2 : I.b.U0;
  return b;
}

// map[block] = (instance, tile of that instance)
template <int M, bool STEP = false>
__global__ SWF_FLUX_BOUNDS(M) void k_xflux_batch(const BatchInst *inst, const int2 *map, int stage, int slot_in,
                                                 int parity) {
  __shared__ XSmem sm;
  const int2 e = map[blockIdx.x];
  const BatchInst &I = inst[e.x];
  const int gx = xf_gx(I.k), gy = xf_gy(I.k);
  xflux_tile<M, false, STEP>(I.k, batch_stage_bufs(I, stage), stage, slot_in, parity, e.y % gx, e.y / gx, gx, gy,
                             sm);
}

template <int M, bool STEP = false>
__global__ SWF_STAGE_BOUNDS(M) void k_ystage_batch(const BatchInst *inst, const int2 *map, int stage, int slot_in,
                                                  int parity, int flags) {
  __shared__ YSmem sm;
  const int2 e = map[blockIdx.x];
  const BatchInst &I = inst[e.x];
  const int gx = ub_gx(I.k);
  ystage_tile<M, false, STEP>(I.k, batch_stage_bufs(I, stage), stage, slot_in, parity, flags, e.y % gx, e.y / gx,
                              gx, sm);
}

// ---------------------------------------------------------------------------
// The step diffraction program (ex6_shock_diffraction/main.f90)
// ---------------------------------------------------------------------------

// Its bc (:888-942) on one domain, in the program's order: fixed gas at rest
// (gamma, 0, 0, 1/gamma1) below and to the right, zero gradient above, the
// post-shock inflow on the left, then the loop meant for the step's face, which
// reads i after its loop has ended (i = 1) and resets column 1 to (1, 0, 0,
// 1/gamma1) in every row below 6*total.  Value of cell (i, j) after these passes;
// false for a cell they leave alone.
SWF_D bool step_bc_value(const KCfg &k, const Bufs &b, int i, int j, double v[4]) {
  const int T = k.step_total;
  if (i <= 0) {
    v[0] = k.in_rho; v[1] = k.in_mx; v[2] = 0.; v[3] = k.in_e;
  } else if (i == 1 && j < 6 * T && 1 < T) {
    v[0] = 1.; v[1] = 0.; v[2] = 0.; v[3] = 1. / k.gamma1;
  } else if (j <= 0 || i > k.nx) {
    v[0] = k.gamma; v[1] = 0.; v[2] = 0.; v[3] = 1. / k.gamma1;
  } else if (j > k.ny) {
    load_ij(k, b, i, k.ny, v);
  } else {
    return false;
  }
  return true;
}

SWF_D void step_bc_cell(const KCfg &k, long long t, int &i, int &j) {
  const long long W = k.nx + 6, nrow = 3 * W, nside = 3LL * k.ny;
  if (t < nrow) {
    j = -2 + (int)(t / W);
    i = -2 + (int)(t % W);
  } else if (t < 2 * nrow) {
    j = k.ny + 1 + (int)((t - nrow) / W);
    i = -2 + (int)((t - nrow) % W);
  } else if (t < 2 * nrow + nside) {
    const long long s = t - 2 * nrow;
    j = 1 + (int)(s / 3);
    i = -2 + (int)(s % 3);
  } else if (t < 2 * nrow + 2 * nside) {
    const long long s = t - 2 * nrow - nside;
    j = 1 + (int)(s / 3);
    i = k.nx + 1 + (int)(s % 3);
  } else {
    i = 1;
    j = 1 + (int)(t - 2 * nrow - 2 * nside);
  }
}

// the bc by nthreads threads from thread t0 (the top rows copy interior row ny,
// which the bc itself does not write)
SWF_D void step_frame(const KCfg &k, const Bufs &b, long long t0, long long nthreads) {
  const long long n = step_bc_cells(k);
  for (long long t = t0; t < n; t += nthreads) {
    int i, j;
    step_bc_cell(k, t, i, j);
    double v[4];
    if (step_bc_value(k, b, i, j, v)) store_ij(k, b, i, j, v);
  }
}

#ifdef SWF_COMMON_KERNELS
// Mode-independent grid kernels, compiled once (swf_k_exact.hip).

// End of an RK stage, one workgroup: sweep_lim when checkpositivity flagged the
// stage, then the alpha max of the new state (findalpha) from the tile partials
// (recomputed for the tiles a sweep changed) and the prescribed boundary states.
// The gate, the flag and the partials are loaded together (one memory round
// trip when no sweep is needed).
struct StageEndShared {
  SweepShared<SW_T> sh;
  double red[SW_T / 64];
  long long zc[2 * (SW_T / 64)];
};

SWF_D void stage_end_body(const KCfg &k, const Bufs &b, int stage, int slot_out, int parity, int serial_ubar,
                          StageEndShared &se) {
  SweepShared<SW_T> &sh = se.sh;
  double *red = se.red;
  long long *zc = se.zc;
  DevScalars *g = b.g;
  const int tid = threadIdx.x;
  const int nblk = ub_gx(k) * ub_gy(k);
  const int act = g->active[parity], st = g->status, flagged = g->postag, ubt = g->ubar_tiles;
  double a = 0.0;
  for (int q = tid; q < nblk; q += SW_T) a = fmax2(a, b.apart[q]);
  if (!act || st != ST_OK) return;
  if (k.use_limiter == 1) {
    // den_count / pre_count of the stage's flux limiter, from the workgroup pairs
    long long den = 0, pre = 0;
    const int np = zs_pairs(k);
    for (int q = tid; q < np; q += SW_T) {
      den += b.zs[2 * q];
      pre += b.zs[2 * q + 1];
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      den += __shfl_xor(den, off, 64);
      pre += __shfl_xor(pre, off, 64);
    }
    if ((tid & 63) == 0) {
      zc[2 * (tid >> 6)] = den;
      zc[2 * (tid >> 6) + 1] = pre;
    }
    __syncthreads();
    if (tid == 0) {
      for (int w = 0; w < SW_T / 64; ++w) {
        g->den_count += zc[2 * w];
        g->pre_count += zc[2 * w + 1];
      }
    }
  }
  if (k.use_limiter == 2 && flagged) {
    sweep_body<SW_T>(k, b, serial_ubar != 0, (ubt && !serial_ubar) ? b.upart : nullptr, sh, tid);
    after_sweep<SW_T>(k, b, sh, red, tid);
    a = 0.0;
    for (int q = tid; q < nblk; q += SW_T) a = fmax2(a, b.apart[q]);
  }
  // the step program's bc follows checkpositivity and the sweep (:736-769)
  if (k.bc == BC_STEP) step_frame(k, b, tid, SW_T);
  a = block_max<SW_T>(a, red, tid);
  if (tid == 0) {
    g->alpha_bits[slot_out] = (unsigned long long)__double_as_longlong(fmax2(a, k.alpha_const));
    // sweeps tend to follow sweeps: then the next stage kernel sums weight*u per tile
    g->ubar_tiles = (k.use_limiter == 2 && flagged && !serial_ubar) ? 1 : 0;
  }
}

__global__ __launch_bounds__(SW_T) void k_stage_end(KCfg k, Bufs b, int stage, int slot_out, int parity,
                                                    int serial_ubar) {
  __shared__ StageEndShared se;
  stage_end_body(k, b, stage, slot_out, parity, serial_ubar, se);
}

// one workgroup per instance
__global__ __launch_bounds__(SW_T) void k_stage_end_batch(const BatchInst *inst, int stage, int slot_out, int parity,
                                                          int serial_ubar) {
  __shared__ StageEndShared se;
  const BatchInst &I = inst[blockIdx.x];
  stage_end_body(I.k, batch_stage_bufs(I, stage), stage, slot_out, parity, serial_ubar, se);
}

// the time loop's start for every instance (k_begin_run of each)
__global__ void k_begin_run_batch(const BatchInst *inst, int n, const double *time, const double *tfinal) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  DevScalars *g = inst[i].b.g;
  g->time[0] = time[i];
  g->tfinal = tfinal[i];
  g->active[0] = time_continue_q(time[i], tfinal[i]);
  g->steps_done = 0;
}

__global__ void k_batch_status(const BatchInst *inst, int n, BatchStatus *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const DevScalars *g = inst[i].b.g;
  const long long s = g->steps_done;
  out[i].time = g->time[s & 1];
  out[i].steps_done = s;
  out[i].status = g->status;
  out[i].active = g->active[s & 1];
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
  if (k.bc == BC_STEP) {
    const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
    double a = 0.0;
    if (t < step_bc_cells(k)) {
      int i, j;
      step_bc_cell(k, t, i, j);
      double v[4];
      if (step_bc_value(k, b, i, j, v)) {
        store_ij(k, b, i, j, v);
        a = cell_alpha(k, v[0], v[1], v[2], v[3]);
      }
    }
    if (slot_out >= 0) alpha_atomic_max(&g->alpha_bits[slot_out], a);
    return;
  }
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

// Per-line alphas (ex6_shock_diffraction/main.f90:217-219, :449-451): findalpha's
// alpha(i,j) of the stage input over every cell, the max of each row and each
// column of a 64 x 8 tile (bx, by) into lpart (rows first: la_gx x (ny+6), then
// columns: la_gy x (nx+6)), then line_alpha_reduce_block.
SWF_D void line_alpha_block(const KCfg &k, const Bufs &b, int slot, int parity, int bx, int by,
                            double (&cred)[4][LA_W]) {
  DevScalars *g = b.g;
  if (parity >= 0 && (!g->active[parity] || g->status != ST_OK)) return;
  if (bx == 0 && by == 0 && threadIdx.x == 0) g->alpha_bits[slot] = 0ull;  // +0.0
  const int tx = threadIdx.x & (LA_W - 1), ty = threadIdx.x / LA_W;
  const int i = -2 + bx * LA_W + tx;
  double colmax = 0.0;
  for (int r = ty; r < LA_H; r += 4) {
    const int j = -2 + by * LA_H + r;
    double a = 0.0;
    if (i <= k.nx + 3 && j <= k.ny + 3) {
      const long long o = cidx(k, i, j);
      a = cell_alpha(k, b.Uin[o], b.Uin[k.plane + o], b.Uin[2 * k.plane + o], b.Uin[3 * k.plane + o]);
    }
    colmax = fmax2(colmax, a);
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      const double o = __shfl_xor(a, off, 64);
      a = (o > a) ? o : a;
    }
    if (tx == 0 && j <= k.ny + 3) b.lpart[(long long)bx * (k.ny + 6) + (j + 2)] = a;
  }
  cred[ty][tx] = colmax;
  __syncthreads();
  if (ty == 0 && i <= k.nx + 3) {
    const double m = fmax2(fmax2(cred[0][tx], cred[1][tx]), fmax2(cred[2][tx], cred[3][tx]));
    b.lpart[(long long)la_gx(k) * (k.ny + 6) + (long long)by * (k.nx + 6) + (i + 2)] = m;
  }
}
__global__ __launch_bounds__(256) void k_line_alpha(KCfg k, Bufs b, int slot, int parity) {
  __shared__ double cred[4][LA_W];
  line_alpha_block(k, b, slot, parity, blockIdx.x, blockIdx.y, cred);
}

// arow, acol and alphamax from the tile partials: 8 lanes per line (rows first,
// then columns), each block's max into alpha_bits[slot] (cleared by line_alpha_block)
SWF_D void line_alpha_reduce_block(const KCfg &k, const Bufs &b, int slot, int parity, int blk, double (&red)[4]) {
  DevScalars *g = b.g;
  if (parity >= 0 && (!g->active[parity] || g->status != ST_OK)) return;
  const int nyg = k.ny + 6, nxg = k.nx + 6, gxa = la_gx(k), gya = la_gy(k);
  const int t = blk * 256 + threadIdx.x;
  const int line = t / LA_LANES, r = t % LA_LANES;
  double m = 0.0;
  if (line < nyg) {
    for (int q = r; q < gxa; q += LA_LANES) m = fmax2(m, b.lpart[(long long)q * nyg + line]);
  } else if (line < nyg + nxg) {
    const int c = line - nyg;
    const double *cp = b.lpart + (long long)gxa * nyg + c;
    for (int q = r; q < gya; q += LA_LANES) m = fmax2(m, cp[(long long)q * nxg]);
  }
#pragma unroll
  for (int off = LA_LANES / 2; off >= 1; off >>= 1) {
    const double o = __shfl_xor(m, off, 64);
    m = (o > m) ? o : m;
  }
  if (r == 0) {
    if (line < nyg)
      b.arow[line] = m;
    else if (line < nyg + nxg)
      b.acol[line - nyg] = m;
  }
#pragma unroll
  for (int off = 32; off >= LA_LANES; off >>= 1) {
    const double o = __shfl_xor(m, off, 64);
    m = (o > m) ? o : m;
  }
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double bm = fmax2(fmax2(red[0], red[1]), fmax2(red[2], red[3]));
    atomicMax(&g->alpha_bits[slot], (unsigned long long)__double_as_longlong(bm));  // alpha >= 0: bits order
  }
}
__global__ __launch_bounds__(256) void k_line_alpha_reduce(KCfg k, Bufs b, int slot, int parity) {
  __shared__ double red[4];
  line_alpha_reduce_block(k, b, slot, parity, blockIdx.x, red);
}

// The step's walls inside wenofd (ex6_shock_diffraction/main.f90:188-197, :417-431),
// on the stage input: dir 0, before the x-pass, the cells (T-2..T, j) mirror
// (T+3..T+1, j) with m1 negated in rows 0 .. min(ny+1, 6T); dir 1, before the
// y-pass, the cells (i, 6T-2..6T) mirror (i, 6T+3..6T+1) with m2 negated in
// columns 1..T.  Sources and targets are disjoint.
SWF_D void step_reflect_block(const KCfg &k, const Bufs &b, int dir, int parity, int blk) {
  const DevScalars *g = b.g;
  if (parity >= 0 && (!g->active[parity] || g->status != ST_OK)) return;
  const int T = k.step_total;
  const int t = blk * 256 + threadIdx.x;
  if (t >= step_reflect_cells(k, dir)) return;
  int di, dj, si, sj;
  if (dir == 0) {
    const int r = t / 3, q = t - 3 * r;  // rows 0 .. min(ny+1, 6T)
    dj = sj = r;
    di = T - q;
    si = T + 1 + q;
  } else {
    const int r = t / T, c = t - r * T;
    di = si = 1 + c;
    dj = 6 * T - r;
    sj = 6 * T + 1 + r;
  }
  const long long so = cidx(k, si, sj), d = cidx(k, di, dj);
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const double v = b.Uin[c * k.plane + so];
    b.Uin[c * k.plane + d] = (c == 1 + dir) ? -v : v;
  }
}
__global__ __launch_bounds__(256) void k_step_reflect(KCfg k, Bufs b, int dir, int parity) {
  step_reflect_block(k, b, dir, parity, blockIdx.x);
}

// The same three for all step-program instances of a batch in one launch each:
// map[block] = (instance, block of that instance's launch)
__global__ __launch_bounds__(256) void k_line_alpha_batch(const BatchInst *inst, const int2 *map, int stage, int slot,
                                                          int parity) {
  __shared__ double cred[4][LA_W];
  const int2 e = map[blockIdx.x];
  const BatchInst &I = inst[e.x];
  const int gx = la_gx(I.k);
  line_alpha_block(I.k, batch_stage_bufs(I, stage), slot, parity, e.y % gx, e.y / gx, cred);
}
__global__ __launch_bounds__(256) void k_line_alpha_reduce_batch(const BatchInst *inst, const int2 *map, int stage,
                                                                 int slot, int parity) {
  __shared__ double red[4];
  const int2 e = map[blockIdx.x];
  const BatchInst &I = inst[e.x];
  line_alpha_reduce_block(I.k, batch_stage_bufs(I, stage), slot, parity, e.y, red);
}
__global__ __launch_bounds__(256) void k_step_reflect_batch(const BatchInst *inst, const int2 *map, int stage, int dir,
                                                            int parity) {
  const int2 e = map[blockIdx.x];
  const BatchInst &I = inst[e.x];
  step_reflect_block(I.k, batch_stage_bufs(I, stage), dir, parity, e.y);
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
                              double *__restrict__ left, long long n, int e12) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  double s[5];
#pragma unroll
  for (int q = 0; q < 5; ++q) s[q] = a[5 * t + q];
  const WenoEps we = e12 ? kWeps12 : kWeps6;
  right[t] = weno_right<M>(&s[2], we);
  left[t] = weno_left<M>(&s[2], we);
}

}  // namespace swf
