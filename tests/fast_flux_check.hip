// Host check of the FP64, FP32-bulk and FP32-full interface fluxes (half_flux_fast of
// sweeping_fd_weno_amd/csrc/swf_device.h, compiled for the host) against the
// bit-exact path (half_flux_exact, which the GPU tests pin to the reference) on
// Sedov states produced by the oracle.
//
//   fast_flux_check <n> <steps>
//
// For the states after 0..steps oracle steps, prints per state
//   state <s> l1 <fp64> <fp32> <fp32f> worst <fp64> <fp32> <fp32f> emin <E32> <E32f>
// with l1 = sum|F_mode - F_exact| / sum|F_exact| over all x and y interfaces,
// worst = max |F_mode - F_exact| / (|f+| + |f-|) per interface and component, and
// emin = the smallest energy after one forward-Euler stage (dt = 0.5 dx/alpha)
// with the FP32-bulk / FP32-full fluxes: the ambient gas holds E = 1e-12, so a
// flux error above ~1e-12 shows up as a negative energy there.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include <vector>

#include "swf_device.h"
extern "C" {
#include "swf_oracle.h"
}

using namespace swf;

namespace {

constexpr int NP = 6;  // stencil points of one interface: i-2 .. i+3

struct Line {
  double fp[4 * NP], fn[4 * NP];  // split fluxes, component-major
  Prim pr[3];                     // Roe inputs of points i, i+1, i+2
};

template <int M, int DIR>
void stage_line(const KCfg &k, const double u[NP][4], double alpha, Line &L) {
  for (int d = 0; d < NP; ++d) {
    double fp[4], fn[4];
    Prim p;
    if (M == MODE_EXACT) {
      split_point<MODE_EXACT, DIR>(k, u[d], alpha, 1.0 / alpha, fp, fn);
      roe_prims<MODE_EXACT, DIR>(k, u[d], p.w, p.vx, p.vy, p.h);
    } else {
      stage_point_fast<DIR>(k, u[d], 1.0 / alpha, fp, fn, p);
    }
    for (int c = 0; c < 4; ++c) {
      L.fp[c * NP + d] = fp[c];
      L.fn[c * NP + d] = fn[c];
    }
    if (d >= 2 && d <= 4) L.pr[d - 2] = p;
  }
}

// f+ and f- halves of the interface between points 2 and 3 of the line
template <int M, int DIR>
void halves(const KCfg &k, const Line &L, double a[4], double c[4]) {
  if (M == MODE_EXACT) {
    half_flux_exact<DIR, true, 1, NP>(k, L.pr[0], L.pr[1], &L.fp[2], a);
    half_flux_exact<DIR, false, 1, NP>(k, L.pr[1], L.pr[2], &L.fn[3], c);
  } else {
    half_flux_fast<M, DIR, true, 1, NP>(k, L.pr[0], L.pr[1], &L.fp[2], a);
    half_flux_fast<M, DIR, false, 1, NP>(k, L.pr[1], L.pr[2], &L.fn[3], c);
  }
}

}  // namespace

int main(int argc, char **argv) {
  const int n = argc > 1 ? atoi(argv[1]) : 64;
  const int steps = argc > 2 ? atoi(argv[2]) : 10;
  const int W = n + 6;
  swo_config oc = {};
  oc.nx = oc.ny = n;
  oc.dx = oc.dy = 1.3 / n;
  oc.gamma = 1.4;
  oc.cfl = 0.5;
  oc.eps = 1e-13;
  oc.eps2 = 1e-8;
  oc.delta = 1e-12;
  oc.bc = SWO_BC_SEDOV;
  oc.use_limiter = 2;
  oc.swp_num = 20000;
  oc.pbar_check = 1;
  oc.fail_stop = 1;
  std::vector<double> u((size_t)4 * W * W);
  swo_init_sedov(&oc, 0.244816, 1e-12, u.data());
  swo_ctx *o = swo_create(&oc);
  swo_set_state(o, u.data());
  swo_bc(o);
  KCfg k = {};
  k.nx = k.ny = n;
  k.gamma = 1.4;
  k.gamma1 = 1.4 - 1.;
  k.ig1 = 1. / k.gamma1;
  k.eps = 1e-13;
  k.eps2 = 1e-8;
  auto at = [&](int c, int i, int j) { return u[c + 4 * ((size_t)(i + 2) + (size_t)W * (j + 2))]; };
  double t = 0.0;
  for (int s = 0; s <= steps; ++s) {
    swo_get_state(o, u.data());
    const double alpha = swo_alpha_max(o), dx = oc.dx, dt = 0.5 * dx / alpha;
    double sum64 = 0, sum32 = 0, sum3f = 0, sumx = 0, worst64 = 0, worst32 = 0, worst3f = 0;
    // net fluxes of the FP32 modes for the energy check: F(i+1/2, j), G(i, j+1/2)
    std::vector<double> F((size_t)4 * W * W, 0.0), G((size_t)4 * W * W, 0.0);
    std::vector<double> Ff((size_t)4 * W * W, 0.0), Gf((size_t)4 * W * W, 0.0);
    for (int dir = 0; dir < 2; ++dir)
      for (int line = 1; line <= n; ++line)
        for (int kk = 0; kk <= n; ++kk) {
          double st[NP][4];
          for (int d = 0; d < NP; ++d)
            for (int c = 0; c < 4; ++c) st[d][c] = dir == 0 ? at(c, kk - 2 + d, line) : at(c, line, kk - 2 + d);
          Line Lx, Lf;
          double ax[4], cx[4], a64[4], c64[4], a32[4], c32[4], a3f[4], c3f[4];
          if (dir == 0) {
            stage_line<MODE_EXACT, 0>(k, st, alpha, Lx);
            stage_line<MODE_FP64, 0>(k, st, alpha, Lf);
            halves<MODE_EXACT, 0>(k, Lx, ax, cx);
            halves<MODE_FP64, 0>(k, Lf, a64, c64);
            halves<MODE_FP32, 0>(k, Lf, a32, c32);
            halves<MODE_FP32F, 0>(k, Lf, a3f, c3f);
          } else {
            stage_line<MODE_EXACT, 1>(k, st, alpha, Lx);
            stage_line<MODE_FP64, 1>(k, st, alpha, Lf);
            halves<MODE_EXACT, 1>(k, Lx, ax, cx);
            halves<MODE_FP64, 1>(k, Lf, a64, c64);
            halves<MODE_FP32, 1>(k, Lf, a32, c32);
            halves<MODE_FP32F, 1>(k, Lf, a3f, c3f);
          }
          for (int c = 0; c < 4; ++c) {
            const double fx = alpha * ax[c] - alpha * cx[c];
            const double f64 = alpha * a64[c] - alpha * c64[c], f32 = alpha * a32[c] - alpha * c32[c];
            const double f3f = alpha * a3f[c] - alpha * c3f[c];
            const double scale = fabs(alpha * ax[c]) + fabs(alpha * cx[c]) + 1e-300;
            sumx += fabs(fx);
            sum64 += fabs(f64 - fx);
            sum32 += fabs(f32 - fx);
            sum3f += fabs(f3f - fx);
            worst64 = fmax(worst64, fabs(f64 - fx) / scale);
            worst32 = fmax(worst32, fabs(f32 - fx) / scale);
            worst3f = fmax(worst3f, fabs(f3f - fx) / scale);
            if (!(f32 == f32)) worst32 = INFINITY;
            if (!(f3f == f3f)) worst3f = INFINITY;
            const size_t o0 = c + 4 * (dir == 0 ? ((size_t)(kk + 2) + (size_t)W * (line + 2))
                                                : ((size_t)(line + 2) + (size_t)W * (kk + 2)));
            (dir == 0 ? F : G)[o0] = f32;
            (dir == 0 ? Ff : Gf)[o0] = f3f;
          }
        }
    double emin = INFINITY, eminf = INFINITY;
    for (int j = 1; j <= n; ++j)
      for (int i = 1; i <= n; ++i) {
        const size_t o3 = 3 + 4 * ((size_t)(i + 2) + (size_t)W * (j + 2));
        const double e = u[o3] + dt * (-((F[o3] - F[o3 - 4]) / dx) - (G[o3] - G[o3 - 4 * (size_t)W]) / dx);
        const double ef = u[o3] + dt * (-((Ff[o3] - Ff[o3 - 4]) / dx) - (Gf[o3] - Gf[o3 - 4 * (size_t)W]) / dx);
        emin = (e == e) ? fmin(emin, e) : -INFINITY;
        eminf = (ef == ef) ? fmin(eminf, ef) : -INFINITY;
      }
    printf("state %d l1 %.3e %.3e %.3e worst %.3e %.3e %.3e emin %.3e %.3e\n", s, sum64 / sumx, sum32 / sumx,
           sum3f / sumx, worst64, worst32, worst3f, emin, eminf);
    if (s < steps) {
      double dts;
      if (swo_step(o, t, 1.0, &dts) != SWO_OK) {
        printf("oracle step failed\n");
        return 1;
      }
      t += dts;
    }
  }
  swo_destroy(o);
  return 0;
}
