// swf_1d.h - the 1D program's time loop (ex2_sedov_1d/main.f90) in one kernel.
//
// The 1D line is small (the reference runs 800 cells): one workgroup holds the
// whole state in registers and LDS and runs the complete time loop of
// `DO WHILE (time <= time_terminal - 1d-10)` (:38-45) with `rk3` (:298-332) in a
// single launch, so a step costs barriers, not kernel launches.  Semantics of the
// 1D program that differ from the 2D ones:
//   - eigenvectors of the cell's own state u(m), not a Roe average (:102-133);
//   - no boundary conditions: the ghost cells keep their values, RHS is 0 there
//     (:263-280), but the RK combination still runs over them;
//   - the sweep (:538-727) walks every cell including the ghosts, its density
//     sweep (:729-789) runs unconditionally with delta = 1d-13, the mean state is
//     dot_product(u, weights) summed left to right, and a sweep that does not
//     converge within swp_num passes is silent;
//   - dt = dx*cfl/alphamax without a clamp to the final time.
// Layout: the state u(3, -2:n+3) arrives component fastest (the program's array);
// cell idx = i + 2 belongs to thread idx % NT.  LDS holds the state (SoA, for the
// sweep), the split fluxes hp/hn (reused for the reconstructed hpf/hnf) and the
// reduction slots: 9 * N + 128 doubles.
#pragma once

#include "swf_device.h"

namespace swf {

constexpr int D1_NT = 512;          // threads of the 1D workgroup (at most)
constexpr int D1_MAXN = 4 * D1_NT;  // cells with ghosts one workgroup holds

struct K1Cfg {
  int n, N;  // interior cells; cells including the ghosts (n + 6)
  int use_limiter, swp_num, track_cap;
  double dx, idx, gamma, gamma1, eps, weight, dt_num;  // dt_num = dx*cfl
};

struct D1Scalars {
  double time, tfinal, dt;
  long long steps_done;
  long long sweep_count, den_count, pre_count, stage_sweeps;
  long long total_sweep, max_sweep, total_neg, max_neg, total_pos, max_pos;
};

struct B1 {
  double *u;       // u(3, -2:n+3)
  D1Scalars *g;
  long long *trk;  // q, negative pairs, positive neighbours per sweep: 3 x track_cap
};

inline size_t d1_lds_bytes(int N) { return sizeof(double) * (9 * (size_t)N + 128); }

// findalpha (:285-296): |m/rho| + sqrt(gamma*p/rho), p = gamma1*(E - m**2/rho/2.)
SWF_D double alpha1d(const K1Cfg &k, const double *v) {
  const double pres = k.gamma1 * (v[2] - ((v[1] * v[1]) / v[0]) / 2.);
  const double sound = sqrt((k.gamma * pres) / v[0]);
  return fabs(v[1] / v[0]) + sound;
}

// checkpositivity (:380-394): rho <= eps or E - m**2/rho/2. <= eps
SWF_D bool bad1d(const K1Cfg &k, const double *v) {
  return v[0] <= k.eps || (v[2] - ((v[1] * v[1]) / v[0]) / 2.) <= k.eps;
}

// fflux (:59-73) split with the global alpha (:81-84): vx = m/rho when
// rho > 1d-8 (a binary128 literal; for binary64 rho that is rho >= the double
// nearest 1d-8, which lies above it), else 0
template <int M>
SWF_D void split1d(const K1Cfg &k, const double *u, double a, double ia, double *hp, double *hn) {
  const double rho = u[0], mx = u[1], e = u[2];
  const double vx = (rho >= qconst::e8_hi) ? mx / rho : 0.0;
  const double pres = k.gamma1 * (e - (0.5 * rho) * (vx * vx));
  const double f[3] = {mx, rho * (vx * vx) + pres, (e + pres) * vx};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double fa = (M == MODE_EXACT) ? f[c] / a : f[c] * ia;
    hp[c] = 0.5 * (u[c] + fa);
    hn[c] = 0.5 * (u[c] - fa);
  }
}

// wenofd's loop body at point m (:93-244): local eigenvectors, projection of the
// 5-point stencils of hp and hn, WENO of each characteristic field, back-projection.
// `sh` points at hp component 0 of cell m; components N apart, hn at 3N.
template <int M>
SWF_D void weno1d(const K1Cfg &k, const double *um, const double *sh, int N, double *fp, double *fn) {
  const double g1 = k.gamma1;
  const double r = um[0], mx = um[1], e = um[2];
  const double vxm = mx / r;
  const double hm = (g1 * (e - (0.5 * (mx * mx)) / r) + e) / r;
  const double qm = (0.5 * vxm) * vxm;
  const double cm = sqrt(g1 * (hm - qm));
  double t0 = vxm * cm;
  const double evr[3][3] = {{1.0, 1.0, 1.0}, {vxm - cm, vxm, vxm + cm}, {hm - t0, qm, hm + t0}};
  const double rcm = 1. / cm;
  const double b1 = (g1 * rcm) * rcm, b2 = qm * b1;
  t0 = vxm * rcm;
  const double t1 = b1 * vxm, t2 = 0.5 * b1;
  const double evl[3][3] = {{0.5 * (b2 + t0), -0.5 * (t1 + rcm), t2},
                            {1. - b2, t1, -b1},
                            {0.5 * (b2 - t0), -0.5 * (t1 - rcm), t2}};
  // one characteristic field at a time (bounded register use); the field's row
  // of evl is selected, not indexed, so that nothing lives in scratch
  double ur0 = 0., ur1 = 0., ur2 = 0., ul0 = 0., ul1 = 0., ul2 = 0.;
#pragma unroll 1
  for (int c = 0; c < 3; ++c) {
    const double l0 = c == 0 ? evl[0][0] : (c == 1 ? evl[1][0] : evl[2][0]);
    const double l1 = c == 0 ? evl[0][1] : (c == 1 ? evl[1][1] : evl[2][1]);
    const double l2 = c == 0 ? evl[0][2] : (c == 1 ? evl[1][2] : evl[2][2]);
    double cp[5], cn[5];
#pragma unroll
    for (int j = 0; j < 5; ++j) {
      cp[j] = (l0 * sh[j - 2] + l1 * sh[N + j - 2]) + l2 * sh[2 * N + j - 2];
      cn[j] = (l0 * sh[3 * N + j - 2] + l1 * sh[4 * N + j - 2]) + l2 * sh[5 * N + j - 2];
    }
    const double r = weno_right<M>(&cp[2], kWeps6), l = weno_left<M>(&cn[2], kWeps6);
    ur0 = c == 0 ? r : ur0;
    ur1 = c == 1 ? r : ur1;
    ur2 = c == 2 ? r : ur2;
    ul0 = c == 0 ? l : ul0;
    ul1 = c == 1 ? l : ul1;
    ul2 = c == 2 ? l : ul2;
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    fp[c] = (evr[c][0] * ur0 + evr[c][1] * ur1) + evr[c][2] * ur2;
    fn[c] = (evr[c][0] * ul0 + evr[c][1] * ul1) + evr[c][2] * ul2;
  }
}

// ---------------------------------------------------------------------------
// The sweep (:538-789), run by one wave on the LDS state su (SoA, pitch N).
// A walk visits its positions in order; a flagged position j changes j and its
// successor.  Each 64-position chunk evaluates the flags of its positions at once
// (ballot); a position changed by the step before it is re-tested on the spot,
// every other position keeps its ballot flag, so the visits are those of the
// serial loop.  All lanes run the serial steps on the same values.
// ---------------------------------------------------------------------------

// the sweep's flag and pressure: gamma1*(E - 0.5*m**2/rho) (:596, :645)
SWF_D double pres1d(const K1Cfg &k, double r, double m, double e) { return k.gamma1 * (e - (0.5 * (m * m)) / r); }
SWF_D bool pflag1d(const K1Cfg &k, double r, double m, double e) {
  return pres1d(k, r, m, e) <= k.eps || r <= k.eps;
}

struct Sweep1dCounts {
  long long tpos, tneg;
};

// nod_lim (:488-532); note its pressure is gamma1*(E - 0.5*(m**2/rho))
SWF_D void nodlim1d(const K1Cfg &k, const double *v1, const double *v2, const double *vb, double *out,
                    Sweep1dCounts &cnt) {
  const double g1 = k.gamma1, eps = k.eps;
  const double p1 = g1 * (v1[2] - 0.5 * ((v1[1] * v1[1]) / v1[0]));
  const double p2 = g1 * (v2[2] - 0.5 * ((v2[1] * v2[1]) / v2[0]));
  double t;
  if (p2 > eps) {
    t = (p1 - eps) / (p1 - p2);
    cnt.tpos++;
  } else {
    const double pb = g1 * (vb[2] - 0.5 * ((vb[1] * vb[1]) / vb[0]));
    const double t_1 = (p1 - eps) / (p1 - pb);
    const double d12 = sqrt(((v2[2] - v1[2]) * (v2[2] - v1[2]) + (v2[1] - v1[1]) * (v2[1] - v1[1])) +
                            (v2[0] - v1[0]) * (v2[0] - v1[0]));
    const double d1b = sqrt(((vb[2] - v1[2]) * (vb[2] - v1[2]) + (vb[1] - v1[1]) * (vb[1] - v1[1])) +
                            (vb[0] - v1[0]) * (vb[0] - v1[0]));
    const double tt = (t_1 * d1b) / d12;
    t = (0.25 < tt) ? 0.25 : tt;  // min(t_1*d_1_bar/d_1_2, 0.25)
    cnt.tneg++;
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) out[c] = (1 - t) * v1[c] + t * v2[c];
}

// One walk.  PRES: the pressure sweep (flag pflag1d, nod_lim, all components);
// else the density sweep (flag rho < 1d-13, density only).  FWD: j = -2..n+2
// towards j+1; else j = n+3..-1 towards j-1.  Lane 0 performs the serial steps
// and their re-tests (broadcast to the wave); the chunk flags are every lane's.
template <bool PRES, bool FWD>
SWF_D void walk1d(const K1Cfg &k, double *su, int N, const double *ub, double wr, Sweep1dCounts &cnt, int lane) {
  const double delta = 1e-13;
  const int len = N - 1;
  for (int base = 0; base < len; base += 64) {
    const int w0 = base + lane;
    bool f = false;
    if (w0 < len) {
      const int p = FWD ? w0 : N - 1 - w0;
      f = PRES ? pflag1d(k, su[p], su[N + p], su[2 * N + p]) : su[p] < delta;
    }
    unsigned long long mask = __ballot(f);
    int d = -1;  // walk index changed by the previous step (re-tested)
    while (true) {
      int w;
      if (d >= 0 && d < base + 64 && d < len) {
        w = d;
        d = -1;
        int now = 0;
        if (lane == 0) {
          const int p = FWD ? w : N - 1 - w;
          now = PRES ? pflag1d(k, su[p], su[N + p], su[2 * N + p]) : su[p] < delta;
        }
        if (!__shfl(now, 0)) continue;
      } else {
        if (mask == 0ull) break;
        w = base + __builtin_ctzll(mask);
        mask &= mask - 1ull;
      }
      if (lane == 0) {
        const int p = FWD ? w : N - 1 - w;
        const int pn = FWD ? p + 1 : p - 1;
        if (PRES) {
          const double vj[3] = {su[p], su[N + p], su[2 * N + p]};
          const double vp[3] = {su[pn], su[N + pn], su[2 * N + pn]};
          double vt[3];
          nodlim1d(k, vj, vp, ub, vt, cnt);
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            su[c * N + p] = vt[c];
            su[c * N + pn] = vp[c] + wr * (vj[c] - vt[c]);
          }
        } else {
          const double vj = su[p], vn = su[pn];
          su[pn] = vn + wr * (vj - delta);
          su[p] = delta;
        }
      }
      d = w + 1;
      if (d < base + 64) mask &= ~(1ull << (d - base));
    }
    __ballot(1);  // the wave sees lane 0's last writes before the next chunk's flags
  }
}

// sweep_lim (:538-727) with scalar_sweep (:729-789); updates the statistics
template <int NWAVE>
SWF_D void sweep1d(const K1Cfg &k, double *su, int N, D1Scalars &st, long long *trk, int lane) {
  const double w = k.weight, wr = w / w;
  Sweep1dCounts cnt{0, 0};
  walk1d<false, true>(k, su, N, nullptr, wr, cnt, lane);
  walk1d<false, false>(k, su, N, nullptr, wr, cnt, lane);
  st.den_count++;
  // dot_product(u(c,:), weights), left to right from 0 (:577-579)
  double ub[3] = {0., 0., 0.};
  for (int i = 0; i < N; ++i) {
#pragma unroll
    for (int c = 0; c < 3; ++c) ub[c] = ub[c] + su[c * N + i] * w;
  }
  st.sweep_count++;
  const long long sc = st.sweep_count;
  cnt.tpos = cnt.tneg = 0;
  for (int q = 1; q <= k.swp_num; ++q) {
    walk1d<true, true>(k, su, N, ub, wr, cnt, lane);
    walk1d<true, false>(k, su, N, ub, wr, cnt, lane);
    double pmin = 1e20, rmin = 1e20;
    for (int i = lane; i < N; i += 64) {
      const double pr = pres1d(k, su[i], su[N + i], su[2 * N + i]);
      pmin = (pr < pmin) ? pr : pmin;
      rmin = (su[i] < rmin) ? su[i] : rmin;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const double op = __shfl_xor(pmin, off), orr = __shfl_xor(rmin, off);
      pmin = (op < pmin) ? op : pmin;
      rmin = (orr < rmin) ? orr : rmin;
    }
    st.pre_count++;
    if (pmin > k.eps && rmin > k.eps) {
      cnt.tneg = __shfl(cnt.tneg, 0);
      cnt.tpos = __shfl(cnt.tpos, 0);
      if (lane == 0 && sc - 1 < k.track_cap) {
        trk[sc - 1] = q;
        trk[k.track_cap + sc - 1] = cnt.tneg;
        trk[2 * k.track_cap + sc - 1] = cnt.tpos;
      }
      st.total_sweep += q;
      st.total_neg += cnt.tneg;
      st.total_pos += cnt.tpos;
      if (q > st.max_sweep) st.max_sweep = q;
      if (cnt.tneg > st.max_neg) st.max_neg = cnt.tneg;
      if (cnt.tpos > st.max_pos) st.max_pos = cnt.tpos;
      return;
    }
  }
}

// max of alpha and OR of the positivity flag over the workgroup; slots by parity
SWF_D void block_max_or(double &a, int &f, double *red, int parity, int tid) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_xor(a, off);
    a = (o > a) ? o : a;
    f |= __shfl_xor(f, off);
  }
  double *slot = red + 64 * parity;
  if ((tid & 63) == 0) {
    slot[tid >> 6] = a;
    slot[32 + (tid >> 6)] = (double)f;
  }
  __syncthreads();
  const int nw = blockDim.x >> 6;
  a = slot[0];
  f = slot[32] != 0.0;
  for (int i = 1; i < nw; ++i) {
    a = (slot[i] > a) ? slot[i] : a;
    f |= slot[32 + i] != 0.0;
  }
}

// The time loop: up to max_steps steps while time <= tfinal - 1d-10 (binary128).
template <int M, int CPT>
__global__ void __launch_bounds__(D1_NT) k_run1d(K1Cfg k, B1 b, long long max_steps) {
  HIP_DYNAMIC_SHARED(double, lds)
  const int N = k.N, n = k.n;
  const int tid = threadIdx.x, NT = blockDim.x;
  double *su = lds;
  double *sh = lds + 3 * N;
  double *red = lds + 9 * N;
  D1Scalars st = *b.g;

  double u[CPT][3], u0[CPT][3];
#pragma unroll
  for (int q = 0; q < CPT; ++q) {
    const int ix = tid + q * NT;
#pragma unroll
    for (int c = 0; c < 3; ++c) u[q][c] = (ix < N) ? b.u[3 * ix + c] : 1.0;
  }
  int parity = 0;
  double amax = 0.0;
  int bad = 0;
#pragma unroll
  for (int q = 0; q < CPT; ++q)
    if (tid + q * NT < N) {
      const double a = alpha1d(k, u[q]);
      amax = (a > amax) ? a : amax;
    }
  block_max_or(amax, bad, red, parity, tid);
  parity ^= 1;

  double t = st.time, dt = st.dt;
  long long steps = 0;
  while (steps < max_steps && time_continue_q(t, st.tfinal)) {
    for (int s = 1; s <= 3; ++s) {
      if (s == 1) {
        dt = k.dt_num / amax;
#pragma unroll
        for (int q = 0; q < CPT; ++q)
#pragma unroll
          for (int c = 0; c < 3; ++c) u0[q][c] = u[q][c];
      }
      const double ia = 1.0 / amax;
#pragma unroll
      for (int q = 0; q < CPT; ++q) {
        const int ix = tid + q * NT;
        if (ix < N) {
          double hp[3], hn[3];
          split1d<M>(k, u[q], amax, ia, hp, hn);
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            sh[c * N + ix] = hp[c];
            sh[(3 + c) * N + ix] = hn[c];
          }
        }
      }
      __syncthreads();
      double fp[CPT][3], fn[CPT][3];
#pragma unroll
      for (int q = 0; q < CPT; ++q) {
        const int ix = tid + q * NT;
        if (ix >= 2 && ix <= n + 3) weno1d<M>(k, u[q], sh + ix, N, fp[q], fn[q]);
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < CPT; ++q) {
        const int ix = tid + q * NT;
        if (ix >= 2 && ix <= n + 3) {
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            sh[c * N + ix] = fp[q][c];
            sh[(3 + c) * N + ix] = fn[q][c];
          }
        }
      }
      __syncthreads();
      // scheme (:263-280), the RK combination (:298-332), checkpositivity and the
      // next findalpha on the new state
      double anew = 0.0;
      int flag = 0;
#pragma unroll
      for (int q = 0; q < CPT; ++q) {
        const int ix = tid + q * NT;
        if (ix < N) {
          const int i = ix - 2;
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            double rhs = 0.0;
            if (i >= 1 && i <= n) {
              const double fl0 = amax * (sh[c * N + ix - 1] - sh[(3 + c) * N + ix]);
              const double fl1 = amax * (sh[c * N + ix] - sh[(3 + c) * N + ix + 1]);
              rhs = (M == MODE_EXACT) ? -(fl1 - fl0) / k.dx : (fl0 - fl1) * k.idx;
            }
            const double x = u[q][c] + dt * rhs;
            double un;
            if (s == 1)
              un = x;
            else if (s == 2)
              un = (M == MODE_EXACT) ? rk2_q(u0[q][c], x) : 0.75 * u0[q][c] + 0.25 * x;
            else
              un = (M == MODE_EXACT) ? rk3_q(u0[q][c], x) : (u0[q][c] + 2.0 * x) * (1.0 / 3.0);
            u[q][c] = un;
          }
          if (i >= 1 && i <= n && bad1d(k, u[q])) flag = 1;
          const double a = alpha1d(k, u[q]);
          anew = (a > anew) ? a : anew;
        }
      }
      block_max_or(anew, flag, red, parity, tid);
      parity ^= 1;
      amax = anew;
      if (flag && k.use_limiter == 2) {
#pragma unroll
        for (int q = 0; q < CPT; ++q) {
          const int ix = tid + q * NT;
          if (ix < N)
#pragma unroll
            for (int c = 0; c < 3; ++c) su[c * N + ix] = u[q][c];
        }
        __syncthreads();
        if (tid < 64) sweep1d<D1_NT / 64>(k, su, N, st, b.trk, tid);
        if (tid == 0) st.stage_sweeps++;
        __syncthreads();
        amax = 0.0;
        int none = 0;
#pragma unroll
        for (int q = 0; q < CPT; ++q) {
          const int ix = tid + q * NT;
          if (ix < N) {
#pragma unroll
            for (int c = 0; c < 3; ++c) u[q][c] = su[c * N + ix];
            const double a = alpha1d(k, u[q]);
            amax = (a > amax) ? a : amax;
          }
        }
        block_max_or(amax, none, red, parity, tid);
        parity ^= 1;
      }
    }
    t = t + dt;
    ++steps;
  }
#pragma unroll
  for (int q = 0; q < CPT; ++q) {
    const int ix = tid + q * NT;
    if (ix < N)
#pragma unroll
      for (int c = 0; c < 3; ++c) b.u[3 * ix + c] = u[q][c];
  }
  if (tid == 0) {
    st.time = t;
    st.dt = dt;
    st.steps_done += steps;
    *b.g = st;
  }
}

}  // namespace swf
