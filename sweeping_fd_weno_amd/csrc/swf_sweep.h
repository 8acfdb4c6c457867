// swf_sweep.h - the use_limiter = 2 sweeping post-process on the GPU, with the
// reference's sequential semantics (ex4_sedov_2d/main.f90:1359-1396 nod_lim,
// :1401-1761 sweep_lim, :1765-1833 scalar_sweep, :1836-1895 sweeping_order).
//
// The reference scans all nx*ny cells in snake order, forward then backward, and
// adjusts a flagged cell against its successor (forward) or predecessor
// (backward).  Only flagged cells and the cascades they start do work, so the
// GPU version works on bitmaps of flagged snake positions:
//   - the RK update kernel sets a bit for every cell flagged after the stage;
//   - a pass walks the set bits in order.  A walk starting at a flagged position
//     adjusts it, hands the difference to the neighbour, and continues while the
//     neighbour is flagged; it stops at the first unflagged neighbour, which is
//     exactly where the reference's loop resumes.  Walks are computed
//     speculatively in parallel (reading unmodified cells only), then one thread
//     picks, in order, the walks that the sequential loop would really start
//     (those beginning after the previous real walk ended), and only those write.
//   - after a forward pass only cells adjusted in it (and the last position) can
//     still be flagged, so the walk itself records the next pass's bitmap; the
//     convergence test (global min P and rho) reduces to "no cell left flagged".
// Statistics (sweep_count, den_count, pre_count, per-trigger q / negative /
// positive adjustment counts) are accumulated exactly as the reference does.
// One workgroup of NT threads runs the whole sweep (all q iterations): its own
// single-workgroup kernel (swf_sweep) or the stage-end kernel.  A pass first
// compacts the set bits of its bitmap into an ordered list in LDS (one block
// scan), so its cost follows the number of flagged cells, not the grid size.
#pragma once

#include "swf_device.h"

namespace swf {

constexpr int SW_WCH = 256;          // bitmap words per scan chunk (8192 positions)
constexpr int SW_DIRTY_WORDS = 256;  // dirty-tile bitmap: up to 8192 update tiles

SWF_D void cell_of(const KCfg &k, int order, int p, int &i, int &j) {
  // cellI / cellII as one computation with selected operands (no address of i or
  // j chosen at run time, which would put them in scratch memory)
  const int n = order == 0 ? k.nx : k.ny;  // length of a snake line
  const int a = p / n + 1;                 // line
  const int r = p - (a - 1) * n;
  const int c = (a & 1) ? r + 1 : n - r;   // cell within the line
  i = order == 0 ? c : a;
  j = order == 0 ? a : c;
}
SWF_D int pos_of(const KCfg &k, int order, int i, int j) { return order == 0 ? posI(k, i, j) : posII(k, i, j); }

// A position on a snake path kept as (line, cell within the line), stepped
// without division: odd lines run up, even lines down, and a line end steps to
// the next line at the same cell.
struct Snake {
  int a, c;
};
SWF_D Snake snake_at(const KCfg &k, int order, int p) {
  const int n = order == 0 ? k.nx : k.ny;
  const int a = p / n + 1, r = p - (a - 1) * n;
  return Snake{a, (a & 1) ? r + 1 : n - r};
}
template <bool FWD>
SWF_D void snake_step(Snake &s, int n) {
  const bool up = s.a & 1;
  if (FWD) {
    if (up ? s.c < n : s.c > 1)
      s.c += up ? 1 : -1;
    else
      ++s.a;
  } else {
    if (up ? s.c > 1 : s.c < n)
      s.c += up ? -1 : 1;
    else
      --s.a;
  }
}
SWF_D void snake_ij(const Snake &s, int order, int &i, int &j) {
  i = order == 0 ? s.c : s.a;
  j = order == 0 ? s.a : s.c;
}

SWF_D void set_bit(unsigned *m, int p) { atomicOr(&m[p >> 5], 1u << (p & 31)); }
SWF_D void clear_bit(unsigned *m, int p) { atomicAnd(&m[p >> 5], ~(1u << (p & 31))); }


template <int NT>
struct SweepShared {
  union {
    int list[SW_WCH * 32];  // flagged positions of one scan chunk
    double prod[2][4][NT];  // u_bar products, two blocks of NT cells (serial u_bar)
  };
  int ends[NT];
  int woff[NT];  // first replay step of each slot's genuine walk
  int wscan[SW_WCH];
  unsigned dirty[SW_DIRTY_WORDS];
  unsigned char genuine[NT];
  unsigned char psum_bad[NT];  // the slot's speculative walk failed ex6's pressure-sum check
  int nlist, last_end, flag, all_dirty, nsteps;
  int chain0, chainlast;  // the genuine walks' first and last slot (select_genuine)
  // the positions a pass leaves flagged, in the pass's order, as its replay writes
  // them (the next pass's candidates or the convergence test's list); out_ok is 0
  // when the list is incomplete (a walk that wrote as it went, or overflow)
  int list2[SW_WCH * 32];
  int nout, out_ok, nscan;
  long long tpos, tneg;
  double ubar[4];
  double red[4][NT / 64];
};

template <int NT>
SWF_D void mark_dirty(const KCfg &k, SweepShared<NT> &sh, int i, int j) {
  const int blk = ub_of(k, i, j);
  if (blk < SW_DIRTY_WORDS * 32)
    atomicOr(&sh.dirty[blk >> 5], 1u << (blk & 31));
  else
    sh.all_dirty = 1;
}

SWF_D void load_ij(const KCfg &k, const Bufs &b, int i, int j, double v[4]) {
  const long long o = cidx(k, i, j);
#pragma unroll
  for (int c = 0; c < 4; ++c) v[c] = b.U[c * k.plane + o];
}
SWF_D void store_ij(const KCfg &k, const Bufs &b, int i, int j, const double v[4]) {
  const long long o = cidx(k, i, j);
#pragma unroll
  for (int c = 0; c < 4; ++c) b.U[c * k.plane + o] = v[c];
}

// nod_lim (ex4_sedov_2d/main.f90:1359-1396); returns true for the positive branch.
// p1, p2, pb: the pressures of v1, v2 and vb, each g1*(E - (0.5*(m1^2 + m2^2))/rho)
// as sweep_pres forms them (the reference's expression), passed in so that a walk
// forms each cell's pressure once.
// fast: the divisions by refined reciprocals and the square roots by sqrt64 (the
// fast modes' arithmetic; each within 1 ulp), else IEEE operations as the reference.
template <bool FAST>
SWF_D bool nod_lim_t(double eps, double p1, double p2, double pb, const double v1[4], const double v2[4],
                     const double vb[4], double out[4]) {
#pragma clang fp contract(off)
  constexpr bool fast = FAST;
  double t;
  bool pos;
  if (p2 > eps) {
    t = fast ? (p1 - eps) * rcp64(p1 - p2) : (p1 - eps) / (p1 - p2);
    pos = true;
  } else {
    double s12, s1b;
    if (fast) {  // the sums of squares fused, in the same order
      double d[4], e[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        d[c] = v2[c] - v1[c];
        e[c] = vb[c] - v1[c];
      }
      s12 = fma(d[0], d[0], fma(d[1], d[1], fma(d[2], d[2], d[3] * d[3])));
      s1b = fma(e[0], e[0], fma(e[1], e[1], fma(e[2], e[2], e[3] * e[3])));
    } else {
      s12 = (((v2[3] - v1[3]) * (v2[3] - v1[3]) + (v2[2] - v1[2]) * (v2[2] - v1[2])) +
             (v2[1] - v1[1]) * (v2[1] - v1[1])) +
            (v2[0] - v1[0]) * (v2[0] - v1[0]);
      s1b = (((vb[3] - v1[3]) * (vb[3] - v1[3]) + (vb[2] - v1[2]) * (vb[2] - v1[2])) +
             (vb[1] - v1[1]) * (vb[1] - v1[1])) +
            (vb[0] - v1[0]) * (vb[0] - v1[0]);
    }
    if (fast) {
      const double t_1 = (p1 - eps) * rcp64(p1 - pb);
      t = fmin2((t_1 * sqrt64(s1b)) * rcp64(sqrt64(s12)), 0.25);
    } else {
      double t_1 = (p1 - eps) / (p1 - pb);
      double d12 = sqrt(s12);
      double d1b = sqrt(s1b);
      t = fmin2((t_1 * d1b) / d12, 0.25);
    }
    pos = false;
  }
  const double omt = 1. - t;
#pragma unroll
  for (int c = 0; c < 4; ++c) out[c] = fast ? fma(t, v2[c], omt * v1[c]) : omt * v1[c] + t * v2[c];
  return pos;
}
SWF_D bool nod_lim(double eps, double p1, double p2, double pb, const double v1[4], const double v2[4],
                   const double vb[4], double out[4], bool fast = false) {
  return fast ? nod_lim_t<true>(eps, p1, p2, pb, v1, v2, vb, out) : nod_lim_t<false>(eps, p1, p2, pb, v1, v2, vb, out);
}

// A snake cursor that also carries the cell's offset in a plane (cidx), stepped by
// additions: along a line by dc (1 along a row, the pitch along a column), to the
// next line by da.
struct SnakeO {
  int a, c;
  unsigned o;
};
SWF_D SnakeO snakeo_at(const KCfg &k, int order, int p) {
  const Snake s = snake_at(k, order, p);
  int i, j;
  snake_ij(s, order, i, j);
  return SnakeO{s.a, s.c, (unsigned)cidx(k, i, j)};
}
template <bool FWD>
SWF_D void snakeo_step(SnakeO &s, int n, unsigned dc, unsigned da) {
  const bool up = s.a & 1;
  if (FWD) {
    if (up ? s.c < n : s.c > 1) {
      s.c += up ? 1 : -1;
      s.o = up ? s.o + dc : s.o - dc;
    } else {
      ++s.a;
      s.o += da;
    }
  } else {
    if (up ? s.c > 1 : s.c < n) {
      s.c += up ? -1 : 1;
      s.o = up ? s.o - dc : s.o + dc;
    } else {
      --s.a;
      s.o -= da;
    }
  }
}

// Pressure-sweep walk from position f in `order` that writes as it goes (a genuine
// walk longer than the SW_REC steps the speculative record holds).  FWD: neighbour
// f+1, else f-1.  Each step adjusts the current cell against its neighbour
// (nod_lim), stores the adjusted cell, flags it for the next pass when still
// flagged, counts its branch, and hands the difference to the neighbour; the walk
// stops at the first neighbour no longer flagged (or at the pass's last position)
// and stores it.  The path is loaded two cells ahead; each cell's pressure is
// formed once (sweep_pres = nod_lim's form).  Returns the stop position.
template <bool FWD, int NT>
SWF_D int write_walk(const KCfg &k, const Bufs &b, int order, int f, const double ub[4], double pb,
                     unsigned *next_map, long long *npos, long long *nneg, SweepShared<NT> *sh) {
#pragma clang fp contract(off)
  const int N = k.nx * k.ny;
  const int last = FWD ? N - 1 : 0;
  const int nline = order == 0 ? k.nx : k.ny;
  const double g1 = k.gamma1, eps = k.eps, wr = k.weight / k.weight;
  int p = f, i, j;
  Snake cur = snake_at(k, order, p);  // the walk's current cell
  snake_ij(cur, order, i, j);
  double v[4];
  load_ij(k, b, i, j, v);
  double a0[4], a1[4];
  int q = p;  // the cell loaded last, two ahead of p (clamped at `last`)
  Snake pre = cur;
  {
    double(*slots[2])[4] = {&a0, &a1};
#pragma unroll
    for (int d = 0; d < 2; ++d) {
      if (q != last) {
        snake_step<FWD>(pre, nline);
        q += FWD ? 1 : -1;
      }
      int ci, cj;
      snake_ij(pre, order, ci, cj);
      load_ij(k, b, ci, cj, *slots[d]);
    }
  }
  double pv = sweep_pres(k, v[0], v[1], v[2], v[3]);
  // one step with the neighbour in `slot_v`, which then receives the cell two ahead;
  // returns true at the stop position
  auto step = [&](double(&slot_v)[4]) __attribute__((always_inline)) -> bool {
    const int nb = FWD ? p + 1 : p - 1;
    double vn[4], vt[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) vn[c] = slot_v[c];
    {
      if (q != last) {
        snake_step<FWD>(pre, nline);
        q += FWD ? 1 : -1;
      }
      int ci, cj;
      snake_ij(pre, order, ci, cj);
      load_ij(k, b, ci, cj, slot_v);
    }
    const double pn = sweep_pres(k, vn[0], vn[1], vn[2], vn[3]);
    const bool pos = nod_lim(eps, pv, pn, pb, v, vn, ub, vt, k.fast_div != 0);
    double vn2[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) vn2[c] = vn[c] + wr * (v[c] - vt[c]);
    // the step program's check of an order-I adjustment (ex6_shock_diffraction/
    // main.f90:1418-1426, :1477-1485): total pressure of the pair must not rise
    if (order == 0 && k.psum_check &&
        pv + pn > (sweep_pres(k, vt[0], vt[1], vt[2], vt[3]) + sweep_pres(k, vn2[0], vn2[1], vn2[2], vn2[3])) + eps)
      b.g->status = ST_PSUM;
    store_ij(k, b, i, j, vt);
    mark_dirty(k, *sh, i, j);
    if (sweep_flag(k, vt[0], vt[1], vt[2], vt[3])) set_bit(next_map, p);
    if (pos)
      ++*npos;
    else
      ++*nneg;
    p = nb;
    snake_step<FWD>(cur, nline);
    snake_ij(cur, order, i, j);
#pragma unroll
    for (int c = 0; c < 4; ++c) v[c] = vn2[c];
    pv = sweep_pres(k, v[0], v[1], v[2], v[3]);
    const bool edge = FWD ? (p == N - 1) : (p == 0);
    if (edge || !((pv <= eps) || (v[0] <= eps))) {  // stop at the first cell not flagged
      store_ij(k, b, i, j, v);
      mark_dirty(k, *sh, i, j);
      return true;
    }
    return false;
  };
  while (true) {
    if (step(a0)) return p;
    if (step(a1)) return p;
  }
}

// Speculative pressure-sweep walk from position f: the same steps as write_walk
// (the precision mode FAST and ex6's pressure-sum check PS fixed at compile time),
// but it writes no cell: it returns the stop position and records each step's
// adjusted cell and branch (wrec, wpos) and, in the row after the last step, the
// value handed to the stop position, up to SW_REC rows; psum_hit is set when a step
// fails ex6's check (it counts only if the walk turns out to be genuine).  The
// longest walk of a pass is its critical path, and a step is issue-bound (one wave
// issues its 70-120 instructions in order, 4-8 cycles each: about 0.45 us per step
// on the Mach 2000 jet's walks), so the step carries nothing the chain does not
// need: no run-time mode branches, cell offsets stepped by additions, and the
// record of step s at wrec[(s * NT + slot) * 4 + c], 32 contiguous bytes per slot.
template <bool FWD, int NT, bool FAST, bool PS>
SWF_D int spec_walk(const KCfg &k, const Bufs &b, int order, int f, const double ub[4], double pb, int slot,
                    bool *psum_hit) {
#pragma clang fp contract(off)
  static_assert(NT <= SW_T, "walk records hold SW_T slots");
  const int N = k.nx * k.ny;
  const int last = FWD ? N - 1 : 0;
  const int nline = order == 0 ? k.nx : k.ny;
  const unsigned dc = order == 0 ? 1u : (unsigned)k.P, da = order == 0 ? (unsigned)k.P : 1u;
  // the reference's weight ratio (weight / weight) is exactly 1 for the uniform
  // weights of every program here, so the hand-over is vn + (v - vt) bit for bit
  const double g1 = k.gamma1, eps = k.eps;
  const double *U0 = b.U, *U1 = U0 + k.plane, *U2 = U1 + k.plane, *U3 = U2 + k.plane;
  auto load = [&](unsigned o, double v[4]) __attribute__((always_inline)) {
    v[0] = U0[o];
    v[1] = U1[o];
    v[2] = U2[o];
    v[3] = U3[o];
  };
  int p = f;
  SnakeO pre = snakeo_at(k, order, p);  // the cell loaded last, up to 2 ahead of p
  double v[4];
  load(pre.o, v);
  int q = p;
  // cells p+1, p+2 (p-1, p-2 backward) in two slots used in rotation (the loop body is
  // two steps); measured against deeper rings (4, 6, 8 cells, each slot refilled after
  // its last use), which issued more instructions per step than they saved in waits
  double a0[4], a1[4];
  if (q != last) {
    snakeo_step<FWD>(pre, nline, dc, da);
    q += FWD ? 1 : -1;
  }
  load(pre.o, a0);
  if (q != last) {
    snakeo_step<FWD>(pre, nline, dc, da);
    q += FWD ? 1 : -1;
  }
  load(pre.o, a1);
  double pv = sweep_pres_t<FAST>(g1, v[0], v[1], v[2], v[3]);
  unsigned posw = 0u;
  int s = 0;
  double *rp = b.wrec + 4 * (long long)slot;
  auto step = [&](double(&slot_v)[4]) __attribute__((always_inline)) -> bool {
    double vn[4], vt[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) vn[c] = slot_v[c];
    if (q != last) {
      snakeo_step<FWD>(pre, nline, dc, da);
      q += FWD ? 1 : -1;
    }
    load(pre.o, slot_v);
    const double pn = sweep_pres_t<FAST>(g1, vn[0], vn[1], vn[2], vn[3]);
    const bool pos = nod_lim_t<FAST>(eps, pv, pn, pb, v, vn, ub, vt);
    double vn2[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) vn2[c] = vn[c] + (v[c] - vt[c]);
    // ex6's check of an order-I adjustment (ex6_shock_diffraction/main.f90:1418-1426,
    // :1477-1485): the pair's total pressure must not rise
    if (PS && pv + pn > (sweep_pres_t<FAST>(g1, vt[0], vt[1], vt[2], vt[3]) +
                         sweep_pres_t<FAST>(g1, vn2[0], vn2[1], vn2[2], vn2[3])) +
                            eps)
      *psum_hit = true;
    if (s < SW_REC) {
      double *ra = (double *)__builtin_assume_aligned(rp, 32);  // two 16-byte stores
#pragma unroll
      for (int c = 0; c < 4; ++c) ra[c] = vt[c];
      if (pos) posw |= 1u << (s & 31);
      if ((s & 31) == 31) {
        b.wpos[(s >> 5) * NT + slot] = posw;
        posw = 0u;
      }
    }
    rp += 4 * NT;
    p += FWD ? 1 : -1;
#pragma unroll
    for (int c = 0; c < 4; ++c) v[c] = vn2[c];
    pv = sweep_pres_t<FAST>(g1, v[0], v[1], v[2], v[3]);
    if (p == last || !((pv <= eps) || (v[0] <= eps))) {  // stop at the first cell not flagged
      if (s + 1 < SW_REC) {
#pragma unroll
        for (int c = 0; c < 4; ++c) rp[c] = v[c];
        b.wpos[((s + 1) >> 5) * NT + slot] = posw;  // the block of steps not yet stored
      }
      return true;
    }
    ++s;
    return false;
  };
  while (true) {
    if (step(a0)) return p;
    if (step(a1)) return p;
  }
}

// Step st of a genuine walk from f with adj adjustments, from the speculative
// record of slot `slot`: st < adj writes the adjusted cell f+st (f-st backward),
// flags it for the next pass and counts its branch; st == adj writes the value
// handed to the stop position.  The same stores, flags and counts as the walk.
// Returns the position it flags for the next pass, else -1.
template <bool FWD, int NT>
SWF_D int replay_step(const KCfg &k, const Bufs &b, int order, int f, int st, int adj, int slot, unsigned *next_map,
                      long long *npos, long long *nneg, SweepShared<NT> *sh) {
#pragma clang fp contract(off)
  double v[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) v[c] = b.wrec[((long long)st * NT + slot) * 4 + c];
  const unsigned pw = b.wpos[(st >> 5) * NT + slot];
  const int p = FWD ? f + st : f - st;
  int i, j;
  cell_of(k, order, p, i, j);
  store_ij(k, b, i, j, v);
  mark_dirty(k, *sh, i, j);
  if (st == adj) return -1;
  const bool fl = sweep_flag(k, v[0], v[1], v[2], v[3]);
  if (fl) set_bit(next_map, p);
  if ((pw >> (st & 31)) & 1u)
    ++*npos;
  else
    ++*nneg;
  return fl ? p : -1;
}

// Density-sweep walk (scalar_sweep) in order I; dir +1 forward, -1 backward.
template <bool WRITE, int NT>
SWF_D int swalk(const KCfg &k, const Bufs &b, int f, int dir, SweepShared<NT> *sh) {
#pragma clang fp contract(off)
  const int N = k.nx * k.ny;
  const double delta = k.delta, wr = k.weight / k.weight;
  int p = f, i, j;
  cellI(k, p, i, j);
  double v = b.U[cidx(k, i, j)];
  while (true) {
    const int nb = p + dir;
    int ni, nj;
    cellI(k, nb, ni, nj);
    const double vn = b.U[cidx(k, ni, nj)];
    const double vn2 = vn + wr * (v - delta);
    if (WRITE) {
      const long long o = cidx(k, i, j);
      b.U[o] = delta;
      mark_dirty(k, *sh, i, j);
      if (sweep_flag(k, delta, b.U[k.plane + o], b.U[2 * k.plane + o], b.U[3 * k.plane + o]))
        set_bit(b.bmP, p);
      else
        clear_bit(b.bmP, p);
    }
    p = nb;
    i = ni;
    j = nj;
    v = vn2;
    const bool edge = (dir > 0) ? (p == N - 1) : (p == 0);
    if (edge || !(v < delta)) {
      if (WRITE) {
        const long long o = cidx(k, i, j);
        b.U[o] = v;
        mark_dirty(k, *sh, i, j);
        if (sweep_flag(k, v, b.U[k.plane + o], b.U[2 * k.plane + o], b.U[3 * k.plane + o]))
          set_bit(b.bmP, p);
        else
          clear_bit(b.bmP, p);
      }
      return p;
    }
  }
}

// Collect the set bits of words [w0, w0+SW_WCH) into sh.list (ascending).
template <int NT>
SWF_D void collect_chunk(const KCfg &k, const unsigned *map, int w0, SweepShared<NT> &sh, int tid) {
  static_assert(NT >= SW_WCH, "the scan needs one thread per bitmap word");
  unsigned word = 0;
  if (tid < SW_WCH) {
    word = (w0 + tid < k.nwords) ? map[w0 + tid] : 0u;
    sh.wscan[tid] = __popc(word);
  }
  __syncthreads();
  for (int off = 1; off < SW_WCH; off <<= 1) {  // inclusive scan (Hillis-Steele)
    int v = 0;
    if (tid < SW_WCH && tid >= off) v = sh.wscan[tid - off];
    __syncthreads();
    if (tid < SW_WCH) sh.wscan[tid] += v;
    __syncthreads();
  }
  if (tid < SW_WCH) {
    int o = sh.wscan[tid] - __popc(word);
    const int base = (w0 + tid) * 32;
    while (word) {
      const int bit = __ffs(word) - 1;
      word &= word - 1;
      sh.list[o++] = base + bit;
    }
  }
  if (tid == SW_WCH - 1) sh.nlist = sh.wscan[SW_WCH - 1];
  __syncthreads();
}

// Inclusive prefix sum of x over the 64 lanes of a wave, every lane active:
// row shifts by 1, 2, 4 and 8 scan each row of 16 lanes; lane 15 broadcast to
// rows 1 and 3, then lane 31 to rows 2 and 3, carry the rows' totals.  (Six DPP
// adds; the same scan by lane shuffles took 1.7 times as long per block scan.)
SWF_D int wave_inclusive_scan(int x) {
  x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false);  // row_shr:1
  x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false);  // row_shr:2
  x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false);  // row_shr:4
  x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false);  // row_shr:8
  x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);  // row_bcast:15
  x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);  // row_bcast:31
  return x;
}

// Exclusive prefix sum of one int per thread over the workgroup, in thread
// order; the total goes to *total (in LDS).  Uses sh.wscan.  Called by every
// thread of the workgroup.
template <int NT>
SWF_D int block_exclusive_scan(int x, SweepShared<NT> &sh, int *total, int tid) {
  static_assert(NT / 64 <= SW_WCH, "wave totals live in sh.wscan");
  static_assert(NT % 64 == 0, "whole waves");
  const int incl = wave_inclusive_scan(x);
  if ((tid & 63) == 63) sh.wscan[tid >> 6] = incl;
  __syncthreads();
  if (tid == 0) {
    int run = 0;
    for (int q = 0; q < NT / 64; ++q) {
      const int t = sh.wscan[q];
      sh.wscan[q] = run;
      run += t;
    }
    *total = run;
  }
  __syncthreads();
  return sh.wscan[tid >> 6] + incl - x;
}

// Every set position of `map`, ascending, into sh.list when there are at most
// SW_WCH*32 of them; returns their number, or -1 (sh.list not filled) when there
// are more.  Thread t scans the contiguous words [t*wpt, (t+1)*wpt), so thread
// order is position order, and one block-wide exclusive scan of the per-thread
// counts gives each thread its place in the list.
template <int NT>
SWF_D int collect_all(const KCfg &k, const unsigned *map, SweepShared<NT> &sh, int tid) {
  constexpr int BW = 32;  // words loaded together (one batch for up to 2^19 positions)
  const int wpt = (k.nwords + NT - 1) / NT;
  const int w0 = tid * wpt, w1 = min(w0 + wpt, k.nwords);
  unsigned v[BW];
  int cnt = 0;
  for (int w = w0; w < w1; w += BW) {
#pragma unroll
    for (int q = 0; q < BW; ++q) v[q] = (w + q < w1) ? map[w + q] : 0u;
#pragma unroll
    for (int q = 0; q < BW; ++q) cnt += __popc(v[q]);
  }
  const int first = block_exclusive_scan<NT>(cnt, sh, &sh.nlist, tid);
  const int total = sh.nlist;
  if (total > SW_WCH * 32) return -1;
  if (cnt) {
    int o = first;
    for (int w = w0; w < w1; w += BW) {
      if (wpt > BW) {  // more than one batch: load this one again
#pragma unroll
        for (int q = 0; q < BW; ++q) v[q] = (w + q < w1) ? map[w + q] : 0u;
      }
#pragma unroll
      for (int q = 0; q < BW; ++q) {
        unsigned word = v[q];
        while (word) {
          const int bit = __ffs(word) - 1;
          word &= word - 1;
          sh.list[o++] = (w + q) * 32 + bit;
        }
      }
    }
  }
  __syncthreads();
  return total;
}

// The walks the sequential loop really starts (genuine), from the speculative ends
// (sh.ends, -1 where no walk was computed) of the m = min(NT, n - s0) candidates of
// this chunk of the ordered list: the first candidate beyond the previous walk's end
// is genuine, and so is the first candidate beyond each genuine walk's end.  Every
// slot finds that successor of its own walk by a binary search over the chunk's
// ordered positions (into sh.woff, free until the replay); the genuine walks are the
// chain of successors from the first candidate beyond the previous pass's end, marked
// by pointer doubling (round r marks the successors 2^r links ahead of the marked
// slots), ceil(log2 m) rounds instead of one dependent hop per genuine walk.
template <bool FWD, int NT>
SWF_D void select_genuine(SweepShared<NT> &sh, const int *cl, int s0, int n, int tid) {
  const int m = min(NT, n - s0);
  auto pos = [&](int q) { return FWD ? cl[s0 + q] : cl[n - 1 - (s0 + q)]; };
  auto beyond = [&](int lo, int e) {  // first slot in [lo, m) past position e, or m
    int hi = m;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      const int fm = pos(mid);
      if (FWD ? fm > e : fm < e)
        hi = mid;
      else
        lo = mid + 1;
    }
    return lo;
  };
  const bool mine = tid < m && sh.ends[tid] >= 0;
  // the successor link of each slot (m: none; a slot without a walk ends the chain)
  int nxt = mine ? beyond(tid + 1, sh.ends[tid]) : m;
  if (tid < m) sh.woff[tid] = nxt;
  if (tid == 0) {
    const int q0 = beyond(0, sh.last_end);
    sh.chain0 = (q0 < m && sh.ends[q0] >= 0) ? q0 : -1;
    sh.chainlast = -1;
  }
  __syncthreads();
  bool marked = mine && tid == sh.chain0;
  for (int span = 1; span < m; span <<= 1) {  // marked: within span - 1 links of the start
    const int nn = (nxt < m) ? sh.woff[nxt] : m;
    __syncthreads();
    if (marked && nxt < m) sh.genuine[nxt] = 1;
    if (tid < m) sh.woff[tid] = nn;
    __syncthreads();
    marked = marked || (tid < m && sh.genuine[tid]);
    nxt = nn;
  }
  if (marked && sh.ends[tid] >= 0) {
    sh.genuine[tid] = 1;
    atomicMax(&sh.chainlast, tid);
  } else if (tid < m) {
    sh.genuine[tid] = 0;
  }
  __syncthreads();
  if (tid == 0 && sh.chainlast >= 0) sh.last_end = sh.ends[sh.chainlast];
  __syncthreads();
}

// Walks of one pass from the n positions in cl (ascending; a backward pass takes
// them from the end), NT candidates at a time.  ol (or null): the positions the
// pass leaves flagged are appended to it in the pass's order (sh.nout, sh.out_ok).
template <bool FWD, int NT>
SWF_D void pressure_walks(const KCfg &k, const Bufs &b, int order, int n, unsigned *next, const double ub[4],
                          double pb, SweepShared<NT> &sh, int tid, const int *cl, int *ol) {
  const int N = k.nx * k.ny;
  for (int s0 = 0; s0 < n; s0 += NT) {
    const int idx = s0 + tid;
    int f = -1;
    if (idx < n) f = FWD ? cl[idx] : cl[n - 1 - idx];
    const int le = sh.last_end;
    const bool cand = (f >= 0) && (FWD ? (f > le && f <= N - 2) : (f < le && f >= 1));
    bool hit = false;
    int e = -1;
    if (cand) {
      const bool ps = order == 0 && k.psum_check;
      if (k.fast_div)
        e = ps ? spec_walk<FWD, NT, true, true>(k, b, order, f, ub, pb, tid, &hit)
               : spec_walk<FWD, NT, true, false>(k, b, order, f, ub, pb, tid, &hit);
      else
        e = ps ? spec_walk<FWD, NT, false, true>(k, b, order, f, ub, pb, tid, &hit)
               : spec_walk<FWD, NT, false, false>(k, b, order, f, ub, pb, tid, &hit);
    }
    sh.ends[tid] = e;
    sh.psum_bad[tid] = hit ? 1 : 0;
    sh.genuine[tid] = 0;
    __syncthreads();
    select_genuine<FWD, NT>(sh, cl, s0, n, tid);
    // genuine walks: recorded ones are replayed by the whole workgroup (their
    // steps are independent), longer ones are walked again by their thread
    long long np = 0, nn = 0;
    int len = 0;  // cells written: the adjustments and the stop position
    if (sh.genuine[tid] && sh.psum_bad[tid]) b.g->status = ST_PSUM;
    if (sh.genuine[tid]) {
      const int end = sh.ends[tid];
      const int adj = FWD ? end - f : f - end;
      if (adj < SW_REC) {
        len = adj + 1;
      } else {
        write_walk<FWD, NT>(k, b, order, f, ub, pb, next, &np, &nn, &sh);
        sh.out_ok = 0;  // its flags are in the bitmap only
      }
    }
    sh.woff[tid] = block_exclusive_scan<NT>(len, sh, &sh.nsteps, tid);
    __syncthreads();
    const int total = sh.nsteps;
    const int m = min(NT, n - s0);
    // the replay in rounds of NT steps; step x is the x-th written cell of the chunk
    // in the pass's order, so the flagged ones, compacted round by round, extend ol
    for (int x0 = 0; x0 < total; x0 += NT) {
      const int x = x0 + tid;
      int fp = -1;
      if (x < total) {
        int lo = 0, hi = m - 1;  // the last slot whose replay starts at or before x
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (sh.woff[mid] <= x)
            lo = mid;
          else
            hi = mid - 1;
        }
        const int w = lo, st = x - sh.woff[w];
        const int fw = FWD ? cl[s0 + w] : cl[n - 1 - (s0 + w)];
        const int endw = sh.ends[w];
        const int adj = FWD ? endw - fw : fw - endw;
        fp = replay_step<FWD, NT>(k, b, order, fw, st, adj, w, next, &np, &nn, &sh);
      }
      if (ol) {
        const int base = sh.nout;
        const int r = block_exclusive_scan<NT>(fp >= 0 ? 1 : 0, sh, &sh.nscan, tid);
        if (fp >= 0) {
          if (base + r < SW_WCH * 32)
            ol[base + r] = fp;
          else
            sh.out_ok = 0;
        }
        __syncthreads();
        if (tid == 0) sh.nout = base + sh.nscan;
        __syncthreads();
      }
    }
    if (np) atomicAdd((unsigned long long *)&sh.tpos, (unsigned long long)np);
    if (nn) atomicAdd((unsigned long long *)&sh.tneg, (unsigned long long)nn);
    __syncthreads();
  }
}

// One directional pass of the pressure sweep over the flagged positions in
// `map` (order `order`).  Adjusted-but-still-flagged cells are recorded in `next`.
// More flagged positions than the list holds: chunk by chunk of SW_WCH words.
// cl_in (or null): the pass's flagged positions, n_in of them, ascending (the
// previous pass's output list, or the previous order's set remapped), used instead
// of scanning `map`, which is then not read.  ol (or null): see
// pressure_walks; it must not be sh.list when the pass scans `map` into sh.list.
template <bool FWD, int NT>
SWF_D void pressure_pass(const KCfg &k, const Bufs &b, int order, const unsigned *map, unsigned *next,
                         const double ub[4], double pb, SweepShared<NT> &sh, int tid, const int *cl_in, int n_in,
                         int *ol) {
  const int N = k.nx * k.ny;
  if (tid == 0) {
    sh.last_end = FWD ? -1 : N;
    sh.nout = 0;
    sh.out_ok = ol != nullptr;
  }
  __syncthreads();
  const int n = cl_in ? n_in : collect_all<NT>(k, map, sh, tid);
  if (n >= 0) {
    pressure_walks<FWD, NT>(k, b, order, n, next, ub, pb, sh, tid, cl_in ? cl_in : sh.list, ol);
  } else {
    if (tid == 0) sh.out_ok = 0;
    const int nchunks = (k.nwords + SW_WCH - 1) / SW_WCH;
    for (int ci = 0; ci < nchunks; ++ci) {
      const int c = FWD ? ci : (nchunks - 1 - ci);
      collect_chunk<NT>(k, map, c * SW_WCH, sh, tid);
      pressure_walks<FWD, NT>(k, b, order, sh.nlist, next, ub, pb, sh, tid, sh.list, nullptr);
    }
  }
  // the position the pass never examines (N-1 forward, 0 backward), the last of the
  // pass's order
  if (tid == 0) {
    const int pe = FWD ? N - 1 : 0;
    int i, j;
    cell_of(k, order, pe, i, j);
    double v[4];
    load_ij(k, b, i, j, v);
    if (sweep_flag(k, v[0], v[1], v[2], v[3])) {
      set_bit(next, pe);
      if (ol) {
        if (sh.nout < SW_WCH * 32)
          ol[sh.nout++] = pe;
        else
          sh.out_ok = 0;
      }
    }
  }
  __syncthreads();
}

// Forward walks of the density sweep from the n positions in sh.list.
template <int NT>
SWF_D void density_walks(const KCfg &k, const Bufs &b, int n, SweepShared<NT> &sh, int tid) {
  const int N = k.nx * k.ny;
  for (int s0 = 0; s0 < n; s0 += NT) {
    const int idx = s0 + tid;
    const int f = (idx < n) ? sh.list[idx] : -1;
    const int le = sh.last_end;
    const bool cand = (f >= 0) && f > le && f <= N - 2;
    sh.ends[tid] = cand ? swalk<false, NT>(k, b, f, +1, nullptr) : -1;
    sh.genuine[tid] = 0;
    __syncthreads();
    select_genuine<true, NT>(sh, sh.list, s0, n, tid);
    if (sh.genuine[tid]) swalk<true, NT>(k, b, f, +1, &sh);
    __syncthreads();
  }
}

// Convergence test of a cell left flagged after a backward pass (:1714-1744):
// flagged means P <= eps or rho <= eps, so it fails, except under ex5's test
// (P >= eps and rho >= eps, :1162), which a cell at exactly eps passes.
SWF_D bool converged_at(const KCfg &k, const Bufs &b, int order, int p) {
  if (!k.success_ge) return false;
  int i, j;
  cell_of(k, order, p, i, j);
  double v[4];
  load_ij(k, b, i, j, v);
  const double pr = sweep_pres(k, v[0], v[1], v[2], v[3]);
  return !((pr < k.eps) || (v[0] < k.eps));
}

template <int NT>
SWF_D void clear_map(const KCfg &k, unsigned *m, int tid) {
  for (int w = tid; w < k.nwords; w += NT) m[w] = 0u;
  __syncthreads();
}

// u_bar = sum of weights*u over the interior.  Exact mode: the reference's serial
// order (j outer, i inner, ex4_sedov_2d/main.f90:1431-1438), one lane per
// component.  Otherwise a fixed-order parallel sum: of the stage kernel's
// per-tile sums when upart is given, else over the cells.
template <int NT>
SWF_D void compute_ubar(const KCfg &k, const Bufs &b, bool serial, const double *upart, SweepShared<NT> &sh,
                        int tid) {
#pragma clang fp contract(off)
  const long long N = (long long)k.nx * k.ny;
  const double w = k.weight;
  if (serial) {
    // Lane c of wave 0 adds component c's products in the reference's order, one
    // block of NT cells at a time from LDS, with the next sixteen values read before
    // the current sixteen are added; meanwhile the other waves form the next block's
    // products in the other buffer (one barrier per block).
    static_assert(NT >= 128 && NT % 64 == 0, "a summing wave and at least one staging wave");
    const long long nb = (N + NT - 1) / NT;
    auto stage = [&](long long blk, int buf, int t0, int stride) {
      for (int t = t0; t < NT; t += stride) {
        const long long cell = blk * NT + t;
        if (cell < N) {
          const int i = 1 + (int)(cell % k.nx), j = 1 + (int)(cell / k.nx);
          const long long o = cidx(k, i, j);
#pragma unroll
          for (int c = 0; c < 4; ++c) sh.prod[buf][c][t] = w * b.U[c * k.plane + o];
        }
      }
    };
    stage(0, 0, tid, NT);
    __syncthreads();
    double acc = 0.0;
    for (long long blk = 0; blk < nb; ++blk) {
      const int cur = (int)(blk & 1);
      if (tid < 64) {
        if (tid < 4) {
          const double *pr = sh.prod[cur][tid];
          const int m = (int)min((long long)NT, N - blk * NT);
          int q = 0;
          if (m == NT) {
            constexpr int G = 16;  // sixteen dependent adds cover one LDS read latency
            double x[G];
#pragma unroll
            for (int e = 0; e < G; ++e) x[e] = pr[e];
            for (q = G; q < NT; q += G) {
              double y[G];
#pragma unroll
              for (int e = 0; e < G; ++e) y[e] = pr[q + e];
#pragma unroll
              for (int e = 0; e < G; ++e) acc = acc + x[e];
#pragma unroll
              for (int e = 0; e < G; ++e) x[e] = y[e];
            }
#pragma unroll
            for (int e = 0; e < G; ++e) acc = acc + x[e];
          } else {
            for (; q < m; ++q) acc = acc + pr[q];
          }
        }
      } else if (blk + 1 < nb) {
        stage(blk + 1, 1 - cur, tid - 64, NT - 64);
      }
      __syncthreads();
    }
    if (tid < 4) sh.ubar[tid] = acc;
  } else {
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    if (upart) {
      const int nblk = ub_gx(k) * ub_gy(k);
      for (int q = tid; q < nblk; q += NT) {
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[c] = acc[c] + upart[4 * q + c];
      }
    } else {
      for (long long t = tid; t < N; t += NT) {
        const int i = 1 + (int)(t % k.nx), j = 1 + (int)(t / k.nx);
        const long long o = cidx(k, i, j);
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[c] = acc[c] + w * b.U[c * k.plane + o];
      }
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      double v = acc[c];
      for (int off = 32; off >= 1; off >>= 1) v = v + __shfl_xor(v, off, 64);
      if ((tid & 63) == 0) sh.red[c][tid >> 6] = v;
    }
    __syncthreads();
    if (tid < 4) {
      double s = 0.0;
      for (int q = 0; q < NT / 64; ++q) s = s + sh.red[tid][q];
      sh.ubar[tid] = s;
    }
  }
  __syncthreads();
}

// The set of positions of `order` in bitmap A (listed in sh.list when nf >= 0) as
// positions of the other order, set in bitmap B (cleared beforehand).
template <int NT>
SWF_D void remap_bits(const KCfg &k, int order, const unsigned *A, unsigned *B, int nf, SweepShared<NT> &sh,
                      int tid) {
  if (nf >= 0) {
    for (int x = tid; x < nf; x += NT) {
      int i, j;
      cell_of(k, order, sh.list[x], i, j);
      set_bit(B, pos_of(k, 1 - order, i, j));
    }
  } else {
    for (int w = tid; w < k.nwords; w += NT) {
      unsigned word = A[w];
      while (word) {
        const int bit = __ffs(word) - 1;
        word &= word - 1;
        int i, j;
        cell_of(k, order, w * 32 + bit, i, j);
        set_bit(B, pos_of(k, 1 - order, i, j));
      }
    }
  }
  __syncthreads();
}

// The sweep of one stage (sweep_lim incl. scalar_sweep), run by one workgroup of
// NT threads.  Marks the update blocks whose cells it changed in sh.dirty.
// upart: the stage kernel's per-tile sums of weight*u of the state, or null.
template <int NT>
SWF_D void sweep_body(const KCfg &k, const Bufs &b, bool serial_ubar, const double *upart, SweepShared<NT> &sh,
                      int tid) {
  DevScalars *g = b.g;
  const int N = k.nx * k.ny;
  for (int w = tid; w < SW_DIRTY_WORDS; w += NT) sh.dirty[w] = 0u;
  if (tid == 0) {
    g->stage_sweeps++;
    sh.tpos = 0;
    sh.tneg = 0;
    sh.all_dirty = 0;
  }
  __syncthreads();
  // ---- scalar_sweep (density), order I
  if (g->rho_any) {
    if (tid == 0) sh.last_end = -1;
    __syncthreads();
    const int n = collect_all<NT>(k, b.bmR, sh, tid);
    if (n >= 0) {
      density_walks<NT>(k, b, n, sh, tid);
    } else {
      const int nchunks = (k.nwords + SW_WCH - 1) / SW_WCH;
      for (int c = 0; c < nchunks; ++c) {
        collect_chunk<NT>(k, b.bmR, c * SW_WCH, sh, tid);
        density_walks<NT>(k, b, sh.nlist, sh, tid);
      }
    }
    // backward pass: after the forward pass only position N-1 can be below delta
    if (tid == 0) {
      int i, j;
      cellI(k, N - 1, i, j);
      if (b.U[cidx(k, i, j)] < k.delta) swalk<true, NT>(k, b, N - 1, -1, &sh);
      g->den_count++;
    }
    __syncthreads();
  }
  // ---- u_bar and the P(u_bar) > 0 precondition
  // the density sweep changes cells after the stage kernel summed them
  compute_ubar<NT>(k, b, serial_ubar, g->rho_any ? nullptr : upart, sh, tid);
  double ub[4] = {sh.ubar[0], sh.ubar[1], sh.ubar[2], sh.ubar[3]};
  const double pb = sweep_pres(k, ub[0], ub[1], ub[2], ub[3]);  // P(u_bar)
  if (k.pbar_check) {
    if (pb <= 0.) {
      if (tid == 0) {
        g->status = ST_PBAR;
        g->pbar = pb;
      }
      __syncthreads();
      return;
    }
  }
  long long sc = 0;
  if (tid == 0) {
    g->sweep_count++;
    sc = g->sweep_count;
    sh.flag = 0;
  }
  unsigned *A = b.bmP, *B = b.bmA;
  int qdone = 0;
  int nin = -1;  // the forward pass's candidates are sh.list[0, nin) when nin >= 0, else map A
  for (int q = 1; q <= k.swp_num; ++q) {
    const int order = (q & 1) ? 0 : 1;
    clear_map<NT>(k, B, tid);
    // the forward pass lists what it leaves flagged (ascending) in sh.list2; the
    // backward pass takes its candidates from there and lists what it leaves flagged
    // in sh.list for the convergence test and the next order (the bitmaps are kept
    // as well, for the passes whose lists are incomplete)
    pressure_pass<true, NT>(k, b, order, A, B, ub, pb, sh, tid, nin >= 0 ? sh.list : nullptr, nin, sh.list2);
    const bool fwd_list = sh.out_ok;
    const int n2 = sh.nout;
    clear_map<NT>(k, A, tid);
    pressure_pass<false, NT>(k, b, order, B, A, ub, pb, sh, tid, fwd_list ? sh.list2 : nullptr, n2,
                             fwd_list ? sh.list : nullptr);
    const bool bwd_list = fwd_list && sh.out_ok;
    const int n3 = sh.nout;
    // convergence test (:1714-1744): no cell left with P <= eps or rho <= eps
    // (ex5: P < eps or rho < eps)
    if (tid == 0) {
      sh.flag = 0;
      g->pre_count++;
    }
    __syncthreads();
    // the flagged positions once, as a list when they fit
    const int nf = bwd_list ? n3 : collect_all<NT>(k, A, sh, tid);
    if (nf >= 0) {
      for (int x = tid; x < nf; x += NT)
        if (!converged_at(k, b, order, sh.list[x])) atomicOr(&sh.flag, 1);
    } else {
      for (int w = tid; w < k.nwords; w += NT) {
        unsigned word = A[w];
        while (word) {
          const int bit = __ffs(word) - 1;
          word &= word - 1;
          if (!converged_at(k, b, order, w * 32 + bit)) atomicOr(&sh.flag, 1);
        }
      }
    }
    __syncthreads();
    if (!sh.flag) {
      qdone = q;
      break;
    }
    if (q == k.swp_num) break;
    // flagged set of the next pass, in the other snake order: up to NT positions as
    // an ascending list in sh.list (each thread places its position by counting the
    // smaller ones; map A is then not read), more as a bitmap
    if (nf >= 0 && nf <= NT) {
      int key = 0;
      if (tid < nf) {
        int i, j;
        cell_of(k, order, sh.list[tid], i, j);
        key = pos_of(k, 1 - order, i, j);
        sh.list2[tid] = key;
      }
      __syncthreads();
      if (tid < nf) {
        int r = 0;
        for (int y = 0; y < nf; ++y) r += sh.list2[y] < key ? 1 : 0;
        sh.list[r] = key;
      }
      __syncthreads();
      nin = nf;
    } else {
      nin = -1;
      clear_map<NT>(k, B, tid);
      remap_bits<NT>(k, order, A, B, nf, sh, tid);
    }
    unsigned *t = A;
    A = B;
    B = t;
  }
  if (tid == 0) {
    long long qv = qdone, nn = qdone ? sh.tneg : 0, np = qdone ? sh.tpos : 0;
    if (!qdone) {
      g->fail_q = k.swp_num;
      if (k.fail_stop) g->status = ST_SWEEP_FAIL;
    }
    if (sc - 1 < k.track_cap) {
      b.trk[sc - 1] = qv;
      b.trk[k.track_cap + sc - 1] = nn;
      b.trk[2 * k.track_cap + sc - 1] = np;
    }
    g->total_sweep += qv;
    g->total_neg += nn;
    g->total_pos += np;
    if (qv > g->max_sweep) g->max_sweep = qv;
    if (nn > g->max_neg) g->max_neg = nn;
    if (np > g->max_pos) g->max_pos = np;
  }
  __syncthreads();
}

#ifdef SWF_COMMON_KERNELS
// Stand-alone sweep of the current state (the sweep_lim operator on its own).
__global__ __launch_bounds__(SW_T) void k_sweep(KCfg k, Bufs b, int serial_ubar) {
  __shared__ SweepShared<SW_T> sh;
  if (k.use_limiter == 2) sweep_body<SW_T>(k, b, serial_ubar != 0, nullptr, sh, threadIdx.x);
}
#endif

}  // namespace swf
