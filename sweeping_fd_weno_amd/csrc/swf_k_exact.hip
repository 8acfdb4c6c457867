// swf_k_exact.hip - FP64-exact stage kernels and the mode-independent
// kernels (sweep, boundary conditions, alpha, layout transposes).  Compiled with
// -ffp-contract=off: every product and sum is rounded separately, as in the
// reference build (x86-64 SSE2 without FMA).
#define SWF_COMMON_KERNELS 1
#include <algorithm>

#include "swf_kernels.h"
#include "swf_launch.h"

namespace swf {

template <int M, int CPT>
static hipError_t run1d(const K1Cfg &k, const B1 &b, hipStream_t s, long long max_steps) {
  const size_t lds = d1_lds_bytes(k.N);
  if (lds > 65536) {
    const hipError_t e = hipFuncSetAttribute((const void *)k_run1d<M, CPT>,
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL((k_run1d<M, CPT>), dim3(1), dim3(run1d_threads(k.N)), lds, s, k, b, max_steps);
  return hipGetLastError();
}

hipError_t launch_run1d_m0(const K1Cfg &k, const B1 &b, hipStream_t s, long long max_steps) {
  return run1d_cpt(k.N) == 2 ? run1d<MODE_EXACT, 2>(k, b, s, max_steps) : run1d<MODE_EXACT, 4>(k, b, s, max_steps);
}
hipError_t launch_run1d(int mode, const K1Cfg &k, const B1 &b, hipStream_t s, long long max_steps) {
  return mode == MODE_EXACT ? launch_run1d_m0(k, b, s, max_steps) : launch_run1d_m12(mode, k, b, s, max_steps);
}

void launch_xflux_m0(const KCfg &k, const Bufs &b, hipStream_t s, int stage, int slot_in, int parity) {
  if (k.use_limiter == 1)
    hipLaunchKernelGGL((k_xflux<MODE_EXACT, true>), grid_xflux(k), dim3(256), 0, s, k, b, stage, slot_in, parity);
  else if (k.alpha_lines)
    hipLaunchKernelGGL((k_xflux<MODE_EXACT, false, true>), grid_xflux(k), dim3(256), 0, s, k, b, stage, slot_in,
                       parity);
  else
    hipLaunchKernelGGL((k_xflux<MODE_EXACT, false>), grid_xflux(k), dim3(256), 0, s, k, b, stage, slot_in, parity);
}
void launch_ystage_m0(const KCfg &k, const Bufs &b, hipStream_t s, int stage, int slot_in, int parity, int flags) {
  if (k.use_limiter == 1)
    hipLaunchKernelGGL((k_ystage<MODE_EXACT, true>), grid_ystage(k), dim3(YS_T), 0, s, k, b, stage, slot_in, parity,
                       flags);
  else if (k.alpha_lines)
    hipLaunchKernelGGL((k_ystage<MODE_EXACT, false, true>), grid_ystage(k), dim3(YS_T), 0, s, k, b, stage, slot_in,
                       parity, flags);
  else
    hipLaunchKernelGGL((k_ystage<MODE_EXACT, false>), grid_ystage(k), dim3(YS_T), 0, s, k, b, stage, slot_in, parity,
                       flags);
}
void launch_zs_edges_m0(const KCfg &k, const Bufs &b, hipStream_t s, int slot_in, int parity) {
  hipLaunchKernelGGL(k_zs_edges<MODE_EXACT>, dim3(grid_zs_edges(k)), dim3(256), 0, s, k, b, slot_in, parity);
}
void launch_weno_points_m0(const double *a, double *r, double *l, long long n, int e12, hipStream_t s) {
  hipLaunchKernelGGL(k_weno_points<MODE_EXACT>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, r, l, n, e12);
}

void launch_xflux(int mode, const KCfg &k, const Bufs &b, hipStream_t s, int stage, int slot_in, int parity) {
  if (mode == MODE_EXACT)
    launch_xflux_m0(k, b, s, stage, slot_in, parity);
  else
    launch_xflux_m12(mode, k, b, s, stage, slot_in, parity);
}
void launch_ystage(int mode, const KCfg &k, const Bufs &b, hipStream_t s, int stage, int slot_in, int parity,
                   int flags) {
  if (mode == MODE_EXACT)
    launch_ystage_m0(k, b, s, stage, slot_in, parity, flags);
  else
    launch_ystage_m12(mode, k, b, s, stage, slot_in, parity, flags);
}
void launch_xflux_batch_m0(const BatchInst *inst, const int2 *map, int nblocks, hipStream_t s, int stage,
                           int slot_in, int parity, int step) {
  if (step)
    hipLaunchKernelGGL((k_xflux_batch<MODE_EXACT, true>), dim3(nblocks), dim3(256), 0, s, inst, map, stage, slot_in,
                       parity);
  else
    hipLaunchKernelGGL(k_xflux_batch<MODE_EXACT>, dim3(nblocks), dim3(256), 0, s, inst, map, stage, slot_in, parity);
}
void launch_ystage_batch_m0(const BatchInst *inst, const int2 *map, int nblocks, hipStream_t s, int stage,
                            int slot_in, int parity, int flags, int step) {
  if (step)
    hipLaunchKernelGGL((k_ystage_batch<MODE_EXACT, true>), dim3(nblocks), dim3(YS_T), 0, s, inst, map, stage, slot_in,
                       parity, flags);
  else
    hipLaunchKernelGGL(k_ystage_batch<MODE_EXACT>, dim3(nblocks), dim3(YS_T), 0, s, inst, map, stage, slot_in, parity,
                       flags);
}
// step: the blocks of step-program instances (their wenofd variant)
void launch_xflux_batch(int mode, const BatchInst *inst, const int2 *map, int nblocks, hipStream_t s, int stage,
                        int slot_in, int parity, int step) {
  if (mode == MODE_EXACT)
    launch_xflux_batch_m0(inst, map, nblocks, s, stage, slot_in, parity, step);
  else
    launch_xflux_batch_m12(mode, inst, map, nblocks, s, stage, slot_in, parity, step);
}
void launch_ystage_batch(int mode, const BatchInst *inst, const int2 *map, int nblocks, hipStream_t s, int stage,
                         int slot_in, int parity, int flags, int step) {
  if (mode == MODE_EXACT)
    launch_ystage_batch_m0(inst, map, nblocks, s, stage, slot_in, parity, flags, step);
  else
    launch_ystage_batch_m12(mode, inst, map, nblocks, s, stage, slot_in, parity, flags, step);
}
void launch_stage_end_batch(const BatchInst *inst, int n, hipStream_t s, int stage, int slot_out, int parity,
                            int serial_ubar) {
  hipLaunchKernelGGL(k_stage_end_batch, dim3(n), dim3(SW_T), 0, s, inst, stage, slot_out, parity, serial_ubar);
}
void launch_begin_run_batch(const BatchInst *inst, int n, const double *time, const double *tfinal, hipStream_t s) {
  hipLaunchKernelGGL(k_begin_run_batch, dim3((n + 63) / 64), dim3(64), 0, s, inst, n, time, tfinal);
}
void launch_batch_status(const BatchInst *inst, int n, BatchStatus *out, hipStream_t s) {
  hipLaunchKernelGGL(k_batch_status, dim3((n + 63) / 64), dim3(64), 0, s, inst, n, out);
}
void launch_zs_edges(int mode, const KCfg &k, const Bufs &b, hipStream_t s, int slot_in, int parity) {
  if (mode == MODE_EXACT)
    launch_zs_edges_m0(k, b, s, slot_in, parity);
  else
    launch_zs_edges_m12(mode, k, b, s, slot_in, parity);
}
void launch_weno_points(int mode, const double *a, double *r, double *l, long long n, int e12, hipStream_t s) {
  if (mode == MODE_EXACT)
    launch_weno_points_m0(a, r, l, n, e12, s);
  else
    launch_weno_points_m12(mode, a, r, l, n, e12, s);
}

// checkpositivity flag and sweep bitmaps of the current interior (no RK update)
__global__ __launch_bounds__(256) void k_flags(KCfg k, Bufs b) {
  const int i = 1 + blockIdx.x * 64 + threadIdx.x, j = 1 + blockIdx.y * 4 + threadIdx.y;
  if (i > k.nx || j > k.ny) return;
  const long long o = cidx(k, i, j);
  double v[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) v[c] = b.U[c * k.plane + o];
  const double ke = (v[3] - ((v[2] * v[2]) / v[0]) / 2.) - ((v[1] * v[1]) / v[0]) / 2.;
  const bool bad = k.poscheck_le ? (v[0] <= k.eps || ke <= k.eps) : (v[0] < k.eps || ke < k.eps);
  if (bad) atomicOr(&b.g->postag, 1);
  const int p = posI(k, i, j);
  if (sweep_flag(k, v[0], v[1], v[2], v[3])) atomicOr(&b.bmP[p >> 5], 1u << (p & 31));
  if (v[0] < k.delta) {
    atomicOr(&b.bmR[p >> 5], 1u << (p & 31));
    atomicOr(&b.g->rho_any, 1);
  }
}

__global__ void k_clear_flags(KCfg k, Bufs b) {
  for (int w = blockIdx.x * blockDim.x + threadIdx.x; w < k.nwords; w += gridDim.x * blockDim.x) {
    b.bmR[w] = 0u;
    b.bmP[w] = 0u;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    b.g->postag = 0;
    b.g->rho_any = 0;
  }
}

__global__ void k_begin_run(Bufs b, double time, double tfinal, int active) {
  b.g->time[0] = time;
  b.g->tfinal = tfinal;
  b.g->active[0] = active;
  b.g->steps_done = 0;
}
__global__ void k_set_alpha(Bufs b, int slot, unsigned long long bits, int p, int active) {
  b.g->alpha_bits[slot] = bits;
  if (p >= 0) b.g->active[p] = active;
}
__global__ void k_upload_reset(Bufs b) {
  b.g->alpha_bits[0] = 0ull;
  b.g->swept = 0;
}
void launch_begin_run(const Bufs &b, hipStream_t s, double time, double tfinal, int active) {
  hipLaunchKernelGGL(k_begin_run, dim3(1), dim3(1), 0, s, b, time, tfinal, active);
}
void launch_set_alpha(const Bufs &b, hipStream_t s, int slot, unsigned long long bits, int p, int active) {
  hipLaunchKernelGGL(k_set_alpha, dim3(1), dim3(1), 0, s, b, slot, bits, p, active);
}
void launch_upload_reset(const Bufs &b, hipStream_t s) {
  hipLaunchKernelGGL(k_upload_reset, dim3(1), dim3(1), 0, s, b);
}

void launch_flags(const KCfg &k, const Bufs &b, hipStream_t s) {
  hipLaunchKernelGGL(k_clear_flags, dim3(64), dim3(256), 0, s, k, b);
  hipLaunchKernelGGL(k_flags, grid_cells(k), dim3(64, 4), 0, s, k, b);
}
void launch_stage_end(const KCfg &k, const Bufs &b, hipStream_t s, int stage, int slot_out, int parity,
                      int serial_ubar) {
  hipLaunchKernelGGL(k_stage_end, dim3(1), dim3(SW_T), 0, s, k, b, stage, slot_out, parity, serial_ubar);
}
void launch_sweep(const KCfg &k, const Bufs &b, hipStream_t s, int serial_ubar) {
  hipLaunchKernelGGL(k_sweep, dim3(1), dim3(SW_T), 0, s, k, b, serial_ubar);
}
void launch_alpha_all(const KCfg &k, const Bufs &b, hipStream_t s, int slot_out) {
  hipLaunchKernelGGL(k_alpha_all, dim3((k.nx + 6 + 63) / 64, (k.ny + 6 + 3) / 4), dim3(64, 4), 0, s, k, b, slot_out);
}
void launch_bc(const KCfg &k, const Bufs &b, hipStream_t s, int slot_out, int parity) {
  const long long total = k.bc == BC_STEP ? step_bc_cells(k) : 2LL * 3 * (k.nx + 6) + 2LL * 3 * k.ny;
  hipLaunchKernelGGL(k_bc, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, k, b, slot_out, parity);
}
void launch_line_alpha(const KCfg &k, const Bufs &b, hipStream_t s, int slot, int parity) {
  hipLaunchKernelGGL(k_line_alpha, dim3(la_gx(k), la_gy(k)), dim3(256), 0, s, k, b, slot, parity);
  hipLaunchKernelGGL(k_line_alpha_reduce, dim3(la_reduce_blocks(k)), dim3(256), 0, s, k, b, slot, parity);
}
void launch_step_reflect(const KCfg &k, const Bufs &b, hipStream_t s, int dir, int parity) {
  const int n = step_reflect_cells(k, dir);
  hipLaunchKernelGGL(k_step_reflect, dim3((n + 255) / 256), dim3(256), 0, s, k, b, dir, parity);
}
void launch_line_alpha_batch(const BatchInst *inst, const int2 *lmap, int nl, const int2 *rmap, int nr,
                             hipStream_t s, int stage, int slot, int parity) {
  if (nl) hipLaunchKernelGGL(k_line_alpha_batch, dim3(nl), dim3(256), 0, s, inst, lmap, stage, slot, parity);
  if (nr) hipLaunchKernelGGL(k_line_alpha_reduce_batch, dim3(nr), dim3(256), 0, s, inst, rmap, stage, slot, parity);
}
void launch_step_reflect_batch(const BatchInst *inst, const int2 *map, int n, hipStream_t s, int stage, int dir,
                               int parity) {
  if (n) hipLaunchKernelGGL(k_step_reflect_batch, dim3(n), dim3(256), 0, s, inst, map, stage, dir, parity);
}
void launch_from_fortran(const KCfg &k, const double *src, double *dst, hipStream_t s) {
  hipLaunchKernelGGL(k_from_fortran, dim3(1024), dim3(256), 0, s, k, src, dst);
}
void launch_to_fortran(const KCfg &k, const double *src, double *dst, hipStream_t s) {
  hipLaunchKernelGGL(k_to_fortran, dim3(1024), dim3(256), 0, s, k, src, dst);
}

// the detonation program: one RK stage in the reference's order (rk :810-888)
void launch7_stage(const ex7::K7 &k, const ex7::B7 &b, hipStream_t s, int stage) {
  using namespace ex7;
  const dim3 frame((k.nx + 6 + 63) / 64, (k.ny + 6 + 3) / 4), cells((k.nx + 63) / 64, (k.ny + 3) / 4);
  const dim3 blk(64, 4);
  // wenofd: findalpha and the line alphas, the x walls and pass, the y walls and pass
  hipLaunchKernelGGL(k7_alpha, frame, blk, 0, s, k, b);
  hipLaunchKernelGGL(k7_lines, dim3((k.nx + k.ny + 12 + 3) / 4), dim3(256), 0, s, k, b);
  if (stage == 1) hipLaunchKernelGGL(k7_dt, dim3(1), dim3(64), 0, s, k, b);
  const int nw0 = 3 * ((k.ny + 1 < k.step * k.total ? k.ny + 1 : k.step * k.total) + 1);
  hipLaunchKernelGGL(k7_walls, dim3((nw0 + 255) / 256), dim3(256), 0, s, k, b, 0);
  hipLaunchKernelGGL(k7_flux<0>, dim3(((k.nx + 2) * (k.ny + 2) + 127) / 128), dim3(128), 0, s, k, b);
  hipLaunchKernelGGL(k7_walls, dim3((3 * k.total + 255) / 256), dim3(256), 0, s, k, b, 1);
  hipLaunchKernelGGL(k7_flux<1>, dim3(((k.nx + 2) * (k.ny + 2) + 127) / 128), dim3(128), 0, s, k, b);
  // scheme, RK, checkpositivity; the sweep when flagged; bc
  hipLaunchKernelGGL(k7_clear, dim3(1), dim3(1), 0, s, b);
  hipLaunchKernelGGL(k7_stage, cells, blk, 0, s, k, b, stage);
  hipLaunchKernelGGL(k7_sweep, dim3(1), dim3(SW7_T), 0, s, k, b);
  hipLaunchKernelGGL(k7_bc, frame, blk, 0, s, k, b, 0);
}
void launch7_bc(const ex7::K7 &k, const ex7::B7 &b, hipStream_t s) {
  hipLaunchKernelGGL(ex7::k7_bc, dim3((k.nx + 6 + 63) / 64, (k.ny + 6 + 3) / 4), dim3(64, 4), 0, s, k, b, 1);
}
void launch7_sweep(const ex7::K7 &k, const ex7::B7 &b, hipStream_t s) {
  hipLaunchKernelGGL(ex7::k7_sweep, dim3(1), dim3(ex7::SW7_T), 0, s, k, b);
}
void launch7_from_fortran(const ex7::K7 &k, const double *src, double *dst, hipStream_t s) {
  hipLaunchKernelGGL(ex7::k7_from_fortran, dim3(256), dim3(256), 0, s, k, src, dst);
}
void launch7_to_fortran(const ex7::K7 &k, const double *src, double *dst, hipStream_t s) {
  hipLaunchKernelGGL(ex7::k7_to_fortran, dim3(256), dim3(256), 0, s, k, src, dst);
}

}  // namespace swf
