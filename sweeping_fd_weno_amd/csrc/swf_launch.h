// swf_launch.h - host-side launchers of the device kernels (internal).
#pragma once

#include <hip/hip_runtime.h>

#include "swf_1d.h"
#include "swf_device.h"
#include "swf_ex7.h"

namespace swf {

// stage kernels of one precision mode (mode 0 in swf_k_exact.hip, modes 1 and 2
// in swf_k_fast.hip)
void launch_xflux(int mode, const KCfg &k, const Bufs &b, hipStream_t s, int stage, int slot_in, int parity);
void launch_ystage(int mode, const KCfg &k, const Bufs &b, hipStream_t s, int stage, int slot_in, int parity,
                   int flags);
void launch_zs_edges(int mode, const KCfg &k, const Bufs &b, hipStream_t s, int slot_in, int parity);
void launch_weno_points(int mode, const double *a, double *r, double *l, long long n, int e12, hipStream_t s);

void launch_xflux_m0(const KCfg &k, const Bufs &b, hipStream_t s, int stage, int slot_in, int parity);
void launch_ystage_m0(const KCfg &k, const Bufs &b, hipStream_t s, int stage, int slot_in, int parity, int flags);
void launch_zs_edges_m0(const KCfg &k, const Bufs &b, hipStream_t s, int slot_in, int parity);
void launch_weno_points_m0(const double *a, double *r, double *l, long long n, int e12, hipStream_t s);
void launch_xflux_m12(int mode, const KCfg &k, const Bufs &b, hipStream_t s, int stage, int slot_in, int parity);
void launch_ystage_m12(int mode, const KCfg &k, const Bufs &b, hipStream_t s, int stage, int slot_in, int parity,
                       int flags);
void launch_zs_edges_m12(int mode, const KCfg &k, const Bufs &b, hipStream_t s, int slot_in, int parity);
void launch_weno_points_m12(int mode, const double *a, double *r, double *l, long long n, int e12, hipStream_t s);

// batch kernels: blocks of the x-flux / stage launches from the schedule maps
void launch_xflux_batch(int mode, const BatchInst *inst, const int2 *map, int nblocks, hipStream_t s, int stage,
                        int slot_in, int parity, int step = 0);
void launch_ystage_batch(int mode, const BatchInst *inst, const int2 *map, int nblocks, hipStream_t s, int stage,
                         int slot_in, int parity, int flags, int step = 0);
void launch_xflux_batch_m0(const BatchInst *inst, const int2 *map, int nblocks, hipStream_t s, int stage,
                           int slot_in, int parity, int step = 0);
void launch_ystage_batch_m0(const BatchInst *inst, const int2 *map, int nblocks, hipStream_t s, int stage,
                            int slot_in, int parity, int flags, int step = 0);
void launch_xflux_batch_m12(int mode, const BatchInst *inst, const int2 *map, int nblocks, hipStream_t s, int stage,
                            int slot_in, int parity, int step = 0);
void launch_ystage_batch_m12(int mode, const BatchInst *inst, const int2 *map, int nblocks, hipStream_t s, int stage,
                             int slot_in, int parity, int flags, int step = 0);
void launch_stage_end_batch(const BatchInst *inst, int n, hipStream_t s, int stage, int slot_out, int parity,
                            int serial_ubar);
void launch_begin_run_batch(const BatchInst *inst, int n, const double *time, const double *tfinal, hipStream_t s);
void launch_batch_status(const BatchInst *inst, int n, BatchStatus *out, hipStream_t s);

// mode-independent kernels (swf_k_exact.hip)
void launch_sweep(const KCfg &k, const Bufs &b, hipStream_t s, int serial_ubar);
void launch_stage_end(const KCfg &k, const Bufs &b, hipStream_t s, int stage, int slot_out, int parity,
                      int serial_ubar);
void launch_alpha_all(const KCfg &k, const Bufs &b, hipStream_t s, int slot_out);
void launch_bc(const KCfg &k, const Bufs &b, hipStream_t s, int slot_out, int parity);
void launch_flags(const KCfg &k, const Bufs &b, hipStream_t s);
void launch_begin_run(const Bufs &b, hipStream_t s, double time, double tfinal, int active);
void launch_set_alpha(const Bufs &b, hipStream_t s, int slot, unsigned long long bits, int p, int active);
void launch_upload_reset(const Bufs &b, hipStream_t s);
// the step program: per-line alphas of the stage input (two kernels) and the walls
void launch_line_alpha(const KCfg &k, const Bufs &b, hipStream_t s, int slot, int parity);
void launch_step_reflect(const KCfg &k, const Bufs &b, hipStream_t s, int dir, int parity);
// the same over all step-program instances of a batch, one launch each (block maps
// of (instance, block))
void launch_line_alpha_batch(const BatchInst *inst, const int2 *lmap, int nl, const int2 *rmap, int nr,
                             hipStream_t s, int stage, int slot, int parity);
void launch_step_reflect_batch(const BatchInst *inst, const int2 *map, int n, hipStream_t s, int stage, int dir,
                               int parity);
void launch_from_fortran(const KCfg &k, const double *src, double *dst, hipStream_t s);
void launch_to_fortran(const KCfg &k, const double *src, double *dst, hipStream_t s);

// the detonation program (swf_ex7.h, compiled in swf_k_exact.hip): one RK stage s
// (findalpha, walls, both passes, scheme with the source, RK, checkpositivity, the
// sweep when flagged, bc), the bc alone, the sweep alone (when postag is set), and the
// layout transposes of 5 components
void launch7_stage(const ex7::K7 &k, const ex7::B7 &b, hipStream_t s, int stage);
void launch7_bc(const ex7::K7 &k, const ex7::B7 &b, hipStream_t s);
void launch7_sweep(const ex7::K7 &k, const ex7::B7 &b, hipStream_t s);
void launch7_from_fortran(const ex7::K7 &k, const double *src, double *dst, hipStream_t s);
void launch7_to_fortran(const ex7::K7 &k, const double *src, double *dst, hipStream_t s);

// the 1D time loop (swf_1d.h): one workgroup of nt threads, cpt cells per thread
// and lds bytes of dynamic LDS; mode 0 in swf_k_exact.hip, 1 and 2 in swf_k_fast.hip
hipError_t launch_run1d(int mode, const K1Cfg &k, const B1 &b, hipStream_t s, long long max_steps);
hipError_t launch_run1d_m0(const K1Cfg &k, const B1 &b, hipStream_t s, long long max_steps);
hipError_t launch_run1d_m12(int mode, const K1Cfg &k, const B1 &b, hipStream_t s, long long max_steps);
// threads per 1D workgroup for N cells: two or four cells per thread, whole waves
inline int run1d_cpt(int N) { return N <= 2 * D1_NT ? 2 : 4; }
inline int run1d_threads(int N) {
  const int per = (N + run1d_cpt(N) - 1) / run1d_cpt(N);
  return (per + 63) / 64 * 64;
}

// x-flux tiles: 64 centres x 4 rows, XT_C interfaces per row (swf_kernels.h);
// stage tiles: UB_W x UB_H cells
inline dim3 grid_xflux(const KCfg &k) { return dim3(xf_gx(k), xf_gy(k)); }
inline unsigned grid_zs_edges(const KCfg &k) { return (unsigned)((zs_edge_halves(k) + 255) / 256); }
inline dim3 grid_ystage(const KCfg &k) { return dim3(ub_gx(k), ub_gy(k)); }
inline dim3 grid_cells(const KCfg &k) { return dim3((k.nx + 63) / 64, (k.ny + 3) / 4); }

}  // namespace swf
