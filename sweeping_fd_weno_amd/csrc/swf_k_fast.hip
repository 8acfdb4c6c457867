// swf_k_fast.hip - flux and update kernels of the FP64 and FP32-bulk modes
// (compiled with floating-point contraction enabled).
#include "swf_kernels.h"
#include "swf_launch.h"

namespace swf {

void launch_flux_m12(int mode, const KCfg &k, const Bufs &b, hipStream_t s, int slot_in, int slot_reset, int parity,
                     int clear_maps) {
  if (mode == MODE_FP64)
    hipLaunchKernelGGL(k_flux<MODE_FP64>, dim3(flux_blocks(k)), dim3(256), 0, s, k, b, slot_in, slot_reset, parity,
                       clear_maps);
  else
    hipLaunchKernelGGL(k_flux<MODE_FP32>, dim3(flux_blocks(k)), dim3(256), 0, s, k, b, slot_in, slot_reset, parity,
                       clear_maps);
}
void launch_update_m12(int mode, const KCfg &k, const Bufs &b, hipStream_t s, int stage, int slot_in, int slot_out,
                       int parity, int rhs_only) {
  if (mode == MODE_FP64)
    hipLaunchKernelGGL(k_update<MODE_FP64>, grid_cells(k), dim3(64, 4), 0, s, k, b, stage, slot_in, slot_out, parity,
                       rhs_only);
  else
    hipLaunchKernelGGL(k_update<MODE_FP32>, grid_cells(k), dim3(64, 4), 0, s, k, b, stage, slot_in, slot_out, parity,
                       rhs_only);
}
void launch_weno_points_m12(int mode, const double *a, double *r, double *l, long long n, hipStream_t s) {
  const int blocks = (int)((n + 255) / 256);
  if (mode == MODE_FP64)
    hipLaunchKernelGGL(k_weno_points<MODE_FP64>, dim3(blocks), dim3(256), 0, s, a, r, l, n);
  else
    hipLaunchKernelGGL(k_weno_points<MODE_FP32>, dim3(blocks), dim3(256), 0, s, a, r, l, n);
}

}  // namespace swf
