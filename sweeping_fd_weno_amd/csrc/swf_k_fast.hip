// swf_k_fast.hip - stage kernels of the FP64 and FP32-bulk modes (compiled with
// floating-point contraction enabled).
#include "swf_kernels.h"
#include "swf_launch.h"

namespace swf {

void launch_xflux_m12(int mode, const KCfg &k, const Bufs &b, hipStream_t s, int stage, int slot_in, int parity) {
  if (mode == MODE_FP64)
    hipLaunchKernelGGL(k_xflux<MODE_FP64>, grid_xflux(k), dim3(256), 0, s, k, b, stage, slot_in, parity);
  else
    hipLaunchKernelGGL(k_xflux<MODE_FP32>, grid_xflux(k), dim3(256), 0, s, k, b, stage, slot_in, parity);
}
void launch_ystage_m12(int mode, const KCfg &k, const Bufs &b, hipStream_t s, int stage, int slot_in, int parity,
                       int flags) {
  if (mode == MODE_FP64)
    hipLaunchKernelGGL(k_ystage<MODE_FP64>, grid_ystage(k), dim3(256), 0, s, k, b, stage, slot_in, parity, flags);
  else
    hipLaunchKernelGGL(k_ystage<MODE_FP32>, grid_ystage(k), dim3(256), 0, s, k, b, stage, slot_in, parity, flags);
}
void launch_weno_points_m12(int mode, const double *a, double *r, double *l, long long n, hipStream_t s) {
  const int blocks = (int)((n + 255) / 256);
  if (mode == MODE_FP64)
    hipLaunchKernelGGL(k_weno_points<MODE_FP64>, dim3(blocks), dim3(256), 0, s, a, r, l, n);
  else
    hipLaunchKernelGGL(k_weno_points<MODE_FP32>, dim3(blocks), dim3(256), 0, s, a, r, l, n);
}

}  // namespace swf
