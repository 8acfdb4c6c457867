// swf_k_fast.hip - stage kernels of the FP64, FP32-bulk and FP32-full modes
// (compiled with floating-point contraction enabled).
#include <type_traits>

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

hipError_t launch_run1d_m12(int mode, const K1Cfg &k, const B1 &b, hipStream_t s, long long max_steps) {
  if (mode == MODE_FP64)
    return run1d_cpt(k.N) == 2 ? run1d<MODE_FP64, 2>(k, b, s, max_steps) : run1d<MODE_FP64, 4>(k, b, s, max_steps);
  return run1d_cpt(k.N) == 2 ? run1d<MODE_FP32, 2>(k, b, s, max_steps) : run1d<MODE_FP32, 4>(k, b, s, max_steps);
}

// the flux limiter (use_limiter = 1) has kernels of its own, so that the others
// carry none of its registers
// so has the step program's wenofd (alpha_lines: per-line alphas, 1d-12 in the weights)
template <int M, bool LIM>
static void xflux(const KCfg &k, const Bufs &b, hipStream_t s, int stage, int slot_in, int parity) {
  if (!LIM && k.alpha_lines)
    hipLaunchKernelGGL((k_xflux<M, false, true>), grid_xflux(k), dim3(256), 0, s, k, b, stage, slot_in, parity);
  else
    hipLaunchKernelGGL((k_xflux<M, LIM>), grid_xflux(k), dim3(256), 0, s, k, b, stage, slot_in, parity);
}
template <int M, bool LIM>
static void ystage(const KCfg &k, const Bufs &b, hipStream_t s, int stage, int slot_in, int parity, int flags) {
  if (!LIM && k.alpha_lines)
    hipLaunchKernelGGL((k_ystage<M, false, true>), grid_ystage(k), dim3(YS_T), 0, s, k, b, stage, slot_in, parity,
                       flags);
  else
    hipLaunchKernelGGL((k_ystage<M, LIM>), grid_ystage(k), dim3(YS_T), 0, s, k, b, stage, slot_in, parity, flags);
}

// f<M>() for the fast mode `mode` (1, 2 or 3)
template <class F>
static void by_mode(int mode, F f) {
  if (mode == MODE_FP64)
    f(std::integral_constant<int, MODE_FP64>());
  else if (mode == MODE_FP32)
    f(std::integral_constant<int, MODE_FP32>());
  else
    f(std::integral_constant<int, MODE_FP32F>());
}

void launch_xflux_m12(int mode, const KCfg &k, const Bufs &b, hipStream_t s, int stage, int slot_in, int parity) {
  const bool lim = k.use_limiter == 1;
  by_mode(mode, [&](auto m) {
    constexpr int M = decltype(m)::value;
    (lim ? xflux<M, true> : xflux<M, false>)(k, b, s, stage, slot_in, parity);
  });
}
void launch_ystage_m12(int mode, const KCfg &k, const Bufs &b, hipStream_t s, int stage, int slot_in, int parity,
                       int flags) {
  const bool lim = k.use_limiter == 1;
  by_mode(mode, [&](auto m) {
    constexpr int M = decltype(m)::value;
    (lim ? ystage<M, true> : ystage<M, false>)(k, b, s, stage, slot_in, parity, flags);
  });
}
void launch_zs_edges_m12(int mode, const KCfg &k, const Bufs &b, hipStream_t s, int slot_in, int parity) {
  by_mode(mode, [&](auto m) {
    constexpr int M = decltype(m)::value;
    hipLaunchKernelGGL(k_zs_edges<M>, dim3(grid_zs_edges(k)), dim3(256), 0, s, k, b, slot_in, parity);
  });
}
template <int M, bool STEP>
static void xflux_batch(const BatchInst *inst, const int2 *map, int nblocks, hipStream_t s, int stage, int slot_in,
                        int parity) {
  hipLaunchKernelGGL((k_xflux_batch<M, STEP>), dim3(nblocks), dim3(256), 0, s, inst, map, stage, slot_in, parity);
}
template <int M, bool STEP>
static void ystage_batch(const BatchInst *inst, const int2 *map, int nblocks, hipStream_t s, int stage, int slot_in,
                         int parity, int flags) {
  hipLaunchKernelGGL((k_ystage_batch<M, STEP>), dim3(nblocks), dim3(YS_T), 0, s, inst, map, stage, slot_in, parity,
                     flags);
}
void launch_xflux_batch_m12(int mode, const BatchInst *inst, const int2 *map, int nblocks, hipStream_t s, int stage,
                            int slot_in, int parity, int step) {
  by_mode(mode, [&](auto m) {
    constexpr int M = decltype(m)::value;
    (step ? xflux_batch<M, true> : xflux_batch<M, false>)(inst, map, nblocks, s, stage, slot_in, parity);
  });
}
void launch_ystage_batch_m12(int mode, const BatchInst *inst, const int2 *map, int nblocks, hipStream_t s, int stage,
                             int slot_in, int parity, int flags, int step) {
  by_mode(mode, [&](auto m) {
    constexpr int M = decltype(m)::value;
    (step ? ystage_batch<M, true> : ystage_batch<M, false>)(inst, map, nblocks, s, stage, slot_in, parity, flags);
  });
}
void launch_weno_points_m12(int mode, const double *a, double *r, double *l, long long n, int e12, hipStream_t s) {
  const int blocks = (int)((n + 255) / 256);
  by_mode(mode, [&](auto m) {
    constexpr int M = decltype(m)::value;
    hipLaunchKernelGGL(k_weno_points<M>, dim3(blocks), dim3(256), 0, s, a, r, l, n, e12);
  });
}

}  // namespace swf
