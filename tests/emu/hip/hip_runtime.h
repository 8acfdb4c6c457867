// CPU emulation of the HIP runtime subset that libswf uses (test infrastructure:
// tests/emu/Makefile builds the library's own sources against this header with
// g++).  Kernels run block by block on one OS thread; each thread of a block is
// a fiber (ucontext) that runs until its next __syncthreads or its end, so the
// threads of a block interleave deterministically.  __shfl_xor, __shfl_up and
// __ballot exchange through a buffer between two wave-level synchronisations
// (every thread of a wave reaches them together); __syncthreads waits for every
// live thread of the block; atomics are plain read-modify-writes; streams are
// synchronous; a captured graph records the launches and replays them; dynamic
// LDS (HIP_DYNAMIC_SHARED) is one buffer of the launch's size, filled with 0xA5.
// Device memory is host memory filled with 0xA5 bytes at allocation.
#pragma once
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <functional>
#include <vector>

#define __host__
#define __device__
#define __global__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static
#define __align__(n) alignas(n)
#define HIP_SYMBOL(x) (&(x))

struct int2 {
  int x, y;
};
inline int2 make_int2(int x, int y) { return int2{x, y}; }

struct dim3 {
  unsigned x, y, z;
  dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};

namespace emu {
extern dim3 tIdx, bIdx, bDim, gDim;
void barrier();
void wave_sync();
int lin_tid();
extern unsigned char *dyn_lds;
void set_dynamic_lds(size_t bytes);
extern unsigned long long shfl_buf[1024];
void launch(dim3 g, dim3 b, std::function<void()> body);
}  // namespace emu
#define threadIdx (emu::tIdx)
#define blockIdx (emu::bIdx)
#define blockDim (emu::bDim)
#define gridDim (emu::gDim)

inline void __syncthreads() { emu::barrier(); }
#define HIP_DYNAMIC_SHARED(type, var) type *var = (type *)emu::dyn_lds;

inline unsigned long long __ballot(int pred) {
  const int t = emu::lin_tid();
  emu::shfl_buf[t] = pred != 0;
  emu::wave_sync();
  unsigned long long m = 0;
  const int w0 = t & ~63;
  for (int l = 0; l < 64 && w0 + l < (int)(emu::bDim.x * emu::bDim.y * emu::bDim.z); ++l)
    if (emu::shfl_buf[w0 + l]) m |= 1ull << l;
  emu::wave_sync();
  return m;
}

template <class T>
T __shfl_xor(T v, int off, int width = 64) {
  static_assert(sizeof(T) <= 8, "shuffle of at most 8 bytes");
  const int t = emu::lin_tid();
  unsigned long long w = 0;
  memcpy(&w, &v, sizeof(T));
  emu::shfl_buf[t] = w;
  emu::wave_sync();
  const int src = (t & ~(width - 1)) | ((t & (width - 1)) ^ off);
  unsigned long long r = emu::shfl_buf[src];
  emu::wave_sync();
  T out;
  memcpy(&out, &r, sizeof(T));
  return out;
}

template <class T>
T __shfl(T v, int src, int width = 64) {
  static_assert(sizeof(T) <= 8, "shuffle of at most 8 bytes");
  const int t = emu::lin_tid();
  unsigned long long w = 0;
  memcpy(&w, &v, sizeof(T));
  emu::shfl_buf[t] = w;
  emu::wave_sync();
  unsigned long long r = emu::shfl_buf[(t & ~(width - 1)) | (src & (width - 1))];
  emu::wave_sync();
  T out;
  memcpy(&out, &r, sizeof(T));
  return out;
}

inline int min(int a, int b) { return a < b ? a : b; }
inline int max(int a, int b) { return a > b ? a : b; }
inline long long min(long long a, long long b) { return a < b ? a : b; }
inline long long max(long long a, long long b) { return a > b ? a : b; }
template <class T>
T __shfl_up(T v, unsigned d, int width = 64) {
  static_assert(sizeof(T) <= 8, "shuffle of at most 8 bytes");
  const int t = emu::lin_tid();
  unsigned long long w = 0;
  memcpy(&w, &v, sizeof(T));
  emu::shfl_buf[t] = w;
  emu::wave_sync();
  const int lane = t & (width - 1);
  const int src = lane >= (int)d ? t - (int)d : t;
  unsigned long long r = emu::shfl_buf[src];
  emu::wave_sync();
  T out;
  memcpy(&out, &r, sizeof(T));
  return out;
}

template <class T>
T __shfl_down(T v, unsigned d, int width = 64) {
  static_assert(sizeof(T) <= 8, "shuffle of at most 8 bytes");
  const int t = emu::lin_tid();
  unsigned long long w = 0;
  memcpy(&w, &v, sizeof(T));
  emu::shfl_buf[t] = w;
  emu::wave_sync();
  const int lane = t & (width - 1);
  const int nthr = (int)(emu::bDim.x * emu::bDim.y * emu::bDim.z);
  const int src = (lane + (int)d < width && t + (int)d < nthr) ? t + (int)d : t;
  unsigned long long r = emu::shfl_buf[src];
  emu::wave_sync();
  T out;
  memcpy(&out, &r, sizeof(T));
  return out;
}

// DPP move of an int (the controls this build uses: row_shr:1-15, row_bcast:15,
// row_bcast:31); rows outside row_mask, and lanes whose source lies outside the
// row (shift) or before row 1 / row 2 (broadcasts), take `old` (bound_ctrl off)
inline int __builtin_amdgcn_update_dpp(int old, int src, int ctrl, int row_mask, int bank_mask,
                                       bool bound_ctrl) {
  (void)bank_mask;
  const int t = emu::lin_tid();
  emu::shfl_buf[t] = (unsigned long long)(unsigned)src;
  emu::wave_sync();
  const int lane = t & 63, row = lane >> 4, w0 = t & ~63;
  int s = -1;
  if (ctrl > 0x110 && ctrl < 0x120) {
    const int n = ctrl - 0x110;
    if ((lane & 15) >= n) s = lane - n;
  } else if (ctrl == 0x142) {
    if (row >= 1) s = row * 16 - 1;
  } else if (ctrl == 0x143) {
    if (row >= 2) s = 31;
  } else {
    abort();
  }
  int r = old;
  if ((row_mask >> row) & 1) {
    if (s >= 0)
      r = (int)(unsigned)emu::shfl_buf[w0 + s];
    else if (bound_ctrl)
      r = 0;
  }
  emu::wave_sync();
  return r;
}

// the value of the wave's first lane (every lane active in this build's uses)
inline int __builtin_amdgcn_readfirstlane(int v) { return __shfl(v, 0); }
inline int __ffs(unsigned x) { return __builtin_ffs((int)x); }
inline int __popc(unsigned x) { return __builtin_popcount(x); }
inline long long __double_as_longlong(double d) { long long r; memcpy(&r, &d, 8); return r; }
inline double __longlong_as_double(long long l) { double r; memcpy(&r, &l, 8); return r; }

// single OS thread: atomics are plain read-modify-writes
template <class T> T atomicOr(T *p, T v) { T o = *p; *p = o | v; return o; }
template <class T> T atomicAnd(T *p, T v) { T o = *p; *p = o & v; return o; }
template <class T> T atomicAdd(T *p, T v) { T o = *p; *p = o + v; return o; }
template <class T> T atomicMax(T *p, T v) { T o = *p; if (v > o) *p = v; return o; }

typedef int hipError_t;
enum { hipSuccess = 0, hipErrorOutOfMemory = 2, hipErrorInvalidValue = 1 };
typedef struct emu_stream *hipStream_t;
typedef struct emu_graph *hipGraph_t;
typedef struct emu_graph *hipGraphExec_t;
typedef struct emu_event *hipEvent_t;
typedef struct emu_node *hipGraphNode_t;
enum hipMemcpyKind { hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice, hipMemcpyHostToHost };
enum { hipStreamNonBlocking = 1 };
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal, hipStreamCaptureModeThreadLocal };
enum hipDeviceAttribute_t { hipDeviceAttributeMaxSharedMemoryPerBlock = 1 };
enum hipFuncAttribute { hipFuncAttributeMaxDynamicSharedMemorySize = 1 };

hipError_t hipSetDevice(int);
hipError_t hipDeviceGetAttribute(int *, hipDeviceAttribute_t, int);
hipError_t hipFuncSetAttribute(const void *, hipFuncAttribute, int);
hipError_t hipGetDeviceCount(int *);
hipError_t hipStreamCreateWithFlags(hipStream_t *, unsigned);
hipError_t hipStreamDestroy(hipStream_t);
hipError_t hipStreamSynchronize(hipStream_t);
hipError_t hipMalloc(void *, size_t);
template <class T> hipError_t hipMalloc(T **p, size_t n) { return hipMalloc((void *)p, n); }
hipError_t hipFree(void *);
hipError_t hipMemcpy(void *, const void *, size_t, hipMemcpyKind);
hipError_t hipMemcpyAsync(void *, const void *, size_t, hipMemcpyKind, hipStream_t = nullptr);
// a __device__ variable is host memory here: its address is the symbol's
inline hipError_t hipGetSymbolAddress(void **p, const void *symbol) {
  *p = const_cast<void *>(symbol);
  return hipSuccess;
}
hipError_t hipMemsetAsync(void *, int, size_t, hipStream_t = nullptr);
hipError_t hipGetLastError();
const char *hipGetErrorString(hipError_t);
hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode);
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t *);
hipError_t hipGraphInstantiate(hipGraphExec_t *, hipGraph_t, hipGraphNode_t *, char *, size_t);
hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t);
hipError_t hipGraphExecDestroy(hipGraphExec_t);
hipError_t hipGraphDestroy(hipGraph_t);
hipError_t hipEventCreate(hipEvent_t *);
hipError_t hipEventDestroy(hipEvent_t);
hipError_t hipEventRecord(hipEvent_t, hipStream_t = nullptr);
hipError_t hipEventSynchronize(hipEvent_t);
hipError_t hipEventElapsedTime(float *, hipEvent_t, hipEvent_t);

namespace emu {
void trace_launch(const void *fn, dim3 g, dim3 b);
template <class K, class... A>
void launch_k(K k, dim3 g, dim3 b, size_t lds, hipStream_t, A... a) {
  trace_launch((const void *)k, g, b);
  launch(g, b, [=]() {
    if (threadIdx.x == 0 && threadIdx.y == 0 && threadIdx.z == 0) set_dynamic_lds(lds);
    k(a...);
  });
}
}  // namespace emu
#define hipLaunchKernelGGL(K, ...) emu::launch_k(K, __VA_ARGS__)
