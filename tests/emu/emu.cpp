// Runtime of the CPU emulation declared in hip/hip_runtime.h: the fiber
// scheduler of a block, launches (recorded while a stream captures), and the
// memory, stream, graph and event calls.  EMU_TRACE=1 prints every launch.
#include "hip/hip_runtime.h"
#include <stdio.h>
#include <stdlib.h>
#include <ucontext.h>
#include <chrono>
#include <dlfcn.h>
#include <cxxabi.h>

namespace emu {
dim3 tIdx, bIdx, bDim, gDim;
unsigned long long shfl_buf[1024];

namespace {
constexpr size_t kStack = 256 * 1024;
struct Fiber {
  ucontext_t ctx;
  char *stack = nullptr;
  int state = 0;  // 0 runnable, 1 at barrier, 2 done
  dim3 tid;
};
std::vector<Fiber> fibers;
ucontext_t sched;
int cur = -1;
std::function<void()> *body_ptr = nullptr;
bool capturing = false;
std::vector<std::function<void()>> *capture_list = nullptr;

void trampoline() {
  (*body_ptr)();
  fibers[cur].state = 2;
  swapcontext(&fibers[cur].ctx, &sched);
}

void run_block(std::function<void()> &body) {
  const int n = bDim.x * bDim.y * bDim.z;
  if ((int)fibers.size() < n) fibers.resize(n);
  body_ptr = &body;
  for (int t = 0; t < n; ++t) {
    Fiber &f = fibers[t];
    if (!f.stack) f.stack = (char *)malloc(kStack);
    getcontext(&f.ctx);
    f.ctx.uc_stack.ss_sp = f.stack;
    f.ctx.uc_stack.ss_size = kStack;
    f.ctx.uc_link = nullptr;
    makecontext(&f.ctx, trampoline, 0);
    f.state = 0;
    f.tid = dim3(t % bDim.x, (t / bDim.x) % bDim.y, t / (bDim.x * bDim.y));
  }
  // run every runnable fiber to its next yield; then release the waves whose
  // live fibers all wait at a wave-level exchange, else (all live fibers at
  // __syncthreads) the block barrier
  const int nw = (n + 63) / 64;
  while (true) {
    for (int t = 0; t < n; ++t) {
      Fiber &f = fibers[t];
      if (f.state != 0) continue;
      cur = t;
      tIdx = f.tid;
      swapcontext(&sched, &f.ctx);
    }
    int live = 0, at_block = 0, at_wave = 0;
    for (int t = 0; t < n; ++t) {
      live += fibers[t].state != 2;
      at_block += fibers[t].state == 1;
      at_wave += fibers[t].state == 3;
    }
    if (live == 0) break;
    bool released = false;
    for (int w = 0; w < nw && at_wave; ++w) {
      int wl = 0, ww = 0;
      for (int t = w * 64; t < n && t < (w + 1) * 64; ++t) {
        wl += fibers[t].state != 2;
        ww += fibers[t].state == 3;
      }
      if (ww > 0 && ww == wl) {
        for (int t = w * 64; t < n && t < (w + 1) * 64; ++t)
          if (fibers[t].state == 3) fibers[t].state = 0;
        released = true;
      }
    }
    if (released) continue;
    if (at_block == live) {
      for (int t = 0; t < n; ++t)
        if (fibers[t].state == 1) fibers[t].state = 0;
      continue;
    }
    fprintf(stderr, "emu: deadlock in block (%u,%u): %d live, %d at __syncthreads, %d at wave exchanges\n",
            bIdx.x, bIdx.y, live, at_block, at_wave);
    abort();
  }
  cur = -1;
}
}  // namespace

int lin_tid() { return cur; }

static int trace_level = -1;
void trace_launch(const void *fn, dim3 g, dim3 b) {
  if (trace_level < 0) trace_level = getenv("EMU_TRACE") ? atoi(getenv("EMU_TRACE")) : 0;
  if (!trace_level || capturing) return;
  Dl_info info;
  const char *name = "?";
  char *dem = nullptr;
  if (dladdr(fn, &info) && info.dli_sname) {
    int st = 0;
    dem = abi::__cxa_demangle(info.dli_sname, nullptr, nullptr, &st);
    name = dem ? dem : info.dli_sname;
  }
  fprintf(stderr, "launch %s grid %u,%u block %u\n", name, g.x, g.y, b.x * b.y * b.z);
  free(dem);
}

void barrier() {
  if (cur < 0) return;
  fibers[cur].state = 1;
  swapcontext(&fibers[cur].ctx, &sched);
}

void wave_sync() {
  if (cur < 0) return;
  fibers[cur].state = 3;
  swapcontext(&fibers[cur].ctx, &sched);
}

unsigned char *dyn_lds = nullptr;
static std::vector<unsigned char> dyn_lds_buf;
void set_dynamic_lds(size_t bytes) {
  dyn_lds_buf.assign(bytes ? bytes : 8, 0xA5);  // like uninitialised LDS
  dyn_lds = dyn_lds_buf.data();
}

void launch(dim3 g, dim3 b, std::function<void()> body) {
  if (capturing) {
    capture_list->push_back([=]() { launch(g, b, body); });
    return;
  }
  gDim = g;
  bDim = b;
  for (unsigned z = 0; z < g.z; ++z)
    for (unsigned y = 0; y < g.y; ++y)
      for (unsigned x = 0; x < g.x; ++x) {
        bIdx = dim3(x, y, z);
        std::function<void()> bb = body;
        run_block(bb);
      }
}
}  // namespace emu

struct emu_graph {
  std::vector<std::function<void()>> ops;
};
struct emu_event {
  std::chrono::steady_clock::time_point t;
};

hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipDeviceGetAttribute(int *v, hipDeviceAttribute_t a, int) {
  *v = a == hipDeviceAttributeMaxSharedMemoryPerBlock ? 160 * 1024 : 0;
  return hipSuccess;
}
hipError_t hipFuncSetAttribute(const void *, hipFuncAttribute, int) { return hipSuccess; }
hipError_t hipGetDeviceCount(int *n) { *n = 1; return hipSuccess; }
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned) { *s = (hipStream_t)0x1; return hipSuccess; }
hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipMalloc(void *pp, size_t n) {
  void *p = aligned_alloc(256, (n + 255) / 256 * 256);
  if (!p) return hipErrorOutOfMemory;
  memset(p, 0xA5, n);  // like uninitialised device memory
  *(void **)pp = p;
  return hipSuccess;
}
hipError_t hipFree(void *p) { free(p); return hipSuccess; }
hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind) { memcpy(d, s, n); return hipSuccess; }
hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind, hipStream_t) {
  if (emu::capturing) { fprintf(stderr, "emu: memcpy during capture\n"); abort(); }
  memcpy(d, s, n);
  return hipSuccess;
}
hipError_t hipMemsetAsync(void *d, int v, size_t n, hipStream_t) {
  if (emu::capturing) { fprintf(stderr, "emu: memset during capture\n"); abort(); }
  memset(d, v, n);
  return hipSuccess;
}
hipError_t hipGetLastError() { return hipSuccess; }
const char *hipGetErrorString(hipError_t) { return "emulated HIP error"; }
hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) {
  emu::capturing = true;
  emu::capture_list = new std::vector<std::function<void()>>();
  return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t *g) {
  emu::capturing = false;
  *g = new emu_graph{std::move(*emu::capture_list)};
  delete emu::capture_list;
  emu::capture_list = nullptr;
  return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, hipGraphNode_t *, char *, size_t) {
  *e = new emu_graph(*g);
  return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t) {
  for (auto &op : e->ops) op();
  return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t e) { delete e; return hipSuccess; }
hipError_t hipGraphDestroy(hipGraph_t g) { delete g; return hipSuccess; }
hipError_t hipEventCreate(hipEvent_t *e) { *e = new emu_event; return hipSuccess; }
hipError_t hipEventDestroy(hipEvent_t e) { delete e; return hipSuccess; }
hipError_t hipEventRecord(hipEvent_t e, hipStream_t) { e->t = std::chrono::steady_clock::now(); return hipSuccess; }
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b) {
  *ms = std::chrono::duration<float, std::milli>(b->t - a->t).count();
  return hipSuccess;
}
