// swf_api.hip - the C-ABI of libswf.so (include/swf.h): solver contexts, state
// transfer, stage sequencing and statistics.  Each RK stage is a fixed sequence
// of kernel launches on the context's stream; the time-step, sweep and step
// gates live on the device, so steps are enqueued without host synchronisation.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/swf.h"
#include "swf_launch.h"

using namespace swf;

namespace {

thread_local char g_last_error[256] = "";

int hip_fail(hipError_t e, const char *what) {
  snprintf(g_last_error, sizeof(g_last_error), "%s: %s", what, hipGetErrorString(e));
  return SWF_E_HIP;
}

#define HIPC(call)                                   \
  do {                                               \
    hipError_t e_ = (call);                          \
    if (e_ != hipSuccess) return hip_fail(e_, #call); \
  } while (0)

constexpr int kTrackCap = 1 << 18;
constexpr int kGraphSteps = 16;   // steps per captured graph (even: parity restarts at 0)
constexpr int kCheckEvery = 256;  // steps between host checks of the device time gate

}  // namespace

struct swf_ctx {
  swf_config cfg;
  KCfg k;
  Bufs b;              // outside a stage: Uin == U == U0 == S0
  double *S1, *S2;     // the other two state buffers of the stage cycle
  int mode;
  hipStream_t stream;
  double *d_fortran;   // staging buffer in the Fortran layout
  size_t n_fortran;    // doubles in u(4,-2:nx+3,-2:ny+3)
  long long steps_total;
  hipGraphExec_t gexec;  // kGraphSteps steps, captured on first use
};

// The kernels' precision mode for a configuration.  SWF_FP32_FULL keeps the
// FP32-bulk form (FP64 projections and candidates) for the jets, whose Mach 2000
// makes the characteristic projections ill-conditioned (condition ~ Mach^2), and for
// the step program, chaotic on its coarse grids (FP32 projections: 2.3e-6 relative
// L1 density after 10 steps at 104 x 88, against 1e-6).
static int kernel_mode(int precision, int bc) {
  if (precision == SWF_FP32_FULL && (bc == SWF_BC_JET_HALF || bc == SWF_BC_JET_FULL || bc == SWF_BC_STEP))
    return MODE_FP32;
  return precision;
}

static int check_launch(const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, what);
  return SWF_OK;
}

static int set_device(const swf_ctx *c) {
  HIPC(hipSetDevice(c->cfg.device));
  return SWF_OK;
}

extern "C" {

const char *swf_last_error(void) { return g_last_error; }

const char *swf_strerror(int code) {
  switch (code) {
    case SWF_OK: return "ok";
    case SWF_E_PBAR: return "ERROR: P(u_bar) <= 0.";
    case SWF_E_SWEEP_FAIL: return "SWEEP FAIL: swp_num reached";
    case SWF_E_PSUM: return "Sweep check failed: total pressure increased after adjustment";
    case SWF_E_ORDER2_EDGE: return "the detonation sweep's order-II pass reads its order array at position 0";
    case SWF_E_INVALID: return "invalid argument";
    case SWF_E_HIP: return "HIP runtime error";
    case SWF_E_NOMEM: return "out of device memory";
    case SWF_E_NODEVICE: return "no HIP device";
    default: return "unknown error";
  }
}

int swf_time_continue(double time, double tfinal) { return time_continue_q(time, tfinal); }

int swf_create(const swf_config *cfg, swf_ctx **out) {
  if (!cfg || !out) return SWF_E_INVALID;
  *out = nullptr;
  if (cfg->nx < 4 || cfg->ny < 4 || !(cfg->dx > 0) || !(cfg->dy > 0) || cfg->precision < 0 ||
      cfg->precision > SWF_FP32_FULL || cfg->bc < 0 || cfg->bc > 4 || cfg->use_limiter < 0 || cfg->use_limiter > 2 ||
      (cfg->use_limiter == 2 && cfg->swp_num < 1))
    return SWF_E_INVALID;
  // the step program: its bc and walls need the step inside the grid; the per-line
  // alphas are not combined with the flux limiter's edge kernels
  if ((cfg->bc == SWF_BC_STEP) != (cfg->step_total > 0) ||
      (cfg->step_total > 0 && (cfg->step_total < 2 || cfg->nx < cfg->step_total + 3 ||
                               cfg->ny < 6 * cfg->step_total + 3)) ||
      (cfg->alpha_lines && cfg->use_limiter == 1) || (cfg->alpha_lines != 0) != (cfg->weno_eps12 != 0)) {
    snprintf(g_last_error, sizeof(g_last_error),
             "swf_create: step_total > 0 exactly with bc SWF_BC_STEP, 2 <= step_total, nx >= step_total + 3, "
             "ny >= 6*step_total + 3; alpha_lines and weno_eps12 together (the step program's wenofd), not with "
             "use_limiter 1");
    return SWF_E_INVALID;
  }
  if ((long long)cfg->nx * cfg->ny >= (1LL << 31) - 64) return SWF_E_INVALID;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= cfg->device || cfg->device < 0) {
    snprintf(g_last_error, sizeof(g_last_error), "no HIP device %d (found %d)", cfg->device, ndev);
    return SWF_E_NODEVICE;
  }
  swf_ctx *c = new (std::nothrow) swf_ctx();
  if (!c) return SWF_E_NOMEM;
  c->cfg = *cfg;
  c->mode = kernel_mode(cfg->precision, cfg->bc);
  KCfg &k = c->k;
  k.fast_div = c->mode != MODE_EXACT;
  k.nx = cfg->nx;
  k.ny = cfg->ny;
  k.P = row_pitch(cfg->nx);
  k.plane = (long long)(cfg->ny + 6) * k.P;
  k.nwords = (int)(((long long)cfg->nx * cfg->ny + 31) / 32);
  k.dx = cfg->dx;
  k.dy = cfg->dy;
  k.idx = 1. / cfg->dx;
  k.idy = 1. / cfg->dy;
  k.ig1 = 1. / (cfg->gamma - 1.);
  k.gamma = cfg->gamma;
  k.gamma1 = cfg->gamma - 1.;
  k.cfl = cfg->cfl;
  k.eps = cfg->eps;
  k.eps2 = cfg->eps2;
  k.delta = cfg->delta;
  k.weight = 1. / ((double)cfg->nx * (double)cfg->ny);
  // ex3: dt = dx**(5./3.)*cfl/alphamax, the numerator being a run constant
  k.dt_num = pow(cfg->dx, 5. / 3.) * cfg->cfl;
  k.xlo = cfg->xlo;
  k.ylo = cfg->ylo;
  k.jet_rho = cfg->jet_rho;
  k.jet_mx = cfg->jet_rho * cfg->jet_u;
  k.amb_rho = cfg->amb_rho;
  k.amb_e = cfg->amb_p / k.gamma1;
  // ex5_mach2000_jet/main.f90:711: 0.4127/gamma1 + 0.5*u(2)**2/u(1)
  k.jet_e = k.amb_e + (0.5 * (k.jet_mx * k.jet_mx)) / k.jet_rho;
  k.jet_halfwidth = cfg->jet_halfwidth;
  // findalpha over the prescribed ghost states (the jet's inflow and ambient
  // columns); every other ghost is an image of an interior cell with equal alpha
  k.alpha_const = 0.0;
  if (cfg->bc == SWF_BC_JET_HALF || cfg->bc == SWF_BC_JET_FULL) {
    k.findalpha_ex3 = cfg->findalpha_ex3;
    k.eps2 = cfg->eps2;
    k.gamma = cfg->gamma;
    k.gamma1 = cfg->gamma - 1.;
    k.alpha_const = fmax2(cell_alpha(k, k.jet_rho, k.jet_mx, 0.0, k.jet_e),
                          cell_alpha(k, k.amb_rho, 0.0, 0.0, k.amb_e));
  }
  k.bc = cfg->bc;
  k.step_total = cfg->step_total;
  k.alpha_lines = cfg->alpha_lines;
  k.weno_eps12 = cfg->weno_eps12;
  k.psum_check = cfg->psum_check;
  k.in_rho = cfg->in_rho;
  // ex6_shock_diffraction/main.f90:934-937: ule*dle, ple/gamma1 + 0.5*dle*ule**2
  k.in_mx = cfg->in_u * cfg->in_rho;
  k.in_e = cfg->in_p / (cfg->gamma - 1.) + (0.5 * cfg->in_rho) * (cfg->in_u * cfg->in_u);
  k.use_limiter = cfg->use_limiter;
  k.swp_num = cfg->swp_num;
  k.track_cap = kTrackCap;
  k.alpha_in_scheme = cfg->alpha_in_scheme;
  k.roe_eps_double = cfg->roe_eps_double;
  k.ydens_plain = cfg->ydens_plain;
  k.eps2_quad_literal = cfg->eps2_quad_literal;
  k.findalpha_ex3 = cfg->findalpha_ex3;
  k.dt_rule_dx53 = cfg->dt_rule_dx53;
  k.poscheck_le = cfg->poscheck_le;
  k.success_ge = cfg->success_ge;
  k.pbar_check = cfg->pbar_check;
  k.fail_stop = cfg->fail_stop;
  c->n_fortran = (size_t)4 * (cfg->nx + 6) * (cfg->ny + 6);

  int rc = SWF_OK;
  hipError_t e = hipSetDevice(cfg->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  const size_t pbytes = sizeof(double) * 4 * (size_t)k.plane;
  double *pool = nullptr;
  if (e == hipSuccess) e = hipMalloc(&pool, 5 * pbytes);
  unsigned *maps = nullptr;
  if (e == hipSuccess) e = hipMalloc(&maps, sizeof(unsigned) * 4 * (size_t)k.nwords);
  DevScalars *g = nullptr;
  if (e == hipSuccess) e = hipMalloc(&g, sizeof(DevScalars));
  long long *trk = nullptr;
  if (e == hipSuccess) e = hipMalloc(&trk, sizeof(long long) * 3 * (size_t)kTrackCap);
  double *stage = nullptr;
  if (e == hipSuccess) e = hipMalloc(&stage, sizeof(double) * c->n_fortran);
  // stage-tile alpha partials, the tiles' weight*u sums (4 each), then the flux
  // limiter's (den, pre) count pairs
  const size_t nblk = (size_t)ub_gx(k) * ub_gy(k);
  const size_t part_bytes = sizeof(double) * 5 * nblk + sizeof(int) * 2 * (size_t)zs_pairs(k);
  // then the per-line alphas and their tile partials (alpha_lines)
  const size_t nline = (size_t)(k.nx + 6) + (k.ny + 6);
  const size_t nlpart = (size_t)la_gx(k) * (k.ny + 6) + (size_t)la_gy(k) * (k.nx + 6);
  const size_t line_bytes = cfg->alpha_lines ? sizeof(double) * (nline + nlpart) : 0;
  double *apart = nullptr;
  if (e == hipSuccess) e = hipMalloc(&apart, part_bytes + line_bytes);
  if (e == hipSuccess) e = hipMemsetAsync(apart, 0, part_bytes, c->stream);
  // the sweep's walk records (use_limiter = 2)
  double *wrec = nullptr;
  const size_t rec_bytes = sizeof(double) * 4 * SW_REC * SW_T + sizeof(unsigned) * (SW_REC / 32) * SW_T;
  if (e == hipSuccess && k.use_limiter == 2) e = hipMalloc(&wrec, rec_bytes);
  if (e == hipSuccess) e = hipMemsetAsync(pool, 0, 5 * pbytes, c->stream);
  if (e == hipSuccess) e = hipMemsetAsync(maps, 0, sizeof(unsigned) * 4 * (size_t)k.nwords, c->stream);
  if (e == hipSuccess) e = hipMemsetAsync(g, 0, sizeof(DevScalars), c->stream);
  if (e == hipSuccess) e = hipMemsetAsync(trk, 0, sizeof(long long) * 3 * (size_t)kTrackCap, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) {
    rc = (e == hipErrorOutOfMemory) ? SWF_E_NOMEM : hip_fail(e, "swf_create");
    if (pool) (void)hipFree(pool);
    if (maps) (void)hipFree(maps);
    if (g) (void)hipFree(g);
    if (trk) (void)hipFree(trk);
    if (stage) (void)hipFree(stage);
    if (apart) (void)hipFree(apart);
    if (wrec) (void)hipFree(wrec);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return rc;
  }
  const size_t pd = 4 * (size_t)k.plane;
  c->b.U = pool;
  c->b.Uin = pool;
  c->b.U0 = pool;
  c->S1 = pool + pd;
  c->S2 = pool + 2 * pd;
  c->b.Fx = pool + 3 * pd;
  c->b.R = pool + 4 * pd;
  c->b.bmR = maps;
  c->b.bmP = maps + k.nwords;
  c->b.bmA = maps + 2 * (size_t)k.nwords;
  c->b.bmB = maps + 3 * (size_t)k.nwords;
  c->b.g = g;
  c->b.trk = trk;
  c->b.apart = apart;
  c->b.upart = apart + nblk;
  c->b.wrec = wrec;
  c->b.wpos = wrec ? reinterpret_cast<unsigned *>(wrec + 4 * SW_REC * SW_T) : nullptr;
  c->b.zs = reinterpret_cast<int *>(apart + 5 * nblk);
  if (cfg->alpha_lines) {
    double *lines = reinterpret_cast<double *>(reinterpret_cast<char *>(apart) + part_bytes);
    c->b.arow = lines;
    c->b.acol = lines + (k.ny + 6);
    c->b.lpart = lines + nline;
  }
  c->d_fortran = stage;
  *out = c;
  return SWF_OK;
}

void swf_destroy(swf_ctx *c) {
  if (!c) return;
  (void)hipSetDevice(c->cfg.device);
  (void)hipStreamSynchronize(c->stream);
  (void)hipFree(c->b.U);
  (void)hipFree(c->b.bmR);
  (void)hipFree(c->b.g);
  (void)hipFree(c->b.trk);
  (void)hipFree(c->d_fortran);
  (void)hipFree(c->b.apart);
  if (c->b.wrec) (void)hipFree(c->b.wrec);
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  (void)hipStreamDestroy(c->stream);
  delete c;
}

int swf_upload(swf_ctx *c, const double *u) {
  if (!c || !u) return SWF_E_INVALID;
  int rc = set_device(c);
  if (rc) return rc;
  HIPC(hipMemcpyAsync(c->d_fortran, u, sizeof(double) * c->n_fortran, hipMemcpyHostToDevice, c->stream));
  launch_from_fortran(c->k, c->d_fortran, c->b.U, c->stream);
  launch_upload_reset(c->b, c->stream);
  launch_alpha_all(c->k, c->b, c->stream, 0);
  HIPC(hipStreamSynchronize(c->stream));
  return check_launch("swf_upload");
}

int swf_download(swf_ctx *c, double *u) {
  if (!c || !u) return SWF_E_INVALID;
  int rc = set_device(c);
  if (rc) return rc;
  // the stepping kernels keep only the edge ghosts the stencils read; the full
  // ghost frame (corners included) is the reference's `bc` applied to the interior
  launch_bc(c->k, c->b, c->stream, -1, -1);
  launch_to_fortran(c->k, c->b.U, c->d_fortran, c->stream);
  HIPC(hipMemcpyAsync(u, c->d_fortran, sizeof(double) * c->n_fortran, hipMemcpyDeviceToHost, c->stream));
  HIPC(hipStreamSynchronize(c->stream));
  return check_launch("swf_download");
}

// buffers of stage s: stencils from S0, S1, S2, results into S1, S2, S0
static Bufs stage_bufs(const swf_ctx *c, int s) {
  Bufs b = c->b;
  double *S0 = c->b.U0;
  b.Uin = (s == 1) ? S0 : (s == 2) ? c->S1 : c->S2;
  b.U = (s == 1) ? c->S1 : (s == 2) ? c->S2 : S0;
  return b;
}

// one RK stage (s = 1, 2, 3) of one step with parity p: wenofd x-pass; the
// y-pass fused with scheme, the RK combination, checkpositivity and bc images;
// then one workgroup for sweep_lim (when flagged), findalpha of the new state
// and, at s = 3, the time.  With the flux limiter (use_limiter = 1) the halves
// outside the flux tiles are counted from the stage input, and the new state gets
// its full ghost frame, which those halves read at the next stage.
static void enqueue_stage(swf_ctx *c, int s, int p) {
  const int slot_in = s - 1, slot_out = s % 3;
  const Bufs b = stage_bufs(c, s);
  // the step program: findalpha's per-line maxima of the stage input, then the
  // wall images written into it before each pass
  if (c->k.alpha_lines) launch_line_alpha(c->k, b, c->stream, slot_in, p);
  if (c->k.step_total > 0) launch_step_reflect(c->k, b, c->stream, 0, p);
  launch_xflux(c->mode, c->k, b, c->stream, s, slot_in, p);
  if (c->k.step_total > 0) launch_step_reflect(c->k, b, c->stream, 1, p);
  launch_ystage(c->mode, c->k, b, c->stream, s, slot_in, p, 0);
  if (c->k.use_limiter == 1) {
    launch_zs_edges(c->mode, c->k, b, c->stream, slot_in, p);
    launch_bc(c->k, b, c->stream, -1, p);
  }
  launch_stage_end(c->k, b, c->stream, s, slot_out, p, c->mode == MODE_EXACT ? 1 : 0);
}

static void enqueue_step(swf_ctx *c, int p) {
  enqueue_stage(c, 1, p);
  enqueue_stage(c, 2, p);
  enqueue_stage(c, 3, p);
}

static int read_scalars(swf_ctx *c, DevScalars *h) {
  HIPC(hipMemcpyAsync(h, c->b.g, sizeof(DevScalars), hipMemcpyDeviceToHost, c->stream));
  HIPC(hipStreamSynchronize(c->stream));
  return SWF_OK;
}

int swf_step(swf_ctx *c, double time, double tfinal, double *dt_out) {
  if (!c) return SWF_E_INVALID;
  int rc = set_device(c);
  if (rc) return rc;
  launch_begin_run(c->b, c->stream, time, tfinal, 1);
  enqueue_step(c, 0);
  rc = check_launch("swf_step");
  if (rc) return rc;
  DevScalars h;
  rc = read_scalars(c, &h);
  if (rc) return rc;
  c->steps_total += 1;
  if (dt_out) *dt_out = h.dt;
  return h.status;
}

// SWF_NO_GRAPH=1 in the environment: every step launched on the stream, no graph
// captured or launched (for tools that do not handle graph dispatches); read once.
static bool no_graph() {
  static const bool v = [] {
    const char *e = getenv("SWF_NO_GRAPH");
    return e && e[0] == '1';
  }();
  return v;
}

static int ensure_graph(swf_ctx *c) {
  if (c->gexec || no_graph()) return SWF_OK;
  hipGraph_t graph = nullptr;
  HIPC(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  for (int q = 0; q < kGraphSteps; ++q) enqueue_step(c, q & 1);
  HIPC(hipStreamEndCapture(c->stream, &graph));
  hipError_t e = hipGraphInstantiate(&c->gexec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) {
    c->gexec = nullptr;
    return hip_fail(e, "hipGraphInstantiate");
  }
  return SWF_OK;
}

int swf_advance(swf_ctx *c, double *time_inout, double tfinal, int64_t max_steps, int64_t *steps_out) {
  if (!c || !time_inout || max_steps < 0) return SWF_E_INVALID;
  int rc = set_device(c);
  if (rc) return rc;
  rc = ensure_graph(c);
  if (rc) return rc;
  launch_begin_run(c->b, c->stream, *time_inout, tfinal, time_continue_q(*time_inout, tfinal));
  int64_t launched = 0, since_check = 0;
  DevScalars h;
  while (launched < max_steps) {
    if (max_steps - launched >= kGraphSteps && !no_graph()) {
      HIPC(hipGraphLaunch(c->gexec, c->stream));
      launched += kGraphSteps;
      since_check += kGraphSteps;
    } else {
      // remainder: plain launches continue the parity sequence
      for (; launched < max_steps; ++launched, ++since_check) enqueue_step(c, (int)(launched & 1));
    }
    rc = check_launch("swf_advance");
    if (rc) return rc;
    if (since_check >= kCheckEvery || launched >= max_steps) {
      since_check = 0;
      rc = read_scalars(c, &h);
      if (rc) return rc;
      if (h.steps_done < launched || h.status != ST_OK) break;
    }
  }
  rc = read_scalars(c, &h);
  if (rc) return rc;
  *time_inout = h.time[h.steps_done & 1];
  if (steps_out) *steps_out = h.steps_done;
  c->steps_total += h.steps_done;
  return h.status;
}

int swf_bc(swf_ctx *c) {
  if (!c) return SWF_E_INVALID;
  int rc = set_device(c);
  if (rc) return rc;
  launch_bc(c->k, c->b, c->stream, -1, -1);
  // findalpha of the next stage sees the new ghosts (a jet's inflow states among
  // them): the state's alpha is taken again over the whole frame
  launch_set_alpha(c->b, c->stream, 0, 0ull, -1, 0);
  launch_alpha_all(c->k, c->b, c->stream, 0);
  HIPC(hipStreamSynchronize(c->stream));
  return check_launch("swf_bc");
}

int swf_alpha_max(swf_ctx *c, double *alpha_out) {
  if (!c || !alpha_out) return SWF_E_INVALID;
  int rc = set_device(c);
  if (rc) return rc;
  launch_set_alpha(c->b, c->stream, 3, 0ull, -1, 0);
  launch_alpha_all(c->k, c->b, c->stream, 3);
  rc = check_launch("swf_alpha_max");
  if (rc) return rc;
  DevScalars h;
  rc = read_scalars(c, &h);
  if (rc) return rc;
  memcpy(alpha_out, &h.alpha_bits[3], sizeof(double));
  return SWF_OK;
}

int swf_rhs(swf_ctx *c, double alphamax, double *rhs_out) {
  if (!c || !rhs_out || !(alphamax > 0)) return SWF_E_INVALID;
  int rc = set_device(c);
  if (rc) return rc;
  unsigned long long bits;
  memcpy(&bits, &alphamax, sizeof(bits));
  // save the stage-1 alpha slot and the gate, use slot 0 for the query
  DevScalars h0;
  rc = read_scalars(c, &h0);
  if (rc) return rc;
  launch_set_alpha(c->b, c->stream, 0, bits, 0, 1);
  launch_xflux(c->mode, c->k, c->b, c->stream, 0, 0, 0);
  launch_ystage(c->mode, c->k, c->b, c->stream, 1, 0, 0, YS_RHS_ONLY);
  launch_to_fortran(c->k, c->b.R, c->d_fortran, c->stream);
  HIPC(hipMemcpyAsync(rhs_out, c->d_fortran, sizeof(double) * c->n_fortran, hipMemcpyDeviceToHost, c->stream));
  launch_set_alpha(c->b, c->stream, 0, h0.alpha_bits[0], 0, h0.active[0]);
  HIPC(hipStreamSynchronize(c->stream));
  return check_launch("swf_rhs");
}

int swf_check_positivity(swf_ctx *c, int32_t *postag_out) {
  if (!c || !postag_out) return SWF_E_INVALID;
  int rc = set_device(c);
  if (rc) return rc;
  launch_flags(c->k, c->b, c->stream);
  rc = check_launch("swf_check_positivity");
  if (rc) return rc;
  DevScalars h;
  rc = read_scalars(c, &h);
  if (rc) return rc;
  *postag_out = h.postag;
  return SWF_OK;
}

int swf_sweep(swf_ctx *c) {
  if (!c) return SWF_E_INVALID;
  int rc = set_device(c);
  if (rc) return rc;
  launch_flags(c->k, c->b, c->stream);
  launch_sweep(c->k, c->b, c->stream, c->mode == MODE_EXACT ? 1 : 0);
  rc = check_launch("swf_sweep");
  if (rc) return rc;
  DevScalars h;
  rc = read_scalars(c, &h);
  if (rc) return rc;
  return h.status;
}

int swf_stats(swf_ctx *c, swf_sweep_stats *s) {
  if (!c || !s) return SWF_E_INVALID;
  int rc = set_device(c);
  if (rc) return rc;
  DevScalars h;
  rc = read_scalars(c, &h);
  if (rc) return rc;
  s->sweep_count = h.sweep_count;
  s->den_count = h.den_count;
  s->pre_count = h.pre_count;
  s->total_sweep = h.total_sweep;
  s->max_sweep = h.max_sweep;
  s->total_neg = h.total_neg;
  s->max_neg = h.max_neg;
  s->total_pos = h.total_pos;
  s->max_pos = h.max_pos;
  s->stage_sweeps = h.stage_sweeps;
  s->steps = c->steps_total;
  return SWF_OK;
}

int64_t swf_track(swf_ctx *c, int64_t *q, int64_t *neg, int64_t *pos, int64_t cap) {
  if (!c) return SWF_E_INVALID;
  if (set_device(c)) return SWF_E_HIP;
  DevScalars h;
  if (read_scalars(c, &h)) return SWF_E_HIP;
  const int64_t n = h.sweep_count < cap ? h.sweep_count : cap;
  const int64_t m = n < kTrackCap ? n : kTrackCap;
  if (m > 0 && q && neg && pos) {
    if (hipMemcpy(q, c->b.trk, sizeof(int64_t) * m, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(neg, c->b.trk + kTrackCap, sizeof(int64_t) * m, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(pos, c->b.trk + 2 * kTrackCap, sizeof(int64_t) * m, hipMemcpyDeviceToHost) != hipSuccess)
      return SWF_E_HIP;
  }
  return h.sweep_count;
}

int swf_time_kernel(swf_ctx *c, int32_t which, int32_t nreps, double *avg_ms) {
  if (!c || !avg_ms || nreps < 1 || which < 0 || which > 1) return SWF_E_INVALID;
  int rc = set_device(c);
  if (rc) return rc;
  // which = 0: the x-flux kernel (reads S0 and alpha slot 0, writes Fx);
  // which = 1: the stage-1 kernel (reads S0 and Fx, writes S1, the stage flags
  // and alpha partials, which the next stage recomputes before any use)
  DevScalars h0;
  rc = read_scalars(c, &h0);
  if (rc) return rc;
  launch_set_alpha(c->b, c->stream, 0, h0.alpha_bits[0], 0, 1);
  hipEvent_t e0, e1;
  HIPC(hipEventCreate(&e0));
  HIPC(hipEventCreate(&e1));
  const Bufs b1 = stage_bufs(c, 1);
  auto launch = [&]() {
    if (which == 0)
      launch_xflux(c->mode, c->k, b1, c->stream, 0, 0, 0);
    else
      launch_ystage(c->mode, c->k, b1, c->stream, 1, 0, 0, YS_NO_SCALARS);
  };
  launch_xflux(c->mode, c->k, b1, c->stream, 0, 0, 0);  // Fx of the current state, untimed
  launch();                                              // one untimed launch
  HIPC(hipEventRecord(e0, c->stream));
  for (int r = 0; r < nreps; ++r) launch();
  HIPC(hipEventRecord(e1, c->stream));
  HIPC(hipEventSynchronize(e1));
  float ms = 0.f;
  HIPC(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  launch_set_alpha(c->b, c->stream, 0, h0.alpha_bits[0], 0, h0.active[0]);
  HIPC(hipStreamSynchronize(c->stream));
  *avg_ms = (double)ms / nreps;
  return check_launch("swf_time_kernel");
}

int swf_sync(swf_ctx *c) {
  if (!c) return SWF_E_INVALID;
  int rc = set_device(c);
  if (rc) return rc;
  HIPC(hipStreamSynchronize(c->stream));
  return SWF_OK;
}

int swf_weno_points(const double *a, double *right, double *left, int64_t n, int32_t precision, int32_t device) {
  const int e12 = (precision & SWF_WENO_EPS12) != 0;
  precision &= ~SWF_WENO_EPS12;
  if (!a || !right || !left || n < 0 || precision < 0 || precision > SWF_FP32_FULL) return SWF_E_INVALID;
  if (n == 0) return SWF_OK;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return SWF_E_NODEVICE;
  HIPC(hipSetDevice(device));
  double *d = nullptr;
  HIPC(hipMalloc(&d, sizeof(double) * 7 * (size_t)n));
  hipError_t e = hipMemcpy(d, a, sizeof(double) * 5 * (size_t)n, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    launch_weno_points(precision, d, d + 5 * n, d + 6 * n, n, e12, 0);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(right, d + 5 * n, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(left, d + 6 * n, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (e != hipSuccess) return hip_fail(e, "swf_weno_points");
  return SWF_OK;
}

}  // extern "C"

// The staging of swf_weno_points_noalloc: stencils (5 per point), then the right and
// left values, SWF_WENO_NOALLOC_MAX points.  Host code uses only the address that
// hipGetSymbolAddress returns for the current device, never this name.
static __device__ __align__(256) double scratch[7 * (size_t)SWF_WENO_NOALLOC_MAX];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

// this device's address of `scratch`, taken once per device with it current
int scratch_address(int device, double **out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void *ptr = nullptr;
    HIPC(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch_ptr[device] = static_cast<double *>(ptr);
  }
  *out = g_scratch_ptr[device];
  return SWF_OK;
}
}  // namespace

extern "C" {

int swf_weno_points_noalloc(const double *a, double *right, double *left, int64_t n, int32_t precision,
                            int32_t device) {
  const int e12 = (precision & SWF_WENO_EPS12) != 0;
  precision &= ~SWF_WENO_EPS12;
  if (!a || !right || !left || n < 0 || precision < 0 || precision > SWF_FP32_FULL) return SWF_E_INVALID;
  if (n > SWF_WENO_NOALLOC_MAX) {
    snprintf(g_last_error, sizeof(g_last_error), "swf_weno_points_noalloc: %lld stencils exceed the maximum %d",
             (long long)n, SWF_WENO_NOALLOC_MAX);
    return SWF_E_INVALID;
  }
  if (n == 0) return SWF_OK;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev || device >= kMaxDevices)
    return SWF_E_NODEVICE;
  HIPC(hipSetDevice(device));
  double *d = nullptr;
  int rc = scratch_address(device, &d);
  if (rc) return rc;
  hipError_t e = hipMemcpy(d, a, sizeof(double) * 5 * (size_t)n, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    launch_weno_points(precision, d, d + 5 * n, d + 6 * n, n, e12, 0);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(right, d + 5 * n, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(left, d + 6 * n, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(e, "swf_weno_points_noalloc");
  return SWF_OK;
}

// ---------------------------------------------------------------------------
// Batches of independent instances
// ---------------------------------------------------------------------------

}  // extern "C"

struct swf_batch {
  int n = 0, mode = 0, schedule = 0;
  swf_ctx **ctx = nullptr;     // the instances (their own buffers, time and statistics)
  BatchInst *d_inst = nullptr;
  int2 *d_xmap = nullptr, *d_ymap = nullptr;  // block -> (instance, tile)
  int nxb = 0, nyb = 0;
  // the same for the step-program instances (their own wenofd variant), and their indices
  int2 *d_xmap_s = nullptr, *d_ymap_s = nullptr;
  int nxb_s = 0, nyb_s = 0;
  // their per-line alpha tiles and reductions, and their wall images before the x-
  // and the y-pass: one launch each over every step-program instance
  int2 *d_lamap = nullptr, *d_lrmap = nullptr, *d_rf0map = nullptr, *d_rf1map = nullptr;
  int nla = 0, nlr = 0, nrf0 = 0, nrf1 = 0;
  std::vector<int> steps;
  double *d_time = nullptr, *d_tfinal = nullptr;
  BatchStatus *d_status = nullptr;
  hipStream_t stream = nullptr;
  hipGraphExec_t gexec = nullptr;
  int device = 0;
};

namespace {

// The order of the tiles of all instances in a batch launch.  0: instance by
// instance; 1: round robin over the instances (tile t of every instance before
// tile t+1); 2: round robin with the largest instances first.
// step: the map of the step-program instances, else of the others.
void batch_map(const swf_batch *bt, bool ytiles, int schedule, bool step, std::vector<int2> &map) {
  map.clear();
  std::vector<int> cnt(bt->n), order(bt->n);
  for (int i = 0; i < bt->n; ++i) {
    const KCfg &k = bt->ctx[i]->k;
    cnt[i] = ((k.alpha_lines != 0) != step) ? 0 : ytiles ? ub_gx(k) * ub_gy(k) : xf_gx(k) * xf_gy(k);
    order[i] = i;
  }
  if (schedule == 2)
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cnt[a] > cnt[b]; });
  if (schedule == 0) {
    for (int i = 0; i < bt->n; ++i)
      for (int t = 0; t < cnt[i]; ++t) map.push_back(make_int2(i, t));
    return;
  }
  int most = 0;
  for (int i = 0; i < bt->n; ++i) most = cnt[i] > most ? cnt[i] : most;
  for (int t = 0; t < most; ++t)
    for (int q = 0; q < bt->n; ++q) {
      const int i = order[q];
      if (t < cnt[i]) map.push_back(make_int2(i, t));
    }
}

int upload_map(const std::vector<int2> &m, int2 **d, int *n) {
  if (*d) (void)hipFree(*d);
  *d = nullptr;
  *n = (int)m.size();
  if (m.empty()) return SWF_OK;
  HIPC(hipMalloc(d, sizeof(int2) * m.size()));
  HIPC(hipMemcpy(*d, m.data(), sizeof(int2) * m.size(), hipMemcpyHostToDevice));
  return SWF_OK;
}

int batch_set_maps(swf_batch *bt, int schedule) {
  std::vector<int2> xm, ym, xs, ys;
  batch_map(bt, false, schedule, false, xm);
  batch_map(bt, true, schedule, false, ym);
  batch_map(bt, false, schedule, true, xs);
  batch_map(bt, true, schedule, true, ys);
  // the step-program instances' small launches, instance by instance
  std::vector<int2> la, lr, r0, r1;
  for (int i = 0; i < bt->n; ++i) {
    const KCfg &k = bt->ctx[i]->k;
    if (k.alpha_lines) {
      for (int q = 0; q < la_gx(k) * la_gy(k); ++q) la.push_back(make_int2(i, q));
      for (int q = 0; q < la_reduce_blocks(k); ++q) lr.push_back(make_int2(i, q));
    }
    if (k.step_total > 0) {
      for (int q = 0; q < (step_reflect_cells(k, 0) + 255) / 256; ++q) r0.push_back(make_int2(i, q));
      for (int q = 0; q < (step_reflect_cells(k, 1) + 255) / 256; ++q) r1.push_back(make_int2(i, q));
    }
  }
  int rc = upload_map(xm, &bt->d_xmap, &bt->nxb);
  if (!rc) rc = upload_map(ym, &bt->d_ymap, &bt->nyb);
  if (!rc) rc = upload_map(xs, &bt->d_xmap_s, &bt->nxb_s);
  if (!rc) rc = upload_map(ys, &bt->d_ymap_s, &bt->nyb_s);
  if (!rc) rc = upload_map(la, &bt->d_lamap, &bt->nla);
  if (!rc) rc = upload_map(lr, &bt->d_lrmap, &bt->nlr);
  if (!rc) rc = upload_map(r0, &bt->d_rf0map, &bt->nrf0);
  if (!rc) rc = upload_map(r1, &bt->d_rf1map, &bt->nrf1);
  if (rc) return rc;
  bt->steps.clear();
  for (int i = 0; i < bt->n; ++i)
    if (bt->ctx[i]->k.alpha_lines || bt->ctx[i]->k.step_total > 0) bt->steps.push_back(i);
  bt->schedule = schedule;
  if (bt->gexec) {
    (void)hipGraphExecDestroy(bt->gexec);
    bt->gexec = nullptr;
  }
  return SWF_OK;
}

// the flux launches of one stage over every instance: the step-program instances
// (per-line alphas, 1d-12) have kernels of their own, in the mode their contexts use
void batch_flux_launches(swf_batch *bt, bool y, int s, int slot_in, int p, int flags) {
  const int ms = kernel_mode(bt->mode, SWF_BC_STEP);
  if (y) {
    if (bt->nyb) launch_ystage_batch(bt->mode, bt->d_inst, bt->d_ymap, bt->nyb, bt->stream, s, slot_in, p, flags, 0);
    if (bt->nyb_s) launch_ystage_batch(ms, bt->d_inst, bt->d_ymap_s, bt->nyb_s, bt->stream, s, slot_in, p, flags, 1);
  } else {
    if (bt->nxb) launch_xflux_batch(bt->mode, bt->d_inst, bt->d_xmap, bt->nxb, bt->stream, s, slot_in, p, 0);
    if (bt->nxb_s) launch_xflux_batch(ms, bt->d_inst, bt->d_xmap_s, bt->nxb_s, bt->stream, s, slot_in, p, 1);
  }
}

// one RK stage of every instance: the flux and stage-end launches over all of
// them, and over the step-program instances their per-line alphas and wall images
void batch_enqueue_stage(swf_batch *bt, int s, int p) {
  const int slot_in = s - 1, slot_out = s % 3;
  launch_line_alpha_batch(bt->d_inst, bt->d_lamap, bt->nla, bt->d_lrmap, bt->nlr, bt->stream, s, slot_in, p);
  launch_step_reflect_batch(bt->d_inst, bt->d_rf0map, bt->nrf0, bt->stream, s, 0, p);
  batch_flux_launches(bt, false, s, slot_in, p, 0);
  launch_step_reflect_batch(bt->d_inst, bt->d_rf1map, bt->nrf1, bt->stream, s, 1, p);
  batch_flux_launches(bt, true, s, slot_in, p, 0);
  launch_stage_end_batch(bt->d_inst, bt->n, bt->stream, s, slot_out, p, bt->mode == MODE_EXACT ? 1 : 0);
}

void batch_enqueue_step(swf_batch *bt, int p) {
  batch_enqueue_stage(bt, 1, p);
  batch_enqueue_stage(bt, 2, p);
  batch_enqueue_stage(bt, 3, p);
}

int batch_ensure_graph(swf_batch *bt) {
  if (bt->gexec || no_graph()) return SWF_OK;
  hipGraph_t graph = nullptr;
  HIPC(hipStreamBeginCapture(bt->stream, hipStreamCaptureModeThreadLocal));
  for (int q = 0; q < kGraphSteps; ++q) batch_enqueue_step(bt, q & 1);
  HIPC(hipStreamEndCapture(bt->stream, &graph));
  hipError_t e = hipGraphInstantiate(&bt->gexec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate (batch)");
  return SWF_OK;
}

}  // namespace

extern "C" {

int swf_batch_create(const swf_config *cfgs, int32_t n, swf_batch **out) {
  if (!cfgs || !out || n < 1) return SWF_E_INVALID;
  *out = nullptr;
  for (int i = 0; i < n; ++i) {
    if (cfgs[i].precision != cfgs[0].precision || cfgs[i].device != cfgs[0].device || cfgs[i].use_limiter == 1)
      return SWF_E_INVALID;
    // the jets run the FP32-bulk form (kernel_mode); the other instances share their kernels
    if (cfgs[i].precision == SWF_FP32_FULL && kernel_mode(SWF_FP32_FULL, cfgs[i].bc) != SWF_FP32_FULL &&
        cfgs[i].bc != SWF_BC_STEP) {
      snprintf(g_last_error, sizeof(g_last_error), "swf_batch_create: SWF_FP32_FULL batches take no jet instances");
      return SWF_E_INVALID;
    }
  }
  swf_batch *bt = new (std::nothrow) swf_batch;
  if (!bt) return SWF_E_NOMEM;
  bt->n = n;
  bt->mode = cfgs[0].precision;
  bt->device = cfgs[0].device;
  bt->ctx = new (std::nothrow) swf_ctx *[n]();
  int rc = bt->ctx ? SWF_OK : SWF_E_NOMEM;
  for (int i = 0; rc == SWF_OK && i < n; ++i) rc = swf_create(&cfgs[i], &bt->ctx[i]);
  if (rc == SWF_OK) {
    std::vector<BatchInst> h(n);
    for (int i = 0; i < n; ++i) {
      h[i].k = bt->ctx[i]->k;
      h[i].b = bt->ctx[i]->b;
      h[i].S1 = bt->ctx[i]->S1;
      h[i].S2 = bt->ctx[i]->S2;
    }
    hipError_t e = hipSetDevice(bt->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&bt->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&bt->d_inst, sizeof(BatchInst) * n);
    if (e == hipSuccess) e = hipMemcpy(bt->d_inst, h.data(), sizeof(BatchInst) * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc(&bt->d_time, sizeof(double) * 2 * n);
    if (e == hipSuccess) e = hipMalloc(&bt->d_status, sizeof(BatchStatus) * n);
    if (e != hipSuccess) rc = (e == hipErrorOutOfMemory) ? SWF_E_NOMEM : hip_fail(e, "swf_batch_create");
    bt->d_tfinal = bt->d_time ? bt->d_time + n : nullptr;
  }
  if (rc == SWF_OK) rc = batch_set_maps(bt, 1);
  if (rc != SWF_OK) {
    swf_batch_destroy(bt);
    return rc;
  }
  *out = bt;
  return SWF_OK;
}

void swf_batch_destroy(swf_batch *bt) {
  if (!bt) return;
  (void)hipSetDevice(bt->device);
  if (bt->stream) (void)hipStreamSynchronize(bt->stream);
  if (bt->gexec) (void)hipGraphExecDestroy(bt->gexec);
  (void)hipFree(bt->d_inst);
  (void)hipFree(bt->d_xmap);
  (void)hipFree(bt->d_ymap);
  if (bt->d_xmap_s) (void)hipFree(bt->d_xmap_s);
  if (bt->d_ymap_s) (void)hipFree(bt->d_ymap_s);
  for (int2 *m : {bt->d_lamap, bt->d_lrmap, bt->d_rf0map, bt->d_rf1map})
    if (m) (void)hipFree(m);
  (void)hipFree(bt->d_time);
  (void)hipFree(bt->d_status);
  if (bt->stream) (void)hipStreamDestroy(bt->stream);
  if (bt->ctx)
    for (int i = 0; i < bt->n; ++i) swf_destroy(bt->ctx[i]);
  delete[] bt->ctx;
  delete bt;
}

swf_ctx *swf_batch_instance(swf_batch *bt, int32_t i) {
  if (!bt || i < 0 || i >= bt->n) return nullptr;
  return bt->ctx[i];
}

int swf_batch_set_schedule(swf_batch *bt, int32_t schedule) {
  if (!bt || schedule < 0 || schedule > 2) return SWF_E_INVALID;
  HIPC(hipSetDevice(bt->device));
  return batch_set_maps(bt, schedule);
}

int swf_batch_advance(swf_batch *bt, double *time_inout, const double *tfinal, int64_t max_steps, int64_t *steps_out,
                      int32_t *status_out) {
  if (!bt || !time_inout || !tfinal || max_steps < 0) return SWF_E_INVALID;
  HIPC(hipSetDevice(bt->device));
  int rc = batch_ensure_graph(bt);
  if (rc) return rc;
  const int n = bt->n;
  std::vector<BatchStatus> h(n);
  HIPC(hipMemcpyAsync(bt->d_time, time_inout, sizeof(double) * n, hipMemcpyHostToDevice, bt->stream));
  HIPC(hipMemcpyAsync(bt->d_tfinal, tfinal, sizeof(double) * n, hipMemcpyHostToDevice, bt->stream));
  launch_begin_run_batch(bt->d_inst, n, bt->d_time, bt->d_tfinal, bt->stream);
  int64_t launched = 0, since_check = 0;
  while (launched < max_steps) {
    if (max_steps - launched >= kGraphSteps && !no_graph()) {
      HIPC(hipGraphLaunch(bt->gexec, bt->stream));
      launched += kGraphSteps;
      since_check += kGraphSteps;
    } else {
      for (; launched < max_steps; ++launched, ++since_check) batch_enqueue_step(bt, (int)(launched & 1));
    }
    rc = check_launch("swf_batch_advance");
    if (rc) return rc;
    if (since_check >= kCheckEvery || launched >= max_steps) {
      since_check = 0;
      launch_batch_status(bt->d_inst, n, bt->d_status, bt->stream);
      HIPC(hipMemcpyAsync(h.data(), bt->d_status, sizeof(BatchStatus) * n, hipMemcpyDeviceToHost, bt->stream));
      HIPC(hipStreamSynchronize(bt->stream));
      bool any = false;  // an instance still stepping
      for (int i = 0; i < n; ++i) any = any || (h[i].status == ST_OK && h[i].steps_done >= launched);
      if (!any) break;
    }
  }
  launch_batch_status(bt->d_inst, n, bt->d_status, bt->stream);
  HIPC(hipMemcpyAsync(h.data(), bt->d_status, sizeof(BatchStatus) * n, hipMemcpyDeviceToHost, bt->stream));
  HIPC(hipStreamSynchronize(bt->stream));
  int worst = SWF_OK;
  for (int i = 0; i < n; ++i) {
    time_inout[i] = h[i].time;
    if (steps_out) steps_out[i] = h[i].steps_done;
    if (status_out) status_out[i] = h[i].status;
    bt->ctx[i]->steps_total += h[i].steps_done;
    if (h[i].status != SWF_OK && worst == SWF_OK) worst = h[i].status;
  }
  return worst;
}

int swf_batch_tune(swf_batch *bt, int32_t reps, int32_t *schedule_out, float *ms_out) {
  if (!bt || reps < 1) return SWF_E_INVALID;
  HIPC(hipSetDevice(bt->device));
  // the stage-1 kernels of every instance without side effects on the state:
  // x-fluxes into Fx, stage results into S1, no scalars; every instance's
  // step gate is opened for the timing and restored after it
  std::vector<DevScalars> h0(bt->n);
  for (int i = 0; i < bt->n; ++i) {
    int rc = read_scalars(bt->ctx[i], &h0[i]);
    if (rc) return rc;
    launch_set_alpha(bt->ctx[i]->b, bt->stream, 0, h0[i].alpha_bits[0], 0, 1);
  }
  hipEvent_t e0, e1;
  HIPC(hipEventCreate(&e0));
  HIPC(hipEventCreate(&e1));
  float best = -1.f;
  int pick = 0;
  for (int sched = 0; sched < 3; ++sched) {
    int rc = batch_set_maps(bt, sched);
    if (rc) return rc;
    // the step-program instances' per-line alphas of their current state
    for (int i : bt->steps)
      if (bt->ctx[i]->k.alpha_lines) launch_line_alpha(bt->ctx[i]->k, bt->ctx[i]->b, bt->stream, 3, -1);
    for (int w = 0; w < 2; ++w) {
      batch_flux_launches(bt, false, 0, 0, 0, 0);
      batch_flux_launches(bt, true, 1, 0, 0, YS_NO_SCALARS);
    }
    HIPC(hipEventRecord(e0, bt->stream));
    for (int r = 0; r < reps; ++r) {
      batch_flux_launches(bt, false, 0, 0, 0, 0);
      batch_flux_launches(bt, true, 1, 0, 0, YS_NO_SCALARS);
    }
    HIPC(hipEventRecord(e1, bt->stream));
    HIPC(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPC(hipEventElapsedTime(&ms, e0, e1));
    ms /= reps;
    if (ms_out) ms_out[sched] = ms;
    if (best < 0.f || ms < best) {
      best = ms;
      pick = sched;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  for (int i = 0; i < bt->n; ++i) launch_set_alpha(bt->ctx[i]->b, bt->stream, 0, h0[i].alpha_bits[0], 0, h0[i].active[0]);
  HIPC(hipStreamSynchronize(bt->stream));
  int rc = check_launch("swf_batch_tune");
  if (rc) return rc;
  if (schedule_out) *schedule_out = pick;
  return batch_set_maps(bt, pick);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// The 1D program: one workgroup runs the time loop (swf_1d.h)
// ---------------------------------------------------------------------------

namespace {
constexpr int kTrackCap1d = 1 << 16;
constexpr long long kStepsPerLaunch1d = 4096;  // bounds one launch's duration
}  // namespace

struct swf1d_ctx {
  swf1d_config cfg;
  K1Cfg k;
  B1 b;
  int mode;
  hipStream_t stream;
  hipEvent_t e0, e1;
  double last_ms;
  long long steps_total;
};

extern "C" {

void swf1d_destroy(swf1d_ctx *c) {
  if (!c) return;
  (void)hipSetDevice(c->cfg.device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->b.u) (void)hipFree(c->b.u);
  if (c->b.g) (void)hipFree(c->b.g);
  if (c->b.trk) (void)hipFree(c->b.trk);
  if (c->e0) (void)hipEventDestroy(c->e0);
  if (c->e1) (void)hipEventDestroy(c->e1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int swf1d_create(const swf1d_config *cfg, swf1d_ctx **out) {
  if (!cfg || !out) return SWF_E_INVALID;
  *out = nullptr;
  const int N = cfg->n + 6;
  if (cfg->n < 1 || N > D1_MAXN || !(cfg->dx > 0.0) || !(cfg->gamma > 1.0) || !(cfg->cfl > 0.0) ||
      (cfg->use_limiter != 0 && cfg->use_limiter != 2) || (cfg->use_limiter == 2 && cfg->swp_num < 1) ||
      cfg->precision < SWF_FP64_EXACT || cfg->precision > SWF_FP32_BULK) {
    snprintf(g_last_error, sizeof(g_last_error),
             "swf1d_create: need 1 <= n <= %d, dx > 0, gamma > 1, cfl > 0, use_limiter 0 or 2 (swp_num >= 1), "
             "precision 0..2",
             D1_MAXN - 6);
    return SWF_E_INVALID;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= cfg->device || cfg->device < 0) return SWF_E_NODEVICE;
  HIPC(hipSetDevice(cfg->device));
  int lds_max = 0;
  HIPC(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, cfg->device));
  if ((size_t)lds_max < d1_lds_bytes(N)) {
    snprintf(g_last_error, sizeof(g_last_error), "swf1d_create: n = %d needs %zu bytes of LDS, the device has %d",
             cfg->n, d1_lds_bytes(N), lds_max);
    return SWF_E_INVALID;
  }
  swf1d_ctx *c = new (std::nothrow) swf1d_ctx();
  if (!c) return SWF_E_NOMEM;
  c->cfg = *cfg;
  c->mode = cfg->precision;
  K1Cfg &k = c->k;
  k.n = cfg->n;
  k.N = N;
  k.use_limiter = cfg->use_limiter;
  k.swp_num = cfg->swp_num;
  k.track_cap = kTrackCap1d;
  k.dx = cfg->dx;
  k.idx = 1.0 / cfg->dx;
  k.gamma = cfg->gamma;
  k.gamma1 = cfg->gamma - 1.;
  k.eps = cfg->eps;
  // weights(i) = 1/(nSteps + 6) with nSteps = n as REAL (ex2_sedov_1d/main.f90:5,34-36)
  k.weight = 1. / ((double)cfg->n + 6.);
  k.dt_num = cfg->dx * cfg->cfl;
  if (hipMalloc(&c->b.u, sizeof(double) * 3 * N) != hipSuccess ||
      hipMalloc(&c->b.g, sizeof(D1Scalars)) != hipSuccess ||
      hipMalloc(&c->b.trk, sizeof(long long) * 3 * kTrackCap1d) != hipSuccess) {
    swf1d_destroy(c);
    return SWF_E_NOMEM;
  }
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&c->e0) != hipSuccess || hipEventCreate(&c->e1) != hipSuccess ||
      hipMemsetAsync(c->b.g, 0, sizeof(D1Scalars), c->stream) != hipSuccess ||
      hipMemsetAsync(c->b.trk, 0, sizeof(long long) * 3 * kTrackCap1d, c->stream) != hipSuccess ||
      hipMemsetAsync(c->b.u, 0, sizeof(double) * 3 * N, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess) {
    const hipError_t e = hipGetLastError();
    hip_fail(e, "swf1d_create");
    swf1d_destroy(c);
    return SWF_E_HIP;
  }
  *out = c;
  return SWF_OK;
}

int swf1d_upload(swf1d_ctx *c, const double *u) {
  if (!c || !u) return SWF_E_INVALID;
  HIPC(hipSetDevice(c->cfg.device));
  HIPC(hipMemcpyAsync(c->b.u, u, sizeof(double) * 3 * c->k.N, hipMemcpyHostToDevice, c->stream));
  HIPC(hipStreamSynchronize(c->stream));
  return SWF_OK;
}

int swf1d_download(swf1d_ctx *c, double *u) {
  if (!c || !u) return SWF_E_INVALID;
  HIPC(hipSetDevice(c->cfg.device));
  HIPC(hipMemcpyAsync(u, c->b.u, sizeof(double) * 3 * c->k.N, hipMemcpyDeviceToHost, c->stream));
  HIPC(hipStreamSynchronize(c->stream));
  return SWF_OK;
}

static int read_scalars1d(swf1d_ctx *c, D1Scalars *h) {
  HIPC(hipMemcpyAsync(h, c->b.g, sizeof(D1Scalars), hipMemcpyDeviceToHost, c->stream));
  HIPC(hipStreamSynchronize(c->stream));
  return SWF_OK;
}

int swf1d_advance(swf1d_ctx *c, double *time_inout, double tfinal, int64_t max_steps, int64_t *steps_out) {
  if (!c || !time_inout || max_steps < 0) return SWF_E_INVALID;
  HIPC(hipSetDevice(c->cfg.device));
  D1Scalars h;
  int rc = read_scalars1d(c, &h);
  if (rc) return rc;
  h.time = *time_inout;
  h.tfinal = tfinal;
  HIPC(hipMemcpyAsync(c->b.g, &h, sizeof(D1Scalars), hipMemcpyHostToDevice, c->stream));
  long long done = 0;
  const long long start = h.steps_done;
  HIPC(hipEventRecord(c->e0, c->stream));
  while (done < max_steps) {
    const long long chunk = std::min<long long>(max_steps - done, kStepsPerLaunch1d);
    const hipError_t e = launch_run1d(c->mode, c->k, c->b, c->stream, chunk);
    if (e != hipSuccess) return hip_fail(e, "k_run1d launch");
    rc = read_scalars1d(c, &h);
    if (rc) return rc;
    const long long took = h.steps_done - start - done;
    done += took;
    if (took < chunk) break;  // the time test ended the loop
  }
  HIPC(hipEventRecord(c->e1, c->stream));
  HIPC(hipEventSynchronize(c->e1));
  float ms = 0.f;
  HIPC(hipEventElapsedTime(&ms, c->e0, c->e1));
  c->last_ms = ms;
  c->steps_total += done;
  *time_inout = h.time;
  if (steps_out) *steps_out = done;
  return check_launch("swf1d_advance");
}

int swf1d_stats(swf1d_ctx *c, swf_sweep_stats *s) {
  if (!c || !s) return SWF_E_INVALID;
  HIPC(hipSetDevice(c->cfg.device));
  D1Scalars h;
  const int rc = read_scalars1d(c, &h);
  if (rc) return rc;
  s->sweep_count = h.sweep_count;
  s->den_count = h.den_count;
  s->pre_count = h.pre_count;
  s->total_sweep = h.total_sweep;
  s->max_sweep = h.max_sweep;
  s->total_neg = h.total_neg;
  s->max_neg = h.max_neg;
  s->total_pos = h.total_pos;
  s->max_pos = h.max_pos;
  s->stage_sweeps = h.stage_sweeps;
  s->steps = c->steps_total;
  return SWF_OK;
}

int64_t swf1d_track(swf1d_ctx *c, int64_t *q, int64_t *neg, int64_t *pos, int64_t cap) {
  if (!c) return SWF_E_INVALID;
  if (hipSetDevice(c->cfg.device) != hipSuccess) return SWF_E_HIP;
  D1Scalars h;
  if (read_scalars1d(c, &h)) return SWF_E_HIP;
  const int64_t n = h.sweep_count < cap ? h.sweep_count : cap;
  const int64_t m = n < kTrackCap1d ? n : kTrackCap1d;
  if (m > 0 && q && neg && pos) {
    if (hipMemcpy(q, c->b.trk, sizeof(int64_t) * m, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(neg, c->b.trk + kTrackCap1d, sizeof(int64_t) * m, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(pos, c->b.trk + 2 * kTrackCap1d, sizeof(int64_t) * m, hipMemcpyDeviceToHost) != hipSuccess)
      return SWF_E_HIP;
  }
  return h.sweep_count;
}

int swf1d_last_ms(swf1d_ctx *c, double *ms) {
  if (!c || !ms) return SWF_E_INVALID;
  *ms = c->last_ms;
  return SWF_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// The detonation program (swf_ex7.h)
// ---------------------------------------------------------------------------
namespace {
constexpr int kTrackCap7 = 1 << 16;
}

struct swf7_ctx {
  swf7_config cfg;
  ex7::K7 k;
  ex7::B7 b;       // outside a step: Uin == U == U0 == S0
  double *S1, *S2;
  hipStream_t stream;
  double *pool;    // every device array of the context
  double *d_fortran;
  size_t n_fortran;
  long long steps_total;
};

static int status7(int st) {
  switch (st) {
    case ex7::ST7_OK: return SWF_OK;
    case ex7::ST7_PBAR: return SWF_E_PBAR;
    case ex7::ST7_SWEEP_FAIL: return SWF_E_SWEEP_FAIL;
    case ex7::ST7_PSUM: return SWF_E_PSUM;
    default: return SWF_E_ORDER2_EDGE;
  }
}

static int read7(swf7_ctx *c, ex7::Dev7 *h) {
  HIPC(hipMemcpyAsync(h, c->b.g, sizeof(ex7::Dev7), hipMemcpyDeviceToHost, c->stream));
  HIPC(hipStreamSynchronize(c->stream));
  return SWF_OK;
}

extern "C" {

int swf7_create(const swf7_config *cfg, swf7_ctx **out) {
  if (!cfg || !out) return SWF_E_INVALID;
  *out = nullptr;
  if (cfg->total < 2 || cfg->step < 1 || 5LL * cfg->total * 5LL * cfg->total >= (1LL << 31) ||
      cfg->step * cfg->total + 3 > 5 * cfg->total || !(cfg->dx > 0) || !(cfg->dy > 0) || cfg->swp_num < 1)
    return SWF_E_INVALID;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= cfg->device || cfg->device < 0) {
    snprintf(g_last_error, sizeof(g_last_error), "no HIP device %d (found %d)", cfg->device, ndev);
    return SWF_E_NODEVICE;
  }
  swf7_ctx *c = new (std::nothrow) swf7_ctx();
  if (!c) return SWF_E_NOMEM;
  c->cfg = *cfg;
  ex7::K7 &k = c->k;
  k.nx = k.ny = 5 * cfg->total;
  k.P = row_pitch(k.nx);
  k.plane = (long long)(k.ny + 6) * k.P;
  k.total = cfg->total;
  k.step = cfg->step;
  k.swp_num = cfg->swp_num;
  k.track_cap = kTrackCap7;
  k.dx = cfg->dx;
  k.dy = cfg->dy;
  k.gamma = cfg->gamma;
  k.gamma1 = cfg->gamma - 1.;
  k.eps = cfg->eps;
  k.eps2 = cfg->eps2;
  k.q = cfg->q_const;
  k.tc = cfg->t_const;
  k.kc = cfg->k_const;
  k.cfl = cfg->cfl;
  {
    // weights(i, j) = 1./(nSteps_x*nSteps_y) with nSteps the REAL(8) grid sizes
    const double n = (double)k.nx;
    k.weight = 1. / (n * n);
  }
  c->n_fortran = (size_t)ex7::NC * (k.nx + 6) * (k.ny + 6);
  // 3 states, 4 flux arrays of 5 planes each, alpha, line maxima, the Fortran staging
  const size_t st = (size_t)ex7::NC * k.plane;
  const size_t ndbl = 7 * st + (size_t)k.plane + (size_t)(k.nx + k.ny + 12) + c->n_fortran;
  hipError_t e = hipSetDevice(cfg->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc(&c->pool, sizeof(double) * ndbl);
  ex7::Dev7 *g = nullptr;
  long long *trk = nullptr;
  unsigned *bm = nullptr;
  int *lst = nullptr;
  const long long npos = (long long)k.nx * k.ny;
  if (e == hipSuccess) e = hipMalloc(&g, sizeof(ex7::Dev7));
  if (e == hipSuccess) e = hipMalloc(&trk, sizeof(long long) * 3 * kTrackCap7);
  if (e == hipSuccess) e = hipMalloc(&bm, sizeof(unsigned) * ((npos + 31) / 32));
  if (e == hipSuccess) e = hipMalloc(&lst, sizeof(int) * npos);
  if (e == hipSuccess) e = hipMemsetAsync(c->pool, 0, sizeof(double) * ndbl, c->stream);
  if (e == hipSuccess) e = hipMemsetAsync(g, 0, sizeof(ex7::Dev7), c->stream);
  // the program's sweep_track / neg_track / pos_track start at 0; a sweep that stops
  // leaves its entry there
  if (e == hipSuccess) e = hipMemsetAsync(trk, 0, sizeof(long long) * 3 * kTrackCap7, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) {
    const int rc = (e == hipErrorOutOfMemory) ? SWF_E_NOMEM : hip_fail(e, "swf7_create");
    if (c->pool) (void)hipFree(c->pool);
    if (g) (void)hipFree(g);
    if (trk) (void)hipFree(trk);
    if (bm) (void)hipFree(bm);
    if (lst) (void)hipFree(lst);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return rc;
  }
  double *p = c->pool;
  c->b.U = c->b.Uin = c->b.U0 = p;
  c->S1 = p + st;
  c->S2 = p + 2 * st;
  c->b.fpf = p + 3 * st;
  c->b.fnf = p + 4 * st;
  c->b.gpf = p + 5 * st;
  c->b.gnf = p + 6 * st;
  c->b.al = p + 7 * st;
  c->b.arow = c->b.al + k.plane;
  c->b.acol = c->b.arow + (k.ny + 6);
  c->d_fortran = c->b.acol + (k.nx + 6);
  c->b.g = g;
  c->b.trk = trk;
  c->b.bm = bm;
  c->b.lst = lst;
  *out = c;
  return SWF_OK;
}

void swf7_destroy(swf7_ctx *c) {
  if (!c) return;
  (void)hipSetDevice(c->cfg.device);
  (void)hipStreamSynchronize(c->stream);
  (void)hipFree(c->pool);
  (void)hipFree(c->b.g);
  (void)hipFree(c->b.trk);
  (void)hipFree(c->b.bm);
  (void)hipFree(c->b.lst);
  (void)hipStreamDestroy(c->stream);
  delete c;
}

int swf7_upload(swf7_ctx *c, const double *u) {
  if (!c || !u) return SWF_E_INVALID;
  HIPC(hipSetDevice(c->cfg.device));
  HIPC(hipMemcpyAsync(c->d_fortran, u, sizeof(double) * c->n_fortran, hipMemcpyHostToDevice, c->stream));
  launch7_from_fortran(c->k, c->d_fortran, c->b.U, c->stream);
  HIPC(hipStreamSynchronize(c->stream));
  return check_launch("swf7_upload");
}

int swf7_download(swf7_ctx *c, double *u) {
  if (!c || !u) return SWF_E_INVALID;
  HIPC(hipSetDevice(c->cfg.device));
  launch7_to_fortran(c->k, c->b.U, c->d_fortran, c->stream);
  HIPC(hipMemcpyAsync(u, c->d_fortran, sizeof(double) * c->n_fortran, hipMemcpyDeviceToHost, c->stream));
  HIPC(hipStreamSynchronize(c->stream));
  return check_launch("swf7_download");
}

int swf7_bc(swf7_ctx *c) {
  if (!c) return SWF_E_INVALID;
  HIPC(hipSetDevice(c->cfg.device));
  launch7_bc(c->k, c->b, c->stream);
  HIPC(hipStreamSynchronize(c->stream));
  return check_launch("swf7_bc");
}

// The time loop: each step's three stages on the device, then the host adds dt to the
// time (time = time + dt in binary64, :99) and tests the loop condition.
int swf7_advance(swf7_ctx *c, double *time_inout, double tfinal, int64_t max_steps, int64_t *steps_out) {
  if (!c || !time_inout || max_steps < 0) return SWF_E_INVALID;
  HIPC(hipSetDevice(c->cfg.device));
  double time = *time_inout;
  int64_t n = 0;
  ex7::Dev7 h;
  int rc = read7(c, &h);
  if (rc) return rc;
  if (h.status != ex7::ST7_OK) return status7(h.status);
  h.tfinal = tfinal;
  h.active = 1;
  while (n < max_steps && time_continue_q(time, tfinal)) {
    h.time = time;
    HIPC(hipMemcpyAsync(c->b.g, &h, sizeof(ex7::Dev7), hipMemcpyHostToDevice, c->stream));
    for (int s = 1; s <= 3; ++s) {
      ex7::B7 b = c->b;
      b.Uin = (s == 1) ? c->b.U : (s == 2) ? c->S1 : c->S2;
      b.U = (s == 1) ? c->S1 : (s == 2) ? c->S2 : c->b.U;
      launch7_stage(c->k, b, c->stream, s);
    }
    rc = check_launch("swf7_advance");
    if (rc) return rc;
    rc = read7(c, &h);
    if (rc) return rc;
    if (h.status != ex7::ST7_OK) break;
    time = time + h.dt;
    h.steps_done += 1;
    ++n;
  }
  h.active = 0;
  HIPC(hipMemcpyAsync(c->b.g, &h, sizeof(ex7::Dev7), hipMemcpyHostToDevice, c->stream));
  HIPC(hipStreamSynchronize(c->stream));
  *time_inout = time;
  if (steps_out) *steps_out = n;
  c->steps_total += n;
  return status7(h.status);
}

// sweep_lim on the current state: the positivity flag set, the sweep kernel alone
int swf7_sweep(swf7_ctx *c) {
  if (!c) return SWF_E_INVALID;
  HIPC(hipSetDevice(c->cfg.device));
  ex7::Dev7 h;
  int rc = read7(c, &h);
  if (rc) return rc;
  if (h.status != ex7::ST7_OK) return status7(h.status);
  h.postag = 1;
  h.active = 1;
  HIPC(hipMemcpyAsync(c->b.g, &h, sizeof(ex7::Dev7), hipMemcpyHostToDevice, c->stream));
  launch7_sweep(c->k, c->b, c->stream);
  rc = check_launch("swf7_sweep");
  if (rc) return rc;
  rc = read7(c, &h);
  if (rc) return rc;
  return status7(h.status);
}

int swf7_stats(swf7_ctx *c, swf_sweep_stats *s) {
  if (!c || !s) return SWF_E_INVALID;
  HIPC(hipSetDevice(c->cfg.device));
  ex7::Dev7 h;
  const int rc = read7(c, &h);
  if (rc) return rc;
  *s = swf_sweep_stats();
  s->sweep_count = h.sweep_count;
  s->den_count = h.den_count;
  s->pre_count = h.pre_count;
  s->stage_sweeps = h.sweep_count;
  s->steps = c->steps_total;
  const long long m = h.sweep_count < kTrackCap7 ? h.sweep_count : kTrackCap7;
  if (m > 0) {
    std::vector<long long> t(3 * (size_t)m);
    HIPC(hipMemcpy(t.data(), c->b.trk, sizeof(long long) * 3 * m, hipMemcpyDeviceToHost));
    for (long long i = 0; i < m; ++i) {
      s->total_sweep += t[3 * i];
      s->total_neg += t[3 * i + 1];
      s->total_pos += t[3 * i + 2];
      s->max_sweep = i == 0 || t[3 * i] > s->max_sweep ? t[3 * i] : s->max_sweep;
      s->max_neg = i == 0 || t[3 * i + 1] > s->max_neg ? t[3 * i + 1] : s->max_neg;
      s->max_pos = i == 0 || t[3 * i + 2] > s->max_pos ? t[3 * i + 2] : s->max_pos;
    }
  }
  return SWF_OK;
}

int64_t swf7_track(swf7_ctx *c, int64_t *q, int64_t *neg, int64_t *pos, int64_t cap) {
  if (!c) return SWF_E_INVALID;
  if (hipSetDevice(c->cfg.device) != hipSuccess) return SWF_E_HIP;
  ex7::Dev7 h;
  if (read7(c, &h)) return SWF_E_HIP;
  const int64_t n = h.sweep_count < cap ? h.sweep_count : cap;
  const int64_t m = n < kTrackCap7 ? n : kTrackCap7;
  if (m > 0 && q && neg && pos) {
    std::vector<long long> t(3 * (size_t)m);
    if (hipMemcpy(t.data(), c->b.trk, sizeof(long long) * 3 * m, hipMemcpyDeviceToHost) != hipSuccess)
      return SWF_E_HIP;
    for (int64_t i = 0; i < m; ++i) {
      q[i] = t[3 * i];
      neg[i] = t[3 * i + 1];
      pos[i] = t[3 * i + 2];
    }
  }
  return h.sweep_count;
}

}  // extern "C"
