/*
 * swf.h - C-ABI of the MI355X sweeping FD-WENO library (libswf.so).
 *
 * Drop-in boundary for the reference's time-stepping hot path.  The reference
 * (dcgriffi/Sweeping_FD_WENO) has no library API: its "operators" are argument-less
 * internal procedures of each example PROGRAM that share host-associated arrays.
 * Each entry point below replaces one of them (ex4_sedov_2d/main.f90 unless noted):
 *
 *   swf_create / swf_destroy   the PARAMETER block and array declarations (:15-67)
 *   swf_upload / swf_download  the global state array u(4,-2:nx+3,-2:ny+3) (:21)
 *   swf_step                   one call of `rk` (:685-777) inside the time loop (:112-117)
 *   swf_advance                the time loop `DO WHILE (time <= time_terminal - 1d-10)` (:112)
 *   swf_alpha_max              `findalpha` (:652-683)
 *   swf_rhs                    `wenofd` + `scheme` (:179-650)
 *   swf_check_positivity       `checkpositivity` (:779-801)
 *   swf_sweep                  `sweep_lim` incl. `scalar_sweep` (:1401-1833)
 *   swf_bc                     `bc` (:803-996, physical part; ex6: :888-942)
 *   swf_stats / swf_track      the statistics printed by `output` (:1326-1351)
 *   swf_weno_points            weno_right / weno_left (ex4_sedov_2d/weno5.f90:86-146)
 *   swf_time_continue          the binary128 loop test of :112
 *   swf_batch_*                many instances advanced together (no counterpart: the
 *                              reference runs one problem per program; see below)
 *
 * Arrays crossing the boundary are host memory in the reference's Fortran layout
 * u(4, -2:nx+3, -2:ny+3): column-major, component fastest.  The caller owns them;
 * the context owns its device buffers.  All calls on one context are ordered on
 * the context's private HIP stream; swf_download, swf_stats and every call that
 * returns a host value synchronize it.  One context per host thread.
 *
 * Return codes: 0 OK; > 0 numerical outcomes of the reference's `stop` paths
 * (SWF_E_PBAR, SWF_E_SWEEP_FAIL); < 0 usage or HIP errors.  swf_strerror()
 * returns a static description.
 */
#ifndef SWF_H
#define SWF_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* precision modes */
#define SWF_FP64_EXACT 0 /* bit-identical to the reference (binary128 sites emulated) */
#define SWF_FP64 1       /* plain binary64 */
#define SWF_FP32_BULK 2  /* FP32 smoothness indicators / weights, FP64 state and sweep */
#define SWF_FP32_FULL 3  /* FP32-bulk with the eigen-projections, candidate values, reconstruction and
                          * back-projection in FP32 too (Roe averages, centre values, RK, pressure,
                          * positivity and the sweep in FP64): the Sedov and periodic (vortex)
                          * problems.  The jets (Mach 2000: ill-conditioned projections) and the step
                          * program (chaotic on coarse grids) run the SWF_FP32_BULK form under it.
                          * 2D contexts and batches (no jet instances); swf1d_create takes 0..2 */

/* boundary conditions */
#define SWF_BC_SEDOV 0    /* ex4: reflective x=0, y=0; zero-gradient x=xm, y=ym */
#define SWF_BC_PERIODIC 1 /* ex3 */
#define SWF_BC_JET_HALF 2 /* ex5: inflow at x=0 for y<=w, reflective y=0, zero-gradient elsewhere */
#define SWF_BC_JET_FULL 3 /* full-domain jet: inflow at x=0 for |y|<w, zero-gradient elsewhere */
#define SWF_BC_STEP 4     /* ex6: gas at rest below and right, zero-gradient above, post-shock
                             inflow left, column 1 reset below 6*step_total (the program's bc) */

/* return codes */
#define SWF_OK 0
#define SWF_E_PBAR 1        /* "ERROR: P(u_bar) <= 0." (ex4_sedov_2d/main.f90:1447-1454) */
#define SWF_E_SWEEP_FAIL 2  /* "SWEEP FAIL: swp_num =" (ex4_sedov_2d/main.f90:1747-1752) */
#define SWF_E_PSUM 3        /* "Sweep check failed: total pressure increased after adjustment"
                               (ex6_shock_diffraction/main.f90:1418-1426) */
#define SWF_E_INVALID (-1)
#define SWF_E_HIP (-2)
#define SWF_E_NOMEM (-3)
#define SWF_E_NODEVICE (-4)

typedef struct swf_config {
  int32_t nx, ny;           /* interior cells (total_x, total_y) */
  double xlo, ylo, dx, dy;  /* grid origin and spacing as the driver computed them */
  double gamma;             /* gamma1 = gamma - 1 is formed in binary64 */
  double cfl;
  double eps;               /* positivity threshold, 1d-13 */
  double eps2;              /* velocity cut-off, 1d-8 */
  double delta;             /* density sweep threshold, 1d-12 */
  int32_t bc;               /* SWF_BC_* */
  int32_t use_limiter;      /* 2 = sweeping post-process, 1 = Zhang-Shu flux limiter
                               (den_count / pre_count in swf_sweep_stats), 0 = none */
  int32_t swp_num;          /* sweep iteration cap */
  /* variant switches: 0 reproduces ex4_sedov_2d; the ex3/ex5 programs set: */
  int32_t alpha_in_scheme;   /* ex3/ex5: scheme multiplies by alphamax */
  int32_t roe_eps_double;    /* ex3/ex5: Roe weight / sound speed use 1.e-15 (binary64) */
  int32_t ydens_plain;       /* ex3/ex5: y-pass 1./den instead of 1./(den+1d-15) */
  int32_t eps2_quad_literal; /* ex3/ex5: fflux tests rho < 1d-8 (binary128 literal) */
  int32_t findalpha_ex3;     /* ex3 findalpha expression */
  int32_t dt_rule_dx53;      /* ex3: dt = dx**(5/3)*cfl/alpha, no clamp to time_terminal */
  int32_t poscheck_le;       /* ex3: checkpositivity with <= */
  int32_t success_ge;        /* ex5: sweep success with >= */
  int32_t pbar_check;        /* ex4: stop on P(u_bar) <= 0 */
  int32_t fail_stop;         /* ex4: SWEEP FAIL stops (else print and continue) */
  double jet_rho, jet_u, amb_rho, amb_p, jet_halfwidth; /* jet boundary data */
  /* the step diffraction program (ex6_shock_diffraction/main.f90), on one domain: */
  int32_t step_total;        /* > 0: the step of step_total x 6*step_total cells, its walls
                                reflected inside wenofd (:188-197, :417-431); bc SWF_BC_STEP */
  int32_t alpha_lines;       /* wenofd splits each row (x) / column (y) with its own alpha max
                                (:217-219, :449-451) */
  int32_t weno_eps12;        /* the program's weno5.f90: 1d-12 in the weights (:104-106);
                                alpha_lines and weno_eps12 are set together */
  int32_t psum_check;        /* sweep stops (SWF_E_PSUM) when an order-I pair adjustment raises
                                the pair's total pressure (:1418-1426) */
  double in_rho, in_u, in_p; /* post-shock inflow state (dle, ule, ple of :1104-1109) */
  int32_t precision;         /* SWF_FP64_EXACT / SWF_FP64 / SWF_FP32_BULK */
  int32_t device;            /* HIP device ordinal */
} swf_config;

typedef struct swf_sweep_stats {
  int64_t sweep_count, den_count, pre_count;
  int64_t total_sweep, max_sweep, total_neg, max_neg, total_pos, max_pos;
  int64_t stage_sweeps; /* RK stages in which the sweep ran */
  int64_t steps;        /* steps advanced on this context */
} swf_sweep_stats;

typedef struct swf_ctx swf_ctx;

int swf_create(const swf_config *cfg, swf_ctx **out);
void swf_destroy(swf_ctx *ctx);

/* The state's alpha (findalpha over every cell of the frame) is taken at upload
 * and again by swf_bc; the first stage's dt follows from it. */
int swf_upload(swf_ctx *ctx, const double *u);
int swf_download(swf_ctx *ctx, double *u);

/* One SSP-RK3 step at `time` (dt decided on the device); dt_out may be NULL. */
int swf_step(swf_ctx *ctx, double time, double tfinal, double *dt_out);
/* Up to max_steps steps while time <= tfinal - 1d-10; *time_inout advances,
 * *steps_out receives the number of steps taken.  Steps are enqueued without
 * host synchronisation (device-side dt and sweep gates). */
int swf_advance(swf_ctx *ctx, double *time_inout, double tfinal, int64_t max_steps, int64_t *steps_out);

int swf_bc(swf_ctx *ctx);
int swf_alpha_max(swf_ctx *ctx, double *alpha_out);
int swf_rhs(swf_ctx *ctx, double alphamax, double *rhs_out);
int swf_check_positivity(swf_ctx *ctx, int32_t *postag_out);
int swf_sweep(swf_ctx *ctx);

int swf_stats(swf_ctx *ctx, swf_sweep_stats *out);
/* per-trigger (q, negative pairs, positive neighbours); returns sweep_count */
int64_t swf_track(swf_ctx *ctx, int64_t *q, int64_t *neg, int64_t *pos, int64_t cap);

/* weno_right / weno_left of n host stencils a(-2..2) on the device; precision may
 * carry SWF_WENO_EPS12: the weights of ex6_shock_diffraction/weno5.f90 (1d-12) */
#define SWF_WENO_EPS12 0x10
int swf_weno_points(const double *a, double *right, double *left, int64_t n, int32_t precision,
                    int32_t device);

/* swf_weno_points without any device allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The device staging of
 * the stencils and results lives in one array of the library per GPU, of
 * SWF_WENO_NOALLOC_MAX stencils (3.5 MiB, held until the process ends); larger n
 * returns SWF_E_INVALID.  Results equal swf_weno_points'.  Calls on one GPU share
 * that array: calls that overlap on one GPU (from several host threads) race. */
#define SWF_WENO_NOALLOC_MAX 65536
int swf_weno_points_noalloc(const double *a, double *right, double *left, int64_t n, int32_t precision,
                            int32_t device);

/* Average duration (ms) of one launch of a hot kernel on the context's stream,
 * measured with HIP events over `nreps` back-to-back launches on the current state
 * (which: 0 = the x-flux kernel, the x-pass of `wenofd`; 1 = the stage kernel, the
 * y-pass of `wenofd` with `scheme`, the RK combination, `checkpositivity` and the
 * ghost images, run as stage 1 into a scratch state).  No counterpart in the
 * reference: measurement support. */
int swf_time_kernel(swf_ctx *ctx, int32_t which, int32_t nreps, double *avg_ms);

/*
 * Batches.  No counterpart in the reference, which runs one problem per program:
 * n independent instances (any grids and problems, one precision mode, use_limiter
 * 0 or 2) advanced together, each RK stage one launch of each stage kernel over the
 * tiles of every instance and one sweep workgroup per instance.  Each instance
 * keeps its own time step, time, sweep and statistics; its results are those of
 * the same configuration run alone (the same kernels).
 *   swf_batch_create      one swf_ctx per instance (swf_create of cfgs[i])
 *   swf_batch_instance    instance i's context: swf_upload, swf_bc, swf_download,
 *                         swf_stats and swf_track work on it; it belongs to the batch
 *   swf_batch_advance     every instance's time loop up to max_steps steps: time_inout,
 *                         tfinal, steps_out, status_out are arrays of n; returns the
 *                         first non-zero status (0 if none)
 *   swf_batch_set_schedule / swf_batch_tune
 *                         the order of all instances' tiles in a launch (0: instance by
 *                         instance, 1: round robin, 2: round robin largest first);
 *                         swf_batch_tune times the stage kernels under each (reps
 *                         launches, no effect on the states), keeps the fastest and
 *                         reports it and the three times (ms_out[3], may be NULL)
 */
typedef struct swf_batch swf_batch;
int swf_batch_create(const swf_config *cfgs, int32_t n, swf_batch **out);
void swf_batch_destroy(swf_batch *batch);
swf_ctx *swf_batch_instance(swf_batch *batch, int32_t i);
int swf_batch_advance(swf_batch *batch, double *time_inout, const double *tfinal, int64_t max_steps,
                      int64_t *steps_out, int32_t *status_out);
int swf_batch_set_schedule(swf_batch *batch, int32_t schedule);
int swf_batch_tune(swf_batch *batch, int32_t reps, int32_t *schedule_out, float *ms_out);

/*
 * The 1D program (ex2_sedov_1d/main.f90): one line of n cells with three ghost
 * cells each side, state u(3, -2:n+3) in the program's layout (component fastest).
 *   swf1d_create / swf1d_destroy   the PARAMETER block, arrays and weights (:4-36)
 *   swf1d_upload / swf1d_download  the state array u (:10)
 *   swf1d_advance                  the time loop `DO WHILE (time <= time_terminal - 1d-10)`
 *                                  around `rk3` (:38-45, :298-332), i.e. wenofd, scheme, the
 *                                  RK combination, checkpositivity and sweep_lim per stage
 *   swf1d_stats / swf1d_track      the statistics printed by `output` (:487-511)
 * Semantics of that program: eigenvectors of each cell's own state; no boundary
 * conditions (ghost cells keep their values apart from the RK combination and the
 * sweep); the sweep walks every cell including the ghosts, its density pass runs
 * unconditionally with delta = 1d-13; a sweep that reaches swp_num is silent (no
 * error code); dt = dx*cfl/alphamax is not clamped to tfinal.  use_limiter 1 (the
 * flux limiter branch of that program's wenofd) is not provided: SWF_E_INVALID.
 * One workgroup runs the whole loop; n + 6 <= 2048 (and the device's LDS per
 * workgroup must hold 72*(n+6) + 1024 bytes).
 */
typedef struct swf1d_config {
  int32_t n;            /* interior cells (ex2: 800) */
  double dx;            /* cell size (ex2: 4/n) */
  double gamma, cfl;    /* gamma1 = gamma - 1 is formed in binary64 */
  double eps;           /* positivity threshold, 1d-13 */
  int32_t use_limiter;  /* 2 = sweeping post-process, 0 = none */
  int32_t swp_num;      /* sweep iteration cap */
  int32_t precision;    /* SWF_FP64_EXACT / SWF_FP64 / SWF_FP32_BULK */
  int32_t device;       /* HIP device ordinal */
} swf1d_config;

typedef struct swf1d_ctx swf1d_ctx;
int swf1d_create(const swf1d_config *cfg, swf1d_ctx **out);
void swf1d_destroy(swf1d_ctx *ctx);
int swf1d_upload(swf1d_ctx *ctx, const double *u);
int swf1d_download(swf1d_ctx *ctx, double *u);
/* Up to max_steps steps while time <= tfinal - 1d-10; *time_inout advances and
 * *steps_out (may be NULL) receives the steps taken.  Synchronous. */
int swf1d_advance(swf1d_ctx *ctx, double *time_inout, double tfinal, int64_t max_steps, int64_t *steps_out);
int swf1d_stats(swf1d_ctx *ctx, swf_sweep_stats *out);
/* per-trigger (q, negative pairs, positive neighbours); returns sweep_count */
int64_t swf1d_track(swf1d_ctx *ctx, int64_t *q, int64_t *neg, int64_t *pos, int64_t cap);
/* Device time (ms, HIP events) of the kernels of the last swf1d_advance.  No
 * counterpart in the reference: measurement support. */
int swf1d_last_ms(swf1d_ctx *ctx, double *ms);

/*
 * The detonation program, ex7_detonation/main.f90, on one domain (cut_x = cut_y = 1),
 * FP64-exact: a fifth conserved variable rho*Y (unburnt mass fraction), pressure
 * gamma1*(E - |m|^2/(2 rho) - q_const*rho*Y), the source -k_const*rho*Y*exp(-t_const*rho/p)
 * in the rho*Y equation, the 5x5 characteristic system in wenofd, per-line alphas, the
 * step's walls written into u inside wenofd, its bc and its sweep.  Entry points:
 *   swf7_create / swf7_destroy   the PARAMETER block and arrays (:11-56)
 *   swf7_upload / swf7_download  u(5,-2:nx+3,-2:ny+3) (:20), Fortran layout
 *   swf7_bc                      `bc` (:915-1071), as `initdata` ends (:1259)
 *   swf7_advance                 the time loop `DO WHILE (time <= time_terminal - 1d-10)`
 *                                with `rk` (:97-102, :810-888)
 *   swf7_sweep                   `sweep_lim` with `scalar_sweep` (:1433-1806) on the current
 *                                state, as `rk` calls it when checkpositivity flags a stage
 *   swf7_stats / swf7_track      the statistics printed by `output` (:1364-1393)
 * Return codes as above; SWF_E_ORDER2_EDGE when the program would read its sweep order
 * array at position 0 (an order-II forward pass flagging the first cell, :1619), which
 * has no defined result.
 */
#define SWF_E_ORDER2_EDGE 4

typedef struct swf7_config {
  int32_t total;  /* nx = ny = 5*total */
  int32_t step;   /* the step spans columns 1..total-1 and rows 1..step*total-1 (2) */
  double dx, dy;  /* xlen/nx, ylen/ny as initdata forms them */
  double gamma, q_const, t_const, k_const;  /* 1.2, 50, 50, 2566.4 */
  double eps, eps2, cfl;                    /* 1d-13, 1d-8, 0.88 */
  int32_t swp_num;                          /* 20000 */
  int32_t device;
} swf7_config;

typedef struct swf7_ctx swf7_ctx;

int swf7_create(const swf7_config *cfg, swf7_ctx **out);
void swf7_destroy(swf7_ctx *ctx);
int swf7_upload(swf7_ctx *ctx, const double *u);
int swf7_download(swf7_ctx *ctx, double *u);
int swf7_bc(swf7_ctx *ctx);
int swf7_advance(swf7_ctx *ctx, double *time_inout, double tfinal, int64_t max_steps, int64_t *steps_out);
int swf7_sweep(swf7_ctx *ctx);
int swf7_stats(swf7_ctx *ctx, swf_sweep_stats *out);
/* per sweep (q, negative pairs, positive neighbours); returns sweep_count */
int64_t swf7_track(swf7_ctx *ctx, int64_t *q, int64_t *neg, int64_t *pos, int64_t cap);

int swf_time_continue(double time, double tfinal);
int swf_sync(swf_ctx *ctx);
const char *swf_strerror(int code);
/* last HIP error text of this thread ("" if none) */
const char *swf_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
