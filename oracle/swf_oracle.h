/*
 * swf_oracle.h - CPU restatement of the reference sweeping FD-WENO hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the oracle that the HIP library is checked
 * against.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it; the product library never links or calls it.
 *
 * It restates, operation by operation and in the reference's evaluation order,
 *   ex4_sedov_2d/main.f90   (Sedov, reflective/outflow, binary128 Roe/cm/y-density sites)
 *   ex3_vortex_accuracy_2d/main.f90 (periodic vortex, alpha applied in `scheme`)
 *   ex5_mach2000_jet/main.f90 (jet inflow, `>=` sweep success test)
 *   ex4_sedov_2d/weno5.f90:86-146 (weno_left / weno_right)
 *   ex2_sedov_1d/main.f90   (the 1D program: swo1d_*)
 * Every `d`-exponent literal of the reference is REAL(16) under -fdefault-real-8,
 * so those sites are evaluated here in __float128 and rounded once to double.
 *
 * State layout (the reference's, and the drop-in boundary's): Fortran
 * u(4, -2:nx+3, -2:ny+3), column-major, component fastest.
 */
#ifndef SWF_ORACLE_H
#define SWF_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

enum swo_bc { SWO_BC_SEDOV = 0, SWO_BC_PERIODIC = 1, SWO_BC_JET_HALF = 2, SWO_BC_JET_FULL = 3, SWO_BC_STEP = 4 };

typedef struct swo_config {
    int32_t nx, ny;
    double xlo, ylo, dx, dy;
    double gamma;
    double cfl;
    double eps;      /* positivity threshold, 1d-13 */
    double eps2;     /* velocity cut-off in fflux/findalpha, 1d-8 */
    double delta;    /* density sweep threshold, 1d-12 */
    int32_t bc;      /* enum swo_bc */
    int32_t use_limiter;  /* 2 = sweeping post-process, 1 = Zhang-Shu flux limiter, 0 = none */
    int32_t swp_num;
    /* variant switches (0 = ex4 behaviour) */
    int32_t alpha_in_scheme;   /* ex3/ex5: fluxes unscaled in wenofd, scheme multiplies by alphamax */
    int32_t roe_eps_double;    /* ex3/ex5: Roe/cm use the double literal 1.e-15 */
    int32_t ydens_plain;       /* ex3/ex5: y-pass t0 = 1./den (ex4: 1./(den+1d-15)) */
    int32_t eps2_quad_literal; /* ex3/ex5: fflux tests rho < 1d-8 (binary128 literal) */
    int32_t findalpha_ex3;     /* ex3 findalpha form */
    int32_t dt_rule_dx53;      /* ex3: dt = dx**(5/3)*cfl/alpha, no clamp */
    int32_t poscheck_le;       /* ex3: checkpositivity uses <= */
    int32_t success_ge;        /* ex5: sweep success uses >= */
    int32_t pbar_check;        /* ex4: stop when P(u_bar) <= 0 */
    int32_t fail_stop;         /* ex4: SWEEP FAIL stops the run */
    /* jet boundary data (ex5 bc) */
    double jet_rho, jet_u, amb_rho, amb_p, jet_halfwidth;
    /* the step diffraction program (ex6_shock_diffraction/main.f90) */
    int32_t step_total;  /* > 0: a step of step_total x 6*step_total cells at the lower left,
                            its walls reflected inside wenofd (SWO_BC_STEP) */
    int32_t alpha_lines; /* wenofd splits with the alpha max of each row (x) / column (y) */
    int32_t weno_eps12;  /* weno5.f90 of ex6: 1d-12 in place of 1d-6 in the weights */
    int32_t psum_check;  /* sweep stops when an order-I pair adjustment raises P_i + P_p */
    double in_rho, in_u, in_p; /* post-shock inflow state (ex6 initdata) */
} swo_config;

typedef struct swo_stats {
    int64_t sweep_count, den_count, pre_count;
    int64_t total_sweep, max_sweep, total_neg, max_neg, total_pos, max_pos;
    int64_t stage_sweeps;       /* stages in which the sweep ran */
} swo_stats;

typedef struct swo_ctx swo_ctx;

/* status codes (shared meaning with the HIP library) */
#define SWO_OK 0
#define SWO_PBAR_NONPOSITIVE 1
#define SWO_SWEEP_FAIL 2
#define SWO_PSUM_INCREASE 3

swo_ctx *swo_create(const swo_config *cfg);
void swo_destroy(swo_ctx *c);
void swo_set_state(swo_ctx *c, const double *u);
void swo_get_state(const swo_ctx *c, double *u);
/* One SSP-RK3 step (reference `rk`); dt_out receives dt. */
int swo_step(swo_ctx *c, double time, double tfinal, double *dt_out);
/* Apply the boundary condition (reference `bc`). */
void swo_bc(swo_ctx *c);
/* findalpha over all cells including ghosts */
double swo_alpha_max(swo_ctx *c);
/* spatial operator: RHS(4,-2:nx+3,-2:ny+3) for the current state and alphamax */
void swo_rhs(swo_ctx *c, double alphamax, double *rhs_out);
/* checkpositivity flag of the current state */
int swo_check_positivity(const swo_ctx *c);
/* sweep_lim on the current interior (includes scalar_sweep); returns status */
int swo_sweep(swo_ctx *c);
void swo_get_stats(const swo_ctx *c, swo_stats *s);
int64_t swo_track(const swo_ctx *c, int64_t *q, int64_t *neg, int64_t *pos, int64_t cap);
/* Reference time-loop test `time <= time_terminal - 1d-10` in binary128. */
int swo_time_continue(double time, double tfinal);

/* weno5.f90 point reconstructions (n stencils of 5 values a(-2..2)) */
void swo_weno_right_n(const double *a, double *out, int64_t n);
void swo_weno_left_n(const double *a, double *out, int64_t n);
/* the same with weno5.f90 of ex6 (1d-12 in the weights) */
void swo_weno_eps12_n(const double *a, double *right, double *left, int64_t n);
/* nod_lim (ex4_sedov_2d/main.f90:1359-1396); returns 1 for the positive branch, 0 negative */
int swo_nod_lim(const double *v1, const double *v2, const double *vbar, double gamma1,
                double eps, double *out);

/* initial data (driver side) */
void swo_init_sedov(const swo_config *cfg, double e_corner_num, double e_amb, double *u);
void swo_init_vortex(const swo_config *cfg, double mag, double t, double *u);
void swo_vortex_exact(const swo_config *cfg, double mag, double t, double *u);
void swo_init_jet(const swo_config *cfg, double *u);
/* ex6 initdata (:1068-1151) on a single domain: pre-shock gas (gamma, 0, 0, 1/gamma1)
 * and the post-shock state of a Mach `mach` shock for i <= total/2, j >= 6*total.
 * Writes in_rho/in_u/in_p into *cfg (the inflow state of SWO_BC_STEP). */
void swo_init_step(swo_config *cfg, double mach, double *u);

/* The 1D program ex2_sedov_1d/main.f90 (built with -fdefault-real-8): u(3, -2:n+3). */
typedef struct swo1d_config {
    int32_t n;
    double xlo, dx, gamma, cfl, eps;
    int32_t use_limiter;  /* 2 = sweeping (scalar + pressure sweeps), 0 = none */
    int32_t swp_num;
} swo1d_config;
typedef struct swo1d_stats {
    int64_t sweep_count, den_count, pre_count, total_sweep, max_sweep, total_neg, max_neg, total_pos, max_pos,
        steps;
} swo1d_stats;
typedef struct swo1d_ctx swo1d_ctx;
swo1d_ctx *swo1d_create(const swo1d_config *cfg);
void swo1d_destroy(swo1d_ctx *ctx);
void swo1d_set_state(swo1d_ctx *ctx, const double *u);
void swo1d_get_state(const swo1d_ctx *ctx, double *u);
void swo1d_get_stats(const swo1d_ctx *ctx, swo1d_stats *out);
int64_t swo1d_track(const swo1d_ctx *ctx, int64_t *q, int64_t *neg, int64_t *pos, int64_t cap);
void swo1d_init_sedov(const swo1d_config *cfg, int shipped_energy, double e_amb, double *u);
int swo1d_run(swo1d_ctx *ctx, double *time_inout, double tfinal, int64_t max_steps, int64_t *steps_out);

/* The detonation program's sweep (ex7_detonation/main.f90 :1399-1862) on its state
 * u(5, -2:n+3, -2:n+3), n = 5 total (swf_oracle_ex7.c).  counts[0..2] += sweep_count,
 * den_count, pre_count; track = (q, negative pairs, positive neighbours) of this sweep.
 * Returns 0, or the stop: 1 P(u_bar) <= 0, 2 SWEEP FAIL, 3 the pressure-sum check,
 * 4 the order-II pass read before its order arrays. */
int swo7_sweep(double *u, int32_t total, double gamma, double q_const, double eps, int32_t swp_num,
               int64_t *counts, int64_t *track);

#ifdef __cplusplus
}
#endif
#endif
