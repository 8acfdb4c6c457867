/*
 * swf_oracle.c - CPU restatement of the reference sweeping FD-WENO hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): see swf_oracle.h.
 *
 * Conventions used throughout, chosen to reproduce the reference build
 * (gfortran/flang, `-O3 -fdefault-real-8`, x86-64 SSE2, no FMA):
 *   - every double expression is written in the reference's left-to-right order,
 *     structural zeros and ones of the eigenvector matrices included, so that the
 *     sign of zero and every rounding match;
 *   - every `d`-literal site is evaluated in __float128 (REAL(16)) and rounded once
 *     to double on assignment (`D()` below);
 *   - this file is compiled with -ffp-contract=off (no FMA).
 */
#include "swf_oracle.h"

#include <math.h>
#include <quadmath.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef __float128 quad;
#define Q(x) ((quad)(x))
static inline double D(quad x) { return (double)x; }

/* Fortran MAX / MIN on two reals as lowered for non-NaN operands. */
static inline double fmax2(double a, double b) { return (b > a) ? b : a; }
static inline double fmin2(double a, double b) { return (b < a) ? b : a; }

/* ------------------------------------------------------------------------- */
/* weno5.f90 (ex4_sedov_2d/weno5.f90:3-15, 86-146)                            */
/* ------------------------------------------------------------------------- */

/* c(-1:2, 0:2), RESHAPE column-major of the 12 decimal literals (weno5.f90:3-14).
 * The literals are default REAL = REAL(8) under -fdefault-real-8. */
static const double WC[4][3] = {
    /* c(-1,0..2) */ {1.83333333333333333333333333333333, -1.16666666666666666666666666666667,
                      0.333333333333333333333333333333333},
    /* c( 0,0..2) */ {0.333333333333333333333333333333333, 0.833333333333333333333333333333333,
                      -0.166666666666666666666666666666667},
    /* c( 1,0..2) */ {-0.166666666666666666666666666666667, 0.833333333333333333333333333333333,
                      0.333333333333333333333333333333333},
    /* c( 2,0..2) */ {0.333333333333333333333333333333333, -1.16666666666666666666666666666667,
                      1.83333333333333333333333333333333},
};
#define CW(r, col) WC[(r) + 1][(col)]

/* beta(1..3), weno5.f90:100-102 and :134-136.  `a` points at average(0). */
static void weno_beta(const double *a, double beta[3]) {
    const quad q13_12 = Q(13) / Q(12); /* 13d0/12d0 folded in REAL(16) */
    const quad q025 = Q(0.25);
    quad x, y;
    /* beta(1) = 13/12*(a0-2a1+a2)**2 + 0.25*(3a0-4a1+a2)**2, all REAL(16) */
    x = (Q(a[0]) - Q(2) * Q(a[1])) + Q(a[2]);
    y = (Q(3) * Q(a[0]) - Q(4) * Q(a[1])) + Q(a[2]);
    beta[0] = D(q13_12 * (x * x) + q025 * (y * y));
    /* beta(2) = 13/12*(a-1-2a0+a1)**2 + 0.25*(a-1 - a1)**2 ; the second
     * square has no d-literal inside, so it is REAL(8) then promoted. */
    x = (Q(a[-1]) - Q(2) * Q(a[0])) + Q(a[1]);
    {
        double yd = a[-1] - a[1];
        double y2 = yd * yd;
        beta[1] = D(q13_12 * (x * x) + q025 * Q(y2));
    }
    /* beta(3) = 13/12*(a-2-2a-1+a0)**2 + 0.25*(a-2-4a-1+3a0)**2 */
    x = (Q(a[-2]) - Q(2) * Q(a[-1])) + Q(a[0]);
    y = (Q(a[-2]) - Q(4) * Q(a[-1])) + Q(3) * Q(a[0]);
    beta[2] = D(q13_12 * (x * x) + q025 * (y * y));
}

/* alpha = d/(1d-6+beta)**2 with d a REAL(8) literal, evaluated in REAL(16); the
 * weno5.f90 of ex6_shock_diffraction has 1d-12 in its place (weps). */
static inline double weno_alpha(double d, double beta, quad weps) {
    quad e = weps + Q(beta);
    return D(Q(d) / (e * e));
}

/* weno_right (weno5.f90:120-146): value at x_{i+1/2}; `a` points at average(0) */
static double weno_right(const double *a, quad weps) {
    double um[3], beta[3], al[3], ww[3];
    for (int k = 1; k <= 3; ++k) {
        int is = k - 3;
        double s = 0.0;
        for (int i = 0; i <= 2; ++i) s = s + CW(-is, i) * a[is + i];
        um[k - 1] = s;
    }
    weno_beta(a, beta);
    al[0] = weno_alpha(0.3, beta[0], weps);
    al[1] = weno_alpha(0.6, beta[1], weps);
    al[2] = weno_alpha(0.1, beta[2], weps);
    double sum = (al[0] + al[1]) + al[2];
    ww[0] = al[0] / sum;
    ww[1] = al[1] / sum;
    ww[2] = al[2] / sum;
    return (ww[0] * um[2] + ww[1] * um[1]) + ww[2] * um[0];
}

/* weno_left (weno5.f90:86-113): value at x_{i-1/2} */
static double weno_left(const double *a, quad weps) {
    double up[3], beta[3], al[3], ww[3];
    for (int k = 1; k <= 3; ++k) {
        int is = k - 3;
        double s = 0.0;
        for (int i = 0; i <= 2; ++i) s = s + CW(-is - 1, i) * a[is + i];
        up[k - 1] = s;
    }
    weno_beta(a, beta);
    al[0] = weno_alpha(0.1, beta[0], weps);
    al[1] = weno_alpha(0.6, beta[1], weps);
    al[2] = weno_alpha(0.3, beta[2], weps);
    double sum = (al[0] + al[1]) + al[2];
    ww[0] = al[0] / sum;
    ww[1] = al[1] / sum;
    ww[2] = al[2] / sum;
    return (ww[0] * up[2] + ww[1] * up[1]) + ww[2] * up[0];
}

void swo_weno_right_n(const double *a, double *out, int64_t n) {
    for (int64_t k = 0; k < n; ++k) out[k] = weno_right(a + 5 * k + 2, 1e-6Q);
}
void swo_weno_left_n(const double *a, double *out, int64_t n) {
    for (int64_t k = 0; k < n; ++k) out[k] = weno_left(a + 5 * k + 2, 1e-6Q);
}
/* ex6_shock_diffraction/weno5.f90:104-106,138-140: 1d-12 */
void swo_weno_eps12_n(const double *a, double *right, double *left, int64_t n) {
    for (int64_t k = 0; k < n; ++k) {
        right[k] = weno_right(a + 5 * k + 2, 1e-12Q);
        left[k] = weno_left(a + 5 * k + 2, 1e-12Q);
    }
}

/* ------------------------------------------------------------------------- */
/* context                                                                    */
/* ------------------------------------------------------------------------- */

struct swo_ctx {
    swo_config cfg;
    int nx, ny, nxg, nyg; /* nxg = nx+6 */
    double gamma1;
    double weight;        /* weights(i,j) = 1./(nSteps_x*nSteps_y) */
    double *u, *u0, *rhs, *fpf, *fnf, *gpf, *gnf;
    double *fp, *fn;      /* one line of split fluxes, 4*(max(nx,ny)+6) */
    double *pw, *pvx, *pvy, *ph;
    int32_t *ordI_i, *ordI_j, *ordII_i, *ordII_j;
    double *rho_t;
    /* statistics */
    int64_t sweep_count, den_count, pre_count, stage_sweeps;
    int64_t *trk_q, *trk_neg, *trk_pos;
    int64_t trk_cap;
    int64_t t_pos_count, t_neg_count;
    quad weps;            /* the WENO weights' 1d-6 (1d-12 in ex6) */
    double *alpha;        /* findalpha's alpha(i,j) (per-line alphas) */
    double *aline;        /* alpha max of one row / column */
};

#define U4(arr, c, i, j) (arr)[(c) + 4 * (((i) + 2) + C->nxg * ((j) + 2))]

static void sweeping_order(swo_ctx *C) {
    /* ex4_sedov_2d/main.f90:1836-1895 */
    int nx = C->nx, ny = C->ny;
    int64_t k = 0;
    for (int j = 1; j <= ny; ++j) {
        if (j % 2 == 1)
            for (int i = 1; i <= nx; ++i) { C->ordI_i[k] = i; C->ordI_j[k] = j; ++k; }
        else
            for (int i = nx; i >= 1; --i) { C->ordI_i[k] = i; C->ordI_j[k] = j; ++k; }
    }
    k = 0;
    for (int i = 1; i <= nx; ++i) {
        if (i % 2 == 1)
            for (int j = 1; j <= ny; ++j) { C->ordII_i[k] = i; C->ordII_j[k] = j; ++k; }
        else
            for (int j = ny; j >= 1; --j) { C->ordII_i[k] = i; C->ordII_j[k] = j; ++k; }
    }
}

swo_ctx *swo_create(const swo_config *cfg) {
    swo_ctx *C = (swo_ctx *)calloc(1, sizeof(swo_ctx));
    C->cfg = *cfg;
    C->nx = cfg->nx;
    C->ny = cfg->ny;
    C->nxg = cfg->nx + 6;
    C->nyg = cfg->ny + 6;
    C->gamma1 = cfg->gamma - 1.;
    C->weight = 1. / ((double)cfg->nx * (double)cfg->ny);
    size_t n = (size_t)4 * C->nxg * C->nyg;
    C->u = (double *)calloc(n, sizeof(double));
    C->u0 = (double *)calloc(n, sizeof(double));
    C->rhs = (double *)calloc(n, sizeof(double));
    C->fpf = (double *)calloc(n, sizeof(double));
    C->fnf = (double *)calloc(n, sizeof(double));
    C->gpf = (double *)calloc(n, sizeof(double));
    C->gnf = (double *)calloc(n, sizeof(double));
    int m = (cfg->nx > cfg->ny ? cfg->nx : cfg->ny) + 6;
    C->fp = (double *)calloc((size_t)4 * m, sizeof(double));
    C->fn = (double *)calloc((size_t)4 * m, sizeof(double));
    C->pw = (double *)calloc(m, sizeof(double));
    C->pvx = (double *)calloc(m, sizeof(double));
    C->pvy = (double *)calloc(m, sizeof(double));
    C->ph = (double *)calloc(m, sizeof(double));
    size_t N = (size_t)cfg->nx * cfg->ny;
    C->ordI_i = (int32_t *)malloc(N * sizeof(int32_t));
    C->ordI_j = (int32_t *)malloc(N * sizeof(int32_t));
    C->ordII_i = (int32_t *)malloc(N * sizeof(int32_t));
    C->ordII_j = (int32_t *)malloc(N * sizeof(int32_t));
    C->rho_t = (double *)malloc(N * sizeof(double));
    C->weps = cfg->weno_eps12 ? 1e-12Q : 1e-6Q;
    C->alpha = (double *)calloc((size_t)C->nxg * C->nyg, sizeof(double));
    C->aline = (double *)calloc((size_t)(C->nxg > C->nyg ? C->nxg : C->nyg), sizeof(double));
    C->trk_cap = 1024;
    C->trk_q = (int64_t *)calloc(C->trk_cap, sizeof(int64_t));
    C->trk_neg = (int64_t *)calloc(C->trk_cap, sizeof(int64_t));
    C->trk_pos = (int64_t *)calloc(C->trk_cap, sizeof(int64_t));
    sweeping_order(C);
    return C;
}

void swo_destroy(swo_ctx *C) {
    if (!C) return;
    free(C->u); free(C->u0); free(C->rhs);
    free(C->fpf); free(C->fnf); free(C->gpf); free(C->gnf);
    free(C->fp); free(C->fn); free(C->pw); free(C->pvx); free(C->pvy); free(C->ph);
    free(C->ordI_i); free(C->ordI_j); free(C->ordII_i); free(C->ordII_j);
    free(C->rho_t);
    free(C->trk_q); free(C->trk_neg); free(C->trk_pos);
    free(C->alpha); free(C->aline);
    free(C);
}

void swo_set_state(swo_ctx *C, const double *u) {
    memcpy(C->u, u, sizeof(double) * 4 * (size_t)C->nxg * C->nyg);
}
void swo_get_state(const swo_ctx *C, double *u) {
    memcpy(u, C->u, sizeof(double) * 4 * (size_t)C->nxg * C->nyg);
}

/* ------------------------------------------------------------------------- */
/* boundary conditions                                                        */
/* ------------------------------------------------------------------------- */

static double cell_y(const swo_ctx *C, int j) { return C->cfg.ylo + j * C->cfg.dy - C->cfg.dy / 2.; }

void swo_bc(swo_ctx *C) {
    double *u = C->u;
    int nx = C->nx, ny = C->ny;
    switch (C->cfg.bc) {
    case SWO_BC_SEDOV:
        /* ex4_sedov_2d/main.f90:916-994 at cut_x = cut_y = 1.  The periodic
         * self-exchange that precedes these loops (:857-911) is overwritten in
         * every ghost cell by the four physical passes below, in this order. */
        for (int j = -2; j <= ny + 3; ++j)
            for (int i = 1; i <= 3; ++i) {
                U4(u, 0, 1 - i, j) = U4(u, 0, i, j);
                U4(u, 1, 1 - i, j) = -U4(u, 1, i, j);
                U4(u, 2, 1 - i, j) = U4(u, 2, i, j);
                U4(u, 3, 1 - i, j) = U4(u, 3, i, j);
            }
        for (int j = 1; j <= 3; ++j)
            for (int i = -2; i <= nx + 3; ++i) {
                U4(u, 0, i, 1 - j) = U4(u, 0, i, j);
                U4(u, 1, i, 1 - j) = U4(u, 1, i, j);
                U4(u, 2, i, 1 - j) = -U4(u, 2, i, j);
                U4(u, 3, i, 1 - j) = U4(u, 3, i, j);
            }
        for (int ig = 1; ig <= 3; ++ig)
            for (int j = -2; j <= ny + 3; ++j)
                for (int c = 0; c < 4; ++c) U4(u, c, nx + ig, j) = U4(u, c, nx, j);
        for (int ig = 1; ig <= 3; ++ig)
            for (int i = -2; i <= nx + 3; ++i)
                for (int c = 0; c < 4; ++c) U4(u, c, i, ny + ig) = U4(u, c, i, ny);
        break;
    case SWO_BC_PERIODIC:
        /* ex3_vortex_accuracy_2d/main.f90:653-669 */
        for (int j = -2; j <= ny + 3; ++j)
            for (int i = 1; i <= 3; ++i)
                for (int c = 0; c < 4; ++c) {
                    U4(u, c, 1 - i, j) = U4(u, c, nx + 1 - i, j);
                    U4(u, c, nx + i, j) = U4(u, c, i, j);
                }
        for (int j = 1; j <= 3; ++j)
            for (int i = -2; i <= nx + 3; ++i)
                for (int c = 0; c < 4; ++c) {
                    U4(u, c, i, 1 - j) = U4(u, c, i, ny + 1 - j);
                    U4(u, c, i, ny + j) = U4(u, c, i, j);
                }
        break;
    case SWO_BC_JET_HALF:
    case SWO_BC_JET_FULL: {
        /* ex5_mach2000_jet/main.f90:686-741 (half domain, reflective y = 0);
         * the full-domain variant replaces the reflection by outflow and the
         * inflow test y <= 0.05 by |y| < 0.05. */
        const double g1 = C->gamma1;
        for (int j = -2; j <= ny + 3; ++j) {
            double yj = cell_y(C, j);
            int in = (C->cfg.bc == SWO_BC_JET_HALF) ? (yj <= C->cfg.jet_halfwidth)
                                                    : (fabs(yj) < C->cfg.jet_halfwidth);
            for (int i = 1; i <= 3; ++i) {
                if (in) {
                    U4(u, 0, 1 - i, j) = C->cfg.jet_rho;
                    U4(u, 1, 1 - i, j) = C->cfg.jet_rho * C->cfg.jet_u;
                    U4(u, 2, 1 - i, j) = 0.;
                    U4(u, 3, 1 - i, j) = C->cfg.amb_p / g1 +
                                          0.5 * (U4(u, 1, 1 - i, j) * U4(u, 1, 1 - i, j)) / U4(u, 0, 1 - i, j);
                } else {
                    U4(u, 0, 1 - i, j) = C->cfg.amb_rho;
                    U4(u, 1, 1 - i, j) = 0.;
                    U4(u, 2, 1 - i, j) = 0.;
                    U4(u, 3, 1 - i, j) = C->cfg.amb_p / g1;
                }
            }
        }
        if (C->cfg.bc == SWO_BC_JET_HALF) {
            for (int j = 1; j <= 3; ++j)
                for (int i = -2; i <= nx + 3; ++i) {
                    for (int c = 0; c < 4; ++c) U4(u, c, i, 1 - j) = U4(u, c, i, j);
                    U4(u, 2, i, 1 - j) = -U4(u, 2, i, j);
                }
        } else {
            for (int j = 1; j <= 3; ++j)
                for (int i = -2; i <= nx + 3; ++i)
                    for (int c = 0; c < 4; ++c) U4(u, c, i, 1 - j) = U4(u, c, i, 1);
        }
        for (int ig = 1; ig <= 3; ++ig)
            for (int j = -2; j <= ny + 3; ++j)
                for (int c = 0; c < 4; ++c) U4(u, c, nx + ig, j) = U4(u, c, nx, j);
        for (int ig = 1; ig <= 3; ++ig)
            for (int i = -2; i <= nx + 3; ++i)
                for (int c = 0; c < 4; ++c) U4(u, c, i, ny + ig) = U4(u, c, i, ny);
        break;
    }
    case SWO_BC_STEP: {
        /* ex6_shock_diffraction/main.f90:888-942 on one domain (all four faces are
         * physical; the self exchange before them is overwritten in every ghost
         * cell), in this order: fixed quiescent gas (gamma, 0, 0, 1/gamma1) below and
         * to the right, zero gradient above, the post-shock inflow on the left; then
         * the loop that was meant for the step's face reads i after its loop has
         * ended (i = 1) and resets column 1 to (1, 0, 0, 1/gamma1) in rows below
         * 6*total. */
        const double g1 = C->gamma1, g = C->cfg.gamma;
        for (int jg = -2; jg <= 0; ++jg)
            for (int i = -2; i <= nx + 3; ++i) {
                U4(u, 0, i, jg) = g;
                U4(u, 1, i, jg) = 0.;
                U4(u, 2, i, jg) = 0.;
                U4(u, 3, i, jg) = 1. / g1;
            }
        for (int ig = 1; ig <= 3; ++ig)
            for (int j = -2; j <= ny + 3; ++j) {
                U4(u, 0, nx + ig, j) = g;
                U4(u, 1, nx + ig, j) = 0.;
                U4(u, 2, nx + ig, j) = 0.;
                U4(u, 3, nx + ig, j) = 1. / g1;
            }
        for (int ig = 1; ig <= 3; ++ig)
            for (int i = -2; i <= nx + 3; ++i)
                for (int c = 0; c < 4; ++c) U4(u, c, i, ny + ig) = U4(u, c, i, ny);
        const double dle = C->cfg.in_rho, ule = C->cfg.in_u, ple = C->cfg.in_p;
        for (int ig = -2; ig <= 0; ++ig)
            for (int j = -2; j <= ny + 3; ++j) {
                U4(u, 0, ig, j) = dle;
                U4(u, 1, ig, j) = ule * dle;
                U4(u, 2, ig, j) = 0.;
                U4(u, 3, ig, j) = ple / g1 + (0.5 * dle) * (ule * ule);
            }
        const int T = C->cfg.step_total;
        if (1 < T)
            for (int j = -2; j <= ny + 3; ++j)
                if (j < 6 * T) {
                    U4(u, 0, 1, j) = 1.;
                    U4(u, 1, 1, j) = 0.;
                    U4(u, 2, 1, j) = 0.;
                    U4(u, 3, 1, j) = 1. / g1;
                }
        break;
    }
    default:
        break;
    }
}

/* ------------------------------------------------------------------------- */
/* findalpha, fluxes, wenofd, scheme                                          */
/* ------------------------------------------------------------------------- */

double swo_alpha_max(swo_ctx *C) {
    const double g = C->cfg.gamma, g1 = C->gamma1, eps2 = C->cfg.eps2;
    const double *u = C->u;
    double amax = 0.;
    for (int i = -2; i <= C->nx + 3; ++i)
        for (int j = -2; j <= C->ny + 3; ++j) {
            double u1 = U4(u, 0, i, j), u2 = U4(u, 1, i, j), u3 = U4(u, 2, i, j), u4 = U4(u, 3, i, j);
            double den = u1;
            if (C->cfg.findalpha_ex3) {
                /* ex3_vortex_accuracy_2d/main.f90:549-557 */
                double pres = g1 * ((u4 - ((u2 * u2) / u1) / 2.) - ((u3 * u3) / u1) / 2.);
                double sound = sqrt((g * pres) / den);
                amax = fmax2(fmax2(amax, fabs(u2 / u1) + sound), fabs(u3 / u1) + sound);
            } else {
                /* ex4_sedov_2d/main.f90:652-683 */
                double vx, vy;
                if (den >= eps2) {
                    vx = u2 / den;
                    vy = u3 / den;
                } else {
                    vx = 0.;
                    vy = 0.;
                }
                double pres = g1 * ((u4 - (0.5 * den) * (vx * vx)) - (0.5 * den) * (vy * vy));
                double sound = sqrt((g * pres) / den);
                double a = fmax2(fabs(u2 / u1) + sound, fabs(u3 / u1) + sound);
                C->alpha[(i + 2) + (size_t)C->nxg * (j + 2)] = a;
                amax = fmax2(amax, a);
            }
        }
    return amax;
}

/* rho < eps2 test of fflux/gflux: REAL(8) parameter (ex4) or the 1d-8 literal (ex3/ex5) */
static inline int below_eps2(const swo_ctx *C, double rho) {
    if (C->cfg.eps2_quad_literal) return Q(rho) < 1e-8Q;
    return rho < C->cfg.eps2;
}

/* fflux / gflux (ex4_sedov_2d/main.f90:125-177); dir 0 = x, 1 = y */
static void euler_flux(const swo_ctx *C, const double *a, int dir, double f[4]) {
    double rho = a[0], mx = a[1], my = a[2], e = a[3];
    double vx = mx / rho, vy = my / rho;
    if (below_eps2(C, rho)) vx = 0.;
    if (below_eps2(C, rho)) vy = 0.;
    double pres = C->gamma1 * ((e - (0.5 * rho) * (vx * vx)) - (0.5 * rho) * (vy * vy));
    if (dir == 0) {
        f[0] = mx;
        f[1] = rho * (vx * vx) + pres;
        f[2] = (rho * vx) * vy;
        f[3] = (e + pres) * vx;
    } else {
        f[0] = my;
        f[1] = (rho * vx) * vy;
        f[2] = rho * (vy * vy) + pres;
        f[3] = (e + pres) * vy;
    }
}

/* primitive data for the Roe average; `ydens` selects the y-pass reciprocal */
static void roe_prims(const swo_ctx *C, const double *a, int ypass, double *w, double *vx, double *vy,
                      double *h) {
    double den = a[0], xmt = a[1], ymt = a[2], eng = a[3];
    double t0;
    if (ypass && !C->cfg.ydens_plain)
        t0 = D(Q(1.) / (Q(den) + 1e-15Q)); /* ex4_sedov_2d/main.f90:426 */
    else
        t0 = 1. / den; /* :206 */
    double vex = xmt * t0, vey = ymt * t0;
    double pre = C->gamma1 * (eng - 0.5 * (xmt * vex + ymt * vey));
    *w = sqrt(fabs(den));
    *vx = vex;
    *vy = vey;
    *h = (pre + eng) * t0;
}

static inline double roe_weight(const swo_ctx *C, double wa, double wb) {
    if (C->cfg.roe_eps_double) return wa / ((wa + wb) + 1.e-15); /* ex3 :153 */
    return D(Q(wa) / (Q(wa + wb) + 1e-15Q));                      /* ex4 :234 */
}
static inline double cm_eps(const swo_ctx *C, double s) {
    if (C->cfg.roe_eps_double) return s + 1.e-15; /* ex3 :159 */
    return D(Q(s) + 1e-15Q);                      /* ex4 :240 */
}

static inline double dot4(const double e[4], const double *f) {
    return ((e[0] * f[0] + e[1] * f[1]) + e[2] * f[2]) + e[3] * f[3];
}

/* One interface: Roe matrices at (a|b), projection of the 5-point stencils
 * fp/fn centred on `a`, 8 reconstructions, back-projection.  `fps`/`fns` point at
 * the stencil centre; stride between points is 4 doubles.  dir 0 = x, 1 = y. */
static void interface_flux(const swo_ctx *C, int dir, double wa, double wb, double vxa, double vxb,
                           double vya, double vyb, double ha, double hb, const double *fps,
                           const double *fns, double fpo[4], double fno[4]) {
    const double g1 = C->gamma1;
    double evr[4][4], evl[4][4];
    double t0 = roe_weight(C, wa, wb);
    double t1 = 1. - t0;
    double vxm = t0 * vxa + t1 * vxb;
    double vym = t0 * vya + t1 * vyb;
    double hm = t0 * ha + t1 * hb;
    double qm = 0.5 * (vxm * vxm + vym * vym);
    double cm = cm_eps(C, sqrt(fabs(g1 * (hm - qm))));
    if (dir == 0) {
        /* ex4_sedov_2d/main.f90:241-280 */
        t0 = vxm * cm;
        evr[0][0] = 1.0; evr[0][1] = 0.0; evr[0][2] = 1.0; evr[0][3] = 1.0;
        evr[1][0] = vxm - cm; evr[1][1] = 0.0; evr[1][2] = vxm; evr[1][3] = vxm + cm;
        evr[2][0] = vym; evr[2][1] = 1.0; evr[2][2] = vym; evr[2][3] = vym;
        evr[3][0] = hm - t0; evr[3][1] = vym; evr[3][2] = qm; evr[3][3] = hm + t0;
        double rcm = 1. / cm;
        double b1 = (g1 * rcm) * rcm;
        double b2 = qm * b1;
        t0 = vxm * rcm;
        t1 = b1 * vxm;
        double t2 = 0.5 * b1;
        double t3 = b1 * vym;
        evl[0][0] = 0.5 * (b2 + t0); evl[0][1] = -(0.5 * (t1 + rcm)); evl[0][2] = -(0.5 * t3); evl[0][3] = t2;
        evl[1][0] = -vym; evl[1][1] = 0.0; evl[1][2] = 1.0; evl[1][3] = 0.0;
        evl[2][0] = 1. - b2; evl[2][1] = t1; evl[2][2] = t3; evl[2][3] = -b1;
        evl[3][0] = 0.5 * (b2 - t0); evl[3][1] = -(0.5 * (t1 - rcm)); evl[3][2] = -(0.5 * t3); evl[3][3] = t2;
    } else {
        /* ex4_sedov_2d/main.f90:462-501 */
        t0 = vym * cm;
        evr[0][0] = 1.0; evr[0][1] = 0.0; evr[0][2] = 1.0; evr[0][3] = 1.0;
        evr[1][0] = vxm; evr[1][1] = 1.0; evr[1][2] = vxm; evr[1][3] = vxm;
        evr[2][0] = vym - cm; evr[2][1] = 0.0; evr[2][2] = vym; evr[2][3] = vym + cm;
        evr[3][0] = hm - t0; evr[3][1] = vxm; evr[3][2] = qm; evr[3][3] = hm + t0;
        double rcm = 1. / cm;
        double b1 = (g1 * rcm) * rcm;
        double b2 = qm * b1;
        t0 = vym * rcm;
        t1 = b1 * vym;
        double t2 = 0.5 * b1;
        double t3 = b1 * vxm;
        evl[0][0] = 0.5 * (b2 + t0); evl[0][1] = -(0.5 * t3); evl[0][2] = -(0.5 * (t1 + rcm)); evl[0][3] = t2;
        evl[1][0] = -vxm; evl[1][1] = 1.0; evl[1][2] = 0.0; evl[1][3] = 0.0;
        evl[2][0] = 1. - b2; evl[2][1] = t3; evl[2][2] = t1; evl[2][3] = -b1;
        evl[3][0] = 0.5 * (b2 - t0); evl[3][1] = -(0.5 * t3); evl[3][2] = -(0.5 * (t1 - rcm)); evl[3][3] = t2;
    }
    /* projection (:285-290), reconstruction (:294-297), back-projection (:301-304) */
    double cup[4][5], cun[4][5], ur[4], ul[4];
    for (int ig = -2; ig <= 2; ++ig)
        for (int k = 0; k < 4; ++k) {
            cup[k][ig + 2] = dot4(evl[k], fps + 4 * ig);
            cun[k][ig + 2] = dot4(evl[k], fns + 4 * ig);
        }
    for (int k = 0; k < 4; ++k) {
        ur[k] = weno_right(&cup[k][2], C->weps);
        ul[k] = weno_left(&cun[k][2], C->weps);
    }
    for (int k = 0; k < 4; ++k) {
        fpo[k] = dot4(evr[k], ur);
        fno[k] = dot4(evr[k], ul);
    }
}

/* Zhang-Shu positivity limiter of one reconstructed split flux F against the
 * split flux `ave` of its centre cell (use_limiter == 1;
 * ex4_sedov_2d/main.f90:306-412 x-pass, :527-629 y-pass; ex3 :241-330, :444-530).
 * Density first: theta from the two points F and (ave - F/12)/(11/12); then
 * pressure, with the density component of both points updated and the other
 * components as before the density step.  The ex4 y-pass clamps the pressure
 * theta of gp at 0 (:573); theta >= 0 for finite data, so the clamp is omitted.
 * Returns 1/2 in bit 0/1 when the density/pressure step changed F. */
static int zs_limit(double g1, const double ave[4], double F[4]) {
    int changed = 0;
    const double d_ave = ave[0], m_ave = ave[1], n_ave = ave[2], e_ave = ave[3];
    const double p_ave = g1 * ((e_ave - (0.5 * (m_ave * m_ave)) / d_ave) - (0.5 * (n_ave * n_ave)) / d_ave);
    double point[2][4];
    double theta = 1.;
    for (int c = 0; c < 4; ++c) {
        point[0][c] = F[c];
        point[1][c] = (ave[c] - F[c] / 12.) / (11. / 12.);
    }
    for (int ig = 0; ig < 2; ++ig) {
        const double d_l = point[ig][0];
        if (d_l < 0.) theta = fmin2(theta, fabs(d_ave - 0.) / fabs(d_ave - d_l));
    }
    if (theta < 1.) {
        F[0] = theta * F[0] + (1. - theta) * d_ave;
        changed |= 1;
    }
    theta = 1.;
    point[0][0] = F[0];
    point[1][0] = (ave[0] - F[0] / 12.) / (11. / 12.);
    for (int ig = 0; ig < 2; ++ig) {
        const double d_l = point[ig][0], m_l = point[ig][1], n_l = point[ig][2], e_l = point[ig][3];
        const double p_l = g1 * ((e_l - ((m_l * m_l) / d_l) / 2.) - ((n_l * n_l) / d_l) / 2.);
        if (p_l < 0.) theta = fmin2(theta, fabs(p_ave - 0.) / fabs(p_ave - p_l));
    }
    if (theta < 1.) {
        F[0] = theta * F[0] + (1. - theta) * d_ave;
        F[1] = theta * F[1] + (1. - theta) * m_ave;
        F[2] = theta * F[2] + (1. - theta) * n_ave;
        F[3] = theta * F[3] + (1. - theta) * e_ave;
        changed |= 2;
    }
    return changed;
}

static void zs_count(swo_ctx *C, int changed) {
    if (changed & 1) C->den_count++;
    if (changed & 2) C->pre_count++;
}

/* wenofd (ex4_sedov_2d/main.f90:179-637).  Without the flux limiter, the rows
 * and columns that `scheme` reads; with it (use_limiter == 1) the reference's
 * full ranges (x: rows 0..ny+1, y: columns 0..nx+1, cells 0..n+1), whose every
 * limited half counts in den_count / pre_count.  Fills fpf/fnf and gpf/gnf. */
/* alpha max of row j (x-pass) or column i (y-pass) over every cell of the line,
 * ghosts included, from 0 (ex6_shock_diffraction/main.f90:217-219, :449-451) */
static double line_alpha(const swo_ctx *C, int dir, int line) {
    double a = 0.;
    if (dir == 0)
        for (int i = -2; i <= C->nx + 3; ++i) a = fmax2(a, C->alpha[(i + 2) + (size_t)C->nxg * (line + 2)]);
    else
        for (int j = -2; j <= C->ny + 3; ++j) a = fmax2(a, C->alpha[(line + 2) + (size_t)C->nxg * (j + 2)]);
    return a;
}

/* The step's walls inside wenofd (ex6_shock_diffraction/main.f90:188-197 before the
 * x-pass, :417-431 before the y-pass): the three cells inside the vertical wall
 * x = total mirror the three outside it (m1 negated) on rows 0..ny+1 up to 6*total;
 * the three rows inside the top y = 6*total mirror the three above it (m2 negated)
 * in columns 1..total.  The state keeps these values (they enter the RK update). */
static void step_reflect(swo_ctx *C, int dir) {
    double *u = C->u;
    const int T = C->cfg.step_total;
    if (dir == 0) {
        for (int j = 0; j <= C->ny + 1 && j <= 6 * T; ++j)
            for (int i = 0; i <= 2; ++i) {
                for (int c = 0; c < 4; ++c) U4(u, c, T - i, j) = U4(u, c, T + 1 + i, j);
                U4(u, 1, T - i, j) = -U4(u, 1, T + 1 + i, j);
            }
    } else {
        const int j = 6 * T;
        if (j < -2 || j > C->ny + 3) return;
        for (int i = 0; i <= 2; ++i)
            for (int k = 1; k <= T; ++k) {
                for (int c = 0; c < 4; ++c) U4(u, c, k, j - i) = U4(u, c, k, j + 1 + i);
                U4(u, 2, k, j - i) = -U4(u, 2, k, j + 1 + i);
            }
    }
}

static void wenofd(swo_ctx *C, double alphamax) {
    const double *u = C->u;
    const int nx = C->nx, ny = C->ny;
    const int scale = !C->cfg.alpha_in_scheme;
    double f[4];
    const int zs = C->cfg.use_limiter == 1;
    const double amax_all = alphamax;
    if (C->cfg.step_total > 0) step_reflect(C, 0);
    /* x sweep over rows j = 1..ny (0..ny+1 with the limiter) */
    for (int j = zs ? 0 : 1; j <= (zs ? ny + 1 : ny); ++j) {
        alphamax = C->cfg.alpha_lines ? line_alpha(C, 0, j) : amax_all;
        for (int i = 0; i <= nx + 2; ++i)
            roe_prims(C, &U4(u, 0, i, j), 0, &C->pw[i], &C->pvx[i], &C->pvy[i], &C->ph[i]);
        for (int i = -2; i <= nx + 3; ++i) {
            const double *a = &U4(u, 0, i, j);
            euler_flux(C, a, 0, f);
            for (int c = 0; c < 4; ++c) {
                C->fp[4 * (i + 2) + c] = 0.5 * (a[c] + f[c] / alphamax);
                C->fn[4 * (i + 2) + c] = 0.5 * (a[c] - f[c] / alphamax);
            }
        }
        for (int i = 0; i <= nx + 1; ++i) {
            double fpo[4], fno[4];
            interface_flux(C, 0, C->pw[i], C->pw[i + 1], C->pvx[i], C->pvx[i + 1], C->pvy[i], C->pvy[i + 1],
                           C->ph[i], C->ph[i + 1], &C->fp[4 * (i + 2)], &C->fn[4 * (i + 2)], fpo, fno);
            if (zs) {
                zs_count(C, zs_limit(C->gamma1, &C->fp[4 * (i + 2)], fpo));
                zs_count(C, zs_limit(C->gamma1, &C->fn[4 * (i + 2)], fno));
            }
            for (int c = 0; c < 4; ++c) {
                U4(C->fpf, c, i, j) = scale ? alphamax * fpo[c] : fpo[c];
                U4(C->fnf, c, i, j) = scale ? alphamax * fno[c] : fno[c];
            }
        }
    }
    if (C->cfg.step_total > 0) step_reflect(C, 1);
    /* y sweep over columns i = 1..nx (0..nx+1 with the limiter) */
    for (int i = zs ? 0 : 1; i <= (zs ? nx + 1 : nx); ++i) {
        alphamax = C->cfg.alpha_lines ? line_alpha(C, 1, i) : amax_all;
        for (int j = 0; j <= ny + 2; ++j)
            roe_prims(C, &U4(u, 0, i, j), 1, &C->pw[j], &C->pvx[j], &C->pvy[j], &C->ph[j]);
        for (int j = -2; j <= ny + 3; ++j) {
            const double *a = &U4(u, 0, i, j);
            euler_flux(C, a, 1, f);
            for (int c = 0; c < 4; ++c) {
                C->fp[4 * (j + 2) + c] = 0.5 * (a[c] + f[c] / alphamax);
                C->fn[4 * (j + 2) + c] = 0.5 * (a[c] - f[c] / alphamax);
            }
        }
        for (int j = 0; j <= ny + 1; ++j) {
            double fpo[4], fno[4];
            interface_flux(C, 1, C->pw[j], C->pw[j + 1], C->pvx[j], C->pvx[j + 1], C->pvy[j], C->pvy[j + 1],
                           C->ph[j], C->ph[j + 1], &C->fp[4 * (j + 2)], &C->fn[4 * (j + 2)], fpo, fno);
            if (zs) {
                zs_count(C, zs_limit(C->gamma1, &C->fp[4 * (j + 2)], fpo));
                zs_count(C, zs_limit(C->gamma1, &C->fn[4 * (j + 2)], fno));
            }
            for (int c = 0; c < 4; ++c) {
                U4(C->gpf, c, i, j) = scale ? alphamax * fpo[c] : fpo[c];
                U4(C->gnf, c, i, j) = scale ? alphamax * fno[c] : fno[c];
            }
        }
    }
}

/* scheme (ex4_sedov_2d/main.f90:639-650; ex3_vortex_accuracy_2d/main.f90:536-546) */
static void scheme(swo_ctx *C, double alphamax) {
    const double dx = C->cfg.dx, dy = C->cfg.dy;
    memset(C->rhs, 0, sizeof(double) * 4 * (size_t)C->nxg * C->nyg);
    for (int i = 1; i <= C->nx; ++i)
        for (int j = 1; j <= C->ny; ++j)
            for (int c = 0; c < 4; ++c) {
                double X = (U4(C->fpf, c, i, j) - U4(C->fnf, c, i + 1, j)) -
                           (U4(C->fpf, c, i - 1, j) - U4(C->fnf, c, i, j));
                double Y = (U4(C->gpf, c, i, j) - U4(C->gnf, c, i, j + 1)) -
                           (U4(C->gpf, c, i, j - 1) - U4(C->gnf, c, i, j));
                double r;
                if (C->cfg.alpha_in_scheme)
                    r = -((alphamax * X) / dx) - (alphamax * Y) / dy;
                else
                    r = -(X / dx) - Y / dy;
                U4(C->rhs, c, i, j) = r;
            }
}

void swo_rhs(swo_ctx *C, double alphamax, double *rhs_out) {
    wenofd(C, alphamax);
    scheme(C, alphamax);
    memcpy(rhs_out, C->rhs, sizeof(double) * 4 * (size_t)C->nxg * C->nyg);
}

/* checkpositivity (ex4_sedov_2d/main.f90:779-801; ex3 uses <=) */
int swo_check_positivity(const swo_ctx *C) {
    const double eps = C->cfg.eps;
    const double *u = C->u;
    int postag = 0;
    for (int i = 1; i <= C->nx; ++i)
        for (int j = 1; j <= C->ny; ++j) {
            double u1 = U4(u, 0, i, j), u2 = U4(u, 1, i, j), u3 = U4(u, 2, i, j), u4 = U4(u, 3, i, j);
            double ke = (u4 - ((u3 * u3) / u1) / 2.) - ((u2 * u2) / u1) / 2.;
            if (C->cfg.poscheck_le) {
                if (u1 <= eps) postag = 1;
                if (ke <= eps) postag = 1;
            } else {
                if (u1 < eps) postag = 1;
                if (ke < eps) postag = 1;
            }
        }
    return postag;
}

/* ------------------------------------------------------------------------- */
/* sweeping post-process                                                      */
/* ------------------------------------------------------------------------- */

static inline double sweep_pressure(double g1, const double *v) {
    return g1 * (v[3] - (0.5 * (v[1] * v[1] + v[2] * v[2])) / v[0]);
}

/* nod_lim (ex4_sedov_2d/main.f90:1359-1396) */
static int nod_lim(double g1, double eps, const double *v1, const double *v2, const double *vb, double *out) {
    double p1 = g1 * (v1[3] - (0.5 * (v1[1] * v1[1] + v1[2] * v1[2])) / v1[0]);
    double p2 = g1 * (v2[3] - (0.5 * (v2[1] * v2[1] + v2[2] * v2[2])) / v2[0]);
    double t;
    int pos;
    if (p2 > eps) {
        t = (p1 - eps) / (p1 - p2);
        pos = 1;
    } else {
        double pb = g1 * (vb[3] - (0.5 * (vb[1] * vb[1] + vb[2] * vb[2])) / vb[0]);
        double t_1 = (p1 - eps) / (p1 - pb);
        double d12 = sqrt((((v2[3] - v1[3]) * (v2[3] - v1[3]) + (v2[2] - v1[2]) * (v2[2] - v1[2])) +
                           (v2[1] - v1[1]) * (v2[1] - v1[1])) +
                          (v2[0] - v1[0]) * (v2[0] - v1[0]));
        double d1b = sqrt((((vb[3] - v1[3]) * (vb[3] - v1[3]) + (vb[2] - v1[2]) * (vb[2] - v1[2])) +
                           (vb[1] - v1[1]) * (vb[1] - v1[1])) +
                          (vb[0] - v1[0]) * (vb[0] - v1[0]));
        t = fmin2((t_1 * d1b) / d12, 0.25);
        pos = 0;
    }
    for (int k = 0; k < 4; ++k) out[k] = (1. - t) * v1[k] + t * v2[k];
    return pos;
}

int swo_nod_lim(const double *v1, const double *v2, const double *vbar, double gamma1, double eps,
                double *out) {
    return nod_lim(gamma1, eps, v1, v2, vbar, out);
}

static void load4(const swo_ctx *C, const double *u, int i, int j, double v[4]) {
    for (int c = 0; c < 4; ++c) v[c] = U4(u, c, i, j);
}
static void store4(swo_ctx *C, double *u, int i, int j, const double v[4]) {
    for (int c = 0; c < 4; ++c) U4(u, c, i, j) = v[c];
}

/* scalar_sweep (ex4_sedov_2d/main.f90:1765-1833) */
static void scalar_sweep(swo_ctx *C) {
    double *u = C->u;
    const int nx = C->nx, ny = C->ny;
    const int64_t N = (int64_t)nx * ny;
    double rho_min = 1e20;
    for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) rho_min = fmin2(U4(u, 0, i, j), rho_min);
    const double delta = D(1e-12Q);
    if (rho_min < delta) {
        double *rt = C->rho_t; /* rho_t(i,j) at (i-1) + nx*(j-1) */
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) rt[(i - 1) + (int64_t)nx * (j - 1)] = U4(u, 0, i, j);
        const double w = C->weight;
#define RT(k) rt[(C->ordI_i[k] - 1) + (int64_t)nx * (C->ordI_j[k] - 1)]
        for (int64_t k = 0; k < N - 1; ++k) {
            double v_i = RT(k), v_ip1 = RT(k + 1);
            if (v_i < delta) {
                RT(k + 1) = v_ip1 + (w / w) * (v_i - delta);
                RT(k) = delta;
            }
        }
        for (int64_t k = N - 1; k >= 1; --k) {
            double v_i = RT(k), v_im1 = RT(k - 1);
            if (v_i < delta) {
                RT(k - 1) = v_im1 + (w / w) * (v_i - delta);
                RT(k) = delta;
            }
        }
#undef RT
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) U4(u, 0, i, j) = rt[(i - 1) + (int64_t)nx * (j - 1)];
        C->den_count++;
    }
}

static void track_push(swo_ctx *C, int64_t q, int64_t neg, int64_t pos) {
    if (C->sweep_count > C->trk_cap) {
        int64_t nc = C->trk_cap * 2;
        C->trk_q = (int64_t *)realloc(C->trk_q, nc * sizeof(int64_t));
        C->trk_neg = (int64_t *)realloc(C->trk_neg, nc * sizeof(int64_t));
        C->trk_pos = (int64_t *)realloc(C->trk_pos, nc * sizeof(int64_t));
        memset(C->trk_q + C->trk_cap, 0, (nc - C->trk_cap) * sizeof(int64_t));
        memset(C->trk_neg + C->trk_cap, 0, (nc - C->trk_cap) * sizeof(int64_t));
        memset(C->trk_pos + C->trk_cap, 0, (nc - C->trk_cap) * sizeof(int64_t));
        C->trk_cap = nc;
    }
    C->trk_q[C->sweep_count - 1] = q;
    C->trk_neg[C->sweep_count - 1] = neg;
    C->trk_pos[C->sweep_count - 1] = pos;
}

/* one directional pass of sweep_lim over an order (:1468-1586).  check: ex6's test
 * of each pair adjustment (ex6_shock_diffraction/main.f90:1418-1426,1477-1485 in the
 * order-I passes): P(v_i) + P(v_p) > P(v_i~) + P(v_p~) + eps stops the program. */
static int sweep_pass(swo_ctx *C, const int32_t *oi, const int32_t *oj, int forward, const double ub[4],
                      int check) {
    double *u = C->u;
    const double g1 = C->gamma1, eps = C->cfg.eps, w = C->weight;
    const int64_t N = (int64_t)C->nx * C->ny;
    double vi[4], vp[4], vit[4], vpt[4];
    for (int64_t s = 0; s < N - 1; ++s) {
        int64_t k = forward ? s : (N - 1 - s);
        int64_t p = forward ? k + 1 : k - 1;
        load4(C, u, oi[k], oj[k], vi);
        double pres = g1 * (vi[3] - (0.5 * (vi[1] * vi[1] + vi[2] * vi[2])) / vi[0]);
        if (pres <= eps || vi[0] <= eps) {
            load4(C, u, oi[p], oj[p], vp);
            if (nod_lim(g1, eps, vi, vp, ub, vit))
                C->t_pos_count++;
            else
                C->t_neg_count++;
            for (int c = 0; c < 4; ++c) vpt[c] = vp[c] + (w / w) * (vi[c] - vit[c]);
            store4(C, u, oi[k], oj[k], vit);
            store4(C, u, oi[p], oj[p], vpt);
            if (check && pres + sweep_pressure(g1, vp) > (sweep_pressure(g1, vit) + sweep_pressure(g1, vpt)) + eps)
                return SWO_PSUM_INCREASE;
        }
    }
    return SWO_OK;
}

/* sweep_lim (ex4_sedov_2d/main.f90:1401-1761) */
int swo_sweep(swo_ctx *C) {
    double *u = C->u;
    const int nx = C->nx, ny = C->ny;
    const double g1 = C->gamma1, eps = C->cfg.eps, w = C->weight;
    C->stage_sweeps++;
    scalar_sweep(C);
    double ra = 0.0, m1a = 0.0, m2a = 0.0, ea = 0.0;
    for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) {
            ra = ra + w * U4(u, 0, i, j);
            m1a = m1a + w * U4(u, 1, i, j);
            m2a = m2a + w * U4(u, 2, i, j);
            ea = ea + w * U4(u, 3, i, j);
        }
    double ub[4] = {ra, m1a, m2a, ea};
    if (C->cfg.pbar_check) {
        if (g1 * (ub[3] - (0.5 * (ub[1] * ub[1] + ub[2] * ub[2])) / ub[0]) <= 0.) return SWO_PBAR_NONPOSITIVE;
    }
    C->sweep_count++;
    C->t_pos_count = 0;
    C->t_neg_count = 0;
    track_push(C, 0, 0, 0);
    for (int q = 1; q <= C->cfg.swp_num; ++q) {
        if (q % 2 == 1) {
            int rc = sweep_pass(C, C->ordI_i, C->ordI_j, 1, ub, C->cfg.psum_check);
            if (rc) return rc;
            rc = sweep_pass(C, C->ordI_i, C->ordI_j, 0, ub, C->cfg.psum_check);
            if (rc) return rc;
        } else {
            sweep_pass(C, C->ordII_i, C->ordII_j, 1, ub, 0);
            sweep_pass(C, C->ordII_i, C->ordII_j, 0, ub, 0);
        }
        double rho_min = 1e20, p_min = 1e20;
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                double v[4];
                load4(C, u, i, j, v);
                p_min = fmin2(sweep_pressure(g1, v), p_min);
                rho_min = fmin2(v[0], rho_min);
            }
        C->pre_count++;
        int ok = C->cfg.success_ge ? (p_min >= eps && rho_min >= eps) : (p_min > eps && rho_min > eps);
        if (ok) {
            track_push(C, q, C->t_neg_count, C->t_pos_count);
            return SWO_OK;
        }
        if (q == C->cfg.swp_num) {
            if (C->cfg.fail_stop) return SWO_SWEEP_FAIL;
        }
    }
    return SWO_OK;
}

/* ------------------------------------------------------------------------- */
/* rk (ex4_sedov_2d/main.f90:685-777)                                         */
/* ------------------------------------------------------------------------- */

static int post_stage(swo_ctx *C) {
    int rc = SWO_OK;
    if (swo_check_positivity(C) && C->cfg.use_limiter == 2) rc = swo_sweep(C);
    if (rc != SWO_OK) return rc;
    swo_bc(C);
    return SWO_OK;
}

int swo_step(swo_ctx *C, double time, double tfinal, double *dt_out) {
    const size_t n = (size_t)4 * C->nxg * C->nyg;
    double *u = C->u, *u0 = C->u0, *rhs = C->rhs;
    double am = swo_alpha_max(C);
    wenofd(C, am);
    double dt;
    if (C->cfg.dt_rule_dx53) {
        dt = (pow(C->cfg.dx, 5. / 3.) * C->cfg.cfl) / am;
    } else {
        double dtx = (C->cfg.dx * C->cfg.cfl) / am;
        double dty = (C->cfg.dy * C->cfg.cfl) / am;
        dt = fmin2(fmin2(dtx, dty), tfinal - time);
    }
    if (dt_out) *dt_out = dt;
    memcpy(u0, u, n * sizeof(double));
    scheme(C, am);
    for (size_t k = 0; k < n; ++k) u[k] = u[k] + dt * rhs[k];
    int rc = post_stage(C);
    if (rc) return rc;

    am = swo_alpha_max(C);
    wenofd(C, am);
    scheme(C, am);
    for (size_t k = 0; k < n; ++k) {
        double x = u[k] + dt * rhs[k];
        u[k] = D(Q(0.75) * Q(u0[k]) + Q(0.25) * Q(x));
    }
    rc = post_stage(C);
    if (rc) return rc;

    am = swo_alpha_max(C);
    wenofd(C, am);
    scheme(C, am);
    for (size_t k = 0; k < n; ++k) {
        double x = u[k] + dt * rhs[k];
        u[k] = D((Q(u0[k]) + Q(2) * Q(x)) / Q(3));
    }
    return post_stage(C);
}

int swo_time_continue(double time, double tfinal) { return Q(time) <= Q(tfinal) - 1e-10Q; }

void swo_get_stats(const swo_ctx *C, swo_stats *s) {
    memset(s, 0, sizeof(*s));
    s->sweep_count = C->sweep_count;
    s->den_count = C->den_count;
    s->pre_count = C->pre_count;
    s->stage_sweeps = C->stage_sweeps;
    for (int64_t k = 0; k < C->sweep_count; ++k) {
        s->total_sweep += C->trk_q[k];
        s->total_neg += C->trk_neg[k];
        s->total_pos += C->trk_pos[k];
        if (C->trk_q[k] > s->max_sweep) s->max_sweep = C->trk_q[k];
        if (C->trk_neg[k] > s->max_neg) s->max_neg = C->trk_neg[k];
        if (C->trk_pos[k] > s->max_pos) s->max_pos = C->trk_pos[k];
    }
}

int64_t swo_track(const swo_ctx *C, int64_t *q, int64_t *neg, int64_t *pos, int64_t cap) {
    int64_t n = C->sweep_count < cap ? C->sweep_count : cap;
    for (int64_t k = 0; k < n; ++k) {
        q[k] = C->trk_q[k];
        neg[k] = C->trk_neg[k];
        pos[k] = C->trk_pos[k];
    }
    return C->sweep_count;
}

/* ------------------------------------------------------------------------- */
/* initial data                                                               */
/* ------------------------------------------------------------------------- */

#define UI(c, i, j) u[(c) + 4 * (((i) + 2) + (cfg->nx + 6) * ((j) + 2))]

/* ex4_sedov_2d/main.f90:1121-1189 at cut_x = cut_y = 1 (ghosts filled by swo_bc) */
void swo_init_sedov(const swo_config *cfg, double e_corner_num, double e_amb, double *u) {
    const double dx = cfg->dx, dy = cfg->dy;
    for (int j = -2; j <= cfg->ny + 3; ++j)
        for (int i = -2; i <= cfg->nx + 3; ++i) {
            UI(0, i, j) = 1.;
            UI(1, i, j) = 0.;
            UI(2, i, j) = 0.;
            UI(3, i, j) = e_amb;
        }
    for (int j = 1; j <= cfg->ny; ++j)
        for (int i = 1; i <= cfg->nx; ++i) {
            double gx = cfg->xlo + dx * (i - 0.5);
            double gy = cfg->ylo + dy * (j - 0.5);
            if (fabs(gx - 0.0) <= dx / 2.0 && fabs(gy - 0.0) <= dy / 2.0) UI(3, i, j) = e_corner_num / (dx * dy);
        }
}

/* ex3_vortex_accuracy_2d/main.f90:699-718 */
static void vortex_exact(double gamma, double mag, double x, double y, double t, double e[4]) {
    const double g1 = gamma - 1.;
    const double pi = 3.14159265358979323846264338327950;
    double r = sqrt(((x - 5.) - t) * ((x - 5.) - t) + ((y - 5.) - t) * ((y - 5.) - t));
    e[0] = pow(1. - ((((g1 * (mag * mag)) / 8.) / gamma) / (pi * pi)) * exp(1. - r * r), 1. / g1);
    e[1] = e[0] * (1. - (((mag / 2.) / pi) * exp((1. - r * r) / 2.)) * (y - 5.));
    e[2] = e[0] * (1. + (((mag / 2.) / pi) * exp((1. - r * r) / 2.)) * (x - 5.));
    e[3] = (pow(e[0], gamma) / g1 + (0.5 * (e[1] * e[1])) / e[0]) + (0.5 * (e[2] * e[2])) / e[0];
}

void swo_vortex_exact(const swo_config *cfg, double mag, double t, double *u) {
    for (int j = -2; j <= cfg->ny + 3; ++j)
        for (int i = -2; i <= cfg->nx + 3; ++i) {
            double x = cfg->xlo + i * cfg->dx - cfg->dx / 2.;
            double y = cfg->ylo + j * cfg->dy - cfg->dy / 2.;
            double e[4];
            vortex_exact(cfg->gamma, mag, x, y, t, e);
            for (int c = 0; c < 4; ++c) UI(c, i, j) = e[c];
        }
}

void swo_init_vortex(const swo_config *cfg, double mag, double t, double *u) {
    swo_vortex_exact(cfg, mag, t, u);
}

/* ex5_mach2000_jet/main.f90:743-768 (ghosts filled by swo_bc) */
void swo_init_jet(const swo_config *cfg, double *u) {
    const double g1 = cfg->gamma - 1.;
    for (int j = -2; j <= cfg->ny + 3; ++j)
        for (int i = -2; i <= cfg->nx + 3; ++i) {
            UI(0, i, j) = cfg->amb_rho;
            UI(1, i, j) = 0.;
            UI(2, i, j) = 0.;
            UI(3, i, j) = cfg->amb_p / g1;
        }
}

/* ex6_shock_diffraction/main.f90:1068-1151 at cut_x = cut_y = 1: the Rankine-Hugoniot
 * state behind a shock of Mach xm1 = mach into gas at rest (dri = gamma, pri = 1),
 * every expression in the program's order */
void swo_init_step(swo_config *cfg, double mach, double *u) {
    const int nx = cfg->nx, ny = cfg->ny, nxg = nx + 6, T = cfg->step_total;
    const double gamma = cfg->gamma, g1 = gamma - 1.;
    const double ss = mach, dri = gamma, pri = 1.;
    const double ach2 = (((0.0 - ss) * (0.0 - ss)) / gamma / pri) * dri;
    const double ple = pri * (((gamma * ach2) / (gamma + 1.)) * 2. - (gamma - 1.) / (gamma + 1.));
    const double dle = (((dri * (gamma + 1.)) / 2.) * ach2) / (((gamma - 1.) / 2.) * ach2 + 1.);
    const double ule = ((0.0 - ss) / dle) * dri + ss;
    cfg->in_rho = dle;
    cfg->in_u = ule;
    cfg->in_p = ple;
    for (int j = -2; j <= ny + 3; ++j)
        for (int i = -2; i <= nx + 3; ++i) {
            double *a = &u[4 * ((i + 2) + (size_t)nxg * (j + 2))];
            if (i <= T / 2 && j >= T * 6) {
                a[0] = dle;
                a[1] = ule * dle;
                a[2] = 0.;
                a[3] = ple / g1 + (0.5 * dle) * (ule * ule);
            } else {
                a[0] = dri;
                a[1] = 0.;
                a[2] = 0.;
                a[3] = pri / g1;
            }
        }
}

/* ========================================================================= */
/* The 1D program (ex2_sedov_1d/main.f90), built with -fdefault-real-8: the   */
/* state u(3, -2:n+3), local (cell) eigenvectors, no boundary conditions (the */
/* ghost cells keep their initial values: RHS is 0 there), sweep over every   */
/* cell including the ghosts, unconditional scalar (density) sweep, no stop   */
/* on a sweep failure.                                                         */
/* ========================================================================= */

struct swo1d_ctx {
    swo1d_config cfg;
    int n, N;            /* interior cells; cells with ghosts (n + 6) */
    double gamma1, weight, alphamax, dt;
    int postag;
    double *u, *u0, *rhs, *hp, *hn, *hpf, *hnf, *flus;
    swo1d_stats st;
    int64_t *trk_q, *trk_n, *trk_p, trk_cap;
    int64_t t_pos, t_neg;
};

#define U1(a, c, i) ((a)[3 * ((i) + 2) + (c)])

swo1d_ctx *swo1d_create(const swo1d_config *cfg) {
    swo1d_ctx *C = (swo1d_ctx *)calloc(1, sizeof(swo1d_ctx));
    C->cfg = *cfg;
    C->n = cfg->n;
    C->N = cfg->n + 6;
    C->gamma1 = cfg->gamma - 1.;
    /* weights(i) = 1/(nSteps + 6), nSteps = n as REAL (ex2_sedov_1d/main.f90:5,34-36) */
    C->weight = 1. / ((double)cfg->n + 6.);
    double **arrs[] = {&C->u, &C->u0, &C->rhs, &C->hp, &C->hn, &C->hpf, &C->hnf, &C->flus};
    for (int a = 0; a < 8; ++a) *arrs[a] = (double *)calloc((size_t)3 * C->N, sizeof(double));
    C->trk_cap = 1 << 16;
    C->trk_q = (int64_t *)calloc(C->trk_cap, sizeof(int64_t));
    C->trk_n = (int64_t *)calloc(C->trk_cap, sizeof(int64_t));
    C->trk_p = (int64_t *)calloc(C->trk_cap, sizeof(int64_t));
    return C;
}

void swo1d_destroy(swo1d_ctx *C) {
    if (!C) return;
    free(C->u); free(C->u0); free(C->rhs); free(C->hp); free(C->hn); free(C->hpf); free(C->hnf); free(C->flus);
    free(C->trk_q); free(C->trk_n); free(C->trk_p);
    free(C);
}

void swo1d_set_state(swo1d_ctx *C, const double *u) { memcpy(C->u, u, sizeof(double) * 3 * C->N); }
void swo1d_get_state(const swo1d_ctx *C, double *u) { memcpy(u, C->u, sizeof(double) * 3 * C->N); }
void swo1d_get_stats(const swo1d_ctx *C, swo1d_stats *s) { *s = C->st; }
int64_t swo1d_track(const swo1d_ctx *C, int64_t *q, int64_t *neg, int64_t *pos, int64_t cap) {
    const int64_t m = C->st.sweep_count < cap ? C->st.sweep_count : cap;
    for (int64_t k = 0; k < m && k < C->trk_cap; ++k) {
        q[k] = C->trk_q[k];
        neg[k] = C->trk_n[k];
        pos[k] = C->trk_p[k];
    }
    return C->st.sweep_count;
}

/* initdata (:398-420): rho = 1, m = 0, E = e_amb everywhere (the shipped program's
 * 1d-9/gamma1 in binary128 when shipped_energy), E(n/2+1) = 3200000.0/dx */
void swo1d_init_sedov(const swo1d_config *cfg, int shipped_energy, double e_amb, double *u) {
    const int N = cfg->n + 6;
    const double g1 = cfg->gamma - 1.;
    const double e = shipped_energy ? D(1e-9Q / Q(g1)) : e_amb;
    for (int i = 0; i < N; ++i) {
        u[3 * i + 0] = 1.;
        u[3 * i + 1] = 0.;
        u[3 * i + 2] = e;
    }
    u[3 * (cfg->n / 2 + 1 + 2) + 2] = 3200000.0 / cfg->dx;
}

/* findalpha (:285-296) */
static double alpha1d(const swo1d_ctx *C) {
    double amax = 0.;
    for (int i = -2; i <= C->n + 3; ++i) {
        const double den = U1(C->u, 0, i);
        const double pres = C->gamma1 * (U1(C->u, 2, i) - ((U1(C->u, 1, i) * U1(C->u, 1, i)) / U1(C->u, 0, i)) / 2.);
        const double sound = sqrt((C->cfg.gamma * pres) / den);
        amax = fmax2(amax, fabs(U1(C->u, 1, i) / U1(C->u, 0, i)) + sound);
    }
    return amax;
}

/* fflux (:59-73): vx = m/rho when rho > 1d-8 (binary128 literal), else 0 */
static void flux1d(double g1, double rho, double mx, double e, double f[3]) {
    const double vx = (Q(rho) > 1e-8Q) ? mx / rho : 0.;
    const double pres = g1 * (e - (0.5 * rho) * (vx * vx));
    f[0] = mx;
    f[1] = rho * (vx * vx) + pres;
    f[2] = (e + pres) * vx;
}

/* checkpositivity (:380-394) over the interior */
static int pos1d(const swo1d_ctx *C) {
    int tag = 0;
    for (int m = 1; m <= C->n; ++m) {
        if (U1(C->u, 0, m) <= C->cfg.eps) tag = 1;
        if ((U1(C->u, 2, m) - ((U1(C->u, 1, m) * U1(C->u, 1, m)) / U1(C->u, 0, m)) / 2.) <= C->cfg.eps) tag = 1;
    }
    return tag;
}

/* wenofd (:75-244) without the use_limiter = 1 branch: split, then per cell m in
 * 0..n+1 the cell's own eigenvectors, WENO of each characteristic field */
static void wenofd1d(swo1d_ctx *C) {
    const double g1 = C->gamma1;
    C->alphamax = alpha1d(C);
    for (int i = -2; i <= C->n + 3; ++i) {
        double f[3];
        flux1d(g1, U1(C->u, 0, i), U1(C->u, 1, i), U1(C->u, 2, i), f);
        for (int c = 0; c < 3; ++c) {
            U1(C->hp, c, i) = 0.5 * (U1(C->u, c, i) + f[c] / C->alphamax);
            U1(C->hn, c, i) = 0.5 * (U1(C->u, c, i) - f[c] / C->alphamax);
        }
    }
    C->postag = pos1d(C);
    for (int m = 0; m <= C->n + 1; ++m) {
        const double r = U1(C->u, 0, m), mx = U1(C->u, 1, m), e = U1(C->u, 2, m);
        const double vxm = mx / r;
        const double hm = (g1 * (e - (0.5 * (mx * mx)) / r) + e) / r;
        const double qm = (0.5 * vxm) * vxm;
        const double cm = sqrt(g1 * (hm - qm));
        double t0 = vxm * cm;
        const double evr[3][3] = {{1.0, 1.0, 1.0}, {vxm - cm, vxm, vxm + cm}, {hm - t0, qm, hm + t0}};
        const double rcm = 1. / cm;
        const double b1 = (g1 * rcm) * rcm, b2 = qm * b1;
        t0 = vxm * rcm;
        const double t1 = b1 * vxm, t2 = 0.5 * b1;
        const double evl[3][3] = {{0.5 * (b2 + t0), -0.5 * (t1 + rcm), t2},
                                  {1. - b2, t1, -b1},
                                  {0.5 * (b2 - t0), -0.5 * (t1 - rcm), t2}};
        double cup[3][5], cun[3][5];
        for (int j = -2; j <= 2; ++j)
            for (int k = 0; k < 3; ++k) {
                cup[k][j + 2] = (evl[k][0] * U1(C->hp, 0, m + j) + evl[k][1] * U1(C->hp, 1, m + j)) +
                                evl[k][2] * U1(C->hp, 2, m + j);
                cun[k][j + 2] = (evl[k][0] * U1(C->hn, 0, m + j) + evl[k][1] * U1(C->hn, 1, m + j)) +
                                evl[k][2] * U1(C->hn, 2, m + j);
            }
        double ur[3], ul[3];
        for (int k = 0; k < 3; ++k) {
            ur[k] = weno_right(&cup[k][2], 1e-6Q);
            ul[k] = weno_left(&cun[k][2], 1e-6Q);
        }
        for (int k = 0; k < 3; ++k) {
            U1(C->hpf, k, m) = (evr[k][0] * ur[0] + evr[k][1] * ur[1]) + evr[k][2] * ur[2];
            U1(C->hnf, k, m) = (evr[k][0] * ul[0] + evr[k][1] * ul[1]) + evr[k][2] * ul[2];
        }
    }
}

/* scheme (:263-280) */
static void scheme1d(swo1d_ctx *C) {
    for (int i = 1; i <= C->n + 1; ++i)
        for (int c = 0; c < 3; ++c) U1(C->flus, c, i) = C->alphamax * (U1(C->hpf, c, i - 1) - U1(C->hnf, c, i));
    memset(C->rhs, 0, sizeof(double) * 3 * C->N);
    for (int i = 1; i <= C->n; ++i)
        for (int c = 0; c < 3; ++c) U1(C->rhs, c, i) = -(U1(C->flus, c, i + 1) - U1(C->flus, c, i)) / C->cfg.dx;
}

/* nod_lim (:488-532): note the parenthesisation 0.5*(m**2/rho) */
static void nodlim1d(swo1d_ctx *C, const double v1[3], const double v2[3], const double vb[3], double out[3]) {
    const double g1 = C->gamma1, eps = C->cfg.eps;
    const double p1 = g1 * (v1[2] - 0.5 * ((v1[1] * v1[1]) / v1[0]));
    const double p2 = g1 * (v2[2] - 0.5 * ((v2[1] * v2[1]) / v2[0]));
    double t;
    if (p2 > eps) {
        t = (p1 - eps) / (p1 - p2);
        C->t_pos++;
    } else {
        const double pb = g1 * (vb[2] - 0.5 * ((vb[1] * vb[1]) / vb[0]));
        const double t_1 = (p1 - eps) / (p1 - pb);
        const double d12 = sqrt(((v2[2] - v1[2]) * (v2[2] - v1[2]) + (v2[1] - v1[1]) * (v2[1] - v1[1])) +
                                (v2[0] - v1[0]) * (v2[0] - v1[0]));
        const double d1b = sqrt(((vb[2] - v1[2]) * (vb[2] - v1[2]) + (vb[1] - v1[1]) * (vb[1] - v1[1])) +
                                (vb[0] - v1[0]) * (vb[0] - v1[0]));
        t = fmin2((t_1 * d1b) / d12, 0.25);
        C->t_neg++;
    }
    for (int c = 0; c < 3; ++c) out[c] = (1 - t) * v1[c] + t * v2[c];
}

/* the sweep's flag and pressure: gamma1*(E - 0.5*(m**2)/rho) (:562, :645) */
static inline double pres1d(double g1, const double *v) { return g1 * (v[2] - (0.5 * (v[1] * v[1])) / v[0]); }

/* scalar_sweep (:729-789), unconditional inside sweep_lim; delta = 1d-13 */
static void scalar1d(swo1d_ctx *C) {
    const double delta = 1e-13, wr = C->weight / C->weight;
    const int n = C->n;
    for (int j = -2; j <= n + 2; ++j) {
        const double vj = U1(C->u, 0, j), vjp1 = U1(C->u, 0, j + 1);
        if (vj < delta) {
            U1(C->u, 0, j + 1) = vjp1 + wr * (vj - delta);
            U1(C->u, 0, j) = delta;
        }
    }
    for (int j = n + 3; j >= -1; --j) {
        const double vj = U1(C->u, 0, j), vjm1 = U1(C->u, 0, j - 1);
        if (vj < delta) {
            U1(C->u, 0, j - 1) = vjm1 + wr * (vj - delta);
            U1(C->u, 0, j) = delta;
        }
    }
    C->st.den_count++;
}

/* sweep_lim (:538-727) */
static void sweep1d(swo1d_ctx *C) {
    const double g1 = C->gamma1, eps = C->cfg.eps, w = C->weight, wr = w / w;
    const int n = C->n;
    scalar1d(C);
    double ub[3] = {0., 0., 0.};  /* dot_product(rho_t, weights) ... in index order */
    for (int i = -2; i <= n + 3; ++i)
        for (int c = 0; c < 3; ++c) ub[c] = ub[c] + U1(C->u, c, i) * w;
    C->st.sweep_count++;
    const int64_t sc = C->st.sweep_count;
    C->t_pos = C->t_neg = 0;
    for (int q = 1; q <= C->cfg.swp_num; ++q) {
        for (int j = -2; j <= n + 2; ++j) {
            double vj[3] = {U1(C->u, 0, j), U1(C->u, 1, j), U1(C->u, 2, j)};
            if (pres1d(g1, vj) <= eps || vj[0] <= eps) {
                double vp[3] = {U1(C->u, 0, j + 1), U1(C->u, 1, j + 1), U1(C->u, 2, j + 1)}, vt[3];
                nodlim1d(C, vj, vp, ub, vt);
                for (int c = 0; c < 3; ++c) {
                    U1(C->u, c, j) = vt[c];
                    U1(C->u, c, j + 1) = vp[c] + wr * (vj[c] - vt[c]);
                }
            }
        }
        for (int j = n + 3; j >= -1; --j) {
            double vj[3] = {U1(C->u, 0, j), U1(C->u, 1, j), U1(C->u, 2, j)};
            if (pres1d(g1, vj) <= eps || vj[0] <= eps) {
                double vp[3] = {U1(C->u, 0, j - 1), U1(C->u, 1, j - 1), U1(C->u, 2, j - 1)}, vt[3];
                nodlim1d(C, vj, vp, ub, vt);
                for (int c = 0; c < 3; ++c) {
                    U1(C->u, c, j) = vt[c];
                    U1(C->u, c, j - 1) = vp[c] + wr * (vj[c] - vt[c]);
                }
            }
        }
        double pmin = 1e20, rmin = 1e20;
        for (int i = -2; i <= n + 3; ++i) {
            const double v[3] = {U1(C->u, 0, i), U1(C->u, 1, i), U1(C->u, 2, i)};
            pmin = fmin2(pres1d(g1, v), pmin);
            rmin = fmin2(v[0], rmin);
        }
        C->st.pre_count++;
        if (pmin > eps && rmin > eps) {
            C->postag = 0;
            if (sc - 1 < C->trk_cap) {
                C->trk_q[sc - 1] = q;
                C->trk_n[sc - 1] = C->t_neg;
                C->trk_p[sc - 1] = C->t_pos;
            }
            C->st.total_sweep += q;
            C->st.total_neg += C->t_neg;
            C->st.total_pos += C->t_pos;
            if (q > C->st.max_sweep) C->st.max_sweep = q;
            if (C->t_neg > C->st.max_neg) C->st.max_neg = C->t_neg;
            if (C->t_pos > C->st.max_pos) C->st.max_pos = C->t_pos;
            return;
        }
    }
}

/* rk3 (:298-332): dt = dx*cfl/alphamax, no clamp to the final time */
static void rk3_1d(swo1d_ctx *C) {
    const int N3 = 3 * C->N;
    for (int s = 1; s <= 3; ++s) {
        wenofd1d(C);
        if (s == 1) {
            C->dt = (C->cfg.dx * C->cfg.cfl) / C->alphamax;
            C->postag = 0;
            memcpy(C->u0, C->u, sizeof(double) * N3);
        }
        scheme1d(C);
        for (int k = 0; k < N3; ++k) {
            const double x = C->u[k] + C->dt * C->rhs[k];
            if (s == 1)
                C->u[k] = x;
            else if (s == 2)
                C->u[k] = D(Q(0.75) * Q(C->u0[k]) + Q(0.25) * Q(x));
            else
                C->u[k] = D((Q(C->u0[k]) + Q(2) * Q(x)) / Q(3));
        }
        C->postag = pos1d(C);
        if (C->postag == 1 && C->cfg.use_limiter == 2) sweep1d(C);
    }
}

int swo1d_run(swo1d_ctx *C, double *time_inout, double tfinal, int64_t max_steps, int64_t *steps_out) {
    double t = *time_inout;
    int64_t k = 0;
    while (k < max_steps && swo_time_continue(t, tfinal)) {
        rk3_1d(C);
        t = t + C->dt;
        ++k;
    }
    C->st.steps += k;
    *time_inout = t;
    if (steps_out) *steps_out = k;
    return 0;
}
