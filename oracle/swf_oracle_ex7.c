/*
 * swf_oracle_ex7.c - CPU restatement of the detonation program's sweep
 * (ex7_detonation/main.f90: sweep_lim :1433-1737, scalar_sweep :1740-1806,
 * nod_lim :1399-1429, sweeping_order :1808-1862), serial, in the reference's loop
 * order.
 *
 * TEST INFRASTRUCTURE ONLY: the checker of the library's workgroup-parallel sweep on
 * crafted states (tests/); the product library never links or calls it.
 *
 * State: the program's u(5, -2:n+3, -2:n+3) in the Fortran layout (component
 * fastest), n = 5 total; the sweep works on the interior (the program's utemp
 * gathered from one rank).  Uniform weights 1/(n*n) as the program's nSteps.
 * Every literal of these procedures is REAL(8) under -fdefault-real-8 apart from
 * 1d20 and 1d-12, which are exact or only compared, so plain double arithmetic
 * without contraction reproduces the reference build (x86-64 SSE2).
 */
#include <math.h>
#include <stdint.h>

#include "swf_oracle.h"

typedef struct {
    int n;
    double *u;
    double gamma1, q, eps, w;
    long long pos, neg;
} sw7;

static double *U7(const sw7 *s, int c, int i, int j) {
    return s->u + c + 5 * ((long long)(i + 2) + (long long)(s->n + 6) * (j + 2));
}

/* sweeping_order (:1808-1862): order I row by row, odd rows left to right; order II
 * column by column, odd columns bottom to top.  p is 1-based. */
static void ord_I(int n, long long p, int *i, int *j) {
    const long long q = p - 1;
    *j = (int)(q / n) + 1;
    const int r = (int)(q - (long long)(*j - 1) * n);
    *i = (*j % 2 == 1) ? r + 1 : n - r;
}
static void ord_II(int n, long long p, int *i, int *j) {
    const long long q = p - 1;
    *i = (int)(q / n) + 1;
    const int r = (int)(q - (long long)(*i - 1) * n);
    *j = (*i % 2 == 1) ? r + 1 : n - r;
}

/* gamma1*(v(4) - 0.5*v(2)*v(2)/v(1) - 0.5*v(3)**2/v(1) - q_const*v(5)), the form of
 * the flag tests (:1503) */
static double p_flag(const sw7 *s, const double v[5]) {
    return s->gamma1 * (((v[3] - ((0.5 * v[1]) * v[1]) / v[0]) - (0.5 * (v[2] * v[2])) / v[0]) - s->q * v[4]);
}
/* gamma1*(v(4) - 0.5*v(2)**2/v(1) - 0.5*v(3)**2/v(1) - q_const*v(5)) (nod_lim :1405) */
static double p_sq(const sw7 *s, const double v[5]) {
    return s->gamma1 * (((v[3] - (0.5 * (v[1] * v[1])) / v[0]) - (0.5 * (v[2] * v[2])) / v[0]) - s->q * v[4]);
}

/* nod_lim (:1399-1429) */
static void nod_lim7(sw7 *s, const double *v1, const double *v2, const double *vb, double *out) {
    const double p1 = p_sq(s, v1), p2 = p_sq(s, v2);
    double t;
    if (p2 > s->eps) {
        t = (p1 - s->eps) / (p1 - p2);
        s->pos++;
    } else {
        const double pb = p_sq(s, vb);
        const double t1 = (p1 - s->eps) / (p1 - pb);
        double a12 = 0., a1b = 0.;
        /* (v2(5)-v1(5))**2 + (v2(4)-v1(4))**2 + ... + (v2(1)-v1(1))**2, left to right */
        for (int c = 4; c >= 0; --c) {
            const double a = v2[c] - v1[c], d = vb[c] - v1[c];
            a12 = (c == 4) ? a * a : a12 + a * a;
            a1b = (c == 4) ? d * d : a1b + d * d;
        }
        const double x = (t1 * sqrt(a1b)) / sqrt(a12);
        t = (x < 0.25) ? x : 0.25; /* min(x, 0.25) */
        s->neg++;
    }
    for (int c = 0; c < 5; ++c) out[c] = (1. - t) * v1[c] + t * v2[c];
}

/* min(x, m) as the program's loops fold it */
static double min7(double x, double m) { return (m < x) ? m : x; }

/* One flagged-pair adjustment (:1505-1540): cell a, neighbour p, component 5 of each
 * at a5 / p5.  Returns 0 when the cell is not flagged, 1 when adjusted, 2 when the
 * order-I pressure-sum check fails (check). */
static int adjust7(sw7 *s, const double *ub, int ia, int ja, int ip, int jp, int ia5, int ja5, int ip5, int jp5,
                   int check) {
    double vi[5], vp[5], vit[5], vpt[5];
    for (int c = 0; c < 4; ++c) vi[c] = *U7(s, c, ia, ja);
    vi[4] = *U7(s, 4, ia5, ja5);
    const double pvi = p_flag(s, vi);
    if (!(pvi <= s->eps || vi[0] <= s->eps)) return 0;
    for (int c = 0; c < 4; ++c) vp[c] = *U7(s, c, ip, jp);
    vp[4] = *U7(s, 4, ip5, jp5);
    nod_lim7(s, vi, vp, ub, vit);
    for (int c = 0; c < 5; ++c) vpt[c] = vp[c] + (s->w / s->w) * (vi[c] - vit[c]);
    for (int c = 0; c < 4; ++c) *U7(s, c, ia, ja) = vit[c];
    *U7(s, 4, ia5, ja5) = vit[4];
    for (int c = 0; c < 4; ++c) *U7(s, c, ip, jp) = vpt[c];
    *U7(s, 4, ip5, jp5) = vpt[4];
    if (check) {
        const double pvp = p_flag(s, vp), pvit = p_sq(s, vit), pvpt = p_sq(s, vpt);
        if ((pvi + pvp) > (pvit + pvpt) + s->eps) return 2;
    }
    return 1;
}

int swo7_sweep(double *u, int32_t total, double gamma, double q_const, double eps, int32_t swp_num,
               int64_t *counts, int64_t *track) {
    sw7 s;
    s.n = 5 * total;
    s.u = u;
    s.gamma1 = gamma - 1.;
    s.q = q_const;
    s.eps = eps;
    s.w = 1. / ((double)s.n * (double)s.n);
    s.pos = s.neg = 0;
    const int n = s.n;
    const long long N = (long long)n * n;
    int i, j, i2, j2, i5, j5, ip5, jp5;
    /* scalar_sweep (:1740-1806) */
    double rmin = 1e20;
    for (j = 1; j <= n; ++j)
        for (i = 1; i <= n; ++i) rmin = min7(*U7(&s, 0, i, j), rmin);
    if (rmin < 1e-12) {
        for (long long p = 1; p <= N - 1; ++p) {
            ord_I(n, p, &i, &j);
            ord_I(n, p + 1, &i2, &j2);
            const double vi = *U7(&s, 0, i, j), vn = *U7(&s, 0, i2, j2);
            if (vi < 1e-12) {
                *U7(&s, 0, i2, j2) = vn + (s.w / s.w) * (vi - 1e-12);
                *U7(&s, 0, i, j) = 1e-12;
            }
        }
        for (long long p = N; p >= 2; --p) {
            ord_I(n, p, &i, &j);
            ord_I(n, p - 1, &i2, &j2);
            const double vi = *U7(&s, 0, i, j), vn = *U7(&s, 0, i2, j2);
            if (vi < 1e-12) {
                *U7(&s, 0, i2, j2) = vn + (s.w / s.w) * (vi - 1e-12);
                *U7(&s, 0, i, j) = 1e-12;
            }
        }
        counts[1]++;
    }
    /* u_bar (:1461-1475), j outer, i inner */
    double ub[5] = {0., 0., 0., 0., 0.};
    for (j = 1; j <= n; ++j)
        for (i = 1; i <= n; ++i)
            for (int c = 0; c < 5; ++c) ub[c] = ub[c] + s.w * *U7(&s, c, i, j);
    if (s.gamma1 * (((ub[3] - (0.5 * (ub[1] * ub[1])) / ub[0]) - (0.5 * (ub[2] * ub[2])) / ub[0]) - s.q * ub[4]) <= 0.)
        return 1;
    counts[0]++;
    for (int q = 1; q <= swp_num; ++q) {
        if (q % 2 == 1) {
            for (long long p = 1; p <= N - 1; ++p) {
                ord_I(n, p, &i, &j);
                ord_I(n, p + 1, &i2, &j2);
                if (adjust7(&s, ub, i, j, i2, j2, i, j, i2, j2, 1) == 2) {
                    track[0] = q;
                    return 3;
                }
            }
            for (long long p = N; p >= 2; --p) {
                ord_I(n, p, &i, &j);
                ord_I(n, p - 1, &i2, &j2);
                if (adjust7(&s, ub, i, j, i2, j2, i, j, i2, j2, 1) == 2) {
                    track[0] = q;
                    return 3;
                }
            }
        } else {
            /* order II, component 5 at the order-I positions p and p - 1 (:1604-1727) */
            for (long long p = 1; p <= N - 1; ++p) {
                ord_II(n, p, &i, &j);
                ord_II(n, p + 1, &i2, &j2);
                ord_I(n, p, &i5, &j5);
                double v[5];
                for (int c = 0; c < 4; ++c) v[c] = *U7(&s, c, i, j);
                v[4] = *U7(&s, 4, i5, j5);
                if (!(p_flag(&s, v) <= s.eps || v[0] <= s.eps)) continue;
                if (p == 1) { /* sweepI_i(0): outside the order arrays */
                    track[0] = q;
                    return 4;
                }
                ord_I(n, p - 1, &ip5, &jp5);
                adjust7(&s, ub, i, j, i2, j2, i5, j5, ip5, jp5, 0);
            }
            for (long long p = N; p >= 2; --p) {
                ord_II(n, p, &i, &j);
                ord_II(n, p - 1, &i2, &j2);
                ord_I(n, p, &i5, &j5);
                ord_I(n, p - 1, &ip5, &jp5);
                adjust7(&s, ub, i, j, i2, j2, i5, j5, ip5, jp5, 0);
            }
        }
        /* the convergence test (:1690-1712): pressure without the reaction term */
        double pmin = 1e20;
        rmin = 1e20;
        for (j = 1; j <= n; ++j)
            for (i = 1; i <= n; ++i) {
                const double r = *U7(&s, 0, i, j), m1 = *U7(&s, 1, i, j), m2 = *U7(&s, 2, i, j),
                             e = *U7(&s, 3, i, j);
                pmin = min7(s.gamma1 * (e - (0.5 * (m1 * m1 + m2 * m2)) / r), pmin);
                rmin = min7(r, rmin);
            }
        counts[2]++;
        if (pmin > s.eps && rmin > s.eps) {
            track[0] = q;
            track[1] = s.neg;
            track[2] = s.pos;
            return 0;
        }
    }
    track[0] = swp_num;
    return 2;
}
