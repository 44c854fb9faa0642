/*
 * Oracle: continuous photon-energy samplers (TEST INFRASTRUCTURE).
 *
 * Restates src/energy_sampler_mod.f90: a piecewise-linear non-negative
 * density with its analytically integrated CDF, the adaptive Planck
 * refinement, and the quadratic inverse.
 *
 * Pinned by: the closed-form Planck band integral (the sampler total and CDF
 * against scipy quadrature to the accuracy the reference's own gate uses,
 * tests/energy_sampler/check_energy_sampler.py:149-168); the constant-density
 * inverse E = Emin + (Emax-Emin) u; the CDF round trip; and, when oracle/_ref
 * is built, the knot set and CDF of the reference module compiled from its
 * own source.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#include "mc_oracle.h"

#define MAX_PLANCK_KNOTS 200000
#define KB_EV 8.617333262145e-5

void orc_sampler_free(orc_sampler *s)
{
    free(s->energy); free(s->density); free(s->cdf);
    memset(s, 0, sizeof(*s));
}

/* build_tabulated_sampler (src/energy_sampler_mod.f90:34-64) */
int orc_build_tabulated(orc_sampler *s, const double *e, const double *f, int n)
{
    if (n < 2) return -1;
    for (int i = 0; i < n; ++i) if (f[i] < 0.0) return -2;
    for (int i = 0; i < n - 1; ++i) if (e[i + 1] <= e[i]) return -3;
    s->n = n;
    s->energy = (double *)malloc(sizeof(double) * n);
    s->density = (double *)malloc(sizeof(double) * n);
    s->cdf = (double *)malloc(sizeof(double) * n);
    memcpy(s->energy, e, sizeof(double) * n);
    memcpy(s->density, f, sizeof(double) * n);
    s->cdf[0] = 0.0;
    for (int i = 0; i < n - 1; ++i) {
        double area = 0.5 * (f[i] + f[i + 1]) * (e[i + 1] - e[i]);
        s->cdf[i + 1] = s->cdf[i] + area;
    }
    double total = s->cdf[n - 1];
    if (total <= 0.0) return -4;
    s->total_integral = total;
    for (int i = 0; i < n; ++i) s->cdf[i] = s->cdf[i] / total;
    s->cdf[n - 1] = 1.0;
    return 0;
}

/* planck_energy_shape (src/energy_sampler_mod.f90:277-290).  E**3 is the
 * integer power, evaluated as (E*E)*E. */
double orc_planck_shape(double E, double T)
{
    double x = E / (KB_EV * T);
    double e3 = (E * E) * E;
    if (x <= 0.0) return 0.0;
    if (x < 1.0e-3) return e3 / (x * (1.0 + 0.5 * x + x * x / 6.0));
    if (x < 700.0) return e3 / (exp(x) - 1.0);
    return e3 * exp(-x);
}

/* rough_planck_integral (src/energy_sampler_mod.f90:321-334) */
static double rough_planck_integral(double T, double emin, double emax)
{
    const int n = 256;
    double total = 0.0;
    for (int i = 1; i <= n; ++i) {
        double e0 = emin + (emax - emin) * (double)(i - 1) / (double)n;
        double e1 = emin + (emax - emin) * (double)i / (double)n;
        total += 0.5 * (orc_planck_shape(e0, T) + orc_planck_shape(e1, T)) * (e1 - e0);
    }
    return total;
}

typedef struct { double *e, *f; int n, cap, overflow; } knots_t;

/* refine_planck_interval (src/energy_sampler_mod.f90:159-185) */
static void refine(double T, double e0, double e1, double f0, double f1, double tol,
                   double scale, double span, int depth, knots_t *w)
{
    double em = 0.5 * (e0 + e1);
    double fm = orc_planck_shape(em, T);
    double coarse = 0.5 * (f0 + f1) * (e1 - e0);
    double fine = 0.25 * (f0 + 2.0 * fm + f1) * (e1 - e0);
    double err = fabs(fine - coarse) / 3.0;
    double allowance = tol * (fabs(fine) + scale * (e1 - e0) / span);
    if (err <= allowance || depth >= 50) {
        if (w->n >= w->cap) { w->overflow = 1; return; }
        w->e[w->n] = e1;
        w->f[w->n] = f1;
        w->n++;
    } else {
        refine(T, e0, em, f0, fm, tol, scale, span, depth + 1, w);
        refine(T, em, e1, fm, f1, tol, scale, span, depth + 1, w);
    }
}

static int cmp_dbl(const void *a, const void *b)
{
    double x = *(const double *)a, y = *(const double *)b;
    return (x > y) - (x < y);
}

/* build_planck_sampler (src/energy_sampler_mod.f90:106-157) */
int orc_build_planck(orc_sampler *s, double T, double emin, double emax, double rtol,
                     const double *thr, int nthr)
{
    if (T <= 0.0 || emin < 0.0 || emax <= emin) return -1;
    double tol = rtol > 0.0 ? rtol : 1.0e-8;
    double *mand = (double *)malloc(sizeof(double) * (nthr + 2));
    int nm = 0;
    mand[nm++] = emin;
    for (int i = 0; i < nthr; ++i)
        if (thr[i] > emin && thr[i] < emax) mand[nm++] = thr[i];
    mand[nm++] = emax;
    qsort(mand, nm, sizeof(double), cmp_dbl);       /* sort_unique */
    int nk = 1;
    for (int i = 1; i < nm; ++i)
        if (mand[i] > mand[nk - 1]) mand[nk++] = mand[i];

    knots_t w;
    w.cap = MAX_PLANCK_KNOTS;
    w.e = (double *)malloc(sizeof(double) * w.cap);
    w.f = (double *)malloc(sizeof(double) * w.cap);
    w.overflow = 0;
    double scale = rough_planck_integral(T, emin, emax);
    double span = emax - emin;
    w.n = 1;
    w.e[0] = mand[0];
    w.f[0] = orc_planck_shape(mand[0], T);
    for (int i = 0; i < nk - 1; ++i)
        refine(T, mand[i], mand[i + 1], orc_planck_shape(mand[i], T),
               orc_planck_shape(mand[i + 1], T), tol, scale, span, 0, &w);
    int rc = w.overflow ? -5 : orc_build_tabulated(s, w.e, w.f, w.n);
    free(w.e); free(w.f); free(mand);
    return rc;
}

/* sample_energy_cdf (src/energy_sampler_mod.f90:187-238) */
double orc_sample_energy(const orc_sampler *s, double u)
{
    double target = u < 0.0 ? 0.0 : (u > 1.0 ? 1.0 : u);
    if (target <= 0.0) return s->energy[0];
    if (target >= 1.0) return s->energy[s->n - 1];
    int lo = 1, hi = s->n;                 /* 1-based like the reference */
    while (hi - lo > 1) {
        int mid = (lo + hi) / 2;
        if (s->cdf[mid - 1] <= target) lo = mid; else hi = mid;
    }
    double interval_area = s->cdf[hi - 1] - s->cdf[lo - 1];
    if (interval_area <= 0.0) return s->energy[hi - 1];
    double dx = s->energy[hi - 1] - s->energy[lo - 1];
    double y0 = s->density[lo - 1], y1 = s->density[hi - 1];
    double local_area = (target - s->cdf[lo - 1]) / interval_area * (0.5 * (y0 + y1) * dx);
    double slope = (y1 - y0) / dx;
    double sc = fabs(y0) > fabs(y1) ? fabs(y0) : fabs(y1);
    if (sc < DBL_MIN) sc = DBL_MIN;
    double offset;
    if (fabs(slope * dx) <= 64.0 * DBL_EPSILON * sc) {
        offset = y0 > 0.0 ? local_area / y0 : 0.0;
    } else {
        double q = y0 * y0 + 2.0 * slope * local_area;
        double disc = sqrt(q > 0.0 ? q : 0.0);
        offset = 2.0 * local_area / (y0 + disc);
    }
    if (offset < 0.0) offset = 0.0;
    if (offset > dx) offset = dx;
    return s->energy[lo - 1] + offset;
}

/* energy_cdf_value (src/energy_sampler_mod.f90:240-275) */
double orc_energy_cdf_value(const orc_sampler *s, double E)
{
    if (E <= s->energy[0]) return 0.0;
    if (E >= s->energy[s->n - 1]) return 1.0;
    int lo = 1, hi = s->n;
    while (hi - lo > 1) {
        int mid = (lo + hi) / 2;
        if (s->energy[mid - 1] <= E) lo = mid; else hi = mid;
    }
    double dx = s->energy[hi - 1] - s->energy[lo - 1];
    double off = E - s->energy[lo - 1];
    double y0 = s->density[lo - 1];
    double slope = (s->density[hi - 1] - y0) / dx;
    double tot = 0.5 * (y0 + s->density[hi - 1]) * dx;
    double partial = y0 * off + 0.5 * slope * off * off;
    if (tot > 0.0)
        return s->cdf[lo - 1] + (s->cdf[hi - 1] - s->cdf[lo - 1]) * partial / tot;
    return s->cdf[lo - 1];
}
