/*
 * Oracle: per-cell H/He ionization balance at fixed T (TEST INFRASTRUCTURE).
 *
 * Restates src/ion_balance_mod.f90; the metal-electron and charge-exchange
 * branches of solve_ion_cell and charge_neutrality_ne are in orc_metals.c:
 *   solve_ion_cell (no registry)   :53-150
 *   hydrogen_balance_residual(_grid):280-343
 *   gas_equilibrium_update          :352-434
 *
 * Pinned by: the converged state satisfying the three balance equations and
 * charge neutrality to the solver tolerance (the reference's
 * tests/charge_exchange/check_cell_convergence.f90 gate: one call reaches the
 * converged state to 1e-9); the closed-form pure-H fixed point
 * x_HI = alpha ne / (Gamma + (C + alpha) ne) with ne = nH (1 - x_HI)
 * (a quadratic), and the relaxation/metric formulas by direct evaluation.
 */
#include <math.h>
#include <float.h>
#include <pthread.h>
#include <stdlib.h>
#include "mc_oracle.h"

void orc_solve_ion_cell(double gH, double gHe1, double gHe2, double nH, double T, int caseA,
                        double yHe, double *xHI, double *xHeI, double *xHeII, double *ne_out)
{
    double aH, aHe2, aHe3;
    if (caseA) { aH = orc_alphaA_HII(T); aHe2 = orc_alphaA_HeII(T); aHe3 = orc_alphaA_HeIII(T); }
    else { aH = orc_alphaB_HII(T); aHe2 = orc_alphaB_HeII(T); aHe3 = orc_alphaB_HeIII(T); }
    double cH = orc_ci_HI(T), cHe1 = orc_ci_HeI(T), cHe2 = orc_ci_HeII(T);
    double xHeIII = 1.0 - *xHeI - *xHeII;
    if (xHeIII < 0.0) xHeIII = 0.0;
    double ne = nH * ((1.0 - *xHI) + yHe * (*xHeII + 2.0 * xHeIII));
    if (ne < 1.0e-12 * nH) ne = 1.0e-12 * nH;
    for (int it = 1; it <= 200; ++it) {
        double ne_old = ne;
        double denH = gH + (cH + aH) * ne;
        *xHI = denH > 0.0 ? (aH * ne) / denH : 1.0;
        double r1 = (gHe1 + cHe1 * ne) / (aHe2 * ne);
        double r2 = (gHe2 + cHe2 * ne) / (aHe3 * ne);
        if (r1 > 1.0e150) r1 = 1.0e150;
        if (r2 > 1.0e150) r2 = 1.0e150;
        *xHeI = 1.0 / (1.0 + r1 + r1 * r2);
        *xHeII = *xHeI * r1;
        xHeIII = *xHeII * r2;
        ne = nH * ((1.0 - *xHI) + yHe * (*xHeII + 2.0 * xHeIII));
        double nn = 0.5 * (ne + ne_old);
        ne = nn > 1.0e-12 * nH ? nn : 1.0e-12 * nH;
        if (fabs(ne - ne_old) <= 1.0e-10 * ne) break;
    }
    *ne_out = ne;
}

static double leaf_vol(const orc_ctx *c, int i)
{
    double s = c->car ? c->dcell : 2.0 * c->ch[c->icell_of_leaf[i] - 1];
    return (s * s) * s;
}

/* gas_equilibrium_update (src/ion_balance_mod.f90:352-434) */
void orc_equilibrium_update(orc_ctx *c, double *max_dx, double *dx_vol)
{
    const int ci_prev = orc_ci_model_set(c->p.ci_dere);
    const double T = c->p.te_fixed, yHe = c->p.He_abund, w = c->p.ion_relax;
    double mdx = 0.0, sdx = 0.0, sx = 0.0;
    for (int i = 0; i < c->nleaf; ++i) {
        double nH = c->nH[i];
        if (nH <= 0.0) { c->ne[i] = 0.0; continue; }
        double xHI = c->xHI[i], xHeI = c->xHeI[i], xHeII = c->xHeII[i], ne;
        orc_solve_ion_cell_leaf(c, i, c->gHI[i], c->gHeI[i], c->gHeII[i], nH, T, c->p.caseA, yHe,
                                &xHI, &xHeI, &xHeII, &ne);
        double nHI = (1.0 - w) * c->xHI[i] + w * xHI;
        double nHeI = (1.0 - w) * c->xHeI[i] + w * xHeI;
        double nHeII = (1.0 - w) * c->xHeII[i] + w * xHeII;
        double xHeIII = 1.0 - nHeI - nHeII;
        if (xHeIII < 0.0) xHeIII = 0.0;
        double d = fabs(nHI - c->xHI[i]);
        if (d > mdx) mdx = d;
        double v = leaf_vol(c, i);
        sdx = sdx + v * d;
        sx = sx + v * (1.0 - nHI);
        c->xHI[i] = nHI;
        c->xHeI[i] = nHeI;
        c->xHeII[i] = nHeII;
        c->ne[i] = nH * ((1.0 - nHI) + yHe * (nHeII + 2.0 * xHeIII));
        if (c->p.metal_ne && c->use_metals && c->nel > 0)
            c->ne[i] = orc_charge_neutrality_ne(c, i, T, nH, c->ne[i], nHI, ne);
    }
    *max_dx = mdx;
    *dx_vol = sdx / (sx > DBL_MIN ? sx : DBL_MIN);
    orc_ci_model_set(ci_prev);
}

/* hydrogen_balance_residual_grid (src/ion_balance_mod.f90:280-343) */
void orc_residual_grid(orc_ctx *c, double *eps_max, double *eps_vol)
{
    const int ci_prev = orc_ci_model_set(c->p.ci_dere);
    double emax = 0.0, sev = 0.0, sv = 0.0;
    for (int i = 0; i < c->nleaf; ++i) {
        if (c->nH[i] <= 0.0) continue;
        if (1.0 - c->xHI[i] < 1.0e-10) continue;
        double T = c->Te[i];
        double rin = 0.0, rout = 0.0, nem;
        if (c->use_metals && c->nel > 0 && c->p.cx_two_way) {
            orc_mcache cc;
            orc_species_ne_prepare(c, i, T, &cc);
            orc_species_metal_sums(c, &cc, c->nH[i], c->ne[i], c->nH[i] * c->xHI[i],
                                   c->nH[i] * (1.0 - c->xHI[i]), &nem, &rin, &rout);
        }
        double aH = c->p.caseA ? orc_alphaA_HII(T) : orc_alphaB_HII(T);
        double cH = orc_ci_HI(T);
        double nHI = c->nH[i] * c->xHI[i], nHII = c->nH[i] * (1.0 - c->xHI[i]);
        double loss = nHI * (c->gHI[i] + cH * c->ne[i] + rout);
        double gain = (aH * c->ne[i] + rin) * nHII;
        double eps = fabs(loss - gain) / (loss > DBL_MIN ? loss : DBL_MIN);
        if (eps > emax) emax = eps;
        double v = leaf_vol(c, i);
        sev = sev + v * eps;
        sv = sv + v;
    }
    *eps_max = emax;
    *eps_vol = sev / (sv > DBL_MIN ? sv : DBL_MIN);
    orc_ci_model_set(ci_prev);
}

/* cooling_setup (src/cooling_mod.f90:67-92): H I fit terms (from
 * data/atomic/cooling_h_1.txt) and the gbar_ff table */
void orc_cooling_setup(orc_ctx *c, int nfit, const double *A, const double *Ti)
{
    c->cool.nfit = nfit;
    for (int i = 0; i < nfit; ++i) { c->cool.A[i] = A[i]; c->cool.Ti[i] = Ti[i]; }
    orc_gff_table(c->cool.logT, c->cool.gff);
}

/* net_rate (src/thermal_mod.f90:38-115) with the metal terms, no grain
 * heating */
double orc_net_rate(const orc_ctx *c, int il, double T, double *xHI, double *xHeI,
                    double *xHeII, double *ne)
{
    const int ci_prev = orc_ci_model_set(c->p.ci_dere);
    const double nH = c->nH[il], y = c->p.He_abund;
    orc_solve_ion_cell_leaf(c, il, c->gHI[il], c->gHeI[il], c->gHeII[il], nH, T, c->p.caseA, y,
                            xHI, xHeI, xHeII, ne);
    double heat = nH * (*xHI * c->hHI[il] + y * (*xHeI * c->hHeI[il] + *xHeII * c->hHeII[il]));
    double net = heat - orc_cooling_total(&c->cool, T, nH, *ne, *xHI, *xHeI, *xHeII, y,
                                          c->p.caseA);
    if (c->use_metals) {
        double nHI = nH * *xHI, nHII = nH * (1.0 - *xHI);
        net = net - orc_metal_cooling(c, il, T, nH, *ne, nHI, nHII);
        net = net - orc_metal_cooling_H(c, il, T, nH, *ne, nHI, nHII);
        if (c->p.metal_freefree) net = net - orc_metal_freefree(c, il, T, nH, *ne, nHI, nHII);
        if (c->p.metal_heat) net = net + orc_metal_heating(c, il, T, nH, *ne, nHI, nHII);
    }
    orc_ci_model_set(ci_prev);
    return net;
}

/* gas_thermal_update (src/thermal_mod.f90:124-249): the per-leaf
 * bisection of T_e with the ionization solved at each trial T, then the
 * relaxation of x.  Leaves are solved by solve_threads workers into
 * per-leaf results; the metrics are then summed in leaf order, as the
 * reference's loop sums them on one process. */
typedef struct {
    orc_ctx *c;
    int lo, hi;
    double *te, *x3, *ne;            /* per leaf: T_e, (xHI xHeI xHeII), ne */
} thermal_work;

static void thermal_leaf(const orc_ctx *c, int il, double *te_out, double *x_out, double *ne_out)
{
    double xHI = c->xHI[il], xHeI = c->xHeI[il], xHeII = c->xHeII[il], ne, te;
    double tlo = c->p.te_min, thi = c->p.te_max;
    double net_lo = orc_net_rate(c, il, tlo, &xHI, &xHeI, &xHeII, &ne);
    double net_hi = orc_net_rate(c, il, thi, &xHI, &xHeI, &xHeII, &ne);
    if (net_lo <= 0.0) {
        te = tlo;
    } else if (net_hi >= 0.0) {
        te = thi;
    } else {
        for (int it = 1; it <= 60; ++it) {
            double tmid = sqrt(tlo * thi);
            double net_mid = orc_net_rate(c, il, tmid, &xHI, &xHeI, &xHeII, &ne);
            if (net_mid > 0.0) tlo = tmid; else thi = tmid;
            if (thi / tlo - 1.0 < 1.0e-5) break;
        }
        te = sqrt(tlo * thi);
    }
    orc_net_rate(c, il, te, &xHI, &xHeI, &xHeII, &ne);
    *te_out = te;
    x_out[0] = xHI; x_out[1] = xHeI; x_out[2] = xHeII;
    *ne_out = ne;
}

static void *thermal_worker(void *arg)
{
    thermal_work *w = (thermal_work *)arg;
    for (int il = w->lo; il < w->hi; ++il)
        if (w->c->nH[il] > 0.0) thermal_leaf(w->c, il, &w->te[il], &w->x3[3 * il], &w->ne[il]);
    return NULL;
}

void orc_set_solve_threads(orc_ctx *c, int n) { c->solve_threads = n < 1 ? 1 : n; }

void orc_thermal_update(orc_ctx *c, double *max_dx, double *max_dte, double *dx_vol,
                        double *dte_vol)
{
    const int ci_prev = orc_ci_model_set(c->p.ci_dere);
    const double w = c->p.ion_relax, y = c->p.He_abund;
    const int nl = c->nleaf;
    double *te = (double *)malloc(sizeof(double) * (nl + 1));
    double *x3 = (double *)malloc(sizeof(double) * 3 * (nl + 1));
    double *nes = (double *)malloc(sizeof(double) * (nl + 1));
    int nt = c->solve_threads < 1 ? 1 : c->solve_threads;
    if (nt > nl) nt = nl > 0 ? nl : 1;
    thermal_work *wk = (thermal_work *)calloc(nt, sizeof(thermal_work));
    pthread_t *th = (pthread_t *)calloc(nt, sizeof(pthread_t));
    for (int t = 0; t < nt; ++t) {
        wk[t].c = c; wk[t].te = te; wk[t].x3 = x3; wk[t].ne = nes;
        wk[t].lo = (int)((int64_t)nl * t / nt);
        wk[t].hi = (int)((int64_t)nl * (t + 1) / nt);
    }
    for (int t = 1; t < nt; ++t) pthread_create(&th[t], NULL, thermal_worker, &wk[t]);
    thermal_worker(&wk[0]);
    for (int t = 1; t < nt; ++t) pthread_join(th[t], NULL);
    double mdx = 0.0, mdt = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0, s4 = 0.0;
    for (int il = 0; il < nl; ++il) {
        const double nH = c->nH[il];
        if (nH <= 0.0) { c->ne[il] = 0.0; continue; }
        const double t_e = te[il], ne = nes[il];
        double nHI = (1.0 - w) * c->xHI[il] + w * x3[3 * il];
        double nHeI = (1.0 - w) * c->xHeI[il] + w * x3[3 * il + 1];
        double nHeII = (1.0 - w) * c->xHeII[il] + w * x3[3 * il + 2];
        double xHeIII = 1.0 - nHeI - nHeII;
        if (xHeIII < 0.0) xHeIII = 0.0;
        double ne_new = nH * ((1.0 - nHI) + y * (nHeII + 2.0 * xHeIII));
        if (c->p.metal_ne && c->use_metals && c->nel > 0)
            ne_new = orc_charge_neutrality_ne(c, il, t_e, nH, ne_new, nHI, ne);
        double d = fabs(nHI - c->xHI[il]);
        if (d > mdx) mdx = d;
        double dt = fabs(t_e - c->Te[il]) / c->Te[il];
        if (dt > mdt) mdt = dt;
        double v = leaf_vol(c, il);
        s1 = s1 + v * d;
        s2 = s2 + v * (1.0 - nHI);
        s3 = s3 + v * ne_new * fabs(t_e - c->Te[il]);
        s4 = s4 + v * ne_new * t_e;
        c->xHI[il] = nHI; c->xHeI[il] = nHeI; c->xHeII[il] = nHeII;
        c->ne[il] = ne_new; c->Te[il] = t_e;
    }
    free(te); free(x3); free(nes); free(wk); free(th);
    *max_dx = mdx; *max_dte = mdt;
    *dx_vol = s1 / (s2 > DBL_MIN ? s2 : DBL_MIN);
    *dte_vol = s3 / (s4 > DBL_MIN ? s4 : DBL_MIN);
    orc_ci_model_set(ci_prev);
}
