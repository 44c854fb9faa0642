/*
 * Oracle: per-cell H/He ionization balance at fixed T (TEST INFRASTRUCTURE).
 *
 * Restates src/ion_balance_mod.f90 without the metal registry
 * (use_metals off, so the charge-exchange and metal-electron terms are
 * identically zero):
 *   solve_ion_cell                 :53-150
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
    const double T = c->p.te_fixed, yHe = c->p.He_abund, w = c->p.ion_relax;
    double mdx = 0.0, sdx = 0.0, sx = 0.0;
    for (int i = 0; i < c->nleaf; ++i) {
        double nH = c->nH[i];
        if (nH <= 0.0) { c->ne[i] = 0.0; continue; }
        double xHI = c->xHI[i], xHeI = c->xHeI[i], xHeII = c->xHeII[i], ne;
        orc_solve_ion_cell(c->gHI[i], c->gHeI[i], c->gHeII[i], nH, T, c->p.caseA, yHe, &xHI,
                           &xHeI, &xHeII, &ne);
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
    }
    *max_dx = mdx;
    *dx_vol = sdx / (sx > DBL_MIN ? sx : DBL_MIN);
}

/* hydrogen_balance_residual_grid (src/ion_balance_mod.f90:280-343) */
void orc_residual_grid(orc_ctx *c, double *eps_max, double *eps_vol)
{
    double emax = 0.0, sev = 0.0, sv = 0.0;
    for (int i = 0; i < c->nleaf; ++i) {
        if (c->nH[i] <= 0.0) continue;
        if (1.0 - c->xHI[i] < 1.0e-10) continue;
        double T = c->Te[i];
        double aH = c->p.caseA ? orc_alphaA_HII(T) : orc_alphaB_HII(T);
        double cH = orc_ci_HI(T);
        double nHI = c->nH[i] * c->xHI[i], nHII = c->nH[i] * (1.0 - c->xHI[i]);
        double loss = nHI * (c->gHI[i] + cH * c->ne[i]);
        double gain = (aH * c->ne[i]) * nHII;
        double eps = fabs(loss - gain) / (loss > DBL_MIN ? loss : DBL_MIN);
        if (eps > emax) emax = eps;
        double v = leaf_vol(c, i);
        sev = sev + v * eps;
        sv = sv + v;
    }
    *eps_max = emax;
    *eps_vol = sev / (sv > DBL_MIN ? sv : DBL_MIN);
}
