/*
 * Oracle: model set-up -- band edges, source samplers, grid geometry, leaf
 * state, dust optics (TEST INFRASTRUCTURE).
 *
 * Restates src/ion_band_mod.f90 (ion_setup, build_aligned_edges, ion_bin_of,
 * ion_qmc_ndim, setup_continuous_source_sampler for Planck point sources and
 * a Planck external field), src/octree_mod.f90 (car geometry, amr_build_tree,
 * amr_build_neighbors, amr_find_leaf, amr_cell_exit, amr_next_leaf, car_next),
 * src/grid_mod_amr.f90:85-108 (namelist car grid), src/gas_state_mod.f90:41-76
 * (initial state) and src/gas_opacity_mod.f90:217-290 (dust optics at E).
 *
 * Pinned by: band sum rules (edges pinned at the He I/He II thresholds,
 * bin counts by largest remainder over log widths) and brute force on small
 * cases in the tests; octree walk against the car walk on single-level trees
 * (bit identical, src/define.f90:665-672) and brute-force point location; dust
 * optics against direct interpolation of the table.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#include "mc_oracle.h"

orc_ctx *orc_create(const orc_params *p)
{
    orc_ctx *c = (orc_ctx *)calloc(1, sizeof(orc_ctx));
    c->p = *p;
    orc_qmc_setup(&c->qmc, p->qmc_seed);
    return c;
}

void orc_destroy(orc_ctx *c)
{
    if (!c) return;
    free(c->eedge);
    for (int i = 0; i <= ORC_MAXSRC; ++i) orc_sampler_free(&c->samp[i]);
    free(c->parent); free(c->children); free(c->level); free(c->ileaf);
    free(c->icell_of_leaf); free(c->neighbor);
    free(c->cx); free(c->cy); free(c->cz); free(c->ch);
    free(c->nH); free(c->xHI); free(c->xHeI); free(c->xHeII); free(c->ne); free(c->Te);
    free(c->rhokap);
    free(c->dlam); free(c->dalb); free(c->dg); free(c->dcext);
    free(c->hhe); free(c->jt); free(c->dust_abs);
    free(c->gHI); free(c->gHeI); free(c->gHeII); free(c->hHI); free(c->hHeI); free(c->hHeII);
    free(c->milne); free(c->hei); free(c->dif_ch); free(c->dif_cdf);
    free(c->mraw); free(c->mgam); free(c->mheat); free(c->mfrac);
    orc_dustir_free(c->ir);
    free(c);
}

/* build_aligned_edges (src/ion_band_mod.f90:1053-1214): H/He thresholds
 * (priority 1) and the active metal thresholds (priority 0) strictly inside
 * (lo, hi), merged within |ln ratio| < 0.01, metal ones dropped (most
 * redundant first) while they outnumber the bins */
static int aligned_edges(double *eedge, int nnu, double lo, double hi, int nmet,
                         const double *meth)
{
    const double MERGE_TOL = 0.01;
    double cand[80], thr[80], points[82], w[81];
    int cpri[80], tpri[80], nb[81];
    int ncand = 0;
    const double th[3] = {ORC_ETH_HI, ORC_ETH_HEI, ORC_ETH_HEII};
    for (int i = 0; i < 3; ++i)
        if (th[i] > lo && th[i] < hi) { cand[ncand] = th[i]; cpri[ncand] = 1; ++ncand; }
    for (int i = 0; i < nmet && i < 64; ++i)
        if (meth[i] > 0.0 && meth[i] > lo && meth[i] < hi) {
            cand[ncand] = meth[i]; cpri[ncand] = 0; ++ncand;
        }
    for (int i = 1; i < ncand; ++i) {       /* insertion sort */
        double t = cand[i]; int tp = cpri[i], j = i - 1;
        while (j >= 0 && cand[j] > t) { cand[j + 1] = cand[j]; cpri[j + 1] = cpri[j]; --j; }
        cand[j + 1] = t; cpri[j + 1] = tp;
    }
    int nthr = 0;
    for (int i = 0; i < ncand; ++i) {
        int dup = nthr >= 1 && fabs(log(cand[i] / thr[nthr - 1])) < MERGE_TOL;
        if (dup) {
            if (cpri[i] > tpri[nthr - 1]) { thr[nthr - 1] = cand[i]; tpri[nthr - 1] = cpri[i]; }
        } else {
            thr[nthr] = cand[i]; tpri[nthr] = cpri[i]; ++nthr;
        }
    }
    while (nthr + 1 > nnu) {                /* 2c: drop the most redundant metal edge */
        int best = -1; double frac = 1.0e30;
        for (int i = 0; i < nthr; ++i) {
            if (tpri[i] != 0) continue;
            double t = i == 0 ? log(thr[i] / lo) : log(thr[i] / thr[i - 1]);
            if (i == nthr - 1) t = fmin(t, log(hi / thr[i]));
            else t = fmin(t, log(thr[i + 1] / thr[i]));
            if (t < frac) { frac = t; best = i; }
        }
        if (best < 0) break;
        for (int i = best; i < nthr - 1; ++i) { thr[i] = thr[i + 1]; tpri[i] = tpri[i + 1]; }
        --nthr;
    }
    points[0] = lo;
    for (int i = 0; i < nthr; ++i) points[i + 1] = thr[i];
    points[nthr + 1] = hi;
    int nseg = nthr + 1;
    if (nnu < nseg) return -1;
    double wsum = 0.0;
    for (int j = 0; j < nseg; ++j) { w[j] = log(points[j + 1]) - log(points[j]); wsum += w[j]; }
    int ntot = 0;
    for (int j = 0; j < nseg; ++j) {
        int f = (int)((double)nnu * w[j] / wsum);
        nb[j] = f > 1 ? f : 1;
        ntot += nb[j];
    }
    int r = nnu - ntot;
    while (r > 0) {
        int best = 0; double frac = -1.0;
        for (int j = 0; j < nseg; ++j) {
            double t = (double)nnu * w[j] / wsum - (double)nb[j];
            if (t > frac) { frac = t; best = j; }
        }
        nb[best]++; r--;
    }
    while (r < 0) {
        int best = -1; double frac = -1.0e30;
        for (int j = 0; j < nseg; ++j) {
            if (nb[j] <= 1) continue;
            double t = (double)nb[j] - (double)nnu * w[j] / wsum;
            if (t > frac) { frac = t; best = j; }
        }
        if (best < 0) break;
        nb[best]--; r++;
    }
    int ipt = 0;
    eedge[0] = points[0];
    for (int j = 0; j < nseg; ++j)
        for (int ib = 1; ib <= nb[j]; ++ib) {
            ++ipt;
            eedge[ipt] = exp(log(points[j]) + (log(points[j + 1]) - log(points[j])) *
                                                  (double)ib / (double)nb[j]);
        }
    return 0;
}

/* ion_setup + setup_continuous_source_sampler (src/ion_band_mod.f90:122-353,
 * 437-495, 516-615): Planck spectra only, no FUV extension */
int orc_setup_band(orc_ctx *c)
{
    const orc_params *p = &c->p;
    int nnu = p->nnu_ion;
    c->nnu_band = nnu;
    c->eedge = (double *)malloc(sizeof(double) * (nnu + 1));
    if (p->align_edges) {
        if (aligned_edges(c->eedge, nnu, p->eion_min, p->eion_max, p->nmetal, p->metal_eth))
            return -1;
    } else {
        double lo = log(p->eion_min), hi = log(p->eion_max);
        for (int i = 0; i <= nnu; ++i)
            c->eedge[i] = exp(lo + (hi - lo) * (double)i / (double)nnu);
    }
    c->ncomp = p->nsource + (p->ext_on ? 1 : 0);
    c->multi_src = c->ncomp >= 2;
    double band_L[ORC_MAXSRC + 1];
    /* photoionization_threshold_knots (src/ion_band_mod.f90:408-420) */
    double knots[3 + 64] = {ORC_ETH_HI, ORC_ETH_HEI, ORC_ETH_HEII};
    int nknots = 3;
    for (int i = 0; i < p->nmetal && i < 64; ++i) knots[nknots++] = p->metal_eth[i];
    for (int ic = 0; ic < c->ncomp; ++ic) {
        int ik = ic < p->nsource ? ic + 1 : 0;          /* points first, then external */
        c->comp_kind[ic] = ik;
        double T = ik >= 1 ? p->tstar[ik - 1] : p->ext_tstar;
        int rc = orc_build_planck(&c->samp[ic], T, p->eion_min, p->eion_max,
                                  p->source_cdf_tol, knots, nknots);
        if (rc) return rc;
        double f_ion = 1.0 - orc_energy_cdf_value(&c->samp[ic], p->eion_min);
        if (ik >= 1) {
            band_L[ic] = p->lum[ik - 1] / f_ion;
        } else {
            double asurf;
            if (p->ext_geometry == ORC_GEOM_EXT_SPH) {
                double r = p->rmax * p->distance2cm;
                asurf = ORC_FOURPI * (r * r);
            } else {
                double xr = 2.0 * p->xmax * p->distance2cm, yr = 2.0 * p->ymax * p->distance2cm,
                       zr = 2.0 * p->zmax * p->distance2cm;
                asurf = 2.0 * (xr * yr + yr * zr + zr * xr);
            }
            band_L[ic] = ORC_PI * p->ext_intensity * asurf / f_ion;
        }
    }
    c->Ltot = 0.0;
    for (int ic = 0; ic < c->ncomp; ++ic) c->Ltot += band_L[ic];
    if (c->multi_src) {
        c->src_cdf[0] = band_L[0];
        for (int ic = 1; ic < c->ncomp; ++ic) c->src_cdf[ic] = c->src_cdf[ic - 1] + band_L[ic];
        double last = c->src_cdf[c->ncomp - 1];
        for (int ic = 0; ic < c->ncomp; ++ic) c->src_cdf[ic] = c->src_cdf[ic] / last;
    }
    /* ion_qmc_ndim (src/ion_band_mod.f90:654-663) */
    c->qmc_ndim = (!c->multi_src && p->nsource == 1) ? 3 : 7;
    return 0;
}

/* ion_bin_of with include_fuv (src/ion_band_mod.f90:624-646) */
int orc_bin_of(const orc_ctx *c, double E)
{
    int lo = 1, hi = c->nnu_band;
    while (lo < hi) {
        int mid = (lo + hi + 1) / 2;
        if (c->eedge[mid - 1] <= E) lo = mid; else hi = mid - 1;
    }
    if (lo < 1) lo = 1;
    if (lo > c->nnu_band) lo = c->nnu_band;
    return lo;      /* 1-based */
}

static void alloc_leaf_arrays(orc_ctx *c)
{
    int n = c->nleaf, nnu = c->nnu_band > 0 ? c->nnu_band : 1;
    c->nH = (double *)calloc(n, sizeof(double));
    c->xHI = (double *)calloc(n, sizeof(double));
    c->xHeI = (double *)calloc(n, sizeof(double));
    c->xHeII = (double *)calloc(n, sizeof(double));
    c->ne = (double *)calloc(n, sizeof(double));
    c->Te = (double *)calloc(n, sizeof(double));
    c->rhokap = (double *)calloc(n, sizeof(double));
    c->hhe = (double *)calloc((size_t)n * 6, sizeof(double));
    c->jt = (double *)calloc((size_t)n * nnu, sizeof(double));
    c->dust_abs = (double *)calloc(n, sizeof(double));
    c->gHI = (double *)calloc(n, sizeof(double));
    c->gHeI = (double *)calloc(n, sizeof(double));
    c->gHeII = (double *)calloc(n, sizeof(double));
    c->hHI = (double *)calloc(n, sizeof(double));
    c->hHeI = (double *)calloc(n, sizeof(double));
    c->hHeII = (double *)calloc(n, sizeof(double));
}

/* namelist car grid (src/grid_mod_amr.f90:85-108) + amr_build_car
 * (src/octree_mod.f90:648-674) */
int orc_grid_car(orc_ctx *c)
{
    const orc_params *p = &c->p;
    c->car = 1;
    c->xmin = -p->xmax; c->xmaxb = p->xmax; c->xrange = c->xmaxb - c->xmin;
    c->ymin = -p->ymax; c->ymaxb = p->ymax; c->yrange = c->ymaxb - c->ymin;
    c->zmin = -p->zmax; c->zmaxb = p->zmax; c->zrange = c->zmaxb - c->zmin;
    c->dcell = (c->xmaxb - c->xmin) / (double)p->nx;
    c->nleaf = c->ncells = p->nx * p->ny * p->nz;
    alloc_leaf_arrays(c);
    return 0;
}

/* amr_find_cell_at_level (src/octree_mod.f90:678-698); indices 1-based */
static int find_cell_at_level(const orc_ctx *c, double x, double y, double z, int lev)
{
    if (x < c->xmin || x > c->xmaxb || y < c->ymin || y > c->ymaxb || z < c->zmin ||
        z > c->zmaxb)
        return 0;
    int ic = 1;
    for (;;) {
        if (c->level[ic - 1] >= lev) return ic;
        if (c->ileaf[ic - 1] > 0) return ic;
        int o = 1;
        if (x >= c->cx[ic - 1]) o += 1;
        if (y >= c->cy[ic - 1]) o += 2;
        if (z >= c->cz[ic - 1]) o += 4;
        int ch = c->children[(ic - 1) * 8 + o - 1];
        if (ch == 0) return ic;
        ic = ch;
    }
}

static int is_ancestor(const orc_ctx *c, int anc, int desc)
{
    int k = desc;
    while (k > 0) {
        k = c->parent[k - 1];
        if (k == anc) return 1;
    }
    return 0;
}

/* amr_build_tree + amr_build_neighbors (src/octree_mod.f90:347-497) */
int orc_grid_octree(orc_ctx *c, int nleaf, const double *xl, const double *yl,
                    const double *zl, const int *lev, double half)
{
    c->car = 0;
    c->xmin = -half; c->xmaxb = half; c->xrange = 2.0 * half;
    c->ymin = -half; c->ymaxb = half; c->yrange = 2.0 * half;
    c->zmin = -half; c->zmaxb = half; c->zrange = 2.0 * half;
    double Lbox = c->xmaxb - c->xmin;
    int cap = nleaf * 2 > 16 ? nleaf * 2 : 16;
    int *par = (int *)calloc(cap, sizeof(int)), *chl = (int *)calloc((size_t)cap * 8, sizeof(int));
    int *lv = (int *)calloc(cap, sizeof(int)), *il_of = (int *)calloc(cap, sizeof(int));
    double *cx = (double *)calloc(cap, sizeof(double)), *cy = (double *)calloc(cap, sizeof(double));
    double *cz = (double *)calloc(cap, sizeof(double)), *ch = (double *)calloc(cap, sizeof(double));
    int *cell_of = (int *)calloc(nleaf, sizeof(int));
    int nc = 1;
    cx[0] = (c->xmin + c->xmaxb) * 0.5; cy[0] = (c->ymin + c->ymaxb) * 0.5;
    cz[0] = (c->zmin + c->zmaxb) * 0.5; ch[0] = Lbox * 0.5;
    for (int il = 0; il < nleaf; ++il) {
        int ic = 1;
        for (int l = 0; l < lev[il]; ++l) {
            int o = 1;
            if (xl[il] >= cx[ic - 1]) o += 1;
            if (yl[il] >= cy[ic - 1]) o += 2;
            if (zl[il] >= cz[ic - 1]) o += 4;
            if (chl[(ic - 1) * 8 + o - 1] == 0) {
                if (nc + 1 > cap) {
                    int nc2 = cap * 2;
                    par = (int *)realloc(par, sizeof(int) * nc2);
                    chl = (int *)realloc(chl, sizeof(int) * 8 * (size_t)nc2);
                    lv = (int *)realloc(lv, sizeof(int) * nc2);
                    il_of = (int *)realloc(il_of, sizeof(int) * nc2);
                    cx = (double *)realloc(cx, sizeof(double) * nc2);
                    cy = (double *)realloc(cy, sizeof(double) * nc2);
                    cz = (double *)realloc(cz, sizeof(double) * nc2);
                    ch = (double *)realloc(ch, sizeof(double) * nc2);
                    memset(par + cap, 0, sizeof(int) * (nc2 - cap));
                    memset(chl + (size_t)cap * 8, 0, sizeof(int) * 8 * (size_t)(nc2 - cap));
                    memset(lv + cap, 0, sizeof(int) * (nc2 - cap));
                    memset(il_of + cap, 0, sizeof(int) * (nc2 - cap));
                    cap = nc2;
                }
                ++nc;
                int n = nc - 1;
                par[n] = ic;
                lv[n] = l + 1;
                il_of[n] = 0;
                ch[n] = ch[ic - 1] * 0.5;
                int ix = (o - 1) % 2, iy = ((o - 1) / 2) % 2, iz = (o - 1) / 4;
                cx[n] = cx[ic - 1] + (double)(2 * ix - 1) * ch[n];
                cy[n] = cy[ic - 1] + (double)(2 * iy - 1) * ch[n];
                cz[n] = cz[ic - 1] + (double)(2 * iz - 1) * ch[n];
                chl[(ic - 1) * 8 + o - 1] = nc;
            }
            ic = chl[(ic - 1) * 8 + o - 1];
        }
        il_of[ic - 1] = il + 1;
        cell_of[il] = ic;
    }
    c->ncells = nc;
    c->nleaf = nleaf;
    c->parent = par; c->children = chl; c->level = lv; c->ileaf = il_of;
    c->cx = cx; c->cy = cy; c->cz = cz; c->ch = ch; c->icell_of_leaf = cell_of;
    c->neighbor = (int *)calloc((size_t)nc * 6, sizeof(int));
    for (int ic = 1; ic <= nc; ++ic) {
        double x = cx[ic - 1], y = cy[ic - 1], z = cz[ic - 1], hp = 2.0 * ch[ic - 1];
        int l = lv[ic - 1];
        int *nb = c->neighbor + (size_t)(ic - 1) * 6;
        if (x + hp <= c->xmaxb) nb[0] = find_cell_at_level(c, x + hp, y, z, l);
        if (x - hp >= c->xmin) nb[1] = find_cell_at_level(c, x - hp, y, z, l);
        if (y + hp <= c->ymaxb) nb[2] = find_cell_at_level(c, x, y + hp, z, l);
        if (y - hp >= c->ymin) nb[3] = find_cell_at_level(c, x, y - hp, z, l);
        if (z + hp <= c->zmaxb) nb[4] = find_cell_at_level(c, x, y, z + hp, l);
        if (z - hp >= c->zmin) nb[5] = find_cell_at_level(c, x, y, z - hp, l);
        for (int f = 0; f < 6; ++f)
            if (nb[f] > 0 && nb[f] != ic && is_ancestor(c, nb[f], ic)) nb[f] = 0;
    }
    alloc_leaf_arrays(c);
    return 0;
}

/* gas_state_setup (src/gas_state_mod.f90:41-76) and the dust law
 * 'global_dgr' is applied by the caller through rhokap */
int orc_state_init(orc_ctx *c, const double *nH, double xHI0, double xHeI0, double xHeII0)
{
    for (int il = 0; il < c->nleaf; ++il) {
        c->nH[il] = nH[il];
        c->xHI[il] = xHI0;
        c->xHeI[il] = xHeI0;
        c->xHeII[il] = xHeII0;
        double xHeIII = 1.0 - xHeI0 - xHeII0;
        if (xHeIII < 0.0) xHeIII = 0.0;
        c->ne[il] = nH[il] * ((1.0 - xHI0) + c->p.He_abund * (xHeII0 + 2.0 * xHeIII));
        c->Te[il] = c->p.te_fixed;
    }
    return 0;
}

/* ---- dust optics (src/gas_opacity_mod.f90:69-290) ----------------------- */
static double interp_lin(const double *x, const double *y, int n, double x0)
{
    if (x0 <= x[0]) return y[0];
    if (x0 >= x[n - 1]) return y[n - 1];
    int i = 0;
    for (; i < n - 1; ++i) if (x0 < x[i + 1]) break;
    double w = (x0 - x[i]) / (x[i + 1] - x[i]);
    return y[i] * (1.0 - w) + y[i + 1] * w;
}

static double interp_log(const double *x, const double *y, int n, double x0)
{
    if (x0 <= x[0]) return y[0];
    if (x0 >= x[n - 1]) return y[n - 1];
    int i = 0;
    for (; i < n - 1; ++i) if (x0 < x[i + 1]) break;
    double w = log(x0 / x[i]) / log(x[i + 1] / x[i]);
    double a = y[i] > DBL_MIN ? y[i] : DBL_MIN, b = y[i + 1] > DBL_MIN ? y[i + 1] : DBL_MIN;
    return exp(log(a) * (1.0 - w) + log(b) * w);
}

/* ion_dust_read sorting + C_ext(lambda_ref) (src/gas_opacity_mod.f90:119-160) */
int orc_set_dust_table(orc_ctx *c, int n, const double *lam, const double *alb,
                       const double *g, const double *cext)
{
    c->dn = n;
    c->dlam = (double *)malloc(sizeof(double) * n);
    c->dalb = (double *)malloc(sizeof(double) * n);
    c->dg = (double *)malloc(sizeof(double) * n);
    c->dcext = (double *)malloc(sizeof(double) * n);
    memcpy(c->dlam, lam, sizeof(double) * n);
    memcpy(c->dalb, alb, sizeof(double) * n);
    memcpy(c->dg, g, sizeof(double) * n);
    memcpy(c->dcext, cext, sizeof(double) * n);
    for (int i = 1; i < n; ++i) {          /* sort4: stable insertion sort */
        double tx = c->dlam[i], t1 = c->dalb[i], t2 = c->dg[i], t3 = c->dcext[i];
        int j = i - 1;
        while (j >= 0 && c->dlam[j] > tx) {
            c->dlam[j + 1] = c->dlam[j]; c->dalb[j + 1] = c->dalb[j];
            c->dg[j + 1] = c->dg[j]; c->dcext[j + 1] = c->dcext[j];
            --j;
        }
        c->dlam[j + 1] = tx; c->dalb[j + 1] = t1; c->dg[j + 1] = t2; c->dcext[j + 1] = t3;
    }
    c->dcref = interp_log(c->dlam, c->dcext, n, c->p.lambda_ref);
    return 0;
}

/* ion_dust_{sabs,ssca,g}_at (src/gas_opacity_mod.f90:217-238) */
void orc_dust_at(const orc_ctx *c, double E, double *sabs, double *ssca, double *g)
{
    double lb = 1.23984 / E;
    double a = interp_lin(c->dlam, c->dalb, c->dn, lb);
    double ce = interp_log(c->dlam, c->dcext, c->dn, lb);
    *sabs = (1.0 - a) * ce / c->dcref;
    *ssca = a * ce / c->dcref;
    *g = interp_lin(c->dlam, c->dg, c->dn, lb);
}
