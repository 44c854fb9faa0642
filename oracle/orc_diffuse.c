/*
 * Oracle: explicit diffuse ground-recombination field (TEST INFRASTRUCTURE).
 *
 * Restates:
 *   src/milne_recomb_spectrum_mod.f90:124-350  Milne ground-continuum tables,
 *                                              in-band energy per
 *                                              recombination, CDF inversion
 *   src/diffuse_mod.f90:60-121, 327-355        per-leaf channel luminosities,
 *                                              leaf CDF, packet count,
 *                                              ground_mean_energy /
 *                                              ground_photon_energy
 *   src/diffuse_mod.f90:95-105, 308-324       the He I excited channel
 *                                              (par%hei_diffuse): channel-4
 *                                              luminosity, branch energies
 *   src/hei_cascade_mod.f90:62-81, 115-119    branch fractions, 584 A closure
 *   src/hei_twophoton_mod.f90:81-211          two-photon moments and sampler
 * The diffuse launch itself (gen_diffuse_photon_qmc, diffuse_mod.f90:221-296)
 * sits beside the stellar launch in orc_transport.c.
 *
 * Pinned by: the reference modules compiled from source (oracle/_ref: Milne
 * tables and sampler, He I branch fractions and 584 A closure bit for bit);
 * the two-photon moments the reference's module states (0.5564 photons and
 * 8.9589 eV per decay, mean 16.100 eV) and the exact integral of the
 * tabulated shape; the reference's own sampler gates
 * (tests/milne/check_milne.f90: range, energy-weighted mean against
 * quadrature, <1/E> <E>_n = 1, quantiles, continuity across table rows)
 * redone against scipy quadrature; the Milne alpha_1 against alpha_A -
 * alpha_B; and the closed-form channel luminosity of a uniform slab.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#include "mc_oracle.h"

#define MILNE_TLO 1.0e2
#define MILNE_THI 1.0e6
#define MILNE_XMAX 60.0

static double sigma_ch(int ich, double E)
{
    return ich == 0 ? orc_sigma_HI(E) : (ich == 1 ? orc_sigma_HeI(E) : orc_sigma_HeII(E));
}

/* photon-number emissivity shape n(x) = E^2 sigma(E) e^-x, E = E_th + x kT */
static double shape(const orc_milne *m, int ich, double x, double kT)
{
    double E = m->eth[ich] + x * kT;
    return ((E * E) * sigma_ch(ich, E)) * exp(-x);
}

/* milne_setup (src/milne_recomb_spectrum_mod.f90:124-206): cumulative
 * Simpson on the quadratically clustered x nodes, truncated at the band edge */
void orc_milne_setup(orc_milne *m, double eion_max)
{
    enum { NT = ORC_MILNE_NT, NX = ORC_MILNE_NX };
    m->eth[0] = ORC_ETH_HI; m->eth[1] = ORC_ETH_HEI; m->eth[2] = ORC_ETH_HEII;
    m->eion_max = eion_max;
    for (int ix = 1; ix <= NX; ++ix) {
        double t = (double)(ix - 1) / (double)(NX - 1);
        m->xg[ix - 1] = (MILNE_XMAX * t) * t;
    }
    const double l0 = log10(MILNE_TLO), l1 = log10(MILNE_THI);
    for (int it = 1; it <= NT; ++it)
        m->lgT[it - 1] = l0 + ((l1 - l0) * (double)(it - 1)) / (double)(NT - 1);
    for (int ich = 0; ich < 3; ++ich) {
        const double eth = m->eth[ich];
        for (int it = 0; it < NT; ++it) {
            double *cdf = m->cdf[ich][it];
            const double T = pow(10.0, m->lgT[it]);
            const double kT = ORC_KBOLTZ * T / ORC_EV2ERG;
            double xhi = (eion_max - eth) / kT;
            if (xhi > MILNE_XMAX) xhi = MILNE_XMAX;
            if (xhi <= 0.0) {
                for (int ix = 0; ix < NX; ++ix) cdf[ix] = 1.0;
                m->xbar[ich][it] = 0.0;
                m->fin[ich][it] = 0.0;
                continue;
            }
            double sn = 0.0, se = 0.0, sxn = 0.0;
            cdf[0] = 0.0;
            for (int ix = 0; ix < NX - 1; ++ix) {
                const double xa = m->xg[ix];
                const double xb = m->xg[ix + 1] < xhi ? m->xg[ix + 1] : xhi;
                if (xb > xa) {
                    const double xm = 0.5 * (xa + xb);
                    const double fa = shape(m, ich, xa, kT), fm = shape(m, ich, xm, kT),
                                 fb = shape(m, ich, xb, kT);
                    const double h6 = (xb - xa) / 6.0;
                    sn = sn + h6 * ((fa + 4.0 * fm) + fb);
                    se = se + h6 * ((fa * (eth + xa * kT) + (4.0 * fm) * (eth + xm * kT)) +
                                    fb * (eth + xb * kT));
                    sxn = sxn + h6 * ((fa * xa + (4.0 * fm) * xm) + fb * xb);
                }
                cdf[ix + 1] = se;
            }
            double tail = sn;
            if (xhi < MILNE_XMAX) {
                double xa = xhi;
                while (xa < MILNE_XMAX) {
                    double xb = xa + 0.25 < MILNE_XMAX ? xa + 0.25 : MILNE_XMAX;
                    double xm = 0.5 * (xa + xb);
                    tail = tail + ((xb - xa) / 6.0) *
                                      ((shape(m, ich, xa, kT) + 4.0 * shape(m, ich, xm, kT)) +
                                       shape(m, ich, xb, kT));
                    xa = xb;
                }
            }
            if (se > 0.0)
                for (int ix = 0; ix < NX; ++ix) cdf[ix] = cdf[ix] / se;
            else
                for (int ix = 0; ix < NX; ++ix) cdf[ix] = 1.0;
            cdf[NX - 1] = 1.0;
            m->xbar[ich][it] = sxn / (sn > DBL_MIN ? sn : DBL_MIN);
            m->fin[ich][it] = sn / (tail > DBL_MIN ? tail : DBL_MIN);
        }
    }
}

/* milne_tbracket (src/milne_recomb_spectrum_mod.f90:291-300); it 0-based */
static void tbracket(const orc_milne *m, double T, int *it, double *w)
{
    double s = (log10(T > DBL_MIN ? T : DBL_MIN) - m->lgT[0]) / (m->lgT[1] - m->lgT[0]);
    if (s < 0.0) s = 0.0;
    if (s > (double)(ORC_MILNE_NT - 1)) s = (double)(ORC_MILNE_NT - 1);
    int i = (int)s + 1;
    if (i > ORC_MILNE_NT - 1) i = ORC_MILNE_NT - 1;
    *w = s - (double)(i - 1);
    *it = i - 1;
}

/* milne_energy_per_recomb (src/milne_recomb_spectrum_mod.f90:307-318) [eV] */
double orc_milne_energy_per_recomb(const orc_milne *m, int ich, double T)
{
    int it;
    double w;
    tbracket(m, T, &it, &w);
    double xbar = (1.0 - w) * m->xbar[ich][it] + w * m->xbar[ich][it + 1];
    double fkeep = (1.0 - w) * m->fin[ich][it] + w * m->fin[ich][it + 1];
    double kT = ORC_KBOLTZ * T / ORC_EV2ERG;
    return fkeep * (m->eth[ich] + xbar * kT);
}

/* milne_sample_energy (src/milne_recomb_spectrum_mod.f90:320-350): CDF
 * interpolated in log T, inverted once by bisection + linear interpolation */
double orc_milne_sample_energy(const orc_milne *m, int ich, double T, double u)
{
    int it;
    double w;
    tbracket(m, T, &it, &w);
    const double *c0 = m->cdf[ich][it], *c1 = m->cdf[ich][it + 1];
    double kT = ORC_KBOLTZ * T / ORC_EV2ERG;
    double uu = u < 0.0 ? 0.0 : (u > 1.0 ? 1.0 : u);
    int lo = 0, hi = ORC_MILNE_NX - 1;
    double ca = (1.0 - w) * c0[lo] + w * c1[lo];
    double cb = (1.0 - w) * c0[hi] + w * c1[hi];
    while (hi - lo > 1) {
        int mid = (lo + 1 + hi + 1) / 2 - 1;
        double cm = (1.0 - w) * c0[mid] + w * c1[mid];
        if (uu > cm) { lo = mid; ca = cm; } else { hi = mid; cb = cm; }
    }
    double x = cb > ca ? m->xg[lo] + ((m->xg[hi] - m->xg[lo]) * (uu - ca)) / (cb - ca) : m->xg[lo];
    double E = m->eth[ich] + x * kT;
    double emax = m->eion_max * (1.0 - 1.0e-12);
    return E < emax ? E : emax;
}

/* ---- He I excited channel (par%hei_diffuse) ---------------------------- */
#define HEI_E_TOT 20.62            /* 2^1S - 1^1S [eV] */
#define HEI_E3S 19.82              /* 2^3S single photon [eV] */
#define HEI_E1P 21.22              /* 2^1P 584 A [eV] */

/* alpha_fit (src/hei_cascade_mod.f90:62-65): the nebular fits clamped to
 * [3e3, 2e4] K */
static double hei_alpha_fit(double T, double a, double b)
{
    double t = T > 3.0e3 ? T : 3.0e3;
    t = t < 2.0e4 ? t : 2.0e4;
    return a * pow(t / 1.0e4, b);
}

/* hei_branch_fractions (src/hei_cascade_mod.f90:72-81) */
void orc_hei_branch_fractions(double T, double *f3s, double *f1s, double *f1p)
{
    const double a3 = hei_alpha_fit(T, 2.10e-13, -0.381);
    const double a1s = hei_alpha_fit(T, 2.06e-14, -0.451);
    const double a1p = hei_alpha_fit(T, 4.17e-14, -0.695);
    const double sum = (a3 + a1s) + a1p;
    *f3s = a3 / sum;
    *f1s = a1s / sum;
    *f1p = a1p / sum;
}

/* hei_584_ionizes_H (src/hei_cascade_mod.f90:115-119) */
double orc_hei_584_ionizes_H(double T, double xHI, double xHeI)
{
    const double t = T > DBL_MIN ? T : DBL_MIN;
    const double he = xHeI > 0.0 ? xHeI : 0.0;
    const double h = xHI > DBL_MIN ? xHI : DBL_MIN;
    return 1.0 / (1.0 + ((77.0 / sqrt(t)) * he) / h);
}

/* shape_A (src/hei_twophoton_mod.f90:81-92): linear in the table, 0 outside */
static double hei_shape(const orc_hei2ph *h, double y)
{
    const int n = h->n_pts;
    if (y <= h->tab_y[0] || y >= h->tab_y[n - 1]) return 0.0;
    int i = 0;
    while (i < n - 2 && h->tab_y[i + 1] < y) ++i;
    return h->tab_A[i] + ((h->tab_A[i + 1] - h->tab_A[i]) * (y - h->tab_y[i])) /
                             (h->tab_y[i + 1] - h->tab_y[i]);
}

/* hei_2ph_setup (src/hei_twophoton_mod.f90:98-183) from the table rows:
 * Simpson moments of the ionizing window [13.598, min(20.62, eion_max)] eV
 * and the CDF of the energy-weighted density y A(y) on 1024 nodes */
int orc_hei2ph_setup(orc_hei2ph *h, int n, const double *y, const double *A, double eion_max)
{
    if (n < 4 || n > ORC_HEI2PH_MAXPTS) return -1;
    memset(h, 0, sizeof(*h));
    h->n_pts = n;
    for (int i = 0; i < n; ++i) { h->tab_y[i] = y[i]; h->tab_A[i] = A[i]; }
    double tot = 0.0;
    for (int i = 0; i < n - 1; ++i) {
        const double ya = h->tab_y[i], yb = h->tab_y[i + 1], ym = 0.5 * (ya + yb);
        tot = tot + ((yb - ya) / 6.0) * ((h->tab_A[i] + 4.0 * hei_shape(h, ym)) + h->tab_A[i + 1]);
    }
    if (!(tot > 0.0)) return -1;
    const double e_hi = HEI_E_TOT < eion_max ? HEI_E_TOT : eion_max;
    h->y_lo = ORC_ETH_HI / HEI_E_TOT;
    h->y_hi = e_hi / HEI_E_TOT;
    const int N = ORC_HEI2PH_NCDF;
    if (h->y_hi <= h->y_lo) {
        for (int i = 0; i < N; ++i) h->cdf_e[i] = 1.0;
        h->ready = 1;
        return 0;
    }
    const double span = h->y_hi - h->y_lo;
    double acc_n = 0.0, acc_e = 0.0;
    for (int i = 1; i <= N - 1; ++i) {
        const double ya = h->y_lo + (span * (double)(i - 1)) / (double)(N - 1);
        const double yb = h->y_lo + (span * (double)i) / (double)(N - 1);
        const double ym = 0.5 * (ya + yb);
        const double fa = hei_shape(h, ya), fm = hei_shape(h, ym), fb = hei_shape(h, yb);
        acc_n = acc_n + ((yb - ya) / 6.0) * ((fa + 4.0 * fm) + fb);
        acc_e = acc_e + (((yb - ya) / 6.0) * (((fa * ya) + (4.0 * fm) * ym) + fb * yb)) * HEI_E_TOT;
    }
    h->n_per_decay = (2.0 * acc_n) / tot;
    h->e_per_decay = (2.0 * acc_e) / tot;
    h->grid_y[0] = h->y_lo;
    h->cdf_e[0] = 0.0;
    double seg = 0.0;
    for (int i = 1; i <= N - 1; ++i) {
        const double ya = h->y_lo + (span * (double)(i - 1)) / (double)(N - 1);
        const double yb = h->y_lo + (span * (double)i) / (double)(N - 1);
        const double ym = 0.5 * (ya + yb);
        const double fa = hei_shape(h, ya) * ya, fm = hei_shape(h, ym) * ym,
                     fb = hei_shape(h, yb) * yb;
        seg = seg + ((yb - ya) / 6.0) * ((fa + 4.0 * fm) + fb);
        h->grid_y[i] = yb;
        h->cdf_e[i] = seg;
    }
    for (int i = 0; i < N; ++i) h->cdf_e[i] = seg > 0.0 ? h->cdf_e[i] / seg : 1.0;
    h->cdf_e[N - 1] = 1.0;
    h->ready = 1;
    return 0;
}

/* hei_2ph_sample_energy (src/hei_twophoton_mod.f90:190-211) [eV] */
double orc_hei2ph_sample_energy(const orc_hei2ph *h, double u)
{
    const double uu = u < 0.0 ? 0.0 : (u > 1.0 ? 1.0 : u);
    int lo = 0, hi = ORC_HEI2PH_NCDF - 1;
    while (hi - lo > 1) {
        const int mid = (lo + hi) / 2;
        if (uu > h->cdf_e[mid]) lo = mid; else hi = mid;
    }
    double y;
    if (h->cdf_e[hi] > h->cdf_e[lo])
        y = h->grid_y[lo] + ((h->grid_y[hi] - h->grid_y[lo]) * (uu - h->cdf_e[lo])) /
                                (h->cdf_e[hi] - h->cdf_e[lo]);
    else
        y = h->grid_y[lo];
    return y * HEI_E_TOT;
}

/* hei_branch_energies (src/diffuse_mod.f90:308-324): energy radiated per
 * He II recombination in each H-ionizing branch [eV] */
void orc_hei_branch_energies(const orc_hei2ph *h, double T, double xHI, double xHeI,
                             double *e3s, double *e1p, double *e1s)
{
    double f3s, f1s, f1p;
    orc_hei_branch_fractions(T, &f3s, &f1s, &f1p);
    const double p584 = orc_hei_584_ionizes_H(T, xHI, xHeI);
    *e3s = f3s * HEI_E3S;
    *e1p = (f1p * p584) * HEI_E1P;
    *e1s = (f1s + f1p * (1.0 - p584)) * h->e_per_decay;
}

int orc_diffuse_hei_setup(orc_ctx *c, int n, const double *y, const double *A)
{
    if (!c->hei) c->hei = (orc_hei2ph *)calloc(1, sizeof(orc_hei2ph));
    if (!c->hei) return -1;
    return orc_hei2ph_setup(c->hei, n, y, A, c->p.eion_max);
}

int orc_diffuse_setup(orc_ctx *c)
{
    if (!c->p.diffuse_field) return 0;
    if (!c->milne) c->milne = (orc_milne *)calloc(1, sizeof(orc_milne));
    if (!c->milne) return -1;
    orc_milne_setup(c->milne, c->p.eion_max);
    free(c->dif_ch); free(c->dif_cdf);
    c->dif_ch = (double *)calloc((size_t)c->nleaf * 4, sizeof(double));
    c->dif_cdf = (double *)calloc((size_t)c->nleaf, sizeof(double));
    return c->dif_ch && c->dif_cdf ? 0 : -1;
}

/* ground_mean_energy (src/diffuse_mod.f90:327-339) [eV] */
static double ground_mean_energy(const orc_ctx *c, int ich, double eth, double kT, double T)
{
    if (c->p.diffuse_model == 2) return orc_milne_energy_per_recomb(c->milne, ich, T);
    if (c->p.diffuse_model == 1) return eth + kT;
    return eth;
}

static double leaf_half_d(const orc_ctx *c, int il)
{
    return c->car ? 0.5 * c->dcell : c->ch[c->icell_of_leaf[il - 1] - 1];
}

/* diffuse_build (src/diffuse_mod.f90:60-121), channels 1-4; lpacket =
 * ion_Ltot / nphotons */
void orc_diffuse_build(orc_ctx *c)
{
    const double y = c->p.He_abund, d2cm = c->p.distance2cm;
    const double eth[3] = {ORC_ETH_HI, ORC_ETH_HEI, ORC_ETH_HEII};
    const int hei = c->p.hei_diffuse && c->hei && c->hei->ready;
    for (int il = 1; il <= c->nleaf; ++il) {
        double *ch = c->dif_ch + (size_t)(il - 1) * 4;
        const double nH = c->nH[il - 1];
        if (nH <= 0.0) { ch[0] = ch[1] = ch[2] = ch[3] = 0.0; continue; }
        const double T = c->Te[il - 1], ne = c->ne[il - 1];
        const double s = (2.0 * leaf_half_d(c, il)) * d2cm;
        const double vol = (s * s) * s;
        const double kT = ORC_KBOLTZ * T / ORC_EV2ERG;
        double x3 = (1.0 - c->xHeI[il - 1]) - c->xHeII[il - 1];
        if (x3 < 0.0) x3 = 0.0;
        const double nHII = nH * (1.0 - c->xHI[il - 1]);
        const double nHeII = (nH * y) * c->xHeII[il - 1];
        const double nHeIII = (nH * y) * x3;
        double a1 = orc_alphaA_HII(T) - orc_alphaB_HII(T);
        ch[0] = ((((ne * nHII) * a1) * ground_mean_energy(c, 0, eth[0], kT, T)) * ORC_EV2ERG) * vol;
        a1 = orc_alphaA_HeII(T) - orc_alphaB_HeII(T);
        ch[1] = ((((ne * nHeII) * a1) * ground_mean_energy(c, 1, eth[1], kT, T)) * ORC_EV2ERG) * vol;
        a1 = orc_alphaA_HeIII(T) - orc_alphaB_HeIII(T);
        ch[2] = ((((ne * nHeIII) * a1) * ground_mean_energy(c, 2, eth[2], kT, T)) * ORC_EV2ERG) * vol;
        ch[3] = 0.0;
        if (hei) {
            double be3s, be1p, be1s;
            a1 = orc_alphaB_HeII(T);
            orc_hei_branch_energies(c->hei, T, c->xHI[il - 1], c->xHeI[il - 1], &be3s, &be1p,
                                    &be1s);
            ch[3] = ((((ne * nHeII) * a1) * vol) * ORC_EV2ERG) * ((be3s + be1p) + be1s);
        }
    }
    double L = 0.0;
    for (int il = 0; il < c->nleaf; ++il) {
        const double *ch = c->dif_ch + (size_t)il * 4;
        L = L + (((ch[0] + ch[1]) + ch[2]) + ch[3]);
        c->dif_cdf[il] = L;
    }
    c->diffuse_lum = L;
    if (L > 0.0) {
        for (int il = 0; il < c->nleaf; ++il) c->dif_cdf[il] = c->dif_cdf[il] / L;
        c->diffuse_nphot = (int64_t)llround(L / (c->Ltot / (double)c->p.nphotons));
    } else {
        c->diffuse_nphot = 0;
    }
}
