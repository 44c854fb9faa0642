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
 * The diffuse launch itself (gen_diffuse_photon_qmc, diffuse_mod.f90:221-296)
 * sits beside the stellar launch in orc_transport.c.  The He I excited
 * channel (par%hei_diffuse, off by default) is not restated.
 *
 * Pinned by: the reference module compiled from source (oracle/_ref, tables
 * and sampler bit for bit); the reference's own sampler gates
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

int orc_diffuse_setup(orc_ctx *c)
{
    if (!c->p.diffuse_field) return 0;
    if (!c->milne) c->milne = (orc_milne *)calloc(1, sizeof(orc_milne));
    if (!c->milne) return -1;
    orc_milne_setup(c->milne, c->p.eion_max);
    free(c->dif_ch); free(c->dif_cdf);
    c->dif_ch = (double *)calloc((size_t)c->nleaf * 3, sizeof(double));
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

/* diffuse_build (src/diffuse_mod.f90:60-121), channels 1-3; lpacket =
 * ion_Ltot / nphotons */
void orc_diffuse_build(orc_ctx *c)
{
    const double y = c->p.He_abund, d2cm = c->p.distance2cm;
    const double eth[3] = {ORC_ETH_HI, ORC_ETH_HEI, ORC_ETH_HEII};
    for (int il = 1; il <= c->nleaf; ++il) {
        double *ch = c->dif_ch + (size_t)(il - 1) * 3;
        const double nH = c->nH[il - 1];
        if (nH <= 0.0) { ch[0] = ch[1] = ch[2] = 0.0; continue; }
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
    }
    double L = 0.0;
    for (int il = 0; il < c->nleaf; ++il) {
        const double *ch = c->dif_ch + (size_t)il * 3;
        L = L + (((ch[0] + ch[1]) + ch[2]) + 0.0);
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
