/*
 * Oracle: trace-metal registry and cascade (TEST INFRASTRUCTURE).
 *
 * Restates src/species_mod.f90 with the pieces of src/photo_xsec.f90,
 * src/recomb_mod.f90, src/cooling_mod.f90 and src/ion_balance_mod.f90 it
 * calls:
 *   species_setup (flat registry, threshold sort)      species_mod :143-228
 *   species_stage_cache_refresh                        :247-261
 *   species_packet_cross_sections, species_sigma       :263-275, :763-776
 *   species_cached_opacity                             :291-301 (orc_transport.c)
 *   species_rate_path_score                            :329-343 (orc_transport.c)
 *   alpha_rec, cx_rate                                 :428-486
 *   species_gamma_compute                              :493-512
 *   species_ne_prepare, species_metal_sums             :517-619
 *   species_ne_cached                                  :646-652
 *   metal_cooling_H, metal_heating                     :704-752
 *   species_fractions, voronov_ci                      :782-821
 *   metal_cooling, metal_freefree                      :825-868
 *   sigma_vy95                                         photo_xsec :73-89
 *   ci_dere_ratio                                      recomb_mod :288-311
 *   solve_ion_cell (metal branches),
 *   hydrogen_neutral_fraction, charge_neutrality_ne    ion_balance_mod :53-254
 * ci_model 'dere_hybrid' multiplies every Voronov rate, H and He included,
 * by ci_dere_ratio (the model's ci_dere switch, set for each model-level call).
 * cooling_model 'local_ne': each stage's fit is suppressed by its (T, n_e)
 * table when one is set (orc_set_stage_table; nlevel_cooling_apply,
 * src/species_mod.f90:836-840), else it enters unsuppressed ('low_density').
 *
 * Pinned by (tests/test_oracle_metals.py): sigma_vfky96 / sigma_vy95,
 * rr_badnell, voronov and ci_dere_ratio against the reference's
 * photo_xsec / recomb_mod compiled from source (oracle/_ref); the DR sum of
 * alpha_rec against the reference's He II DR (dr_HeII_badnell) fed the same
 * coefficients; the printed threshold cross sections of the element files;
 * charge-exchange detailed balance of the first transitions
 * (tests/charge_exchange/check_detailed_balance.f90: ratio within 35% at
 * 5e3-2e4 K where min k >= 1e-15); species_fractions against a dense linear
 * solve of the stage balance equations; the metal sums against the fractions
 * they are built from; hydrogen_neutral_fraction's root against the residual
 * it zeroes; one solve_ion_cell call reaching its own fixed point
 * (tests/charge_exchange/check_cell_convergence.f90); the two-level
 * H-impact cooling's low- and high-density limits.
 */
#include <math.h>
#include <float.h>
#include <stdlib.h>
#include <string.h>
#include "mc_oracle.h"

#define TINEST DBL_MIN

/* sigma_vy95 (src/photo_xsec.f90:73-89) */
double orc_sigma_vy95(double E, double Eth, double E0, double s0, double ya, double P,
                      double yw, double l)
{
    if (E < Eth) return 0.0;
    double y = E / E0;
    double Q = 5.5 + l - 0.5 * P;
    double ym = y - 1.0;
    double Fy = (ym * ym + yw * yw) * pow(y, -Q) * pow(1.0 + sqrt(y / ya), -P);
    return s0 * Fy * 1.0e-18;
}

/* species_sigma (src/species_mod.f90:763-776); it 0-based */
double orc_species_sigma(const orc_element *el, int it, double E)
{
    const double *p = el->photo[it];
    if (el->photo_form[it] == 2) return orc_sigma_vy95(E, p[0], p[1], p[2], p[3], p[4], p[5], p[6]);
    return orc_sigma_vfky96(E, p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7]);
}

/* ci_dere_ratio (src/recomb_mod.f90:288-311); stage = transition, 1-based */
double orc_ci_dere_ratio(int ci_dere, int Z, int stage)
{
    if (!ci_dere) return 1.0;
    double tab[4] = {1.0, 1.0, 1.0, 1.0};
    int n = 0;
    switch (Z) {
    case 1:  tab[0] = 0.9063; n = 1; break;
    case 2:  tab[0] = 1.0389; tab[1] = 1.0686; n = 2; break;
    case 6:  tab[0] = 1.0499; tab[1] = 0.913; tab[2] = 1.0377; n = 3; break;
    case 7:  tab[0] = 1.0421; tab[1] = 1.1966; n = 2; break;
    case 8:  tab[0] = 1.041; tab[1] = 1.1181; n = 2; break;
    case 10: tab[0] = 0.8089; tab[1] = 1.1395; n = 2; break;
    case 12: tab[0] = 0.3793; tab[1] = 0.9857; n = 2; break;
    case 14: tab[0] = 0.7328; tab[1] = 0.8798; tab[2] = 0.4492; tab[3] = 0.8221; n = 4; break;
    case 16: tab[0] = 1.3572; tab[1] = 0.8925; tab[2] = 0.8119; n = 3; break;
    case 17: tab[0] = 0.5412; tab[1] = 0.8428; tab[2] = 0.9237; tab[3] = 0.819; n = 4; break;
    case 18: tab[0] = 0.9242; tab[1] = 0.8644; tab[2] = 0.9752; n = 3; break;
    case 20: tab[0] = 0.7652; tab[1] = 1.1668; tab[2] = 1.0422; tab[3] = 0.8705; n = 4; break;
    case 26: tab[0] = 0.9904; tab[1] = 1.0568; tab[2] = 1.824; n = 3; break;
    default: break;
    }
    return (stage >= 1 && stage <= n) ? tab[stage - 1] : 1.0;
}

/* voronov_ci (src/species_mod.f90:815-821): c = dE, P, A, X, K */
double orc_voronov_ci(const double c[5], double T)
{
    double U = c[0] * ORC_EV2ERG / (ORC_KBOLTZ * T);
    return c[2] * (1.0 + c[1] * sqrt(U)) * pow(U, c[4]) * exp(-U) / (c[3] + U);
}

/* alpha_rec (src/species_mod.f90:428-447); it 0-based */
double orc_alpha_rec(const orc_element *el, int it, double T)
{
    double a;
    if (el->has_rr2[it]) {
        a = el->rr2[it][0] * pow(T / 1.0e4, -el->rr2[it][1]);
    } else {
        const double *r = el->rr[it];
        double s0 = sqrt(T / r[2]), s1 = sqrt(T / r[3]);
        double b = r[1] + r[4] * exp(-r[5] / T);
        a = r[0] / (s0 * pow(1.0 + s0, 1.0 - b) * pow(1.0 + s1, 1.0 + b));
    }
    for (int k = 0; k < el->ndr[it]; ++k)
        a = a + pow(T, -1.5) * el->dr_c[it][k] * exp(-el->dr_e[it][k] / T);
    if (el->has_dr2[it]) {
        const double *d = el->dr2[it];
        a = a + d[0] * pow(T, -1.5) * exp(-d[2] / T) * (1.0 + d[1] * exp(-d[3] / T));
    }
    return a;
}

/* cx_rate (src/species_mod.f90:452-486) */
double orc_cx_rate(int form, const double c[6], double T)
{
    double t4 = T / 1.0e4, k, lnT;
    switch (form) {
    case 1:
        k = c[0] * pow(T / 300.0, c[1]);
        if (c[2] != 0.0) k = k * exp(-c[2] / t4);
        return k;
    case 2:
        return c[0] * pow(t4, c[1]) + c[2] * pow(t4, c[3]);
    case 3:
        return (c[0] * pow(t4, c[1]) + c[2] * pow(t4, c[3])) * exp(-c[4] / T);
    case 4:
        lnT = log(T);
        return exp(c[0] + lnT * (c[1] + lnT * (c[2] + lnT * (c[3] + lnT * c[4]))));
    case 5:
        lnT = log(T);
        return exp(c[0] + lnT * (c[1] + lnT * (c[2] + lnT * (c[3] + lnT * c[4]))) - c[5] / t4);
    case 6:
        t4 = fmin(fmax(t4, 0.5), 5.0);
        k = c[0] * pow(t4, c[1]) * (1.0 + c[2] * exp(c[3] * t4));
        return fmax(k, 0.0);
    case 7:
        t4 = fmin(fmax(t4, 0.5), 5.0);
        k = c[0] * pow(t4, c[1]) * (1.0 + c[2] * exp(c[3] * t4)) * exp(-c[4] / t4);
        return fmax(k, 0.0);
    default:
        return 0.0;
    }
}

/* species_setup (src/species_mod.f90:143-228): registry, tallies and the
 * threshold-sorted flat transition list (stable insertion sort, :205-226) */
int orc_species_setup(orc_ctx *c, int nel, const orc_element *els)
{
    free(c->mraw); free(c->mgam); free(c->mheat); free(c->mfrac);
    c->mraw = c->mgam = c->mheat = c->mfrac = NULL;
    c->use_metals = 0; c->nel = 0; c->ntr = 0; c->nst = 0;
    if (nel <= 0) return 0;
    if (nel > ORC_MAX_EL) return -1;
    int nt = 0, ns = 0;
    for (int ie = 0; ie < nel; ++ie) {
        c->el[ie] = els[ie];
        if (els[ie].nstage < 1 || els[ie].nstage > ORC_MAX_ST) return -2;
        c->tr_off[ie] = nt;
        c->st_off[ie] = ns;
        for (int it = 0; it < els[ie].nstage - 1; ++it) {
            if (nt >= ORC_MAX_TR) return -3;
            c->flat_ie[nt] = ie; c->flat_it[nt] = it; c->flat_eth[nt] = els[ie].eth[it];
            ++nt;
        }
        ns += els[ie].nstage;
    }
    for (int i = 1; i < nt; ++i) {
        int ie = c->flat_ie[i], it = c->flat_it[i];
        double e = c->flat_eth[i];
        int j = i - 1;
        while (j >= 0 && c->flat_eth[j] > e) {
            c->flat_ie[j + 1] = c->flat_ie[j]; c->flat_it[j + 1] = c->flat_it[j];
            c->flat_eth[j + 1] = c->flat_eth[j];
            --j;
        }
        c->flat_ie[j + 1] = ie; c->flat_it[j + 1] = it; c->flat_eth[j + 1] = e;
    }
    c->nel = nel; c->ntr = nt; c->nst = ns;
    size_t nl = (size_t)c->nleaf;
    c->mraw = (double *)calloc(nl * nt * 2 + 1, sizeof(double));
    c->mgam = (double *)calloc(nl * nt + 1, sizeof(double));
    c->mheat = (double *)calloc(nl * nt + 1, sizeof(double));
    c->mfrac = (double *)calloc(nl * ns + 1, sizeof(double));
    c->use_metals = 1;
    return 0;
}

/* species_packet_cross_sections (src/species_mod.f90:263-275) */
void orc_species_packet_cross_sections(const orc_ctx *c, double E, double *sigma, int *ntr)
{
    int n = 0;
    for (int k = 0; k < c->ntr; ++k) {
        if (E < c->flat_eth[k]) break;
        sigma[k] = orc_species_sigma(&c->el[c->flat_ie[k]], c->flat_it[k], E);
        n = k + 1;
    }
    *ntr = n;
}

/* species_fractions (src/species_mod.f90:782-812); ie, il 0-based */
void orc_species_fractions(const orc_ctx *c, int ie, int il, double T, double ne, double nHI,
                           double nHII, double *frac)
{
    const orc_element *el = &c->el[ie];
    const int ns = el->nstage;
    double r[ORC_MAX_ST];
    for (int it = 0; it < ns - 1; ++it) {
        double rion = c->mgam[(size_t)il * c->ntr + c->tr_off[ie] + it] +
                      ne * orc_voronov_ci(el->ci[it], T) * orc_ci_dere_ratio(c->p.ci_dere, el->Z, it + 1);
        if (el->cxi_form[it] > 0) rion = rion + nHII * orc_cx_rate(el->cxi_form[it], el->cxi[it], T);
        double rrec = ne * orc_alpha_rec(el, it, T);
        if (el->cxr_form[it] > 0) rrec = rrec + nHI * orc_cx_rate(el->cxr_form[it], el->cxr[it], T);
        r[it] = rion / fmax(rrec, TINEST);
    }
    double s = 1.0, prod = 1.0;
    for (int i = 0; i < ns - 1; ++i) {
        prod = prod * r[i];
        s = s + prod;
    }
    frac[0] = 1.0 / s;
    for (int i = 1; i < ns; ++i) frac[i] = frac[i - 1] * r[i - 1];
}

/* species_ne_prepare (src/species_mod.f90:517-536) */
void orc_species_ne_prepare(const orc_ctx *c, int il, double T, orc_mcache *cc)
{
    for (int ie = 0; ie < c->nel; ++ie) {
        const orc_element *el = &c->el[ie];
        for (int it = 0; it < el->nstage - 1; ++it) {
            int t = c->tr_off[ie] + it;
            cc->gam[t] = c->mgam[(size_t)il * c->ntr + t];
            cc->ci[t] = orc_voronov_ci(el->ci[it], T) * orc_ci_dere_ratio(c->p.ci_dere, el->Z, it + 1);
            cc->rec[t] = orc_alpha_rec(el, it, T);
            cc->cxi[t] = el->cxi_form[it] > 0 ? orc_cx_rate(el->cxi_form[it], el->cxi[it], T) : 0.0;
            cc->cxr[t] = el->cxr_form[it] > 0 ? orc_cx_rate(el->cxr_form[it], el->cxr[it], T) : 0.0;
        }
    }
}

/* species_metal_sums (src/species_mod.f90:566-619) */
void orc_species_metal_sums(const orc_ctx *c, const orc_mcache *cc, double nH, double ne,
                            double nHI, double nHII, double *nem_out, double *rin_out,
                            double *rout_out)
{
    double nem = 0.0, rin = 0.0, rout = 0.0, r[ORC_MAX_ST];
    for (int ie = 0; ie < c->nel; ++ie) {
        const orc_element *el = &c->el[ie];
        const int ns = el->nstage, t0 = c->tr_off[ie];
        for (int it = 0; it < ns - 1; ++it) {
            int t = t0 + it;
            double rion = cc->gam[t] + ne * cc->ci[t];
            if (cc->cxi[t] > 0.0) rion = rion + nHII * cc->cxi[t];
            double rrec = ne * cc->rec[t];
            if (cc->cxr[t] > 0.0) rrec = rrec + nHI * cc->cxr[t];
            r[it] = rion / fmax(rrec, TINEST);
        }
        double s = 1.0, prod = 1.0;
        for (int i = 0; i < ns - 1; ++i) {
            prod = prod * r[i];
            s = s + prod;
        }
        double nX = el->abund * nH, fr = 1.0 / s;
        if (ns > 1) rin = rin + cc->cxi[t0] * nX * fr;
        for (int i = 2; i <= ns; ++i) {
            fr = fr * r[i - 2];
            nem = nem + el->abund * nH * fr * (double)(i - 1);
            rout = rout + cc->cxr[t0 + i - 2] * nX * fr;
            if (i < ns) rin = rin + cc->cxi[t0 + i - 1] * nX * fr;
        }
    }
    *nem_out = nem; *rin_out = rin; *rout_out = rout;
}

/* hydrogen_neutral_fraction (src/ion_balance_mod.f90:177-208): Illinois
 * false position on R(x) = x [B + R_in(x) + R_out(x)] - [A + R_in(x)] */
static double hydrogen_neutral_fraction(const orc_ctx *c, const orc_mcache *cc, double nH,
                                        double ne, double A, double B)
{
    double nem, rin, rout;
    double xa = 0.0;
    orc_species_metal_sums(c, cc, nH, ne, 0.0, nH, &nem, &rin, &rout);
    double Ra = -(A + rin);
    double xb = 1.0;
    orc_species_metal_sums(c, cc, nH, ne, nH, 0.0, &nem, &rin, &rout);
    double Rb = B + rout;
    double scale = A + B + fabs(Ra) + fabs(Rb);
    double x = xb;
    for (int k = 1; k <= 80; ++k) {
        if (Rb == Ra) break;
        x = xb - Rb * (xb - xa) / (Rb - Ra);
        x = fmin(fmax(x, fmin(xa, xb)), fmax(xa, xb));
        orc_species_metal_sums(c, cc, nH, ne, nH * x, nH * (1.0 - x), &nem, &rin, &rout);
        double Rx = x * (B + rin + rout) - (A + rin);
        if (fabs(Rx) <= 1.0e-15 * scale) break;
        if (fabs(xb - xa) <= 1.0e-15) break;
        if (Rx * Rb < 0.0) {
            xa = xb; Ra = Rb;
        } else {
            Ra = 0.5 * Ra;
        }
        xb = x; Rb = Rx;
    }
    return x;
}

/* solve_ion_cell (src/ion_balance_mod.f90:53-150) for leaf il (0-based),
 * with the metal-electron and charge-exchange terms when the registry is
 * loaded */
void orc_solve_ion_cell_leaf(const orc_ctx *c, int il, double gH, double gHe1, double gHe2,
                             double nH, double T, int caseA, double yHe, double *xHI,
                             double *xHeI, double *xHeII, double *ne_out)
{
    const int ci_prev = orc_ci_model_set(c->p.ci_dere);
    int with_ne = 0, with_cx = 0;
    if (il >= 0 && c->use_metals && c->nel > 0) {
        with_ne = c->p.metal_ne;
        with_cx = c->p.cx_two_way;
    }
    orc_mcache cc;
    if (with_ne || with_cx) orc_species_ne_prepare(c, il, T, &cc);
    double aH, aHe2, aHe3;
    if (caseA) { aH = orc_alphaA_HII(T); aHe2 = orc_alphaA_HeII(T); aHe3 = orc_alphaA_HeIII(T); }
    else { aH = orc_alphaB_HII(T); aHe2 = orc_alphaB_HeII(T); aHe3 = orc_alphaB_HeIII(T); }
    double cH = orc_ci_HI(T), cHe1 = orc_ci_HeI(T), cHe2 = orc_ci_HeII(T);
    double xHeIII = fmax(0.0, 1.0 - *xHeI - *xHeII);
    double ne = nH * ((1.0 - *xHI) + yHe * (*xHeII + 2.0 * xHeIII));
    ne = fmax(ne, 1.0e-12 * nH);
    double nem = 0.0, rin = 0.0, rout = 0.0;
    for (int it = 1; it <= 200; ++it) {
        double ne_old = ne;
        if (with_ne || with_cx) {
            orc_species_metal_sums(c, &cc, nH, ne_old, nH * *xHI, nH * (1.0 - *xHI), &nem, &rin,
                                   &rout);
            if (!with_cx) { rin = 0.0; rout = 0.0; }
        }
        if (with_cx) {
            *xHI = hydrogen_neutral_fraction(c, &cc, nH, ne, aH * ne, gH + (cH + aH) * ne);
        } else {
            double denH = gH + (cH + aH) * ne + rin + rout;
            *xHI = denH > 0.0 ? (aH * ne + rin) / denH : 1.0;
        }
        double r1 = fmin((gHe1 + cHe1 * ne) / (aHe2 * ne), 1.0e150);
        double r2 = fmin((gHe2 + cHe2 * ne) / (aHe3 * ne), 1.0e150);
        *xHeI = 1.0 / (1.0 + r1 + r1 * r2);
        *xHeII = *xHeI * r1;
        xHeIII = *xHeII * r2;
        ne = nH * ((1.0 - *xHI) + yHe * (*xHeII + 2.0 * xHeIII));
        if (with_ne) ne = ne + nem;
        ne = fmax(0.5 * (ne + ne_old), 1.0e-12 * nH);
        if (fabs(ne - ne_old) <= 1.0e-10 * ne) break;
    }
    *ne_out = ne;
    orc_ci_model_set(ci_prev);
}

/* charge_neutrality_ne (src/ion_balance_mod.f90:235-254) */
double orc_charge_neutrality_ne(const orc_ctx *c, int il, double T, double nH, double ne_HHe,
                                double xHI, double ne_seed)
{
    orc_mcache cc;
    orc_species_ne_prepare(c, il, T, &cc);
    double nHI = nH * xHI, nHII = nH * (1.0 - xHI);
    double ne = fmax(ne_seed, 1.0e-12 * nH), nem, rin, rout;
    for (int it = 1; it <= 200; ++it) {
        double ne_old = ne;
        orc_species_metal_sums(c, &cc, nH, ne_old, nHI, nHII, &nem, &rin, &rout);
        ne = ne_HHe + nem;
        ne = fmax(0.5 * (ne + ne_old), 1.0e-12 * nH);
        if (fabs(ne - ne_old) <= 1.0e-10 * ne) break;
    }
    return ne;
}

/* metal_cooling (src/species_mod.f90:825-846), cooling_model 'low_density':
 * cooling_eval = T^-1/2 sum A_i exp(-T_i/T) (src/cooling_mod.f90:125-134) */
double orc_metal_cooling(const orc_ctx *c, int il, double T, double nH, double ne, double nHI,
                         double nHII)
{
    double cool = 0.0, frac[ORC_MAX_ST];
    for (int ie = 0; ie < c->nel; ++ie) {
        const orc_element *el = &c->el[ie];
        orc_species_fractions(c, ie, il, T, ne, nHI, nHII, frac);
        double nel = el->abund * nH;
        for (int i = 0; i < el->nstage; ++i) {
            if (el->ncool[i] <= 0) continue;
            double lam = 0.0;
            for (int k = 0; k < el->ncool[i]; ++k) lam = lam + el->cool_A[i][k] * exp(-el->cool_T[i][k] / T);
            lam = lam / sqrt(T);
            cool = cool + ne * nel * frac[i] * orc_nlcool_apply(c->st_tab[ie][i], lam, T, ne);
        }
    }
    return cool;
}

/* metal_cooling_H (src/species_mod.f90:704-733): [C II] 158 um and [O I]
 * 63 um excited by H I collisions, two-level saturation */
double orc_metal_cooling_H(const orc_ctx *c, int il, double T, double nH, double ne,
                           double nHI, double nHII)
{
    const double kb = 1.380649e-16, A_CII = 2.321e-6, A_OI = 8.865e-5;
    double cool = 0.0, frac[ORC_MAX_ST];
    for (int ie = 0; ie < c->nel; ++ie) {
        const orc_element *el = &c->el[ie];
        if (strcmp(el->name, "c") == 0) {
            if (el->nstage < 2) continue;
            orc_species_fractions(c, ie, il, T, ne, nHI, nHII, frac);
            double nion = el->abund * nH * frac[1];
            double qul = 7.6e-10 * pow(T / 100.0, 0.14);
            double qlu = 2.0 * qul * exp(-91.25 / T);
            cool = cool + nion * nHI * qlu * (91.25 * kb) / (1.0 + nHI * (qul + qlu) / A_CII);
        } else if (strcmp(el->name, "o") == 0) {
            orc_species_fractions(c, ie, il, T, ne, nHI, nHII, frac);
            double nion = el->abund * nH * frac[0];
            double qul = 9.2e-11 * pow(T / 100.0, 0.67);
            double qlu = 0.6 * qul * exp(-227.7 / T);
            cool = cool + nion * nHI * qlu * (227.7 * kb) / (1.0 + nHI * (qul + qlu) / A_OI);
        }
    }
    return cool;
}

/* metal_heating (src/species_mod.f90:739-752) */
double orc_metal_heating(const orc_ctx *c, int il, double T, double nH, double ne, double nHI,
                         double nHII)
{
    double heat = 0.0, frac[ORC_MAX_ST];
    for (int ie = 0; ie < c->nel; ++ie) {
        const orc_element *el = &c->el[ie];
        orc_species_fractions(c, ie, il, T, ne, nHI, nHII, frac);
        for (int i = 0; i < el->nstage - 1; ++i)
            heat = heat + el->abund * nH * frac[i] * c->mheat[(size_t)il * c->ntr + c->tr_off[ie] + i];
    }
    return heat;
}

/* metal_freefree (src/species_mod.f90:852-868) */
double orc_metal_freefree(const orc_ctx *c, int il, double T, double nH, double ne, double nHI,
                          double nHII)
{
    double s = 0.0, frac[ORC_MAX_ST];
    for (int ie = 0; ie < c->nel; ++ie) {
        const orc_element *el = &c->el[ie];
        orc_species_fractions(c, ie, il, T, ne, nHI, nHII, frac);
        for (int j = 2; j <= el->nstage; ++j) {
            int zion = j - 1;
            s = s + el->abund * frac[j - 1] * (double)(zion * zion) *
                        orc_gbar_ff(c->cool.logT, c->cool.gff, T, zion);
        }
    }
    return 1.42554e-27 * sqrt(T) * ne * nH * s;
}

/* species_gamma_compute (src/species_mod.f90:493-512): raw (flat order) ->
 * egam / eheat (element order) per leaf */
void orc_species_gamma_compute(orc_ctx *c)
{
    if (!c->use_metals) return;
    double d2 = c->p.distance2cm * c->p.distance2cm;
    for (int il = 0; il < c->nleaf; ++il) {
        double s = c->car ? c->dcell : 2.0 * c->ch[c->icell_of_leaf[il] - 1];
        double vol = (s * s) * s;
        double fac = 1.0 / (vol * d2);
        for (int k = 0; k < c->ntr; ++k) {
            int t = c->tr_off[c->flat_ie[k]] + c->flat_it[k];
            const double *r = c->mraw + ((size_t)il * c->ntr + k) * 2;
            c->mgam[(size_t)il * c->ntr + t] = r[0] * fac;
            c->mheat[(size_t)il * c->ntr + t] = r[1] * fac;
        }
    }
}

/* species_stage_cache_refresh (src/species_mod.f90:247-261) */
void orc_species_stage_cache_refresh(orc_ctx *c)
{
    if (!c->use_metals) return;
    double frac[ORC_MAX_ST];
    for (int ie = 0; ie < c->nel; ++ie)
        for (int il = 0; il < c->nleaf; ++il) {
            double nHI = c->nH[il] * c->xHI[il], nHII = c->nH[il] * (1.0 - c->xHI[il]);
            orc_species_fractions(c, ie, il, c->Te[il], c->ne[il], nHI, nHII, frac);
            for (int i = 0; i < c->el[ie].nstage; ++i)
                c->mfrac[(size_t)il * c->nst + c->st_off[ie] + i] = frac[i];
        }
}

/* gas_opacity_fill (src/gas_opacity_mod.f90:299-306): the stage cache (the
 * dust law is fixed on this path) */
void orc_gas_opacity_fill(orc_ctx *c)
{
    if (c->use_metals && c->nel > 0) orc_species_stage_cache_refresh(c);
}

/* layout check of the shared element record */
int orc_element_size(void) { return (int)sizeof(orc_element); }
