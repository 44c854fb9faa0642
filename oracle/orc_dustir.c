/*
 * Oracle: the dust thermal-infrared band (TEST INFRASTRUCTURE).
 *
 * Restates, for the equilibrium modified-blackbody emission (par%dust_sed
 * off) of the extinction table the ionizing band already holds:
 *   src/dust_temp_mod.f90:39-177           dust_temp_setup (femit on 61
 *                                          log-T points), planck_nu,
 *                                          femit_of, dust_temp_compute
 *   src/dust_emis_transport_mod.f90:100-180 dust_emis_transport_setup
 *                                          (0.1-3000 um grid, band cut)
 *   src/dust_emis_transport_mod.f90:186-201 gas_absorption_at
 *   src/dust_emis_transport_mod.f90:208-366 transport_dust_emission (the
 *                                          self-heating iteration)
 *   src/dust_emis_transport_mod.f90:372-483 leaf spectrum, trapezoid CDF,
 *                                          continuous in-bin sampling
 *   src/dust_ir_band_mod.f90:63-84,182-192 band edges of the escaped
 *                                          spectrum, dust_ir_heating
 *   src/ion_score_mod.f90:319-338          ion_score_apply_dust_heat
 * The packet launch and transport of one pass (launch_and_transport,
 * src/dust_emis_transport_mod.f90:489-546, and the band switch of
 * score_ion_path / ion_score_energy_packet) sit in orc_transport.c beside the
 * walk they reuse.
 *
 * Random numbers.  The reference draws the stratification offset on rank 0
 * and every launch and scattering number from MT19937 in leaf order.  Here
 * every draw comes from orc_rng_uniform on stream 3 keyed by
 * (pass << 40 | packet index k): draw 0 of key (pass << 40 | 0) is the
 * offset xi, draws 0-5 of packet k its launch (x, y, z, cos theta, phi,
 * wavelength), draws 6.. its scattering.  The realisation differs from the
 * reference's; the estimator and every count do not depend on the stream.
 *
 * Pinned by (tests/test_oracle_dustir.py):
 *   - femit against 4 pi s_abs int B_nu dnu by quadrature for a grey table,
 *     and against 4 s_abs sigma_SB T^4 where the table covers the spectrum;
 *     femit_of exact at the knots and a T^4 power law in between (grey);
 *   - dust_temp_compute against the closed-form grey inverse
 *     (heat / (4 s sigma kd))^(1/4) to the bisection tolerance;
 *   - the emission cut: the first node at or beyond the band floor whose gas
 *     cross sections all vanish, the C I / S I metal thresholds with metals;
 *   - sample_linear_spectrum_bin against the inverse of its own linear CDF
 *     (round trip) and its three limiting forms (d -> 0, p0 = 0, p1 = 0);
 *   - the stratified counts: they sum to N and each lies in
 *     {floor, ceil}(N L_il / L_tot);
 *   - the transport: IR energy closure <= 5e-13, and without scattering the
 *     path estimator sum_il heat_ir V equals the extinction ledger (the
 *     identity int kappa_abs e^-tau dl = 1 - e^-tau_edge per packet);
 *   - the sampled wavelengths against the leaf's piecewise-linear density
 *     (chi-square) and the launch positions and directions (moments).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#include "mc_oracle.h"

#define CLIGHT 2.99792458e10

/* ---- dust_temp_mod ------------------------------------------------------ */

/* planck_nu (src/dust_temp_mod.f90:121-130) */
static double planck_nu(double nu, double T)
{
    double x = ORC_HPLANCK * nu / (ORC_KBOLTZ * T);
    if (x < 700.0) {
        double nu3 = nu * nu * nu;
        return 2.0 * ORC_HPLANCK * nu3 / (CLIGHT * CLIGHT) / (exp(x) - 1.0);
    }
    return 0.0;
}

static orc_dustir *ir_state(orc_ctx *c)
{
    if (!c->ir) c->ir = (orc_dustir *)calloc(1, sizeof(orc_dustir));
    return c->ir;
}

/* dust_temp_setup (src/dust_temp_mod.f90:39-118): the equilibrium emission
 * function femit(T) = 4 pi int s_abs(nu) B_nu(T) dnu [erg/s/cm^2 per unit
 * reference extinction] on 61 log-spaced T from 10^0.5 to 10^3.5 K, by the
 * trapezoid over the sorted extinction table (the same rows the ionizing
 * band reads; the table's albedo and C_ext, C_ext(lambda_ref) = dcref) */
int orc_dust_temp_setup(orc_ctx *c)
{
    if (c->dn < 2) return -1;
    orc_dustir *ir = ir_state(c);
    const double *lam = c->dlam, *alb = c->dalb, *cext = c->dcext;
    double cref = c->dcref;
    for (int i = 0; i < ORC_DT_NT; ++i) {
        ir->tgrid[i] = pow(10.0, 0.5 + 3.0 * (double)i / (double)(ORC_DT_NT - 1));
        double T = ir->tgrid[i], f = 0.0;
        for (int k = 0; k < c->dn - 1; ++k) {
            double nu1 = CLIGHT / (lam[k + 1] * 1.0e-4);
            double nu2 = CLIGHT / (lam[k] * 1.0e-4);
            double snu = 0.5 * ((1.0 - alb[k]) * cext[k] + (1.0 - alb[k + 1]) * cext[k + 1]) / cref;
            double bnu = 0.5 * (planck_nu(nu1, T) + planck_nu(nu2, T));
            f = f + ORC_FOURPI * snu * bnu * (nu2 - nu1);
        }
        ir->femit[i] = f;
    }
    ir->temp_ready = 1;
    return 0;
}

/* femit_of (src/dust_temp_mod.f90:132-140): log-log interpolation in T */
double orc_dust_femit_of(const orc_ctx *c, double T)
{
    const orc_dustir *ir = c->ir;
    const double *tg = ir->tgrid;
    double l1 = log10(tg[0]), ln = log10(tg[ORC_DT_NT - 1]);
    double lt = log10(T);
    lt = lt > l1 ? lt : l1;
    lt = lt < ln ? lt : ln;
    int i = (int)((lt - l1) / (3.0 / (double)(ORC_DT_NT - 1))) + 1;
    if (i > ORC_DT_NT - 1) i = ORC_DT_NT - 1;
    /* 1-based i -> 0-based i-1 */
    double li = log10(tg[i - 1]), lj = log10(tg[i]);
    double w = (lt - li) / (lj - li);
    return exp(log(ir->femit[i - 1]) * (1.0 - w) + log(ir->femit[i]) * w);
}

/* dust_temp_compute (src/dust_temp_mod.f90:145-177): per leaf the T_d with
 * femit(T_d) kd = heat_dust, kd = rhokap / distance2cm, by bisection in log T
 * on [tgrid(1), tgrid(NT)] to thi/tlo - 1 < 1e-5 (at most 60 halvings);
 * T_d = 0 where kd or heat_dust is not positive */
void orc_dust_temp_compute(orc_ctx *c, const double *heat)
{
    orc_dustir *ir = c->ir;
    if (!ir->tdust) ir->tdust = (double *)calloc((size_t)c->nleaf, sizeof(double));
    const double tlo0 = ir->tgrid[0], thi0 = ir->tgrid[ORC_DT_NT - 1];
    for (int il = 0; il < c->nleaf; ++il) {
        ir->tdust[il] = 0.0;
        double kd = c->rhokap[il] / c->p.distance2cm;
        if (kd <= 0.0 || heat[il] <= 0.0) continue;
        double target = heat[il] / kd;
        double tlo = tlo0, thi = thi0;
        if (orc_dust_femit_of(c, thi) <= target) {
            ir->tdust[il] = thi;
            continue;
        }
        for (int it = 1; it <= 60; ++it) {
            double tm = sqrt(tlo * thi);
            if (orc_dust_femit_of(c, tm) < target) tlo = tm;
            else thi = tm;
            if (thi / tlo - 1.0 < 1.0e-5) break;
        }
        ir->tdust[il] = sqrt(tlo * thi);
    }
}

/* ion_score_apply_dust_heat (src/ion_score_mod.f90:319-338): EUV grain
 * heating from the dust-absorption path tally of the last ionizing pass */
void orc_dust_heat_compute(const orc_ctx *c, double *heat)
{
    double d2 = c->p.distance2cm * c->p.distance2cm;
    for (int il = 1; il <= c->nleaf; ++il) {
        double h = c->car ? 0.5 * c->dcell : c->ch[c->icell_of_leaf[il - 1] - 1];
        double s = 2.0 * h;
        double fac = 1.0 / ((s * s) * s * d2);
        heat[il - 1] = c->dust_abs[il - 1] * fac * c->rhokap[il - 1] / c->p.distance2cm;
    }
}

/* ---- dust_emis_transport_mod -------------------------------------------- */

/* gas_absorption_at (src/dust_emis_transport_mod.f90:186-201): the total
 * gas photoionization cross section at E, zero below every threshold */
static double gas_absorption_at(const orc_ctx *c, double E)
{
    double s = orc_sigma_HI(E) + orc_sigma_HeI(E) + orc_sigma_HeII(E);
    if (c->use_metals) {
        double sm[ORC_MAX_TR];
        int n = 0;
        orc_species_packet_cross_sections(c, E, sm, &n);
        if (n > 0) {
            double t = 0.0;
            for (int k = 0; k < n; ++k) t = t + sm[k];
            s = s + t;
        }
    }
    return s;
}

/* dust_emis_transport_setup (src/dust_emis_transport_mod.f90:100-180) for
 * the equilibrium spectrum, and dust_ir_band_setup
 * (src/dust_ir_band_mod.f90:63-84): NLAM_BB nodes log-spaced from
 * max(0.1 um, band floor) to 3000 um, s_abs(lambda) of the extinction table,
 * the band cut (nodes 1..kpass zeroed, kpass = first node at or beyond the
 * band floor with no gas absorption) and the geometric-midpoint bin edges of
 * the escaped spectrum */
int orc_dust_emis_transport_setup(orc_ctx *c)
{
    if (c->dn < 2) return -1;
    orc_dustir *ir = ir_state(c);
    double E_floor = c->p.eion_min;
    double lam_floor = 1.23984 / E_floor;
    double lam_start = lam_floor > ORC_LAM_BB_MIN ? lam_floor : ORC_LAM_BB_MIN;
    int n = ORC_NLAM_BB;
    free(ir->lam); free(ir->sabs); free(ir->edge); free(ir->lesc);
    ir->nl = n;
    ir->lam = (double *)malloc(sizeof(double) * n);
    ir->sabs = (double *)malloc(sizeof(double) * n);
    ir->edge = (double *)malloc(sizeof(double) * (n + 1));
    ir->lesc = (double *)calloc((size_t)n, sizeof(double));
    for (int k = 0; k < n; ++k) {
        ir->lam[k] = lam_start * pow(ORC_LAM_BB_MAX / lam_start, (double)k / (double)(n - 1));
        double sa, ss, g;
        orc_dust_at(c, 1.23984 / ir->lam[k], &sa, &ss, &g);
        ir->sabs[k] = sa;
    }
    int kpass = 0;                                 /* 1-based, 0 = none */
    for (int k = 0; k < n; ++k) {
        if (ir->lam[k] < lam_floor) continue;
        if (gas_absorption_at(c, 1.23984 / ir->lam[k]) <= 0.0) {
            kpass = k + 1;
            break;
        }
    }
    if (kpass == 0) return -2;
    ir->nzero = kpass > 1 ? kpass : 0;
    ir->lam_hardest = ir->lam[kpass - 1];
    for (int k = 1; k < n; ++k) ir->edge[k] = sqrt(ir->lam[k - 1] * ir->lam[k]);
    ir->edge[0] = ir->lam[0] * sqrt(ir->lam[0] / ir->lam[1]);
    ir->edge[n] = ir->lam[n - 1] * sqrt(ir->lam[n - 1] / ir->lam[n - 2]);
    ir->frac_discarded = 0.0;
    free(ir->irabs); free(ir->heat_ir); free(ir->khi);
    ir->irabs = (double *)calloc((size_t)c->nleaf, sizeof(double));
    ir->heat_ir = (double *)calloc((size_t)c->nleaf, sizeof(double));
    ir->khi = (int64_t *)calloc((size_t)c->nleaf + 1, sizeof(int64_t));
    ir->setup_ready = 1;
    return 0;
}

/* planck_lambda (src/dust_emis_transport_mod.f90:418-429): B_lambda(T) up to
 * a constant, lambda in um; lam_cm^5 as (l^2)^2 l */
static double planck_lambda(double lam_um, double T)
{
    double lc = lam_um * 1.0e-4;
    double x = ORC_HPLANCK * CLIGHT / (lc * ORC_KBOLTZ * T);
    if (x > 700.0) return 0.0;
    double l2 = lc * lc;
    double l5 = l2 * l2 * lc;
    return 1.0 / (l5 * (exp(x) - 1.0));
}

/* leaf_emission_spectrum + equilibrium_leaf_spectrum + emission_cumulative
 * (src/dust_emis_transport_mod.f90:372-413, 432-457): pdf(k) = s_abs(k)
 * B_lambda(lam(k), T_d), the band-cut nodes zeroed, cdf by the trapezoid;
 * returns esum = cdf(nl) (0 when T_d is not positive) and the emission
 * fraction the cut removed */
double orc_ir_leaf_spectrum(const orc_ctx *c, int il, double *pdf, double *cdf, double *fdisc)
{
    const orc_dustir *ir = c->ir;
    int n = ir->nl;
    *fdisc = 0.0;
    double T = ir->tdust ? ir->tdust[il] : 0.0;
    for (int k = 0; k < n; ++k) pdf[k] = 0.0;
    if (!(T > 0.0)) { for (int k = 0; k < n; ++k) cdf[k] = 0.0; return 0.0; }
    for (int k = 0; k < n; ++k) pdf[k] = ir->sabs[k] * planck_lambda(ir->lam[k], T);
    double esum = 0.0;
    for (int k = 0; k < n - 1; ++k)
        esum = esum + 0.5 * (pdf[k] + pdf[k + 1]) * (ir->lam[k + 1] - ir->lam[k]);
    if (ir->nzero > 0 && esum > 0.0) {
        double full = esum;
        for (int k = 0; k < ir->nzero; ++k) pdf[k] = 0.0;
        esum = 0.0;
        for (int k = 0; k < n - 1; ++k)
            esum = esum + 0.5 * (pdf[k] + pdf[k + 1]) * (ir->lam[k + 1] - ir->lam[k]);
        *fdisc = (full - esum) / full;
    }
    cdf[0] = 0.0;
    for (int k = 0; k < n - 1; ++k)
        cdf[k + 1] = cdf[k] + 0.5 * (pdf[k] + pdf[k + 1]) * (ir->lam[k + 1] - ir->lam[k]);
    return esum;
}

/* sample_linear_spectrum_bin (src/dust_emis_transport_mod.f90:467-483) */
double orc_sample_linear_bin(double lam0, double lam1, double p0, double p1, double u)
{
    double uu = u < 0.0 ? 0.0 : (u > 1.0 ? 1.0 : u);
    if (p0 + p1 <= 0.0) return lam0 + uu * (lam1 - lam0);
    double d = p1 - p0;
    double twoc = uu * (p0 + p1);
    double disc = p0 * p0 + d * twoc;
    double t = twoc / (p0 + sqrt(disc > 0.0 ? disc : 0.0));
    t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
    return lam0 + t * (lam1 - lam0);
}

/* per-leaf cumulative emission luminosity and the stratified packet counts
 * of one pass (src/dust_emis_transport_mod.f90:252-290): khi[il] = the last
 * packet index of leaf il (khi[-1] = 0), k_lo(il) = khi(il-1) + 1; returns
 * L_tot and writes the offset xi */
double orc_ir_counts(orc_ctx *c, int pass, int64_t N, const double *heat, double *xi_out)
{
    orc_dustir *ir = c->ir;
    int nl = c->nleaf;
    double *lcum = (double *)malloc(sizeof(double) * ((size_t)nl + 1));
    lcum[0] = 0.0;
    for (int il = 1; il <= nl; ++il) {
        double h = c->car ? 0.5 * c->dcell : c->ch[c->icell_of_leaf[il - 1] - 1];
        double w = 2.0 * h * c->p.distance2cm;
        double vol = (w * w) * w;
        double hd = heat[il - 1] > 0.0 ? heat[il - 1] : 0.0;
        lcum[il] = lcum[il - 1] + hd * vol;
    }
    double L = lcum[nl];
    double xi = orc_rng_uniform((uint32_t)c->p.qmc_seed, 3, (uint64_t)pass << 40, 0);
    xi = xi > 1.0e-12 ? xi : 1.0e-12;
    xi = xi < 1.0 - 1.0e-12 ? xi : 1.0 - 1.0e-12;
    *xi_out = xi;
    if (L > 0.0)
        for (int il = 1; il <= nl; ++il)
            ir->khi[il - 1] = (int64_t)ceil((double)N * (lcum[il] / L) - xi);
    free(lcum);
    return L;
}

/* dust_ir_heating (src/dust_ir_band_mod.f90:182-192) */
static void ir_heating(const orc_ctx *c, const double *irabs, double *heat_ir)
{
    double d2 = c->p.distance2cm * c->p.distance2cm;
    for (int il = 1; il <= c->nleaf; ++il) {
        double h = c->car ? 0.5 * c->dcell : c->ch[c->icell_of_leaf[il - 1] - 1];
        double s = 2.0 * h;
        double fac = 1.0 / ((s * s) * s * d2);
        heat_ir[il - 1] = irabs[il - 1] * fac * c->rhokap[il - 1] / c->p.distance2cm;
    }
}

/* transport_dust_emission (src/dust_emis_transport_mod.f90:208-366): heat
 * enters as the EUV grain heating and leaves as EUV + reabsorbed IR, with
 * T_d consistent with it.  report[ORC_IR_NREPORT]: passes, converged, L_emit
 * of the last pass, IR emitted / extinguished / escaped, closure, the
 * path-estimator absorbed power, the relative change of the emitted total,
 * the largest emission fraction the band cut removed.  Returns -3 when the
 * closure exceeds 5e-13 (the reference aborts). */
int orc_transport_dust_emission(orc_ctx *c, double *heat, int64_t npacket, int niter,
                                double tol, int nthreads, double *report)
{
    orc_dustir *ir = c->ir;
    if (!ir || !ir->temp_ready || !ir->setup_ready) return -1;
    int nl = c->nleaf;
    double *heat_euv = (double *)malloc(sizeof(double) * nl);
    memcpy(heat_euv, heat, sizeof(double) * nl);
    int64_t N = npacket > 1 ? npacket : 1;
    if (niter < 1) niter = 1;
    double L_prev = 0.0;
    int converged = 0, done = 0, rc = 0;
    for (int k = 0; k < ORC_IR_NREPORT; ++k) report[k] = 0.0;
    for (int iter = 1; iter <= niter; ++iter) {
        done = iter;
        double xi;
        double L_tot = orc_ir_counts(c, iter, N, heat, &xi);
        if (iter == 1) L_prev = L_tot;
        if (!(L_tot > 0.0)) break;
        ir->Ltot = L_tot;
        ir->Lpacket = L_tot / (double)N;
        ir->xi = xi;
        if (orc_ir_transport(c, iter, N, nthreads)) { rc = -2; break; }
        ir_heating(c, ir->irabs, ir->heat_ir);
        for (int il = 0; il < nl; ++il) heat[il] = heat_euv[il] + ir->heat_ir[il];
        orc_dust_temp_compute(c, heat);
        double emitted = ir->emitted, extinct = ir->extinct, escaped = ir->escaped;
        double closure = fabs(extinct + escaped - emitted) / (emitted > DBL_MIN ? emitted : DBL_MIN);
        double L_abs = 0.0, L_new = 0.0;
        for (int il = 1; il <= nl; ++il) {
            double h = c->car ? 0.5 * c->dcell : c->ch[c->icell_of_leaf[il - 1] - 1];
            double w = 2.0 * h * c->p.distance2cm;
            double vol = (w * w) * w;
            L_abs = L_abs + ir->heat_ir[il - 1] * vol;
            L_new = L_new + heat[il - 1] * vol;
        }
        double drel = fabs(L_new - L_prev) / (L_new > DBL_MIN ? L_new : DBL_MIN);
        report[2] = L_tot; report[3] = emitted; report[4] = extinct; report[5] = escaped;
        report[6] = closure; report[7] = L_abs; report[8] = drel;
        if (closure > 5.0e-13) { rc = -3; break; }
        L_prev = L_new;
        if (drel < tol) { converged = 1; break; }
    }
    report[0] = (double)done;
    report[1] = (double)converged;
    report[9] = ir->nzero > 0 ? ir->frac_discarded : 0.0;
    free(heat_euv);
    return rc;
}

/* accessors of the band's state for the tests */
int orc_dust_ir_arrays(const orc_ctx *c, double *tdust, double *heat_ir, double *irabs,
                       double *lesc, double *lam, double *sabs, double *edge, double *femit)
{
    const orc_dustir *ir = c->ir;
    if (!ir) return -1;
    size_t nl = (size_t)c->nleaf;
    if (tdust && ir->tdust) memcpy(tdust, ir->tdust, sizeof(double) * nl);
    if (heat_ir && ir->heat_ir) memcpy(heat_ir, ir->heat_ir, sizeof(double) * nl);
    if (irabs && ir->irabs) memcpy(irabs, ir->irabs, sizeof(double) * nl);
    if (lesc && ir->lesc) memcpy(lesc, ir->lesc, sizeof(double) * ir->nl);
    if (lam && ir->lam) memcpy(lam, ir->lam, sizeof(double) * ir->nl);
    if (sabs && ir->sabs) memcpy(sabs, ir->sabs, sizeof(double) * ir->nl);
    if (edge && ir->edge) memcpy(edge, ir->edge, sizeof(double) * (ir->nl + 1));
    if (femit) {
        memcpy(femit, ir->tgrid, sizeof(double) * ORC_DT_NT);
        memcpy(femit + ORC_DT_NT, ir->femit, sizeof(double) * ORC_DT_NT);
    }
    return 0;
}

void orc_dustir_free(orc_dustir *ir)
{
    if (!ir) return;
    free(ir->lam); free(ir->sabs); free(ir->edge); free(ir->lesc);
    free(ir->tdust); free(ir->heat_ir); free(ir->irabs); free(ir->khi);
    free(ir);
}
