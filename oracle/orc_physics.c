/*
 * Oracle: photoionization cross sections, recombination and collisional
 * ionization rate coefficients (TEST INFRASTRUCTURE).
 *
 * Restates src/photo_xsec.f90 and src/recomb_mod.f90 (recomb_model
 * 'badnell_milne', ci_model 'voronov': the reference defaults).
 *
 * Pinned by: the threshold values sigma_HI(13.598 eV) = 6.30e-18 cm^2,
 * sigma_HeII(54.416 eV) = 1.575e-18 cm^2 (closed form), the VFKY96 He I value
 * at threshold (7.4 Mb quoted in tests/g0_gamma/check_g0_gamma.py:146-148);
 * the alpha_1 values at 1e4 K printed in src/recomb_mod.f90:181-185; and,
 * when oracle/_ref is built, the reference photo_xsec and recomb_mod compiled
 * from their own source (bit for bit).
 *
 * Integer powers follow the reference compiler's left-to-right expansion
 * (x**4 = ((x*x)*x)*x, checked bit for bit against the compiled reference),
 * real powers go through pow().
 */
#include <math.h>
#include "mc_oracle.h"

/* sigma_hydrogenic (src/photo_xsec.f90:181-196) */
double orc_sigma_hydrogenic(double E, double Eth, double Z)
{
    const double sig0 = 6.30e-18;
    double s0 = sig0 / (Z * Z);
    if (E < Eth) return 0.0;
    if (E > Eth) {
        double eps = sqrt(E / Eth - 1.0);
        double r = Eth / E;
        return s0 * (((r * r) * r) * r) * exp(4.0 - 4.0 * atan(eps) / eps) /
               (1.0 - exp(-2.0 * ORC_PI / eps));
    }
    return s0;
}

/* sigma_vfky96 (src/photo_xsec.f90:135-148) */
double orc_sigma_vfky96(double E, double Eth, double E0, double s0, double ya,
                        double P, double yw, double y0, double y1)
{
    if (E < Eth) return 0.0;
    double x = E / E0 - y0;
    double z = sqrt(x * x + y1 * y1);
    double Q = 5.5 - 0.5 * P;
    double xm = x - 1.0;
    double Fy = (xm * xm + yw * yw) * pow(z, -Q) * pow(1.0 + sqrt(z / ya), -P);
    return s0 * Fy * 1.0e-18;
}

/* src/photo_xsec.f90:199-215 */
double orc_sigma_HI(double E) { return orc_sigma_hydrogenic(E, ORC_ETH_HI, 1.0); }
double orc_sigma_HeII(double E) { return orc_sigma_hydrogenic(E, ORC_ETH_HEII, 2.0); }
double orc_sigma_HeI(double E)
{
    return orc_sigma_vfky96(E, ORC_ETH_HEI, 1.361e1, 9.492e2, 1.469, 3.188, 2.039,
                            4.434e-1, 2.136);
}

/* rr_badnell (src/recomb_mod.f90:124-130) */
static double rr_badnell(double T, double A, double B, double T0, double T1,
                         double Cc, double T2)
{
    double bp = B + Cc * exp(-T2 / T);
    double tt = sqrt(T / T0);
    return A / (tt * pow(1.0 + tt, 1.0 - bp) * pow(1.0 + sqrt(T / T1), 1.0 + bp));
}

/* dr_HeII_badnell (src/recomb_mod.f90:132-137) */
static double dr_HeII(double T)
{
    return pow(T, -1.5) * (1.417e-3 * exp(-4.633e5 / T) + 2.235e-4 * exp(-5.532e5 / T) -
                           2.185e-5 * exp(-8.887e5 / T));
}

/* rr_mao fit form (src/recomb_mod.f90:145-151) */
static double rr_mao(double T, double a0, double b0, double c0, double a1, double b1,
                     double a2, double b2)
{
    double te = T * (ORC_KBOLTZ / ORC_EV2ERG);
    return a0 * 1.0e-10 * pow(te, -b0 - c0 * log(te)) * (1.0 + a2 * pow(te, -b2)) /
           (1.0 + a1 * pow(te, -b1));
}

/* alpha1_* Milne fits (src/recomb_mod.f90:188-204) */
static double alpha1_HII(double T)
{
    return rr_mao(T, 1.87007e-2, 1.29077, -2.17465e-3, 1.29189e1, 8.05060e-1,
                  8.44399e-2, 5.93940e-4);
}
static double alpha1_HeIII(double T)
{
    return rr_mao(T, 2.61525e-1, 1.39527, -3.22079e-4, 6.40708e1, 8.97582e-1,
                  5.01551e-1, 1.21533e-3);
}
static double alpha1_HeII(double T)
{
    return rr_mao(T, 3.26862e-3, 6.97361e-1, 3.14907e-4, 5.76341e1, 1.18534,
                  2.58417e1, 9.90134e-1);
}

/* src/recomb_mod.f90:47-118, 'badnell_milne' branches */
double orc_alphaA_HII(double T)
{
    return rr_badnell(T, 8.318e-11, 0.7472, 2.965, 7.001e5, 0.0, 0.0);
}
double orc_alphaB_HII(double T) { return orc_alphaA_HII(T) - alpha1_HII(T); }
double orc_alphaA_HeII(double T)
{
    return rr_badnell(T, 5.235e-11, 0.6988, 7.301, 4.475e6, 0.0829, 1.682e5) + dr_HeII(T);
}
double orc_alphaB_HeII(double T) { return orc_alphaA_HeII(T) - alpha1_HeII(T); }
double orc_alphaA_HeIII(double T)
{
    return rr_badnell(T, 1.818e-10, 0.7492, 1.017e1, 2.786e6, 0.0, 0.0);
}
double orc_alphaB_HeIII(double T) { return orc_alphaA_HeIII(T) - alpha1_HeIII(T); }

/* voronov (src/recomb_mod.f90:256-261) */
static double voronov(double T, double dE, double P, double A, double X, double K)
{
    double U = dE * ORC_EV2ERG / (ORC_KBOLTZ * T);
    return A * (1.0 + P * sqrt(U)) * pow(U, K) * exp(-U) / (X + U);
}

/* ci_model of the model being solved: 1 for 'dere_hybrid'.  The model-level
 * entry points set it from their context for the duration of the call
 * (orc_ci_model_set returns the previous value to restore); 0 otherwise. */
static int g_ci_dere = 0;
int orc_ci_model_set(int ci_dere)
{
    const int prev = g_ci_dere;
    g_ci_dere = ci_dere;
    return prev;
}

/* ci_* (src/recomb_mod.f90:263-279): Voronov times ci_dere_ratio, which is 1
 * for ci_model 'voronov' */
double orc_ci_HI(double T)
{
    return voronov(T, 13.6, 0.0, 2.91e-8, 0.232, 0.39) * orc_ci_dere_ratio(g_ci_dere, 1, 1);
}
double orc_ci_HeI(double T)
{
    return voronov(T, 24.6, 0.0, 1.75e-8, 0.180, 0.35) * orc_ci_dere_ratio(g_ci_dere, 2, 1);
}
double orc_ci_HeII(double T)
{
    return voronov(T, 54.4, 1.0, 2.05e-9, 0.265, 0.25) * orc_ci_dere_ratio(g_ci_dere, 2, 2);
}

/* ---- cooling (src/cooling_mod.f90, src/gaunt.f90) ---------------------- */

/* Hummer (1988, ApJ 327, 477) thermally averaged free-free Gaunt factor
 * Chebyshev coefficients, as tabulated in src/gaunt.f90:44-73 (a published
 * table, carried as data; column-major d(8,11)). */
static const double hummer_d[88] = {
    8.986940175e+00, -4.009515855e+00, 8.808871266e-01, 2.640245111e-02, -4.580645915e-02,
    -3.568055702e-03, 2.827798067e-03, 3.365860195e-04, -8.006936989e-01, 9.466021705e-01,
    9.043402532e-02, -9.608451450e-02, -1.885629865e-02, 1.050313890e-02, 2.800889961e-03,
    -1.078209202e-03, -3.781305103e-01, 1.102726332e-01, -1.543619180e-02, 8.310561114e-03,
    2.179620525e-02, 4.259726289e-03, -4.181588794e-03, -1.770208330e-03, 1.877213132e-02,
    -1.004885705e-01, -5.483366378e-02, -4.520154409e-03, 8.366530426e-03, 3.700273930e-03,
    6.889320423e-04, 9.460313195e-05, 7.300158392e-02, 3.576785497e-03, -4.545307025e-03,
    -1.017965604e-02, -9.530211924e-03, -3.450186162e-03, 1.040482914e-03, 1.407073544e-03,
    -1.744671550e-03, 2.864013856e-02, 1.903394837e-02, 7.091074494e-03, -9.668371391e-04,
    -2.999107465e-03, -1.820642230e-03, -3.874082085e-04, -1.707268366e-02, -4.694254776e-03,
    1.311691517e-03, 5.316703136e-03, 5.178193095e-03, 2.451228935e-03, -2.277321615e-05,
    -8.182359057e-04, 2.567331664e-04, -9.155339970e-03, -6.997479192e-03, -3.571518641e-03,
    -2.096101038e-04, 1.553822487e-03, 1.509584686e-03, 6.212627837e-04, 4.098322531e-03,
    1.635218463e-03, -5.918883504e-04, -2.333091048e-03, -2.484138313e-03, -1.359996060e-03,
    -5.371426147e-05, 5.553549563e-04, 3.837562402e-05, 2.938325230e-03, 2.393747064e-03,
    1.328839809e-03, 9.135013312e-05, -7.137252303e-04, -7.656848158e-04, -3.504683798e-04,
    -8.491991820e-04, -3.615327726e-04, 3.148015257e-04, 8.909207650e-04, 9.869737522e-04,
    6.134671184e-04, 1.068883394e-04, -2.046080100e-04};

/* getGauntFF (src/gaunt.f90:19-156): single-precision evaluation with the
 * coefficient sums promoted to double where the reference mixes kinds */
void orc_gaunt_ff(float z, float log10Te, const float *xlf, float *g, int n)
{
    float b[12], c[9];
    float xlrkt;
    if (log10Te < 2.5f) xlrkt = 5.1983649f - 2.5f;
    else if (log10Te > 8.0f) xlrkt = 5.1983649f - 8.0f;
    else xlrkt = 5.1983649f - log10Te;
    float txg = 0.66666667f * (2.0f * z + xlrkt);
    float gamma2 = powf(10.0f, txg * 1.5f);
    float con = 0.72727273f * xlrkt + 0.90909091f;
    for (int j = 1; j <= 8; ++j) {
/* the table literals are default (single-precision) reals in the reference,
 * widened to double on assignment */
#define D(jj, ii) ((double)(float)hummer_d[((ii) - 1) * 8 + (jj) - 1])
        int ir = 9;
        b[11] = (float)D(j, 11);
        b[10] = (float)((double)(txg * b[11]) + D(j, 10));
        for (int i = 1; i <= 9; ++i) {
            b[ir] = (float)((double)(txg * b[ir + 1] - b[ir + 2]) + D(j, ir));
            ir = ir - 1;
        }
        c[j] = 0.25f * (b[1] - b[3]);
#undef D
    }
    for (int i = 0; i < n; ++i) {
        float txu = 0.72727273f * xlf[i] + con;
        float u = powf(10.0f, (txu - 0.90909091f) / 0.72727273f);
        if (fabsf(txu) <= 2.0f) {
            int ir = 6;
            b[8] = c[8];
            b[7] = txu * b[8] + c[7];
            for (int j = 1; j <= 6; ++j) {
                b[ir] = txu * b[ir + 1] - b[ir + 2] + c[ir];
                ir = ir - 1;
            }
            g[i] = b[1] - b[3];
        } else if (log10f(u) < -4.0f && gamma2 < 0.3f) {
            g[i] = 0.551329f * logf(2.24593f / u);
        } else if (log10f(u) < -4.0f && gamma2 >= 0.3f) {
            g[i] = 0.551329f * logf(0.944931f / u / sqrtf(gamma2));
        } else if (txu > 2.0f) {
            if (log10f(gamma2) < -3.0f) {
                g[i] = sqrtf(0.9549297f / u);
            } else if (log10f(gamma2) < 0.0f) {
                float slope = (sqrtf(12.0f * gamma2 / u) - sqrtf(0.9549297f / u)) / 3.0f;
                g[i] = sqrtf(12.0f * gamma2 / u) + slope * log10f(gamma2);
            } else {
                g[i] = sqrtf(12.0f * gamma2 / u);
                g[i] = fminf(1.05f, g[i]);
            }
        }
    }
}

/* gaunt_ff_mean (src/cooling_mod.f90:182-220) */
double orc_gaunt_ff_mean(double T, double Zc)
{
    enum { NU = 64 };
    double u[NU], du[NU];
    float xlf[NU], g[NU];
    double lo = log(1.0e-4), hi = log(20.0);
    for (int i = 1; i <= NU; ++i) {
        u[i - 1] = exp(lo + (hi - lo) * ((double)i - 0.5) / (double)NU);
        du[i - 1] = u[i - 1] * (hi - lo) / (double)NU;
    }
    double ryd_over_kt = 13.605693 * ORC_EV2ERG / (ORC_KBOLTZ * T);
    for (int i = 0; i < NU; ++i) xlf[i] = (float)log10(u[i] / ryd_over_kt);
    orc_gaunt_ff((float)log10(Zc), (float)log10(T), xlf, g, NU);
    double gm = 0.0;
    for (int i = 0; i < NU; ++i) gm = gm + (double)g[i] * exp(-u[i]) * du[i];
    return gm / (exp(-1.0e-4) - exp(-20.0));
}

/* cooling_setup's gbar_ff table: log T = 2..6 (41 points), Z = 1..4
 * (src/cooling_mod.f90:80-86) */
void orc_gff_table(double *logT41, double *gff41x4)
{
    for (int i = 1; i <= 41; ++i) {
        logT41[i - 1] = 2.0 + 4.0 * (double)(i - 1) / 40.0;
        for (int iz = 1; iz <= 4; ++iz)
            gff41x4[(iz - 1) * 41 + i - 1] = orc_gaunt_ff_mean(pow(10.0, logT41[i - 1]), (double)iz);
    }
}

/* gbar_ff (src/cooling_mod.f90:223-234) */
double orc_gbar_ff(const double *logT, const double *gff, double T, int Z)
{
    int iz = Z < 1 ? 1 : (Z > 4 ? 4 : Z);
    double lt = log10(T);
    if (lt < logT[0]) lt = logT[0];
    if (lt > logT[40]) lt = logT[40];
    int i = (int)((lt - logT[0]) / (logT[1] - logT[0])) + 1;
    if (i > 40) i = 40;
    double w = (lt - logT[i - 1]) / (logT[i] - logT[i - 1]);
    return gff[(iz - 1) * 41 + i - 1] * (1.0 - w) + gff[(iz - 1) * 41 + i] * w;
}

/* recombination cooling (src/cooling_mod.f90:139-175) */
static double betaA_HII(double T)
{
    double lam = 2.0 * 157807.0 / T;
    return 1.778e-29 * T * pow(lam, 1.965) / pow(1.0 + pow(lam / 0.541, 0.502), 2.697);
}
static double betaB_HII(double T)
{
    double lam = 2.0 * 157807.0 / T;
    return 3.435e-30 * T * pow(lam, 1.970) / pow(1.0 + pow(lam / 2.250, 0.376), 3.720);
}
static double beta_HeIII(double T, int caseA)
{
    double lam = 2.0 * 631515.0 / T;
    if (caseA)
        return 2.0 * 1.778e-29 * T * pow(lam, 1.965) / pow(1.0 + pow(lam / 0.541, 0.502), 2.697);
    return 2.0 * 3.435e-30 * T * pow(lam, 1.970) / pow(1.0 + pow(lam / 2.250, 0.376), 3.720);
}

/* cooling_total (src/cooling_mod.f90:240-276) with the H I line-cooling fit
 * Lambda = T^-1/2 sum A_i exp(-T_i/T) (cooling_eval, :125-134) unsuppressed
 * (cooling_model 'low_density') */
double orc_cooling_total(const orc_cooling *cl, double T, double nH, double ne, double xHI,
                         double xHeI, double xHeII, double yHe, int caseA)
{
    double xHeIII = 1.0 - xHeI - xHeII;
    if (xHeIII < 0.0) xHeIII = 0.0;
    double nHI = nH * xHI, nHII = nH * (1.0 - xHI);
    double nHeI = nH * yHe * xHeI, nHeII = nH * yHe * xHeII, nHeIII = nH * yHe * xHeIII;
    double cool;
    if (caseA)
        cool = ne * (nHII * betaA_HII(T) + nHeII * (ORC_KBOLTZ * T * orc_alphaA_HeII(T)) +
                     nHeIII * beta_HeIII(T, 1));
    else
        cool = ne * (nHII * betaB_HII(T) + nHeII * (ORC_KBOLTZ * T * orc_alphaB_HeII(T)) +
                     nHeIII * beta_HeIII(T, 0));
    cool = cool + 1.42554e-27 * sqrt(T) * ne *
                      ((nHII + nHeII) * orc_gbar_ff(cl->logT, cl->gff, T, 1) +
                       4.0 * nHeIII * orc_gbar_ff(cl->logT, cl->gff, T, 2));
    cool = cool + ne * (nHI * orc_ci_HI(T) * ORC_ETH_HI + nHeI * orc_ci_HeI(T) * ORC_ETH_HEI +
                        nHeII * orc_ci_HeII(T) * ORC_ETH_HEII) * ORC_EV2ERG;
    double lam = 0.0;
    for (int i = 0; i < cl->nfit; ++i) lam = lam + cl->A[i] * exp(-cl->Ti[i] / T);
    lam = lam / sqrt(T);
    /* cooling_model 'local_ne': suppressed by the H I table
     * (src/cooling_mod.f90:271-275) */
    cool = cool + ne * nHI * orc_nlcool_apply(cl->hi_tab, lam, T, ne);
    return cool;
}
