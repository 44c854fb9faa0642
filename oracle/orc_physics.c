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

/* ci_* (src/recomb_mod.f90:263-279); the Dere ratio is 1 for ci_model
 * 'voronov' */
double orc_ci_HI(double T) { return voronov(T, 13.6, 0.0, 2.91e-8, 0.232, 0.39); }
double orc_ci_HeI(double T) { return voronov(T, 24.6, 0.0, 1.75e-8, 0.180, 0.35); }
double orc_ci_HeII(double T) { return voronov(T, 54.4, 1.0, 2.05e-9, 0.265, 0.25); }
