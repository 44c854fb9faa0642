// Line-cooling suppression of cooling_model 'local_ne' and the n-level
// atoms it is built from (src/nlevel_mod.f90, src/nlevel_cooling_mod.f90).
//
// Device layout of the tables: one slot per tabulated ion, kNlSlot doubles:
// lamref[kNlT] (the n_e-floor column, the low-density cooling coefficient)
// then Ssup[kNlNe][kNlT] (Lambda_SE(T, n_e) / lamref, T fastest), on the
// grid log10 T = 2 + 0.01 i (401 rows), log10 n_e = -6 + 0.1 j (161 columns).
// The lookup restates nlevel_cooling_apply / cooling_coeff / col_cooling /
// low_density_cooling (src/nlevel_cooling_mod.f90:157-262) operation for
// operation; the (T, n_e) cell and weights, identical for every ion, are
// found once per evaluation (NlPos) and shared by all ions of a leaf.
#pragma once
#include <cfloat>
#include <cmath>

#ifndef __HIPCC__
#define MC_HD
#else
#define MC_HD __host__ __device__
#endif

constexpr int kNlT = 401, kNlNe = 161;
constexpr double kNlT0 = 2.0, kNlDT = 0.01, kNlNe0 = -6.0, kNlDNe = 0.1;
constexpr long kNlSlot = (long)kNlT + (long)kNlT * kNlNe;

struct NlPos {
    int it;          // 0-based T row (it, it + 1)
    int je;          // 0-based n_e column (je, je + 1) when col == 0
    int col;         // 0: between columns je, je + 1; 1: floor column; 2: last column
    double w1, wt;   // weights of the row pair: in 1/T (ln form) and log T (linear form)
    double wn;       // log n_e weight between the columns
};

MC_HD inline double nl_gT(int i) { return kNlT0 + kNlDT * (double)i; }
MC_HD inline double nl_gNe(int j) { return kNlNe0 + kNlDNe * (double)j; }

MC_HD inline NlPos nl_pos(double T, double ne)
{
#pragma clang fp contract(off)
    NlPos p;
    const double lt = fmin(fmax(log10(T), nl_gT(0)), nl_gT(kNlT - 1));
    int i = (int)((lt - nl_gT(0)) / kNlDT) + 1;
    if (i > kNlT - 1) i = kNlT - 1;
    p.it = i - 1;
    p.w1 = (1.0 / pow(10.0, lt) - 1.0 / pow(10.0, nl_gT(p.it))) /
           (1.0 / pow(10.0, nl_gT(p.it + 1)) - 1.0 / pow(10.0, nl_gT(p.it)));
    p.wt = fmin(fmax((lt - nl_gT(p.it)) / (nl_gT(p.it + 1) - nl_gT(p.it)), 0.0), 1.0);
    const double lne = log10(ne);
    p.je = 0;
    p.wn = 0.0;
    if (lne <= nl_gNe(0)) {
        p.col = 1;
    } else if (lne >= nl_gNe(kNlNe - 1)) {
        p.col = 2;
    } else {
        int j = (int)((lne - nl_gNe(0)) / kNlDNe) + 1;
        if (j > kNlNe - 1) j = kNlNe - 1;
        p.col = 0;
        p.je = j - 1;
        p.wn = (lne - nl_gNe(p.je)) / (nl_gNe(p.je + 1) - nl_gNe(p.je));
    }
    return p;
}

// col_cooling: Lambda on column j between the rows it, it + 1
MC_HD inline double nl_col(const double *tab, const NlPos &p, int j)
{
#pragma clang fp contract(off)
    const double *lr = tab, *S = tab + kNlT + (long)j * kNlT;
    const double Li = lr[p.it] * S[p.it];
    const double Lip = lr[p.it + 1] * S[p.it + 1];
    if (Li > 0.0 && Lip > 0.0) return exp(log(Li) + p.w1 * (log(Lip) - log(Li)));
    return Li * (1.0 - p.wt) + Lip * p.wt;
}

// cooling_coeff: Lambda_SE(T, n_e)
MC_HD inline double nl_coeff(const double *tab, const NlPos &p)
{
#pragma clang fp contract(off)
    if (p.col == 1) return nl_col(tab, p, 0);
    if (p.col == 2) return nl_col(tab, p, kNlNe - 1);
    const double L0 = nl_col(tab, p, p.je), L1 = nl_col(tab, p, p.je + 1);
    if (L0 > 0.0 && L1 > 0.0) return exp(log(L0) + p.wn * (log(L1) - log(L0)));
    return L0 * (1.0 - p.wn) + L1 * p.wn;
}

// nlevel_cooling_apply: the fit lam_fit suppressed by the ion's table (tab
// null: no table, lam_fit unchanged)
MC_HD inline double nl_apply(const double *tab, const NlPos &p, double lam_fit)
{
#pragma clang fp contract(off)
    if (!tab) return lam_fit;
    const double lamref = nl_col(tab, p, 0);
    const double S = nl_coeff(tab, p) / fmax(lamref, DBL_MIN);
    const double fcov = fmin(1.0, lamref / fmax(lam_fit, DBL_MIN));
    return lam_fit * (fcov * S + (1.0 - fcov));
}

// ---- line luminosities (mc_lines_compute) ----------------------------------
constexpr int kLnLev = 16, kLnRad = 48, kLnUps = 96, kLnMaxLines = 512;

// one metal ion's atom, oriented on the host: radiative rows as (lower,
// upper) 0-based with dE [erg]; collision rows as (lower, upper), g_u,
// g_u / g_l, dE [erg] and the Upsilon fit
struct DevLineIon {
    int32_t nlev, nrad, nups, out0;    // out0: index of the ion's first line
    int32_t t, up;                     // stage fraction: up ? up[t] : lo[t] of m_fracs
    double abund;
    int32_t rlo[kLnRad], rhi[kLnRad];
    double A[kLnRad], rde[kLnRad];
    int32_t clo[kLnUps], chi[kLnUps], ttype[kLnUps], logf[kLnUps], nc[kLnUps];
    double gu[kLnUps], gr[kLnUps], cde[kLnUps];
    double cups[kLnUps], dery[kLnUps], st_lo[kLnUps], st_hi[kLnUps];
    double coef[kLnUps][16];
};

// one recombination-line table, etab[line][iT][iD]
struct DevLineTab {
    int32_t nT, nD, nlin, out0;
    double tgrid[MC_LT_MAXG], dgrid[MC_LT_MAXG];
    double etab[MC_LT_MAXL * MC_LT_MAXG * MC_LT_MAXG];
};

