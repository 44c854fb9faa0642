// Trace-metal cascade on the device (included by mc_kernels.hip inside its
// anonymous namespace, after the H/He rate coefficients).
//
// One lane per leaf.  The (leaf, T) rate cache of species_ne_prepare
// (src/species_mod.f90:114-128, 517-536) lives in registers: every loop over
// transitions is unrolled to MC_GPU_MAXTR with a wave-uniform bound (the
// registry size), so the cache arrays are indexed by compile-time constants
// and the per-transition registry fields are scalar loads.
#pragma once
#include "mc_nlevel.h"

// photoionisation cross sections m_vfky96 / m_vy95 / m_sigma: mc_internal.h

// voronov_ci x ci_dere_ratio (src/species_mod.f90:815-821, 524-525)
__device__ __forceinline__ double m_ci(const DevMetalTab *mt, int t, double T)
{
#pragma clang fp contract(off)
    const double *c = mt->ci[t];
    const double U = c[0] * MC_EV2ERG / (MC_KBOLTZ * T);
    return c[2] * (1.0 + c[1] * sqrt(U)) * pow(U, c[4]) * exp(-U) / (c[3] + U) * mt->ci_ratio[t];
}

// alpha_rec: Badnell RR (or the RR2 power law) + Badnell DR terms + the
// Shull & Van Steenberg DR2 form (src/species_mod.f90:428-447)
__device__ __forceinline__ double m_alpha_rec(const DevMetalTab *mt, int t, double T, double tm15)
{
#pragma clang fp contract(off)
    double a;
    if (mt->has_rr2[t]) {
        a = mt->rr2[t][0] * pow(T / 1.0e4, -mt->rr2[t][1]);
    } else {
        const double *r = mt->rr[t];
        const double s0 = sqrt(T / r[2]), s1 = sqrt(T / r[3]);
        const double b = r[1] + r[4] * exp(-r[5] / T);
        a = r[0] / (s0 * pow(1.0 + s0, 1.0 - b) * pow(1.0 + s1, 1.0 + b));
    }
    const int nd = mt->ndr[t];
    for (int k = 0; k < nd; ++k) a = a + tm15 * mt->dr_c[t][k] * exp(-mt->dr_e[t][k] / T);
    if (mt->has_dr2[t]) {
        const double *d = mt->dr2[t];
        a = a + d[0] * tm15 * exp(-d[2] / T) * (1.0 + d[1] * exp(-d[3] / T));
    }
    return a;
}

// cx_rate forms 1-7 (src/species_mod.f90:452-486)
__device__ __forceinline__ double m_cx(int form, const double *c, double T)
{
#pragma clang fp contract(off)
    double t4 = T / 1.0e4, lnT, k;
    switch (form) {
    case 1:
        k = c[0] * pow(T / 300.0, c[1]);
        if (c[2] != 0.0) k = k * exp(-c[2] / t4);
        return k;
    case 2: return c[0] * pow(t4, c[1]) + c[2] * pow(t4, c[3]);
    case 3: return (c[0] * pow(t4, c[1]) + c[2] * pow(t4, c[3])) * exp(-c[4] / T);
    case 4:
        lnT = log(T);
        return exp(c[0] + lnT * (c[1] + lnT * (c[2] + lnT * (c[3] + lnT * c[4]))));
    case 5:
        lnT = log(T);
        return exp(c[0] + lnT * (c[1] + lnT * (c[2] + lnT * (c[3] + lnT * c[4]))) - c[5] / t4);
    case 6:
        t4 = fmin(fmax(t4, 0.5), 5.0);
        return fmax(c[0] * pow(t4, c[1]) * (1.0 + c[2] * exp(c[3] * t4)), 0.0);
    case 7:
        t4 = fmin(fmax(t4, 0.5), 5.0);
        return fmax(c[0] * pow(t4, c[1]) * (1.0 + c[2] * exp(c[3] * t4)) * exp(-c[4] / t4), 0.0);
    default: return 0.0;
    }
}

// (leaf, T) rate cache
template <int MT> struct MCache {
    double gam[MT], ci[MT], rec[MT], cxi[MT], cxr[MT];
};

// the four T-dependent coefficients of transition t (out of line: the
// unrolled cache fill below stays small enough to unroll completely)
struct MCoef { double ci, rec, cxi, cxr; };

__device__ __noinline__ MCoef m_coef(int t, double T, double tm15)
{
#pragma clang fp contract(off)
    const DevMetalTab *mt = &c_mt;
    MCoef q;
    q.ci = m_ci(mt, t, T);
    q.rec = m_alpha_rec(mt, t, T, tm15);
    q.cxi = mt->cxi_form[t] > 0 ? m_cx(mt->cxi_form[t], mt->cxi[t], T) : 0.0;
    q.cxr = mt->cxr_form[t] > 0 ? m_cx(mt->cxr_form[t], mt->cxr[t], T) : 0.0;
    return q;
}

// species_ne_prepare: coefficients of every transition at T, the leaf's
// metal Gammas (element order) from gam
template <int MT>
__device__ __forceinline__ void m_prepare(const DevMetalTab *mt, const double *gam, double T,
                                          MCache<MT> &cc)
{
#pragma clang fp contract(off)
    const int ntr = mt->ntr;
    const double tm15 = pow(T, -1.5);
#pragma unroll
    for (int t = 0; t < MT; ++t) {
        if (t < ntr) {
            const MCoef q = m_coef(t, T, tm15);
            cc.gam[t] = gam[t];
            cc.ci[t] = q.ci;
            cc.rec[t] = q.rec;
            cc.cxi[t] = q.cxi;
            cc.cxr[t] = q.cxr;
        } else {
            cc.gam[t] = 0.0; cc.ci[t] = 0.0; cc.rec[t] = 0.0; cc.cxi[t] = 0.0; cc.cxr[t] = 0.0;
        }
    }
}

// ratio chain of one transition: n_{i+1}/n_i = R_ion / R_rec
// (species_fractions / species_metal_sums, src/species_mod.f90:598-605)
template <int MT>
__device__ __forceinline__ double m_ratio(const MCache<MT> &cc, int t, double ne, double nHI,
                                          double nHII)
{
#pragma clang fp contract(off)
    double rion = cc.gam[t] + ne * cc.ci[t];
    if (cc.cxi[t] > 0.0) rion = rion + nHII * cc.cxi[t];
    double rrec = ne * cc.rec[t];
    if (cc.cxr[t] > 0.0) rrec = rrec + nHI * cc.cxr[t];
    return rion / fmax(rrec, DBL_MIN);
}

// Stage fractions from the ratio chains (species_fractions,
// src/species_mod.f90:782-812): for every transition u, lo[u] is the fraction
// of its lower stage and up[u] of its upper stage: lo = 1/s for an element's
// first transition and P(u-1)/s otherwise, up = P(u)/s, with P the running
// product of the element's ratios and s = 1 + sum P.  Every index is a
// compile-time constant after unrolling, so lo/up stay in registers.
template <int MT>
__device__ __forceinline__ void m_fracs(const DevMetalTab *mt, const MCache<MT> &cc, double ne,
                                        double nHI, double nHII, double (&lo)[MT],
                                        double (&up)[MT])
{
#pragma clang fp contract(off)
    const int ntr = mt->ntr;
    double s = 1.0, prod = 1.0;
#pragma unroll
    for (int t = 0; t < MT; ++t) {
        lo[t] = 0.0;
        up[t] = 0.0;
        if (t < ntr) {
            const int first = mt->t_first[t];
            if (first == t) { s = 1.0; prod = 1.0; }
            prod = prod * m_ratio(cc, t, ne, nHI, nHII);
            s = s + prod;
            up[t] = prod;
            if (mt->t_last[t]) {
                const double inv = 1.0 / s;
#pragma unroll
                for (int j = MC_MAX_ST - 2; j >= 0; --j) {
                    const int u = t - j;      // compile-time after unrolling
                    if (u >= 0 && u >= first) {
                        lo[u >= 0 ? u : 0] = u > first ? up[u > 0 ? u - 1 : 0] : 1.0;
                    }
                }
#pragma unroll
                for (int j = 0; j < MC_MAX_ST - 1; ++j) {
                    const int u = t - j;
                    if (u >= 0 && u >= first) {
                        lo[u >= 0 ? u : 0] = lo[u >= 0 ? u : 0] * inv;
                        up[u >= 0 ? u : 0] = up[u >= 0 ? u : 0] * inv;
                    }
                }
            }
        }
    }
}

// species_metal_sums (src/species_mod.f90:566-619): metal electrons and the
// two charge-exchange rates on hydrogen, summed element by element and stage
// by stage as the reference does
template <int MT>
__device__ __forceinline__ void m_sums(const DevMetalTab *mt, const MCache<MT> &cc, double nH,
                                       double ne, double nHI, double nHII, double &nem,
                                       double &rin, double &rout)
{
#pragma clang fp contract(off)
    double lo[MT], up[MT];
    m_fracs<MT>(mt, cc, ne, nHI, nHII, lo, up);
    const int ntr = mt->ntr;
    double a = 0.0, b = 0.0, c = 0.0;
#pragma unroll
    for (int u = 0; u < MT; ++u) {
        if (u < ntr) {
            const double nX = mt->t_ab[u] * nH;
            b = b + cc.cxi[u] * nX * lo[u];
            a = a + mt->t_ab[u] * nH * up[u] * (double)(mt->t_it[u] + 1);
            c = c + cc.cxr[u] * nX * up[u];
        }
    }
    nem = a; rin = b; rout = c;
}

// Registry layouts known at compile time: transitions per element.
template <int LAY> struct MLay;
template <> struct MLay<MC_LAYOUT_CNONeSAr> {
    static constexpr int nel = 6;
    __device__ static constexpr int nt(int e) { return e == 0 || e >= 4 ? 3 : 2; }
};
template <> struct MLay<MC_LAYOUT_CNONeS> {
    static constexpr int nel = 5;
    __device__ static constexpr int nt(int e) { return e == 0 || e == 4 ? 3 : 2; }
};
template <class L> __device__ constexpr int mlay_off(int e)
{
    int s = 0;
    for (int k = 0; k < e; ++k) s += L::nt(k);
    return s;
}

// species_metal_sums for a compiled-in layout, in the reference's own
// arithmetic (ratios, s, fr = 1/s then fr = fr r per stage, the sums in its
// order); only the element abundances are read from the registry
template <int LAY, int MT>
__device__ __forceinline__ void m_sums_fixed(const MCache<MT> &cc, double nH, double ne,
                                             double nHI, double nHII, double &nem, double &rin,
                                             double &rout)
{
#pragma clang fp contract(off)
    using L = MLay<LAY>;
    double a = 0.0, b = 0.0, c = 0.0;
#pragma unroll
    for (int e = 0; e < L::nel; ++e) {
        const int t0 = mlay_off<L>(e), n = L::nt(e);
        double r[MC_MAX_ST - 1];
        double s = 1.0, prod = 1.0;
#pragma unroll
        for (int i = 0; i < MC_MAX_ST - 1; ++i) {
            if (i < n) {
                r[i] = m_ratio(cc, t0 + i, ne, nHI, nHII);
                prod = prod * r[i];
                s = s + prod;
            }
        }
        const double A = c_mt.el_ab[e];
        const double nX = A * nH;
        double fr = 1.0 / s;
        b = b + cc.cxi[t0] * nX * fr;
#pragma unroll
        for (int i = 0; i < MC_MAX_ST - 1; ++i) {
            if (i < n) {
                fr = fr * r[i];
                a = a + A * nH * fr * (double)(i + 1);
                c = c + cc.cxr[t0 + i] * nX * fr;
                if (i + 1 < n) b = b + cc.cxi[t0 + i + 1] * nX * fr;
            }
        }
    }
    nem = a; rin = b; rout = c;
}

template <int MT, int LAY>
__device__ __forceinline__ void m_sums_l(const MCache<MT> &cc, double nH, double ne, double nHI,
                                         double nHII, double &nem, double &rin, double &rout)
{
#pragma clang fp contract(off)
    if constexpr (LAY == MC_LAYOUT_GENERIC) m_sums<MT>(&c_mt, cc, nH, ne, nHI, nHII, nem, rin, rout);
    else m_sums_fixed<LAY, MT>(cc, nH, ne, nHI, nHII, nem, rin, rout);
}

// hydrogen_neutral_fraction (src/ion_balance_mod.f90:177-208)
template <int MT, int LAY>
__device__ double m_hydrogen_neutral_fraction(const MCache<MT> &cc, double nH, double ne,
                                              double A, double B, double xw = -1.0)
{
#pragma clang fp contract(off)
    double nem, rin, rout;
    double xa = 0.0, Ra, xb = 1.0, Rb, scale = 0.0;
    bool warm = false;
    if (xw > 0.0 && xw < 1.0) {
        // warm-started form: a bracket of +-10% of the distance to the
        // nearer end around the previous pass's root, when the residual
        // changes sign across it (else the reference's [0, 1])
        const double d = 0.1 * fmin(xw, 1.0 - xw);
        const double wa = xw - d, wb = xw + d;
        m_sums_l<MT, LAY>(cc, nH, ne, nH * wa, nH * (1.0 - wa), nem, rin, rout);
        const double ra = wa * (B + rin + rout) - (A + rin), rin_a = rin;
        m_sums_l<MT, LAY>(cc, nH, ne, nH * wb, nH * (1.0 - wb), nem, rin, rout);
        const double rb = wb * (B + rin + rout) - (A + rin);
        if (ra < 0.0 && rb > 0.0) {
            xa = wa; Ra = ra; xb = wb; Rb = rb;
            // the tolerance keeps the size of the end-point residuals
            // |R(0)| + |R(1)| of the full bracket, from the sums at wa, wb
            scale = 2.0 * (A + B) + rin_a + rout;
            warm = true;
        }
    }
    if (!warm) {
        m_sums_l<MT, LAY>(cc, nH, ne, 0.0, nH, nem, rin, rout);
        Ra = -(A + rin);
        m_sums_l<MT, LAY>(cc, nH, ne, nH, 0.0, nem, rin, rout);
        Rb = B + rout;
        scale = A + B + fabs(Ra) + fabs(Rb);
    }
    double x = xb;
    for (int k = 0; k < 80; ++k) {
        if (Rb == Ra) break;
        x = xb - Rb * (xb - xa) / (Rb - Ra);
        x = fmin(fmax(x, fmin(xa, xb)), fmax(xa, xb));
        m_sums_l<MT, LAY>(cc, nH, ne, nH * x, nH * (1.0 - x), nem, rin, rout);
        const double Rx = x * (B + rin + rout) - (A + rin);
        if (fabs(Rx) <= 1.0e-15 * scale) break;
        if (fabs(xb - xa) <= 1.0e-15) break;
        if (Rx * Rb < 0.0) { xa = xb; Ra = Rb; }
        else Ra = 0.5 * Ra;
        xb = x; Rb = Rx;
    }
    return x;
}

// metal switches of the cell solve and the thermal balance
struct MetalSw {
    const double *gam, *heat;    // nleaf x ntr, element order
    int32_t ntr, metal_ne, cx_two_way, metal_heat, metal_ff;
    // ci_model 'dere_hybrid': the H / He collisional ionisation rates carry
    // the Dere ratio (ci_HI / ci_HeI / ci_HeII, src/recomb_mod.f90:263-279;
    // the metal ones hold it in the registry's ci_ratio)
    int32_t ci_dere;
    // warm-started solve only (0 = the reference's rule): stop the n_e
    // iteration after this many consecutive passes whose hydrogen is below
    // resolution (x_HII < 1e-10); see solve_cell_m
    int32_t noise_cap;
    // cooling_model 'local_ne' tables of the thermal update (null: the
    // stage fits unsuppressed)
    const double *nlt;
};

// solve_ion_cell with the metal-electron and two-way charge-exchange terms
// (src/ion_balance_mod.f90:53-150), the cache prepared at T
template <int MT, int LAY>
__device__ double solve_cell_m(const MetalSw &ms, const MCache<MT> &cc, const RateSet &r,
                               double gH, double gHe1, double gHe2, double nH, double yHe,
                               double &xHI, double &xHeI, double &xHeII)
{
#pragma clang fp contract(off)
    const bool use = ms.metal_ne || ms.cx_two_way;
    double xHeIII = fmax(0.0, 1.0 - xHeI - xHeII);
    double ne = fmax(nH * ((1.0 - xHI) + yHe * (xHeII + 2.0 * xHeIII)), 1.0e-12 * nH);
    double nem = 0.0, rin = 0.0, rout = 0.0;
    int unresolved = 0;
    for (int it = 0; it < 200; ++it) {
        const double ne_old = ne;
        if (use) {
            m_sums_l<MT, LAY>(cc, nH, ne_old, nH * xHI, nH * (1.0 - xHI), nem, rin, rout);
            if (!ms.cx_two_way) { rin = 0.0; rout = 0.0; }
        }
        if (ms.cx_two_way) {
            xHI = m_hydrogen_neutral_fraction<MT, LAY>(cc, nH, ne, r.aH * ne,
                                                       gH + (r.cH + r.aH) * ne,
                                                       ms.noise_cap > 0 ? xHI : -1.0);
        } else {
            const double den = gH + (r.cH + r.aH) * ne + rin + rout;
            xHI = den > 0.0 ? (r.aH * ne + rin) / den : 1.0;
        }
        const double r1 = fmin((gHe1 + r.cHe1 * ne) / (r.aHe2 * ne), 1.0e150);
        const double r2 = fmin((gHe2 + r.cHe2 * ne) / (r.aHe3 * ne), 1.0e150);
        xHeI = 1.0 / (1.0 + r1 + r1 * r2);
        xHeII = xHeI * r1;
        xHeIII = xHeII * r2;
        ne = nH * ((1.0 - xHI) + yHe * (xHeII + 2.0 * xHeIII));
        if (ms.metal_ne) ne = ne + nem;
        ne = fmax(0.5 * (ne + ne_old), 1.0e-12 * nH);
        if (fabs(ne - ne_old) <= 1.0e-10 * ne) break;
        // In a leaf without ionising flux the charge-exchange root leaves
        // 1 - x_HI ~ 1e-13 undetermined (any x in a band ~1e-12 wide zeroes
        // the balance in floating point), so n_e wanders inside a noise band
        // and the reference's test above is never met: it runs all 200
        // passes.  The warm-started form stops once the band is reached.
        if (ms.noise_cap > 0) {
            unresolved = 1.0 - xHI < 1.0e-10 ? unresolved + 1 : 0;
            if (unresolved >= ms.noise_cap) break;
        }
    }
    return ne;
}

// charge_neutrality_ne (src/ion_balance_mod.f90:235-254), the cache prepared
// at the written T
template <int MT, int LAY>
__device__ double charge_neutrality_m(const MCache<MT> &cc, double nH, double ne_HHe, double xHI,
                                      double ne_seed)
{
#pragma clang fp contract(off)
    const double nHI = nH * xHI, nHII = nH * (1.0 - xHI);
    double ne = fmax(ne_seed, 1.0e-12 * nH), nem, rin, rout;
    for (int it = 0; it < 200; ++it) {
        const double ne_old = ne;
        m_sums_l<MT, LAY>(cc, nH, ne_old, nHI, nHII, nem, rin, rout);
        ne = fmax(0.5 * ((ne_HHe + nem) + ne_old), 1.0e-12 * nH);
        if (fabs(ne - ne_old) <= 1.0e-10 * ne) break;
    }
    return ne;
}

// Metal terms of net_rate at the solved state (src/thermal_mod.f90:53-71),
// from the stage fractions lo/up of m_fracs:
// metal_cooling (cooling_model 'low_density': the stage fits
// T^-1/2 sum A_i exp(-T_i/T) unsuppressed), metal_cooling_H ([C II] 158 um
// and [O I] 63 um by H I impact), metal_freefree and metal_heating
// (src/species_mod.f90:704-868).  gff[z-1] = gbar_ff(T, z) for z = 1..4.
// stage line-cooling fit T^-1/2 sum A_i exp(-T_i/T) (cooling_eval,
// src/cooling_mod.f90:125-134)
__device__ __noinline__ double m_fit(int st, double T)
{
#pragma clang fp contract(off)
    const DevMetalTab *mt = &c_mt;
    double lam = 0.0;
    const int nf = mt->s_ncool[st];
    for (int k = 0; k < nf; ++k) lam = lam + mt->s_A[st][k] * exp(-mt->s_T[st][k] / T);
    return lam / sqrt(T);
}

// H-impact two-level cooling of one ion, [C II] 158 um (cII) or [O I] 63 um
// (src/species_mod.f90:713-730)
__device__ __noinline__ double m_cool_h(int cII, double nion, double nHI, double T)
{
#pragma clang fp contract(off)
    if (cII) {
        const double qul = 7.6e-10 * pow(T / 100.0, 0.14);
        const double qlu = 2.0 * qul * exp(-91.25 / T);
        return nion * nHI * qlu * (91.25 * 1.380649e-16) / (1.0 + nHI * (qul + qlu) / 2.321e-6);
    }
    const double qul = 9.2e-11 * pow(T / 100.0, 0.67);
    const double qlu = 0.6 * qul * exp(-227.7 / T);
    return nion * nHI * qlu * (227.7 * 1.380649e-16) / (1.0 + nHI * (qul + qlu) / 8.865e-5);
}

// a stage's line-cooling coefficient: its fit, suppressed by the stage's
// (T, n_e) table under cooling_model 'local_ne' (src/species_mod.f90:836-840)
__device__ __forceinline__ double m_cool_coef(const MetalSw &ms, const NlPos &np, int st, double T)
{
#pragma clang fp contract(off)
    const double f = m_fit(st, T);
    const int sl = c_mt.s_nl[st];
    return (ms.nlt && sl >= 0) ? nl_apply(ms.nlt + (long)sl * kNlSlot, np, f) : f;
}

template <int MT>
__device__ __forceinline__ void m_thermal_terms(const MetalSw &ms, const double (&lo)[MT],
                                                const double (&up)[MT], const double *heat,
                                                double T, double nH, double ne, double nHI,
                                                const double gff[4], double &mcool,
                                                double &mcoolH, double &mff, double &mheat)
{
#pragma clang fp contract(off)
    const DevMetalTab *mt = &c_mt;
    const double sqT = sqrt(T);
    const int ntr = mt->ntr;
    double c1 = 0.0, c2 = 0.0, sff = 0.0, h = 0.0;
    NlPos np{};
    if (ms.nlt) np = nl_pos(T, ne);
#pragma unroll
    for (int u = 0; u < MT; ++u) {
        if (u < ntr) {
            const double nel = mt->t_ab[u] * nH;
            const int sl = mt->t_st[u], su = sl + 1;
            if (mt->t_first[u] == u) {        // the element's neutral stage
                if (mt->s_ncool[sl] > 0) c1 = c1 + ne * nel * lo[u] * m_cool_coef(ms, np, sl, T);
                if (mt->t_oI[u]) c2 = c2 + m_cool_h(0, nel * lo[u], nHI, T);
            }
            if (mt->s_ncool[su] > 0) c1 = c1 + ne * nel * up[u] * m_cool_coef(ms, np, su, T);
            if (mt->t_cII[u]) c2 = c2 + m_cool_h(1, nel * up[u], nHI, T);
            const int z = mt->s_charge[su];
            const double gz = z <= 1 ? gff[0] : (z == 2 ? gff[1] : (z == 3 ? gff[2] : gff[3]));
            sff = sff + mt->t_ab[u] * up[u] * (double)(z * z) * gz;
            h = h + nel * lo[u] * heat[u];
        }
    }
    mcool = c1;
    mcoolH = c2;
    mff = ms.metal_ff ? 1.42554e-27 * sqT * ne * nH * sff : 0.0;
    mheat = ms.metal_heat ? h : 0.0;
}
