// Energy-moment path tallies of the mixed-form metal walk (host set-up).
//
// Every rate tally the walk keeps for a scored segment is the segment's
// weighted path length pl times a coefficient that depends on the packet's
// energy alone: sigma_X(E)/(E ev2erg) and sigma_X(E)(1 - E_X/E) for H I,
// He I, He II (ion_score_mod hhe_rate_path, src/ion_score_mod.f90:137-231)
// and for every metal transition below E (species_rate_path_score,
// src/species_mod.f90:329-343), plus 1 for the band J tally
// (src/jtally_mod.f90).  Inside one band bin no threshold is crossed (the
// band edges are the thresholds), so every coefficient is a smooth function
// of ln E there.  Each bin is cut into sub-bins, equal in ln E, on which
// every coefficient equals its degree-3 Chebyshev interpolant in the
// sub-bin coordinate t in [-1, 1] to a relative 1e-7 (checked on 48 interior
// points per sub-bin); a segment then adds pl T_m(t), m = 0..3, to its
// leaf's four moments of its sub-bin, and after the walk
//   tally_q(leaf) = sum_s sum_m a[s][q][m] M[leaf][s][m]
// gives every coefficient sum.  Four atomics per segment instead of one per
// nonzero coefficient (~19 at 15 metal transitions).  Degree 3 needs ~2x the
// sub-bins of degree 5 for the same 1e-7 (218 against ~100 on the C N O Ne S
// Ar band), which costs only moment memory; degree 2 does not reach 1e-7
// within 512 sub-bins per band bin.  The FP64 form keeps the reference's
// per-coefficient sums.
#include <cfloat>
#include <cstdio>
#include <algorithm>
#include <cmath>
#include <vector>

#include "mc_internal.h"

namespace {

constexpr int kMomDeg = MC_MOM_NM - 1;   // Chebyshev degree
constexpr int kMomMaxSub = 128;          // sub-bins per band bin at most
constexpr int kMomCheck = 48;            // interior check points per sub-bin
constexpr double kMomTol = 1.0e-7;       // relative to the sub-bin's max |c|

// the walk's coefficient q of a packet of energy E, in its own arithmetic:
// q = 0..5 H/He (coef rows), q = 6 + 2k + {0,1} metal transition k in
// threshold order (metal_coeff)
// grain absorption cross section per H at E, in the launch's arithmetic
// (ion_dust_*_at, src/gas_opacity_mod.f90:217-290: linear in lambda for the
// albedo, log-log for C_ext, between the sorted table rows)
double dust_sabs(const mc_ctx *c, double E)
{
    const int n = c->dn;
    const double *lam = c->dlam.data(), *alb = c->dalb.data(), *ce = c->dcext.data();
    const double x0 = 1.23984 / E;
    double al, cx;
    if (x0 <= lam[0]) {
        al = alb[0]; cx = ce[0];
    } else if (x0 >= lam[n - 1]) {
        al = alb[n - 1]; cx = ce[n - 1];
    } else {
        int lo = 0, hi = n - 1;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (lam[mid] <= x0) lo = mid; else hi = mid;
        }
        const double w = (x0 - lam[lo]) / (lam[lo + 1] - lam[lo]);
        al = alb[lo] * (1.0 - w) + alb[lo + 1] * w;
        const double wl = std::log(x0 / lam[lo]) / std::log(lam[lo + 1] / lam[lo]);
        cx = std::exp(std::log(std::fmax(ce[lo], DBL_MIN)) * (1.0 - wl) +
                      std::log(std::fmax(ce[lo + 1], DBL_MIN)) * wl);
    }
    return (1.0 - al) * cx / c->dcref;
}

void coeffs(const mc_ctx *c, double E, double *out)
{
    const double he = E * MC_EV2ERG;
    const double sHI = sigma_hydrogenic(E, MC_ETH_HI, 1.0);
    const double sHeI = sigma_HeI(E);
    const double sHeII = sigma_hydrogenic(E, MC_ETH_HEII, 2.0);
    out[0] = sHI / he;
    out[1] = sHeI / he;
    out[2] = sHeII / he;
    out[3] = sHI * std::fmax(1.0 - MC_ETH_HI / E, 0.0);
    out[4] = sHeI * std::fmax(1.0 - MC_ETH_HEI / E, 0.0);
    out[5] = sHeII * std::fmax(1.0 - MC_ETH_HEII / E, 0.0);
    const double rhe = 1.0 / he, rE = 1.0 / E;
    for (int k = 0; k < c->ntr; ++k) {
        double c1 = 0.0, c2 = 0.0;
        if (E >= c->mtab.flat_eth[k]) {
            const double sg = m_sigma(&c->mtab, c->mtab.flat_t[k], E);
            c1 = sg * rhe;
            c2 = sg * std::fmax(1.0 - c->mtab.flat_eth[k] * rE, 0.0);
        }
        out[6 + 2 * k] = c1;
        out[7 + 2 * k] = c2;
    }
    if (c->mom.ndust) out[6 + 2 * c->ntr] = dust_sabs(c, E);
}

// Chebyshev T_0..T_deg at t
void cheb(double t, double *T)
{
    T[0] = 1.0;
    T[1] = t;
    for (int m = 2; m <= kMomDeg; ++m) T[m] = 2.0 * t * T[m - 1] - T[m - 2];
}

// interpolant of every coefficient on the ln-E interval [ua, ub]: a[q][m];
// returns the largest relative check error
double fit_sub(const mc_ctx *c, int nq, double ua, double ub, std::vector<double> &a)
{
    const int nn = kMomDeg + 1;
    const double mid = 0.5 * (ua + ub), half = 0.5 * (ub - ua);
    std::vector<double> v((size_t)nn * nq), w(nq), T(nn);
    a.assign((size_t)nq * nn, 0.0);
    for (int i = 0; i < nn; ++i) {
        const double ti = std::cos(MC_PI * (i + 0.5) / nn);
        coeffs(c, std::exp(mid + half * ti), v.data() + (size_t)i * nq);
        cheb(ti, T.data());
        for (int q = 0; q < nq; ++q)
            for (int m = 0; m < nn; ++m)
                a[(size_t)q * nn + m] += (m == 0 ? 1.0 : 2.0) / nn * v[(size_t)i * nq + q] * T[m];
    }
    std::vector<double> err(nq, 0.0), big(nq, 0.0);
    for (int j = 0; j < kMomCheck; ++j) {
        const double t = -1.0 + 2.0 * (j + 0.5) / kMomCheck;
        coeffs(c, std::exp(mid + half * t), w.data());
        cheb(t, T.data());
        for (int q = 0; q < nq; ++q) {
            double s = 0.0;
            for (int m = 0; m < nn; ++m) s += a[(size_t)q * nn + m] * T[m];
            err[q] = std::max(err[q], std::fabs(s - w[q]));
            big[q] = std::max(big[q], std::fabs(w[q]));
        }
    }
    double worst = 0.0;
    for (int q = 0; q < nq; ++q)
        if (big[q] > 0.0) worst = std::max(worst, err[q] / big[q]);
        else if (err[q] > 0.0) worst = HUGE_VAL;
    return worst;
}

}  // namespace

// Build the sub-bins and interpolation coefficients of the context's band
// and metal registry.  A band bin is first cut at every H/He and metal
// threshold inside it (the band need not carry all of them as edges), then
// each piece into sub-bins.  mom.nsub = 0 (the walk keeps per-coefficient
// sums) when a piece cannot be resolved within kMomMaxSub sub-bins.
int mc_mom_build(mc_ctx *c)
{
    MomTab &m = c->mom;
    m.nsub = 0;
    m.fit_err = 0.0;
    m.built_ntr = c->ntr;
    m.ndust = (c->p.ion_add_dust && c->dn >= 2) ? 1 : 0;
    m.built_dust = m.ndust;
    m.seg.clear();
    m.bseg.assign((size_t)c->nnu * 2, 0.0);
    m.band.clear();
    m.coef.clear();
    const int nq = 6 + 2 * c->ntr + m.ndust;
    m.nq = nq;
    std::vector<double> thr = {MC_ETH_HI, MC_ETH_HEI, MC_ETH_HEII};
    for (int k = 0; k < c->ntr; ++k) thr.push_back(c->mtab.flat_eth[k]);
    // grain absorption is smooth in ln E between the rows of its table only
    if (m.ndust)
        for (double l : c->dlam) thr.push_back(1.23984 / l);
    std::sort(thr.begin(), thr.end());
    std::vector<double> a, coef;
    auto fail = [&]() {
        m.nsub = 0;
        m.seg.clear();
        m.coef.clear();
        m.band.clear();
        return 0;
    };
    for (int b = 0; b < c->nnu; ++b) {
        const double lo = c->eedge[b], hi = c->eedge[b + 1];
        if (!(lo > 0.0) || !(hi > lo)) return fail();
        std::vector<double> bp = {lo};
        for (double e : thr)
            if (e > lo && e < hi && e > bp.back()) bp.push_back(e);
        bp.push_back(hi);
        m.bseg[(size_t)b * 2] = (double)(m.seg.size() / 5);
        m.bseg[(size_t)b * 2 + 1] = (double)(bp.size() - 1);
        for (size_t g = 0; g + 1 < bp.size(); ++g) {
            const double ul = std::log(bp[g]), uh = std::log(bp[g + 1]);
            int n = 1;
            double worst = 0.0;
            for (; n <= kMomMaxSub; ++n) {
                coef.clear();
                worst = 0.0;
                for (int j = 0; j < n && worst <= kMomTol; ++j) {
                    const double ua = ul + (uh - ul) * j / n, ub = ul + (uh - ul) * (j + 1) / n;
                    worst = std::max(worst, fit_sub(c, nq, ua, ub, a));
                    coef.insert(coef.end(), a.begin(), a.end());
                }
                if (worst <= kMomTol) break;
            }
            if (n > kMomMaxSub || m.nsub + n > MC_MOM_MAXSUB) return fail();
            m.seg.insert(m.seg.end(), {bp[g], ul, n / (uh - ul), (double)n, (double)m.nsub});
            for (int j = 0; j < n; ++j) m.band.push_back(b);
            m.coef.insert(m.coef.end(), coef.begin(), coef.end());
            m.nsub += n;
            m.fit_err = std::max(m.fit_err, worst);
        }
    }
    return 0;
}

void mc_mom_free(mc_ctx *c)
{
    for (void *p : {(void *)c->d_mom, (void *)c->d_mom_coef, (void *)c->d_mom_tab,
                    (void *)c->d_mom_bseg, (void *)c->d_mom_band})
        if (p) (void)hipFree(p);
    c->d_mom = c->d_mom_coef = c->d_mom_tab = c->d_mom_bseg = nullptr;
    c->d_mom_band = nullptr;
    c->mom_cap = 0;
    c->mom.built_ntr = -1;
    c->mom.built_dust = -1;
    c->mom.nsub = 0;
}

// Tables on the device and moments for every leaf, built on first use with
// this band and registry.  Returns 1 when the walk uses the moments, 0 when
// it keeps the per-coefficient sums, < 0 on a device error.
int mc_mom_ready(mc_ctx *c)
{
    if (!c->mom_on) return 0;
    const int ndust = (c->p.ion_add_dust && c->dn >= 2) ? 1 : 0;
    if (c->mom.built_ntr != c->ntr || c->mom.built_dust != ndust) {
        mc_mom_free(c);
        if (mc_mom_build(c)) return -1;
    }
    if (c->mom.nsub == 0) return 0;
    if (!c->d_mom_tab) {
        const MomTab &m = c->mom;
        if (hipMalloc(&c->d_mom_tab, sizeof(double) * m.seg.size()) != hipSuccess ||
            hipMalloc(&c->d_mom_bseg, sizeof(double) * m.bseg.size()) != hipSuccess ||
            hipMalloc(&c->d_mom_coef, sizeof(double) * m.coef.size()) != hipSuccess ||
            hipMalloc(&c->d_mom_band, sizeof(int32_t) * m.band.size()) != hipSuccess) {
            mc_set_error("energy-moment tables: hipMalloc failed");
            return -1;
        }
        if (hipMemcpy(c->d_mom_tab, m.seg.data(), sizeof(double) * m.seg.size(),
                      hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(c->d_mom_bseg, m.bseg.data(), sizeof(double) * m.bseg.size(),
                      hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(c->d_mom_coef, m.coef.data(), sizeof(double) * m.coef.size(),
                      hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(c->d_mom_band, m.band.data(), sizeof(int32_t) * m.band.size(),
                      hipMemcpyHostToDevice) != hipSuccess) {
            mc_set_error("energy-moment tables: hipMemcpy failed");
            return -1;
        }
    }
    if (c->mom.nsub == 0) return 0;
    // the leaves' rows, then the replica rows of the hot leaves
    const size_t rows = (size_t)c->nleaf + (size_t)c->st.nrep * c->st.nhot;
    const size_t need = sizeof(double) * rows * c->mom.nsub * MC_MOM_NM;
    if (c->mom_cap < need) {
        if (c->d_mom) (void)hipFree(c->d_mom);
        c->d_mom = nullptr;
        c->mom_cap = 0;
        // the rows are zeroed and read once per pass (2 x need bytes of
        // traffic): beyond half the free device memory, or when the
        // allocation fails, the walk keeps the per-coefficient sums
        size_t fr = 0, tot = 0;
        const bool fits = hipMemGetInfo(&fr, &tot) == hipSuccess && need <= fr / 2;
        if (!fits || hipMalloc(&c->d_mom, need) != hipSuccess) {
            (void)hipGetLastError();
            c->d_mom = nullptr;
            if (!c->mom_warned)
                std::fprintf(stderr,
                             "mochii_gpu: energy-moment tallies need %.3g GB (%zu rows x %d "
                             "sub-bins), more than half the free device memory (%.3g GB); the "
                             "walk keeps the per-coefficient sums\n",
                             (double)need * 1e-9, rows, c->mom.nsub, (double)fr * 1e-9);
            c->mom_warned = 1;
            return 0;
        }
        c->mom_cap = need;
    }
    return 1;
}
