// Host-side set-up of a transport context: band edges, the continuous
// source samplers, the Sobol/Owen tables and the octree build.  Runs once per
// context (or per grid), so it is plain FP64 C++ on the host.
#include <thread>
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <functional>

#include "mc_internal.h"

// ---------------------------------------------------------------------------
// Sobol direction numbers and Owen keys (reference: src/qmc_mod.f90:73-139).
// Dimension 1 is van der Corput; dimensions 2..12 expand the published
// Joe-Kuo new-joe-kuo-6 rows (degree, coefficient bits, initial m_k).
void mc_qmc_tables(int seed, uint32_t V[32][MC_QMC_DIM], uint32_t key[MC_QMC_DIM][2])
{
    static const int deg[MC_QMC_DIM] = {0, 1, 2, 3, 3, 4, 4, 5, 5, 5, 5, 5};
    static const int cof[MC_QMC_DIM] = {0, 0, 1, 1, 2, 1, 4, 2, 4, 7, 11, 13};
    static const int m0[MC_QMC_DIM][5] = {
        {0}, {1}, {1, 3}, {1, 3, 1}, {1, 1, 1}, {1, 1, 3, 3}, {1, 3, 5, 13},
        {1, 1, 5, 5, 17}, {1, 1, 5, 5, 5}, {1, 1, 7, 11, 19}, {1, 1, 5, 1, 1},
        {1, 1, 1, 3, 11}};
    for (int b = 0; b < 32; ++b) V[b][0] = 0x80000000u >> b;
    for (int d = 1; d < MC_QMC_DIM; ++d) {
        const int s = deg[d];
        for (int b = 0; b < s; ++b) V[b][d] = (uint32_t)m0[d][b] << (31 - b);
        for (int b = s; b < 32; ++b) {
            uint32_t v = V[b - s][d] ^ (V[b - s][d] >> s);
            for (int j = 1; j < s; ++j)
                if ((cof[d] >> (s - 1 - j)) & 1) v ^= V[b - j][d];
            V[b][d] = v;
        }
    }
    auto avalanche = [](uint32_t x) {
        x ^= x >> 16; x *= 0x7FEB352Du; x ^= x >> 15; x *= 0x846CA68Bu; x ^= x >> 16;
        return x;
    };
    for (int d = 0; d < MC_QMC_DIM; ++d) {
        const uint32_t b = (uint32_t)seed + (uint32_t)(d + 1) * 0x9E3779B9u;
        key[d][0] = avalanche(b);
        key[d][1] = avalanche(b + 0x85EBCA6Bu);
    }
}

// ---------------------------------------------------------------------------
// Band edges.  With ion_align_edges the H/He thresholds (priority 1) and the
// active metal thresholds (priority 0) strictly inside the band split it into
// sub-segments: near-coincident thresholds (|ln ratio| < 0.01) merge keeping
// the H/He member; while the thresholds outnumber the bins the most redundant
// metal one (smallest log gap to a neighbour) is dropped; the bins are shared
// out by log width, floor >= 1, remainder by largest fractional part
// (reference: build_aligned_edges, src/ion_band_mod.f90:1053-1214).
int mc_setup_band(mc_ctx *c)
{
    const mc_params &p = c->p;
    const int n = p.nnu_ion;
    if (n < 1 || n > MC_MAXBINS) {
        mc_set_error("nnu_ion must be in [1, 64]");
        return -1;
    }
    if (p.nmetal < 0 || p.nmetal > MC_MAX_METAL_TR) {
        mc_set_error("nmetal must be in [0, 64]");
        return -1;
    }
    c->nnu = n;
    c->eedge.assign(n + 1, 0.0);
    const double lo = p.eion_min, hi = p.eion_max;
    if (!p.ion_align_edges) {
        for (int i = 0; i <= n; ++i)
            c->eedge[i] = std::exp(std::log(lo) + (std::log(hi) - std::log(lo)) * (double)i / n);
        return 0;
    }
    std::vector<std::pair<double, int>> cand;
    for (double t : {MC_ETH_HI, MC_ETH_HEI, MC_ETH_HEII})
        if (t > lo && t < hi) cand.push_back({t, 1});
    for (int i = 0; i < p.nmetal; ++i)
        if (p.metal_eth[i] > 0.0 && p.metal_eth[i] > lo && p.metal_eth[i] < hi)
            cand.push_back({p.metal_eth[i], 0});
    std::stable_sort(cand.begin(), cand.end(),
                     [](const std::pair<double, int> &x, const std::pair<double, int> &y) {
                         return x.first < y.first;
                     });
    std::vector<double> cut;
    std::vector<int> pri;
    for (const auto &cd : cand) {
        if (!cut.empty() && std::fabs(std::log(cd.first / cut.back())) < 0.01) {
            if (cd.second > pri.back()) { cut.back() = cd.first; pri.back() = cd.second; }
            continue;
        }
        cut.push_back(cd.first);
        pri.push_back(cd.second);
    }
    while ((int)cut.size() + 1 > n) {
        int best = -1;
        double gap = 1.0e30;
        const int m = (int)cut.size();
        for (int i = 0; i < m; ++i) {
            if (pri[i] != 0) continue;
            double t = i == 0 ? std::log(cut[i] / lo) : std::log(cut[i] / cut[i - 1]);
            t = std::min(t, i == m - 1 ? std::log(hi / cut[i]) : std::log(cut[i + 1] / cut[i]));
            if (t < gap) { gap = t; best = i; }
        }
        if (best < 0) break;
        cut.erase(cut.begin() + best);
        pri.erase(pri.begin() + best);
    }
    std::vector<double> pts;
    pts.push_back(lo);
    pts.insert(pts.end(), cut.begin(), cut.end());
    pts.push_back(hi);
    const int nseg = (int)pts.size() - 1;
    if (n < nseg) {
        mc_set_error("nnu_ion too small for the aligned H/He thresholds");
        return -1;
    }
    std::vector<double> w(nseg);
    double wsum = 0.0;
    for (int j = 0; j < nseg; ++j) {
        w[j] = std::log(pts[j + 1]) - std::log(pts[j]);
        wsum += w[j];
    }
    std::vector<int> nb(nseg);
    int tot = 0;
    for (int j = 0; j < nseg; ++j) {
        nb[j] = std::max(1, (int)((double)n * w[j] / wsum));
        tot += nb[j];
    }
    for (int r = n - tot; r > 0; --r) {
        int best = 0;
        double fr = -1.0;
        for (int j = 0; j < nseg; ++j) {
            const double f = (double)n * w[j] / wsum - (double)nb[j];
            if (f > fr) { fr = f; best = j; }
        }
        nb[best]++;
    }
    for (int r = n - tot; r < 0;) {
        // only reached when the floor of 1 overshoots; r re-evaluated below
        int best = -1;
        double ex = -1.0e30;
        for (int j = 0; j < nseg; ++j) {
            if (nb[j] <= 1) continue;
            const double e = (double)nb[j] - (double)n * w[j] / wsum;
            if (e > ex) { ex = e; best = j; }
        }
        if (best < 0) break;
        nb[best]--;
        ++r;
    }
    int k = 0;
    c->eedge[0] = pts[0];
    for (int j = 0; j < nseg; ++j) {
        const double a = std::log(pts[j]), b = std::log(pts[j + 1]);
        for (int ib = 1; ib <= nb[j]; ++ib) c->eedge[++k] = std::exp(a + (b - a) * (double)ib / nb[j]);
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Planck shape E^3/(e^x - 1) with the small- and large-x forms
// (reference: planck_energy_shape, src/energy_sampler_mod.f90:277-290).
static double planck_shape(double E, double T)
{
    const double x = E / (8.617333262145e-5 * T);
    const double e3 = (E * E) * E;
    if (x <= 0.0) return 0.0;
    if (x < 1.0e-3) return e3 / (x * (1.0 + 0.5 * x + x * x / 6.0));
    if (x < 700.0) return e3 / (std::exp(x) - 1.0);
    return e3 * std::exp(-x);
}

// Adaptive piecewise-linear Planck density on [emin, emax] with mandatory
// knots at the thresholds; an interval is split at its midpoint until the
// trapezoid/Simpson difference is within rtol of the local and the average
// band integral (reference: build_planck_sampler / refine_planck_interval,
// src/energy_sampler_mod.f90:106-185).  Appends knots to e/f.
static int planck_knots(double T, double emin, double emax, double rtol, int nmet,
                        const double *meth, std::vector<double> &e, std::vector<double> &f)
{
    std::vector<double> mand = {emin, emax};
    for (double t : {MC_ETH_HI, MC_ETH_HEI, MC_ETH_HEII})
        if (t > emin && t < emax) mand.push_back(t);
    // photoionization_threshold_knots (src/ion_band_mod.f90:408-420)
    for (int i = 0; i < nmet; ++i)
        if (meth[i] > emin && meth[i] < emax) mand.push_back(meth[i]);
    std::sort(mand.begin(), mand.end());
    mand.erase(std::unique(mand.begin(), mand.end()), mand.end());
    double scale = 0.0;
    for (int i = 1; i <= 256; ++i) {
        const double a = emin + (emax - emin) * (double)(i - 1) / 256.0;
        const double b = emin + (emax - emin) * (double)i / 256.0;
        scale += 0.5 * (planck_shape(a, T) + planck_shape(b, T)) * (b - a);
    }
    const double span = emax - emin;
    e.push_back(mand[0]);
    f.push_back(planck_shape(mand[0], T));
    bool overflow = false;
    std::function<void(double, double, double, double, int)> split =
        [&](double e0, double e1, double f0, double f1, int depth) {
            const double em = 0.5 * (e0 + e1), fm = planck_shape(em, T);
            const double coarse = 0.5 * (f0 + f1) * (e1 - e0);
            const double fine = 0.25 * (f0 + 2.0 * fm + f1) * (e1 - e0);
            const double err = std::fabs(fine - coarse) / 3.0;
            const double allow = rtol * (std::fabs(fine) + scale * (e1 - e0) / span);
            if (err <= allow || depth >= 50) {
                if (e.size() >= 200000) { overflow = true; return; }
                e.push_back(e1);
                f.push_back(f1);
                return;
            }
            split(e0, em, f0, fm, depth + 1);
            split(em, e1, fm, f1, depth + 1);
        };
    for (size_t i = 0; i + 1 < mand.size(); ++i)
        split(mand[i], mand[i + 1], planck_shape(mand[i], T), planck_shape(mand[i + 1], T), 0);
    if (overflow) {
        mc_set_error("Planck sampler knot limit");
        return -1;
    }
    return 0;
}

// One sampler per source component, its band luminosity, the component CDF
// and the packet luminosity (reference: setup_continuous_source_sampler,
// src/ion_band_mod.f90:275-353; Planck spectra, no FUV extension, so the
// ionizing fraction is 1 and band_L is the component's own luminosity).
int mc_setup_samplers(mc_ctx *c)
{
    const mc_params &p = c->p;
    c->ncomp = p.nsource + (p.ext_on ? 1 : 0);
    if (c->ncomp < 1 || p.nsource < 0 || p.nsource > MC_MAXSRC) {
        mc_set_error("need 0..16 point sources and at least one source component");
        return -1;
    }
    c->multi = c->ncomp >= 2;
    c->samp_e.clear(); c->samp_f.clear(); c->samp_c.clear();
    c->samp_off.assign(1, 0);
    std::vector<double> bandL(c->ncomp);
    for (int ic = 0; ic < c->ncomp; ++ic) {
        const int kind = ic < p.nsource ? ic + 1 : 0;
        c->comp_kind[ic] = kind;
        const double T = kind ? p.tstar[kind - 1] : p.ext_tstar;
        if (!(T > 0.0)) {
            mc_set_error("source temperature must be positive");
            return -1;
        }
        std::vector<double> e, f;
        if (planck_knots(T, p.eion_min, p.eion_max, p.source_cdf_tol > 0 ? p.source_cdf_tol : 1e-8,
                         p.nmetal, p.metal_eth, e, f))
            return -1;
        std::vector<double> cdf(e.size(), 0.0);
        for (size_t i = 0; i + 1 < e.size(); ++i)
            cdf[i + 1] = cdf[i] + 0.5 * (f[i] + f[i + 1]) * (e[i + 1] - e[i]);
        const double tot = cdf.back();
        for (auto &v : cdf) v = v / tot;
        cdf.back() = 1.0;
        c->samp_e.insert(c->samp_e.end(), e.begin(), e.end());
        c->samp_f.insert(c->samp_f.end(), f.begin(), f.end());
        c->samp_c.insert(c->samp_c.end(), cdf.begin(), cdf.end());
        c->samp_off.push_back((int)c->samp_e.size());
        if (kind) {
            bandL[ic] = p.src_lum[kind - 1];
        } else {
            double area;
            if (p.ext_geometry == MC_EXT_SPH) {
                const double r = p.rmax * p.distance2cm;
                area = MC_FOURPI * (r * r);
            } else {
                const double xr = 2 * p.xmax * p.distance2cm, yr = 2 * p.ymax * p.distance2cm,
                             zr = 2 * p.zmax * p.distance2cm;
                area = 2.0 * (xr * yr + yr * zr + zr * xr);
            }
            bandL[ic] = MC_PI * p.ext_intensity * area;
        }
    }
    c->Ltot = 0.0;
    for (double L : bandL) c->Ltot += L;
    double acc = 0.0;
    for (int ic = 0; ic < c->ncomp; ++ic) {
        acc += bandL[ic];
        c->src_cdf[ic] = acc;
    }
    for (int ic = 0; ic < c->ncomp; ++ic) c->src_cdf[ic] /= acc;
    c->qmc_ndim = (!c->multi && p.nsource == 1) ? 3 : 7;
    return 0;
}

// ---------------------------------------------------------------------------
// Octree from a leaf list (reference: amr_build_tree and
// amr_build_neighbors, src/octree_mod.f90:347-497).  Cells are numbered in
// creation order (0-based); the neighbour of a face is the cell of the same
// level (or the coarser leaf) across it, 0-based, -1 outside or when the
// lookup lands on an ancestor.
int mc_build_octree_host(mc_ctx *c, int nleaf, const double *xl, const double *yl,
                         const double *zl, const int32_t *lev, double half)
{
    c->car = 0;
    c->xmin = c->ymin = c->zmin = -half;
    c->xmaxb = c->ymaxb = c->zmaxb = half;
    auto &cx = c->hcx, &cy = c->hcy, &cz = c->hcz, &ch = c->hch;
    auto &il = c->hileaf, &kid = c->hchildren, &lv = c->hlevel, &par = c->hparent;
    cx.assign(1, 0.0); cy.assign(1, 0.0); cz.assign(1, 0.0); ch.assign(1, half);
    il.assign(1, -1); kid.assign(8, -1); lv.assign(1, 0); par.assign(1, -1);
    c->hcell_of_leaf.assign(nleaf, -1);
    for (int i = 0; i < nleaf; ++i) {
        int cell = 0;
        for (int l = 0; l < lev[i]; ++l) {
            const int o = (xl[i] >= cx[cell] ? 1 : 0) + (yl[i] >= cy[cell] ? 2 : 0) +
                          (zl[i] >= cz[cell] ? 4 : 0);
            if (kid[cell * 8 + o] < 0) {
                const int n = (int)cx.size();
                const double h = ch[cell] * 0.5;
                cx.push_back(cx[cell] + (double)(2 * (o & 1) - 1) * h);
                cy.push_back(cy[cell] + (double)(2 * ((o >> 1) & 1) - 1) * h);
                cz.push_back(cz[cell] + (double)(2 * (o >> 2) - 1) * h);
                ch.push_back(h);
                il.push_back(-1);
                lv.push_back(l + 1);
                par.push_back(cell);
                kid.insert(kid.end(), 8, -1);
                kid[cell * 8 + o] = n;
            }
            cell = kid[cell * 8 + o];
        }
        if (il[cell] >= 0) {
            mc_set_error("two leaves map to one octree cell");
            return -1;
        }
        il[cell] = i;
        c->hcell_of_leaf[i] = cell;
    }
    const int nc = (int)cx.size();
    for (int i = 0; i < nleaf; ++i)
        if (c->hcell_of_leaf[i] < 0) {
            mc_set_error("leaf outside the tree");
            return -1;
        }
    auto at_level = [&](double x, double y, double z, int level) {
        if (x < -half || x > half || y < -half || y > half || z < -half || z > half) return -1;
        int cell = 0;
        for (;;) {
            if (lv[cell] >= level || il[cell] >= 0) return cell;
            const int o = (x >= cx[cell] ? 1 : 0) + (y >= cy[cell] ? 2 : 0) + (z >= cz[cell] ? 4 : 0);
            const int k = kid[cell * 8 + o];
            if (k < 0) return cell;
            cell = k;
        }
    };
    auto ancestor = [&](int a, int d) {
        for (int k = par[d]; k >= 0; k = par[k])
            if (k == a) return true;
        return false;
    };
    c->hneighbor.assign((size_t)nc * 6, -1);
    // the cells' face neighbours are independent lookups: host threads
    auto nbr_range = [&](int i0, int i1) {
        for (int i = i0; i < i1; ++i) {
            const double x = cx[i], y = cy[i], z = cz[i], hp = 2.0 * ch[i];
            int *nb = &c->hneighbor[(size_t)i * 6];
            if (x + hp <= half) nb[0] = at_level(x + hp, y, z, lv[i]);
            if (x - hp >= -half) nb[1] = at_level(x - hp, y, z, lv[i]);
            if (y + hp <= half) nb[2] = at_level(x, y + hp, z, lv[i]);
            if (y - hp >= -half) nb[3] = at_level(x, y - hp, z, lv[i]);
            if (z + hp <= half) nb[4] = at_level(x, y, z + hp, lv[i]);
            if (z - hp >= -half) nb[5] = at_level(x, y, z - hp, lv[i]);
            for (int f = 0; f < 6; ++f)
                if (nb[f] >= 0 && nb[f] != i && ancestor(nb[f], i)) nb[f] = -1;
        }
    };
    const int nth = nc < 65536 ? 1
                               : (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (int t = 1; t < nth; ++t)
        th.emplace_back(nbr_range, (int)((int64_t)nc * t / nth), (int)((int64_t)nc * (t + 1) / nth));
    nbr_range(0, (int)((int64_t)nc / nth));
    for (auto &t : th) t.join();
    c->ncells = nc;
    c->nleaf = nleaf;
    return 0;
}

// ---------------------------------------------------------------------------
// Thermally averaged free-free Gaunt factor (reference: gaunt_ff_mean and
// getGauntFF, src/cooling_mod.f90:182-220, src/gaunt.f90).  Hummer (1988,
// ApJ 327, 477) two-dimensional Chebyshev fit in single precision; the
// coefficients are single-precision values widened to double, as in the
// reference's table.  Setup only.
static const float kHummer[88] = {
    8.986940175e+00f, -4.009515855e+00f, 8.808871266e-01f, 2.640245111e-02f, -4.580645915e-02f,
    -3.568055702e-03f, 2.827798067e-03f, 3.365860195e-04f, -8.006936989e-01f, 9.466021705e-01f,
    9.043402532e-02f, -9.608451450e-02f, -1.885629865e-02f, 1.050313890e-02f, 2.800889961e-03f,
    -1.078209202e-03f, -3.781305103e-01f, 1.102726332e-01f, -1.543619180e-02f, 8.310561114e-03f,
    2.179620525e-02f, 4.259726289e-03f, -4.181588794e-03f, -1.770208330e-03f, 1.877213132e-02f,
    -1.004885705e-01f, -5.483366378e-02f, -4.520154409e-03f, 8.366530426e-03f, 3.700273930e-03f,
    6.889320423e-04f, 9.460313195e-05f, 7.300158392e-02f, 3.576785497e-03f, -4.545307025e-03f,
    -1.017965604e-02f, -9.530211924e-03f, -3.450186162e-03f, 1.040482914e-03f, 1.407073544e-03f,
    -1.744671550e-03f, 2.864013856e-02f, 1.903394837e-02f, 7.091074494e-03f, -9.668371391e-04f,
    -2.999107465e-03f, -1.820642230e-03f, -3.874082085e-04f, -1.707268366e-02f, -4.694254776e-03f,
    1.311691517e-03f, 5.316703136e-03f, 5.178193095e-03f, 2.451228935e-03f, -2.277321615e-05f,
    -8.182359057e-04f, 2.567331664e-04f, -9.155339970e-03f, -6.997479192e-03f, -3.571518641e-03f,
    -2.096101038e-04f, 1.553822487e-03f, 1.509584686e-03f, 6.212627837e-04f, 4.098322531e-03f,
    1.635218463e-03f, -5.918883504e-04f, -2.333091048e-03f, -2.484138313e-03f, -1.359996060e-03f,
    -5.371426147e-05f, 5.553549563e-04f, 3.837562402e-05f, 2.938325230e-03f, 2.393747064e-03f,
    1.328839809e-03f, 9.135013312e-05f, -7.137252303e-04f, -7.656848158e-04f, -3.504683798e-04f,
    -8.491991820e-04f, -3.615327726e-04f, 3.148015257e-04f, 8.909207650e-04f, 9.869737522e-04f,
    6.134671184e-04f, 1.068883394e-04f, -2.046080100e-04f};

static void gaunt_ff(float z, float lt, const float *xlf, float *g, int n)
{
    const float xl = lt < 2.5f ? 5.1983649f - 2.5f : (lt > 8.0f ? 5.1983649f - 8.0f : 5.1983649f - lt);
    const float txg = 0.66666667f * (2.0f * z + xl);
    const float gamma2 = std::pow(10.0f, txg * 1.5f);
    const float con = 0.72727273f * xl + 0.90909091f;
    float c[9];
    for (int j = 0; j < 8; ++j) {
        // Clenshaw sum over the 11 gamma^2 coefficients of row j
        float b[13];
        b[11] = kHummer[10 * 8 + j];
        b[10] = (float)((double)(txg * b[11]) + (double)kHummer[9 * 8 + j]);
        for (int ir = 9; ir >= 1; --ir)
            b[ir] = (float)((double)(txg * b[ir + 1] - b[ir + 2]) + (double)kHummer[(ir - 1) * 8 + j]);
        c[j + 1] = 0.25f * (b[1] - b[3]);
    }
    for (int i = 0; i < n; ++i) {
        const float txu = 0.72727273f * xlf[i] + con;
        const float u = std::pow(10.0f, (txu - 0.90909091f) / 0.72727273f);
        if (std::fabs(txu) <= 2.0f) {
            float b[9];
            b[8] = c[8];
            b[7] = txu * b[8] + c[7];
            for (int ir = 6; ir >= 1; --ir) b[ir] = txu * b[ir + 1] - b[ir + 2] + c[ir];
            g[i] = b[1] - b[3];
        } else if (std::log10(u) < -4.0f && gamma2 < 0.3f) {
            g[i] = 0.551329f * std::log(2.24593f / u);
        } else if (std::log10(u) < -4.0f) {
            g[i] = 0.551329f * std::log(0.944931f / u / std::sqrt(gamma2));
        } else if (txu > 2.0f) {
            if (std::log10(gamma2) < -3.0f) {
                g[i] = std::sqrt(0.9549297f / u);
            } else if (std::log10(gamma2) < 0.0f) {
                const float sl = (std::sqrt(12.0f * gamma2 / u) - std::sqrt(0.9549297f / u)) / 3.0f;
                g[i] = std::sqrt(12.0f * gamma2 / u) + sl * std::log10(gamma2);
            } else {
                g[i] = std::fmin(1.05f, std::sqrt(12.0f * gamma2 / u));
            }
        }
    }
}

// gbar_ff table on log T = 2..6 (41 points) for net charge Z = 1..4: the
// 64-point midpoint rule in ln u over u in [1e-4, 20] with weight e^-u,
// normalised by the integral of e^-u over the same range
void mc_gff_table(double logT[41], double gff[4][41])
{
    const int nu = 64;
    const double lo = std::log(1.0e-4), hi = std::log(20.0);
    double u[64], du[64];
    for (int i = 0; i < nu; ++i) {
        u[i] = std::exp(lo + (hi - lo) * ((double)(i + 1) - 0.5) / nu);
        du[i] = u[i] * (hi - lo) / nu;
    }
    for (int it = 0; it < 41; ++it) {
        logT[it] = 2.0 + 4.0 * (double)it / 40.0;
        const double T = std::pow(10.0, logT[it]);
        const double r = 13.605693 * MC_EV2ERG / (MC_KBOLTZ * T);
        float xlf[64], g[64];
        for (int i = 0; i < nu; ++i) xlf[i] = (float)std::log10(u[i] / r);
        for (int iz = 1; iz <= 4; ++iz) {
            gaunt_ff((float)std::log10((double)iz), (float)std::log10(T), xlf, g, nu);
            double s = 0.0;
            for (int i = 0; i < nu; ++i) s += (double)g[i] * std::exp(-u[i]) * du[i];
            gff[iz - 1][it] = s / (std::exp(-1.0e-4) - std::exp(-20.0));
        }
    }
}

// ---- Milne ground-continuum tables ------------------------------------------
// milne_setup (src/milne_recomb_spectrum_mod.f90:124-206): for each channel
// and temperature node, cumulative Simpson of the photon emissivity
// n(x) = E^2 sigma(E) e^-x and of E n(x) over x nodes 60 t^2 (t uniform on
// [0,1]), truncated at the band top; the energy CDF is normalised to 1, and
// <x> and the in-band photon fraction (against the integral to x = 60) are
// kept for the energy radiated per recombination.
#pragma clang fp contract(off)
static double milne_shape(int ich, double eth, double x, double kT)
{
    const double E = eth + x * kT;
    const double s = ich == 0 ? sigma_hydrogenic(E, MC_ETH_HI, 1.0)
                              : (ich == 1 ? sigma_HeI(E) : sigma_hydrogenic(E, MC_ETH_HEII, 2.0));
    return ((E * E) * s) * std::exp(-x);
}

void mc_milne_tables(double eion_max, double *cdf, double *xbar, double *fin, double *xg,
                     double *lgT)
{
    const int NT = MC_MILNE_NT, NX = MC_MILNE_NX;
    const double xmax = 60.0;
    const double eth[3] = {MC_ETH_HI, MC_ETH_HEI, MC_ETH_HEII};
    for (int i = 0; i < NX; ++i) {
        const double t = (double)i / (double)(NX - 1);
        xg[i] = (xmax * t) * t;
    }
    const double l0 = std::log10(1.0e2), l1 = std::log10(1.0e6);
    for (int i = 0; i < NT; ++i) lgT[i] = l0 + ((l1 - l0) * (double)i) / (double)(NT - 1);
    for (int ich = 0; ich < 3; ++ich) {
        for (int it = 0; it < NT; ++it) {
            double *row = cdf + ((size_t)ich * NT + it) * NX;
            const double kT = MC_KBOLTZ * std::pow(10.0, lgT[it]) / MC_EV2ERG;
            const double xhi = std::min(xmax, (eion_max - eth[ich]) / kT);
            double &xb_out = xbar[ich * NT + it], &fin_out = fin[ich * NT + it];
            if (xhi <= 0.0) {
                for (int i = 0; i < NX; ++i) row[i] = 1.0;
                xb_out = 0.0;
                fin_out = 0.0;
                continue;
            }
            double sn = 0.0, se = 0.0, sxn = 0.0;
            row[0] = 0.0;
            for (int i = 0; i + 1 < NX; ++i) {
                const double xa = xg[i], xb = std::min(xg[i + 1], xhi);
                if (xb > xa) {
                    const double xm = 0.5 * (xa + xb);
                    const double fa = milne_shape(ich, eth[ich], xa, kT);
                    const double fm = milne_shape(ich, eth[ich], xm, kT);
                    const double fb = milne_shape(ich, eth[ich], xb, kT);
                    const double h6 = (xb - xa) / 6.0;
                    sn = sn + h6 * ((fa + 4.0 * fm) + fb);
                    se = se + h6 * ((fa * (eth[ich] + xa * kT) + (4.0 * fm) * (eth[ich] + xm * kT)) +
                                    fb * (eth[ich] + xb * kT));
                    sxn = sxn + h6 * ((fa * xa + (4.0 * fm) * xm) + fb * xb);
                }
                row[i + 1] = se;
            }
            double tail = sn;
            for (double xa = xhi; xa < xmax;) {
                const double xb = std::min(xa + 0.25, xmax), xm = 0.5 * (xa + xb);
                tail = tail + ((xb - xa) / 6.0) *
                                  ((milne_shape(ich, eth[ich], xa, kT) +
                                    4.0 * milne_shape(ich, eth[ich], xm, kT)) +
                                   milne_shape(ich, eth[ich], xb, kT));
                xa = xb;
            }
            for (int i = 0; i < NX; ++i) row[i] = se > 0.0 ? row[i] / se : 1.0;
            row[NX - 1] = 1.0;
            xb_out = sxn / std::max(sn, DBL_MIN);
            fin_out = sn / std::max(tail, DBL_MIN);
        }
    }
}

// hei_2ph_setup (src/hei_twophoton_mod.f90:98-183): the He I 2^1S pair
// spectrum A(y), y = E / 20.62 eV, linear between table rows and zero
// outside them (shape_A, :81-92).  Simpson over each table interval gives the
// normalisation (photons per decay = 2 over the whole pair spectrum); Simpson
// over 1023 equal panels of the H-ionizing window gives the in-window photon
// and energy moments and the cumulative energy-weighted density y A(y) that
// the sampler inverts.  Same operation order as the reference, contraction
// off, so the moments agree with it to the last bit.
namespace {
struct Hei2phShape {
    int n;
    const double *y, *A;
    double operator()(double v) const
    {
        if (v <= y[0] || v >= y[n - 1]) return 0.0;
        int i = 0;
        while (i < n - 2 && y[i + 1] < v) ++i;
        return A[i] + ((A[i + 1] - A[i]) * (v - y[i])) / (y[i + 1] - y[i]);
    }
};
}  // namespace

int mc_hei2ph_tables(int n, const double *y, const double *A, double eion_max, double *grid_y,
                     double *cdf_e, double *n_per, double *e_per)
{
    if (n < 4 || n > MC_HEI2PH_MAXPTS) return -1;
    for (int i = 0; i + 1 < n; ++i)
        if (!(y[i + 1] > y[i])) return -2;
    const Hei2phShape f{n, y, A};
    double tot = 0.0;
    for (int i = 0; i + 1 < n; ++i) {
        const double ya = y[i], yb = y[i + 1], ym = 0.5 * (ya + yb);
        tot = tot + ((yb - ya) / 6.0) * ((A[i] + 4.0 * f(ym)) + A[i + 1]);
    }
    if (!(tot > 0.0)) return -3;
    const int N = MC_HEI2PH_NCDF;
    const double y_lo = MC_ETH_HI / MC_HEI_E_TOT;
    const double y_hi = std::min(MC_HEI_E_TOT, eion_max) / MC_HEI_E_TOT;
    *n_per = 0.0;
    *e_per = 0.0;
    if (y_hi <= y_lo) {            // window closed by the band edge
        for (int i = 0; i < N; ++i) {
            grid_y[i] = 0.0;
            cdf_e[i] = 1.0;
        }
        return 0;
    }
    const double span = y_hi - y_lo;
    auto node = [&](int i) { return y_lo + (span * (double)i) / (double)(N - 1); };
    double acc_n = 0.0, acc_e = 0.0, seg = 0.0;
    grid_y[0] = y_lo;
    cdf_e[0] = 0.0;
    for (int i = 1; i < N; ++i) {
        const double ya = node(i - 1), yb = node(i), ym = 0.5 * (ya + yb), h6 = (yb - ya) / 6.0;
        const double fa = f(ya), fm = f(ym), fb = f(yb);
        acc_n = acc_n + h6 * ((fa + 4.0 * fm) + fb);
        acc_e = acc_e + (h6 * (((fa * ya) + (4.0 * fm) * ym) + fb * yb)) * MC_HEI_E_TOT;
        seg = seg + h6 * (((fa * ya) + 4.0 * (fm * ym)) + fb * yb);
        grid_y[i] = yb;
        cdf_e[i] = seg;
    }
    *n_per = (2.0 * acc_n) / tot;
    *e_per = (2.0 * acc_e) / tot;
    for (int i = 0; i < N; ++i) cdf_e[i] = seg > 0.0 ? cdf_e[i] / seg : 1.0;
    cdf_e[N - 1] = 1.0;
    return 0;
}
