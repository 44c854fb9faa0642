/*
 * Oracle: one ionizing-band transport pass (TEST INFRASTRUCTURE).
 *
 * Restates, for the configurations on this path:
 *   src/ionizing_field_mod.f90:72-170      pass driver (stellar packets)
 *   src/ion_band_mod.f90:1880-1935,2232-2385 QMC launch and emitters
 *   src/ion_packet_mod.f90:21-87            packet cross sections, opacity
 *   src/species_mod.f90:263-343             metal cross sections, cached
 *                                           opacity and path tallies
 *   src/raytrace_amr.f90:249-676            transport_ion_packet, edge and
 *                                           tau walks (octree and car DDA)
 *   src/ion_score_mod.f90:137-252           path scoring and energy ledger
 *   src/gas_rates_mod.f90:82-153,
 *   src/ion_score_mod.f90:298-338           rate normalisation
 * The MPI data parallelism over packets is restated as threads over packet
 * ids with private tallies reduced in a fixed order (deterministic for a
 * given thread count).  Post-launch draws (dust scattering) come from the
 * counter-based generator orc_rng_uniform keyed by (qmc_seed, stream, id,
 * draw number) instead of MT19937.
 *
 * Pinned by: the analytic single-ray expectation
 * w (e^-tau_in - e^-tau_out)/kappa per cell and its sum rule
 * sum_segments = w (1 - e^-tau_edge)/kappa in a uniform medium; brute-force
 * cell enumeration on small grids; the energy-closure identity; the
 * octree walk on a single-level tree against the car walk; Henyey-Greenstein
 * moments <cos theta> = g; and end to end the analytic Stromgren radius.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#include <pthread.h>
#include "mc_oracle.h"

#define WMIN 1.0e-4
#define PSURV 0.1
#define MAX_WALK_STEPS 20000000

typedef struct {
    double x, y, z, kx, ky, kz;
    double wgt, Lpacket, E;
    double lam;                        /* wavelength [um] (infrared band) */
    int inu, icell, nscatt, inside;
    int64_t id;
    int istream;
    uint32_t nk;                       /* draws consumed from the generator */
    double sHI, sHeI, sHeII, dabs, dsca, dg;
    int nmet;                          /* metal transitions below E (flat order) */
    double smet[ORC_MAX_TR];
} pkt_t;

typedef struct {
    double *hhe, *jt, *dust, *mraw;
    int64_t seg_direct, seg_scattered;
    long double e_emit, e_abs, e_esc;
    /* infrared band (transported_band == BAND_DUST_IR): grain-absorption path
     * tally, escaped luminosity per emission-grid bin and its edges */
    int ir, nl;
    double *irabs, *lesc;
    const double *edge;
} tally_t;

static double rnd(const orc_ctx *c, pkt_t *ph)
{
    return orc_rng_uniform((uint32_t)c->p.qmc_seed, ph->istream, (uint64_t)ph->id, ph->nk++);
}

/* ---- geometry (src/octree_mod.f90) -------------------------------------- */
static int car_index(const orc_ctx *c, double x, double y, double z)
{
    int ix = (int)((x - c->xmin) / c->dcell);
    int iy = (int)((y - c->ymin) / c->dcell);
    int iz = (int)((z - c->zmin) / c->dcell);
    const int nx = c->p.nx, ny = c->p.ny, nz = c->p.nz;
    ix = ix < 0 ? 0 : (ix > nx - 1 ? nx - 1 : ix);
    iy = iy < 0 ? 0 : (iy > ny - 1 ? ny - 1 : iy);
    iz = iz < 0 ? 0 : (iz > nz - 1 ? nz - 1 : iz);
    return 1 + ix + nx * (iy + ny * iz);
}

/* amr_find_leaf (src/octree_mod.f90:69-93) */
static int find_leaf(const orc_ctx *c, double x, double y, double z)
{
    if (x < c->xmin || x > c->xmaxb || y < c->ymin || y > c->ymaxb || z < c->zmin ||
        z > c->zmaxb)
        return 0;
    if (c->car) return car_index(c, x, y, z);
    int ic = 1;
    for (;;) {
        if (c->ileaf[ic - 1] > 0) return c->ileaf[ic - 1];
        int o = 1;
        if (x >= c->cx[ic - 1]) o += 1;
        if (y >= c->cy[ic - 1]) o += 2;
        if (z >= c->cz[ic - 1]) o += 4;
        ic = c->children[(ic - 1) * 8 + o - 1];
        if (ic == 0) return 0;
    }
}

static void cell_geom(const orc_ctx *c, int icell, double *cx, double *cy, double *cz,
                      double *h)
{
    if (c->car) {
        int nxy = c->p.nx * c->p.ny;
        int iz = (icell - 1) / nxy, rem = (icell - 1) - iz * nxy;
        int iy = rem / c->p.nx, ix = rem - iy * c->p.nx;
        *cx = c->xmin + ((double)ix + 0.5) * c->dcell;
        *cy = c->ymin + ((double)iy + 0.5) * c->dcell;
        *cz = c->zmin + ((double)iz + 0.5) * c->dcell;
        *h = 0.5 * c->dcell;
    } else {
        *cx = c->cx[icell - 1]; *cy = c->cy[icell - 1]; *cz = c->cz[icell - 1];
        *h = c->ch[icell - 1];
    }
}

/* amr_cell_exit (src/octree_mod.f90:309-340): first minimum on ties */
static void cell_exit(const orc_ctx *c, int icell, double x, double y, double z, double kx,
                      double ky, double kz, double *t_exit, int *iface)
{
    double cx, cy, cz, h, t[6];
    cell_geom(c, icell, &cx, &cy, &cz, &h);
    if (kx > 0.0) { t[0] = (cx + h - x) / kx; t[1] = DBL_MAX; }
    else if (kx < 0.0) { t[0] = DBL_MAX; t[1] = (cx - h - x) / kx; }
    else { t[0] = DBL_MAX; t[1] = DBL_MAX; }
    if (ky > 0.0) { t[2] = (cy + h - y) / ky; t[3] = DBL_MAX; }
    else if (ky < 0.0) { t[2] = DBL_MAX; t[3] = (cy - h - y) / ky; }
    else { t[2] = DBL_MAX; t[3] = DBL_MAX; }
    if (kz > 0.0) { t[4] = (cz + h - z) / kz; t[5] = DBL_MAX; }
    else if (kz < 0.0) { t[4] = DBL_MAX; t[5] = (cz - h - z) / kz; }
    else { t[4] = DBL_MAX; t[5] = DBL_MAX; }
    int f = 0;
    for (int i = 1; i < 6; ++i) if (t[i] < t[f]) f = i;
    *iface = f + 1;
    *t_exit = t[f];
}

/* car_next (src/octree_mod.f90:620-641) */
static int car_next(const orc_ctx *c, int icell, int iface)
{
    int nx = c->p.nx, ny = c->p.ny, nz = c->p.nz, nxy = nx * ny;
    int iz = (icell - 1) / nxy, rem = (icell - 1) - iz * nxy;
    int iy = rem / nx, ix = rem - iy * nx;
    switch (iface) {
    case 1: ix++; break; case 2: ix--; break; case 3: iy++; break;
    case 4: iy--; break; case 5: iz++; break; default: iz--; break;
    }
    if (ix < 0 || ix >= nx || iy < 0 || iy >= ny || iz < 0 || iz >= nz) return 0;
    return 1 + ix + nx * (iy + ny * iz);
}

/* amr_next_leaf (src/octree_mod.f90:518-559) */
static int next_leaf(const orc_ctx *c, int icell, int iface, double x, double y, double z)
{
    if (c->car) return car_next(c, icell, iface);
    int ng = c->neighbor[(size_t)(icell - 1) * 6 + iface - 1];
    if (ng == 0) return 0;
    while (c->ileaf[ng - 1] == 0) {
        int o = 1;
        switch (iface) {
        case 1: if (y >= c->cy[ng - 1]) o += 2; if (z >= c->cz[ng - 1]) o += 4; break;
        case 2: o += 1; if (y >= c->cy[ng - 1]) o += 2; if (z >= c->cz[ng - 1]) o += 4; break;
        case 3: if (x >= c->cx[ng - 1]) o += 1; if (z >= c->cz[ng - 1]) o += 4; break;
        case 4: if (x >= c->cx[ng - 1]) o += 1; o += 2; if (z >= c->cz[ng - 1]) o += 4; break;
        case 5: if (x >= c->cx[ng - 1]) o += 1; if (y >= c->cy[ng - 1]) o += 2; break;
        default: if (x >= c->cx[ng - 1]) o += 1; if (y >= c->cy[ng - 1]) o += 2; o += 4; break;
        }
        int child = c->children[(ng - 1) * 8 + o - 1];
        if (child == 0) break;
        ng = child;
    }
    return c->ileaf[ng - 1];
}

static int leaf_cell(const orc_ctx *c, int il) { return c->car ? il : c->icell_of_leaf[il - 1]; }

static double leaf_half(const orc_ctx *c, int il)
{
    return c->car ? 0.5 * c->dcell : c->ch[c->icell_of_leaf[il - 1] - 1];
}

/* ---- packet physics (src/ion_packet_mod.f90:21-87) ---------------------- */
static void build_physics(const orc_ctx *c, pkt_t *ph)
{
    ph->sHI = orc_sigma_HI(ph->E);
    ph->sHeI = orc_sigma_HeI(ph->E);
    ph->sHeII = orc_sigma_HeII(ph->E);
    ph->nmet = 0;
    if (c->use_metals) orc_species_packet_cross_sections(c, ph->E, ph->smet, &ph->nmet);
    ph->dabs = ph->dsca = ph->dg = 0.0;
    if (c->p.add_dust) orc_dust_at(c, ph->E, &ph->dabs, &ph->dsca, &ph->dg);
}

static double opacity(const orc_ctx *c, const pkt_t *ph, int il)
{
    int i = il - 1;
    /* species_cached_opacity (src/species_mod.f90:291-301) */
    double metal = 0.0;
    if (c->use_metals && c->p.ion_metal_abs)
        for (int k = 0; k < ph->nmet; ++k) {
            int ie = c->flat_ie[k];
            metal = metal + c->el[ie].abund *
                                c->mfrac[(size_t)i * c->nst + c->st_off[ie] + c->flat_it[k]] *
                                ph->smet[k];
        }
    double kap = c->nH[i] * (c->xHI[i] * ph->sHI +
                             c->p.He_abund * (c->xHeI[i] * ph->sHeI + c->xHeII[i] * ph->sHeII) +
                             metal) *
                 c->p.distance2cm;
    if (c->p.add_dust) {
        kap = kap + c->rhokap[i] * ph->dabs;
        if (c->p.dust_scatter) kap = kap + c->rhokap[i] * ph->dsca;
    }
    return kap;
}

/* score_ion_path (src/ion_score_mod.f90:137-231) */
static void score(const orc_ctx *c, tally_t *t, const pkt_t *ph, int il, double pl)
{
    if (pl == 0.0) return;
    int i = il - 1;
    if (t->ir) {
        /* the infrared branch (src/ion_score_mod.f90:149-158): grain
         * absorption into the band's own estimator (dust_ir_absorb_path,
         * src/dust_ir_band_mod.f90:102-107), nothing else */
        t->irabs[i] += pl * ph->dabs;
        if (ph->nscatt > 0) t->seg_scattered++; else t->seg_direct++;
        return;
    }
    t->jt[(size_t)i * c->nnu_band + ph->inu - 1] += pl;
    if (c->p.add_dust) t->dust[i] += pl * ph->dabs;
    if (ph->nscatt > 0) t->seg_scattered++; else t->seg_direct++;
    double E = ph->E, coeff[6];
    coeff[0] = ph->sHI / (E * ORC_EV2ERG);
    coeff[1] = ph->sHeI / (E * ORC_EV2ERG);
    coeff[2] = ph->sHeII / (E * ORC_EV2ERG);
    double f;
    f = 1.0 - ORC_ETH_HI / E;   coeff[3] = ph->sHI * (f > 0.0 ? f : 0.0);
    f = 1.0 - ORC_ETH_HEI / E;  coeff[4] = ph->sHeI * (f > 0.0 ? f : 0.0);
    f = 1.0 - ORC_ETH_HEII / E; coeff[5] = ph->sHeII * (f > 0.0 ? f : 0.0);
    double *h = t->hhe + (size_t)i * 6;
    for (int k = 0; k < 6; ++k) h[k] = h[k] + pl * coeff[k];
    /* species_rate_path_score (src/species_mod.f90:329-343) */
    if (c->use_metals) {
        double *r = t->mraw + (size_t)i * c->ntr * 2;
        for (int k = 0; k < ph->nmet; ++k) {
            double f = 1.0 - c->flat_eth[k] / E;
            r[2 * k] = r[2 * k] + pl * ph->smet[k] / (E * ORC_EV2ERG);
            r[2 * k + 1] = r[2 * k + 1] + pl * ph->smet[k] * (f > 0.0 ? f : 0.0);
        }
    }
}

/* ion_score_energy_packet (src/ion_score_mod.f90:233-252) */
static void ledger(tally_t *t, const pkt_t *ph, double tau_edge)
{
    double L = ph->Lpacket * ph->wgt;
    double tr = exp(-(tau_edge > 0.0 ? tau_edge : 0.0));
    t->e_emit += (long double)L;
    t->e_abs += (long double)(L * (1.0 - tr));
    t->e_esc += (long double)(L * tr);
    if (!t->ir) return;
    /* dust_ir_escape_packet (src/dust_ir_band_mod.f90:123-150): the escaped
     * luminosity into the bin of the packet's wavelength, the first bin k
     * with lambda < edge(k+1); outside the outer edges nothing is binned */
    double lesc = L * tr;
    int n = t->nl;
    if (n <= 0 || lesc <= 0.0) return;
    if (ph->lam <= t->edge[0] || ph->lam >= t->edge[n]) return;
    int lo = 1, hi = n;
    while (lo < hi) {
        int mid = (lo + hi) / 2;
        if (ph->lam < t->edge[mid]) hi = mid;
        else lo = mid + 1;
    }
    t->lesc[lo - 1] += lesc;
}

/* ---- Cartesian DDA (src/raytrace_amr.f90:495-676) ----------------------- */
typedef struct { int ix, iy, iz, stp[3]; double tmax[3], tdel[3]; } dda_t;

static void dda_init(const orc_ctx *c, double x, double y, double z, double kx, double ky,
                     double kz, dda_t *d)
{
    double dc = c->dcell;
    double pos[3] = {x, y, z}, kk[3] = {kx, ky, kz};
    double org[3] = {c->xmin, c->ymin, c->zmin};
    int n[3] = {c->p.nx, c->p.ny, c->p.nz}, idx[3];
    for (int a = 0; a < 3; ++a) {
        int i = (int)((pos[a] - org[a]) / dc);
        idx[a] = i < 0 ? 0 : (i > n[a] - 1 ? n[a] - 1 : i);
    }
    d->ix = idx[0]; d->iy = idx[1]; d->iz = idx[2];
    for (int a = 0; a < 3; ++a) {
        if (kk[a] > 0.0) {
            d->stp[a] = 1;
            d->tmax[a] = (org[a] + (double)(idx[a] + 1) * dc - pos[a]) / kk[a];
            d->tdel[a] = dc / kk[a];
        } else if (kk[a] < 0.0) {
            d->stp[a] = -1;
            d->tmax[a] = (org[a] + (double)idx[a] * dc - pos[a]) / kk[a];
            d->tdel[a] = -dc / kk[a];
        } else {
            d->stp[a] = 0;
            d->tmax[a] = DBL_MAX;
            d->tdel[a] = DBL_MAX;
        }
    }
}

static int dda_minloc(const dda_t *d)
{
    int a = 0;
    if (d->tmax[1] < d->tmax[a]) a = 1;
    if (d->tmax[2] < d->tmax[a]) a = 2;
    return a;
}

/* advance index along axis a; returns 0 when the ray left the box */
static int dda_step(const orc_ctx *c, dda_t *d, int a)
{
    if (a == 0) { d->ix += d->stp[0]; return !(d->ix < 0 || d->ix >= c->p.nx); }
    if (a == 1) { d->iy += d->stp[1]; return !(d->iy < 0 || d->iy >= c->p.ny); }
    d->iz += d->stp[2];
    return !(d->iz < 0 || d->iz >= c->p.nz);
}

/* raytrace_ion_to_edge_car (src/raytrace_amr.f90:534-592) */
static double edge_car(const orc_ctx *c, tally_t *t, const pkt_t *ph)
{
    dda_t d;
    dda_init(c, ph->x, ph->y, ph->z, ph->kx, ph->ky, ph->kz, &d);
    double wl = ph->wgt * ph->Lpacket, expo = 1.0, tau = 0.0, t_cur = 0.0;
    for (int nstep = 1; nstep <= MAX_WALK_STEPS; ++nstep) {
        int il = 1 + d.ix + c->p.nx * (d.iy + c->p.ny * d.iz);
        int a = dda_minloc(&d);
        double seg = d.tmax[a] - t_cur;
        double kap = opacity(c, ph, il);
        if (kap * seg > 0.0) {
            double eo = expo * exp(-kap * seg);
            score(c, t, ph, il, wl * (expo - eo) / kap);
            expo = eo;
        } else {
            score(c, t, ph, il, wl * expo * seg);
        }
        tau = tau + kap * seg;
        if (tau >= ORC_TAU_HUGE) break;
        t_cur = d.tmax[a];
        d.tmax[a] = d.tmax[a] + d.tdel[a];
        if (!dda_step(c, &d, a)) break;
    }
    return tau;
}

/* raytrace_ion_to_tau_car (src/raytrace_amr.f90:594-676) */
static void tau_car(const orc_ctx *c, tally_t *t, pkt_t *ph, double tau_in)
{
    dda_t d;
    dda_init(c, ph->x, ph->y, ph->z, ph->kx, ph->ky, ph->kz, &d);
    int do_tally = ph->nscatt > 0;
    double wl = ph->wgt * ph->Lpacket, tau = 0.0, t_cur = 0.0;
    for (int nstep = 1;; ++nstep) {
        if (nstep > MAX_WALK_STEPS) { ph->inside = 0; break; }
        int il = 1 + d.ix + c->p.nx * (d.iy + c->p.ny * d.iz);
        int a = dda_minloc(&d);
        double seg = d.tmax[a] - t_cur;
        double kap = opacity(c, ph, il);
        if (tau + seg * kap >= tau_in) {
            double ds = kap > 0.0 ? (tau_in - tau) / kap : seg;
            if (do_tally) score(c, t, ph, il, wl * ds);
            t_cur = t_cur + ds;
            break;
        }
        if (do_tally) score(c, t, ph, il, wl * seg);
        tau = tau + seg * kap;
        t_cur = d.tmax[a];
        d.tmax[a] = d.tmax[a] + d.tdel[a];
        if (!dda_step(c, &d, a)) { ph->inside = 0; break; }
    }
    ph->x = ph->x + t_cur * ph->kx;
    ph->y = ph->y + t_cur * ph->ky;
    ph->z = ph->z + t_cur * ph->kz;
    if (ph->inside) ph->icell = 1 + d.ix + c->p.nx * (d.iy + c->p.ny * d.iz);
}

/* raytrace_ion_to_edge_amr (src/raytrace_amr.f90:422-486) */
static double edge_amr(const orc_ctx *c, tally_t *t, const pkt_t *ph)
{
    double x = ph->x, y = ph->y, z = ph->z, kx = ph->kx, ky = ph->ky, kz = ph->kz;
    int il = ph->icell;
    double wl = ph->wgt * ph->Lpacket, expo = 1.0, tau = 0.0;
    if (il <= 0) {
        il = find_leaf(c, x, y, z);
        if (il <= 0) return 0.0;
    }
    for (int nstep = 1; nstep <= MAX_WALK_STEPS; ++nstep) {
        int ic = leaf_cell(c, il), f;
        double te;
        cell_exit(c, ic, x, y, z, kx, ky, kz, &te, &f);
        double kap = opacity(c, ph, il);
        if (kap * te > 0.0) {
            double eo = expo * exp(-kap * te);
            score(c, t, ph, il, wl * (expo - eo) / kap);
            expo = eo;
        } else {
            score(c, t, ph, il, wl * expo * te);
        }
        tau = tau + kap * te;
        if (tau >= ORC_TAU_HUGE) break;
        x = x + te * kx; y = y + te * ky; z = z + te * kz;
        int nl = next_leaf(c, ic, f, x, y, z);
        if (nl <= 0) break;
        il = nl;
    }
    return tau;
}

/* raytrace_ion_to_tau_amr (src/raytrace_amr.f90:348-420) */
static void tau_amr(const orc_ctx *c, tally_t *t, pkt_t *ph, double tau_in)
{
    int do_tally = ph->nscatt > 0;
    double wl = ph->wgt * ph->Lpacket;
    double x = ph->x, y = ph->y, z = ph->z, kx = ph->kx, ky = ph->ky, kz = ph->kz;
    int il = ph->icell;
    if (il <= 0) {
        il = find_leaf(c, x, y, z);
        if (il <= 0) { ph->inside = 0; return; }
    }
    double tau = 0.0;
    int nstep = 0;
    while (ph->inside) {
        if (++nstep > MAX_WALK_STEPS) { ph->inside = 0; break; }
        int ic = leaf_cell(c, il), f;
        double te;
        cell_exit(c, ic, x, y, z, kx, ky, kz, &te, &f);
        double kap = opacity(c, ph, il);
        if (tau + te * kap >= tau_in) {
            double ds = kap > 0.0 ? (tau_in - tau) / kap : te;
            if (do_tally) score(c, t, ph, il, wl * ds);
            x = x + ds * kx; y = y + ds * ky; z = z + ds * kz;
            break;
        }
        if (do_tally) score(c, t, ph, il, wl * te);
        tau = tau + te * kap;
        x = x + te * kx; y = y + te * ky; z = z + te * kz;
        int nl = next_leaf(c, ic, f, x, y, z);
        if (nl <= 0) { ph->inside = 0; break; }
        il = nl;
    }
    ph->x = x; ph->y = y; ph->z = z;
    ph->icell = il;
}

static int use_car_dda(const orc_ctx *c) { return c->car && !c->p.car_walk_amr; }

/* transport_ion_packet (src/raytrace_amr.f90:249-323) */
static void transport(const orc_ctx *c, tally_t *t, pkt_t *ph)
{
    build_physics(c, ph);
    double tau_edge = use_car_dda(c) ? edge_car(c, t, ph) : edge_amr(c, t, ph);
    ledger(t, ph, tau_edge);
    if (!(c->p.add_dust && c->p.dust_scatter)) return;
    if (tau_edge <= 1.0e-12) return;
    double w1 = 1.0 - exp(-tau_edge);
    ph->wgt = ph->wgt * w1;
    double tau = -log(1.0 - rnd(c, ph) * w1);
    if (use_car_dda(c)) tau_car(c, t, ph, tau); else tau_amr(c, t, ph, tau);
    while (ph->inside) {
        int il = ph->icell;
        if (il <= 0) break;
        double kap = opacity(c, ph, il);
        double a_eff = c->rhokap[il - 1] * ph->dsca / (kap > DBL_MIN ? kap : DBL_MIN);
        ph->wgt = ph->wgt * a_eff;
        if (ph->wgt < WMIN) {
            if (rnd(c, ph) > PSURV) break;
            ph->wgt = ph->wgt / PSURV;
        }
        double gg = ph->dg, cost;
        if (fabs(gg) > 1.0e-3) {
            double q = (1.0 - gg * gg) / (1.0 + gg * (2.0 * rnd(c, ph) - 1.0));
            cost = (1.0 + gg * gg - q * q) / (2.0 * gg);
        } else {
            cost = 2.0 * rnd(c, ph) - 1.0;
        }
        cost = cost < -1.0 ? -1.0 : (cost > 1.0 ? 1.0 : cost);
        double sint = sqrt(1.0 - cost * cost);
        double phi = ORC_TWOPI * rnd(c, ph);
        double kx = ph->kx, ky = ph->ky, kz = ph->kz;
        double norm = sqrt(kx * kx + ky * ky + kz * kz);
        if (norm <= 0.0) break;
        kx = kx / norm; ky = ky / norm; kz = kz / norm;
        double ux, uy, uz;
        if (fabs(kz) < 0.99) { ux = -ky; uy = kx; uz = 0.0; }
        else { ux = 0.0; uy = -kz; uz = ky; }
        norm = sqrt(ux * ux + uy * uy + uz * uz);
        ux = ux / norm; uy = uy / norm; uz = uz / norm;
        double vx = ky * uz - kz * uy, vy = kz * ux - kx * uz, vz = kx * uy - ky * ux;
        double cp = cos(phi), sp = sin(phi);
        ph->kx = sint * (cp * ux + sp * vx) + cost * kx;
        ph->ky = sint * (cp * uy + sp * vy) + cost * ky;
        ph->kz = sint * (cp * uz + sp * vz) + cost * kz;
        norm = sqrt(ph->kx * ph->kx + ph->ky * ph->ky + ph->kz * ph->kz);
        ph->kx = ph->kx / norm; ph->ky = ph->ky / norm; ph->kz = ph->kz / norm;
        ph->nscatt++;
        double r = rnd(c, ph);
        tau = -log(r > DBL_MIN ? r : DBL_MIN);
        if (use_car_dda(c)) tau_car(c, t, ph, tau); else tau_amr(c, t, ph, tau);
    }
}

/* ---- launch (src/ion_band_mod.f90:1880-1935, 2232-2385) ----------------- */
static void ortho_frame(double nx, double ny, double nz, double *t, double *b)
{
    double tx, ty, tz;
    if (fabs(nx) <= fabs(ny) && fabs(nx) <= fabs(nz)) { tx = 0.0; ty = nz; tz = -ny; }
    else if (fabs(ny) <= fabs(nz)) { tx = -nz; ty = 0.0; tz = nx; }
    else { tx = ny; ty = -nx; tz = 0.0; }
    double inv = 1.0 / sqrt(tx * tx + ty * ty + tz * tz);
    tx = tx * inv; ty = ty * inv; tz = tz * inv;
    t[0] = tx; t[1] = ty; t[2] = tz;
    b[0] = ny * tz - nz * ty; b[1] = nz * tx - nx * tz; b[2] = nx * ty - ny * tx;
}

static void emit_point(pkt_t *ph, double xs, double ys, double zs, double umu, double uphi)
{
    ph->x = xs; ph->y = ys; ph->z = zs;
    double cost = 2.0 * umu - 1.0;
    double sint = sqrt(1.0 - cost * cost);
    double phi = ORC_TWOPI * uphi;
    ph->kx = sint * cos(phi);
    ph->ky = sint * sin(phi);
    ph->kz = cost;
}

static void emit_ext_sph(const orc_ctx *c, pkt_t *ph, double umu, double uphi, double ucost,
                         double usaz)
{
    double xc = 0.5 * (c->xmin + c->xmaxb), yc = 0.5 * (c->ymin + c->ymaxb),
           zc = 0.5 * (c->zmin + c->zmaxb);
    double cost = 2.0 * ucost - 1.0, sint = sqrt(1.0 - cost * cost), phi = ORC_TWOPI * usaz;
    double rhx = sint * cos(phi), rhy = sint * sin(phi), rhz = cost;
    ph->x = xc + c->p.rmax * rhx; ph->y = yc + c->p.rmax * rhy; ph->z = zc + c->p.rmax * rhz;
    double t[3], b[3];
    ortho_frame(-rhx, -rhy, -rhz, t, b);
    double mu = sqrt(umu), sinm = sqrt(1.0 - mu * mu);
    phi = ORC_TWOPI * uphi;
    double cphi = cos(phi), sphi = sin(phi);
    ph->kx = -mu * rhx + sinm * (cphi * t[0] + sphi * b[0]);
    ph->ky = -mu * rhy + sinm * (cphi * t[1] + sphi * b[1]);
    ph->kz = -mu * rhz + sinm * (cphi * t[2] + sphi * b[2]);
    double span = c->xrange < c->yrange ? c->xrange : c->yrange;
    if (c->zrange < span) span = c->zrange;
    double nudge = 1.0e-6 * span;
    ph->x = ph->x - nudge * rhx; ph->y = ph->y - nudge * rhy; ph->z = ph->z - nudge * rhz;
}

static void emit_ext_rec(const orc_ctx *c, pkt_t *ph, double umu, double uphi, double uface,
                         double uc1, double uc2)
{
    double span = c->xrange < c->yrange ? c->xrange : c->yrange;
    if (c->zrange < span) span = c->zrange;
    double nudge = 1.0e-6 * span;
    double fa[6], fc[6];
    fa[0] = c->yrange * c->zrange; fa[1] = fa[0];
    fa[2] = c->zrange * c->xrange; fa[3] = fa[2];
    fa[4] = c->xrange * c->yrange; fa[5] = fa[4];
    fc[0] = fa[0];
    for (int i = 1; i < 6; ++i) fc[i] = fc[i - 1] + fa[i];
    double fpick = uface * fc[5];
    int f = 6;
    for (int i = 0; i < 6; ++i) if (fpick <= fc[i]) { f = i + 1; break; }
    double mu = sqrt(umu), sinm = sqrt(1.0 - mu * mu), phi = ORC_TWOPI * uphi;
    double cphi = cos(phi), sphi = sin(phi);
    switch (f) {
    case 1: ph->x = c->xmaxb - nudge; ph->y = c->ymin + c->yrange * uc1;
            ph->z = c->zmin + c->zrange * uc2; ph->kx = -mu; ph->ky = sinm * cphi;
            ph->kz = sinm * sphi; break;
    case 2: ph->x = c->xmin + nudge; ph->y = c->ymin + c->yrange * uc1;
            ph->z = c->zmin + c->zrange * uc2; ph->kx = mu; ph->ky = sinm * cphi;
            ph->kz = sinm * sphi; break;
    case 3: ph->x = c->xmin + c->xrange * uc1; ph->y = c->ymaxb - nudge;
            ph->z = c->zmin + c->zrange * uc2; ph->kx = sinm * cphi; ph->ky = -mu;
            ph->kz = sinm * sphi; break;
    case 4: ph->x = c->xmin + c->xrange * uc1; ph->y = c->ymin + nudge;
            ph->z = c->zmin + c->zrange * uc2; ph->kx = sinm * cphi; ph->ky = mu;
            ph->kz = sinm * sphi; break;
    case 5: ph->x = c->xmin + c->xrange * uc1; ph->y = c->ymin + c->yrange * uc2;
            ph->z = c->zmaxb - nudge; ph->kx = sinm * cphi; ph->ky = sinm * sphi;
            ph->kz = -mu; break;
    default: ph->x = c->xmin + c->xrange * uc1; ph->y = c->ymin + c->yrange * uc2;
            ph->z = c->zmin + nudge; ph->kx = sinm * cphi; ph->ky = sinm * sphi;
            ph->kz = mu; break;
    }
}

/* gen_ion_photon_qmc (src/ion_band_mod.f90:1880-1935); id is 1-based */
static void launch(const orc_ctx *c, int64_t id, pkt_t *ph)
{
    double u[ORC_QMC_MAXDIM];
    orc_qmc_uniforms(&c->qmc, (uint64_t)(id - 1), c->qmc_ndim, 1, u);
    const orc_params *p = &c->p;
    double E;
    if (!c->multi_src && p->nsource == 1) {
        emit_point(ph, p->sx[0], p->sy[0], p->sz[0], u[1], u[2]);
        E = orc_sample_energy(&c->samp[0], u[0]);
    } else if (!c->multi_src) {
        if (p->ext_geometry == ORC_GEOM_EXT_SPH) emit_ext_sph(c, ph, u[2], u[3], u[4], u[5]);
        else emit_ext_rec(c, ph, u[2], u[3], u[4], u[5], u[6]);
        E = orc_sample_energy(&c->samp[0], u[1]);
    } else {
        int ic = c->ncomp - 1;
        for (int i = 0; i < c->ncomp; ++i) if (u[0] <= c->src_cdf[i]) { ic = i; break; }
        int ik = c->comp_kind[ic];
        if (ik >= 1) emit_point(ph, p->sx[ik - 1], p->sy[ik - 1], p->sz[ik - 1], u[2], u[3]);
        else if (p->ext_geometry == ORC_GEOM_EXT_SPH) emit_ext_sph(c, ph, u[2], u[3], u[4], u[5]);
        else emit_ext_rec(c, ph, u[2], u[3], u[4], u[5], u[6]);
        E = orc_sample_energy(&c->samp[ic], u[1]);
    }
    ph->inu = orc_bin_of(c, E);
    ph->wgt = 1.0;
    ph->Lpacket = c->Ltot / (double)p->nphotons;
    ph->E = E;
    ph->nscatt = 0;
    ph->inside = 1;
    ph->icell = find_leaf(c, ph->x, ph->y, ph->z);
    ph->id = id;
    ph->istream = 1;
    ph->nk = 0;
}

/* gen_diffuse_photon_qmc (src/diffuse_mod.f90:221-296) with
 * ground_photon_energy (:341-355); channels 1-4, 9 uniforms on stream 2 */
static void launch_diffuse(const orc_ctx *c, int64_t id, pkt_t *ph)
{
    double u[ORC_QMC_MAXDIM];
    orc_qmc_uniforms(&c->qmc, (uint64_t)(id - 1), 9, 2, u);
    int lo = 1, hi = c->nleaf;
    while (lo < hi) {
        int mid = (lo + hi) / 2;
        if (c->dif_cdf[mid - 1] < u[0]) lo = mid + 1; else hi = mid;
    }
    const int il = lo;
    double cx, cy, cz, half;
    cell_geom(c, leaf_cell(c, il), &cx, &cy, &cz, &half);
    ph->x = cx + (2.0 * u[1] - 1.0) * half;
    ph->y = cy + (2.0 * u[2] - 1.0) * half;
    ph->z = cz + (2.0 * u[3] - 1.0) * half;
    double cost = 2.0 * u[4] - 1.0, sint = sqrt(1.0 - cost * cost), phi = ORC_TWOPI * u[5];
    ph->kx = sint * cos(phi);
    ph->ky = sint * sin(phi);
    ph->kz = cost;
    const double *w = c->dif_ch + (size_t)(il - 1) * 4;
    double uu = u[6] * (((w[0] + w[1]) + w[2]) + w[3]);
    int ch;
    double eph;
    if (uu <= w[0]) { ch = 0; eph = ORC_ETH_HI; }
    else if (uu <= w[0] + w[1]) { ch = 1; eph = ORC_ETH_HEI; }
    else if (uu <= (w[0] + w[1]) + w[2]) { ch = 2; eph = ORC_ETH_HEII; }
    else { ch = 3; eph = 0.0; }
    const double T = c->Te[il - 1];
    if (ch <= 2) {
        if (c->p.diffuse_model == 2) {
            eph = orc_milne_sample_energy(c->milne, ch, T, u[7]);
        } else if (c->p.diffuse_model == 1) {
            double kT = ORC_KBOLTZ * T / ORC_EV2ERG;
            eph = eph + kT * (-log(u[7] > DBL_MIN ? u[7] : DBL_MIN));
        }
        if (eph > c->p.eion_max * 0.999) eph = c->p.eion_max * 0.999;
    } else {
        /* the He I cascade: decay branch by energy luminosity, fixed line
         * energies, the two-photon energy from the tabulated shape */
        double be3s, be1p, be1s;
        orc_hei_branch_energies(c->hei, T, c->xHI[il - 1], c->xHeI[il - 1], &be3s, &be1p, &be1s);
        const double ub = u[7] * ((be3s + be1p) + be1s);
        if (ub <= be3s) eph = 19.82;
        else if (ub <= be3s + be1p) eph = 21.22;
        else eph = orc_hei2ph_sample_energy(c->hei, u[8]);
    }
    ph->inu = orc_bin_of(c, eph);
    ph->wgt = 1.0;
    ph->Lpacket = c->Ltot / (double)c->p.nphotons;
    ph->E = eph;
    ph->nscatt = 0;
    ph->inside = 1;
    ph->icell = il;
    ph->id = id;
    ph->istream = 2;
    ph->nk = 0;
}

int orc_launch_diffuse(orc_ctx *c, int64_t id, double *out)
{
    pkt_t ph;
    memset(&ph, 0, sizeof(ph));
    launch_diffuse(c, id, &ph);
    out[0] = ph.x; out[1] = ph.y; out[2] = ph.z;
    out[3] = ph.kx; out[4] = ph.ky; out[5] = ph.kz;
    out[6] = ph.E; out[7] = (double)ph.inu; out[8] = (double)ph.icell;
    return 0;
}

int orc_launch(orc_ctx *c, int64_t id, double *out)
{
    pkt_t ph;
    memset(&ph, 0, sizeof(ph));
    launch(c, id, &ph);
    out[0] = ph.x; out[1] = ph.y; out[2] = ph.z;
    out[3] = ph.kx; out[4] = ph.ky; out[5] = ph.kz;
    out[6] = ph.E; out[7] = (double)ph.inu; out[8] = (double)ph.icell;
    return 0;
}

/* ---- pass driver --------------------------------------------------------- */
typedef struct {
    const orc_ctx *c;
    tally_t t;
    int64_t first, last;      /* stellar ids [first, last) */
    int64_t dfirst, dlast;    /* diffuse ids [dfirst, dlast) */
} work_t;

static void *worker(void *arg)
{
    work_t *w = (work_t *)arg;
    for (int64_t id = w->first; id < w->last; ++id) {
        pkt_t ph;
        launch(w->c, id, &ph);
        transport(w->c, &w->t, &ph);
    }
    for (int64_t id = w->dfirst; id < w->dlast; ++id) {
        pkt_t ph;
        launch_diffuse(w->c, id, &ph);
        transport(w->c, &w->t, &ph);
    }
    return NULL;
}

/* ---- dust thermal-infrared band ------------------------------------------ */

/* launch_and_transport's launch (src/dust_emis_transport_mod.f90:489-544):
 * uniform position in leaf il (0-based), isotropic direction, wavelength bin
 * from the leaf's cumulative by binary search, then continuously inside the
 * bin; draws 0-5 of the packet's key on stream 3.
 * out = x, y, z, kx, ky, kz, lambda [um] */
void orc_ir_launch_state(const orc_ctx *c, int il, const double *pdf, const double *cdf,
                         uint64_t key, double *out)
{
    const orc_dustir *ir = c->ir;
    const uint32_t seed = (uint32_t)c->p.qmc_seed;
    double cx, cy, cz, half;
    cell_geom(c, leaf_cell(c, il + 1), &cx, &cy, &cz, &half);
    out[0] = cx + (2.0 * orc_rng_uniform(seed, 3, key, 0) - 1.0) * half;
    out[1] = cy + (2.0 * orc_rng_uniform(seed, 3, key, 1) - 1.0) * half;
    out[2] = cz + (2.0 * orc_rng_uniform(seed, 3, key, 2) - 1.0) * half;
    double cost = 2.0 * orc_rng_uniform(seed, 3, key, 3) - 1.0;
    double sint = sqrt(1.0 - cost * cost);
    double phi = ORC_TWOPI * orc_rng_uniform(seed, 3, key, 4);
    out[3] = sint * cos(phi);
    out[4] = sint * sin(phi);
    out[5] = cost;
    int n = ir->nl;
    double u = orc_rng_uniform(seed, 3, key, 5) * cdf[n - 1];
    int lo = 1, hi = n - 1;                  /* 1-based bins */
    while (lo < hi) {
        int mid = (lo + hi) / 2;
        if (u <= cdf[mid]) hi = mid;         /* cdf(mid+1), 1-based */
        else lo = mid + 1;
    }
    int k = lo - 1;                          /* 0-based lower node */
    double ulo = cdf[k], ubin = 0.0;
    if (cdf[k + 1] > ulo) ubin = (u - ulo) / (cdf[k + 1] - ulo);
    out[6] = orc_sample_linear_bin(ir->lam[k], ir->lam[k + 1], pdf[k], pdf[k + 1], ubin);
}

typedef struct {
    orc_ctx *c;
    tally_t t;
    int pass;
    int64_t k0, k1;                          /* packet indices, inclusive */
    double fdisc;
} irwork_t;

/* the leaves' packets of one pass (src/dust_emis_transport_mod.f90:267-284):
 * leaf of packet k = the first leaf with khi >= k; its spectrum is rebuilt
 * when the leaf changes; leaves whose spectrum is empty drop their packets */
static void *ir_worker(void *arg)
{
    irwork_t *w = (irwork_t *)arg;
    const orc_ctx *c = w->c;
    const orc_dustir *ir = c->ir;
    if (w->k1 < w->k0) return NULL;
    double *pdf = (double *)malloc(sizeof(double) * ir->nl);
    double *cdf = (double *)malloc(sizeof(double) * ir->nl);
    int lo = 0, hi = c->nleaf - 1;
    while (lo < hi) {
        int mid = (lo + hi) / 2;
        if (ir->khi[mid] >= w->k0) hi = mid; else lo = mid + 1;
    }
    int il = lo, cur = -1;
    double esum = 0.0;
    for (int64_t k = w->k0; k <= w->k1; ++k) {
        while (ir->khi[il] < k) ++il;
        if (il != cur) {
            double fd;
            esum = orc_ir_leaf_spectrum(c, il, pdf, cdf, &fd);
            if (esum > 0.0 && fd > w->fdisc) w->fdisc = fd;
            cur = il;
        }
        if (!(esum > 0.0)) continue;
        uint64_t key = ((uint64_t)w->pass << 40) | (uint64_t)k;
        double o[7];
        orc_ir_launch_state(c, il, pdf, cdf, key, o);
        pkt_t ph;
        memset(&ph, 0, sizeof(ph));
        ph.x = o[0]; ph.y = o[1]; ph.z = o[2];
        ph.kx = o[3]; ph.ky = o[4]; ph.kz = o[5];
        ph.lam = o[6];
        ph.E = 1.23984 / o[6];
        ph.inu = 1;
        ph.icell = il + 1;
        ph.nscatt = 0;
        ph.wgt = 1.0;
        ph.inside = 1;
        ph.Lpacket = ir->Lpacket;
        ph.id = (int64_t)key;
        ph.istream = 3;
        ph.nk = 6;
        transport(c, &w->t, &ph);
    }
    free(pdf); free(cdf);
    return NULL;
}

/* one infrared pass: packets 1..N of the counts in ir->khi, launched and
 * carried by the shared transport with the band's estimators
 * (dust_ir_band_reset / _reduce, src/dust_ir_band_mod.f90:87-95, 153-174) */
int orc_ir_transport(orc_ctx *c, int pass, int64_t N, int nthreads)
{
    orc_dustir *ir = c->ir;
    int nl = c->nleaf, nb = ir->nl;
    if (nthreads < 1) nthreads = 1;
    irwork_t *w = (irwork_t *)calloc(nthreads, sizeof(irwork_t));
    pthread_t *th = (pthread_t *)calloc(nthreads, sizeof(pthread_t));
    memset(ir->irabs, 0, sizeof(double) * nl);
    memset(ir->lesc, 0, sizeof(double) * nb);
    for (int i = 0; i < nthreads; ++i) {
        w[i].c = c;
        w[i].pass = pass;
        w[i].k0 = 1 + N * i / nthreads;
        w[i].k1 = N * (i + 1) / nthreads;
        w[i].t.ir = 1;
        w[i].t.nl = nb;
        w[i].t.edge = ir->edge;
        if (i == 0) {
            w[i].t.irabs = ir->irabs;
            w[i].t.lesc = ir->lesc;
        } else {
            w[i].t.irabs = (double *)calloc((size_t)nl, sizeof(double));
            w[i].t.lesc = (double *)calloc((size_t)nb, sizeof(double));
        }
    }
    for (int i = 1; i < nthreads; ++i) pthread_create(&th[i], NULL, ir_worker, &w[i]);
    ir_worker(&w[0]);
    for (int i = 1; i < nthreads; ++i) pthread_join(th[i], NULL);
    long double ee = w[0].t.e_emit, ea = w[0].t.e_abs, es = w[0].t.e_esc;
    int64_t sd = w[0].t.seg_direct, ss = w[0].t.seg_scattered;
    double fd = w[0].fdisc;
    for (int i = 1; i < nthreads; ++i) {
        for (int k = 0; k < nl; ++k) ir->irabs[k] += w[i].t.irabs[k];
        for (int k = 0; k < nb; ++k) ir->lesc[k] += w[i].t.lesc[k];
        ee += w[i].t.e_emit; ea += w[i].t.e_abs; es += w[i].t.e_esc;
        sd += w[i].t.seg_direct; ss += w[i].t.seg_scattered;
        if (w[i].fdisc > fd) fd = w[i].fdisc;
        free(w[i].t.irabs); free(w[i].t.lesc);
    }
    ir->emitted = (double)ee; ir->extinct = (double)ea; ir->escaped = (double)es;
    ir->seg_direct = sd; ir->seg_scattered = ss;
    if (fd > ir->frac_discarded) ir->frac_discarded = fd;
    free(w); free(th);
    return 0;
}

/* ionizing_field_and_rates (src/ionizing_field_mod.f90:90-164): zero
 * tallies, transport stellar ids [first,last) and, with the diffuse field on,
 * all diffuse packets of the emission built from the current state, reduce */
int orc_transport_pass(orc_ctx *c, int64_t first, int64_t last, int nthreads)
{
    int nl = c->nleaf, nnu = c->nnu_band;
    if (nthreads < 1) nthreads = 1;
    int64_t nd = 0;
    if (c->p.diffuse_field) {
        if (!c->milne && orc_diffuse_setup(c)) return -1;
        orc_diffuse_build(c);
        nd = c->diffuse_nphot;
    }
    work_t *w = (work_t *)calloc(nthreads, sizeof(work_t));
    pthread_t *th = (pthread_t *)calloc(nthreads, sizeof(pthread_t));
    int64_t n = last - first;
    for (int i = 0; i < nthreads; ++i) {
        w[i].c = c;
        w[i].first = first + n * i / nthreads;
        w[i].last = first + n * (i + 1) / nthreads;
        w[i].dfirst = 1 + nd * i / nthreads;
        w[i].dlast = 1 + nd * (i + 1) / nthreads;
        if (i == 0) {
            w[i].t.hhe = c->hhe; w[i].t.jt = c->jt; w[i].t.dust = c->dust_abs;
            w[i].t.mraw = c->mraw;
            memset(c->hhe, 0, sizeof(double) * 6 * (size_t)nl);
            memset(c->jt, 0, sizeof(double) * (size_t)nnu * nl);
            memset(c->dust_abs, 0, sizeof(double) * nl);
            if (c->use_metals) memset(c->mraw, 0, sizeof(double) * (size_t)nl * c->ntr * 2);
        } else {
            w[i].t.hhe = (double *)calloc((size_t)nl * 6, sizeof(double));
            w[i].t.jt = (double *)calloc((size_t)nl * nnu, sizeof(double));
            w[i].t.dust = (double *)calloc(nl, sizeof(double));
            w[i].t.mraw = c->use_metals ? (double *)calloc((size_t)nl * c->ntr * 2, sizeof(double))
                                        : NULL;
        }
    }
    for (int i = 1; i < nthreads; ++i) pthread_create(&th[i], NULL, worker, &w[i]);
    worker(&w[0]);
    for (int i = 1; i < nthreads; ++i) pthread_join(th[i], NULL);
    long double ee = w[0].t.e_emit, ea = w[0].t.e_abs, es = w[0].t.e_esc;
    int64_t sd = w[0].t.seg_direct, ss = w[0].t.seg_scattered;
    for (int i = 1; i < nthreads; ++i) {
        for (size_t k = 0; k < (size_t)nl * 6; ++k) c->hhe[k] += w[i].t.hhe[k];
        for (size_t k = 0; k < (size_t)nl * nnu; ++k) c->jt[k] += w[i].t.jt[k];
        for (int k = 0; k < nl; ++k) c->dust_abs[k] += w[i].t.dust[k];
        if (c->use_metals)
            for (size_t k = 0; k < (size_t)nl * c->ntr * 2; ++k) c->mraw[k] += w[i].t.mraw[k];
        ee += w[i].t.e_emit; ea += w[i].t.e_abs; es += w[i].t.e_esc;
        sd += w[i].t.seg_direct; ss += w[i].t.seg_scattered;
        free(w[i].t.hhe); free(w[i].t.jt); free(w[i].t.dust); free(w[i].t.mraw);
    }
    c->e_emit = (double)ee; c->e_abs = (double)ea; c->e_esc = (double)es;
    c->seg_direct = sd; c->seg_scattered = ss;
    free(w); free(th);
    return 0;
}

/* gas_rates_compute -> ion_score_apply_hhe (src/gas_rates_mod.f90:82-153,
 * src/ion_score_mod.f90:298-317) */
void orc_rates_compute(orc_ctx *c)
{
    double d2 = c->p.distance2cm * c->p.distance2cm;
    for (int il = 1; il <= c->nleaf; ++il) {
        double s = 2.0 * leaf_half(c, il);
        double fac = 1.0 / ((s * s) * s * d2);
        const double *h = c->hhe + (size_t)(il - 1) * 6;
        c->gHI[il - 1] = h[0] * fac; c->gHeI[il - 1] = h[1] * fac; c->gHeII[il - 1] = h[2] * fac;
        c->hHI[il - 1] = h[3] * fac; c->hHeI[il - 1] = h[4] * fac; c->hHeII[il - 1] = h[5] * fac;
    }
    orc_species_gamma_compute(c);
}

/* cells of one ray's direct walk to the box edge (geometry only) */
int orc_trace_ray(orc_ctx *c, double x, double y, double z, double kx, double ky, double kz,
                  int *cells, double *segs, int maxn)
{
    int n = 0;
    if (use_car_dda(c)) {
        dda_t d;
        dda_init(c, x, y, z, kx, ky, kz, &d);
        double t_cur = 0.0;
        while (n < maxn) {
            int il = 1 + d.ix + c->p.nx * (d.iy + c->p.ny * d.iz);
            int a = dda_minloc(&d);
            cells[n] = il; segs[n] = d.tmax[a] - t_cur; ++n;
            t_cur = d.tmax[a];
            d.tmax[a] = d.tmax[a] + d.tdel[a];
            if (!dda_step(c, &d, a)) break;
        }
        return n;
    }
    int il = find_leaf(c, x, y, z);
    if (il <= 0) return 0;
    while (n < maxn) {
        int ic = leaf_cell(c, il), f;
        double te;
        cell_exit(c, ic, x, y, z, kx, ky, kz, &te, &f);
        cells[n] = il; segs[n] = te; ++n;
        x = x + te * kx; y = y + te * ky; z = z + te * kz;
        int nl = next_leaf(c, ic, f, x, y, z);
        if (nl <= 0) break;
        il = nl;
    }
    return n;
}
