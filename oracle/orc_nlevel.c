/*
 * Oracle: n-level atoms, the (T, n_e) line-cooling suppression tables of
 * cooling_model = 'local_ne', and the per-ion line emissivities
 * (TEST INFRASTRUCTURE; no product path links it).
 *
 * Restates
 *   src/nlevel_mod.f90           ups_eval :128-160, emitting_level :174-187,
 *                                nlevel_emissivities :193-253,
 *                                nlevel_cooling_prepT / _solve :277-361,
 *                                nlevel_line_ident :381-393
 *   src/nlevel_cooling_mod.f90   nlevel_cooling_init :86-97 (grid),
 *                                nlevel_cooling_add :105-152 (table build),
 *                                nlevel_cooling_apply :157-171,
 *                                cooling_coeff :198-223, col_cooling :232-245,
 *                                low_density_cooling :251-262
 *   src/sh95_mod.f90             sh95_emis :146-176 (recombination lines)
 * in the reference's arithmetic order (the rate matrix and Gaussian
 * elimination loop orders, the interpolation forms), so the compiled
 * reference modules (oracle/_ref) match it bit for bit.
 *
 * Pinned by: the compiled nlevel_mod / nlevel_cooling_mod (oracle/_ref) on
 * every tabulated ion; a two-level atom's closed-form populations; the
 * low-density limit of the cooling (every excitation out of the ground level
 * radiated); detailed balance (Boltzmann populations) at high density.
 */
#include <float.h>
#include <math.h>
#include <string.h>
#include "mc_oracle.h"

#define KB_CGS       1.380649e-16
#define COLL_PREF    8.629e-6
#define HC_ERG_CM    1.9864458571489287e-16

static double kb_over_ry(void) { return KB_CGS / 2.1798723611035e-11; }

/* ups_eval (src/nlevel_mod.f90:128-160) */
double orc_ups_eval(const orc_nlatom *a, int k, double T)
{
    const double et = kb_over_ry() * T / a->dery[k];
    double st;
    if (a->ttype[k] == 1 || a->ttype[k] == 4)
        st = 1.0 - log(a->cups[k]) / log(et + a->cups[k]);
    else
        st = et / (et + a->cups[k]);
    st = fmin(fmax(st, a->st_lo[k]), a->st_hi[k]);
    double u;
    if (a->st_hi[k] > a->st_lo[k])
        u = 2.0 * (st - a->st_lo[k]) / (a->st_hi[k] - a->st_lo[k]) - 1.0;
    else
        u = 0.0;
    double b0 = 0.0, b1 = 0.0, b2;
    for (int j = a->nc[k]; j >= 2; --j) {
        b2 = b1;
        b1 = b0;
        b0 = 2.0 * u * b1 - b2 + a->coef[k][j - 1];
    }
    double y = u * b0 - b1 + a->coef[k][0];
    if (a->logf[k] == 1) y = pow(10.0, y);
    double ups;
    switch (a->ttype[k]) {
    case 1: ups = y * log(et + exp(1.0)); break;
    case 2: ups = y; break;
    case 3: ups = y / (et + 1.0); break;
    case 4: ups = y * log(et + a->cups[k]); break;
    case 5: ups = y / et; break;
    default: ups = pow(10.0, y); break;
    }
    return fmax(ups, 0.0);
}

/* emitting_level (src/nlevel_mod.f90:174-187): 0-based lo/hi */
void orc_emitting_level(const orc_nlatom *a, int i1, int i2, int *lo, int *hi, double *de)
{
    if (a->e_cm1[i2 - 1] >= a->e_cm1[i1 - 1]) { *lo = i1 - 1; *hi = i2 - 1; }
    else { *lo = i2 - 1; *hi = i1 - 1; }
    *de = (a->e_cm1[*hi] - a->e_cm1[*lo]) * HC_ERG_CM;
}

/* statistical equilibrium of the rate matrix R[i][j] (rate j -> i, row-major
 * with stride ORC_NL_MAXLEV): closure row, Gaussian elimination with partial
 * pivoting, back substitution (src/nlevel_mod.f90:218-246, 311-347) */
static void solve_populations(int n, const double *R, double *pop)
{
    enum { L = ORC_NL_MAXLEV };
    static const int W = L + 1;
    double M[ORC_NL_MAXLEV * (ORC_NL_MAXLEV + 1)];
    for (int i = 0; i < n; ++i)
        for (int j = 0; j <= n; ++j) M[i * W + j] = 0.0;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            if (i != j) {
                M[i * W + j] = R[i * L + j];
                M[i * W + i] = M[i * W + i] - R[j * L + i];
            }
    for (int j = 0; j < n; ++j) M[j] = 1.0;
    M[n] = 1.0;
    for (int i = 0; i < n - 1; ++i) {
        int ip = i;
        for (int j = i + 1; j < n; ++j)
            if (fabs(M[j * W + i]) > fabs(M[ip * W + i])) ip = j;
        if (ip != i)
            for (int j = i; j <= n; ++j) {
                const double f = M[i * W + j];
                M[i * W + j] = M[ip * W + j];
                M[ip * W + j] = f;
            }
        for (int j = i + 1; j < n; ++j) {
            const double fac = M[j * W + i] / M[i * W + i];
            for (int k = i; k <= n; ++k) M[j * W + k] = M[j * W + k] - fac * M[i * W + k];
        }
    }
    for (int i = n - 1; i >= 0; --i) {
        double p = M[i * W + n];
        for (int j = i + 1; j < n; ++j) p = p - M[i * W + j] * pop[j];
        pop[i] = p / M[i * W + i];
    }
}

/* nlevel_cooling_prepT (src/nlevel_mod.f90:277-300): collisional rates
 * without the n_e factor and the radiative A-values, [i][j] = rate j -> i */
void orc_nlevel_prepT(const orc_nlatom *a, double T, double *Rcoll, double *Rrad)
{
    enum { L = ORC_NL_MAXLEV };
    for (int i = 0; i < L * L; ++i) { Rcoll[i] = 0.0; Rrad[i] = 0.0; }
    for (int k = 0; k < a->nups; ++k) {
        int l, u;
        double de;
        orc_emitting_level(a, a->ul_l[k], a->ul_u[k], &l, &u, &de);
        const double ups = orc_ups_eval(a, k, T);
        const double q_ul = COLL_PREF / (a->g[u] * sqrt(T)) * ups;
        const double q_lu = (a->g[u] / a->g[l]) * q_ul * exp(-de / (KB_CGS * T));
        Rcoll[u * L + l] = Rcoll[u * L + l] + q_lu;
        Rcoll[l * L + u] = Rcoll[l * L + u] + q_ul;
    }
    for (int k = 0; k < a->nrad; ++k) {
        int l, u;
        double de;
        orc_emitting_level(a, a->rl[k], a->ru[k], &l, &u, &de);
        Rrad[l * L + u] = Rrad[l * L + u] + a->A[k];
    }
}

/* nlevel_cooling_solve (src/nlevel_mod.f90:303-361): Lambda_SE / n_e */
double orc_nlevel_cooling_solve(const orc_nlatom *a, const double *Rcoll, const double *Rrad,
                                double ne)
{
    enum { L = ORC_NL_MAXLEV };
    double R[L * L], pop[L];
    const int n = a->nlev;
    for (int i = 0; i < L * L; ++i) R[i] = ne * Rcoll[i] + Rrad[i];
    solve_populations(n, R, pop);
    double lam = 0.0;
    for (int k = 0; k < a->nrad; ++k) {
        int lo, hi;
        double de;
        orc_emitting_level(a, a->rl[k], a->ru[k], &lo, &hi, &de);
        lam = lam + pop[hi] * a->A[k] * de;
    }
    return lam / ne;
}

double orc_nlevel_cooling_at(const orc_nlatom *a, double T, double ne)
{
    static __thread double Rc[ORC_NL_MAXLEV * ORC_NL_MAXLEV], Rr[ORC_NL_MAXLEV * ORC_NL_MAXLEV];
    orc_nlevel_prepT(a, T, Rc, Rr);
    return orc_nlevel_cooling_solve(a, Rc, Rr, ne);
}

/* nlevel_emissivities (src/nlevel_mod.f90:193-253): emis[k] = n_u A dE
 * [erg/s per ion] of every radiative row k */
void orc_nlevel_emissivities(const orc_nlatom *a, double T, double ne, double *emis)
{
    enum { L = ORC_NL_MAXLEV };
    double R[L * L], pop[L];
    for (int i = 0; i < L * L; ++i) R[i] = 0.0;
    for (int k = 0; k < a->nups; ++k) {
        int l, u;
        double de;
        orc_emitting_level(a, a->ul_l[k], a->ul_u[k], &l, &u, &de);
        const double ups = orc_ups_eval(a, k, T);
        const double q_ul = COLL_PREF / (a->g[u] * sqrt(T)) * ups;
        const double q_lu = (a->g[u] / a->g[l]) * q_ul * exp(-de / (KB_CGS * T));
        R[u * L + l] = R[u * L + l] + ne * q_lu;
        R[l * L + u] = R[l * L + u] + ne * q_ul;
    }
    for (int k = 0; k < a->nrad; ++k) {
        int l, u;
        double de;
        orc_emitting_level(a, a->rl[k], a->ru[k], &l, &u, &de);
        R[l * L + u] = R[l * L + u] + a->A[k];
    }
    solve_populations(a->nlev, R, pop);
    for (int k = 0; k < a->nrad; ++k) {
        int l, u;
        double de;
        orc_emitting_level(a, a->rl[k], a->ru[k], &l, &u, &de);
        emis[k] = pop[u] * a->A[k] * de;
    }
}

/* nlevel_line_ident (src/nlevel_mod.f90:381-393): vacuum wavelength [A] */
double orc_nlevel_line_ident(const orc_nlatom *a, int k)
{
    const double dcm = fabs(a->e_cm1[a->ru[k] - 1] - a->e_cm1[a->rl[k] - 1]);
    return dcm > 0.0 ? 1.0e8 / dcm : 0.0;
}

/* ---- (T, n_e) cooling tables (src/nlevel_cooling_mod.f90) ---------------- */
static double gT(int i) { return ORC_NLC_TLO + ORC_NLC_DT * (double)i; }      /* 0-based */
static double gNE(int j) { return ORC_NLC_NELO + ORC_NLC_DNE * (double)j; }

/* nlevel_cooling_add's table (src/nlevel_cooling_mod.f90:130-141) */
void orc_nlcool_build(const orc_nlatom *a, orc_nlcool *t)
{
    enum { L = ORC_NL_MAXLEV, NT = ORC_NLC_NT, NNE = ORC_NLC_NNE };
    double Rc[L * L], Rr[L * L];
    for (int i = 0; i < NT; ++i) {
        const double Tk = pow(10.0, gT(i));
        orc_nlevel_prepT(a, Tk, Rc, Rr);
        for (int j = 0; j < NNE; ++j) {
            const double nek = pow(10.0, gNE(j));
            t->Ssup[(size_t)j * NT + i] = orc_nlevel_cooling_solve(a, Rc, Rr, nek);
        }
    }
    for (int i = 0; i < NT; ++i) t->lamref[i] = t->Ssup[i];
    for (int j = 0; j < NNE; ++j)
        for (int i = 0; i < NT; ++i)
            t->Ssup[(size_t)j * NT + i] = t->Ssup[(size_t)j * NT + i] / fmax(t->lamref[i], DBL_MIN);
    t->ready = 1;
}

/* col_cooling (:232-245) at row it (0-based) and column j */
static double col_cooling(const orc_nlcool *t, int it, int j, double w1, double wt)
{
    const double Li = t->lamref[it] * t->Ssup[(size_t)j * ORC_NLC_NT + it];
    const double Lip = t->lamref[it + 1] * t->Ssup[(size_t)j * ORC_NLC_NT + it + 1];
    if (Li > 0.0 && Lip > 0.0) return exp(log(Li) + w1 * (log(Lip) - log(Li)));
    return Li * (1.0 - wt) + Lip * wt;
}

static void t_weights(double T, int *it, double *w1, double *wt)
{
    const double lt = fmin(fmax(log10(T), gT(0)), gT(ORC_NLC_NT - 1));
    int i = (int)((lt - gT(0)) / ORC_NLC_DT) + 1;
    if (i > ORC_NLC_NT - 1) i = ORC_NLC_NT - 1;
    *it = i - 1;
    *w1 = (1.0 / pow(10.0, lt) - 1.0 / pow(10.0, gT(*it))) /
          (1.0 / pow(10.0, gT(*it + 1)) - 1.0 / pow(10.0, gT(*it)));
    *wt = fmin(fmax((lt - gT(*it)) / (gT(*it + 1) - gT(*it)), 0.0), 1.0);
}

/* cooling_coeff (:198-223) */
double orc_nlcool_coeff(const orc_nlcool *t, double T, double ne)
{
    int it;
    double w1, wt;
    t_weights(T, &it, &w1, &wt);
    const double lne = log10(ne);
    if (lne <= gNE(0)) return col_cooling(t, it, 0, w1, wt);
    if (lne >= gNE(ORC_NLC_NNE - 1)) return col_cooling(t, it, ORC_NLC_NNE - 1, w1, wt);
    int je = (int)((lne - gNE(0)) / ORC_NLC_DNE) + 1;
    if (je > ORC_NLC_NNE - 1) je = ORC_NLC_NNE - 1;
    const int j0 = je - 1;
    const double wn = (lne - gNE(j0)) / (gNE(j0 + 1) - gNE(j0));
    const double L0 = col_cooling(t, it, j0, w1, wt), L1 = col_cooling(t, it, j0 + 1, w1, wt);
    if (L0 > 0.0 && L1 > 0.0) return exp(log(L0) + wn * (log(L1) - log(L0)));
    return L0 * (1.0 - wn) + L1 * wn;
}

/* low_density_cooling (:251-262) */
double orc_nlcool_low_density(const orc_nlcool *t, double T)
{
    int it;
    double w1, wt;
    t_weights(T, &it, &w1, &wt);
    return col_cooling(t, it, 0, w1, wt);
}

/* nlevel_cooling_apply (:157-171); t == NULL (no table) returns lam_fit */
double orc_nlcool_apply(const orc_nlcool *t, double lam_fit, double T, double ne)
{
    if (!t || !t->ready) return lam_fit;
    const double lamref = orc_nlcool_low_density(t, T);
    const double S = orc_nlcool_coeff(t, T, ne) / fmax(lamref, DBL_MIN);
    const double fcov = fmin(1.0, lamref / fmax(lam_fit, DBL_MIN));
    return lam_fit * (fcov * S + (1.0 - fcov));
}

/* ---- recombination-line tables (src/sh95_mod.f90:146-176) ---------------- */
double orc_sh95_emis(const orc_sh95 *t, int k, double T, double ne)
{
    const int nT = t->nT, nD = t->nD;
    const double lt = fmin(fmax(T, t->tgrid[0]), t->tgrid[nT - 1]);
    const double ld = fmin(fmax(ne, t->dgrid[0]), t->dgrid[nD - 1]);
    int iT = 0, iD = 0;
    for (iT = 0; iT < nT - 1; ++iT)
        if (lt < t->tgrid[iT + 1]) break;
    if (iT > nT - 2) iT = nT - 2;
    for (iD = 0; iD < nD - 1; ++iD)
        if (ld < t->dgrid[iD + 1]) break;
    if (iD > nD - 2) iD = nD - 2;
    const double wT = log(lt / t->tgrid[iT]) / log(t->tgrid[iT + 1] / t->tgrid[iT]);
    const double wD = log(ld / t->dgrid[iD]) / log(t->dgrid[iD + 1] / t->dgrid[iD]);
    const double *e = t->etab + (size_t)k * nT * nD;    /* [line][T][ne] */
#define E(a, b) e[(size_t)(a) * nD + (b)]
    return exp(log(E(iT, iD)) * (1.0 - wT) * (1.0 - wD) + log(E(iT + 1, iD)) * wT * (1.0 - wD) +
               log(E(iT, iD + 1)) * (1.0 - wT) * wD + log(E(iT + 1, iD + 1)) * wT * wD);
#undef E
}

void orc_set_stage_table(orc_ctx *c, int ie, int stage, const orc_nlcool *t)
{
    if (ie >= 0 && ie < ORC_MAX_EL && stage >= 1 && stage <= ORC_MAX_ST) c->st_tab[ie][stage - 1] = t;
}

void orc_set_hi_table(orc_ctx *c, const orc_nlcool *t) { c->cool.hi_tab = t; }

/* ---- lines_write (src/lines_mod.f90:37-352) ------------------------------ *
 * Line luminosities on the current state, in the reference's table order:
 * the rows of the recombination tables tab[0] (H I, n_e n_p), tab[1]
 * (He II, n_e n_HeIII) and tab[2] (He I, n_e n_HeII), any of them NULL,
 * then for every registry element and stage with an atom
 * (atom[ie * ORC_MAX_ST + stage - 1], NULL = none) its radiative rows.
 * Leaves are split over nthreads threads with private sums, added in
 * thread order.  Returns the number of lines written to L. */
#include <pthread.h>
#include <stdlib.h>

typedef struct {
    const orc_ctx *c;
    const orc_sh95 *const *tab;
    const orc_nlatom *const *atom;
    int lo, hi, nline;
    double *L;
} lines_work;

static double leaf_vol_cm(const orc_ctx *c, int il)
{
    double h;
    if (c->car) h = 0.5 * c->dcell;
    else h = c->ch[c->icell_of_leaf[il] - 1];
    const double w = 2.0 * h * c->p.distance2cm;
    return (w * w) * w;
}

static void *lines_worker(void *arg)
{
    lines_work *w = (lines_work *)arg;
    const orc_ctx *c = w->c;
    double emis[ORC_NL_MAXTR], frac[ORC_MAX_ST];
    for (int il = w->lo; il < w->hi; ++il) {
        const double nH = c->nH[il];
        if (nH <= 0.0) continue;
        const double T = c->Te[il], ne = c->ne[il];
        const double vol = leaf_vol_cm(c, il);
        int out = 0;
        for (int t = 0; t < 3; ++t) {
            const orc_sh95 *tb = w->tab[t];
            if (!tb) continue;
            double x;
            int on = 1;
            if (t == 0) x = 1.0 - c->xHI[il];
            else if (t == 1) {
                x = 1.0 - c->xHeI[il] - c->xHeII[il];
                if (x < 0.0) x = 0.0;
                on = x > 1.0e-30;
            } else {
                x = c->xHeII[il];
                on = x > 1.0e-30;
            }
            if (on) {
                const double pre = t == 0 ? ne * nH * x : ne * nH * c->p.He_abund * x;
                for (int k = 0; k < tb->nlin; ++k) {
                    const double e = pre * orc_sh95_emis(tb, k, T, ne);
                    w->L[out + k] = w->L[out + k] + e * vol;
                }
            }
            out += tb->nlin;
        }
        const double nHI = nH * c->xHI[il], nHII = nH - nHI;
        for (int ie = 0; ie < c->nel; ++ie) {
            const orc_element *el = &c->el[ie];
            int have = 0;
            for (int st = 1; st <= el->nstage; ++st) have |= w->atom[ie * ORC_MAX_ST + st - 1] != NULL;
            if (!have) continue;
            orc_species_fractions(c, ie, il, T, ne, nHI, nHII, frac);
            for (int st = 1; st <= el->nstage; ++st) {
                const orc_nlatom *a = w->atom[ie * ORC_MAX_ST + st - 1];
                if (!a) continue;
                const double nion = el->abund * nH * frac[st - 1];
                if (nion > 1.0e-30) {
                    orc_nlevel_emissivities(a, T, ne, emis);
                    for (int k = 0; k < a->nrad; ++k)
                        w->L[out + k] = w->L[out + k] + emis[k] * nion * vol;
                }
                out += a->nrad;
            }
        }
    }
    return NULL;
}

int orc_lines(const orc_ctx *c, const orc_sh95 *const *tab, const orc_nlatom *const *atom,
              double *L, int nthreads)
{
    int nline = 0;
    for (int t = 0; t < 3; ++t)
        if (tab[t]) nline += tab[t]->nlin;
    for (int ie = 0; ie < c->nel; ++ie)
        for (int st = 1; st <= c->el[ie].nstage; ++st)
            if (atom[ie * ORC_MAX_ST + st - 1]) nline += atom[ie * ORC_MAX_ST + st - 1]->nrad;
    if (nthreads < 1) nthreads = 1;
    lines_work *wk = (lines_work *)calloc(nthreads, sizeof(lines_work));
    pthread_t *th = (pthread_t *)calloc(nthreads, sizeof(pthread_t));
    for (int t = 0; t < nthreads; ++t) {
        wk[t].c = c; wk[t].tab = tab; wk[t].atom = atom; wk[t].nline = nline;
        wk[t].lo = (int)((int64_t)c->nleaf * t / nthreads);
        wk[t].hi = (int)((int64_t)c->nleaf * (t + 1) / nthreads);
        wk[t].L = (double *)calloc(nline + 1, sizeof(double));
    }
    for (int t = 1; t < nthreads; ++t) pthread_create(&th[t], NULL, lines_worker, &wk[t]);
    lines_worker(&wk[0]);
    for (int t = 1; t < nthreads; ++t) pthread_join(th[t], NULL);
    for (int k = 0; k < nline; ++k) {
        double s = 0.0;
        for (int t = 0; t < nthreads; ++t) s = s + wk[t].L[k];
        L[k] = s;
    }
    for (int t = 0; t < nthreads; ++t) free(wk[t].L);
    free(wk);
    free(th);
    return nline;
}
