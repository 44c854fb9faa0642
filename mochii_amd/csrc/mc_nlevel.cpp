// n-level atoms on the host: the statistical-equilibrium solve of
// src/nlevel_mod.f90 (ups_eval :128-160, emitting_level :174-187, the rate
// matrix and Gaussian elimination of nlevel_cooling_prepT / _solve
// :277-361) and the (T, n_e) cooling tables of cooling_model 'local_ne'
// (nlevel_cooling_add, src/nlevel_cooling_mod.f90:105-152), uploaded to the
// device for the thermal update's lookup (mc_nlevel.h).  The rows of a
// table are independent: they are built by host threads.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <thread>

#include "mc_internal.h"
#include "mc_nlevel.h"

namespace {

constexpr double kKb = 1.380649e-16;
constexpr double kCollPref = 8.629e-6;
constexpr double kHcErgCm = 1.9864458571489287e-16;
constexpr int L = MC_NL_MAXLEV;

double ups_eval(const mc_nlevel_atom &a, int k, double T)
{
#pragma clang fp contract(off)
    const double et = (kKb / 2.1798723611035e-11) * T / a.dery[k];
    double st;
    if (a.ttype[k] == 1 || a.ttype[k] == 4)
        st = 1.0 - std::log(a.cups[k]) / std::log(et + a.cups[k]);
    else
        st = et / (et + a.cups[k]);
    st = std::fmin(std::fmax(st, a.st_lo[k]), a.st_hi[k]);
    const double u = a.st_hi[k] > a.st_lo[k]
                         ? 2.0 * (st - a.st_lo[k]) / (a.st_hi[k] - a.st_lo[k]) - 1.0
                         : 0.0;
    // Clenshaw sum of the Chebyshev series
    double b0 = 0.0, b1 = 0.0;
    for (int j = a.nc[k] - 1; j >= 1; --j) {
        const double b2 = b1;
        b1 = b0;
        b0 = 2.0 * u * b1 - b2 + a.coef[k][j];
    }
    double y = u * b0 - b1 + a.coef[k][0];
    if (a.logf[k] == 1) y = std::pow(10.0, y);
    double ups;
    switch (a.ttype[k]) {
    case 1: ups = y * std::log(et + std::exp(1.0)); break;
    case 2: ups = y; break;
    case 3: ups = y / (et + 1.0); break;
    case 4: ups = y * std::log(et + a.cups[k]); break;
    case 5: ups = y / et; break;
    default: ups = std::pow(10.0, y); break;
    }
    return std::fmax(ups, 0.0);
}

// the pair oriented by energy: 0-based lower / upper level and dE [erg]
void emitting(const mc_nlevel_atom &a, int i1, int i2, int &lo, int &hi, double &de)
{
    if (a.e_cm1[i2 - 1] >= a.e_cm1[i1 - 1]) { lo = i1 - 1; hi = i2 - 1; }
    else { lo = i2 - 1; hi = i1 - 1; }
    de = (a.e_cm1[hi] - a.e_cm1[lo]) * kHcErgCm;
}

// collisional (without n_e) and radiative rate matrices, [i][j] = j -> i
void prep_t(const mc_nlevel_atom &a, double T, double *Rc, double *Rr)
{
#pragma clang fp contract(off)
    std::fill(Rc, Rc + L * L, 0.0);
    std::fill(Rr, Rr + L * L, 0.0);
    for (int k = 0; k < a.nups; ++k) {
        int l, u;
        double de;
        emitting(a, a.ul_l[k], a.ul_u[k], l, u, de);
        const double ups = ups_eval(a, k, T);
        const double q_ul = kCollPref / (a.g[u] * std::sqrt(T)) * ups;
        const double q_lu = (a.g[u] / a.g[l]) * q_ul * std::exp(-de / (kKb * T));
        Rc[u * L + l] = Rc[u * L + l] + q_lu;
        Rc[l * L + u] = Rc[l * L + u] + q_ul;
    }
    for (int k = 0; k < a.nrad; ++k) {
        int l, u;
        double de;
        emitting(a, a.rl[k], a.ru[k], l, u, de);
        Rr[l * L + u] = Rr[l * L + u] + a.A[k];
    }
}

// populations of dn/dt = 0 with the closure row replacing level 1's
// equation: Gaussian elimination with partial pivoting (first largest
// pivot), back substitution
void populations(int n, const double *R, double *pop)
{
#pragma clang fp contract(off)
    constexpr int W = L + 1;
    double M[L * W];
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
            if (std::fabs(M[j * W + i]) > std::fabs(M[ip * W + i])) ip = j;
        if (ip != i)
            for (int j = i; j <= n; ++j) std::swap(M[i * W + j], M[ip * W + j]);
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

// Lambda_SE / n_e at (T, n_e) from the rate matrices of T
double cooling_solve(const mc_nlevel_atom &a, const double *Rc, const double *Rr, double ne)
{
#pragma clang fp contract(off)
    double R[L * L], pop[L];
    for (int i = 0; i < L * L; ++i) R[i] = ne * Rc[i] + Rr[i];
    populations(a.nlev, R, pop);
    double lam = 0.0;
    for (int k = 0; k < a.nrad; ++k) {
        int lo, hi;
        double de;
        emitting(a, a.rl[k], a.ru[k], lo, hi, de);
        lam = lam + pop[hi] * a.A[k] * de;
    }
    return lam / ne;
}

// one ion's table: rows i0 .. i1 of Lambda_SE(T_i, n_e,j) into lam[j][i]
void table_rows(const mc_nlevel_atom &a, int i0, int i1, double *lam)
{
    std::vector<double> Rc(L * L), Rr(L * L);
    for (int i = i0; i < i1; ++i) {
        prep_t(a, std::pow(10.0, nl_gT(i)), Rc.data(), Rr.data());
        for (int j = 0; j < kNlNe; ++j)
            lam[(size_t)j * kNlT + i] = cooling_solve(a, Rc.data(), Rr.data(),
                                                      std::pow(10.0, nl_gNe(j)));
    }
}

// the device slot layout of a table from Lambda_SE(T, n_e): lamref = the
// floor column, Ssup = Lambda / max(lamref, tiny)
void to_slot(const std::vector<double> &lam, double *slot)
{
#pragma clang fp contract(off)
    for (int i = 0; i < kNlT; ++i) slot[i] = lam[i];
    for (int j = 0; j < kNlNe; ++j)
        for (int i = 0; i < kNlT; ++i)
            slot[kNlT + (size_t)j * kNlT + i] = lam[(size_t)j * kNlT + i] / std::fmax(slot[i], DBL_MIN);
}

}  // namespace

int find_atom(const mc_ctx *c, const char *name, int stage)
{
    for (size_t k = 0; k < c->nl_atoms.size(); ++k)
        if (c->nl_atoms[k].stage == stage && std::strncmp(c->nl_atoms[k].name, name, 4) == 0)
            return (int)k;
    return -1;
}

extern "C" int mc_nlevel_setup(mc_ctx *c, int n, const mc_nlevel_atom *atoms)
{
    if (!c) {
        mc_set_error("null context");
        return -1;
    }
    if (n < 0 || (n > 0 && !atoms)) {
        mc_set_error("mc_nlevel_setup: bad atom list");
        return -1;
    }
    for (int k = 0; k < n; ++k) {
        const mc_nlevel_atom &a = atoms[k];
        bool ok = a.nlev >= 1 && a.nlev <= MC_NL_MAXLEV && a.nrad >= 0 &&
                  a.nrad <= MC_NL_MAXTR && a.nups >= 0 && a.nups <= MC_NL_MAXTR && a.stage >= 1;
        for (int i = 0; ok && i < a.nrad; ++i)
            ok = a.rl[i] >= 1 && a.rl[i] <= a.nlev && a.ru[i] >= 1 && a.ru[i] <= a.nlev;
        for (int i = 0; ok && i < a.nups; ++i)
            ok = a.ul_l[i] >= 1 && a.ul_l[i] <= a.nlev && a.ul_u[i] >= 1 &&
                 a.ul_u[i] <= a.nlev && a.nc[i] >= 1 && a.nc[i] <= 16 && a.dery[i] > 0.0;
        if (!ok) {
            mc_set_error(std::string("mc_nlevel_setup: bad record for atom ") + a.name + " " +
                         std::to_string(a.stage));
            return -1;
        }
    }
    c->nl_atoms.assign(atoms, atoms + n);
    c->nl_given = 1;
    c->nl_built.clear();              // tables follow the new atoms
    c->nl_dirty = 1;
    return 0;
}

// The (T, n_e) tables of the ions the thermal update cools with under
// cooling_model 'local_ne' (H I, and every registry stage with a cooling
// fit), built on first use with the current registry and uploaded; the
// slots are written into the cooling and registry tables.  Returns < 0 on
// error.
int mc_nl_ready(mc_ctx *c)
{
    c->cool.nlt = nullptr;
    c->cool.nl_hi = -1;
    for (int s = 0; s < MC_GPU_MAXST; ++s) c->mtab.s_nl[s] = -1;
    if (c->p.cooling_model != MC_COOL_LOCAL_NE) return 0;
    if (!c->nl_given) {
        mc_set_error("cooling_model local_ne needs the n-level atoms (mc_nlevel_setup)");
        return -1;
    }
    // the ions: H I, then the registry's stages with a cooling fit
    std::vector<int> want;                      // atom index per slot
    auto slot_of = [&](int ia) -> int {
        for (size_t s = 0; s < want.size(); ++s)
            if (want[s] == ia) return (int)s;
        want.push_back(ia);
        return (int)want.size() - 1;
    };
    const int ih = find_atom(c, "h", 1);
    int hi_slot = ih >= 0 ? slot_of(ih) : -1;
    std::vector<int> st_slot(MC_GPU_MAXST, -1);
    for (int ie = 0, s0 = 0; ie < c->nel; s0 += c->elements[ie].nstage, ++ie) {
        const mc_element &e = c->elements[ie];
        for (int i = 0; i < e.nstage; ++i) {
            if (e.ncool[i] <= 0) continue;
            const int ia = find_atom(c, e.name, i + 1);
            if (ia >= 0) st_slot[s0 + i] = slot_of(ia);
        }
    }
    if (c->nl_dirty || want != c->nl_slots) {
        // build what is not built yet (tables depend on the atom only)
        std::vector<int> todo;
        for (int ia : want)
            if (!c->nl_built.count(ia)) todo.push_back(ia);
        if (!todo.empty()) {
            std::vector<std::vector<double>> lam(todo.size(),
                                                 std::vector<double>((size_t)kNlT * kNlNe));
            const unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
            const int chunk = 8;
            std::vector<std::pair<int, int>> jobs;       // (ion, first row)
            for (size_t t = 0; t < todo.size(); ++t)
                for (int i0 = 0; i0 < kNlT; i0 += chunk) jobs.push_back({(int)t, i0});
            std::atomic<size_t> next(0);
            auto work = [&]() {
                for (size_t q; (q = next++) < jobs.size();) {
                    const int t = jobs[q].first, i0 = jobs[q].second;
                    table_rows(c->nl_atoms[todo[t]], i0, std::min(i0 + chunk, kNlT),
                               lam[t].data());
                }
            };
            std::vector<std::thread> th;
            for (unsigned k = 1; k < nth; ++k) th.emplace_back(work);
            work();
            for (auto &t : th) t.join();
            for (size_t t = 0; t < todo.size(); ++t) {
                std::vector<double> slot(kNlSlot);
                to_slot(lam[t], slot.data());
                c->nl_built[todo[t]] = std::move(slot);
            }
        }
        std::vector<double> h((size_t)std::max<size_t>(want.size(), 1) * kNlSlot, 0.0);
        for (size_t s = 0; s < want.size(); ++s)
            std::copy(c->nl_built[want[s]].begin(), c->nl_built[want[s]].end(),
                      h.begin() + (long)s * kNlSlot);
        if (c->d_nlt) (void)hipFree(c->d_nlt);
        c->d_nlt = nullptr;
        MC_HIP(hipMalloc(&c->d_nlt, sizeof(double) * h.size()));
        MC_HIP(hipMemcpy(c->d_nlt, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
        c->nl_slots = want;
        c->nl_dirty = 0;
    }
    c->cool.nlt = c->d_nlt;
    c->cool.nl_hi = hi_slot;
    for (int s = 0; s < MC_GPU_MAXST; ++s) c->mtab.s_nl[s] = st_slot[s];
    return 0;
}

extern "C" int mc_lines_setup(mc_ctx *c, const mc_line_table *hi, const mc_line_table *heii,
                              const mc_line_table *hei)
{
    if (!c) {
        mc_set_error("null context");
        return -1;
    }
    const mc_line_table *t3[3] = {hi, heii, hei};
    for (int k = 0; k < 3; ++k) {
        const mc_line_table *t = t3[k];
        if (!t) continue;
        bool ok = t->nT >= 2 && t->nT <= MC_LT_MAXG && t->nD >= 2 && t->nD <= MC_LT_MAXG &&
                  t->nlin >= 0 && t->nlin <= MC_LT_MAXL;
        for (int i = 1; ok && i < t->nT; ++i) ok = t->tgrid[i] > t->tgrid[i - 1];
        for (int i = 1; ok && i < t->nD; ++i) ok = t->dgrid[i] > t->dgrid[i - 1];
        for (int l = 0; ok && l < t->nlin; ++l)
            for (int i = 0; ok && i < t->nT; ++i)
                for (int j = 0; ok && j < t->nD; ++j) ok = t->etab[l][i][j] > 0.0;
        if (!ok) {
            mc_set_error("mc_lines_setup: a table needs 2..32 ascending T and n_e nodes, at most "
                         "12 lines, positive emissivities");
            return -1;
        }
    }
    if (!hi) {
        mc_set_error("mc_lines_setup: the H I table is required");
        return -1;
    }
    for (int k = 0; k < 3; ++k) {
        c->lt_on[k] = t3[k] != nullptr;
        if (t3[k]) c->lt[k] = *t3[k];
    }
    return 0;
}

// the metal ions of the line list: every registry (element, stage) with an
// n-level atom, in the reference's order, oriented for the device
int mc_lines_ions(mc_ctx *c, int out0, std::vector<DevLineIon> &ions, std::vector<double> &wl,
                  std::vector<int32_t> &code)
{
    int out = out0;
    for (int ie = 0, t0 = 0; ie < c->nel; t0 += c->elements[ie].nstage - 1, ++ie) {
        const mc_element &e = c->elements[ie];
        for (int st = 1; st <= e.nstage; ++st) {
            const int ia = find_atom(c, e.name, st);
            if (ia < 0) continue;
            const mc_nlevel_atom &a = c->nl_atoms[ia];
            if (a.nlev > kLnLev || a.nrad > kLnRad || a.nups > kLnUps) {
                mc_set_error(std::string("mc_lines_compute: atom ") + e.name + " " +
                             std::to_string(st) + " exceeds the device line solve (16 levels, 48 "
                             "radiative and 96 collision rows)");
                return -1;
            }
            DevLineIon d;
            std::memset(&d, 0, sizeof(d));
            d.nlev = a.nlev; d.nrad = a.nrad; d.nups = a.nups; d.out0 = out;
            // stage st of the element: the lower stage of its first
            // transition (st = 1) or the upper stage of transition st - 2
            d.t = st == 1 ? t0 : t0 + st - 2;
            d.up = st == 1 ? 0 : 1;
            d.abund = e.abund;
            for (int k = 0; k < a.nrad; ++k) {
                int lo, hi;
                double de;
                emitting(a, a.rl[k], a.ru[k], lo, hi, de);
                d.rlo[k] = lo; d.rhi[k] = hi; d.A[k] = a.A[k]; d.rde[k] = de;
                const double dcm = std::fabs(a.e_cm1[a.ru[k] - 1] - a.e_cm1[a.rl[k] - 1]);
                wl.push_back(dcm > 0.0 ? 1.0e8 / dcm : 0.0);
                code.push_back(100 * e.Z + st);
            }
            for (int k = 0; k < a.nups; ++k) {
                int lo, hi;
                double de;
                emitting(a, a.ul_l[k], a.ul_u[k], lo, hi, de);
                d.clo[k] = lo; d.chi[k] = hi; d.cde[k] = de;
                d.gu[k] = a.g[hi];
                d.gr[k] = a.g[hi] / a.g[lo];
                d.ttype[k] = a.ttype[k]; d.logf[k] = a.logf[k]; d.nc[k] = a.nc[k];
                d.cups[k] = a.cups[k]; d.dery[k] = a.dery[k];
                d.st_lo[k] = a.st_lo[k]; d.st_hi[k] = a.st_hi[k];
                for (int j = 0; j < 16; ++j) d.coef[k][j] = a.coef[k][j];
            }
            ions.push_back(d);
            out += a.nrad;
        }
    }
    return out;
}

extern "C" int mc_lines_compute(mc_ctx *c, int maxn, double *L, double *wl, int32_t *ion,
                                int32_t *n)
{
    if (!c) {
        mc_set_error("null context");
        return -1;
    }
    if (!c->state_ready || !c->lt_on[0]) {
        mc_set_error("mc_lines_compute: needs the gas state and the H I line table "
                     "(mc_lines_setup)");
        return -1;
    }
    std::vector<double> Lv, wv;
    std::vector<int32_t> cv;
    const int rc = mc_run_lines(c, Lv, wv, cv);
    if (rc) return rc;
    const int m = std::min<int>(maxn, (int)Lv.size());
    for (int k = 0; k < m; ++k) {
        if (L) L[k] = Lv[k];
        if (wl) wl[k] = wv[k];
        if (ion) ion[k] = cv[k];
    }
    if (n) *n = (int32_t)Lv.size();
    return 0;
}
