// Dust thermal-infrared band: host set-up (femit table, emission grid and
// band cut) and the C-ABI entry points of dust_temp_mod,
// dust_emis_transport_mod and dust_ir_band_mod.  The per-leaf work runs on
// the device (mc_kernels.hip: mc_run_dust_temp, mc_run_dust_emission).
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>

#include "mc_internal.h"

#pragma clang fp contract(off)

namespace {

constexpr double kClight = 2.99792458e10;
constexpr double kHplanck = 6.62607015e-27;
constexpr double kLamBBMin = 0.1, kLamBBMax = 3.0e3;   // um

template <class T>
int dalloc(T **p, size_t n)
{
    *p = nullptr;
    MC_HIP(hipMalloc((void **)p, sizeof(T) * std::max<size_t>(n, 1)));
    return 0;
}

// planck_nu (src/dust_temp_mod.f90:121-130)
double planck_nu(double nu, double T)
{
    const double x = kHplanck * nu / (MC_KBOLTZ * T);
    if (x < 700.0) {
        const double nu3 = nu * nu * nu;
        return 2.0 * kHplanck * nu3 / (kClight * kClight) / (std::exp(x) - 1.0);
    }
    return 0.0;
}

// ion_dust_sabs_at (src/gas_opacity_mod.f90:217-223): (1 - albedo) C_ext /
// C_ext(lambda_ref) at E with the table's lin / log-log interpolation
double dust_sabs_at(const mc_ctx *c, double E)
{
    const int n = (int)c->dlam.size();
    const double *lam = c->dlam.data(), *alb = c->dalb.data(), *ce = c->dcext.data();
    const double x0 = 1.23984 / E;
    double al, cx;
    if (x0 <= lam[0]) {
        al = alb[0]; cx = ce[0];
    } else if (x0 >= lam[n - 1]) {
        al = alb[n - 1]; cx = ce[n - 1];
    } else {
        int i = 0;
        while (i < n - 1 && !(x0 < lam[i + 1])) ++i;
        const double w = (x0 - lam[i]) / (lam[i + 1] - lam[i]);
        al = alb[i] * (1.0 - w) + alb[i + 1] * w;
        const double wl = std::log(x0 / lam[i]) / std::log(lam[i + 1] / lam[i]);
        cx = std::exp(std::log(std::max(ce[i], DBL_MIN)) * (1.0 - wl) +
                      std::log(std::max(ce[i + 1], DBL_MIN)) * wl);
    }
    return (1.0 - al) * cx / c->dcref;
}

// gas_absorption_at (src/dust_emis_transport_mod.f90:186-201): H I, He I,
// He II and every loaded metal transition at E
double gas_absorption_at(const mc_ctx *c, double E)
{
    double s = sigma_hydrogenic(E, MC_ETH_HI, 1.0) + sigma_HeI(E) +
               sigma_hydrogenic(E, MC_ETH_HEII, 2.0);
    if (c->ntr > 0) {
        double t = 0.0;
        int n = 0;
        for (int k = 0; k < c->ntr; ++k) {          // threshold order, early exit
            if (E < c->mtab.flat_eth[k]) break;
            n = k + 1;
        }
        for (int k = 0; k < n; ++k) t = t + m_sigma(&c->mtab, c->mtab.flat_t[k], E);
        if (n > 0) s = s + t;
    }
    return s;
}

void free_ir_device(IrBand &b)
{
    for (void *p : {(void *)b.d_dt, (void *)b.d_lam, (void *)b.d_sabs, (void *)b.d_edge,
                    (void *)b.d_tdust, (void *)b.d_heat_euv, (void *)b.d_heat_ir,
                    (void *)b.d_irabs, (void *)b.d_lum, (void *)b.d_lcum, (void *)b.d_bsum,
                    (void *)b.d_khi, (void *)b.d_kmax, (void *)b.d_lesc, (void *)b.d_led,
                    (void *)b.d_segs, (void *)b.d_wave})
        if (p) (void)hipFree(p);
    b.d_wave = nullptr;
    b.n_wave = 0;
    b.d_dt = b.d_lam = b.d_sabs = b.d_edge = nullptr;
    b.d_tdust = b.d_heat_euv = b.d_heat_ir = b.d_irabs = nullptr;
    b.d_lum = b.d_lcum = b.d_bsum = nullptr;
    b.d_khi = b.d_kmax = nullptr;
    b.d_lesc = b.d_led = nullptr;
    b.d_segs = nullptr;
}

}  // namespace

// dust_temp_setup (src/dust_temp_mod.f90:39-118): femit(T) = 4 pi int
// s_abs(nu) B_nu(T) dnu on 61 log-spaced T (10^0.5 .. 10^3.5 K), trapezoid
// over the sorted extinction table in frequency
int mc_ir_femit_table(mc_ctx *c)
{
    const int n = (int)c->dlam.size();
    if (n < 2) {
        mc_set_error("mc_dust_temp_setup: load the extinction table first (mc_ion_dust_table)");
        return -1;
    }
    IrBand &b = c->ir;
    const double *lam = c->dlam.data(), *alb = c->dalb.data(), *ce = c->dcext.data();
    for (int i = 0; i < MC_DT_NT; ++i) {
        b.tgrid[i] = std::pow(10.0, 0.5 + 3.0 * (double)i / (double)(MC_DT_NT - 1));
        const double T = b.tgrid[i];
        double f = 0.0;
        for (int k = 0; k < n - 1; ++k) {
            const double nu1 = kClight / (lam[k + 1] * 1.0e-4);
            const double nu2 = kClight / (lam[k] * 1.0e-4);
            const double snu =
                0.5 * ((1.0 - alb[k]) * ce[k] + (1.0 - alb[k + 1]) * ce[k + 1]) / c->dcref;
            const double bnu = 0.5 * (planck_nu(nu1, T) + planck_nu(nu2, T));
            f = f + MC_FOURPI * snu * bnu * (nu2 - nu1);
        }
        b.femit[i] = f;
    }
    return 0;
}

// dust_emis_transport_setup (src/dust_emis_transport_mod.f90:100-180) for the
// equilibrium spectrum and dust_ir_band_setup (src/dust_ir_band_mod.f90:63-84)
int mc_ir_emission_grid(mc_ctx *c)
{
    IrBand &b = c->ir;
    const double lam_floor = 1.23984 / c->p.eion_min;
    const double lam_start = std::max(kLamBBMin, lam_floor);
    const int n = MC_NLAM_BB;
    b.nl = n;
    b.lam.assign(n, 0.0);
    b.sabs.assign(n, 0.0);
    b.edge.assign(n + 1, 0.0);
    for (int k = 0; k < n; ++k) {
        b.lam[k] = lam_start * std::pow(kLamBBMax / lam_start, (double)k / (double)(n - 1));
        b.sabs[k] = dust_sabs_at(c, 1.23984 / b.lam[k]);
    }
    int kpass = 0;                                   // 1-based, 0 = none
    for (int k = 0; k < n; ++k) {
        if (b.lam[k] < lam_floor) continue;
        if (gas_absorption_at(c, 1.23984 / b.lam[k]) <= 0.0) {
            kpass = k + 1;
            break;
        }
    }
    if (kpass == 0) {
        mc_set_error("mc_dust_emis_transport_setup: no wavelength of the dust emission grid is "
                     "free of both the transported band and every gas photoionization threshold");
        return -1;
    }
    b.nzero = kpass > 1 ? kpass : 0;
    b.lam_hardest = b.lam[kpass - 1];
    for (int k = 1; k < n; ++k) b.edge[k] = std::sqrt(b.lam[k - 1] * b.lam[k]);
    b.edge[0] = b.lam[0] * std::sqrt(b.lam[0] / b.lam[1]);
    b.edge[n] = b.lam[n - 1] * std::sqrt(b.lam[n - 1] / b.lam[n - 2]);
    return 0;
}

// per-leaf and per-bin device arrays of the band (sized by the grid)
static int ir_alloc(mc_ctx *c)
{
    IrBand &b = c->ir;
    const size_t nl = (size_t)c->nleaf;
    const size_t nb = (nl + 1023) / 1024;             // scan tiles of 1024 leaves
    if (!b.d_dt && dalloc(&b.d_dt, 2 * MC_DT_NT)) return -100;
    if (!b.d_tdust) {
        if (dalloc(&b.d_tdust, nl) || dalloc(&b.d_heat_euv, nl) || dalloc(&b.d_heat_ir, nl) ||
            dalloc(&b.d_irabs, nl) || dalloc(&b.d_lum, nl) || dalloc(&b.d_lcum, nl) ||
            dalloc(&b.d_bsum, nb + 1) || dalloc(&b.d_khi, nl) || dalloc(&b.d_kmax, nb + 1) ||
            dalloc(&b.d_lam, MC_NLAM_BB) ||
            dalloc(&b.d_sabs, MC_NLAM_BB) || dalloc(&b.d_edge, MC_NLAM_BB + 1) ||
            dalloc(&b.d_lesc, MC_NLAM_BB) || dalloc(&b.d_led, 3) || dalloc(&b.d_segs, 4))
            return -100;
        MC_HIP(hipMemset(b.d_tdust, 0, sizeof(double) * nl));
        MC_HIP(hipMemset(b.d_heat_ir, 0, sizeof(double) * nl));
        MC_HIP(hipMemset(b.d_irabs, 0, sizeof(double) * nl));
        MC_HIP(hipMemset(b.d_lesc, 0, sizeof(double) * MC_NLAM_BB));
        MC_HIP(hipMemset(b.d_led, 0, sizeof(double) * 3));
    }
    return 0;
}

void mc_ir_free(mc_ctx *c)
{
    free_ir_device(c->ir);
    c->ir.temp_ready = false;
    c->ir.setup_ready = false;
}

#define CHECK_CTX(c)                                 \
    do {                                             \
        if (!(c)) {                                  \
            mc_set_error("null context");            \
            return -1;                               \
        }                                            \
    } while (0)

extern "C" int mc_dust_temp_setup(mc_ctx *c)
{
    CHECK_CTX(c);
    if (!c->p.ion_add_dust || !c->st.heat_dust) {
        mc_set_error("mc_dust_temp_setup: needs ion_add_dust and a grid");
        return -1;
    }
    if (mc_ir_femit_table(c)) return -1;
    if (ir_alloc(c)) return -100;
    std::vector<double> dt(2 * MC_DT_NT);
    std::copy(c->ir.tgrid, c->ir.tgrid + MC_DT_NT, dt.begin());
    std::copy(c->ir.femit, c->ir.femit + MC_DT_NT, dt.begin() + MC_DT_NT);
    MC_HIP(hipMemcpy(c->ir.d_dt, dt.data(), sizeof(double) * dt.size(), hipMemcpyHostToDevice));
    c->ir.temp_ready = true;
    return 0;
}

extern "C" int mc_dust_temp_compute(mc_ctx *c)
{
    CHECK_CTX(c);
    if (!c->ir.temp_ready) {
        mc_set_error("mc_dust_temp_compute: call mc_dust_temp_setup first");
        return -1;
    }
    if (!c->state_ready) {
        mc_set_error("mc_dust_temp_compute: no gas state (mc_gas_state_setup)");
        return -1;
    }
    return mc_run_dust_temp(c, c->st.heat_dust);
}

extern "C" int mc_dust_emis_transport_setup(mc_ctx *c)
{
    CHECK_CTX(c);
    if (!c->ir.temp_ready) {
        mc_set_error("mc_dust_emis_transport_setup: call mc_dust_temp_setup first");
        return -1;
    }
    if (mc_ir_emission_grid(c)) return -1;
    IrBand &b = c->ir;
    MC_HIP(hipMemcpy(b.d_lam, b.lam.data(), sizeof(double) * b.nl, hipMemcpyHostToDevice));
    MC_HIP(hipMemcpy(b.d_sabs, b.sabs.data(), sizeof(double) * b.nl, hipMemcpyHostToDevice));
    MC_HIP(hipMemcpy(b.d_edge, b.edge.data(), sizeof(double) * (b.nl + 1), hipMemcpyHostToDevice));
    MC_HIP(hipMemset(b.d_segs, 0, sizeof(unsigned long long) * 4));
    b.frac_discarded = 0.0;
    if (mc_packet_capacity(c, (int64_t)std::fmax(c->p.dust_no_photons, 1.0))) return -100;
    b.setup_ready = true;
    return 0;
}

extern "C" int mc_transport_dust_emission(mc_ctx *c, double *report)
{
    CHECK_CTX(c);
    if (!c->ir.setup_ready) {
        mc_set_error("mc_transport_dust_emission: call mc_dust_emis_transport_setup first");
        return -1;
    }
    if (!c->state_ready) {
        mc_set_error("mc_transport_dust_emission: no gas state (mc_gas_state_setup)");
        return -1;
    }
    double rep[MC_DUSTRT_NREPORT];
    const int rc = mc_run_dust_emission(c, rep);
    if (report) std::memcpy(report, rep, sizeof(rep));
    return rc;
}

extern "C" int mc_dust_get(mc_ctx *c, int field, double *dst)
{
    CHECK_CTX(c);
    if (!dst) {
        mc_set_error("mc_dust_get: null destination");
        return -1;
    }
    const double *src = nullptr;
    switch (field) {
    case MC_DUST_HEAT: src = c->st.heat_dust; break;
    case MC_DUST_TEMP: src = c->ir.d_tdust; break;
    case MC_DUST_HEAT_IR: src = c->ir.d_heat_ir; break;
    case MC_DUST_IR_ABS: src = c->ir.d_irabs; break;
    default: mc_set_error("mc_dust_get: unknown field"); return -1;
    }
    if (!src) {
        mc_set_error("mc_dust_get: field not allocated (ion_add_dust off or band not set up)");
        return -1;
    }
    MC_HIP(hipMemcpy(dst, src, sizeof(double) * c->nleaf, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int mc_dust_set(mc_ctx *c, int field, const double *src)
{
    CHECK_CTX(c);
    if (field != MC_DUST_HEAT || !src || !c->st.heat_dust) {
        mc_set_error("mc_dust_set: only MC_DUST_HEAT, with ion_add_dust on");
        return -1;
    }
    MC_HIP(hipMemcpy(c->st.heat_dust, src, sizeof(double) * c->nleaf, hipMemcpyHostToDevice));
    return 0;
}

extern "C" int mc_dust_ir_energy_budget(mc_ctx *c, double *out3)
{
    CHECK_CTX(c);
    if (!out3) return -1;
    for (int k = 0; k < 3; ++k) out3[k] = c->ir.led[k];
    return 0;
}

extern "C" int mc_dust_ir_escaped_spectrum(mc_ctx *c, double *lam, double *Llam, double *dlam)
{
    CHECK_CTX(c);
    const IrBand &b = c->ir;
    if (!b.setup_ready) {
        mc_set_error("mc_dust_ir_escaped_spectrum: band not set up");
        return -1;
    }
    std::vector<double> lesc(b.nl);
    MC_HIP(hipMemcpy(lesc.data(), b.d_lesc, sizeof(double) * b.nl, hipMemcpyDeviceToHost));
    for (int k = 0; k < b.nl; ++k) {
        const double dl = b.edge[k + 1] - b.edge[k];
        if (lam) lam[k] = b.lam[k];
        if (dlam) dlam[k] = dl;
        if (Llam) Llam[k] = lesc[k] / dl;
    }
    return 0;
}

extern "C" int mc_dust_ir_grid(mc_ctx *c, int32_t *nl, double *lam, double *sabs, double *edge,
                               int32_t *nzero, double *lam_hardest, double *femit122)
{
    CHECK_CTX(c);
    const IrBand &b = c->ir;
    if (nl) *nl = b.nl;
    if (b.setup_ready) {
        if (lam) std::copy(b.lam.begin(), b.lam.end(), lam);
        if (sabs) std::copy(b.sabs.begin(), b.sabs.end(), sabs);
        if (edge) std::copy(b.edge.begin(), b.edge.end(), edge);
        if (nzero) *nzero = b.nzero;
        if (lam_hardest) *lam_hardest = b.lam_hardest;
    }
    if (femit122 && b.temp_ready) {
        std::copy(b.tgrid, b.tgrid + MC_DT_NT, femit122);
        std::copy(b.femit, b.femit + MC_DT_NT, femit122 + MC_DT_NT);
    }
    return 0;
}

extern "C" int mc_dust_ir_timing(mc_ctx *c, double *out3, int64_t *segments2)
{
    CHECK_CTX(c);
    if (out3)
        for (int k = 0; k < 3; ++k) out3[k] = c->ir.times[k];
    if (segments2) {
        segments2[0] = c->ir.segs[0];
        segments2[1] = c->ir.segs[1];
    }
    return 0;
}
