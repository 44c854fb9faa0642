/*
 * MoCHII photon-packet transport + cell-balance: FP64 CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a from-scratch C restatement of the
 * reference algorithm (seoncafe/MoCHII, Fortran 90) for the one hot path the
 * HIP library implements.  It exists to check the HIP path: only the tests,
 * the smoke check and the CPU-baseline leg of the benchmark load it.  The
 * product library never links or calls it.
 *
 * Every function cites the reference file:line it restates.  Parity status of
 * each piece (what pins it) is listed in the header comment of the .c file
 * that holds it.
 */
#ifndef MC_ORACLE_H
#define MC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- constants (src/define.f90:12-88) ---------------------------------- */
#define ORC_PI      3.141592653589793238462643383279502884197
#define ORC_TWOPI   6.283185307179586476925286766559005768394
#define ORC_FOURPI  12.56637061435917295385057353311801153679
#define ORC_EV2ERG  1.602176634e-12
#define ORC_HPLANCK 6.62607015e-27
#define ORC_KBOLTZ  1.380649e-16
#define ORC_PC2CM   3.0856776e18
#define ORC_ETH_HI   13.598
#define ORC_ETH_HEI  24.587
#define ORC_ETH_HEII 54.416
#define ORC_TAU_HUGE 745.2
#define ORC_MAXSRC   16
#define ORC_QMC_MAXDIM 12

/* ---- QMC (src/qmc_mod.f90) ---------------------------------------------- */
typedef struct {
    uint32_t V[32][ORC_QMC_MAXDIM];   /* direction numbers V_d(k), k=1..32 */
    uint32_t key[ORC_QMC_MAXDIM][2];  /* Owen keys per dimension and stream */
} orc_qmc;

void     orc_qmc_setup(orc_qmc *q, int seed);
uint32_t orc_sobol_code(const orc_qmc *q, uint64_t idx, int d);   /* d 0-based */
uint32_t orc_owen_scramble(uint32_t x, uint32_t seed);
uint32_t orc_hash32(uint32_t x);
void     orc_qmc_uniforms(const orc_qmc *q, uint64_t idx, int n, int stream, double *u);
void     orc_qmc_raw(const orc_qmc *q, uint64_t idx, int n, double *u);

/* ---- energy sampler (src/energy_sampler_mod.f90) ------------------------ */
typedef struct {
    int n;
    double *energy, *density, *cdf;
    double total_integral;
} orc_sampler;

int    orc_build_tabulated(orc_sampler *s, const double *e, const double *f, int n);
int    orc_build_planck(orc_sampler *s, double T, double emin, double emax,
                        double rtol, const double *thr, int nthr);
double orc_sample_energy(const orc_sampler *s, double u);
double orc_energy_cdf_value(const orc_sampler *s, double E);
double orc_planck_shape(double E, double T);
void   orc_sampler_free(orc_sampler *s);

/* ---- cross sections and rate coefficients ------------------------------- */
double orc_sigma_hydrogenic(double E, double Eth, double Z);
double orc_sigma_vfky96(double E, double Eth, double E0, double s0, double ya,
                        double P, double yw, double y0, double y1);
double orc_sigma_HI(double E);
double orc_sigma_HeI(double E);
double orc_sigma_HeII(double E);
double orc_alphaA_HII(double T);
double orc_alphaB_HII(double T);
double orc_alphaA_HeII(double T);
double orc_alphaB_HeII(double T);
double orc_alphaA_HeIII(double T);
double orc_alphaB_HeIII(double T);
int    orc_ci_model_set(int ci_dere);
double orc_ci_HI(double T);
double orc_ci_HeI(double T);
double orc_ci_HeII(double T);

/* ---- cooling (src/cooling_mod.f90, src/gaunt.f90) ----------------------- */
struct orc_nlcool_s;
typedef struct {
    int nfit;                 /* H I line-cooling fit terms */
    double A[16], Ti[16];
    double logT[41], gff[4 * 41];
    /* cooling_model 'local_ne': the H I (T, n_e) table, NULL = fit unsuppressed */
    const struct orc_nlcool_s *hi_tab;
} orc_cooling;

void   orc_gaunt_ff(float z, float log10Te, const float *xlf, float *g, int n);
double orc_gaunt_ff_mean(double T, double Zc);
void   orc_gff_table(double *logT41, double *gff41x4);
double orc_gbar_ff(const double *logT, const double *gff, double T, int Z);
double orc_cooling_total(const orc_cooling *cl, double T, double nH, double ne, double xHI,
                         double xHeI, double xHeII, double yHe, int caseA);

/* ---- counter-based RNG for the post-launch draws ------------------------ */
double orc_rng_uniform(uint32_t seed, int stream, uint64_t id, uint32_t k);

/* ---- model context ------------------------------------------------------- */
enum { ORC_GRID_CAR = 0, ORC_GRID_AMR = 1 };
enum { ORC_GEOM_POINT = 0, ORC_GEOM_EXT_SPH = 1, ORC_GEOM_EXT_REC = 2 };

typedef struct {
    /* launch */
    int64_t nphotons;
    int     qmc_seed;
    double  eion_min, eion_max, source_cdf_tol;
    int     nnu_ion, align_edges;
    int     nsource;                 /* internal point sources */
    double  tstar[ORC_MAXSRC], lum[ORC_MAXSRC];
    double  sx[ORC_MAXSRC], sy[ORC_MAXSRC], sz[ORC_MAXSRC];
    int     ext_on;                  /* external isotropic field */
    int     ext_geometry;            /* ORC_GEOM_EXT_SPH / _REC */
    double  ext_tstar, ext_intensity, rmax;
    /* grid */
    int     grid_type, car_walk_amr;
    int     nx, ny, nz;
    double  xmax, ymax, zmax;        /* box is [-xmax,xmax] etc. */
    double  distance2cm;
    /* gas */
    double  He_abund, te_fixed, ion_relax, te_min, te_max;
    int     caseA;
    /* dust */
    int     add_dust, dust_scatter;
    double  lambda_ref;
    /* explicit diffuse ground-recombination field (par%diffuse_field,
     * par%diffuse_energy_model: 0 threshold, 1 exponential, 2 milne) */
    int     diffuse_field, diffuse_model;
    /* trace metals: ion_setup's metal thresholds (src/main.f90:66-82) and the
     * par switches (src/define.f90:593-652, 776-801); the registry itself is
     * loaded by orc_species_setup */
    int     nmetal;
    double  metal_eth[64];
    int     ion_metal_abs, metal_ne, metal_heat, metal_freefree;
    int     cx_two_way;              /* par%charge_exchange == 'two_way' */
    int     ci_dere;                 /* par%ci_model == 'dere_hybrid'    */
    int     hei_diffuse;             /* par%hei_diffuse: He I excited channel */
} orc_params;

/* He I 2^1S two-photon continuum above the H I edge
 * (src/hei_twophoton_mod.f90:44-56): the tabulated shape, photons and energy
 * per decay in the ionizing window, the energy-weighted sampling CDF */
#define ORC_HEI2PH_MAXPTS 128
#define ORC_HEI2PH_NCDF 1024
typedef struct {
    int    ready, n_pts;
    double tab_y[ORC_HEI2PH_MAXPTS], tab_A[ORC_HEI2PH_MAXPTS];
    double n_per_decay, e_per_decay, y_lo, y_hi;
    double grid_y[ORC_HEI2PH_NCDF], cdf_e[ORC_HEI2PH_NCDF];
} orc_hei2ph;

/* ---- trace-metal registry (src/species_mod.f90:54-128) -------------------
 * One element as read_element fills it, laid out exactly like mc_element of
 * the C-ABI header so one host-side record serves both. */
#define ORC_MAX_EL 11
#define ORC_MAX_ST 6
#define ORC_MAX_TR 64
typedef struct {
    char    name[4];
    int32_t Z, nstage;
    double  abund;
    double  eth[ORC_MAX_ST];
    int32_t photo_form[ORC_MAX_ST];
    double  photo[ORC_MAX_ST][8];
    double  ci[ORC_MAX_ST][5];
    double  rr[ORC_MAX_ST][6];
    int32_t ndr[ORC_MAX_ST];
    double  dr_c[ORC_MAX_ST][9], dr_e[ORC_MAX_ST][9];
    int32_t has_dr2[ORC_MAX_ST];
    double  dr2[ORC_MAX_ST][4];
    int32_t has_rr2[ORC_MAX_ST];
    double  rr2[ORC_MAX_ST][2];
    int32_t cxi_form[ORC_MAX_ST];
    double  cxi[ORC_MAX_ST][6];
    int32_t cxr_form[ORC_MAX_ST];
    double  cxr[ORC_MAX_ST][6];
    int32_t ncool[ORC_MAX_ST];
    double  cool_A[ORC_MAX_ST][8], cool_T[ORC_MAX_ST][8];
} orc_element;

/* (leaf, T) rate cache of species_ne_prepare (src/species_mod.f90:114-128,
 * 517-536), element-order transitions */
typedef struct {
    double gam[ORC_MAX_TR], ci[ORC_MAX_TR], rec[ORC_MAX_TR], cxi[ORC_MAX_TR], cxr[ORC_MAX_TR];
} orc_mcache;

double orc_sigma_vy95(double E, double Eth, double E0, double s0, double ya, double P,
                      double yw, double l);
double orc_ci_dere_ratio(int ci_dere, int Z, int stage);
double orc_voronov_ci(const double c[5], double T);
double orc_alpha_rec(const orc_element *el, int it, double T);
double orc_cx_rate(int form, const double c[6], double T);
double orc_species_sigma(const orc_element *el, int it, double E);

/* ---- Milne ground-continuum tables (src/milne_recomb_spectrum_mod.f90) -- */
#define ORC_MILNE_NT 192
#define ORC_MILNE_NX 512
typedef struct {
    double lgT[ORC_MILNE_NT], xg[ORC_MILNE_NX];
    double cdf[3][ORC_MILNE_NT][ORC_MILNE_NX];   /* energy-luminosity CDF in x */
    double xbar[3][ORC_MILNE_NT];                /* <x> under the photon PDF   */
    double fin[3][ORC_MILNE_NT];                 /* in-band photon fraction    */
    double eth[3], eion_max;
} orc_milne;

void   orc_milne_setup(orc_milne *m, double eion_max);
double orc_milne_energy_per_recomb(const orc_milne *m, int ich, double T);
double orc_milne_sample_energy(const orc_milne *m, int ich, double T, double u);

/* ---- dust thermal-infrared band (src/dust_temp_mod.f90,
 * src/dust_emis_transport_mod.f90, src/dust_ir_band_mod.f90) ------------- */
#define ORC_DT_NT      61        /* femit temperature knots             */
#define ORC_NLAM_BB    300       /* equilibrium emission grid           */
#define ORC_LAM_BB_MIN 0.1       /* um */
#define ORC_LAM_BB_MAX 3.0e3     /* um */
#define ORC_IR_NREPORT 10
typedef struct {
    double  tgrid[ORC_DT_NT], femit[ORC_DT_NT];
    int     temp_ready, setup_ready;
    int     nl, nzero;           /* emission grid, leading nodes zeroed */
    double *lam, *sabs, *edge;   /* nl, nl, nl+1 [um] */
    double  lam_hardest, frac_discarded;
    double *tdust;               /* nleaf */
    double *irabs, *heat_ir;     /* nleaf: path tally, heating */
    double *lesc;                /* nl: escaped luminosity per bin */
    int64_t *khi;                /* nleaf: last packet index of each leaf */
    double  emitted, extinct, escaped, Ltot, Lpacket, xi;
    int64_t seg_direct, seg_scattered;
} orc_dustir;

typedef struct {
    orc_params p;
    orc_qmc    qmc;
    /* band */
    int     nnu_band;
    double *eedge;                   /* nnu_band+1 */
    int     ncomp, multi_src;
    int     comp_kind[ORC_MAXSRC + 1];
    double  src_cdf[ORC_MAXSRC + 1];
    orc_sampler samp[ORC_MAXSRC + 1];
    double  Ltot;
    int     qmc_ndim;
    /* grid */
    int     nleaf, ncells, car;
    double  xmin, xmaxb, ymin, ymaxb, zmin, zmaxb, xrange, yrange, zrange, dcell;
    int    *parent, *children, *level, *ileaf, *icell_of_leaf, *neighbor;
    double *cx, *cy, *cz, *ch;
    /* leaf state */
    double *nH, *xHI, *xHeI, *xHeII, *ne, *Te, *rhokap;
    /* dust optics table (sorted by wavelength) */
    int     dn;
    double *dlam, *dalb, *dg, *dcext, dcref;
    /* tallies (leaf-major: hhe[il*6+k], jt[il*nnu+inu]) */
    double *hhe, *jt, *dust_abs;
    int64_t seg_direct, seg_scattered;
    double  e_emit, e_abs, e_esc;
    /* rates */
    double *gHI, *gHeI, *gHeII, *hHI, *hHeI, *hHeII;
    /* cooling tables */
    orc_cooling cool;
    /* diffuse field: Milne tables, per-leaf channel luminosities dif_ch
     * (leaf-major, 4 per leaf: H I, He I, He II ground continua, He I
     * excited cascade), leaf CDF, totals of the last build */
    orc_milne *milne;
    orc_hei2ph *hei;                 /* He I two-photon shape (par%hei_diffuse) */
    double *dif_ch, *dif_cdf, diffuse_lum;
    int64_t diffuse_nphot;
    /* trace metals (species_mod module state): registry, element-order
     * transition/stage offsets, the threshold-sorted flat registry
     * (flat_ie/flat_it/flat_eth, src/species_mod.f90:109-112, 196-227),
     * path tallies mraw[(il*ntr + k)*2 + {0,1}] in flat order k, rates
     * mgam/mheat[il*ntr + t] and stage fractions mfrac[il*nst + s] in
     * element order */
    int     use_metals, nel, ntr, nst;
    orc_element el[ORC_MAX_EL];
    int     tr_off[ORC_MAX_EL], st_off[ORC_MAX_EL];
    int     flat_ie[ORC_MAX_TR], flat_it[ORC_MAX_TR];
    double  flat_eth[ORC_MAX_TR];
    double *mraw, *mgam, *mheat, *mfrac;
    /* dust thermal-infrared band, NULL until its setup */
    orc_dustir *ir;
    /* worker threads of the per-leaf thermal update (1 = serial) */
    int     solve_threads;
    /* cooling_model 'local_ne': (T, n_e) table of each registry stage,
     * NULL = the stage's fit unsuppressed */
    const struct orc_nlcool_s *st_tab[ORC_MAX_EL][ORC_MAX_ST];
} orc_ctx;

/* ---- n-level atoms (src/nlevel_mod.f90) ---------------------------------- */
#define ORC_NL_MAXLEV 40
#define ORC_NL_MAXTR  600
typedef struct {
    int    nlev;
    double g[ORC_NL_MAXLEV], e_cm1[ORC_NL_MAXLEV];
    int    nrad;
    int    rl[ORC_NL_MAXTR], ru[ORC_NL_MAXTR];        /* 1-based levels */
    double A[ORC_NL_MAXTR];
    int    nups;
    int    ul_l[ORC_NL_MAXTR], ul_u[ORC_NL_MAXTR], ttype[ORC_NL_MAXTR];
    int    logf[ORC_NL_MAXTR], nc[ORC_NL_MAXTR];
    double cups[ORC_NL_MAXTR], dery[ORC_NL_MAXTR], st_lo[ORC_NL_MAXTR], st_hi[ORC_NL_MAXTR];
    double coef[ORC_NL_MAXTR][16];
} orc_nlatom;

/* (log T, log n_e) cooling table of one ion (src/nlevel_cooling_mod.f90:59-85):
 * log10 T = 2 .. 6 step 0.01, log10 n_e = -6 .. 10 step 0.1; Ssup[j][i] */
#define ORC_NLC_NT  401
#define ORC_NLC_NNE 161
#define ORC_NLC_TLO 2.0
#define ORC_NLC_DT  0.01
#define ORC_NLC_NELO (-6.0)
#define ORC_NLC_DNE 0.1
typedef struct orc_nlcool_s {
    int    ready;
    double lamref[ORC_NLC_NT];
    double Ssup[ORC_NLC_NNE * ORC_NLC_NT];
} orc_nlcool;

/* a recombination-line table (src/sh95_mod.f90:24-31): etab[line][T][ne] */
#define ORC_SH95_MAXG 32
#define ORC_SH95_MAXL 12
typedef struct {
    int    nT, nD, nlin;
    double tgrid[ORC_SH95_MAXG], dgrid[ORC_SH95_MAXG];
    double wls[ORC_SH95_MAXL];
    double etab[ORC_SH95_MAXL * ORC_SH95_MAXG * ORC_SH95_MAXG];
} orc_sh95;

double orc_ups_eval(const orc_nlatom *a, int k, double T);
void   orc_emitting_level(const orc_nlatom *a, int i1, int i2, int *lo, int *hi, double *de);
void   orc_nlevel_prepT(const orc_nlatom *a, double T, double *Rcoll, double *Rrad);
double orc_nlevel_cooling_solve(const orc_nlatom *a, const double *Rcoll, const double *Rrad,
                                double ne);
double orc_nlevel_cooling_at(const orc_nlatom *a, double T, double ne);
void   orc_nlevel_emissivities(const orc_nlatom *a, double T, double ne, double *emis);
double orc_nlevel_line_ident(const orc_nlatom *a, int k);
void   orc_nlcool_build(const orc_nlatom *a, orc_nlcool *t);
double orc_nlcool_coeff(const orc_nlcool *t, double T, double ne);
double orc_nlcool_low_density(const orc_nlcool *t, double T);
double orc_nlcool_apply(const orc_nlcool *t, double lam_fit, double T, double ne);
/* the stage table of registry element ie (0-based), stage 1-based */
void   orc_set_stage_table(orc_ctx *c, int ie, int stage, const orc_nlcool *t);
void   orc_set_hi_table(orc_ctx *c, const orc_nlcool *t);
int    orc_lines(const orc_ctx *c, const orc_sh95 *const *tab, const orc_nlatom *const *atom,
                 double *L, int nthreads);
double orc_sh95_emis(const orc_sh95 *t, int k, double T, double ne);

orc_ctx *orc_create(const orc_params *p);
void     orc_destroy(orc_ctx *c);
int      orc_setup_band(orc_ctx *c);
int      orc_grid_car(orc_ctx *c);
int      orc_grid_octree(orc_ctx *c, int nleaf, const double *xl, const double *yl,
                         const double *zl, const int *lev, double half);
int      orc_state_init(orc_ctx *c, const double *nH, double xHI0, double xHeI0,
                        double xHeII0);
int      orc_set_dust_table(orc_ctx *c, int n, const double *lam, const double *alb,
                            const double *g, const double *cext);
void     orc_dust_at(const orc_ctx *c, double E, double *sabs, double *ssca, double *g);
int      orc_bin_of(const orc_ctx *c, double E);

/* one transport pass over global packet ids [first, last) (1-based ids),
 * with nthreads worker threads; tallies are zeroed first */
int      orc_transport_pass(orc_ctx *c, int64_t first, int64_t last, int nthreads);
void     orc_rates_compute(orc_ctx *c);
void     orc_equilibrium_update(orc_ctx *c, double *max_dx, double *dx_vol);
void     orc_solve_ion_cell(double gH, double gHe1, double gHe2, double nH, double T,
                            int caseA, double yHe, double *xHI, double *xHeI,
                            double *xHeII, double *ne);
void     orc_residual_grid(orc_ctx *c, double *eps_max, double *eps_vol);
void     orc_cooling_setup(orc_ctx *c, int nfit, const double *A, const double *Ti);
double   orc_net_rate(const orc_ctx *c, int il, double T, double *xHI, double *xHeI,
                      double *xHeII, double *ne);
void     orc_thermal_update(orc_ctx *c, double *max_dx, double *max_dte, double *dx_vol,
                            double *dte_vol);
/* threads of the per-leaf solves of orc_thermal_update; the leaves are
 * independent and the metrics are summed in leaf order afterwards, so the
 * result does not depend on the count */
void     orc_set_solve_threads(orc_ctx *c, int n);

/* trace metals (src/species_mod.f90) */
int      orc_species_setup(orc_ctx *c, int nel, const orc_element *els);
void     orc_species_packet_cross_sections(const orc_ctx *c, double E, double *sigma, int *ntr);
void     orc_species_fractions(const orc_ctx *c, int ie, int il, double T, double ne,
                               double nHI, double nHII, double *frac);
void     orc_species_ne_prepare(const orc_ctx *c, int il, double T, orc_mcache *cc);
void     orc_species_metal_sums(const orc_ctx *c, const orc_mcache *cc, double nH, double ne,
                                double nHI, double nHII, double *nem, double *r_cx_in,
                                double *r_cx_out);
double   orc_metal_cooling(const orc_ctx *c, int il, double T, double nH, double ne,
                           double nHI, double nHII);
double   orc_metal_cooling_H(const orc_ctx *c, int il, double T, double nH, double ne,
                             double nHI, double nHII);
double   orc_metal_freefree(const orc_ctx *c, int il, double T, double nH, double ne,
                            double nHI, double nHII);
double   orc_metal_heating(const orc_ctx *c, int il, double T, double nH, double ne,
                           double nHI, double nHII);
void     orc_species_gamma_compute(orc_ctx *c);
void     orc_species_stage_cache_refresh(orc_ctx *c);
void     orc_gas_opacity_fill(orc_ctx *c);
double   orc_charge_neutrality_ne(const orc_ctx *c, int il, double T, double nH,
                                  double ne_HHe, double xHI, double ne_seed);
/* solve_ion_cell with the leaf's metal terms (il >= 0 and metals loaded) */
void     orc_solve_ion_cell_leaf(const orc_ctx *c, int il, double gH, double gHe1, double gHe2,
                                 double nH, double T, int caseA, double yHe, double *xHI,
                                 double *xHeI, double *xHeII, double *ne);

/* dust thermal-infrared band (orc_dustir.c; the pass in orc_transport.c) */
int      orc_dust_temp_setup(orc_ctx *c);
double   orc_dust_femit_of(const orc_ctx *c, double T);
void     orc_dust_temp_compute(orc_ctx *c, const double *heat);
void     orc_dust_heat_compute(const orc_ctx *c, double *heat);
int      orc_dust_emis_transport_setup(orc_ctx *c);
double   orc_ir_leaf_spectrum(const orc_ctx *c, int il, double *pdf, double *cdf,
                              double *fdisc);
double   orc_sample_linear_bin(double lam0, double lam1, double p0, double p1, double u);
double   orc_ir_counts(orc_ctx *c, int pass, int64_t N, const double *heat, double *xi);
void     orc_ir_launch_state(const orc_ctx *c, int il, const double *pdf, const double *cdf,
                             uint64_t key, double *out7);
int      orc_ir_transport(orc_ctx *c, int pass, int64_t N, int nthreads);
int      orc_transport_dust_emission(orc_ctx *c, double *heat, int64_t npacket, int niter,
                                     double tol, int nthreads, double *report);
int      orc_dust_ir_arrays(const orc_ctx *c, double *tdust, double *heat_ir, double *irabs,
                            double *lesc, double *lam, double *sabs, double *edge,
                            double *femit);
void     orc_dustir_free(orc_dustir *ir);

/* diagnostics: leaves crossed by the direct walk of one ray (no tally) */
int      orc_trace_ray(orc_ctx *c, double x, double y, double z, double kx,
                       double ky, double kz, int *cells, double *segs, int maxn);
/* diagnostics: launch state of one stellar packet */
int      orc_launch(orc_ctx *c, int64_t id, double *out9);

/* diffuse field (src/diffuse_mod.f90): tables, per-pass emission build and
 * one diffuse launch (diagnostic; 1-based id on stream 2) */
int      orc_diffuse_setup(orc_ctx *c);
void     orc_diffuse_build(orc_ctx *c);
int      orc_launch_diffuse(orc_ctx *c, int64_t id, double *out9);
/* He I excited channel of the diffuse field (par%hei_diffuse) */
void     orc_hei_branch_fractions(double T, double *f3s, double *f1s, double *f1p);
double   orc_hei_584_ionizes_H(double T, double xHI, double xHeI);
int      orc_hei2ph_setup(orc_hei2ph *h, int n, const double *y, const double *A,
                          double eion_max);
double   orc_hei2ph_sample_energy(const orc_hei2ph *h, double u);
void     orc_hei_branch_energies(const orc_hei2ph *h, double T, double xHI, double xHeI,
                                 double *e3s, double *e1p, double *e1s);
int      orc_diffuse_hei_setup(orc_ctx *c, int n, const double *y, const double *A);

#ifdef __cplusplus
}
#endif
#endif
