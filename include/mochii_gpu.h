/*
 * mochii_gpu.h -- C-ABI of the MI355X (gfx950) photon-packet transport and
 * cell-balance library.
 *
 * The reference (seoncafe/MoCHII) is a Fortran 90 program whose hot path is a
 * set of module procedures with module-global state.  This ABI replaces them
 * at PASS granularity (one call = one transport pass or one grid-wide cell
 * update); the per-packet procedures (gen_ion_photon_qmc, transport_ion_packet,
 * score_ion_path, ...) run inside the GPU kernels.  Each entry point names the
 * reference procedure it stands in for.
 *
 * Conventions
 *   - plain C types only: pointers, sizes, doubles; no runtime-library types;
 *   - every call returns int: 0 = ok, < 0 = error; mc_last_error() describes
 *     the last error of the calling thread (the reference aborts the job via
 *     fatal_error / MPI_ABORT instead, src/utility.f90:382-391);
 *   - the library owns all device-resident state between calls; host arrays
 *     are copied in/out only by the *_set / *_get calls;
 *   - leaf arrays are indexed by the reference's 1-based leaf number il-1;
 *     2-D tallies are returned in the reference's column-major layout, i.e.
 *     hhe(6,nleaf) is 6 contiguous values per leaf, jt_ion(nnu_band,nleaf)
 *     is nnu_band contiguous values per leaf;
 *   - calls are synchronous on return and must come from one host thread
 *     per context (the reference runs one single-threaded process per rank).
 */
#ifndef MOCHII_GPU_H
#define MOCHII_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MC_MAXSRC 16

/* precision forms */
#define MC_FP64  0   /* all-FP64 transport (reference arithmetic) */
#define MC_MIXED 1   /* FP32 cell-local geometry + opacity, FP64 tallies */

/* grid types (par%grid_type) */
#define MC_GRID_CAR 0
#define MC_GRID_AMR 1

/* external-field geometry (par%ext_geometry) */
#define MC_EXT_SPH 1
#define MC_EXT_REC 2

/* diffuse-packet energy models (par%diffuse_energy_model) */
#define MC_DIFFUSE_THRESHOLD   0
#define MC_DIFFUSE_EXPONENTIAL 1
#define MC_DIFFUSE_MILNE       2

/* warm start of the thermal solve (mc_set_warm_start) */
#define MC_WARM_NONE   0   /* the reference's bisection from [te_min, te_max]  */
#define MC_WARM_NEWTON 1   /* safeguarded Newton from the leaf's current T_e    */
#define MC_WARM_MLP    2   /* safeguarded Newton from the network's T_e and x   */

/* walk order of the stellar packets of a pass (mc_set_packet_order) */
#define MC_ORDER_LAUNCH    0   /* launch (packet id) order                      */
#define MC_ORDER_DIRECTION 1   /* sorted by source and direction (default)      */

/* trace-metal registry (src/species_mod.f90:54-99; MAX_EL, MAX_ST there,
 * MAX_METAL_TRANSITIONS in src/define.f90:80) */
#define MC_MAX_EL 11
#define MC_MAX_ST 6
#define MC_MAX_METAL_TR 64
#define MC_MAX_COOLFIT 8

/* par%charge_exchange (src/define.f90:790-801) */
#define MC_CX_TWO_WAY    0   /* hydrogen's side of each reaction in its balance */
#define MC_CX_METAL_ONLY 1   /* charge exchange in the metal cascade only       */

/* par%ci_model (src/define.f90:773-776) */
#define MC_CI_VORONOV     0
#define MC_CI_DERE_HYBRID 1   /* Voronov x the Dere (2007) ratio per stage */

/* par%cooling_model (src/define.f90:781-789) */
#define MC_COOL_LOW_DENSITY 0   /* line-cooling fits unsuppressed             */
#define MC_COOL_LOCAL_NE    1   /* fits x the n-level (T, n_e) suppression
                                   (the reference default): needs the atoms
                                   of mc_nlevel_setup                        */

/* dust thermal-infrared band: per-leaf fields of mc_dust_get / mc_dust_set */
#define MC_DUST_HEAT    0   /* Heat_dust [erg/s/cm^3] (gas_rates_mod heat_dust)     */
#define MC_DUST_TEMP    1   /* T_dust [K] (dust_temp_mod t_dust)                    */
#define MC_DUST_HEAT_IR 2   /* reabsorbed-infrared heating of the last pass          */
#define MC_DUST_IR_ABS  3   /* its path tally ir_abs_path (dust_ir_band_mod)        */
/* report of mc_transport_dust_emission */
#define MC_DUSTRT_NREPORT 10
#define MC_DUSTRT_PASSES     0   /* passes run                                      */
#define MC_DUSTRT_CONVERGED  1   /* 1 when the relative change fell below dust_tol  */
#define MC_DUSTRT_L_EMIT     2   /* emitted luminosity of the last pass [erg/s]     */
#define MC_DUSTRT_EMITTED    3   /* IR ledger of the last pass: emitted,           */
#define MC_DUSTRT_EXTINCT    4   /*   extinguished,                                */
#define MC_DUSTRT_ESCAPED    5   /*   escaped                                      */
#define MC_DUSTRT_CLOSURE    6   /* |extinct + escaped - emitted| / emitted         */
#define MC_DUSTRT_L_ABS_PATH 7   /* grain-absorbed power of the path estimator      */
#define MC_DUSTRT_DREL       8   /* relative change of the total emitted luminosity */
#define MC_DUSTRT_FDISC      9   /* largest emission fraction the band cut removed  */

/* learned warm start: network 8 -> 64 -> 64 -> 64 -> 4 (GELU) */
#define MC_MLP_NPARAM 9180

/* leaf-state fields for mc_state_get / mc_state_set
 * (src/gas_state_mod.f90:23-28, rhokap from src/octree_mod.f90:50) */
#define MC_NH     0
#define MC_XHI    1
#define MC_XHEI   2
#define MC_XHEII  3
#define MC_NE     4
#define MC_TE     5
#define MC_RHOKAP 6

/* Subset of the reference's par (src/define.f90:189-861) this path reads.
 * Fortran character switches become enums/ints; logicals become int (0/1). */
typedef struct mc_params {
    int64_t nphotons;          /* par%nphotons                               */
    int32_t qmc_seed;          /* par%qmc_seed (Owen scramble seed)           */
    int32_t precision;         /* MC_FP64 | MC_MIXED                          */
    double  eion_min, eion_max;/* par%eion_min, par%eion_max [eV]             */
    double  source_cdf_tol;    /* par%source_cdf_tol                          */
    int32_t nnu_ion;           /* par%nnu_ion (diagnostic J bins)             */
    int32_t ion_align_edges;   /* par%ion_align_edges                         */
    int32_t nsource;           /* par%nsource (internal point sources)        */
    int32_t ext_on;            /* external field on (par%ext_intensity > 0)   */
    double  tstar[MC_MAXSRC];  /* Planck T per source (par%tstar/src_tstar)   */
    double  src_lum[MC_MAXSRC];/* band luminosity per source [erg/s]          */
    double  src_x[MC_MAXSRC], src_y[MC_MAXSRC], src_z[MC_MAXSRC];
    int32_t ext_geometry;      /* MC_EXT_SPH | MC_EXT_REC (par%ext_geometry)   */
    int32_t car_walk_amr;      /* par%car_walk == 'amr' on a car grid          */
    double  ext_tstar, ext_intensity, rmax;
    int32_t grid_type;         /* MC_GRID_CAR | MC_GRID_AMR                    */
    int32_t nx, ny, nz;        /* car grid dimensions                          */
    double  xmax, ymax, zmax;  /* box half sizes [code units]                  */
    double  distance2cm;       /* code length -> cm                            */
    double  He_abund;          /* n_He/n_H                                     */
    double  te_fixed, ion_relax, te_min, te_max, gas_tol, gas_tol_te;
    int32_t case_A;            /* par%case_ab == 'A'                           */
    int32_t solve_te;          /* par%solve_te                                 */
    int32_t conv_vol;          /* par%conv_crit == 'vol'                       */
    int32_t gas_niter;         /* par%gas_niter                                */
    double  xHI_init, xHeI_init, xHeII_init;
    int32_t ion_add_dust;      /* par%ion_add_dust                             */
    int32_t ion_dust_scatter;  /* par%ion_dust_scatter                         */
    double  lambda_ref;        /* par%lambda_ref [um]                          */
    int32_t diffuse_field;     /* par%diffuse_field: explicit ground-recombination
                                  packets (forces case A)                      */
    int32_t diffuse_model;     /* par%diffuse_energy_model: MC_DIFFUSE_*       */
    /* active metal photoionisation thresholds [eV] gathered by the driver
     * from the registry (src/main.f90:58-72), the ion_setup(metal_eth,
     * nmetal) arguments: band-edge alignment and sampler knots */
    int32_t nmetal;
    double  metal_eth[MC_MAX_METAL_TR];
    /* metal switches, used once mc_species_setup has loaded a registry */
    int32_t ion_metal_abs;     /* par%ion_metal_abs                            */
    int32_t metal_ne;          /* par%metal_ne                                 */
    int32_t metal_heat;        /* par%metal_heat                               */
    int32_t metal_freefree;    /* par%metal_freefree                           */
    int32_t charge_exchange;   /* par%charge_exchange: MC_CX_*                 */
    int32_t ci_model;          /* par%ci_model: MC_CI_*                        */
    int32_t cooling_model;     /* par%cooling_model: MC_COOL_*                 */
    /* dust thermal-infrared band (src/define.f90:422-433) */
    int32_t dust_niter;        /* par%dust_niter: self-heating passes           */
    double  dust_no_photons;   /* par%dust_no_photons: packets per pass         */
    double  dust_tol;          /* par%dust_tol: relative change that stops it   */
    /* par%hei_diffuse (src/define.f90:634): with diffuse_field, the He I
     * excited-level recombination channel (src/diffuse_mod.f90:95-105);
     * needs mc_diffuse_hei_setup before the first pass */
    int32_t hei_diffuse;
} mc_params;

/* One element of the trace-metal registry, the fields read_element fills
 * from data/atomic/element_<el>.txt (src/species_mod.f90:54-99, 358-422)
 * plus the stage cooling fits of cooling_<el>_<i>.txt (cooling_load,
 * src/cooling_mod.f90:98-123).  Transition it (0-based here) ionises stage
 * it to it+1; arrays over transitions use the first nstage-1 entries, the
 * cooling arrays the first nstage.  photo_form 1: VFKY96 (Eth, E0, s0, ya,
 * P, yw, y0, y1); 2: VY95 outer shell (Eth, E0, s0, ya, P, yw, l). */
typedef struct mc_element {
    char    name[4];           /* "c", "n", "o", "ne", ... (lower case)        */
    int32_t Z, nstage;
    double  abund;             /* n(X)/n(H); the element is active when > 0    */
    double  eth[MC_MAX_ST];
    int32_t photo_form[MC_MAX_ST];
    double  photo[MC_MAX_ST][8];
    double  ci[MC_MAX_ST][5];  /* Voronov dE, P, A, X, K                       */
    double  rr[MC_MAX_ST][6];  /* Badnell A, B, T0, T1, C, T2                  */
    int32_t ndr[MC_MAX_ST];
    double  dr_c[MC_MAX_ST][9], dr_e[MC_MAX_ST][9];
    int32_t has_dr2[MC_MAX_ST];
    double  dr2[MC_MAX_ST][4];
    int32_t has_rr2[MC_MAX_ST];
    double  rr2[MC_MAX_ST][2];
    int32_t cxi_form[MC_MAX_ST];
    double  cxi[MC_MAX_ST][6];
    int32_t cxr_form[MC_MAX_ST];
    double  cxr[MC_MAX_ST][6];
    int32_t ncool[MC_MAX_ST];  /* 0 = no cooling fit for the stage             */
    double  cool_A[MC_MAX_ST][MC_MAX_COOLFIT], cool_T[MC_MAX_ST][MC_MAX_COOLFIT];
} mc_element;

/* One n-level atom, the data nlevel_load reads from
 * data/atomic/nlevel_<name>_<stage>.txt (src/nlevel_mod.f90:61-108): levels
 * (statistical weight g, energy [cm^-1]), radiative rows (lower / upper
 * level, 1-based, A [s^-1]) and Burgess-Tully Chebyshev fits of the
 * collision strengths (levels, type, C, dE [Ry], st_lo, st_hi, logf, nc
 * coefficients). */
#define MC_NL_MAXLEV 40
#define MC_NL_MAXTR  600
typedef struct mc_nlevel_atom {
    char    name[4];           /* element: "h", "c", "o", ... (lower case)     */
    int32_t stage;             /* 1 = neutral                                  */
    int32_t nlev;
    double  g[MC_NL_MAXLEV], e_cm1[MC_NL_MAXLEV];
    int32_t nrad;
    int32_t rl[MC_NL_MAXTR], ru[MC_NL_MAXTR];
    double  A[MC_NL_MAXTR];
    int32_t nups;
    int32_t ul_l[MC_NL_MAXTR], ul_u[MC_NL_MAXTR], ttype[MC_NL_MAXTR];
    int32_t logf[MC_NL_MAXTR], nc[MC_NL_MAXTR];
    double  cups[MC_NL_MAXTR], dery[MC_NL_MAXTR], st_lo[MC_NL_MAXTR], st_hi[MC_NL_MAXTR];
    double  coef[MC_NL_MAXTR][16];
} mc_nlevel_atom;

typedef struct mc_ctx mc_ctx;

/* Last error message of the calling thread ("" if none). */
const char *mc_last_error(void);

/* Set-up: ion_setup (src/ion_band_mod.f90:122-258) + ionizing_field_setup
 * (src/ionizing_field_mod.f90:54-64) + jtally_ion_setup
 * (src/jtally_mod.f90:197-210) + ion_score_setup (src/ion_score_mod.f90:71-109).
 * For grid_type MC_GRID_CAR this also builds the namelist car grid
 * (src/grid_mod_amr.f90:85-108, amr_build_car src/octree_mod.f90:648-674). */
int mc_init(const mc_params *p, mc_ctx **ctx);
int mc_finalize(mc_ctx *ctx);

/* Measured ceiling of the walk's tally traffic: the device rate of
 * scattered FP64 read-modify-writes.  n_ops lanes each add 1.0 to the
 * `width` consecutive doubles (1..16) of a row chosen by a splitmix64 hash
 * of the lane's index among `rows` rows (a rows x width array allocated,
 * zeroed and freed by the call on the current device); one untimed run, then
 * *ms = the timed run's kernel time (HIP events).  checksum (may be null:
 * skips the host read-back) receives the array's sum, 2 n_ops width exactly.
 * No counterpart in the reference. */
int mc_probe_scattered_rmw(int64_t rows, int width, int64_t n_ops, double *ms,
                           double *checksum);

/* The same rows updated the way the walk's tally writers issue them: n_rows
 * row updates, the `width` doubles of a row (1, 2, 4 or 8, so a row lies in
 * one 64-byte line) added by `width` lanes of one instruction.  The device's
 * atomic unit counts requests, one per line an instruction touches, so this
 * completes about `width` times the doubles per second of
 * mc_probe_scattered_rmw at the same row rate.  *ms, checksum (2 n_rows width)
 * as there.  No counterpart in the reference. */
int mc_probe_row_atomics(int64_t rows, int width, int64_t n_rows, double *ms, double *checksum);

/* Leaf list of an AMR input file, generic_amr_read
 * (src/read_generic_amr.f90:34-237) with the file-name resolution of
 * io_resolve_filename (src/iofile_mod.f90:138-170): the text form ("nleaf
 * boxlen" then rows "x y z level nH [...]") or a FITS binary table (columns
 * by name x, y, z, level, nH|dens|gasDen with the reference's index
 * fallback; keywords BOXLEN, ORIGINX/Y/Z), gzip-compressed or not; HDF5 is
 * refused.  Coordinates come back recentred on the box centre as
 * grid_create_amr does (src/grid_mod_amr.f90:51-61), *half = BOXLEN / 2,
 * ready for mc_grid_octree.  With x null or maxn < the leaf count the call
 * only reports *nleaf, *half and has_opt; otherwise it fills the arrays.
 * has_opt[0..2] (may be null): the optional metallicity / xHI / ndust
 * columns are present (their arrays may be null).  Host only, no device. */
int mc_leaf_list_read(const char *path, int64_t maxn, int64_t *nleaf, double *half,
                      double *x, double *y, double *z, int32_t *level, double *nH,
                      double *metallicity, double *xHI, double *ndust, int32_t *has_opt);

/* Octree from a leaf list: amr_build_tree + amr_build_neighbors
 * (src/octree_mod.f90:347-497) as called from grid_create_amr
 * (src/grid_mod_amr.f90:216-219).  Leaf coordinates are box-centred. */
int mc_grid_octree(mc_ctx *ctx, int nleaf, const double *xleaf, const double *yleaf,
                   const double *zleaf, const int32_t *level, double half);

/* gas_state_setup (src/gas_state_mod.f90:41-90): nH per leaf; xHI may be
 * NULL (then par%xHI_init); He fractions from par%xHeI_init/xHeII_init. */
int mc_gas_state_setup(mc_ctx *ctx, const double *nH, const double *xHI);

/* Leaf-state access (the module arrays of gas_state_mod / amr_grid%rhokap). */
int mc_state_set(mc_ctx *ctx, int field, const double *src);
int mc_state_get(mc_ctx *ctx, int field, double *dst);

/* ion_dust_read (src/gas_opacity_mod.f90:69-205): (lambda[um], albedo, <cos>,
 * C_ext/H) rows in any order; the library sorts them. */
int mc_ion_dust_table(mc_ctx *ctx, int n, const double *lam, const double *albedo,
                      const double *gcos, const double *cext);

/* gas_opacity_fill (src/gas_opacity_mod.f90:299-332): refresh the per-leaf
 * opacity inputs from the current state. */
int mc_gas_opacity_fill(mc_ctx *ctx);

/* ionizing_field_and_rates (src/ionizing_field_mod.f90:72-170): one transport
 * pass of all par%nphotons stellar packets and, with diffuse_field on, of the
 * diffuse ground-recombination packets built from the current state
 * (diffuse_build + gen_diffuse_photon_qmc, src/diffuse_mod.f90:60-296, on
 * the second Sobol stream), reduction, gas_rates_compute and the
 * energy-closure check (returns an error if closure > 5e-13, as
 * ion_score_report_energy_closure aborts, src/ion_score_mod.f90:364-387).
 * with_peel is accepted for signature parity and must be 0 (peel-off imaging
 * is not on this path). */
int mc_ionizing_field_and_rates(mc_ctx *ctx, int with_peel);

/* Diffuse field of the last pass: luminosity diffuse_lum [erg/s] and packet
 * count diffuse_nphot (src/diffuse_mod.f90:38-39, printed by
 * ionizing_field_and_rates at src/ionizing_field_mod.f90:141-143). */
int mc_diffuse_info(mc_ctx *ctx, double *lum, int64_t *nphot);

/* hei_2ph_setup (src/hei_twophoton_mod.f90:98-183), called by src/main.f90:
 * 107-112 when diffuse_field and hei_diffuse are on: the He I 2^1S two-photon
 * shape A(y), y = E / 20.62 eV, as the n rows (4 <= n <= 128, ascending y) of
 * data/HeI2phot.dat.  Builds the photon and energy moments of the H-ionizing
 * window [13.598 eV, min(20.62 eV, eion_max)] and the energy-weighted sampling
 * CDF on 1024 nodes that channel 4 of the diffuse field draws from.  The
 * caller reads the file (the rows up to the first that is not two numbers). */
int mc_diffuse_hei_setup(mc_ctx *ctx, int n, const double *y, const double *A);

/* The two-photon moments of the last mc_diffuse_hei_setup: photons and
 * energy [eV] per 2^1S decay inside the window (hei_2ph_photons_per_decay,
 * hei_2ph_energy_per_decay, src/hei_twophoton_mod.f90:67-76), and, when
 * grid_y / cdf_e are not NULL, the 1024 sampling nodes (y) and their CDF. */
int mc_diffuse_hei_info(mc_ctx *ctx, double *n_per_decay, double *e_per_decay, double *grid_y,
                        double *cdf_e);

/* gas_equilibrium_update (src/ion_balance_mod.f90:352-434). */
int mc_gas_equilibrium_update(mc_ctx *ctx, double *max_dx, double *dx_vol);

/* cooling_setup (src/cooling_mod.f90:67-92): the H I line-cooling fit
 * Lambda(T) = T^-1/2 sum_i A_i exp(-T_i/T) (nfit <= 8 terms, the rows of
 * data/atomic/cooling_h_1.txt) and the free-free Gaunt table, which the
 * library builds itself.  With cooling_model MC_COOL_LOCAL_NE the fit is
 * suppressed by the H I atom's (T, n_e) table (mc_nlevel_setup). */
int mc_cooling_setup(mc_ctx *ctx, int nfit, const double *A, const double *Ti);

/* The n-level atoms the library may use (nlevel_load for every ion,
 * src/nlevel_mod.f90:61-108).  With cooling_model MC_COOL_LOCAL_NE the
 * thermal update suppresses the H I line-cooling fit and each metal stage's
 * fit by the ion's (T, n_e) table, built here as nlevel_cooling_add builds
 * it (src/nlevel_cooling_mod.f90:105-152: 401 x 161 statistical-equilibrium
 * solves per ion, cooling_setup and species_setup call it for H I and for
 * every stage with a cooling fit); an ion without an atom keeps its fit
 * unsuppressed, as in the reference.  The atoms also define the lines of
 * mc_lines_compute.  Replaces any earlier set. */
int mc_nlevel_setup(mc_ctx *ctx, int n, const mc_nlevel_atom *atoms);

/* One recombination-line table as sh95_load reads it (src/sh95_mod.f90:93-136):
 * nT temperatures and nD electron densities (ascending), nlin lines with
 * vacuum wavelengths [A] and 4 pi j / (n_e n_ion) [erg cm^3 s^-1] on the
 * grid, etab[line][iT][iD]. */
#define MC_LT_MAXG 32
#define MC_LT_MAXL 12
typedef struct mc_line_table {
    int32_t nT, nD, nlin;
    double  tgrid[MC_LT_MAXG], dgrid[MC_LT_MAXG];
    double  wl[MC_LT_MAXL];
    double  etab[MC_LT_MAXL][MC_LT_MAXG][MC_LT_MAXG];
} mc_line_table;

/* sh95_setup (src/sh95_mod.f90:66-90): the H I and He II recombination-line
 * tables of the case in use and the He I (Porter, case B) table; hi is
 * required by mc_lines_compute, heii and hei may be NULL (no such lines). */
int mc_lines_setup(mc_ctx *ctx, const mc_line_table *hi, const mc_line_table *heii,
                   const mc_line_table *hei);

/* lines_write (src/lines_mod.f90:37-352) on the current state: the line
 * luminosities L = sum over leaves of j n_ion V [erg/s], in the reference's
 * table order: the H I recombination lines (n_e n_p, ion code 101), the He II
 * lines (n_e n_HeIII, 202), the He I lines (n_e n_HeII, 201), then for every
 * registry element and stage with an n-level atom (mc_nlevel_setup) its
 * radiative rows (n-level solve per leaf at (T_e, n_e), n_ion from the stage
 * fractions at the leaf's state and last metal rates; code 100 Z + stage).
 * Writes at most maxn entries of L, the wavelengths wl [A] and the ion codes,
 * and the number of lines to *n (any output pointer may be NULL).  The
 * collisional H I Balmer rows (par%h_coll_effects, off by default) are not
 * computed. */
int mc_lines_compute(mc_ctx *ctx, int maxn, double *L, double *wl, int32_t *ion, int32_t *n);

/* gas_thermal_update (src/thermal_mod.f90:124-249): per-leaf T_e by
 * bisection on log T with the ionisation state solved at each trial T. */
int mc_gas_thermal_update(mc_ctx *ctx, double *max_dx, double *max_dte, double *dx_vol,
                          double *dte_vol);

/* Warm start of the per-leaf solve in mc_gas_thermal_update (MC_WARM_*).
 * No counterpart in the reference (whose solve is the MC_WARM_NONE
 * bisection): every warm mode keeps the reference's edge tests and accepts
 * only a sign-change bracket of the same width thi/tlo - 1 < 1e-5, falling
 * back to bisection from the bracket it has when Newton stalls.  With the
 * metal registry and two-way charge exchange, a warm mode also ends a trial
 * solve's n_e iteration after 2 consecutive passes in which the leaf's
 * hydrogen is unresolved (x_HII < 1e-10, leaves without ionising flux): there
 * the reference's iteration wanders in a noise band of n_e until its
 * 200-pass cap. */
int mc_set_warm_start(mc_ctx *ctx, int mode);

/* Bisection levels a warm start runs first (default MC_WARM_LEVELS_DEFAULT,
 * 0..24): the reference's bisection from [te_min, te_max] for that many
 * levels, then the warm solve inside the level's cell.  A leaf whose net
 * rate crosses zero more than once then ends on the bisection's crossing,
 * or on another one inside that cell: within a factor
 * (te_max/te_min)^(2^-levels) of it in T_e (0.77% for 9 levels over
 * [1e3, 5e4] K).  No counterpart in the reference. */
#define MC_WARM_LEVELS_DEFAULT 9
int mc_set_warm_levels(mc_ctx *ctx, int levels);

/* Work of the last thermal update: {net-rate evaluations spent on the
 * bracketed leaves (those not pinned at te_min / te_max, whose two edge
 * evaluations settle them), bracketed leaves that fell back to bisection,
 * bracketed leaves, all leaves}.  No counterpart in the reference. */
int mc_thermal_stats(mc_ctx *ctx, int64_t *out4);

/* ---- learned warm start (no counterpart in the reference) -----------------
 * Parameter blob (float, MC_MLP_NPARAM): W1[8][64] b1[64] W2[64][64] b2[64]
 * W3[64][64] b3[64] W4[64][4] b4[4] in_mu[8] in_sd[8] out_mu[4] out_sd[4];
 * W[k][n] connects input k to unit n.  Features per leaf (float[8]): log10 of
 * n_H, Gamma_HI, Gamma_HeI, Gamma_HeII, H_HI, H_HeI, H_HeII, n_e (rates as in
 * mc_rates_get, floored at 1e-30 / 1e-40); targets (float[4]): log10 x_HI,
 * x_HeI, x_HeII, T_e.  Standardisation (x - mu) / sd is part of the blob. */
int mc_mlp_set(mc_ctx *ctx, const float *params, int nparam);
int mc_mlp_get(mc_ctx *ctx, float *params);
/* Features of every leaf from the current state and rates (nleaf x 8). */
int mc_mlp_features(mc_ctx *ctx, float *X);
/* Adam training on the device (FP32 forward/backward kernels): steps
 * minibatches of `batch` rows drawn without replacement from (X, Y) (n rows),
 * learning rate lr; loss_hist (may be NULL) receives the minibatch MSE of
 * the standardised outputs per step. */
int mc_mlp_train(mc_ctx *ctx, const float *X, const float *Y, int64_t n, int steps, int batch,
                 float lr, uint32_t seed, float *loss_hist);
/* One FP32 forward/backward over n rows: loss and the gradient in the
 * parameter layout (standardisation entries 0). */
int mc_mlp_grad(mc_ctx *ctx, const float *X, const float *Y, int n, float *loss, float *grad);
/* BF16 matrix-core inference over n feature rows -> n x 4 targets. */
int mc_mlp_infer(mc_ctx *ctx, const float *X, int n, float *out);

/* ---- trace metals ----------------------------------------------------------
 * species_setup (src/species_mod.f90:143-228) + species_rate_path_setup
 * (:308-319): load the registry (n_elements entries with abund > 0, in the
 * reference's element order c n o ne s ar mg fe si cl ca).  Turns the metal
 * path on (par%use_metals): metal opacity in the walk (ion_metal_abs), the
 * raw(2,ntr,nleaf) path tallies, species_gamma_compute after each pass, the
 * metal electrons / charge exchange in the cell solve and the metal cooling
 * and heating terms of the thermal balance.  Refreshes the stage cache when
 * the gas state is already set (species_stage_cache_refresh via
 * gas_opacity_fill, :247-261).  n_elements = 0 turns the metal path off. */
int mc_species_setup(mc_ctx *ctx, int n_elements, const mc_element *elements);

/* Stage fractions stage_frac(ie)%g(nstage, nleaf) of registry element ie
 * (0-based), column-major, as of the last stage-cache refresh
 * (src/species_mod.f90:247-261). */
int mc_species_fractions_get(mc_ctx *ctx, int ie, double *frac);

/* Metal photoionisation and photoheating rates of the last pass:
 * egam(ie)%g(nstage-1, nleaf) [s^-1] and eheat(ie)%g [erg/s]
 * (species_gamma_compute, src/species_mod.f90:493-512); either may be NULL. */
int mc_species_rates_get(mc_ctx *ctx, int ie, double *gam, double *heat);

/* Metal path tallies of the last pass, raw(2, nstage-1, nleaf) of element ie
 * (metal_rate_path(ie)%raw, src/species_mod.f90:102-108, 329-343). */
int mc_species_tallies_get(mc_ctx *ctx, int ie, double *raw);

/* Metal transitions below the packet energy summed over the scored segments
 * of the last pass: each scored segment updates 2 raw entries per such
 * transition (the metal part of the walk's algorithmic tally traffic).  No
 * counterpart in the reference. */
int mc_metal_segment_transitions(mc_ctx *ctx, int64_t *sum_ntr);

/* Global FP64 atomic requests the walk issued for its tallies in the last
 * pass (stellar and diffuse batches: rate sums, band J, grain absorption,
 * metal path tallies, energy moments; after the wave merges and the moment
 * cache): one per 64-byte line an atomic instruction touches, the unit the
 * device's atomic throughput is counted in (the lanes of one instruction
 * that add into one line cost one request).  The walk's achieved atomic rate
 * is this over the walk time.  No counterpart in the reference. */
int mc_walk_atomics(mc_ctx *ctx, int64_t *n);

/* hydrogen_balance_residual_grid (src/ion_balance_mod.f90:313-343). */
int mc_hydrogen_balance_residual_grid(mc_ctx *ctx, double *eps_max, double *eps_vol);

/* Tallies of the last pass in the reference's layouts: hhe(6,nleaf)
 * (ion_score_mod hhe_rate_path), jt_ion(nnu_band,nleaf) (jtally_mod), dust_abs
 * (nleaf), segment counts {direct, scattered} (any pointer may be NULL). */
int mc_tallies_get(mc_ctx *ctx, double *hhe, double *jt_ion, double *dust_abs,
                   int64_t *segments);

/* Rates of the last pass: gamma_HI, gamma_HeI, gamma_HeII, heat_HI, heat_HeI,
 * heat_HeII (src/gas_rates_mod.f90:36-37), 6 arrays of nleaf back to back. */
int mc_rates_get(mc_ctx *ctx, double *rates6);

/* Energy ledger of the last pass {emitted, absorbed, escaped} [erg/s]
 * (ion_score_energy_totals, src/ion_score_mod.f90:357-362). */
int mc_energy_ledger(mc_ctx *ctx, double *out3);

/* Sizes: nleaf, nnu_band; band total ion_Ltot; band edges ion_eedge(nnu+1)
 * (eedge may be NULL). */
int mc_info(mc_ctx *ctx, int32_t *nleaf, int32_t *nnu_band, double *Ltot, double *eedge);

/* Walk order of the stellar packets (MC_ORDER_*).  The tallies are the same
 * sums in either order (to FP64 summation order): with MC_ORDER_DIRECTION
 * the launch records are sorted by source and direction so that a wavefront
 * walks 64 packets through the same leaves and sums their rate tallies
 * before the atomics.  No counterpart in the reference (a CPU rank walks one
 * packet at a time). */
int mc_set_packet_order(mc_ctx *ctx, int order);

/* Launch records per batch of a pass, 124 B of device memory each (halved
 * down to 1M when the device cannot hold them).  A pass launches and walks
 * its packets in batches of this many records; with direction order a batch
 * is sorted as a whole, so larger batches give each wavefront packets of one
 * emitter, sub-bin and direction when many emitters share a pass.  Defaults,
 * set by mc_init from the precision form: MC_MIXED 2^30 (a pass of up to
 * 1.07e9 packets is one batch, at most 133 GB: its energy-moment walk merges
 * a wave's tallies in LDS, so tighter waves issue fewer atomics), MC_FP64
 * 2^26 (its per-coefficient atomics slow down when many waves update the
 * same rows at once: measured 5.18 s per pass at 2^26 against 6.17 s at
 * 2^30 on the 2e7-leaf multifractal model).  The tallies are the same sums
 * for any batch size (to FP64 summation order).  Lower it to bound the
 * records' memory.  max_records >= 1024.  No counterpart in the reference. */
#define MC_LAUNCH_BATCH_MIXED ((int64_t)1 << 30)
#define MC_LAUNCH_BATCH_FP64 ((int64_t)1 << 26)
int mc_set_launch_batch(mc_ctx *ctx, int64_t max_records);

/* Issue form of the mixed walk's energy-moment rows.  The device's atomic
 * unit counts requests, one per 64-byte line an atomic instruction touches.
 * MC_ROWS_LANE: each lane adds its row's MC_MOM_NM sums with atomics of its
 * own (MC_MOM_NM requests a row; the fewest instructions and registers).
 * MC_ROWS_TOGETHER: the rows of the wave's lanes go out transposed, several
 * lanes adding one row in one instruction (one request a row).
 * MC_ROWS_AUTO (default): per-lane until a pass's request rate comes near the
 * device's ceiling, then transposed (the host decides between passes from the
 * measured rate).  The sums are the same in every form (to FP64 summation
 * order).  No counterpart in the reference. */
#define MC_ROWS_AUTO 0
#define MC_ROWS_LANE 1
#define MC_ROWS_TOGETHER 2
int mc_set_row_issue(mc_ctx *ctx, int mode);

/* The caller's MC_ROWS_* mode and the form the last pass used: 1 transposed,
 * 0 per lane, -1 no moment rows (FP64 form, or moments off).  Diagnostic. */
int mc_row_issue(mc_ctx *ctx, int *mode, int *last);

/* Energy-moment tallies of the MC_MIXED walk (on by default): a scored
 * segment adds its path weight times the Chebyshev polynomials T_0..T_3 of
 * its packet's position in an ln-E sub-bin of its band bin to four per-leaf
 * moments instead of one atomic per nonzero rate coefficient (the H/He
 * coefficients, every metal transition's two and, with grains, the grain
 * absorption cross section, whose table rows are sub-bin edges); after the walk
 * each coefficient sum is rebuilt from the moments with the coefficient's
 * interpolant on the sub-bin, which matches it to a relative 1e-7 (the
 * sub-bins are chosen for that on the host).  Off, or when a band bin cannot
 * be resolved, or when the moment rows would not fit the device memory
 * budget, the walk keeps the per-coefficient sums; the MC_FP64 form always
 * does.  No counterpart in the reference. */
int mc_set_tally_moments(mc_ctx *ctx, int on);

/* Sub-bins of the energy-moment tallies over the band (0: not used with this
 * band and registry) and the largest relative interpolation error of any
 * coefficient on its check points.  No counterpart in the reference. */
int mc_tally_moment_info(mc_ctx *ctx, int32_t *nsub, double *fit_err);

/* Switch the transport precision form (MC_FP64 | MC_MIXED) between passes. */
int mc_set_precision(mc_ctx *ctx, int precision);

/* Device times of the last pass/update [ms]: {launch, walk, rates, solve}
 * measured with events on the library's stream. */
int mc_timing(mc_ctx *ctx, double *out4);

/* ---- dust thermal-infrared band --------------------------------------------
 * Heat_dust is device-resident: every ionizing pass with ion_add_dust on
 * refreshes it from the dust-absorption path tally (ion_score_apply_dust_heat,
 * src/ion_score_mod.f90:319-338); mc_dust_set replaces it. */

/* dust_temp_setup (src/dust_temp_mod.f90:39-118): the equilibrium emission
 * function femit(T) of the extinction table loaded by mc_ion_dust_table. */
int mc_dust_temp_setup(mc_ctx *ctx);

/* dust_temp_compute (src/dust_temp_mod.f90:145-177): T_dust of every leaf
 * from the current Heat_dust. */
int mc_dust_temp_compute(mc_ctx *ctx);

/* dust_emis_transport_setup (src/dust_emis_transport_mod.f90:100-180) for the
 * equilibrium modified blackbody (par%dust_sed off): the 0.1-3000 um
 * emission grid, its cut below the band floor and every active gas
 * threshold, and the escaped-spectrum bins (dust_ir_band_setup,
 * src/dust_ir_band_mod.f90:63-84).  Needs ion_add_dust and the femit table. */
int mc_dust_emis_transport_setup(mc_ctx *ctx);

/* transport_dust_emission (src/dust_emis_transport_mod.f90:208-366): up to
 * dust_niter passes of dust_no_photons packets launched from every leaf in
 * proportion to Heat_dust V (stratified counts, continuous wavelengths),
 * carried by the shared walk with grain absorption and scattering and the
 * infrared estimators (src/ion_score_mod.f90:149-158, 236-244); after each
 * pass Heat_dust = EUV part + reabsorbed infrared and T_dust follows.
 * report: MC_DUSTRT_NREPORT values (may be NULL).  Returns an error when the
 * infrared energy closure exceeds 5e-13 (the reference aborts). */
int mc_transport_dust_emission(mc_ctx *ctx, double *report);

/* Per-leaf dust fields (MC_DUST_*) of the band; mc_dust_set accepts
 * MC_DUST_HEAT only. */
int mc_dust_get(mc_ctx *ctx, int field, double *dst);
int mc_dust_set(mc_ctx *ctx, int field, const double *src);

/* dust_ir_energy_budget (src/dust_ir_band_mod.f90:195-199): the infrared
 * ledger of the last pass {emitted, extinguished, escaped} [erg/s]. */
int mc_dust_ir_energy_budget(mc_ctx *ctx, double *out3);

/* dust_ir_escaped_spectrum (src/dust_ir_band_mod.f90:204-214): bin centres
 * lam [um], L_lambda [erg/s/um] and bin widths of the escaped infrared, nl
 * values each (any pointer may be NULL). */
int mc_dust_ir_escaped_spectrum(mc_ctx *ctx, double *lam, double *Llam, double *dlam);

/* The band's set-up: emission grid size *nl, its nodes lam [um] and s_abs
 * (nl each), the escaped-spectrum bin edges (nl + 1), the number of leading
 * nodes the band cut zeroes, the shortest emitted wavelength, and the femit
 * table {tgrid(61), femit(61)} (pointers may be NULL).  Diagnostic. */
int mc_dust_ir_grid(mc_ctx *ctx, int32_t *nl, double *lam, double *sabs, double *edge,
                    int32_t *nzero, double *lam_hardest, double *femit122);

/* Device times of the last mc_transport_dust_emission [ms]: {launch, walk,
 * other}, and its segment counts {direct, scattered}. */
int mc_dust_ir_timing(mc_ctx *ctx, double *out3, int64_t *segments2);

/* qmc_uniforms_stream (src/qmc_mod.f90:200-260): the launch's Owen-scrambled
 * Sobol uniforms u(d) = (code + 0.5) 2^-32 of stream 1 (stellar) or 2
 * (diffuse) for the n consecutive 0-based Sobol indices first_index ..,
 * dims 0 .. ndim-1 (ndim <= 12), computed on the device by the function the
 * launch kernels use; out is row-major n x ndim.  Diagnostic. */
int mc_qmc_uniforms(mc_ctx *ctx, int stream, int64_t first_index, int n, int ndim,
                    double *out);

/* gen_ion_photon_qmc (src/ion_band_mod.f90:1880-1935) for the stellar packet
 * ids first_id .. first_id + n - 1 (1-based, launch order) as the pass's
 * launch kernel computes them: rec is n x 7 {x, y, z, kx, ky, kz, E [eV]},
 * bin_leaf n x 2 {0-based band bin, 0-based start leaf or -1 outside}.
 * n is at most the launch batch (mc_set_launch_batch).  Diagnostic. */
int mc_launch_records(mc_ctx *ctx, int64_t first_id, int n, double *rec, int32_t *bin_leaf);

/* Diagnostic: leaves crossed by the direct walk of one ray to the box edge
 * (geometry only), in the given precision form.  Writes up to maxn leaf
 * numbers (1-based) to cells and their count to *n. */
int mc_trace_ray(mc_ctx *ctx, int precision, const double *pos3, const double *dir3,
                 int32_t *cells, int maxn, int32_t *n);

/* mc_trace_ray for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  Same arguments and result as
 * mc_trace_ray, but the cell list is staged in a library-owned static
 * device array of MC_TRACE_NOALLOC_MAX cells (one per GPU, held until the
 * process ends) instead of a device buffer allocated per call: the call
 * makes no hipMalloc / hipFree.  maxn above MC_TRACE_NOALLOC_MAX is an
 * error.  Calls that overlap on one GPU share that array and race; the
 * array keeps the last call's cells. */
#define MC_TRACE_NOALLOC_MAX 65536
int mc_trace_ray_noalloc(mc_ctx *ctx, int precision, const double *pos3, const double *dir3,
                         int32_t *cells, int maxn, int32_t *n);

#ifdef __cplusplus
}
#endif
#endif
