// Internal declarations of the gfx950 transport / cell-balance library.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <map>
#include <vector>

#include "../../include/mochii_gpu.h"
#include "mc_nlevel.h"

#define MC_MAXBINS 64
#define MC_QMC_DIM 12

// Physical constants (src/define.f90:12-88 of the reference).
#define MC_PI      3.141592653589793238462643383279502884197
#define MC_TWOPI   6.283185307179586476925286766559005768394
#define MC_FOURPI  12.56637061435917295385057353311801153679
#define MC_EV2ERG  1.602176634e-12
#define MC_KBOLTZ  1.380649e-16
#define MC_ETH_HI   13.598
#define MC_ETH_HEI  24.587
#define MC_ETH_HEII 54.416
#define MC_TAU_HUGE 745.2

// Ground-state photoionisation cross sections [cm^2] at E [eV]
// (src/photo_xsec.f90:46-126): hydrogenic (Osterbrock & Ferland eq. 2.4) for
// H I and He II, the Verner et al. (1996) fit for He I.  Host and device use
// the same code (the Milne tables are built on the host).
__host__ __device__ inline double sigma_hydrogenic(double E, double Eth, double Z)
{
    const double s0 = 6.30e-18 / (Z * Z);
    if (E < Eth) return 0.0;
    if (E == Eth) return s0;
    const double eps = sqrt(E / Eth - 1.0), r = Eth / E;
    return s0 * (((r * r) * r) * r) * exp(4.0 - 4.0 * atan(eps) / eps) /
           (1.0 - exp(-2.0 * MC_PI / eps));
}

__host__ __device__ inline double sigma_HeI(double E)
{
    if (E < MC_ETH_HEI) return 0.0;
    const double x = E / 13.61 - 0.4434, z = sqrt(x * x + 2.136 * 2.136);
    const double Q = 5.5 - 0.5 * 3.188, xm = x - 1.0;
    const double F = (xm * xm + 2.039 * 2.039) * pow(z, -Q) * pow(1.0 + sqrt(z / 1.469), -3.188);
    return 949.2 * F * 1.0e-18;
}

// Milne ground-continuum tables (src/milne_recomb_spectrum_mod.f90):
// energy-luminosity CDF in x = (E - E_th)/kT on MC_MILNE_NX quadratically
// clustered nodes for MC_MILNE_NT log-spaced temperatures, per channel
// (0 H I, 1 He I, 2 He II), plus <x> under the photon PDF and the in-band
// photon fraction.
#define MC_MILNE_NT 192
#define MC_MILNE_NX 512
struct DevMilne {
    const double *cdf;       // [3][NT][NX]
    const double *xbar;      // [3][NT]
    const double *fin;       // [3][NT]
    const double *xg;        // [NX]
    double lgT0, lgT1;       // first two log10 T nodes
    double eth[3];
    double eion_max;
    int32_t model;           // 0 threshold, 1 exponential, 2 milne
    // channel 3, He I excited-level recombination (par%hei_diffuse): the
    // two-photon sampling nodes y and their energy CDF, and the energy per
    // decay inside the H-ionizing window [eV]; hei = 0 leaves the channel off
    const double *hei_y;     // [MC_HEI2PH_NCDF]
    const double *hei_cdf;   // [MC_HEI2PH_NCDF]
    double hei_eper;
    int32_t hei;
};

// He I 2^1S two-photon spectrum (src/hei_twophoton_mod.f90): table rows at
// most, sampling nodes over the ionizing window, line and pair energies [eV]
#define MC_HEI2PH_MAXPTS 128
#define MC_HEI2PH_NCDF 1024
#define MC_HEI_E_TOT 20.62
#define MC_HEI_E3S 19.82
#define MC_HEI_E1P 21.22

// the device's atomic request rate: one request per 64-byte line an atomic
// instruction touches, ~2.37e10 per second on MI355X for FP64, FP32 and
// integer adds alike wherever the lines live (L2, MALL or HBM: 1.8e10 for a
// 4 GB footprint), measured with the scattered-RMW probes
#define MC_ATOMIC_REQ_PER_S 2.3e10

// Learned warm start: network sizes (the FP32 parameter blob length
// MC_MLP_NPARAM is part of the C-ABI)
#define MC_MLP_IN 8
#define MC_MLP_HID 64
#define MC_MLP_OUT 4
struct MlpState {
    float *P = nullptr;        // device parameter blob
    bool ready = false;
};

// Trace-metal registry on the device (species_mod module state,
// src/species_mod.f90:54-128), read with wave-uniform indices.  Transitions t
// are in element order (element by element, stage by stage), stages s
// likewise; flat_t lists the transitions by threshold (the packet cache
// order of species_packet_cross_sections).  The cell-solve kernels keep the
// (leaf, T) rate cache in registers with loops unrolled to MC_GPU_MT_FAST
// transitions for registries up to that size (C N O Ne S Ar have 15), and to
// MC_GPU_MAXTR for larger ones (the reference's eleven elements C N O Ne S Ar
// Mg Fe Si Cl Ca have 32; that cache spills to scratch).
#define MC_GPU_MAXTR 32
#define MC_GPU_MT_FAST 16
// Registry layouts with compiled-in element structure (transitions per
// element in registry order); any other registry takes the generic path
#define MC_LAYOUT_GENERIC 0
#define MC_LAYOUT_CNONeSAr 1    // C N O Ne S Ar: 3 2 2 2 3 3 transitions
#define MC_LAYOUT_CNONeS 2      // C N O Ne S:    3 2 2 2 3
#define MC_GPU_MAXST (MC_GPU_MAXTR + MC_MAX_EL)
struct DevMetalTab {
    int32_t nel, ntr, nst;
    int32_t t_first[MC_GPU_MAXTR];   // first transition of t's element
    int32_t t_last[MC_GPU_MAXTR];    // t is its element's last transition
    int32_t t_it[MC_GPU_MAXTR];      // transition index within the element
    int32_t t_st[MC_GPU_MAXTR];      // stage s of t's lower (ionised-from) stage
    int32_t t_flat[MC_GPU_MAXTR];    // position of t in threshold order
    int32_t t_cII[MC_GPU_MAXTR];     // t = C I -> C II (its upper stage is C II)
    int32_t t_oI[MC_GPU_MAXTR];      // t = O I -> O II (its lower stage is O I)
    double  t_ab[MC_GPU_MAXTR];      // abundance of t's element
    double  eth[MC_GPU_MAXTR];
    int32_t photo_form[MC_GPU_MAXTR];
    double  photo[MC_GPU_MAXTR][8];
    double  ci[MC_GPU_MAXTR][5];
    double  ci_ratio[MC_GPU_MAXTR];
    int32_t has_rr2[MC_GPU_MAXTR], ndr[MC_GPU_MAXTR], has_dr2[MC_GPU_MAXTR];
    double  rr[MC_GPU_MAXTR][6], rr2[MC_GPU_MAXTR][2];
    double  dr_c[MC_GPU_MAXTR][9], dr_e[MC_GPU_MAXTR][9], dr2[MC_GPU_MAXTR][4];
    int32_t cxi_form[MC_GPU_MAXTR], cxr_form[MC_GPU_MAXTR];
    double  cxi[MC_GPU_MAXTR][6], cxr[MC_GPU_MAXTR][6];
    int32_t flat_t[MC_GPU_MAXTR];
    double  flat_eth[MC_GPU_MAXTR];
    double  el_ab[MC_MAX_EL];        // abundance of each element
    int32_t s_ncool[MC_GPU_MAXST];
    int32_t s_nl[MC_GPU_MAXST];      // (T, n_e) table slot of the stage, -1: none
    int32_t s_charge[MC_GPU_MAXST];  // stage j of its element has charge j - 1
    double  s_A[MC_GPU_MAXST][MC_MAX_COOLFIT], s_T[MC_GPU_MAXST][MC_MAX_COOLFIT];
};

// Metal photoionisation cross sections [cm^2] (src/photo_xsec.f90:46-89):
// VFKY96 and VY95 forms; host (band cut of the infrared band) and device
// (packet cross sections) use the same code.
__host__ __device__ inline double m_vfky96(double E, const double *p)
{
    if (E < p[0]) return 0.0;
    const double x = E / p[1] - p[6];
    const double z = sqrt(x * x + p[7] * p[7]);
    const double Q = 5.5 - 0.5 * p[4];
    const double xm = x - 1.0;
    const double Fy = (xm * xm + p[5] * p[5]) * pow(z, -Q) * pow(1.0 + sqrt(z / p[3]), -p[4]);
    return p[2] * Fy * 1.0e-18;
}

__host__ __device__ inline double m_vy95(double E, const double *p)
{
    if (E < p[0]) return 0.0;
    const double y = E / p[1];
    const double Q = 5.5 + p[6] - 0.5 * p[4];
    const double ym = y - 1.0;
    const double Fy = (ym * ym + p[5] * p[5]) * pow(y, -Q) * pow(1.0 + sqrt(y / p[3]), -p[4]);
    return p[2] * Fy * 1.0e-18;
}

// species_sigma (src/species_mod.f90:763-776)
__host__ __device__ inline double m_sigma(const DevMetalTab *mt, int t, double E)
{
    return mt->photo_form[t] == 2 ? m_vy95(E, mt->photo[t]) : m_vfky96(E, mt->photo[t]);
}

// Small launch/band tables shared by every kernel of a context
// (read through the scalar/constant cache).
struct DevTables {
    uint32_t V[32][MC_QMC_DIM];      // Sobol direction numbers
    uint32_t key[MC_QMC_DIM][2];     // Owen keys (stellar, diffuse)
    double   eedge[MC_MAXBINS + 1];  // band edges
    double   src_cdf[MC_MAXSRC + 1];
    double   src_pos[MC_MAXSRC + 1][3];
    int32_t  comp_kind[MC_MAXSRC + 1];
    int32_t  samp_off[MC_MAXSRC + 2];// offsets of each component's knots
    int32_t  nnu, ncomp, multi, nsource, qmc_ndim, ext_geometry;
    double   Lpacket;
    double   rmax;
};

// Grid description passed by value to kernels.  The octree node table is in
// breadth-first order: the 8 children of a node are contiguous and the top
// levels form a prefix of the table.
struct DevGrid {
    int32_t car;                      // 1 = uniform Cartesian (DDA), 0 = octree
    int32_t car_walk_amr;             // car grid walked with the face-exit walk
    int32_t nx, ny, nz, nleaf, nnode;
    double  xmin, ymin, zmin, xmax, ymax, zmax, dcell, half;
    const double4 *node_gd;           // centre + half width (code units, FP64)
    const float4  *node_gf;           // centre + half width / half (FP32)
    const int32_t *node_link;         // >= 0: first child; < 0: -(leaf + 1)
    const int32_t *node_nbr;          // 6 face neighbours per node, -1 outside
    const int32_t *leaf_node;         // leaf -> node
};

// Cooling tables for the thermal balance (cooling_setup): gbar_ff on
// log T = 2..6 for Z = 1..4 and the H I line-cooling fit terms.
struct DevCool {
    double gff[4][41];
    double logT0, dlogT;
    int32_t nfit;
    double A[8], Ti[8];
    // cooling_model 'local_ne': (T, n_e) tables (mc_nlevel.h slots) and the
    // slot of H I (-1: the fit unsuppressed)
    const double *nlt;
    int32_t nl_hi;
};

// Packet launch records (SoA), written by the launch kernel.
struct PacketBuf {
    double *x, *y, *z, *kx, *ky, *kz;
    double *E;
    double *sig;       // 6 per packet: sHI, sHeI, sHeII, dabs, dsca, g
    int32_t *inu;      // 0-based band bin
    int32_t *leaf;     // launch leaf (0-based, -1 outside)
    uint32_t *key;     // walk-order key (direction-sorted walks), or nullptr
    int64_t cap;
};

// Per-leaf state and tallies on the device.
struct DevState {
    double *nH, *xHI, *xHeI, *xHeII, *ne, *Te, *rhokap;
    float4 *opac;      // FP32 opacity inputs per leaf (mixed form)
    double *hhe;       // nleaf * 8: 6 rate sums, dust_abs, pad
    double *jt;        // nleaf * nnu
    double *rates;     // 6 * nleaf
    double *ledger;    // 3
    unsigned long long *segs;   // 4: direct, scattered, sum of metal transitions, tally atomics
    unsigned long long *counter;// work counter for persistent kernels
    double *partials;  // per-block metric partials
    // leaves next to the point sources take many updates per pass; their
    // tallies are spread over nrep replicas (folded after the walk) so that
    // concurrent atomics hit different addresses
    int32_t *hot_slot; // nleaf: slot in the hot tables or -1
    int32_t *hot_leaf; // nhot: leaf of each slot
    double *hot_hhe;   // nrep * nhot * 8
    double *hot_jt;    // nrep * nhot * nnu
    int32_t nhot, nrep;
    double *heat_dust; // nleaf: Heat_dust (ion_add_dust on), else nullptr
};

// Dust thermal-infrared band (dust_temp_mod, dust_emis_transport_mod,
// dust_ir_band_mod): the femit table, the equilibrium emission grid with its
// band cut, and the per-leaf arrays of the self-heating iteration.
#define MC_DT_NT 61
#define MC_NLAM_BB 300
// Energy-moment path tallies of the mixed-form metal walk (mc_moments.cpp):
// band bin b holds pieces bseg[2b] .. bseg[2b] + bseg[2b+1] - 1 (cut at the
// thresholds inside it); piece g = seg[5g..5g+4] = {lower energy, its ln,
// n / ln width, n, first sub-bin} is cut into n sub-bins equal in ln E, a
// packet's coordinate in the piece being f = (ln E - seg[5g+1]) seg[5g+2]
// in [0, n).  coef[(s nq + q) MC_MOM_NM + m] are the Chebyshev coefficients
// of coefficient q on sub-bin s, band[s] its band bin.
#define MC_MOM_NM 4          // moments per sub-bin (Chebyshev degree 3)
#define MC_MOM_MAXSUB 512    // sub-bins over the band at most
struct MomTab {
    int nsub = 0, nq = 0, built_ntr = -1;
    int ndust = 0, built_dust = -1;     // grain absorption among the coefficients (0/1)
    double fit_err = 0.0;
    std::vector<double> seg, bseg, coef;
    std::vector<int32_t> band;
};

struct IrBand {
    bool temp_ready = false, setup_ready = false;
    double tgrid[MC_DT_NT], femit[MC_DT_NT];
    int nl = 0, nzero = 0;
    std::vector<double> lam, sabs, edge;      // nl, nl, nl + 1 [um]
    double lam_hardest = 0.0;
    double *d_dt = nullptr;                   // tgrid | femit
    double *d_lam = nullptr, *d_sabs = nullptr, *d_edge = nullptr;
    double *d_tdust = nullptr, *d_heat_euv = nullptr, *d_heat_ir = nullptr;
    double *d_irabs = nullptr;                // grain-absorption path tally
    double *d_lum = nullptr, *d_lcum = nullptr, *d_bsum = nullptr;   // leaf emission scan
    int64_t *d_khi = nullptr;                 // last packet index of each leaf
    int64_t *d_kmax = nullptr;                // tile maxima of its prefix maximum
    double *d_lesc = nullptr;                 // escaped luminosity per bin
    double *d_led = nullptr;                  // emitted, extinguished, escaped
    double *d_wave = nullptr;                 // per-wave ledger slots of the walk
    size_t n_wave = 0;
    unsigned long long *d_segs = nullptr;     // direct, scattered, dropped, fdisc bits
    // last call
    double led[3] = {0, 0, 0};
    int64_t segs[2] = {0, 0};
    double frac_discarded = 0.0;
    double report[MC_DUSTRT_NREPORT];
    double times[3] = {0, 0, 0};              // launch, walk, rest [ms] of the last call
};

struct mc_ctx {
    mc_params p;
    // band + sources (host)
    int nnu;
    std::vector<double> eedge;
    int ncomp, multi;
    int comp_kind[MC_MAXSRC + 1];
    double src_cdf[MC_MAXSRC + 1];
    std::vector<double> samp_e, samp_f, samp_c;
    std::vector<int> samp_off;
    double Ltot;
    int qmc_ndim;
    // dust table (sorted)
    std::vector<double> dlam, dalb, dg, dcext;
    double dcref;
    // grid (host copy of octree arrays)
    int nleaf, ncells, car;
    double xmin, xmaxb, ymin, ymaxb, zmin, zmaxb, dcell;
    std::vector<double> hcx, hcy, hcz, hch;
    std::vector<int32_t> hileaf, hchildren, hneighbor, hcell_of_leaf, hlevel, hparent;
    // device
    hipStream_t stream;
    hipEvent_t ev[8];
    DevTables *d_tab;
    double *d_samp;          // 3 * nknots (energy, density, cdf)
    double *d_dust;          // 4 * dn (lam, alb, g, cext)
    int dn;
    DevGrid grid;
    double4 *d_node_gd;
    float4 *d_node_gf;
    int32_t *d_node_link, *d_node_nbr, *d_leaf_node;
    int n_lds_nodes;          // breadth-first prefix staged in LDS
    DevState st;
    PacketBuf pk;
    int nblocks_part;
    bool state_ready;
    DevCool cool;
    bool cool_ready;
    // diffuse field: Milne tables, per-leaf emission {ch0, ch1, ch2, ch3,
    // total}, leaf CDF, block sums of the scan, totals of the last build
    double *d_milne;          // cdf | xbar | fin | xg
    DevMilne milne;
    // issue form of the mixed walk's moment rows: the caller's choice
    // (MC_ROWS_*), the form for the next pass, the form of the last pass
    // (1 transposed, 0 per lane, -1 no moment rows)
    int row_issue;
    bool rows_together;
    int rows_together_last;
    double *d_hei;            // two-photon grid_y | cdf_e
    double hei_nper, hei_eper;
    bool hei_ready;
    double *d_dif, *d_dcdf, *d_bsum;
    int nbsum;
    double diffuse_lum;
    int64_t diffuse_nphot;
    // learned warm start of the thermal solve
    MlpState mlp;
    int warm_mode;            // MC_WARM_*
    int warm_levels;          // bisection levels before a warm start (mc_set_warm_levels)
    float *d_feat, *d_pred;   // nleaf x 8 features, nleaf x 4 predictions
    int64_t thermal_stats[4]; // evaluations (bracketed leaves), fallbacks, bracketed, leaves
    // direction-sorted walk order of stellar batches
    int packet_order;         // MC_ORDER_*
    int64_t batch_max;        // launch records per batch (mc_set_launch_batch)
    uint32_t *d_key_sorted;
    int32_t *d_perm;
    void *d_sort_tmp;
    size_t sort_tmp_bytes;
    // trace metals: registry (host copy and device table), switches, path
    // tallies mraw[(il*ntr + k)*2 + {0,1}] (threshold order k), rates
    // mgam / mheat[il*ntr + t] and stage fractions mfrac[il*nst + s] (element
    // order), opacity weights abund x fraction of each transition's lower
    // stage mop[il*ntr + k] (FP64) and n_H d2cm x that mopf (FP32, mixed form)
    int nel, ntr, nst;
    std::vector<mc_element> elements;
    DevMetalTab mtab;
    DevMetalTab *d_mtab;
    int metal_layout;         // MC_LAYOUT_* of the registry (cell-solve kernels)
    double *d_mraw, *d_mgam, *d_mheat, *d_mfrac, *d_mop;
    float *d_mopf;
    // dust thermal-infrared band
    IrBand ir;
    // energy-moment tallies (mixed-form metal walk): switch, tables, device
    // moments [leaf][sub-bin][MC_MOM_NM] and tables
    int mom_on;
    MomTab mom;
    double *d_mom, *d_mom_coef, *d_mom_tab, *d_mom_bseg;
    int32_t *d_mom_band;
    size_t mom_cap;
    int mom_warned;                 // fallback to per-coefficient sums reported
    // n-level atoms (mc_nlevel_setup) and the (T, n_e) cooling tables built
    // from them: table per atom index, the device slots in use
    std::vector<mc_nlevel_atom> nl_atoms;
    int nl_given, nl_dirty;
    std::map<int, std::vector<double>> nl_built;
    std::vector<int> nl_slots;
    double *d_nlt;
    // recombination-line tables (mc_lines_setup): H I, He II, He I
    mc_line_table lt[3];
    int lt_on[3];
    // last pass results
    double ledger[3];
    int64_t segs[4];          // direct, scattered, metal transitions over scored segments,
                              // global FP64 tally atomics of the walk
    double times[4];
};

// host set-up (mc_setup.cpp)
int mc_setup_band(mc_ctx *c);
int mc_setup_samplers(mc_ctx *c);
int mc_build_octree_host(mc_ctx *c, int nleaf, const double *xl, const double *yl,
                         const double *zl, const int32_t *lev, double half);
void mc_qmc_tables(int seed, uint32_t V[32][MC_QMC_DIM], uint32_t key[MC_QMC_DIM][2]);
void mc_set_error(const std::string &msg);
void mc_gff_table(double logT[41], double gff[4][41]);
void mc_milne_tables(double eion_max, double *cdf, double *xbar, double *fin, double *xg,
                     double *lgT);
int mc_hei2ph_tables(int n, const double *y, const double *A, double eion_max, double *grid_y,
                     double *cdf_e, double *n_per, double *e_per);

// device work (mc_kernels.hip)
int mc_run_pass(mc_ctx *c);
int mc_run_rates(mc_ctx *c);
int mc_run_equilibrium(mc_ctx *c, double *max_dx, double *dx_vol);
int mc_run_residual(mc_ctx *c, double *eps_max, double *eps_vol);
int mc_run_opacity_fill(mc_ctx *c);
int mc_run_stage_refresh(mc_ctx *c);
int mc_run_thermal(mc_ctx *c, double out4[4]);
int mc_run_mlp_infer(mc_ctx *c, const float *dX, int n, float *dOut);
int mc_run_mlp_features(mc_ctx *c, float *dX);
size_t mc_sort_temp_bytes(int64_t n);
int mc_sort_records(mc_ctx *c, int64_t n);
int mc_run_trace(mc_ctx *c, int precision, const double *pos, const double *dir, int32_t *cells,
                 int maxn, int32_t *n);
int mc_run_dust_temp(mc_ctx *c, const double *d_heat);
int mc_run_trace_noalloc(mc_ctx *c, int precision, const double *pos, const double *dir,
                         int32_t *cells, int maxn, int32_t *n);
int mc_nl_ready(mc_ctx *c);
int find_atom(const mc_ctx *c, const char *name, int stage);
int mc_lines_ions(mc_ctx *c, int out0, std::vector<DevLineIon> &ions, std::vector<double> &wl,
                  std::vector<int32_t> &code);
int mc_run_lines(mc_ctx *c, std::vector<double> &L, std::vector<double> &wl,
                 std::vector<int32_t> &code);
int mc_run_qmc_dump(mc_ctx *c, int stream, int64_t first, int n, int ndim, double *out);
int mc_run_launch_dump(mc_ctx *c, int64_t first_id, int n, double *rec, int32_t *bin_leaf);
int mc_run_dust_emission(mc_ctx *c, double report[MC_DUSTRT_NREPORT]);

// host set-up of the infrared band (mc_dustir.cpp)
int mc_ir_femit_table(mc_ctx *c);
int mc_ir_emission_grid(mc_ctx *c);
void mc_ir_free(mc_ctx *c);
int mc_mom_build(mc_ctx *c);
int mc_mom_ready(mc_ctx *c);
void mc_mom_free(mc_ctx *c);
// launch-record buffers of at least min(n, 4M) packets (mc_capi.cpp)
int mc_packet_capacity(mc_ctx *c, int64_t n);

#define MC_HIP(call)                                                          \
    do {                                                                      \
        hipError_t e_ = (call);                                               \
        if (e_ != hipSuccess) {                                               \
            mc_set_error(std::string(#call) + ": " + hipGetErrorString(e_));  \
            return -100;                                                      \
        }                                                                     \
    } while (0)
