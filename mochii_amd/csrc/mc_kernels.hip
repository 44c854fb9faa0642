// gfx950 kernels of one transport pass and the cell updates.
//
//   k_launch      one thread per packet: Owen-scrambled Sobol point, emitter,
//                 energy by CDF inversion, band bin, cross sections at the
//                 packet energy -> SoA launch records
//   k_walk<P,G>   persistent walk, one packet per lane; a lane whose packet
//                 ends takes the next record at once (wave-aggregated counter),
//                 so waves stay full until the batch drains.  Edge walk with
//                 the analytic path estimator, then (dust scattering on) the
//                 forced interaction, albedo weight, roulette, Henyey-Greenstein
//                 and tau-target flights with the Lucy estimator.
//                 P = FP64 (reference arithmetic) or MIXED (FP32 cell-local
//                 geometry and opacity, FP64 tallies).  G = car DDA or octree.
//   k_rates       leaf-volume normalisation of the tallies
//   k_equilibrium per-leaf H/He balance at fixed T, relaxation, metrics,
//                 and the FP32 opacity inputs for the next pass
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>

#include <mutex>

#include "mc_internal.h"

// The trace-metal registry of the calling context, copied in before every
// call that uses it (calls are synchronous and come from one host thread).
// Read with wave-uniform indices through the scalar cache.
__constant__ DevMetalTab c_mt;

namespace {

constexpr int kBlock = 256;
constexpr double kWmin = 1.0e-4;
constexpr double kPsurv = 0.1;

struct KArgs {
    const DevTables *tab;
    const double *samp;      // energy | density | cdf, each nknots
    int32_t nknots;
    const double *dust;      // lam | alb | g | cext, each dn
    int32_t dn;
    double dcref;
    int32_t add_dust, dust_scatter;
    double He_abund, d2cm;
    uint32_t seed;
    DevGrid g;
    DevState s;
    PacketBuf pk;
    int64_t first_id;        // 1-based id of record 0
    int64_t npk;             // records in this batch
    int32_t stream;          // 1 stellar, 2 diffuse (Sobol keys and post-launch draws)
    const int32_t *perm;     // walk order of the launch records (nullptr: launch order)
    int32_t coherent;        // records sorted by direction: whole-wave refill and
                             // wave-aggregated tallies
    const double *dif;       // per-leaf diffuse emission {ch0, ch1, ch2, total}
    const double *dcdf;      // diffuse leaf CDF
    DevMilne milne;
    // trace metals (the registry itself is c_mt)
    int32_t ntr;             // metal transitions (0: metal path off)
    int32_t metal_abs;       // metal opacity in the walk (par%ion_metal_abs)
    double *mraw;            // path tallies [(il*ntr + k)*2 + {0,1}]
    const double *mop;       // opacity weights abund x stage fraction [il*ntr + k]
    const float *mopf;       // the same x n_H d2cm in FP32 (mixed form)
    double Lpk;              // packet luminosity of the band walked
    // dust infrared band (walk variant M = 2): grain-absorption path tally,
    // escaped luminosity per emission bin and the bins' edges [um]
    double *irabs, *ir_lesc;
    const double *ir_edge;
    int32_t ir_nl;
    double *ir_wave;         // per-wave ledger slots of the infrared walk (5 each)
    // energy-moment tallies of the mixed-form metal walk (nullptr: per-
    // coefficient sums): moments [(s*mom_rows + row)*MC_MOM_NM + m], sub-bin
    // major, row = leaf or nleaf + rep*nhot + slot for a hot leaf's
    // replica; the pieces
    // of the band bins {lower energy, its ln, n / ln width, n, first
    // sub-bin} and each band bin's {first piece, pieces}
    double *mom;
    const double *mom_tab, *mom_bseg;
    int32_t mom_ns;
    int64_t mom_rows;
    int32_t mom_dust;        // grain absorption is among the moments' coefficients
};

// H/He recombination and collisional-ionisation coefficients at one T
struct RateSet { double aH, aHe2, aHe3, cH, cHe1, cHe2; };

#include "mc_metal_dev.h"

// The walk's dynamic LDS, [row][kBlock] doubles per block: rows 0-5 the
// H/He rate-tally coefficients of each lane's packet (coef), rows 6-8 the
// lane's energy-ledger sums (led), row 9 its packet's energy-moment
// sub-bin coordinate (mom_t), then its metal cross sections in threshold
// order (msig), one row per registry transition.
// Kept out of the lane's registers: the walk's live state is register-bound.
constexpr int kCoefRows = 10;    // 6 coefficients + 3 ledger sums + mom_t
__device__ __forceinline__ double &coef(int k)
{
    extern __shared__ double mc_msig[];
    return mc_msig[k * kBlock + threadIdx.x];
}

// the lane's energy-ledger sums {emitted, absorbed, escaped} (rows 6-8)
__device__ __forceinline__ double &led(int k)
{
    extern __shared__ double mc_msig[];
    return mc_msig[(6 + k) * kBlock + threadIdx.x];
}

__device__ __forceinline__ double &mom_t()
{
    extern __shared__ double mc_msig[];
    return mc_msig[9 * kBlock + threadIdx.x];
}

// the coefficient and sub-bin rows of another thread of the block
__device__ __forceinline__ double coef_of(int k, uint32_t tid)
{
    extern __shared__ double mc_msig[];
    return mc_msig[k * kBlock + tid];
}

__device__ __forceinline__ double mom_t_of(uint32_t tid)
{
    extern __shared__ double mc_msig[];
    return mc_msig[9 * kBlock + tid];
}

__device__ __forceinline__ double &msig(int k)
{
    extern __shared__ double mc_msig[];
    return mc_msig[(kCoefRows + k) * kBlock + threadIdx.x];
}

// ---------------------------------------------------------------------------
// launch helpers
__device__ __forceinline__ uint32_t owen_scramble(uint32_t x, uint32_t key)
{
    uint32_t y = __builtin_bitreverse32(x) + key;
    y ^= y * 0x6C50B47Cu;
    y ^= y * 0xB82F1E52u;
    y ^= y * 0xC7AFE638u;
    y ^= y * 0x8D22F6E6u;
    return __builtin_bitreverse32(y);
}

__device__ __forceinline__ double qmc_u(const DevTables *T, uint32_t idx, int d, int stream)
{
    uint32_t code = 0;
    for (uint32_t m = idx; m; m &= m - 1) code ^= T->V[__builtin_ctz(m)][d];
    code = owen_scramble(code, T->key[d][stream]);
    return ((double)code + 0.5) * (1.0 / 4294967296.0);
}

// inverse of the piecewise-linear CDF of knots [lo, hi)
__device__ double sample_energy(const double *samp, int nk, int lo0, int hi0, double u)
{
    const double *e = samp, *f = samp + nk, *c = samp + 2 * nk;
    const double t = fmin(fmax(u, 0.0), 1.0);
    if (t <= 0.0) return e[lo0];
    if (t >= 1.0) return e[hi0 - 1];
    int lo = lo0, hi = hi0 - 1;
    while (hi - lo > 1) {
        const int mid = lo + (hi - lo) / 2;
        if (c[mid] <= t) lo = mid; else hi = mid;
    }
    const double area = c[hi] - c[lo];
    if (area <= 0.0) return e[hi];
    const double dx = e[hi] - e[lo], y0 = f[lo], y1 = f[hi];
    const double local = (t - c[lo]) / area * (0.5 * (y0 + y1) * dx);
    const double slope = (y1 - y0) / dx;
    const double sc = fmax(fmax(fabs(y0), fabs(y1)), DBL_MIN);
    double off;
    if (fabs(slope * dx) <= 64.0 * DBL_EPSILON * sc) {
        off = y0 > 0.0 ? local / y0 : 0.0;
    } else {
        const double disc = sqrt(fmax(0.0, y0 * y0 + 2.0 * slope * local));
        off = 2.0 * local / (y0 + disc);
    }
    return e[lo] + fmin(fmax(off, 0.0), dx);
}

// dust optics at E from the sorted (lambda, albedo, g, C_ext) table
__device__ void dust_at(const KArgs &a, double E, double &sabs, double &ssca, double &g)
{
    const int n = a.dn;
    const double *lam = a.dust, *alb = a.dust + n, *gc = a.dust + 2 * n, *ce = a.dust + 3 * n;
    const double x0 = 1.23984 / E;
    double al, cx, gg;
    if (x0 <= lam[0]) {
        al = alb[0]; cx = ce[0]; gg = gc[0];
    } else if (x0 >= lam[n - 1]) {
        al = alb[n - 1]; cx = ce[n - 1]; gg = gc[n - 1];
    } else {
        int lo = 0, hi = n - 1;   // largest i with lam[i] <= x0
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (lam[mid] <= x0) lo = mid; else hi = mid;
        }
        const double w = (x0 - lam[lo]) / (lam[lo + 1] - lam[lo]);
        al = alb[lo] * (1.0 - w) + alb[lo + 1] * w;
        gg = gc[lo] * (1.0 - w) + gc[lo + 1] * w;
        const double wl = log(x0 / lam[lo]) / log(lam[lo + 1] / lam[lo]);
        cx = exp(log(fmax(ce[lo], DBL_MIN)) * (1.0 - wl) + log(fmax(ce[lo + 1], DBL_MIN)) * wl);
    }
    sabs = (1.0 - al) * cx / a.dcref;
    ssca = al * cx / a.dcref;
    g = gg;
}

__device__ __forceinline__ int car_index(const DevGrid &g, double x, double y, double z)
{
    const int ix = min(max((int)((x - g.xmin) / g.dcell), 0), g.nx - 1);
    const int iy = min(max((int)((y - g.ymin) / g.dcell), 0), g.ny - 1);
    const int iz = min(max((int)((z - g.zmin) / g.dcell), 0), g.nz - 1);
    return ix + g.nx * (iy + g.ny * iz);
}

// leaf containing (x,y,z), -1 outside: raster index on a car grid, root-down
// descent on the octree (amr_find_leaf), reading the first nl nodes from the
// LDS copies when given
__device__ int find_leaf(const DevGrid &g, double x, double y, double z,
                         const double4 *sgd = nullptr, const int32_t *slink = nullptr, int nl = 0)
{
    if (x < g.xmin || x > g.xmax || y < g.ymin || y > g.ymax || z < g.zmin || z > g.zmax)
        return -1;
    if (g.car) return car_index(g, x, y, z);
    int c = 0;
    for (;;) {
        const int link = c < nl ? slink[c] : g.node_link[c];
        if (link < 0) return -link - 1;
        const double4 n = c < nl ? sgd[c] : g.node_gd[c];
        c = link + (x >= n.x ? 1 : 0) + (y >= n.y ? 2 : 0) + (z >= n.z ? 4 : 0);
    }
}

// band bin of E (ion_bin_of, src/ion_band_mod.f90:624-646):
// ion_eedge(bin) <= E < ion_eedge(bin+1), clamped to the band
__device__ __forceinline__ int bin_of(const DevTables *T, double E)
{
    int lo = 0, hi = T->nnu - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (T->eedge[mid] <= E) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// launch record i: position, direction, energy, bin, cross sections at E
// (build_ion_packet_physics), start leaf
__device__ __forceinline__ void store_packet(const KArgs &a, int64_t i, double x, double y, double z,
                                             double kx, double ky, double kz, double E, int leaf)
{
    const PacketBuf &P = a.pk;
    P.x[i] = x; P.y[i] = y; P.z[i] = z;
    P.kx[i] = kx; P.ky[i] = ky; P.kz[i] = kz;
    P.E[i] = E;
    P.inu[i] = bin_of(a.tab, E);
    double sabs = 0.0, ssca = 0.0, gg = 0.0;
    if (a.add_dust) dust_at(a, E, sabs, ssca, gg);
    double *sg = P.sig + 6 * i;
    sg[0] = sigma_hydrogenic(E, MC_ETH_HI, 1.0);
    sg[1] = sigma_HeI(E);
    sg[2] = sigma_hydrogenic(E, MC_ETH_HEII, 2.0);
    sg[3] = sabs; sg[4] = ssca; sg[5] = gg;
    P.leaf[i] = leaf;
}

// energy-moment sub-bin of a packet of energy E in band bin inu and its
// coordinate t in [-1, 1] on the sub-bin, or -1 when E lies outside the bin's
// pieces (pieces start at the thresholds: E >= threshold selects the piece
// where that transition is on, as in the walk's nmet)
__device__ __forceinline__ int mom_subbin(const KArgs &a, int inu, double E, double &t)
{
    const double *bs = a.mom_bseg + 2 * inu;
    int g = (int)bs[0];
    const int ge = g + (int)bs[1];
    while (g + 1 < ge && E >= a.mom_tab[5 * (g + 1)]) ++g;
    const double *mt = a.mom_tab + 5 * g;
    const double f = (log(E) - mt[1]) * mt[2];
    const double nsub = mt[3];
    if (E >= mt[0] && f <= nsub * (1.0 + 1.0e-9)) {
        const int j = min(max((int)floor(f), 0), (int)nsub - 1);
        t = fmin(fmax(2.0 * (f - (double)j) - 1.0, -1.0), 1.0);
        return (int)mt[4] + j;
    }
    return -1;
}

// Walk-order key of a launch record, sorted per batch so that a wavefront
// walks 64 packets that leave one emitter in nearly the same direction:
//   per-coefficient tallies: emitting component (bits 22+), direction on an
//     octahedral map (2^7 cells a side, Morton order), then ln E across the
//     band (8 bits), so a wave group's rate sums span few metal transitions
//     and band bins;
//   energy-moment tallies: component (bits 27+), the packet's moment sub-bin
//     (+1, 8 bits), then the direction (2^9 cells a side), so a wave's lanes
//     share sub-bins and their same-leaf segments merge into one set of
//     moment atomics.
// Measured on the 2e7-leaf multifractal model: the per-coefficient walk is
// slower with the energy first (7.96 vs 5.68 s per pass) and the moment walk
// faster with the sub-bin first (3.91 vs 5.32 s).
constexpr int kKeyDirBits = 7;
constexpr int kKeyEBits = 22 - 2 * kKeyDirBits;
constexpr int kKeyMomDirBits = 9;

__device__ __forceinline__ uint32_t spread16(uint32_t v)
{
    v &= 0xFFFFu;
    v = (v | (v << 8)) & 0x00FF00FFu;
    v = (v | (v << 4)) & 0x0F0F0F0Fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}

__device__ __forceinline__ uint32_t direction_key(const KArgs &a, int comp, double kx, double ky,
                                                  double kz, double E, int inu)
{
    const DevTables *T = a.tab;
    const double n = fabs(kx) + fabs(ky) + fabs(kz);
    double u = kx / n, v = ky / n;
    if (kz < 0.0) {
        const double uu = (1.0 - fabs(v)) * (u >= 0.0 ? 1.0 : -1.0);
        const double vv = (1.0 - fabs(u)) * (v >= 0.0 ? 1.0 : -1.0);
        u = uu; v = vv;
    }
    const int bits = a.mom ? kKeyMomDirBits : kKeyDirBits;
    const double cells = (double)(1 << bits);
    const uint32_t iu = (uint32_t)fmin(fmax((u + 1.0) * 0.5 * cells, 0.0), cells - 1.0);
    const uint32_t iv = (uint32_t)fmin(fmax((v + 1.0) * 0.5 * cells, 0.0), cells - 1.0);
    const uint32_t dir = spread16(iu) | (spread16(iv) << 1);
    if (a.mom) {
        double t;
        const uint32_t sb = (uint32_t)min(mom_subbin(a, inu, E, t) + 1, 255);
        return ((uint32_t)comp << 27) | (sb << (2 * kKeyMomDirBits)) | dir;
    }
    const double elev = (double)((1 << kKeyEBits) - 1);
    const double e0 = T->eedge[0], e1 = T->eedge[T->nnu];
    const uint32_t ie = (uint32_t)fmin(fmax(elev * log(E / e0) / log(e1 / e0), 0.0), elev);
    return ((uint32_t)comp << 22) | (dir << kKeyEBits) | ie;
}

__global__ __launch_bounds__(kBlock) void k_launch(KArgs a, int nlds)
{
    extern __shared__ unsigned char lds_raw[];
    double4 *sgd = reinterpret_cast<double4 *>(lds_raw);
    int32_t *slink = reinterpret_cast<int32_t *>(sgd + nlds);
    for (int k = threadIdx.x; k < nlds; k += kBlock) {
        sgd[k] = a.g.node_gd[k];
        slink[k] = a.g.node_link[k];
    }
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= a.npk) return;
    const DevTables *T = a.tab;
    const int64_t id = a.first_id + i;
    const uint32_t q = (uint32_t)(id - 1);   // index bits beyond 32 carry no direction number
    // a loop bounded by the run-time dimension count would leave u in
    // per-lane scratch memory (56 B written and read back per record):
    // unrolled over the fixed maximum, it stays in registers
    double u[7];
#pragma unroll
    for (int d = 0; d < 7; ++d) u[d] = d < T->qmc_ndim ? qmc_u(T, q, d, 0) : 0.0;
    int comp = 0;
    double umu, uphi, ue;
    if (!T->multi && T->nsource == 1) {
        ue = u[0]; umu = u[1]; uphi = u[2];
    } else {
        if (T->multi) {
            comp = T->ncomp - 1;
            for (int c = 0; c < T->ncomp; ++c)
                if (u[0] <= T->src_cdf[c]) { comp = c; break; }
        }
        ue = u[1]; umu = u[2]; uphi = u[3];
    }
    double x, y, z, kx, ky, kz;
    const int kind = T->comp_kind[comp];
    if (kind >= 1) {
        x = T->src_pos[kind - 1][0]; y = T->src_pos[kind - 1][1]; z = T->src_pos[kind - 1][2];
        const double cost = 2.0 * umu - 1.0, sint = sqrt(1.0 - cost * cost);
        double sp, cp;
        sincos(MC_TWOPI * uphi, &sp, &cp);
        kx = sint * cp; ky = sint * sp; kz = cost;
    } else if (T->ext_geometry == MC_EXT_SPH) {
        // uniform entry point on the bounding sphere, Lambert inward direction
        const DevGrid &g = a.g;
        const double xc = 0.5 * (g.xmin + g.xmax), yc = 0.5 * (g.ymin + g.ymax),
                     zc = 0.5 * (g.zmin + g.zmax);
        const double cost = 2.0 * u[4] - 1.0, sint = sqrt(1.0 - cost * cost);
        double s1, c1;
        sincos(MC_TWOPI * u[5], &s1, &c1);
        const double rx = sint * c1, ry = sint * s1, rz = cost;
        x = xc + T->rmax * rx; y = yc + T->rmax * ry; z = zc + T->rmax * rz;
        const double nx = -rx, ny = -ry, nz = -rz;
        double tx, ty, tz;
        if (fabs(nx) <= fabs(ny) && fabs(nx) <= fabs(nz)) { tx = 0.0; ty = nz; tz = -ny; }
        else if (fabs(ny) <= fabs(nz)) { tx = -nz; ty = 0.0; tz = nx; }
        else { tx = ny; ty = -nx; tz = 0.0; }
        const double inv = 1.0 / sqrt(tx * tx + ty * ty + tz * tz);
        tx *= inv; ty *= inv; tz *= inv;
        const double bx = ny * tz - nz * ty, by = nz * tx - nx * tz, bz = nx * ty - ny * tx;
        const double mu = sqrt(umu), sm = sqrt(1.0 - mu * mu);
        double s2, c2;
        sincos(MC_TWOPI * uphi, &s2, &c2);
        kx = -mu * rx + sm * (c2 * tx + s2 * bx);
        ky = -mu * ry + sm * (c2 * ty + s2 * by);
        kz = -mu * rz + sm * (c2 * tz + s2 * bz);
        const double span = fmin(fmin(g.xmax - g.xmin, g.ymax - g.ymin), g.zmax - g.zmin);
        const double nudge = 1.0e-6 * span;
        x -= nudge * rx; y -= nudge * ry; z -= nudge * rz;
    } else {
        // area-weighted box face, uniform point on it, Lambert inward direction
        const DevGrid &g = a.g;
        const double xr = g.xmax - g.xmin, yr = g.ymax - g.ymin, zr = g.zmax - g.zmin;
        const double nudge = 1.0e-6 * fmin(fmin(xr, yr), zr);
        const double fa = yr * zr, fb = zr * xr, fc = xr * yr;
        const double cum[6] = {fa, 2 * fa, 2 * fa + fb, 2 * fa + 2 * fb, 2 * fa + 2 * fb + fc,
                               2 * fa + 2 * fb + 2 * fc};
        const double pick = u[4] * cum[5];
        int f = 5;
        for (int k = 0; k < 6; ++k) if (pick <= cum[k]) { f = k; break; }
        const double mu = sqrt(umu), sm = sqrt(1.0 - mu * mu);
        double sp, cp;
        sincos(MC_TWOPI * uphi, &sp, &cp);
        const double c1 = u[5], c2 = u[6];
        if (f < 2) {
            x = f == 0 ? g.xmax - nudge : g.xmin + nudge;
            y = g.ymin + yr * c1; z = g.zmin + zr * c2;
            kx = f == 0 ? -mu : mu; ky = sm * cp; kz = sm * sp;
        } else if (f < 4) {
            x = g.xmin + xr * c1; y = f == 2 ? g.ymax - nudge : g.ymin + nudge;
            z = g.zmin + zr * c2;
            kx = sm * cp; ky = f == 2 ? -mu : mu; kz = sm * sp;
        } else {
            x = g.xmin + xr * c1; y = g.ymin + yr * c2;
            z = f == 4 ? g.zmax - nudge : g.zmin + nudge;
            kx = sm * cp; ky = sm * sp; kz = f == 4 ? -mu : mu;
        }
    }
    const double E = sample_energy(a.samp, a.nknots, T->samp_off[comp], T->samp_off[comp + 1], ue);
    store_packet(a, i, x, y, z, kx, ky, kz, E, find_leaf(a.g, x, y, z, sgd, slink, nlds));
    if (a.pk.key) a.pk.key[i] = direction_key(a, comp, kx, ky, kz, E, a.pk.inu[i]);
}

// ---------------------------------------------------------------------------
// counter-based generator for the post-launch draws (same definition as the
// CPU oracle): three splitmix64 finalizer rounds over (seed, id, stream|k)
__host__ __device__ inline uint64_t mix64(uint64_t z)
{
    z ^= z >> 30; z *= 0xBF58476D1CE4E5B9ull;
    z ^= z >> 27; z *= 0x94D049BB133111EBull;
    z ^= z >> 31;
    return z;
}

__host__ __device__ inline double rng_u(uint32_t seed, int stream, uint64_t id, uint32_t k)
{
    uint64_t h = mix64((uint64_t)seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull);
    h = mix64(h ^ id);
    h = mix64(h ^ (((uint64_t)stream << 32) | (uint64_t)k));
    return ((double)(h >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

// ---------------------------------------------------------------------------
// per-lane packet state
struct Lane {
    int phase;               // 0 idle, 1 edge walk, 2 tau flight, 3 drained
    int inu, nscatt, leaf;
    uint32_t nk;
    int64_t id;
    double x, y, z, kx, ky, kz;      // flight origin and direction
    double wgt, E;
    double sHI, sHeI, sHeII, dabs, dsca, dg;
    double wl;
    // walk state, FP64 form: DDA
    int ix, iy, iz, sx, sy, sz;
    double tm[3], td[3], tcur, expo, tau, tau_in;
    // walk state, mixed form: cell-local position in [0,1]^3 (car) or
    // [-1,1]^3 in units of the half width (octree)
    float px, py, pz, fkx, fky, fkz, fsHI, fsHeI, fsHeII, fdust;
    float fexpo, ftau, ftau_in, ftcur;
    int nmet;                // metal transitions below the packet's energy
    int ms;                  // energy-moment sub-bin of the packet (-1: none)
    double rhe, rE;          // 1 / (E ev2erg) and 1 / E of the packet (metal tallies)
    double lam;              // wavelength [um] (infrared band)
    // tally of this step, scored after the step by the whole wave
    int s_il;
    double s_pl;
    // octree walk: current node, current position (FP64 form), node geometry
    int node, leaf0;
    uint32_t nat;            // global FP64 tally atomics this lane issued
    double ox, oy, oz;
    float4 gf;
};

template <int P> struct Prec;
template <> struct Prec<MC_FP64> { static constexpr bool mixed = false; };
template <> struct Prec<MC_MIXED> { static constexpr bool mixed = true; };

// ion_packet_opacity with species_cached_opacity (src/ion_packet_mod.f90:55-86,
// src/species_mod.f90:291-301)
template <int M>
__device__ __forceinline__ double opacity_d(const KArgs &a, const Lane &L, int il)
{
#pragma clang fp contract(off)
    const DevState &s = a.s;
    if constexpr (M == 2) {
        // infrared band: every gas cross section vanishes below the band cut,
        // so n_H (0) d2cm + rhokap sigma is rhokap sigma exactly
        double kap = s.rhokap[il] * L.dabs;
        if (a.dust_scatter) kap = kap + s.rhokap[il] * L.dsca;
        return kap;
    }
    double gas = s.xHI[il] * L.sHI + a.He_abund * (s.xHeI[il] * L.sHeI + s.xHeII[il] * L.sHeII);
    if constexpr (M == 1) {
        if (a.metal_abs) {
            double metal = 0.0;
            const double *w = a.mop + (size_t)il * a.ntr;
            for (int k = 0; k < L.nmet; ++k) metal = metal + w[k] * msig(k);
            gas = gas + metal;
        }
    }
    double kap = s.nH[il] * gas * a.d2cm;
    if (a.add_dust) {
        kap = kap + s.rhokap[il] * L.dabs;
        if (a.dust_scatter) kap = kap + s.rhokap[il] * L.dsca;
    }
    return kap;
}

template <int M>
__device__ __forceinline__ float opacity_f(const KArgs &a, const Lane &L, int il)
{
    if constexpr (M == 2) return (float)a.s.rhokap[il] * L.fdust;   // = opac.w x fdust
    const float4 o = a.s.opac[il];
    float k = o.x * L.fsHI + o.y * L.fsHeI + o.z * L.fsHeII + o.w * L.fdust;
    if constexpr (M == 1) {
        if (a.metal_abs) {
            const float *w = a.mopf + (size_t)il * a.ntr;
            for (int q = 0; q < L.nmet; ++q) k = k + w[q] * (float)msig(q);
        }
    }
    return k;
}

struct Counters {
    double e_emit, e_abs, e_esc; // infrared band only; the ionizing ledger is in LDS (led)
    uint32_t sd, ss;             // per lane and batch: far below 2^32
    uint32_t smet;               // metal transitions summed over scored segments
    // infrared band: packets ledgered (every one at weight 1) and the
    // rounding errors of the extinguished / escaped sums
    unsigned long long n_emit;
    double c_abs, c_esc;
};

// s + x with its rounding error carried in c (Knuth's TwoSum)
__device__ __forceinline__ void two_sum_acc(double &s, double &c, double x)
{
    const double t = s + x;
    const double bp = t - s;
    c += (s - (t - bp)) + (x - bp);
    s = t;
}

// score_ion_path: the packet's path quantity into its leaf.  A coefficient
// that is exactly zero (He I / He II terms below their thresholds) adds +0 to
// the tally, so its atomic is skipped; hot leaves go to the lane's replica.
// score_ion_path for one segment: the record is kept in the lane and the
// atomics are issued by the whole wave after the step (flush_tally), so that
// lanes of a wave that scored the same leaf can be summed first.  Segments
// with zero path luminosity are neither scored nor counted, as in the
// reference (src/ion_score_mod.f90:148).
__device__ __forceinline__ void score(const KArgs &, Lane &L, int il, double pl, Counters &cn)
{
    if (pl == 0.0) return;
    L.s_il = il;
    L.s_pl = pl;
    if (L.nscatt > 0) cn.ss++; else cn.sd++;
    cn.smet += (uint32_t)L.nmet;
}

// entries of the per-wave moment cache: MC_MOM_NM moments and the grain
// absorption kept outside them
constexpr int kAggVals = MC_MOM_NM + 1;

__device__ __forceinline__ void wave_sync_lds()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// tally destination of a leaf: its own rows, or one of the replicas of a hot
// leaf chosen by a thread of the block (tid: the issuing lane, or the first
// lane of the lanes that issue one row together)
__device__ __forceinline__ int hot_rep(const DevState &s, uint32_t tid)
{
    return (int)(((blockIdx.x * kBlock + tid) * 0x9E3779B1u) >> 16) % s.nrep;
}

// one global FP64 tally atomic issued by one lane alone: one request
__device__ __forceinline__ void tally_add(double *p, double v, uint32_t &n)
{
    unsafeAtomicAdd(p, v);
    ++n;
}

__device__ __forceinline__ void tally_rows(const KArgs &a, int il, double *&h, double *&j,
                                           uint32_t tid)
{
    const DevState &s = a.s;
    const int nnu = a.tab->nnu;
    const int hs = s.nhot > 0 ? s.hot_slot[il] : -1;
    if (hs < 0) {
        h = s.hhe + (size_t)il * 8;
        j = s.jt + (size_t)il * nnu;
    } else {
        const size_t slot = (size_t)hot_rep(s, tid) * s.nhot + hs;
        h = s.hot_hhe + slot * 8;
        j = s.hot_jt + slot * nnu;
    }
}

// species_rate_path_score coefficients of transition k (threshold order)
// for the lane's packet (src/species_mod.f90:329-343), with the packet's
// reciprocals of E taken once at load
__device__ __forceinline__ void metal_coeff(int k, const Lane &L, double &c1, double &c2)
{
    const double sg = msig(k);
    c1 = sg * L.rhe;
    c2 = sg * fmax(1.0 - c_mt.flat_eth[k] * L.rE, 0.0);
}

// energy moments of sub-bin ms of a leaf: its own row, or one of the
// replicas of a hot leaf
__device__ __forceinline__ double *mom_row(const KArgs &a, int il, int ms, uint32_t tid)
{
    const DevState &s = a.s;
    const int hs = s.nhot > 0 ? s.hot_slot[il] : -1;
    // replica rows follow the nleaf = mom_rows - nrep nhot leaf rows
    const size_t row = hs < 0 ? (size_t)il
                              : (size_t)a.mom_rows - (size_t)(s.nrep - hot_rep(s, tid)) * s.nhot + hs;
    return a.mom + ((size_t)ms * a.mom_rows + row) * MC_MOM_NM;
}

// The device's atomic unit serves one request per 64-byte line that an
// atomic instruction touches, whatever the number of the instruction's lanes
// in that line: eight lanes adding into eight doubles of one 64-byte row
// complete as fast as one lane adding one double (~2.4e10 requests per
// second device-wide, measured, for FP64, FP32 and integer adds alike).  A
// lane that issues its row's sums alone costs one request per sum; the
// writers below issue the rows of several lanes transposed, W lanes per row
// in one instruction, so that a row costs one request per line it spans.
// Entries are the lanes with `ent` set, in lane order; es is this wave's
// table of 64 source lanes; the rows go out 64 / W per instruction.

constexpr int kRowsLane = 0, kRowsTogether = 1;

#ifndef MC_COOP_MIN
#define MC_COOP_MIN 4
#endif
// rows below which the entry lanes issue their own: with fewer the
// transposed issue costs more instructions than the requests it saves
constexpr int kCoopMin = MC_COOP_MIN;

// rank of each entry lane and the table of source lanes; returns the count
__device__ __forceinline__ int entry_table(bool ent, int lane, int *es)
{
    const unsigned long long em = __ballot(ent);
    if (ent) es[__popcll(em & ((1ull << lane) - 1ull))] = lane;
    wave_sync_lds();
    return __popcll(em);
}

// moment q of a segment of path luminosity pl at sub-bin coordinate t:
// pl T_q(t) by the Chebyshev recurrence
__device__ __forceinline__ double seg_moment(double pl, double t, int q)
{
    const double t2 = 2.0 * t;
    double pm2 = pl, pm1 = 0.5 * t2 * pl;
    for (int j = 2; j <= q; ++j) {
        const double pq = t2 * pm1 - pm2;
        pm2 = pm1;
        pm1 = pq;
    }
    return q == 0 ? pm2 : pm1;
}

// Moment rows.  Entry of a lane: its key (il, sub-bin) as mom_key and
// either the cache slot whose sums go out (eslot >= 0, cleared here) or, with
// eslot < 0, the lane's own segment (its path luminosity and sub-bin
// coordinate).  A row is MC_MOM_NM moments, 32 bytes in one line, and, with
// the grain absorption kept outside the moments, that term in the leaf's
// rate row.  The writing lanes take what they add from the slot or from the
// entry lane, so no lane holds a row in registers.
__device__ __forceinline__ void mom_rows_out(const KArgs &a, const Lane &L, int lane, bool ent,
                                             unsigned long long key, int eslot, double *tv,
                                             int *es, uint32_t &nat)
{
    if (!__ballot(ent)) return;
    const bool dust = a.add_dust && !a.mom_dust;
    if (kCoopMin > 0 && __popcll(__ballot(ent)) < kCoopMin) {
        // few rows: each entry lane issues its own
        if (ent) {
            double *p = mom_row(a, (int)((key - 1ull) >> 20), (int)((key - 1ull) & 0xFFFFFull), threadIdx.x);
            const double t2 = 2.0 * mom_t();
            double pm2 = L.s_pl, pm1 = 0.5 * t2 * L.s_pl;
#pragma unroll
            for (int q = 0; q < MC_MOM_NM; ++q) {
                double val;
                if (eslot >= 0) {
                    val = tv[eslot * kAggVals + q];
                    tv[eslot * kAggVals + q] = 0.0;
                } else if (q < 2) {
                    val = q == 0 ? pm2 : pm1;
                } else {
                    val = t2 * pm1 - pm2;
                    pm2 = pm1;
                    pm1 = val;
                }
                tally_add(p + q, val, nat);
            }
            if (dust) {
                double val = eslot >= 0 ? tv[eslot * kAggVals + MC_MOM_NM] : L.s_pl * L.dabs;
                if (eslot >= 0) tv[eslot * kAggVals + MC_MOM_NM] = 0.0;
                if (val != 0.0) {
                    double *h, *j;
                    tally_rows(a, (int)((key - 1ull) >> 20), h, j, threadIdx.x);
                    tally_add(h + 6, val, nat);
                }
            }
        }
        wave_sync_lds();
        return;
    }
    const int n = entry_table(ent, lane, es);
    const int W = dust ? MC_MOM_NM + 1 : MC_MOM_NM;
    const int R = 64 / W;
    const int r = lane / W, q = lane - r * W;
    const uint32_t tid = (threadIdx.x & ~63u) + (uint32_t)(r * W);
    for (int base = 0; base < n; base += R) {
        const bool on = r < R && base + r < n;
        const int src = on ? es[base + r] : lane;
        const unsigned long long k = (unsigned long long)__shfl((long long)key, src) - 1ull;
        const int sl = __shfl(eslot, src);
        const double pl = __shfl(L.s_pl, src);
        const double da = dust ? __shfl(L.dabs, src) : 0.0;
        if (on) {
            double val;
            if (sl >= 0) {
                double *t = tv + sl * kAggVals + q;
                val = *t;
                *t = 0.0;
            } else if (q < MC_MOM_NM) {
                val = seg_moment(pl, mom_t_of((threadIdx.x & ~63u) + (uint32_t)src), q);
            } else {
                val = pl * da;
            }
            const int il = (int)(k >> 20), ms = (int)(k & 0xFFFFFull);
            if (q < MC_MOM_NM) {
                unsafeAtomicAdd(mom_row(a, il, ms, tid) + q, val);
                if (q == 0) ++nat;
            } else if (val != 0.0) {
                double *h, *j;
                tally_rows(a, il, h, j, tid);
                tally_add(h + 6, val, nat);
            }
        }
    }
    wave_sync_lds();
}

// Per-coefficient rows of the lanes with `ent` set (score_ion_path,
// src/ion_score_mod.f90:137-231): the six rate coefficients and the grain
// absorption into the leaf's 64-byte rate row (skipping exact zeros: He
// terms below their thresholds), the path luminosity into the packet's
// band-J bin; then, with the metal path, each lane's own metal transitions
// (species_rate_path_score, src/species_mod.f90:329-343).
template <int M>
__device__ __forceinline__ void coef_rows_out(const KArgs &a, Lane &L, int lane, bool ent,
                                              int *es)
{
    if (!__ballot(ent)) return;
    const int n = entry_table(ent, lane, es);
    const int r = lane >> 3, q = lane & 7;
    const uint32_t tid = (threadIdx.x & ~63u) + (uint32_t)(r * 8);
    for (int base = 0; base < n; base += 8) {
        const bool on = base + r < n;
        const int src = on ? es[base + r] : lane;
        const int il = __shfl(L.s_il, src), inu = __shfl(L.inu, src);
        const double pl = __shfl(L.s_pl, src);
        const double da = __shfl(L.dabs, src);
        if (on) {
            double *h, *j;
            tally_rows(a, il, h, j, tid);
            if (q == 7) {
                tally_add(j + inu, pl, L.nat);
            } else {
                const double val = q < 6 ? pl * coef_of(q, (threadIdx.x & ~63u) + (uint32_t)src)
                                         : (a.add_dust ? pl * da : 0.0);
                if (val != 0.0) {
                    unsafeAtomicAdd(h + q, val);
                    if (q == 0) ++L.nat;
                }
            }
        }
    }
    wave_sync_lds();
    if constexpr (M == 1) {
        if (ent) {
            const double pl = L.s_pl;
            double *r2 = a.mraw + (size_t)L.s_il * a.ntr * 2;
            for (int k = 0; k < L.nmet; ++k) {
                double c1, c2;
                metal_coeff(k, L, c1, c2);
                tally_add(r2 + 2 * k, pl * c1, L.nat);
                if (c2 != 0.0) tally_add(r2 + 2 * k + 1, pl * c2, L.nat);
            }
        }
    }
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

constexpr int kAggRounds = 4;   // leaf groups tried per step
constexpr int kAggMin = 3;      // smallest group worth a wave reduction
// energy-moment walk: only the step's first leaf group (the largest, as a
// rule) is summed across the wave; measured 83.8 / 57.4 / 57.3 / 66.1 ms
// per pass of the H II region workload at 0 / 1 / 2 / 8 rounds
constexpr int kAggRoundsMom = 1;

// The metal-free moment walk keeps a per-wave LDS cache of (leaf, sub-bin)
// moment rows: 128 slots, each a tag, MC_MOM_NM moments and the grain
// absorption kept outside the moments.  Global atomics cost far more device
// throughput than LDS ones, and the lanes of a wave walk nearby leaves in one
// sub-bin, so a lane adds its segment to the slot that holds its (leaf,
// sub-bin) and the slot's sums go out as one row when another key takes the
// slot or the wave ends.  Per step: a lane whose key is in its slot adds to
// it (a hit, and the slot stays); the lanes whose key is not claim its slot,
// one of them (the owner) takes the old entry out and installs its key, and
// the lanes with the owner's key add to it; a lane whose slot is held by a
// hit of another key, or taken by another key, sends its segment's moments
// out as a row of its own.  The evicted rows and those go out together
// (mom_rows_out); segments without a sub-bin keep the per-coefficient sums.
#ifndef MC_AGG_BITS0
#define MC_AGG_BITS0 7
#endif
// slots per wave: 2^7 (three blocks of four waves per CU fit the LDS); the
// metal walk's LDS also holds its packets' metal cross sections
template <int M> constexpr int agg_slot_bits() { return M == 0 ? MC_AGG_BITS0 : 7; }

// the moment-row key of a (leaf, sub-bin), never 0 (0 marks an empty slot)
__device__ __forceinline__ unsigned long long mom_key(int il, int ms)
{
    return ((((unsigned long long)(unsigned)il) << 20) | (unsigned)ms) + 1ull;
}

// tv: this wave's slots x kAggVals sums (zero in an empty slot), tt its
// tags, tc its claims, es its entry table; M the walk's metal path
template <int M>
__device__ __forceinline__ void flush_mom_cache(const KArgs &a, Lane &L, int lane, double *tv,
                                                unsigned long long *tt, int *tc, int *es)
{
    const bool has = L.s_il >= 0;
    if (!__ballot(has)) return;
    const bool mom = has && L.ms >= 0;
    const unsigned long long key = mom ? mom_key(L.s_il, L.ms) : 0ull;
    const int slot = (int)((key * 0x9E3779B97F4A7C15ull) >> (64 - agg_slot_bits<M>()));
    const bool hit = mom && tt[slot] == key;
    const bool miss = mom && !hit;
    // claims: the missing lanes first, then a hit marks its slot as held
    if (miss) tc[slot] = lane;
    wave_sync_lds();
    if (hit) tc[slot] = 64;
    wave_sync_lds();
    const int claim = miss ? tc[slot] : 64;
    const int owner = claim < 64 ? claim : lane;
    const unsigned long long okey = (unsigned long long)__shfl((long long)key, owner);
    const bool in = hit || (miss && claim < 64 && okey == key);
    // rows out this step: the entry an owner evicts (its slot's sums, the
    // old tag) and the segment of a lane that got no slot; the owner
    // installs its key at once, the sums follow once the old ones are out
    unsigned long long ekey = key;
    int eslot = -1;
    bool ent = mom && !in;
    if (miss && claim == lane) {
        const unsigned long long old = tt[slot];
        if (old != 0ull) {
            ekey = old;
            eslot = slot;
            ent = true;
        }
        tt[slot] = key;
    }
    mom_rows_out(a, L, lane, ent, ekey, eslot, tv, es, L.nat);
    if (in) {
        double *t = tv + slot * kAggVals;
        const double pl = L.s_pl;
        const double t2 = 2.0 * mom_t();
        double pm2 = pl, pm1 = 0.5 * t2 * pl;
        atomicAdd(t + 0, pm2);
        atomicAdd(t + 1, pm1);
#pragma unroll
        for (int q = 2; q < MC_MOM_NM; ++q) {
            const double pq = t2 * pm1 - pm2;
            atomicAdd(t + q, pq);
            pm2 = pm1;
            pm1 = pq;
        }
        if (a.add_dust && !a.mom_dust && L.dabs != 0.0) atomicAdd(t + MC_MOM_NM, pl * L.dabs);
    }
    // segments without a sub-bin: the per-coefficient sums
    coef_rows_out<M>(a, L, lane, has && !mom, es);
    wave_sync_lds();
    L.s_il = -1;
}

// the wave's cached entries out, at the end of its walk
template <int M>
__device__ __forceinline__ void mom_cache_drain(const KArgs &a, const Lane &L, int lane,
                                                double *tv, unsigned long long *tt, int *es,
                                                uint32_t &nat)
{
    for (int s0 = 0; s0 < (1 << agg_slot_bits<M>()); s0 += 64) {
        const int s = s0 + lane;
        const unsigned long long key = tt[s];
        mom_rows_out(a, L, lane, key != 0ull, key, s, tv, es, nat);
        tt[s] = 0ull;
    }
}

// The per-lane issue of the same cache: an owner writes the entry it evicts
// and a lane without a slot its segment's moments, each sum by one atomic
// of its own.  Fewer instructions and registers than the transposed issue,
// so the faster form while the device's atomic unit has room; the host
// switches between the two on the measured request rate (mc_run_pass).
__device__ __forceinline__ void mom_row_lane(const KArgs &a, unsigned long long key, double *t,
                                             uint32_t &nat)
{
    const unsigned long long k = key - 1ull;
    const int il = (int)(k >> 20), ms = (int)(k & 0xFFFFFull);
    double *m = mom_row(a, il, ms, threadIdx.x);
#pragma unroll
    for (int q = 0; q < MC_MOM_NM; ++q) {
        tally_add(m + q, t[q], nat);
        t[q] = 0.0;
    }
    const double d = t[MC_MOM_NM];
    if (d != 0.0) {
        double *h, *j;
        tally_rows(a, il, h, j, threadIdx.x);
        tally_add(h + 6, d, nat);
        t[MC_MOM_NM] = 0.0;
    }
}

// one lane's segment: its moments, or without a sub-bin the per-coefficient
// sums (band J, nonzero rate coefficients, grain absorption, metals)
template <int M>
__device__ __forceinline__ void segment_lane(const KArgs &a, Lane &L)
{
    const double pl = L.s_pl;
    if (L.ms >= 0) {
        double *m = mom_row(a, L.s_il, L.ms, threadIdx.x);
        const double t2 = 2.0 * mom_t();
        double pm2 = pl, pm1 = 0.5 * t2 * pl;
        tally_add(m + 0, pm2, L.nat);
        tally_add(m + 1, pm1, L.nat);
#pragma unroll
        for (int q = 2; q < MC_MOM_NM; ++q) {
            const double pq = t2 * pm1 - pm2;
            tally_add(m + q, pq, L.nat);
            pm2 = pm1;
            pm1 = pq;
        }
        if (a.add_dust && !a.mom_dust && L.dabs != 0.0) {
            double *h, *j;
            tally_rows(a, L.s_il, h, j, threadIdx.x);
            tally_add(h + 6, pl * L.dabs, L.nat);
        }
        return;
    }
    double *h, *j;
    tally_rows(a, L.s_il, h, j, threadIdx.x);
    tally_add(j + L.inu, pl, L.nat);
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const double ck = coef(k);
        if (ck != 0.0) tally_add(h + k, pl * ck, L.nat);
    }
    if (a.add_dust && L.dabs != 0.0) tally_add(h + 6, pl * L.dabs, L.nat);
    if constexpr (M == 1) {
        double *r = a.mraw + (size_t)L.s_il * a.ntr * 2;
        for (int k = 0; k < L.nmet; ++k) {
            double c1, c2;
            metal_coeff(k, L, c1, c2);
            tally_add(r + 2 * k, pl * c1, L.nat);
            if (c2 != 0.0) tally_add(r + 2 * k + 1, pl * c2, L.nat);
        }
    }
}

template <int M>
__device__ __forceinline__ void flush_mom_cache_lane(const KArgs &a, Lane &L, int lane, double *tv,
                                                     unsigned long long *tt, int *tc)
{
    const bool has = L.s_il >= 0;
    if (!__ballot(has)) return;
    const bool mom = has && L.ms >= 0;
    const unsigned long long key = mom ? mom_key(L.s_il, L.ms) : 0ull;
    const int slot = (int)((key * 0x9E3779B97F4A7C15ull) >> (64 - agg_slot_bits<M>()));
    const bool hit = mom && tt[slot] == key;
    const bool miss = mom && !hit;
    if (miss) tc[slot] = lane;
    wave_sync_lds();
    if (hit) tc[slot] = 64;
    wave_sync_lds();
    const int claim = miss ? tc[slot] : 64;
    const int owner = claim < 64 ? claim : lane;
    const unsigned long long okey = (unsigned long long)__shfl((long long)key, owner);
    const bool in = hit || (miss && claim < 64 && okey == key);
    if (miss && claim == lane) {
        double *t = tv + slot * kAggVals;
        const unsigned long long old = tt[slot];
        if (old != 0ull) mom_row_lane(a, old, t, L.nat);
        tt[slot] = key;
    }
    wave_sync_lds();
    if (in) {
        double *t = tv + slot * kAggVals;
        const double pl = L.s_pl;
        const double t2 = 2.0 * mom_t();
        double pm2 = pl, pm1 = 0.5 * t2 * pl;
        atomicAdd(t + 0, pm2);
        atomicAdd(t + 1, pm1);
#pragma unroll
        for (int q = 2; q < MC_MOM_NM; ++q) {
            const double pq = t2 * pm1 - pm2;
            atomicAdd(t + q, pq);
            pm2 = pm1;
            pm1 = pq;
        }
        if (a.add_dust && !a.mom_dust && L.dabs != 0.0) atomicAdd(t + MC_MOM_NM, pl * L.dabs);
    } else if (has) {
        segment_lane<M>(a, L);
    }
    wave_sync_lds();
    L.s_il = -1;
}

template <int M>
__device__ __forceinline__ void mom_cache_drain_lane(const KArgs &a, int lane, double *tv,
                                                     unsigned long long *tt, uint32_t &nat)
{
    for (int s = lane; s < (1 << agg_slot_bits<M>()); s += 64) {
        if (tt[s] != 0ull) {
            mom_row_lane(a, tt[s], tv + s * kAggVals, nat);
            tt[s] = 0ull;
        }
    }
}

// Issue the tallies of the step's segments, called by all 64 lanes.  With
// direction-coherent packets a leaf group (lanes that scored the same leaf)
// is summed across the wave and its rate sums go out as one row from the
// group's lanes; its band-J contributions (keyed by each packet's energy
// bin) are binned in the wave's LDS row first and go out as one atomic per
// nonzero bin.  Lanes outside a large group go out as rows of their own.
// jbin: this wave's nnu LDS bins, zero between calls.
template <int P, int M, int T>
__device__ __forceinline__ void flush_tally(const KArgs &a, Lane &L, int lane, double *jbin,
                                            double *aggv, unsigned long long *aggt, int *aggc,
                                            int *es)
{
    if constexpr (Prec<P>::mixed && M <= 1) {
        if (a.mom) {
            if constexpr (T == kRowsTogether) flush_mom_cache<M>(a, L, lane, aggv, aggt, aggc, es);
            else flush_mom_cache_lane<M>(a, L, lane, aggv, aggt, aggc);
            return;
        }
    }
    if constexpr (M == 2) {
        // infrared band (src/ion_score_mod.f90:149-158): grain absorption
        // into the band's own path tally (dust_ir_absorb_path), nothing else
        if (L.s_il >= 0) unsafeAtomicAdd(a.irabs + L.s_il, L.s_pl * L.dabs);
        L.s_il = -1;
        return;
    }
    bool has = L.s_il >= 0;
    const int nnu = a.tab->nnu;
    constexpr bool kMom = Prec<P>::mixed && M <= 1;
    if (a.coherent) {
        unsigned long long pend = __ballot(has);
        for (int r = 0; r < (kMom ? kAggRoundsMom : kAggRounds) && pend; ++r) {
            const int leader = __ffsll((long long)pend) - 1;
            const int lil = __shfl(L.s_il, leader);
            const bool in = has && L.s_il == lil;
            const unsigned long long grp = __ballot(in);
            pend &= ~grp;
            if (__popcll(grp) < kAggMin) continue;
            const uint32_t ltid = (threadIdx.x & ~63u) + (uint32_t)leader;
            double *h, *j;
            tally_rows(a, lil, h, j, ltid);
            // the group's rate row: sums of the six coefficients and the
            // grain absorption, lane k issuing sum k
            double mine = 0.0;
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                const double v = in ? L.s_pl * coef(k) : 0.0;
                const double sum = __ballot(v != 0.0) ? wave_sum(v) : 0.0;
                if (lane == k) mine = sum;
            }
            if (a.add_dust) {
                const double v = in ? L.s_pl * L.dabs : 0.0;
                const double sum = __ballot(v != 0.0) ? wave_sum(v) : 0.0;
                if (lane == 6) mine = sum;
            }
            if (lane < 7 && mine != 0.0) {
                unsafeAtomicAdd(h + lane, mine);
                if (lane == 0) ++L.nat;
            }
            if constexpr (M == 1) {
                double *r2 = a.mraw + (size_t)lil * a.ntr * 2;
                for (int k = 0; k < a.ntr; ++k) {
                    const bool on = in && k < L.nmet;
                    if (!__ballot(on)) break;
                    double c1 = 0.0, c2 = 0.0;
                    if (on) metal_coeff(k, L, c1, c2);
                    const double s1 = wave_sum(on ? L.s_pl * c1 : 0.0);
                    const double s2 = wave_sum(on ? L.s_pl * c2 : 0.0);
                    // lanes 0 and 1: the pair (c1, c2) of the transition, one line
                    if (lane == 0 || (lane == 1 && s2 != 0.0))
                        unsafeAtomicAdd(r2 + 2 * k + lane, lane == 0 ? s1 : s2);
                    if (lane == 0) ++L.nat;
                }
            }
            // band J of the group: per-wave bins in LDS, then one atomic
            // per nonzero bin
            if (in) atomicAdd(&jbin[L.inu], L.s_pl);
            wave_sync_lds();
            static_assert(MC_MAXBINS <= 64, "one band bin per lane");
            const double jv = lane < nnu ? jbin[lane] : 0.0;
            if (jv != 0.0) {
                unsafeAtomicAdd(j + lane, jv);
                jbin[lane] = 0.0;
            }
            // requests: one per 8-bin line of the row holding a nonzero bin
            const unsigned long long nz = __ballot(jv != 0.0);
            if (lane == 0)
                for (int g = 0; g < 64; g += 8) L.nat += ((nz >> g) & 0xFFull) != 0ull ? 1u : 0u;
            wave_sync_lds();
            if (in) has = false;
        }
    }
    if constexpr (T == kRowsTogether) coef_rows_out<M>(a, L, lane, has, es);
    else if (has) segment_lane<M>(a, L);
    L.s_il = -1;
}

// ---- car DDA, FP64 (reference arithmetic) ---------------------------------
__device__ __forceinline__ void dda_axis_init(double pos, double k, double org, double d, int n, int &idx,
                                      int &st, double &tm, double &td)
{
#pragma clang fp contract(off)
    idx = min(max((int)((pos - org) / d), 0), n - 1);
    if (k > 0.0) {
        st = 1;
        tm = (org + (double)(idx + 1) * d - pos) / k;
        td = d / k;
    } else if (k < 0.0) {
        st = -1;
        tm = (org + (double)idx * d - pos) / k;
        td = -d / k;
    } else {
        st = 0;
        tm = DBL_MAX;
        td = DBL_MAX;
    }
}

__device__ __forceinline__ void dda_init_d(const DevGrid &g, Lane &L)
{
    dda_axis_init(L.x, L.kx, g.xmin, g.dcell, g.nx, L.ix, L.sx, L.tm[0], L.td[0]);
    dda_axis_init(L.y, L.ky, g.ymin, g.dcell, g.ny, L.iy, L.sy, L.tm[1], L.td[1]);
    dda_axis_init(L.z, L.kz, g.zmin, g.dcell, g.nz, L.iz, L.sz, L.tm[2], L.td[2]);
    L.tcur = 0.0;
}

// distance to the nearest face, tm[ax] for the axis dda_axis_d chose (the
// minimum).  Written as a minimum rather than a select on ax: a select of
// per-lane fields becomes a select of their addresses in the optimiser,
// which keeps the whole lane state in scratch memory.
__device__ __forceinline__ double tm_of(const Lane &L, int)
{
    return fmin(L.tm[0], fmin(L.tm[1], L.tm[2]));
}

// advance along the axis of the nearest face; returns false when outside
__device__ __forceinline__ bool dda_advance_d(const DevGrid &g, Lane &L, int ax)
{
    if (ax == 0) { L.tcur = L.tm[0]; L.tm[0] = L.tm[0] + L.td[0]; L.ix += L.sx; return L.ix >= 0 && L.ix < g.nx; }
    if (ax == 1) { L.tcur = L.tm[1]; L.tm[1] = L.tm[1] + L.td[1]; L.iy += L.sy; return L.iy >= 0 && L.iy < g.ny; }
    L.tcur = L.tm[2];
    L.tm[2] = L.tm[2] + L.td[2];
    L.iz += L.sz;
    return L.iz >= 0 && L.iz < g.nz;
}

__device__ __forceinline__ int dda_axis_d(const Lane &L)
{
    int ax = 0;
    if (L.tm[1] < L.tm[0]) ax = 1;
    if (L.tm[2] < fmin(L.tm[0], L.tm[1])) ax = 2;
    return ax;
}

// ---- car walk, mixed form: FP32 position local to the current cell --------
__device__ __forceinline__ void dda_init_f(const DevGrid &g, Lane &L)
{
    const double gx = (L.x - g.xmin) / g.dcell, gy = (L.y - g.ymin) / g.dcell,
                 gz = (L.z - g.zmin) / g.dcell;
    L.ix = min(max((int)gx, 0), g.nx - 1);
    L.iy = min(max((int)gy, 0), g.ny - 1);
    L.iz = min(max((int)gz, 0), g.nz - 1);
    L.px = fminf(fmaxf((float)(gx - L.ix), 0.0f), 1.0f);
    L.py = fminf(fmaxf((float)(gy - L.iy), 0.0f), 1.0f);
    L.pz = fminf(fmaxf((float)(gz - L.iz), 0.0f), 1.0f);
    L.fkx = (float)L.kx; L.fky = (float)L.ky; L.fkz = (float)L.kz;
    L.sx = L.fkx > 0.0f ? 1 : (L.fkx < 0.0f ? -1 : 0);
    L.sy = L.fky > 0.0f ? 1 : (L.fky < 0.0f ? -1 : 0);
    L.sz = L.fkz > 0.0f ? 1 : (L.fkz < 0.0f ? -1 : 0);
    L.ftcur = 0.0f;
}

// distance (cell units) to the exit face of the current cell and its axis
__device__ __forceinline__ float exit_f(const Lane &L, int &ax)
{
    const float inf = __builtin_huge_valf();
    const float tx = L.sx > 0 ? (1.0f - L.px) / L.fkx : (L.sx < 0 ? -L.px / L.fkx : inf);
    const float ty = L.sy > 0 ? (1.0f - L.py) / L.fky : (L.sy < 0 ? -L.py / L.fky : inf);
    const float tz = L.sz > 0 ? (1.0f - L.pz) / L.fkz : (L.sz < 0 ? -L.pz / L.fkz : inf);
    ax = 0;
    float t = tx;
    if (ty < t) { t = ty; ax = 1; }
    if (tz < t) { t = tz; ax = 2; }
    return fmaxf(t, 0.0f);
}

// move by t (cell units) without leaving the cell
__device__ __forceinline__ void move_f(Lane &L, float t)
{
    L.px = fminf(fmaxf(L.px + t * L.fkx, 0.0f), 1.0f);
    L.py = fminf(fmaxf(L.py + t * L.fky, 0.0f), 1.0f);
    L.pz = fminf(fmaxf(L.pz + t * L.fkz, 0.0f), 1.0f);
}

// cross the face on axis ax; returns false when outside the box
__device__ __forceinline__ bool cross_f(const DevGrid &g, Lane &L, int ax, float t)
{
    move_f(L, t);
    if (ax == 0) { L.ix += L.sx; L.px = L.sx > 0 ? 0.0f : 1.0f; return L.ix >= 0 && L.ix < g.nx; }
    if (ax == 1) { L.iy += L.sy; L.py = L.sy > 0 ? 0.0f : 1.0f; return L.iy >= 0 && L.iy < g.ny; }
    L.iz += L.sz;
    L.pz = L.sz > 0 ? 0.0f : 1.0f;
    return L.iz >= 0 && L.iz < g.nz;
}


// ---- face-exit walk (octree, or car with the octree walk) ----------------
__device__ __forceinline__ double4 node_geom_d(const DevGrid &g, int node)
{
    if (g.car) {
        const int nxy = g.nx * g.ny, iz = node / nxy, rem = node - iz * nxy;
        const int iy = rem / g.nx, ix = rem - iy * g.nx;
        return make_double4(g.xmin + ((double)ix + 0.5) * g.dcell,
                            g.ymin + ((double)iy + 0.5) * g.dcell,
                            g.zmin + ((double)iz + 0.5) * g.dcell, 0.5 * g.dcell);
    }
    return g.node_gd[node];
}

__device__ __forceinline__ float4 node_geom_f(const DevGrid &g, int node)
{
    if (g.car) {
        const double4 d = node_geom_d(g, node);
        const double s = 1.0 / g.half;
        return make_float4((float)(d.x * s), (float)(d.y * s), (float)(d.z * s), (float)(d.w * s));
    }
    return g.node_gf[node];
}

__device__ __forceinline__ int leaf_of(const DevGrid &g, int node)
{
    return g.car ? node : -g.node_link[node] - 1;
}

__device__ __forceinline__ int leaf_node_of(const DevGrid &g, int leaf)
{
    return g.car ? leaf : g.leaf_node[leaf];
}

// raster face neighbour (car_next), -1 outside
__device__ __forceinline__ int car_next_node(const DevGrid &g, int node, int f)
{
    const int nxy = g.nx * g.ny, iz = node / nxy, rem = node - iz * nxy;
    int iy = rem / g.nx, ix = rem - iy * g.nx, jz = iz;
    ix += f == 0 ? 1 : (f == 1 ? -1 : 0);
    iy += f == 2 ? 1 : (f == 3 ? -1 : 0);
    jz += f == 4 ? 1 : (f == 5 ? -1 : 0);
    if (ix < 0 || ix >= g.nx || iy < 0 || iy >= g.ny || jz < 0 || jz >= g.nz) return -1;
    return ix + g.nx * (iy + g.ny * jz);
}

// amr_cell_exit: distance to the first face crossed, first minimum on ties in
// the face order +x,-x,+y,-y,+z,-z (f = 0..5)
__device__ __forceinline__ void cell_exit_d(const double4 G, double x, double y, double z,
                                            double kx, double ky, double kz, double &t, int &f)
{
#pragma clang fp contract(off)
    const double tx = kx > 0.0 ? (G.x + G.w - x) / kx : (kx < 0.0 ? (G.x - G.w - x) / kx : DBL_MAX);
    const double ty = ky > 0.0 ? (G.y + G.w - y) / ky : (ky < 0.0 ? (G.y - G.w - y) / ky : DBL_MAX);
    const double tz = kz > 0.0 ? (G.z + G.w - z) / kz : (kz < 0.0 ? (G.z - G.w - z) / kz : DBL_MAX);
    // first minimum over the faces in order; a zero component leaves both of
    // its faces at DBL_MAX, so with every component zero the +x face wins
    t = tx;
    f = kx < 0.0 ? 1 : 0;
    if (ty < t) { t = ty; f = ky < 0.0 ? 3 : 2; }
    if (tz < t) { t = tz; f = kz < 0.0 ? 5 : 4; }
}

// amr_next_leaf: neighbour across face f, then down to the leaf holding the
// exit point; the crossed axis' octant bit is set by topology
__device__ __forceinline__ int next_node_d(const DevGrid &g, int node, int f, double x, double y,
                                           double z)
{
    if (g.car) return car_next_node(g, node, f);
    int ng = g.node_nbr[node * 6 + f];
    if (ng < 0) return -1;
    for (;;) {
        const int link = g.node_link[ng];
        if (link < 0) return ng;
        const double4 c = g.node_gd[ng];
        const int bx = f < 2 ? (f == 1) : (x >= c.x);
        const int by = (f >> 1) == 1 ? (f == 3) : (y >= c.y);
        const int bz = f >= 4 ? (f == 5) : (z >= c.z);
        ng = link + bx + 2 * by + 4 * bz;
    }
}

// mixed form: exit of the local cube [-1,1]^3 (units of the half width)
__device__ __forceinline__ float exit_oct_f(const Lane &L, int &ax)
{
    const float inf = __builtin_huge_valf();
    const float tx = L.sx > 0 ? (1.0f - L.px) / L.fkx : (L.sx < 0 ? (-1.0f - L.px) / L.fkx : inf);
    const float ty = L.sy > 0 ? (1.0f - L.py) / L.fky : (L.sy < 0 ? (-1.0f - L.py) / L.fky : inf);
    const float tz = L.sz > 0 ? (1.0f - L.pz) / L.fkz : (L.sz < 0 ? (-1.0f - L.pz) / L.fkz : inf);
    ax = 0;
    float t = tx;
    if (ty < t) { t = ty; ax = 1; }
    if (tz < t) { t = tz; ax = 2; }
    return fmaxf(t, 0.0f);
}

__device__ __forceinline__ void move_oct_f(Lane &L, float t)
{
    L.px = fminf(fmaxf(L.px + t * L.fkx, -1.0f), 1.0f);
    L.py = fminf(fmaxf(L.py + t * L.fky, -1.0f), 1.0f);
    L.pz = fminf(fmaxf(L.pz + t * L.fkz, -1.0f), 1.0f);
}

// cross the exit face on axis ax; returns false when the ray left the box
__device__ __forceinline__ bool cross_oct_f(const DevGrid &g, Lane &L, int ax, float t)
{
    move_oct_f(L, t);
    // sign of the exit axis by arithmetic, not by a select of lane fields
    const int s = (ax == 0) * L.sx + (ax == 1) * L.sy + (ax == 2) * L.sz;
    const int f = 2 * ax + (s > 0 ? 0 : 1);
    const float4 G = L.gf;
    int ng = g.car ? car_next_node(g, L.node, f) : g.node_nbr[L.node * 6 + f];
    if (ng < 0) return false;
    if (!g.car) {
        for (;;) {
            const int link = g.node_link[ng];
            if (link < 0) break;
            const float4 C = g.node_gf[ng];
            const int bx = ax == 0 ? (s < 0) : ((G.x - C.x) + L.px * G.w >= 0.0f);
            const int by = ax == 1 ? (s < 0) : ((G.y - C.y) + L.py * G.w >= 0.0f);
            const int bz = ax == 2 ? (s < 0) : ((G.z - C.z) + L.pz * G.w >= 0.0f);
            ng = link + bx + 2 * by + 4 * bz;
        }
    }
    const float4 N = node_geom_f(g, ng);
    const float r = 1.0f / N.w;
    L.px = fminf(fmaxf(((G.x - N.x) + L.px * G.w) * r, -1.0f), 1.0f);
    L.py = fminf(fmaxf(((G.y - N.y) + L.py * G.w) * r, -1.0f), 1.0f);
    L.pz = fminf(fmaxf(((G.z - N.z) + L.pz * G.w) * r, -1.0f), 1.0f);
    if (ax == 0) L.px = s > 0 ? -1.0f : 1.0f;
    else if (ax == 1) L.py = s > 0 ? -1.0f : 1.0f;
    else L.pz = s > 0 ? -1.0f : 1.0f;
    L.node = ng;
    L.leaf = leaf_of(g, ng);
    L.gf = N;
    return true;
}

__device__ __forceinline__ void set_dir_f(Lane &L)
{
    L.fkx = (float)L.kx; L.fky = (float)L.ky; L.fkz = (float)L.kz;
    L.sx = L.fkx > 0.0f ? 1 : (L.fkx < 0.0f ? -1 : 0);
    L.sy = L.fky > 0.0f ? 1 : (L.fky < 0.0f ? -1 : 0);
    L.sz = L.fkz > 0.0f ? 1 : (L.fkz < 0.0f ? -1 : 0);
}

// mixed form: local coordinates of (x,y,z) in the given node
__device__ __forceinline__ void local_of(const DevGrid &g, Lane &L, double x, double y, double z)
{
    const double4 D = node_geom_d(g, L.node);
    L.gf = node_geom_f(g, L.node);
    L.px = fminf(fmaxf((float)((x - D.x) / D.w), -1.0f), 1.0f);
    L.py = fminf(fmaxf((float)((y - D.y) / D.w), -1.0f), 1.0f);
    L.pz = fminf(fmaxf((float)((z - D.z) / D.w), -1.0f), 1.0f);
}

// ---------------------------------------------------------------------------
// G selects the walk: kCarDDA (incremental DDA on a uniform grid) or kFaceExit
// (face-exit walk over the octree node table, or over car cells when the
// octree walk is requested on a car grid)
constexpr int kCarDDA = 0, kFaceExit = 1;

// start a tau-target flight from the current position and direction
template <int P, int G>
__device__ __forceinline__ void start_flight(const KArgs &a, Lane &L, double tau_in)
{
    L.wl = L.wgt * a.Lpk;
    if constexpr (Prec<P>::mixed) {
        if constexpr (G == kCarDDA) dda_init_f(a.g, L);
        else set_dir_f(L);                 // position stays local to the node
        L.ftau = 0.0f;
        L.ftau_in = (float)tau_in;
    } else {
        if constexpr (G == kCarDDA) dda_init_d(a.g, L);
        L.tau = 0.0;
        L.tau_in = tau_in;
    }
    L.phase = 2;
}

// scattering event at the end of a flight that stayed inside
// (transport_ion_packet loop body, FP64 in both forms: events are rare)
template <int P, int G, int M>
__device__ __forceinline__ void scatter_event(const KArgs &a, Lane &L)
{
    const int il = L.leaf;
    if (il < 0) { L.phase = 0; return; }
    const double kap = opacity_d<M>(a, L, il);
    L.wgt = L.wgt * (a.s.rhokap[il] * L.dsca / fmax(kap, DBL_MIN));
    if (L.wgt < kWmin) {
        if (rng_u(a.seed, a.stream, L.id, L.nk++) > kPsurv) { L.phase = 0; return; }
        L.wgt = L.wgt / kPsurv;
    }
    const double gg = L.dg;
    double cost;
    if (fabs(gg) > 1.0e-3) {
        const double q = (1.0 - gg * gg) / (1.0 + gg * (2.0 * rng_u(a.seed, a.stream, L.id, L.nk++) - 1.0));
        cost = (1.0 + gg * gg - q * q) / (2.0 * gg);
    } else {
        cost = 2.0 * rng_u(a.seed, a.stream, L.id, L.nk++) - 1.0;
    }
    cost = fmin(fmax(cost, -1.0), 1.0);
    const double sint = sqrt(1.0 - cost * cost);
    const double phi = MC_TWOPI * rng_u(a.seed, a.stream, L.id, L.nk++);
    double kx = L.kx, ky = L.ky, kz = L.kz;
    double nrm = sqrt(kx * kx + ky * ky + kz * kz);
    if (nrm <= 0.0) { L.phase = 0; return; }
    kx /= nrm; ky /= nrm; kz /= nrm;
    double ux, uy, uz;
    if (fabs(kz) < 0.99) { ux = -ky; uy = kx; uz = 0.0; }
    else { ux = 0.0; uy = -kz; uz = ky; }
    nrm = sqrt(ux * ux + uy * uy + uz * uz);
    ux /= nrm; uy /= nrm; uz /= nrm;
    const double vx = ky * uz - kz * uy, vy = kz * ux - kx * uz, vz = kx * uy - ky * ux;
    const double cp = cos(phi), sp = sin(phi);
    double nx = sint * (cp * ux + sp * vx) + cost * kx;
    double ny = sint * (cp * uy + sp * vy) + cost * ky;
    double nz = sint * (cp * uz + sp * vz) + cost * kz;
    nrm = sqrt(nx * nx + ny * ny + nz * nz);
    L.kx = nx / nrm; L.ky = ny / nrm; L.kz = nz / nrm;
    L.nscatt++;
    const double r = rng_u(a.seed, a.stream, L.id, L.nk++);
    start_flight<P, G>(a, L, -log(fmax(r, DBL_MIN)));
}

// end of the direct walk: energy ledger, then the forced first interaction
// (the direct walk works on a copy of the packet: the forced flight starts
// again from the launch point, src/raytrace_amr.f90:259-269)
template <int P, int G, int M>
__device__ __forceinline__ void end_edge(const KArgs &a, Lane &L, double tau_edge, Counters &cn)
{
    const double lum = a.Lpk * L.wgt;
    const double tr = exp(-fmax(tau_edge, 0.0));
    if constexpr (M == 2) {
        // the reference sums this ledger in extended precision
        // (ENERGY_ACC_WP, src/dust_ir_band_mod.f90:40-50): compensated sums
        // here, emitted = packets x L_packet (weight 1 at the ledger)
        cn.n_emit++;
        two_sum_acc(cn.e_abs, cn.c_abs, lum * (1.0 - tr));
        two_sum_acc(cn.e_esc, cn.c_esc, lum * tr);
    } else {
        led(0) += lum;
        led(1) += lum * (1.0 - tr);
        led(2) += lum * tr;
    }
    if constexpr (M == 2) {
        // dust_ir_escape_packet (src/dust_ir_band_mod.f90:123-150): the
        // escaped luminosity into the bin of the packet's wavelength, the
        // first bin k with lambda < edge(k+1); nothing outside the edges
        const double lesc = lum * tr;
        const int n = a.ir_nl;
        if (lesc > 0.0 && L.lam > a.ir_edge[0] && L.lam < a.ir_edge[n]) {
            int lo = 1, hi = n;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (L.lam < a.ir_edge[mid]) hi = mid;
                else lo = mid + 1;
            }
            unsafeAtomicAdd(a.ir_lesc + (lo - 1), lesc);
        }
    }
    if (!(a.add_dust && a.dust_scatter) || tau_edge <= 1.0e-12) { L.phase = 0; return; }
    const double w1 = 1.0 - exp(-tau_edge);
    L.wgt = L.wgt * w1;
    const double tau = -log(1.0 - rng_u(a.seed, a.stream, L.id, L.nk++) * w1);
    if constexpr (G == kFaceExit) {
        L.leaf = L.leaf0;
        L.node = leaf_node_of(a.g, L.leaf0);
        if constexpr (Prec<P>::mixed) local_of(a.g, L, L.x, L.y, L.z);
        else { L.ox = L.x; L.oy = L.y; L.oz = L.z; }
    }
    start_flight<P, G>(a, L, tau);
}

// Path luminosity of one edge-walk segment in the mixed form, wl
// (e^-tau_in - e^-tau_out) / kappa with the segment terms in FP32.  The
// running attenuation is an FP32 product while tau < kTauF32 (e^-80 is well
// inside the normal FP32 range); past it e^-tau_in is taken in FP64 from the
// FP32 optical depth, so that deep, weakly lit leaves keep the FP64 form's
// dynamic range down to the walk's end at tau = 745.2.
constexpr float kTauF32 = 80.0f;

__device__ __forceinline__ double seg_lum_f(Lane &L, double wlc, float kc, float t, float dt)
{
    if (L.ftau < kTauF32) {
        double pl;
        if (dt > 0.0f) {
            pl = wlc * (double)(L.fexpo * -expm1f(-dt) / kc);
            L.fexpo = L.fexpo * expf(-dt);
        } else {
            pl = wlc * (double)(L.fexpo * t);
        }
        return pl;
    }
    const double ein = exp(-(double)L.ftau);
    return dt > 0.0f ? wlc * ein * (double)(-expm1f(-dt) / kc) : wlc * ein * (double)t;
}

// one segment of the current walk (edge walk or tau flight), car DDA
template <int P, int M>
__device__ __forceinline__ void car_step(const KArgs &a, Lane &L, Counters &cn)
{
    const DevGrid &g = a.g;
    const int il = L.ix + g.nx * (L.iy + g.ny * L.iz);
    if constexpr (!Prec<P>::mixed) {
#pragma clang fp contract(off)
        const int ax = dda_axis_d(L);
        const double seg = tm_of(L, ax) - L.tcur;
        const double kap = opacity_d<M>(a, L, il);
        if (L.phase == 1) {
            double pl;
            if (kap * seg > 0.0) {
                const double eo = L.expo * exp(-kap * seg);
                pl = L.wl * (L.expo - eo) / kap;
                L.expo = eo;
            } else {
                pl = L.wl * L.expo * seg;
            }
            score(a, L, il, pl, cn);
            L.tau = L.tau + kap * seg;
            if (L.tau >= MC_TAU_HUGE || !dda_advance_d(g, L, ax))
                end_edge<P, kCarDDA, M>(a, L, L.tau, cn);
        } else {
            if (L.tau + seg * kap >= L.tau_in) {
                const double ds = kap > 0.0 ? (L.tau_in - L.tau) / kap : seg;
                if (L.nscatt > 0) score(a, L, il, L.wl * ds, cn);
                L.tcur = L.tcur + ds;
                L.x = L.x + L.tcur * L.kx; L.y = L.y + L.tcur * L.ky; L.z = L.z + L.tcur * L.kz;
                L.leaf = il;
                scatter_event<P, kCarDDA, M>(a, L);
                return;
            }
            if (L.nscatt > 0) score(a, L, il, L.wl * seg, cn);
            L.tau = L.tau + seg * kap;
            if (!dda_advance_d(g, L, ax)) L.phase = 0;
        }
    } else {
        int ax;
        const float t = exit_f(L, ax);
        const float kc = opacity_f<M>(a, L, il) * (float)g.dcell;   // per cell length
        const float dt = kc * t;
        const double wlc = L.wl * g.dcell;
        if (L.phase == 1) {
            score(a, L, il, seg_lum_f(L, wlc, kc, t, dt), cn);
            L.ftau = L.ftau + dt;
            if (L.ftau >= (float)MC_TAU_HUGE || !cross_f(g, L, ax, t))
                end_edge<P, kCarDDA, M>(a, L, (double)L.ftau, cn);
        } else {
            if (L.ftau + dt >= L.ftau_in) {
                const float ds = kc > 0.0f ? (L.ftau_in - L.ftau) / kc : t;
                if (L.nscatt > 0) score(a, L, il, wlc * (double)ds, cn);
                move_f(L, ds);
                L.x = g.xmin + ((double)L.ix + (double)L.px) * g.dcell;
                L.y = g.ymin + ((double)L.iy + (double)L.py) * g.dcell;
                L.z = g.zmin + ((double)L.iz + (double)L.pz) * g.dcell;
                L.leaf = il;
                scatter_event<P, kCarDDA, M>(a, L);
                return;
            }
            if (L.nscatt > 0) score(a, L, il, wlc * (double)t, cn);
            L.ftau = L.ftau + dt;
            if (!cross_f(g, L, ax, t)) L.phase = 0;
        }
    }
}

// one segment of the face-exit walk (raytrace_ion_to_edge_amr /
// raytrace_ion_to_tau_amr, src/raytrace_amr.f90:348-486)
template <int P, int M>
__device__ __forceinline__ void oct_step(const KArgs &a, Lane &L, Counters &cn)
{
    const DevGrid &g = a.g;
    const int il = L.leaf;
    if constexpr (!Prec<P>::mixed) {
#pragma clang fp contract(off)
        const double4 G = node_geom_d(g, L.node);
        double t;
        int f;
        cell_exit_d(G, L.ox, L.oy, L.oz, L.kx, L.ky, L.kz, t, f);
        const double kap = opacity_d<M>(a, L, il);
        if (L.phase == 1) {
            double pl;
            if (kap * t > 0.0) {
                const double eo = L.expo * exp(-kap * t);
                pl = L.wl * (L.expo - eo) / kap;
                L.expo = eo;
            } else {
                pl = L.wl * L.expo * t;
            }
            score(a, L, il, pl, cn);
            L.tau = L.tau + kap * t;
            if (L.tau >= MC_TAU_HUGE) { end_edge<P, kFaceExit, M>(a, L, L.tau, cn); return; }
            L.ox = L.ox + t * L.kx; L.oy = L.oy + t * L.ky; L.oz = L.oz + t * L.kz;
            const int nn = next_node_d(g, L.node, f, L.ox, L.oy, L.oz);
            if (nn < 0) { end_edge<P, kFaceExit, M>(a, L, L.tau, cn); return; }
            L.node = nn;
            L.leaf = leaf_of(g, nn);
        } else {
            if (L.tau + t * kap >= L.tau_in) {
                const double ds = kap > 0.0 ? (L.tau_in - L.tau) / kap : t;
                if (L.nscatt > 0) score(a, L, il, L.wl * ds, cn);
                L.ox = L.ox + ds * L.kx; L.oy = L.oy + ds * L.ky; L.oz = L.oz + ds * L.kz;
                scatter_event<P, kFaceExit, M>(a, L);
                return;
            }
            if (L.nscatt > 0) score(a, L, il, L.wl * t, cn);
            L.tau = L.tau + t * kap;
            L.ox = L.ox + t * L.kx; L.oy = L.oy + t * L.ky; L.oz = L.oz + t * L.kz;
            const int nn = next_node_d(g, L.node, f, L.ox, L.oy, L.oz);
            if (nn < 0) { L.phase = 0; return; }
            L.node = nn;
            L.leaf = leaf_of(g, nn);
        }
    } else {
        int ax;
        const float t = exit_oct_f(L, ax);                // units of the half width
        const float hw = L.gf.w * (float)g.half;          // half width, code units
        const float kc = opacity_f<M>(a, L, il) * hw;        // per local unit
        const float dt = kc * t;
        const double wlc = L.wl * (double)hw;
        if (L.phase == 1) {
            score(a, L, il, seg_lum_f(L, wlc, kc, t, dt), cn);
            L.ftau = L.ftau + dt;
            if (L.ftau >= (float)MC_TAU_HUGE || !cross_oct_f(g, L, ax, t))
                end_edge<P, kFaceExit, M>(a, L, (double)L.ftau, cn);
        } else {
            if (L.ftau + dt >= L.ftau_in) {
                const float ds = kc > 0.0f ? (L.ftau_in - L.ftau) / kc : t;
                if (L.nscatt > 0) score(a, L, il, wlc * (double)ds, cn);
                move_oct_f(L, ds);
                scatter_event<P, kFaceExit, M>(a, L);
                return;
            }
            if (L.nscatt > 0) score(a, L, il, wlc * (double)t, cn);
            L.ftau = L.ftau + dt;
            if (!cross_oct_f(g, L, ax, t)) L.phase = 0;
        }
    }
}

template <int P, int G, int M>
__device__ __forceinline__ void load_packet(const KArgs &a, Lane &L, int64_t r, Counters &cn)
{
    const PacketBuf &pk = a.pk;
    if (a.perm) r = a.perm[r];
    if constexpr (M == 2) {
        // infrared records are in leaf order and the brightest leaves emit
        // most packets: walk them in a scattered order (r -> r p mod n with
        // the prime p = 2^32 - 5 > n), so that lanes in flight start in
        // different leaves instead of contending on one leaf's tally
        r = (int64_t)(((uint64_t)r * 4294967291ull) % (uint64_t)a.npk);
    }
    L.id = a.first_id + r;
    L.x = pk.x[r]; L.y = pk.y[r]; L.z = pk.z[r];
    L.kx = pk.kx[r]; L.ky = pk.ky[r]; L.kz = pk.kz[r];
    L.inu = pk.inu[r];
    L.leaf = pk.leaf[r];
    if constexpr (M == 2) {
        // infrared record: the wavelength [um] in E's slot
        // (photon%energy_eV = 1.23984 / lambda); leaf -2 marks a packet of a
        // leaf whose spectrum came back empty, dropped without a ledger entry
        L.lam = pk.E[r];
        L.E = 1.23984 / L.lam;
        if (L.leaf == -2) { L.phase = 0; return; }
    } else {
        L.E = pk.E[r];
    }
    L.leaf0 = L.leaf;
    const double *sg = pk.sig + 6 * r;
    L.sHI = sg[0]; L.sHeI = sg[1]; L.sHeII = sg[2];
    L.dabs = sg[3]; L.dsca = sg[4]; L.dg = sg[5];
    const double he = L.E * MC_EV2ERG;
    if constexpr (M != 2) {
        coef(0) = L.sHI / he;
        coef(1) = L.sHeI / he;
        coef(2) = L.sHeII / he;
        coef(3) = L.sHI * fmax(1.0 - MC_ETH_HI / L.E, 0.0);
        coef(4) = L.sHeI * fmax(1.0 - MC_ETH_HEI / L.E, 0.0);
        coef(5) = L.sHeII * fmax(1.0 - MC_ETH_HEII / L.E, 0.0);
    }
    if constexpr (M == 1) {
        // species_packet_cross_sections (src/species_mod.f90:263-275)
        int n = 0;
        const int ntr = a.ntr;
        for (int k = 0; k < ntr; ++k) {
            if (L.E < c_mt.flat_eth[k]) break;
            msig(k) = m_sigma(&c_mt, c_mt.flat_t[k], L.E);
            n = k + 1;
        }
        L.nmet = n;
        L.rhe = 1.0 / he;
        L.rE = 1.0 / L.E;
    }
    if constexpr (M != 2) {
        // the packet's energy-moment sub-bin and coordinate t in [-1, 1];
        // a packet outside its band bin keeps the per-coefficient sums
        L.ms = -1;
        if constexpr (Prec<P>::mixed) {
            if (a.mom) {
                double t = 0.0;
                L.ms = mom_subbin(a, L.inu, L.E, t);
                if (L.ms >= 0) mom_t() = t;
            }
        }
    }
    L.wgt = 1.0;
    L.nscatt = 0;
    L.nk = M == 2 ? 6 : 0;       // an infrared launch used draws 0-5 of its key
    L.wl = a.Lpk;
    L.phase = 1;
    if constexpr (Prec<P>::mixed) {
        L.fsHI = (float)L.sHI; L.fsHeI = (float)L.sHeI; L.fsHeII = (float)L.sHeII;
        L.fdust = a.add_dust ? (float)(L.dabs + (a.dust_scatter ? L.dsca : 0.0)) : 0.0f;
        L.fexpo = 1.0f;
        L.ftau = 0.0f;
    } else {
        L.expo = 1.0;
        L.tau = 0.0;
    }
    if constexpr (G == kCarDDA) {
        if constexpr (Prec<P>::mixed) dda_init_f(a.g, L);
        else dda_init_d(a.g, L);
    } else {
        if (L.leaf < 0) {                  // launched outside the box: nothing to walk
            end_edge<P, G, M>(a, L, 0.0, cn);
            return;
        }
        L.node = leaf_node_of(a.g, L.leaf);
        if constexpr (Prec<P>::mixed) {
            local_of(a.g, L, L.x, L.y, L.z);
            set_dir_f(L);
        } else {
            L.ox = L.x; L.oy = L.y; L.oz = L.z;
        }
    }
}

constexpr int kWalkChunk = 256;   // records a launch-order wave takes at once
#ifndef MC_COHERENT_REFILL
#define MC_COHERENT_REFILL 64
#endif
constexpr int kCoherentRefill = MC_COHERENT_REFILL;   // idle lanes a sorted wave refills at

// lane 0's value as a wave-uniform (scalar) value
__device__ __forceinline__ unsigned long long wave_first(unsigned long long v)
{
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((unsigned long long)hi << 32) | lo;
}

#ifndef MC_WALK_MINB_0
#define MC_WALK_MINB_0 3
#endif
#ifndef MC_WALK_MINB_1
#define MC_WALK_MINB_1 3
#endif
#ifndef MC_WALK_MINB_2
#define MC_WALK_MINB_2 3
#endif
// blocks per CU the walk's registers are sized for: ionizing walk without and
// with the metal path, infrared walk
constexpr int walk_min_blocks(int M)
{
    return M == 2 ? MC_WALK_MINB_2 : (M == 1 ? MC_WALK_MINB_1 : MC_WALK_MINB_0);
}

// T: issue form of the energy-moment rows (mixed form with moments only):
// kRowsTogether (transposed rows) or kRowsLane (per-lane atomics)
template <int P, int G, int M, int T = kRowsTogether>
__global__ __launch_bounds__(kBlock, walk_min_blocks(M)) void k_walk(KArgs a)
{
    __shared__ double jbins[kBlock / 64][MC_MAXBINS];
    double *jbin = jbins[threadIdx.x >> 6];
    for (int b = threadIdx.x & 63; b < MC_MAXBINS; b += 64) jbin[b] = 0.0;
    // the moment walk's per-wave (leaf, sub-bin) cache (mixed form)
    constexpr int kSlots = (Prec<P>::mixed && M <= 1) ? (1 << agg_slot_bits<M>()) : 1;
    __shared__ double aggvs[kBlock / 64][kSlots * kAggVals];
    __shared__ unsigned long long aggts[kBlock / 64][kSlots];
    __shared__ int aggcs[kBlock / 64][kSlots];
    __shared__ int ess[kBlock / 64][64];
    int *es = ess[threadIdx.x >> 6];
    double *aggv = aggvs[threadIdx.x >> 6];
    unsigned long long *aggt = aggts[threadIdx.x >> 6];
    int *aggc = aggcs[threadIdx.x >> 6];
    for (int b = threadIdx.x & 63; b < kSlots * kAggVals; b += 64) aggv[b] = 0.0;
    for (int b = threadIdx.x & 63; b < kSlots; b += 64) aggt[b] = 0ull;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    Lane L;
    L.phase = 0;
    L.s_il = -1;
    L.nmet = 0;
    L.nat = 0u;
    Counters cn = {0.0, 0.0, 0.0, 0u, 0u, 0u, 0ull, 0.0, 0.0};
    if constexpr (M != 2) {
        led(0) = 0.0;
        led(1) = 0.0;
        led(2) = 0.0;
    }
    const int lane = threadIdx.x & 63;
    // launch-order walks: the wave's own queue of records [qn, qe), taken
    // from the batch counter kWalkChunk at a time, so that refilling a lane
    // does not cost an atomic on the one counter every wave shares
    unsigned long long qn = 0, qe = 0;
    bool dry = false;
    for (;;) {
        const unsigned long long need = __ballot(L.phase == 0);
        unsigned long long r = 0;
        bool take = false;
        if (a.coherent) {
            // direction-sorted records: the wave takes consecutive records
            // for its idle lanes once kCoherentRefill of them are idle (or no
            // lane walks), so that its packets stay together
            const int idle = __popcll(need);
            if (idle > 0 && (idle >= kCoherentRefill || !__ballot(L.phase == 1 || L.phase == 2))) {
                unsigned long long base = 0;
                if (lane == 0) base = atomicAdd(a.s.counter, (unsigned long long)idle);
                base = wave_first(base);
                if (L.phase == 0) {
                    r = base + (unsigned long long)__popcll(need & ((1ull << lane) - 1ull));
                    take = true;
                }
            }
        } else if (need != 0ull) {
            if (!dry && qn >= qe) {
                unsigned long long base = 0;
                if (lane == 0) base = atomicAdd(a.s.counter, (unsigned long long)kWalkChunk);
                base = wave_first(base);
                qn = base;
                qe = base + kWalkChunk;
                dry = (int64_t)base >= a.npk;
            }
            if (dry) {
                if (L.phase == 0) L.phase = 3;
            } else {
                const unsigned long long avail = qe - qn;
                const unsigned long long rank = __popcll(need & ((1ull << lane) - 1ull));
                if (L.phase == 0 && rank < avail) {
                    r = qn + rank;
                    take = true;
                }
                const unsigned long long took = __popcll(need);
                qn += took < avail ? took : avail;
            }
        }
        if (take) {
            if ((int64_t)r < a.npk) load_packet<P, G, M>(a, L, (int64_t)r, cn);
            else L.phase = 3;
        }
        const bool busy = L.phase == 1 || L.phase == 2;
        // launch-order walks refill idle lanes until every lane is drained
        // (a record may leave its lane idle at once: dropped, or outside)
        if (!__ballot(busy) && (a.coherent || !__ballot(L.phase == 0))) break;
        if (busy) {
            if constexpr (G == kCarDDA) car_step<P, M>(a, L, cn);
            else oct_step<P, M>(a, L, cn);
        }
        flush_tally<P, M, T>(a, L, lane, jbin, aggv, aggt, aggc, es);
    }
    if constexpr (Prec<P>::mixed && M <= 1) {
        if (a.mom) {
            wave_sync_lds();
            if constexpr (T == kRowsTogether) mom_cache_drain<M>(a, L, lane, aggv, aggt, es, L.nat);
            else mom_cache_drain_lane<M>(a, lane, aggv, aggt, L.nat);
        }
    }
    if constexpr (M == 2) {
        // compensated wave sums of the infrared ledger into this wave's slot
        // {packets, extinguished, its error, escaped, its error}; the host
        // adds the slots in extended precision
        for (int off = 32; off > 0; off >>= 1) {
            const double oa = __shfl_down(cn.e_abs, off), oca = __shfl_down(cn.c_abs, off);
            const double oe = __shfl_down(cn.e_esc, off), oce = __shfl_down(cn.c_esc, off);
            cn.c_abs += oca;
            two_sum_acc(cn.e_abs, cn.c_abs, oa);
            cn.c_esc += oce;
            two_sum_acc(cn.e_esc, cn.c_esc, oe);
            cn.n_emit += __shfl_down(cn.n_emit, off);
        }
        if (lane == 0) {
            double *w = a.ir_wave + ((size_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6)) * 5;
            w[0] = (double)cn.n_emit;
            w[1] = cn.e_abs; w[2] = cn.c_abs;
            w[3] = cn.e_esc; w[4] = cn.c_esc;
        }
    }
    if constexpr (M != 2) {
        cn.e_emit = led(0);
        cn.e_abs = led(1);
        cn.e_esc = led(2);
    }
    if constexpr (M != 2) {
        // the lanes' tally atomics, one counter atomic per wave
        unsigned long long nat = L.nat;
        for (int off = 32; off > 0; off >>= 1) nat += __shfl_down(nat, off);
        if (lane == 0 && nat) atomicAdd(a.s.segs + 3, nat);
    }
    // wave reduction of the ledger and segment counts, one atomic per wave
    for (int off = 32; off > 0; off >>= 1) {
        cn.e_emit += __shfl_down(cn.e_emit, off);
        cn.e_abs += __shfl_down(cn.e_abs, off);
        cn.e_esc += __shfl_down(cn.e_esc, off);
        cn.sd += __shfl_down(cn.sd, off);
        cn.ss += __shfl_down(cn.ss, off);
        cn.smet += __shfl_down(cn.smet, off);
    }
    if (lane == 0 && M != 2) {
        atomicAdd(a.s.ledger + 0, cn.e_emit);
        atomicAdd(a.s.ledger + 1, cn.e_abs);
        atomicAdd(a.s.ledger + 2, cn.e_esc);
    }
    if (lane == 0) {
        atomicAdd(a.s.segs + 0, (unsigned long long)cn.sd);
        atomicAdd(a.s.segs + 1, (unsigned long long)cn.ss);
        if (cn.smet) atomicAdd(a.s.segs + 2, (unsigned long long)cn.smet);
    }
}

// ---------------------------------------------------------------------------
__global__ void k_zero(double *p, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

__global__ void k_reset_counter(unsigned long long *p) { *p = 0ull; }

__global__ void k_zero_counters(DevState s)
{
    if (threadIdx.x < 3) s.ledger[threadIdx.x] = 0.0;
    if (threadIdx.x < 4) s.segs[threadIdx.x] = 0ull;
    if (threadIdx.x == 0) *s.counter = 0ull;
}

// sum the replicas of each hot leaf (fixed order) into the main tallies; the
// main tallies of hot leaves received no direct updates during the walk
__global__ __launch_bounds__(kBlock) void k_fold_hot(DevState s, int nnu)
{
    const int stride = 8 + nnu;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= (int64_t)s.nhot * stride) return;
    const int hs = (int)(i / stride), k = (int)(i % stride);
    const int il = s.hot_leaf[hs];
    double acc = 0.0;
    if (k < 8) {
        for (int r = 0; r < s.nrep; ++r) acc += s.hot_hhe[((size_t)r * s.nhot + hs) * 8 + k];
        s.hhe[(size_t)il * 8 + k] = acc;
    } else {
        const int b = k - 8;
        for (int r = 0; r < s.nrep; ++r) acc += s.hot_jt[((size_t)r * s.nhot + hs) * nnu + b];
        s.jt[(size_t)il * nnu + b] = acc;
    }
}

// Replica rows of the hot leaves' energy moments added into the leaves' own
// rows before k_mom_fold: one thread per (sub-bin, hot slot), coalesced
// over the slots.
__global__ __launch_bounds__(kBlock) void k_mom_fold_hot(DevState s, double *mom, int64_t rows,
                                                         int ns, int nleaf)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= (int64_t)ns * s.nhot) return;
    const int64_t sb = i / s.nhot;
    const int hs = (int)(i - sb * s.nhot);
    double v[MC_MOM_NM];
#pragma unroll
    for (int m = 0; m < MC_MOM_NM; ++m) v[m] = 0.0;
    for (int r = 0; r < s.nrep; ++r) {
        const double *R = mom + ((size_t)sb * rows + nleaf + (size_t)r * s.nhot + hs) * MC_MOM_NM;
#pragma unroll
        for (int m = 0; m < MC_MOM_NM; ++m) v[m] += R[m];
    }
    if (v[0] == 0.0) return;
    double *M = mom + ((size_t)sb * rows + s.hot_leaf[hs]) * MC_MOM_NM;
#pragma unroll
    for (int m = 0; m < MC_MOM_NM; ++m) M[m] += v[m];
}

// Coefficient sums of every leaf from its energy moments, added to the
// per-coefficient sums of the walk (wave groups and replicated leaves):
// tally_q = sum_s sum_m coef[(s nq + q) MC_MOM_NM + m] M[s][m]; band J is
// the sum of M_0 over the band bin's sub-bins.  One thread per leaf over the
// sub-bins in step across the wave: the sub-bin-major moments are read
// coalesced and the interpolation coefficients are wave-uniform loads (the
// hot leaves' replica rows are already in their own rows, k_mom_fold_hot).
template <int KQ>
__global__ __launch_bounds__(kBlock) void k_mom_fold(DevState s, const double *__restrict__ mom,
                                                     int64_t rows, const double *__restrict__ coef,
                                                     const int32_t *__restrict__ band, int ns,
                                                     int ntr, int ndust, int nnu, double *mraw,
                                                     int nleaf)
{
    const int il = blockIdx.x * kBlock + threadIdx.x;
    if (il >= nleaf) return;
    constexpr int kQ = KQ;
    const int nq = 6 + 2 * ntr + ndust;
    double acc[kQ];
#pragma unroll
    for (int q = 0; q < kQ; ++q) acc[q] = 0.0;
    double *jt = s.jt + (size_t)il * nnu;
    for (int sb = 0; sb < ns; ++sb) {
        const double *M = mom + ((size_t)sb * rows + il) * MC_MOM_NM;
        double v[MC_MOM_NM];
#pragma unroll
        for (int m = 0; m < MC_MOM_NM; ++m) v[m] = M[m];
        if (v[0] == 0.0) continue;           // no segment of this sub-bin
        jt[band[sb]] += v[0];
        const double *cf = coef + (size_t)sb * nq * MC_MOM_NM;
#pragma unroll
        for (int q = 0; q < kQ; ++q) {
            if (q < nq) {
                double t = 0.0;
#pragma unroll
                for (int m = 0; m < MC_MOM_NM; ++m) t += cf[q * MC_MOM_NM + m] * v[m];
                acc[q] += t;
            }
        }
    }
    double *h = s.hhe + (size_t)il * 8;
#pragma unroll
    for (int k = 0; k < 6; ++k) h[k] += acc[k];
    double *r = mraw + (size_t)il * ntr * 2;
#pragma unroll
    for (int q = 6; q < kQ; ++q)
        if (q < 6 + 2 * ntr) r[q - 6] += acc[q];
    // grain absorption (the coefficient after the metal ones)
#pragma unroll
    for (int q = 6; q < kQ; ++q)
        if (ndust && q == nq - 1) h[6] += acc[q];
}

// leaf half width (car: constant)
__device__ __forceinline__ double leaf_half(const DevGrid &g, int il)
{
    return g.car ? 0.5 * g.dcell : g.node_gd[g.leaf_node[il]].w;
}

// gas_rates_compute: rate sums / (leaf volume * d2cm^2)
__global__ __launch_bounds__(kBlock) void k_rates(DevGrid g, DevState s, double d2cm)
{
    const int il = blockIdx.x * kBlock + threadIdx.x;
    if (il >= g.nleaf) return;
    const double w = 2.0 * leaf_half(g, il);
    const double fac = 1.0 / ((w * w) * w * (d2cm * d2cm));
    const double *h = s.hhe + (size_t)il * 8;
    for (int k = 0; k < 6; ++k) s.rates[(size_t)k * g.nleaf + il] = h[k] * fac;
    // ion_score_apply_dust_heat (src/ion_score_mod.f90:319-338)
    if (s.heat_dust) s.heat_dust[il] = h[6] * fac * s.rhokap[il] / d2cm;
}

// ---------------------------------------------------------------------------
// recombination and collisional-ionization coefficients (FP64)
__device__ __forceinline__ double rr_badnell(double T, double A, double B, double T0, double T1, double Cc,
                                    double T2)
{
#pragma clang fp contract(off)
    const double bp = B + Cc * exp(-T2 / T), tt = sqrt(T / T0);
    return A / (tt * pow(1.0 + tt, 1.0 - bp) * pow(1.0 + sqrt(T / T1), 1.0 + bp));
}
__device__ __forceinline__ double rr_mao(double T, double a0, double b0, double c0, double a1, double b1,
                                double a2, double b2)
{
#pragma clang fp contract(off)
    const double te = T * (MC_KBOLTZ / MC_EV2ERG);
    return a0 * 1.0e-10 * pow(te, -b0 - c0 * log(te)) * (1.0 + a2 * pow(te, -b2)) /
           (1.0 + a1 * pow(te, -b1));
}
__device__ __forceinline__ double voronov(double T, double dE, double P, double A, double X, double K)
{
#pragma clang fp contract(off)
    const double U = dE * MC_EV2ERG / (MC_KBOLTZ * T);
    return A * (1.0 + P * sqrt(U)) * pow(U, K) * exp(-U) / (X + U);
}

__device__ RateSet rates_at(double T, bool caseA, bool dere)
{
#pragma clang fp contract(off)
    RateSet r;
    const double aA_H = rr_badnell(T, 8.318e-11, 0.7472, 2.965, 7.001e5, 0.0, 0.0);
    const double aA_He2 = rr_badnell(T, 5.235e-11, 0.6988, 7.301, 4.475e6, 0.0829, 1.682e5) +
                          pow(T, -1.5) * (1.417e-3 * exp(-4.633e5 / T) + 2.235e-4 * exp(-5.532e5 / T) -
                                          2.185e-5 * exp(-8.887e5 / T));
    const double aA_He3 = rr_badnell(T, 1.818e-10, 0.7492, 1.017e1, 2.786e6, 0.0, 0.0);
    if (caseA) {
        r.aH = aA_H; r.aHe2 = aA_He2; r.aHe3 = aA_He3;
    } else {
        r.aH = aA_H - rr_mao(T, 1.87007e-2, 1.29077, -2.17465e-3, 1.29189e1, 8.05060e-1, 8.44399e-2,
                             5.93940e-4);
        r.aHe2 = aA_He2 - rr_mao(T, 3.26862e-3, 6.97361e-1, 3.14907e-4, 5.76341e1, 1.18534,
                                 2.58417e1, 9.90134e-1);
        r.aHe3 = aA_He3 - rr_mao(T, 2.61525e-1, 1.39527, -3.22079e-4, 6.40708e1, 8.97582e-1,
                                 5.01551e-1, 1.21533e-3);
    }
    r.cH = voronov(T, 13.6, 0.0, 2.91e-8, 0.232, 0.39);
    r.cHe1 = voronov(T, 24.6, 0.0, 1.75e-8, 0.180, 0.35);
    r.cHe2 = voronov(T, 54.4, 1.0, 2.05e-9, 0.265, 0.25);
    if (dere) {
        // ci_dere_ratio(1, 1), (2, 1), (2, 2) (src/recomb_mod.f90:296-297)
        r.cH = r.cH * 0.9063;
        r.cHe1 = r.cHe1 * 1.0389;
        r.cHe2 = r.cHe2 * 1.0686;
    }
    return r;
}

// solve_ion_cell without metals: damped fixed point on n_e
__device__ double solve_cell(const RateSet &r, double gH, double gHe1, double gHe2, double nH,
                             double yHe, double &xHI, double &xHeI, double &xHeII)
{
#pragma clang fp contract(off)
    double xHeIII = fmax(0.0, 1.0 - xHeI - xHeII);
    double ne = fmax(nH * ((1.0 - xHI) + yHe * (xHeII + 2.0 * xHeIII)), 1.0e-12 * nH);
    for (int it = 0; it < 200; ++it) {
        const double ne_old = ne;
        const double den = gH + (r.cH + r.aH) * ne;
        xHI = den > 0.0 ? (r.aH * ne) / den : 1.0;
        const double r1 = fmin((gHe1 + r.cHe1 * ne) / (r.aHe2 * ne), 1.0e150);
        const double r2 = fmin((gHe2 + r.cHe2 * ne) / (r.aHe3 * ne), 1.0e150);
        xHeI = 1.0 / (1.0 + r1 + r1 * r2);
        xHeII = xHeI * r1;
        xHeIII = xHeII * r2;
        ne = nH * ((1.0 - xHI) + yHe * (xHeII + 2.0 * xHeIII));
        ne = fmax(0.5 * (ne + ne_old), 1.0e-12 * nH);
        if (fabs(ne - ne_old) <= 1.0e-10 * ne) break;
    }
    return ne;
}

__device__ __forceinline__ float4 opac_of(double nH, double xHI, double xHeI, double xHeII, double rk,
                                 double yHe, double d2cm)
{
    return make_float4((float)(nH * xHI * d2cm), (float)(yHe * nH * xHeI * d2cm),
                       (float)(yHe * nH * xHeII * d2cm), (float)rk);
}

template <int N>
__device__ __forceinline__ void block_reduce(double (&v)[N], bool (&is_max)[N], double *out)
{
    __shared__ double sh[kBlock / 64][N];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int k = 0; k < N; ++k)
        for (int off = 32; off > 0; off >>= 1) {
            const double o = __shfl_down(v[k], off);
            v[k] = is_max[k] ? fmax(v[k], o) : v[k] + o;
        }
    if (lane == 0)
        for (int k = 0; k < N; ++k) sh[wv][k] = v[k];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int k = 0; k < N; ++k) {
            double acc = sh[0][k];
            for (int w = 1; w < kBlock / 64; ++w) acc = is_max[k] ? fmax(acc, sh[w][k]) : acc + sh[w][k];
            out[(size_t)blockIdx.x * N + k] = acc;
        }
}

// ---------------------------------------------------------------------------
// explicit diffuse field (src/diffuse_mod.f90, src/milne_recomb_spectrum_mod.f90)

// temperature bracket of the Milne tables (milne_tbracket): node it, weight w
__device__ __forceinline__ void milne_tbracket(const DevMilne &M, double T, int &it, double &w)
{
    double s = (log10(fmax(T, DBL_MIN)) - M.lgT0) / (M.lgT1 - M.lgT0);
    s = fmin(fmax(s, 0.0), (double)(MC_MILNE_NT - 1));
    const int i = min((int)s + 1, MC_MILNE_NT - 1);
    w = s - (double)(i - 1);
    it = i - 1;
}

// in-band energy radiated per ground-state recombination [eV]
// (milne_energy_per_recomb) or the simpler models of ground_mean_energy
__device__ double ground_mean_energy(const DevMilne &M, int ich, double T)
{
    const double kT = MC_KBOLTZ * T / MC_EV2ERG;
    if (M.model == 1) return M.eth[ich] + kT;
    if (M.model != 2) return M.eth[ich];
    int it;
    double w;
    milne_tbracket(M, T, it, w);
    const double *xb = M.xbar + ich * MC_MILNE_NT, *fi = M.fin + ich * MC_MILNE_NT;
    const double xbar = (1.0 - w) * xb[it] + w * xb[it + 1];
    const double fkeep = (1.0 - w) * fi[it] + w * fi[it + 1];
    return fkeep * (M.eth[ich] + xbar * kT);
}

// photon energy of one diffuse packet (ground_photon_energy with
// milne_sample_energy): the energy CDF interpolated in log T, inverted once
__device__ double ground_photon_energy(const DevMilne &M, int ich, double T, double u)
{
    const double kT = MC_KBOLTZ * T / MC_EV2ERG;
    double E;
    if (M.model == 2) {
        int it;
        double w;
        milne_tbracket(M, T, it, w);
        const double *c0 = M.cdf + ((size_t)ich * MC_MILNE_NT + it) * MC_MILNE_NX;
        const double *c1 = c0 + MC_MILNE_NX;
        const double uu = fmin(fmax(u, 0.0), 1.0);
        int lo = 0, hi = MC_MILNE_NX - 1;
        double ca = (1.0 - w) * c0[lo] + w * c1[lo];
        double cb = (1.0 - w) * c0[hi] + w * c1[hi];
        while (hi - lo > 1) {
            const int mid = (lo + hi + 2) / 2 - 1;
            const double cm = (1.0 - w) * c0[mid] + w * c1[mid];
            if (uu > cm) { lo = mid; ca = cm; } else { hi = mid; cb = cm; }
        }
        const double x = cb > ca ? M.xg[lo] + ((M.xg[hi] - M.xg[lo]) * (uu - ca)) / (cb - ca)
                                 : M.xg[lo];
        E = fmin(M.eth[ich] + x * kT, M.eion_max * (1.0 - 1.0e-12));
    } else if (M.model == 1) {
        E = M.eth[ich] + kT * (-log(fmax(u, DBL_MIN)));
    } else {
        E = M.eth[ich];
    }
    return fmin(E, M.eion_max * 0.999);
}

// ---- He I excited-level recombination (par%hei_diffuse) --------------------
// Every case-B He II recombination cascades to n = 2 and decays through
// 2^3S (19.82 eV line), 2^1P (21.22 eV, 584 A) or 2^1S (two-photon pair);
// only the H-ionizing part of each branch enters the band.

// alpha_fit and hei_branch_fractions (src/hei_cascade_mod.f90:62-81): the
// effective recombination coefficients into each branch, power laws in T
// clamped to [3e3, 2e4] K, as fractions of their sum
__device__ void hei_branch_fractions(double T, double &f3s, double &f1s, double &f1p)
{
#pragma clang fp contract(off)
    const double t = fmin(fmax(T, 3.0e3), 2.0e4) / 1.0e4;
    const double a3 = 2.10e-13 * pow(t, -0.381);
    const double a1s = 2.06e-14 * pow(t, -0.451);
    const double a1p = 4.17e-14 * pow(t, -0.695);
    const double sum = (a3 + a1s) + a1p;
    f3s = a3 / sum;
    f1s = a1s / sum;
    f1p = a1p / sum;
}

// hei_branch_energies (src/diffuse_mod.f90:308-324) with hei_584_ionizes_H
// (src/hei_cascade_mod.f90:115-119): energy radiated inside the band per He II
// recombination by each branch [eV].  The trapped 584 A photon ionizes H with
// probability 1 / (1 + 77 T^-1/2 x_HeI / x_HI); otherwise it is converted to
// 2^1S and leaves as the pair continuum.
__device__ void hei_branch_energies(const DevMilne &M, double T, double xHI, double xHeI,
                                    double &e3s, double &e1p, double &e1s)
{
#pragma clang fp contract(off)
    double f3s, f1s, f1p;
    hei_branch_fractions(T, f3s, f1s, f1p);
    const double p584 = 1.0 / (1.0 + ((77.0 / sqrt(fmax(T, DBL_MIN))) * fmax(xHeI, 0.0)) /
                                         fmax(xHI, DBL_MIN));
    e3s = f3s * MC_HEI_E3S;
    e1p = (f1p * p584) * MC_HEI_E1P;
    e1s = (f1s + f1p * (1.0 - p584)) * M.hei_eper;
}

// hei_2ph_sample_energy (src/hei_twophoton_mod.f90:190-211): inverse of the
// energy-weighted pair CDF on its 1024 nodes, linear inside a node [eV]
__device__ double hei_pair_energy(const DevMilne &M, double u)
{
#pragma clang fp contract(off)
    const double uu = fmin(fmax(u, 0.0), 1.0);
    int lo = 0, hi = MC_HEI2PH_NCDF - 1;
    while (hi - lo > 1) {
        const int mid = (lo + hi) / 2;
        if (uu > M.hei_cdf[mid]) lo = mid; else hi = mid;
    }
    const double ca = M.hei_cdf[lo], cb = M.hei_cdf[hi], ya = M.hei_y[lo];
    const double yv = cb > ca ? ya + ((M.hei_y[hi] - ya) * (uu - ca)) / (cb - ca) : ya;
    return yv * MC_HEI_E_TOT;
}

// diffuse_build, channels 1-4 (src/diffuse_mod.f90:60-121): per-leaf energy
// luminosity of the ground continua n_e n_ion alpha_1(T) E_rec(T) V and of the
// He I excited channel n_e n_HeII alpha_B(T) V sum(E_branch) [erg/s], and the
// leaf total; 5 values per leaf
__global__ __launch_bounds__(kBlock) void k_diffuse_build(DevGrid g, DevState s, DevMilne M,
                                                          double y, double d2cm, double *dif)
{
#pragma clang fp contract(off)
    const int il = blockIdx.x * kBlock + threadIdx.x;
    if (il >= g.nleaf) return;
    double *o = dif + (size_t)il * 5;
    const double nH = s.nH[il];
    if (nH <= 0.0) { o[0] = o[1] = o[2] = o[3] = o[4] = 0.0; return; }
    const double T = s.Te[il], ne = s.ne[il];
    const double sd = (2.0 * leaf_half(g, il)) * d2cm, vol = (sd * sd) * sd;
    const double x3 = fmax(0.0, (1.0 - s.xHeI[il]) - s.xHeII[il]);
    const double nHII = nH * (1.0 - s.xHI[il]);
    const double nHeII = (nH * y) * s.xHeII[il], nHeIII = (nH * y) * x3;
    const RateSet rA = rates_at(T, true, false), rB = rates_at(T, false, false);
    const double a1 = rA.aH - rB.aH, a2 = rA.aHe2 - rB.aHe2, a3 = rA.aHe3 - rB.aHe3;
    o[0] = ((((ne * nHII) * a1) * ground_mean_energy(M, 0, T)) * MC_EV2ERG) * vol;
    o[1] = ((((ne * nHeII) * a2) * ground_mean_energy(M, 1, T)) * MC_EV2ERG) * vol;
    o[2] = ((((ne * nHeIII) * a3) * ground_mean_energy(M, 2, T)) * MC_EV2ERG) * vol;
    double c3 = 0.0;
    if (M.hei) {
        double e3s, e1p, e1s;
        hei_branch_energies(M, T, s.xHI[il], s.xHeI[il], e3s, e1p, e1s);
        c3 = ((((ne * nHeII) * rB.aHe2) * vol) * MC_EV2ERG) * ((e3s + e1p) + e1s);
    }
    o[3] = c3;
    o[4] = ((o[0] + o[1]) + o[2]) + c3;
}

// leaf CDF of the diffuse emission: inclusive scan of the leaf totals in
// three passes (tile scan of kScanTile leaves per block -> scan of the tile
// sums in one block -> offset and normalise)
constexpr int kScanTile = 4 * kBlock;

__global__ __launch_bounds__(kBlock) void k_scan_tiles(const double *val, int stride, int n,
                                                       double *cdf, double *bsum)
{
    __shared__ double sh[kBlock];
    const int base = blockIdx.x * kScanTile + threadIdx.x * 4;
    double v[4], run = 0.0;
    for (int k = 0; k < 4; ++k) {
        const int i = base + k;
        run += i < n ? val[(size_t)i * stride] : 0.0;
        v[k] = run;
    }
    sh[threadIdx.x] = run;
    __syncthreads();
    for (int off = 1; off < kBlock; off <<= 1) {     // Hillis-Steele over thread sums
        const double t = threadIdx.x >= off ? sh[threadIdx.x - off] : 0.0;
        __syncthreads();
        sh[threadIdx.x] += t;
        __syncthreads();
    }
    const double before = threadIdx.x > 0 ? sh[threadIdx.x - 1] : 0.0;
    for (int k = 0; k < 4; ++k) {
        const int i = base + k;
        if (i < n) cdf[i] = before + v[k];
    }
    if (threadIdx.x == kBlock - 1) bsum[blockIdx.x] = sh[kBlock - 1];
}

// exclusive scan of the nb tile sums in place; total into bsum[nb]
__global__ __launch_bounds__(1024) void k_scan_sums(double *bsum, int nb)
{
    __shared__ double sh[1024];
    const int per = (nb + 1023) / 1024, lo = threadIdx.x * per, hi = min(lo + per, nb);
    double run = 0.0;
    for (int i = lo; i < hi; ++i) run += bsum[i];
    sh[threadIdx.x] = run;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const double t = threadIdx.x >= off ? sh[threadIdx.x - off] : 0.0;
        __syncthreads();
        sh[threadIdx.x] += t;
        __syncthreads();
    }
    double acc = threadIdx.x > 0 ? sh[threadIdx.x - 1] : 0.0;
    for (int i = lo; i < hi; ++i) {
        const double v = bsum[i];
        bsum[i] = acc;
        acc += v;
    }
    if (threadIdx.x == 1023) bsum[nb] = sh[1023];
}

__global__ __launch_bounds__(kBlock) void k_scan_apply(double *cdf, int n, const double *bsum,
                                                       int nb)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const double tot = bsum[nb];
    const double v = cdf[i] + bsum[i / kScanTile];
    cdf[i] = tot > 0.0 ? v / tot : 0.0;
}

// gen_diffuse_photon_qmc (src/diffuse_mod.f90:221-296): 9 Sobol uniforms of
// the diffuse stream; emitting leaf from the leaf CDF, uniform position in
// it, isotropic direction, channel by luminosity, Milne energy at the leaf T
// or, in the He I channel, the decay branch by energy and the line or pair
// energy
__global__ __launch_bounds__(kBlock) void k_launch_diffuse(KArgs a)
{
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= a.npk) return;
    const DevTables *T = a.tab;
    const uint32_t q = (uint32_t)(a.first_id + i - 1);
    double u[9];
    for (int d = 0; d < 9; ++d) u[d] = qmc_u(T, q, d, 1);
    const DevGrid &g = a.g;
    int lo = 0, hi = g.nleaf - 1;                  // first leaf with cdf >= u
    while (lo < hi) {
        const int mid = (lo + hi) / 2;
        if (a.dcdf[mid] < u[0]) lo = mid + 1; else hi = mid;
    }
    const int il = lo;
    double cx, cy, cz, half;
    if (g.car) {
        const int ix = il % g.nx, iy = (il / g.nx) % g.ny, iz = il / (g.nx * g.ny);
        cx = g.xmin + ((double)ix + 0.5) * g.dcell;
        cy = g.ymin + ((double)iy + 0.5) * g.dcell;
        cz = g.zmin + ((double)iz + 0.5) * g.dcell;
        half = 0.5 * g.dcell;
    } else {
        const double4 n = g.node_gd[g.leaf_node[il]];
        cx = n.x; cy = n.y; cz = n.z; half = n.w;
    }
    const double x = cx + (2.0 * u[1] - 1.0) * half;
    const double y = cy + (2.0 * u[2] - 1.0) * half;
    const double z = cz + (2.0 * u[3] - 1.0) * half;
    const double cost = 2.0 * u[4] - 1.0, sint = sqrt(1.0 - cost * cost);
    double sp, cp;
    sincos(MC_TWOPI * u[5], &sp, &cp);
    const double *w = a.dif + (size_t)il * 5;
    const double uu = u[6] * (((w[0] + w[1]) + w[2]) + w[3]);
    const int ch = uu <= w[0] ? 0 : (uu <= w[0] + w[1] ? 1 : (uu <= (w[0] + w[1]) + w[2] ? 2 : 3));
    const double Te = a.s.Te[il];
    double E;
    if (ch < 3) {
        E = ground_photon_energy(a.milne, ch, Te, u[7]);
    } else {
        double e3s, e1p, e1s;
        hei_branch_energies(a.milne, Te, a.s.xHI[il], a.s.xHeI[il], e3s, e1p, e1s);
        const double ub = u[7] * ((e3s + e1p) + e1s);
        E = ub <= e3s ? MC_HEI_E3S : (ub <= e3s + e1p ? MC_HEI_E1P : hei_pair_energy(a.milne, u[8]));
    }
    store_packet(a, i, x, y, z, sint * cp, sint * sp, cost, E, il);
}

// gas_equilibrium_update at fixed T; partial metrics per block:
// {max|dx|, sum V|dx|, sum V x_HII}
template <int MT, int LAY>
__global__ __launch_bounds__(kBlock) void k_equilibrium(DevGrid g, DevState s, double T, bool caseA,
                                                        double yHe, double w, double d2cm,
                                                        MetalSw ms)
{
#pragma clang fp contract(off)
    const int il = blockIdx.x * kBlock + threadIdx.x;
    double v[3] = {0.0, 0.0, 0.0};
    bool ismax[3] = {true, false, false};
    if (il < g.nleaf) {
        const double nH = s.nH[il];
        if (nH <= 0.0) {
            s.ne[il] = 0.0;
        } else {
            const RateSet r = rates_at(T, caseA, ms.ci_dere != 0);
            const size_t n = g.nleaf;
            double xHI = s.xHI[il], xHeI = s.xHeI[il], xHeII = s.xHeII[il], ne;
            MCache<(MT > 0 ? MT : 1)> cc;
            if constexpr (MT == 0) {
                ne = solve_cell(r, s.rates[il], s.rates[n + il], s.rates[2 * n + il], nH, yHe, xHI,
                                xHeI, xHeII);
            } else {
                m_prepare<MT>((&c_mt), ms.gam + (size_t)il * ms.ntr, T, cc);
                ne = solve_cell_m<MT, LAY>(ms, cc, r, s.rates[il], s.rates[n + il], s.rates[2 * n + il],
                                      nH, yHe, xHI, xHeI, xHeII);
            }
            const double oHI = s.xHI[il];
            const double nHI = (1.0 - w) * oHI + w * xHI;
            const double nHeI = (1.0 - w) * s.xHeI[il] + w * xHeI;
            const double nHeII = (1.0 - w) * s.xHeII[il] + w * xHeII;
            const double xHeIII = fmax(0.0, 1.0 - nHeI - nHeII);
            s.xHI[il] = nHI;
            s.xHeI[il] = nHeI;
            s.xHeII[il] = nHeII;
            double nen = nH * ((1.0 - nHI) + yHe * (nHeII + 2.0 * xHeIII));
            if constexpr (MT != 0)
                if (ms.metal_ne) nen = charge_neutrality_m<MT, LAY>(cc, nH, nen, nHI, ne);
            s.ne[il] = nen;
            s.opac[il] = opac_of(nH, nHI, nHeI, nHeII, s.rhokap[il], yHe, d2cm);
            const double wd = 2.0 * leaf_half(g, il), vol = (wd * wd) * wd;
            const double d = fabs(nHI - oHI);
            v[0] = d;
            v[1] = vol * d;
            v[2] = vol * (1.0 - nHI);
        }
    }
    block_reduce<3>(v, ismax, s.partials);
}

// ---- thermal balance -------------------------------------------------------
__device__ __forceinline__ double gbar_ff(const DevCool &cl, double T, int Z)
{
#pragma clang fp contract(off)
    const int iz = min(max(Z, 1), 4) - 1;
    const double lt = fmin(fmax(log10(T), cl.logT0), cl.logT0 + 40.0 * cl.dlogT);
    const int i = min((int)((lt - cl.logT0) / cl.dlogT), 39);
    const double t0 = cl.logT0 + i * cl.dlogT;
    const double w = (lt - t0) / cl.dlogT;
    return cl.gff[iz][i] * (1.0 - w) + cl.gff[iz][i + 1] * w;
}

// recombination cooling fits (Hui & Gnedin 1997) for H II and He III
__device__ __forceinline__ double beta_hg(double T, double ttr, double z, bool caseA)
{
#pragma clang fp contract(off)
    const double lam = 2.0 * ttr / T;
    return caseA ? z * 1.778e-29 * T * pow(lam, 1.965) / pow(1.0 + pow(lam / 0.541, 0.502), 2.697)
                 : z * 3.435e-30 * T * pow(lam, 1.970) / pow(1.0 + pow(lam / 2.250, 0.376), 3.720);
}

// cooling_total: recombination, free-free, collisional ionisation and H I
// line cooling [erg cm^-3 s^-1]
__device__ double cooling_total(const DevCool &cl, const RateSet &r, double T, double nH,
                                double ne, double xHI, double xHeI, double xHeII, double y,
                                bool caseA)
{
#pragma clang fp contract(off)
    const double xHeIII = fmax(0.0, 1.0 - xHeI - xHeII);
    const double nHI = nH * xHI, nHII = nH * (1.0 - xHI);
    const double nHeI = nH * y * xHeI, nHeII = nH * y * xHeII, nHeIII = nH * y * xHeIII;
    double cool = ne * (nHII * beta_hg(T, 157807.0, 1.0, caseA) + nHeII * (MC_KBOLTZ * T * r.aHe2) +
                        nHeIII * beta_hg(T, 631515.0, 2.0, caseA));
    cool += 1.42554e-27 * sqrt(T) * ne *
            ((nHII + nHeII) * gbar_ff(cl, T, 1) + 4.0 * nHeIII * gbar_ff(cl, T, 2));
    cool += ne * (nHI * r.cH * MC_ETH_HI + nHeI * r.cHe1 * MC_ETH_HEI + nHeII * r.cHe2 * MC_ETH_HEII) *
            MC_EV2ERG;
    double lam = 0.0;
    for (int i = 0; i < cl.nfit; ++i) lam += cl.A[i] * exp(-cl.Ti[i] / T);
    lam = lam / sqrt(T);
    // cooling_model 'local_ne': the fit suppressed by the H I (T, n_e) table
    // (nlevel_cooling_apply, src/cooling_mod.f90:271-275)
    if (cl.nl_hi >= 0) lam = nl_apply(cl.nlt + (long)cl.nl_hi * kNlSlot, nl_pos(T, ne), lam);
    return cool + ne * nHI * lam;
}

struct CellIn { double nH, gH, gHe1, gHe2, hH, hHe1, hHe2; };

// net heating - cooling at trial T, ionisation state solved at T (warm start
// from the state passed in)
__device__ double net_rate(const DevCool &cl, const CellIn &c, double T, bool caseA, bool dere,
                           double y, double &xHI, double &xHeI, double &xHeII, double &ne)
{
#pragma clang fp contract(off)
    const RateSet r = rates_at(T, caseA, dere);
    ne = solve_cell(r, c.gH, c.gHe1, c.gHe2, c.nH, y, xHI, xHeI, xHeII);
    const double heat = c.nH * (xHI * c.hH + y * (xHeI * c.hHe1 + xHeII * c.hHe2));
    return heat - cooling_total(cl, r, T, c.nH, ne, xHI, xHeI, xHeII, y, caseA);
}

// the same balance with ln(heating / cooling) beside it: the Newton variable
// of the warm-started solve.  Near a root where one side varies
// exponentially with 1/T (collisional line cooling) the log ratio is close
// to linear in ln T while the difference is not.
__device__ double net_rate_lr(const DevCool &cl, const CellIn &c, double T, bool caseA, bool dere,
                              double y, double &xHI, double &xHeI, double &xHeII, double &ne,
                              double &lr)
{
#pragma clang fp contract(off)
    const RateSet r = rates_at(T, caseA, dere);
    ne = solve_cell(r, c.gH, c.gHe1, c.gHe2, c.nH, y, xHI, xHeI, xHeII);
    const double heat = c.nH * (xHI * c.hH + y * (xHeI * c.hHe1 + xHeII * c.hHe2));
    const double cool = cooling_total(cl, r, T, c.nH, ne, xHI, xHeI, xHeII, y, caseA);
    lr = log(fmax(heat, DBL_MIN)) - log(fmax(cool, DBL_MIN));
    return heat - cool;
}

// net_rate with the metal terms (src/thermal_mod.f90:38-71); lr (may be
// null) receives ln(heating / cooling) with the metal heating and cooling on
// their sides.  The rate cache lives only through the solve and the stage
// fractions.
struct NetOut { double net, lr, xHI, xHeI, xHeII, ne; };

template <int MT, int LAY>
__device__ __noinline__ NetOut net_rate_mo(const DevCool &cl, const MetalSw ms, int il,
                                           const CellIn c, double T, bool caseA, double y,
                                           double xHI, double xHeI, double xHeII)
{
#pragma clang fp contract(off)
    double ne;
    const RateSet r = rates_at(T, caseA, ms.ci_dere != 0);
    double lo[MT], up[MT];
    {
        MCache<MT> cc;
        m_prepare<MT>((&c_mt), ms.gam + (size_t)il * ms.ntr, T, cc);
        ne = solve_cell_m<MT, LAY>(ms, cc, r, c.gH, c.gHe1, c.gHe2, c.nH, y, xHI, xHeI, xHeII);
        m_fracs<MT>((&c_mt), cc, ne, c.nH * xHI, c.nH * (1.0 - xHI), lo, up);
    }
    const double heat = c.nH * (xHI * c.hH + y * (xHeI * c.hHe1 + xHeII * c.hHe2));
    const double cool = cooling_total(cl, r, T, c.nH, ne, xHI, xHeI, xHeII, y, caseA);
    const double gff[4] = {gbar_ff(cl, T, 1), gbar_ff(cl, T, 2), gbar_ff(cl, T, 3),
                           gbar_ff(cl, T, 4)};
    double mc, mch, mff, mh;
    m_thermal_terms<MT>(ms, lo, up, ms.heat + (size_t)il * ms.ntr, T, c.nH, ne, c.nH * xHI, gff,
                        mc, mch, mff, mh);
    double net = heat - cool;
    net = net - mc;
    net = net - mch;
    if (ms.metal_ff) net = net - mff;
    if (ms.metal_heat) net = net + mh;
    NetOut o;
    o.net = net;
    o.lr = log(fmax(heat + mh, DBL_MIN)) - log(fmax(cool + mc + mch + mff, DBL_MIN));
    o.xHI = xHI; o.xHeI = xHeI; o.xHeII = xHeII; o.ne = ne;
    return o;
}

// gas_thermal_update: T_e in [te_min, te_max] with net(T) = 0 to a verified
// sign-change bracket thi/tlo - 1 < 1e-5, relaxation on x, T taken directly.
//   W = MC_WARM_NONE   the reference's bisection on log T from the full range;
//   W = MC_WARM_NEWTON safeguarded Newton / secant on log T started at the
//                      leaf's current T_e;
//   W = MC_WARM_MLP    the same started from the network's T_e and x guess.
// Both warm forms keep the reference's edge tests (net(te_min) <= 0 pins
// te_min, net(te_max) >= 0 pins te_max), step only inside the current
// bracket, and accept only a bracket of the bisection's width: every leaf is
// judged by the same criterion as the plain solve; a leaf whose Newton steps
// do not close the bracket within kNewtonMax evaluations continues with
// bisection from the bracket it has (counted as a fallback).
// Partials {max|dx|, max|dT|/T, sum V|dx|, sum V x_HII, sum V ne |dT|,
// sum V ne T, net-rate evaluations of bracketed leaves, fallbacks, bracketed
// leaves (those not pinned at te_min / te_max)}
constexpr int kNewtonMax = 24;

template <int W, int MT, int LAY>
__global__ __launch_bounds__(kBlock) void k_thermal(DevGrid g, DevState s, DevCool cl, bool caseA,
                                                    double y, double w, double tmin, double tmax,
                                                    double d2cm, const float *pred, MetalSw ms,
                                                    int levels)
{
#pragma clang fp contract(off)
    const int il = blockIdx.x * kBlock + threadIdx.x;
    double v[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    bool ismax[9] = {true, true, false, false, false, false, false, false, false};
    if (il < g.nleaf) {
        const size_t n = g.nleaf;
        const double nH = s.nH[il];
        if (nH <= 0.0) {
            s.ne[il] = 0.0;
        } else {
            CellIn c = {nH, s.rates[il], s.rates[n + il], s.rates[2 * n + il],
                        s.rates[3 * n + il], s.rates[4 * n + il], s.rates[5 * n + il]};
            double xHI = s.xHI[il], xHeI = s.xHeI[il], xHeII = s.xHeII[il], ne, te;
            auto net_rate = [&](const DevCool &cl_, const CellIn &c_, double T, bool cA, double y_,
                                double &a1, double &a2, double &a3, double &ne_) -> double {
                if constexpr (MT == 0) {
                    return ::net_rate(cl_, c_, T, cA, ms.ci_dere != 0, y_, a1, a2, a3, ne_);
                } else {
                    const NetOut o = net_rate_mo<MT, LAY>(cl_, ms, il, c_, T, cA, y_, a1, a2, a3);
                    a1 = o.xHI; a2 = o.xHeI; a3 = o.xHeII; ne_ = o.ne;
                    return o.net;
                }
            };
            auto net_rate_lr = [&](const DevCool &cl_, const CellIn &c_, double T, bool cA,
                                   double y_, double &a1, double &a2, double &a3, double &ne_,
                                   double &lr) -> double {
                if constexpr (MT == 0) {
                    return ::net_rate_lr(cl_, c_, T, cA, ms.ci_dere != 0, y_, a1, a2, a3, ne_, lr);
                } else {
                    const NetOut o = net_rate_mo<MT, LAY>(cl_, ms, il, c_, T, cA, y_, a1, a2, a3);
                    a1 = o.xHI; a2 = o.xHeI; a3 = o.xHeII; ne_ = o.ne; lr = o.lr;
                    return o.net;
                }
            };
            double tlo = tmin, thi = tmax;
            int evals = 1, fallback = 0, bracketed = 0;
            // the warm-started forms keep the state of an edge evaluation that
            // decides T_e (the leaf is pinned at te_min or te_max) instead of
            // evaluating that edge again, and skip te_max once te_min decides
            bool at_te = false;
            const double nlo = net_rate(cl, c, tlo, caseA, y, xHI, xHeI, xHeII, ne);
            double nhi = -1.0;
            if (W == MC_WARM_NONE || nlo > 0.0) {
                nhi = net_rate(cl, c, thi, caseA, y, xHI, xHeI, xHeII, ne);
                ++evals;
            }
            if (nlo <= 0.0) {
                te = tlo;
                at_te = W != MC_WARM_NONE;
            } else if (nhi >= 0.0) {
                te = thi;
                at_te = W != MC_WARM_NONE;
            } else if constexpr (W == MC_WARM_NONE) {
                bracketed = 1;
                for (int it = 0; it < 60; ++it) {
                    const double tm = sqrt(tlo * thi);
                    ++evals;
                    if (net_rate(cl, c, tm, caseA, y, xHI, xHeI, xHeII, ne) > 0.0) tlo = tm;
                    else thi = tm;
                    if (thi / tlo - 1.0 < 1.0e-5) break;
                }
                te = sqrt(tlo * thi);
            } else {
                bracketed = 1;
                // the reference's bisection for the first `levels` levels:
                // its root lies in the level's cell [tlo, thi], so the warm
                // root found inside that cell is the bisection's, or another
                // crossing within the cell's width ln(te_max/te_min) / 2^levels
                for (int it = 0; it < levels && thi / tlo - 1.0 >= 1.0e-5; ++it) {
                    const double tm = sqrt(tlo * thi);
                    ++evals;
                    if (net_rate(cl, c, tm, caseA, y, xHI, xHeI, xHeII, ne) > 0.0) tlo = tm;
                    else thi = tm;
                }
                double t0 = s.Te[il];
                if constexpr (W == MC_WARM_MLP) {
                    const float *pr = pred + (size_t)il * 4;
                    t0 = exp10((double)pr[3]);
                    xHI = fmin(fmax(exp10((double)pr[0]), 1.0e-12), 1.0);
                    xHeI = fmin(fmax(exp10((double)pr[1]), 1.0e-12), 1.0);
                    xHeII = fmin(fmax(exp10((double)pr[2]), 1.0e-12), 1.0 - xHeI);
                }
                // Newton / secant on g(ln T) = ln(heating / cooling); the bracket
                // side of every evaluation is decided by the sign of the
                // difference, exactly as in the bisection
                const double ltol = log1p(1.0e-5);
                double u = log(fmin(fmax(t0, tlo * 1.0000001), thi * 0.9999999));
                double f;
                if (net_rate_lr(cl, c, exp(u), caseA, y, xHI, xHeI, xHeII, ne, f) > 0.0) tlo = exp(u);
                else thi = exp(u);
                ++evals;
                // slope from a probe 1e-3 in ln T towards the root
                double up = u + (f > 0.0 ? 1.0e-3 : -1.0e-3);
                up = fmin(fmax(up, log(tlo)), log(thi));
                double fp;
                if (net_rate_lr(cl, c, exp(up), caseA, y, xHI, xHeI, xHeII, ne, fp) > 0.0)
                    tlo = fmax(tlo, exp(up));
                else
                    thi = fmin(thi, exp(up));
                ++evals;
                int k = 0;
                while (thi / tlo - 1.0 >= 1.0e-5 && k < kNewtonMax) {
                    const double ul = log(tlo), uh = log(thi);
                    const double slope = up != u ? (f - fp) / (u - up) : 0.0;
                    double un = slope < 0.0 ? u - f / slope : 0.5 * (ul + uh);
                    if (!(un > ul && un < uh)) un = 0.5 * (ul + uh);
                    // a converged step lands within the tolerance: place it half a
                    // tolerance past the estimate so the bracket closes around it
                    if (fabs(un - u) < 0.5 * ltol) un = f > 0.0 ? un + 0.5 * ltol : un - 0.5 * ltol;
                    if (!(un > ul && un < uh)) un = 0.5 * (ul + uh);
                    up = u; fp = f; u = un;
                    if (net_rate_lr(cl, c, exp(u), caseA, y, xHI, xHeI, xHeII, ne, f) > 0.0) tlo = exp(u);
                    else thi = exp(u);
                    ++evals; ++k;
                }
                if (thi / tlo - 1.0 >= 1.0e-5) {
                    fallback = 1;
                    for (int it = 0; it < 60 && thi / tlo - 1.0 >= 1.0e-5; ++it) {
                        const double tm = sqrt(tlo * thi);
                        ++evals;
                        if (net_rate(cl, c, tm, caseA, y, xHI, xHeI, xHeII, ne) > 0.0) tlo = tm;
                        else thi = tm;
                    }
                }
                te = sqrt(tlo * thi);
            }
            if (!at_te) {
                net_rate(cl, c, te, caseA, y, xHI, xHeI, xHeII, ne);
                ++evals;
            }
            const double oHI = s.xHI[il], oT = s.Te[il];
            const double nHI = (1.0 - w) * oHI + w * xHI;
            const double nHeI = (1.0 - w) * s.xHeI[il] + w * xHeI;
            const double nHeII = (1.0 - w) * s.xHeII[il] + w * xHeII;
            const double x3 = fmax(0.0, 1.0 - nHeI - nHeII);
            double nen = nH * ((1.0 - nHI) + y * (nHeII + 2.0 * x3));
            if constexpr (MT != 0) {
                if (ms.metal_ne) {
                    MCache<MT> cc;
                    m_prepare<MT>((&c_mt), ms.gam + (size_t)il * ms.ntr, te, cc);
                    nen = charge_neutrality_m<MT, LAY>(cc, nH, nen, nHI, ne);
                }
            }
            s.xHI[il] = nHI; s.xHeI[il] = nHeI; s.xHeII[il] = nHeII;
            s.ne[il] = nen; s.Te[il] = te;
            s.opac[il] = opac_of(nH, nHI, nHeI, nHeII, s.rhokap[il], y, d2cm);
            const double wd = 2.0 * leaf_half(g, il), vol = (wd * wd) * wd;
            const double d = fabs(nHI - oHI);
            v[0] = d;
            v[1] = fabs(te - oT) / oT;
            v[2] = vol * d;
            v[3] = vol * (1.0 - nHI);
            v[4] = vol * nen * fabs(te - oT);
            v[5] = vol * nen * te;
            v[6] = bracketed ? (double)evals : 0.0;
            v[7] = (double)fallback;
            v[8] = (double)bracketed;
        }
    }
    block_reduce<9>(v, ismax, s.partials);
}

// hydrogen_balance_residual_grid partials {max eps, sum V eps, sum V}
template <int MT, int LAY>
__global__ __launch_bounds__(kBlock) void k_residual(DevGrid g, DevState s, bool caseA, MetalSw ms)
{
#pragma clang fp contract(off)
    const int il = blockIdx.x * kBlock + threadIdx.x;
    double v[3] = {0.0, 0.0, 0.0};
    bool ismax[3] = {true, false, false};
    if (il < g.nleaf && s.nH[il] > 0.0 && 1.0 - s.xHI[il] >= 1.0e-10) {
        const RateSet r = rates_at(s.Te[il], caseA, ms.ci_dere != 0);
        const double nHI = s.nH[il] * s.xHI[il], nHII = s.nH[il] * (1.0 - s.xHI[il]);
        double rin = 0.0, rout = 0.0;
        if constexpr (MT != 0) {
            if (ms.cx_two_way) {
                MCache<MT> cc;
                double nem;
                m_prepare<MT>((&c_mt), ms.gam + (size_t)il * ms.ntr, s.Te[il], cc);
                m_sums_l<MT, LAY>(cc, s.nH[il], s.ne[il], nHI, nHII, nem, rin, rout);
            }
        }
        const double loss = nHI * (s.rates[il] + r.cH * s.ne[il] + rout);
        const double gain = (r.aH * s.ne[il] + rin) * nHII;
        const double eps = fabs(loss - gain) / fmax(loss, DBL_MIN);
        const double wd = 2.0 * leaf_half(g, il), vol = (wd * wd) * wd;
        v[0] = eps; v[1] = vol * eps; v[2] = vol;
    }
    block_reduce<3>(v, ismax, s.partials);
}

__global__ __launch_bounds__(kBlock) void k_opacity_fill(DevGrid g, DevState s, double yHe, double d2cm)
{
    const int il = blockIdx.x * kBlock + threadIdx.x;
    if (il >= g.nleaf) return;
    s.opac[il] = opac_of(s.nH[il], s.xHI[il], s.xHeI[il], s.xHeII[il], s.rhokap[il], yHe, d2cm);
}

// species_gamma_compute (src/species_mod.f90:493-512): path tallies
// (threshold order) -> metal Gammas and photoheating (element order)
__global__ __launch_bounds__(kBlock) void k_metal_rates(DevGrid g, const double *mraw, double *mgam,
                                                        double *mheat, int ntr, double d2cm)
{
#pragma clang fp contract(off)
    const int il = blockIdx.x * kBlock + threadIdx.x;
    if (il >= g.nleaf) return;
    const double w = 2.0 * leaf_half(g, il);
    const double fac = 1.0 / ((w * w) * w * (d2cm * d2cm));
    for (int k = 0; k < ntr; ++k) {
        const int t = c_mt.flat_t[k];
        const double *r = mraw + ((size_t)il * ntr + k) * 2;
        mgam[(size_t)il * ntr + t] = r[0] * fac;
        mheat[(size_t)il * ntr + t] = r[1] * fac;
    }
}

// species_stage_cache_refresh (src/species_mod.f90:247-261) at each leaf's
// (T_e, n_e, x_HI), and the walk's opacity weights: abund x the fraction of
// each transition's lower stage (threshold order), FP64 and, times n_H d2cm,
// FP32.  Leaves without gas get zero weights.
template <int MT>
__global__ __launch_bounds__(kBlock) void k_stage_refresh(DevGrid g, DevState s, MetalSw ms,
                                                          double *mfrac, double *mop, float *mopf,
                                                          int nst, double d2cm)
{
#pragma clang fp contract(off)
    const int il = blockIdx.x * kBlock + threadIdx.x;
    if (il >= g.nleaf) return;
    const double nH = s.nH[il];
    MCache<MT> cc;
    m_prepare<MT>((&c_mt), ms.gam + (size_t)il * ms.ntr, s.Te[il], cc);
    double *fr = mfrac + (size_t)il * nst;
    double *wp = mop + (size_t)il * ms.ntr;
    float *wf = mopf + (size_t)il * ms.ntr;
    const DevMetalTab *mt = &c_mt;
    double lo[MT], up[MT];
    m_fracs<MT>(mt, cc, s.ne[il], nH * s.xHI[il], nH * (1.0 - s.xHI[il]), lo, up);
    const int ntr = mt->ntr;
#pragma unroll
    for (int u = 0; u < MT; ++u) {
        if (u < ntr) {
            if (mt->t_first[u] == u) fr[mt->t_st[u]] = lo[u];
            fr[mt->t_st[u] + 1] = up[u];
            const double wv = nH > 0.0 ? mt->t_ab[u] * lo[u] : 0.0;
            wp[mt->t_flat[u]] = wv;
            wf[mt->t_flat[u]] = (float)(nH * d2cm * wv);
        }
    }
}

// diagnostic: cells of one ray's direct walk (single thread)

// ---------------------------------------------------------------------------
// dust thermal-infrared band (src/dust_temp_mod.f90,
// src/dust_emis_transport_mod.f90, src/dust_ir_band_mod.f90)

// femit_of (src/dust_temp_mod.f90:132-140): log-log interpolation of the
// equilibrium emission function on its 61 temperature knots (dt = tgrid |
// femit)
__device__ double femit_of(const double *dt, double T)
{
#pragma clang fp contract(off)
    const double *tg = dt, *fe = dt + MC_DT_NT;
    const double l1 = log10(tg[0]), ln = log10(tg[MC_DT_NT - 1]);
    const double lt = fmin(fmax(log10(T), l1), ln);
    const int i = min((int)((lt - l1) / (3.0 / (double)(MC_DT_NT - 1))) + 1, MC_DT_NT - 1);
    const double li = log10(tg[i - 1]), lj = log10(tg[i]);
    const double w = (lt - li) / (lj - li);
    return exp(log(fe[i - 1]) * (1.0 - w) + log(fe[i]) * w);
}

// dust_temp_compute (src/dust_temp_mod.f90:145-177): per leaf, bisection in
// log T for femit(T_d) kd = heat_dust, kd = rhokap / distance2cm
__global__ __launch_bounds__(kBlock) void k_dust_temp(int n, const double *rhokap, const double *heat,
                                                      double *tdust, const double *dt, double d2cm)
{
#pragma clang fp contract(off)
    const int il = blockIdx.x * kBlock + threadIdx.x;
    if (il >= n) return;
    const double kd = rhokap[il] / d2cm, h = heat[il];
    double td = 0.0;
    if (kd > 0.0 && h > 0.0) {
        const double target = h / kd;
        double tlo = dt[0], thi = dt[MC_DT_NT - 1];
        if (femit_of(dt, thi) <= target) {
            td = thi;
        } else {
            for (int it = 1; it <= 60; ++it) {
                const double tm = sqrt(tlo * thi);
                if (femit_of(dt, tm) < target) tlo = tm;
                else thi = tm;
                if (thi / tlo - 1.0 < 1.0e-5) break;
            }
            td = sqrt(tlo * thi);
        }
    }
    tdust[il] = td;
}

// emission luminosity of each leaf, max(heat_dust, 0) V
// (src/dust_emis_transport_mod.f90:246-250)
__global__ __launch_bounds__(kBlock) void k_ir_lum(DevGrid g, const double *heat, double d2cm,
                                                   double *lum)
{
#pragma clang fp contract(off)
    const int il = blockIdx.x * kBlock + threadIdx.x;
    if (il >= g.nleaf) return;
    const double w = 2.0 * leaf_half(g, il) * d2cm;
    lum[il] = fmax(heat[il], 0.0) * ((w * w) * w);
}

// stratified counts (src/dust_emis_transport_mod.f90:267-276): the last
// packet index of each leaf, k_hi = ceiling(N C(il)/L_tot - xi) with C the
// inclusive prefix of the leaf luminosities (tile scan + tile offsets);
// L_tot is the last leaf's prefix computed the same way, so the last leaf's
// ratio is 1 exactly and k_hi = N there.  The parallel prefix is not
// monotone to the last ulp at tile and thread seams, so k_hi is clamped to
// [0, N] here and made non-decreasing by an exact prefix maximum after:
// the leaves' ranges [k_hi(il-1)+1, k_hi(il)] then partition 1..N
__global__ __launch_bounds__(kBlock) void k_ir_counts(int n, const double *lcum, const double *bsum,
                                                      int64_t N, double xi, int64_t *khi)
{
#pragma clang fp contract(off)
    const int il = blockIdx.x * kBlock + threadIdx.x;
    if (il >= n) return;
    const double Ltot = lcum[n - 1] + bsum[(n - 1) / kScanTile];
    const double v = lcum[il] + bsum[il / kScanTile];
    const double k = ceil((double)N * (v / Ltot) - xi);
    khi[il] = k <= 0.0 ? 0 : (k >= (double)N ? N : (int64_t)k);
}

// inclusive prefix maximum of non-negative integers in place (exact in any
// order): tiles of kScanTile, then the tiles' exclusive maxima, then apply
__global__ __launch_bounds__(kBlock) void k_imax_tiles(int64_t *a, int n, int64_t *tmax)
{
    __shared__ int64_t sh[kBlock];
    const int base = blockIdx.x * kScanTile + threadIdx.x * 4;
    int64_t v[4], run = 0;
    for (int k = 0; k < 4; ++k) {
        const int i = base + k;
        if (i < n && a[i] > run) run = a[i];
        v[k] = run;
    }
    sh[threadIdx.x] = run;
    __syncthreads();
    for (int off = 1; off < kBlock; off <<= 1) {
        const int64_t t = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
        __syncthreads();
        if (t > sh[threadIdx.x]) sh[threadIdx.x] = t;
        __syncthreads();
    }
    const int64_t before = threadIdx.x > 0 ? sh[threadIdx.x - 1] : 0;
    for (int k = 0; k < 4; ++k) {
        const int i = base + k;
        if (i < n) a[i] = v[k] > before ? v[k] : before;
    }
    if (threadIdx.x == kBlock - 1) tmax[blockIdx.x] = sh[kBlock - 1];
}

__global__ __launch_bounds__(1024) void k_imax_sums(int64_t *tmax, int nb)
{
    __shared__ int64_t sh[1024];
    const int per = (nb + 1023) / 1024, lo = threadIdx.x * per, hi = min(lo + per, nb);
    int64_t run = 0;
    for (int i = lo; i < hi; ++i) run = tmax[i] > run ? tmax[i] : run;
    sh[threadIdx.x] = run;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int64_t t = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
        __syncthreads();
        if (t > sh[threadIdx.x]) sh[threadIdx.x] = t;
        __syncthreads();
    }
    int64_t acc = threadIdx.x > 0 ? sh[threadIdx.x - 1] : 0;
    for (int i = lo; i < hi; ++i) {
        const int64_t v = tmax[i];
        tmax[i] = acc;
        acc = v > acc ? v : acc;
    }
}

__global__ __launch_bounds__(kBlock) void k_imax_apply(int64_t *a, int n, const int64_t *tmax)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const int64_t t = tmax[i / kScanTile];
    if (t > a[i]) a[i] = t;
}

// B_lambda(T) up to a constant factor, lambda in um
// (planck_lambda, src/dust_emis_transport_mod.f90:418-429)
__device__ __forceinline__ double planck_lambda(double lam_um, double T)
{
#pragma clang fp contract(off)
    const double lc = lam_um * 1.0e-4;
    const double x = 6.62607015e-27 * 2.99792458e10 / (lc * MC_KBOLTZ * T);
    if (x > 700.0) return 0.0;
    const double l2 = lc * lc;
    return 1.0 / ((l2 * l2 * lc) * (exp(x) - 1.0));
}

// sample_linear_spectrum_bin (src/dust_emis_transport_mod.f90:467-483)
__device__ __forceinline__ double sample_linear_bin(double lam0, double lam1, double p0, double p1,
                                                    double u)
{
#pragma clang fp contract(off)
    const double uu = fmin(fmax(u, 0.0), 1.0);
    if (p0 + p1 <= 0.0) return lam0 + uu * (lam1 - lam0);
    const double d = p1 - p0, twoc = uu * (p0 + p1);
    const double disc = p0 * p0 + d * twoc;
    double t = twoc / (p0 + sqrt(fmax(disc, 0.0)));
    t = fmin(fmax(t, 0.0), 1.0);
    return lam0 + t * (lam1 - lam0);
}

struct IrLaunch {
    const double *lam, *sabs, *tdust;
    const int64_t *khi;
    int32_t nl, nzero, pass;
    int64_t b0, b1;                 // packet indices of this batch (1-based, inclusive)
    unsigned long long *fdisc;      // bits of the largest discarded fraction
};

// A leaf's emission spectrum in the wave's LDS (leaf_emission_spectrum with
// the band cut, src/dust_emis_transport_mod.f90:372-413, and
// emission_cumulative, :447-457): pdf(k) = s_abs B_lambda(T_d), the trapezoid
// terms in parallel, their running sum in the reference's order as a
// wave-uniform chain over lane reads.  Returns esum = cdf(nl) (0: empty).
__device__ double ir_leaf_spectrum(const IrLaunch &q, double T, double *pdf, double *cdf, int lane)
{
#pragma clang fp contract(off)
    const int n = q.nl;
    if (!(T > 0.0)) return 0.0;
    for (int k = lane; k < n; k += 64) pdf[k] = q.sabs[k] * planck_lambda(q.lam[k], T);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    double full = 0.0;
    if (q.nzero > 0) {                                   // emission before the cut
        for (int k = lane; k < n - 1; k += 64)
            full += 0.5 * (pdf[k] + pdf[k + 1]) * (q.lam[k + 1] - q.lam[k]);
        full = wave_sum(full);
        for (int k = lane; k < q.nzero; k += 64) pdf[k] = 0.0;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    // term j of lane l is t_{64 j + l}; c_{k+1} = c_k + t_k in order of k
    constexpr int kJ = (MC_NLAM_BB + 63) / 64;
    double t[kJ], cv[kJ];
#pragma unroll
    for (int j = 0; j < kJ; ++j) {
        const int k = 64 * j + lane;
        t[j] = k < n - 1 ? 0.5 * (pdf[k] + pdf[k + 1]) * (q.lam[k + 1] - q.lam[k]) : 0.0;
        cv[j] = 0.0;
    }
    double c = 0.0;
#pragma unroll
    for (int j = 0; j < kJ; ++j) {
        const int lmax = min(64, n - 1 - 64 * j);
        const int lo32 = __double2loint(t[j]), hi32 = __double2hiint(t[j]);
        for (int l = 0; l < lmax; ++l) {
            const double tk = __hiloint2double(__builtin_amdgcn_readlane(hi32, l),
                                               __builtin_amdgcn_readlane(lo32, l));
            c = c + tk;
            if (lane == l) cv[j] = c;                    // cdf(64 j + l + 1), 0-based k + 1
        }
    }
    cdf[0] = 0.0;
#pragma unroll
    for (int j = 0; j < kJ; ++j) {
        const int k = 64 * j + lane;
        if (k < n - 1) cdf[k + 1] = cv[j];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const double esum = c;
    if (esum > 0.0 && q.nzero > 0 && lane == 0) {
        const double f = (full - esum) / full;
        if (f > 0.0) atomicMax(q.fdisc, (unsigned long long)__double_as_longlong(f));
    }
    return esum;
}

constexpr int kIrChunk = 256;    // packets per wave of the infrared launch

// first leaf in [lo, n) whose last packet index is >= k (wave-uniform)
__device__ __forceinline__ int ir_leaf_of(const int64_t *khi, int lo, int n, int64_t k)
{
    int hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (khi[mid] >= k) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

// One wave per range of kIrChunk consecutive packets of the batch, so a leaf
// that emits many packets is launched by many waves: for each leaf in the
// range its spectrum (above), then its packets of the range one per lane
// (launch_and_transport's launch, src/dust_emis_transport_mod.f90:489-544):
// uniform position in the leaf, isotropic direction, wavelength bin by binary
// search, continuous wavelength inside it, the grain optics at the photon
// energy.  Draws 0-5 of key (pass << 40 | k) on stream 3.  Packets of a leaf
// whose spectrum is empty get leaf -2 (dropped by the walk).
__global__ __launch_bounds__(kBlock) void k_ir_launch(KArgs a, IrLaunch q)
{
#pragma clang fp contract(off)
    __shared__ double s_pdf[kBlock / 64][MC_NLAM_BB];
    __shared__ double s_cdf[kBlock / 64][MC_NLAM_BB];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t k0 = q.b0 + ((int64_t)blockIdx.x * (kBlock / 64) + wv) * kIrChunk;
    if (k0 > q.b1) return;
    const int64_t k1 = k0 + kIrChunk - 1 < q.b1 ? k0 + kIrChunk - 1 : q.b1;
    const PacketBuf &pk = a.pk;
    const int n = q.nl, nleaf = a.g.nleaf;
    double *pdf = s_pdf[wv], *cdf = s_cdf[wv];
    int il = ir_leaf_of(q.khi, 0, nleaf, k0);
    for (int64_t ks = k0; ks <= k1;) {
        const int64_t kh = q.khi[il] < k1 ? q.khi[il] : k1;
        const double esum = ir_leaf_spectrum(q, q.tdust[il], pdf, cdf, lane);
        const bool empty = !(esum > 0.0);
        double cx, cy, cz, half;
        if (a.g.car) {
            const int ix = il % a.g.nx, iy = (il / a.g.nx) % a.g.ny, iz = il / (a.g.nx * a.g.ny);
            cx = a.g.xmin + ((double)ix + 0.5) * a.g.dcell;
            cy = a.g.ymin + ((double)iy + 0.5) * a.g.dcell;
            cz = a.g.zmin + ((double)iz + 0.5) * a.g.dcell;
            half = 0.5 * a.g.dcell;
        } else {
            const double4 G = a.g.node_gd[a.g.leaf_node[il]];
            cx = G.x; cy = G.y; cz = G.z; half = G.w;
        }
        for (int64_t k = ks + lane; k <= kh; k += 64) {
            const int64_t r = k - q.b0;
            if (empty) {
                pk.leaf[r] = -2;
                continue;
            }
            const uint64_t key = ((uint64_t)q.pass << 40) | (uint64_t)k;
            pk.x[r] = cx + (2.0 * rng_u(a.seed, 3, key, 0) - 1.0) * half;
            pk.y[r] = cy + (2.0 * rng_u(a.seed, 3, key, 1) - 1.0) * half;
            pk.z[r] = cz + (2.0 * rng_u(a.seed, 3, key, 2) - 1.0) * half;
            const double cost = 2.0 * rng_u(a.seed, 3, key, 3) - 1.0;
            const double sint = sqrt(1.0 - cost * cost);
            const double phi = MC_TWOPI * rng_u(a.seed, 3, key, 4);
            pk.kx[r] = sint * cos(phi);
            pk.ky[r] = sint * sin(phi);
            pk.kz[r] = cost;
            const double u = rng_u(a.seed, 3, key, 5) * cdf[n - 1];
            int blo = 1, bhi = n - 1;                         // 1-based bins
            while (blo < bhi) {
                const int mid = (blo + bhi) >> 1;
                if (u <= cdf[mid]) bhi = mid;
                else blo = mid + 1;
            }
            const int j = blo - 1;
            const double ulo = cdf[j];
            const double ubin = cdf[j + 1] > ulo ? (u - ulo) / (cdf[j + 1] - ulo) : 0.0;
            const double lam = sample_linear_bin(q.lam[j], q.lam[j + 1], pdf[j], pdf[j + 1], ubin);
            double sa, ss, gg;
            dust_at(a, 1.23984 / lam, sa, ss, gg);
            pk.E[r] = lam;
            double *sg = pk.sig + 6 * r;
            sg[0] = 0.0; sg[1] = 0.0; sg[2] = 0.0;
            sg[3] = sa; sg[4] = ss; sg[5] = gg;
            pk.inu[r] = 0;
            pk.leaf[r] = il;
        }
        ks = kh + 1;
        if (ks > k1) break;
        // the leaf's spectrum arrays are rewritten by the next leaf
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // the next emitting leaf is usually among the next few: probe them
        // before a binary search over the rest
        int nx = il + 1;
        while (nx < il + 9 && nx < nleaf - 1 && q.khi[nx] < ks) ++nx;
        il = q.khi[nx] >= ks ? nx : ir_leaf_of(q.khi, nx, nleaf, ks);
    }
}

// dust_ir_heating (src/dust_ir_band_mod.f90:182-192) and the new total
// Heat_dust = EUV part + reabsorbed infrared; partials {0, sum heat_ir V,
// sum heat_dust V} per block
__global__ __launch_bounds__(kBlock) void k_ir_heat(DevGrid g, const double *irabs, const double *rhokap,
                                                    const double *heat_euv, double *heat_ir,
                                                    double *heat, double d2cm, double *partials)
{
#pragma clang fp contract(off)
    const int il = blockIdx.x * kBlock + threadIdx.x;
    double v[3] = {0.0, 0.0, 0.0};
    bool ismax[3] = {true, false, false};
    if (il < g.nleaf) {
        const double s2 = 2.0 * leaf_half(g, il);
        const double fac = 1.0 / ((s2 * s2) * s2 * (d2cm * d2cm));
        const double hir = irabs[il] * fac * rhokap[il] / d2cm;
        const double ht = heat_euv[il] + hir;
        heat_ir[il] = hir;
        heat[il] = ht;
        const double w = s2 * d2cm, vol = (w * w) * w;
        v[1] = hir * vol;
        v[2] = ht * vol;
    }
    block_reduce<3>(v, ismax, partials);
}

#include "mc_lines_dev.h"

// The launch's Sobol/Owen uniforms u(index, d) of one stream (0 stellar,
// 1 diffuse) for n consecutive 0-based indices, row-major n x ndim: the
// same qmc_u as k_launch / k_launch_diffuse, for comparison with the
// reference's tabulated sequences.
__global__ __launch_bounds__(kBlock) void k_qmc_dump(const DevTables *T, uint32_t first, int n,
                                                     int ndim, int stream, double *out)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= (int64_t)n * ndim) return;
    const int r = (int)(i / ndim), d = (int)(i % ndim);
    out[i] = qmc_u(T, first + (uint32_t)r, d, stream);
}

template <int P, int G>
__global__ void k_trace(KArgs a, double x, double y, double z, double kx, double ky, double kz,
                        int32_t *cells, int maxn, int32_t *n)
{
    Lane L;
    L.x = x; L.y = y; L.z = z; L.kx = kx; L.ky = ky; L.kz = kz;
    int k = 0;
    const DevGrid &g = a.g;
    if constexpr (G == kCarDDA) {
        if constexpr (Prec<P>::mixed) {
            dda_init_f(g, L);
            while (k < maxn) {
                cells[k++] = 1 + L.ix + g.nx * (L.iy + g.ny * L.iz);
                int ax;
                const float t = exit_f(L, ax);
                if (!cross_f(g, L, ax, t)) break;
            }
        } else {
            dda_init_d(g, L);
            while (k < maxn) {
                cells[k++] = 1 + L.ix + g.nx * (L.iy + g.ny * L.iz);
                if (!dda_advance_d(g, L, dda_axis_d(L))) break;
            }
        }
    } else {
        L.leaf = find_leaf(g, x, y, z);
        if (L.leaf >= 0) {
            L.node = leaf_node_of(g, L.leaf);
            if constexpr (Prec<P>::mixed) {
                local_of(g, L, x, y, z);
                set_dir_f(L);
                while (k < maxn) {
                    cells[k++] = 1 + L.leaf;
                    int ax;
                    const float t = exit_oct_f(L, ax);
                    if (!cross_oct_f(g, L, ax, t)) break;
                }
            } else {
#pragma clang fp contract(off)
                while (k < maxn) {
                    cells[k++] = 1 + L.leaf;
                    double t;
                    int f;
                    cell_exit_d(node_geom_d(g, L.node), x, y, z, kx, ky, kz, t, f);
                    x = x + t * kx; y = y + t * ky; z = z + t * kz;
                    const int nn = next_node_d(g, L.node, f, x, y, z);
                    if (nn < 0) break;
                    L.node = nn;
                    L.leaf = leaf_of(g, nn);
                }
            }
        }
    }
    *n = k;
}

KArgs make_args(mc_ctx *c)
{
    KArgs a;
    a.tab = c->d_tab;
    a.samp = c->d_samp;
    a.nknots = (int32_t)c->samp_e.size();
    a.dust = c->d_dust;
    a.dn = c->dn;
    a.dcref = c->dcref;
    a.add_dust = c->p.ion_add_dust;
    a.dust_scatter = c->p.ion_dust_scatter;
    a.He_abund = c->p.He_abund;
    a.d2cm = c->p.distance2cm;
    a.seed = (uint32_t)c->p.qmc_seed;
    a.g = c->grid;
    a.s = c->st;
    a.pk = c->pk;
    a.first_id = 1;
    a.npk = 0;
    a.stream = 1;
    a.perm = nullptr;
    a.coherent = 0;
    a.dif = c->d_dif;
    a.dcdf = c->d_dcdf;
    a.milne = c->milne;
    a.ntr = c->ntr;
    a.metal_abs = c->p.ion_metal_abs != 0;
    a.mraw = c->d_mraw;
    a.mop = c->d_mop;
    a.mopf = c->d_mopf;
    a.Lpk = c->Ltot / (double)c->p.nphotons;
    a.irabs = nullptr;
    a.ir_lesc = nullptr;
    a.ir_edge = nullptr;
    a.ir_nl = 0;
    a.ir_wave = nullptr;
    a.mom = nullptr;
    a.mom_rows = 0;
    a.mom_dust = 0;
    a.mom_tab = nullptr;
    a.mom_bseg = nullptr;
    a.mom_ns = 0;
    return a;
}

int persistent_blocks(const void *kern, size_t dyn_lds = 0)
{
    int dev = 0, ncu = 0, per = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kern, kBlock, dyn_lds) != hipSuccess ||
        per < 1)
        per = 2;
    return ncu * per;
}

// metal switches of a kernel; thermal = the thermal update (where the
// warm-started solve applies its noise-band exit)
MetalSw make_sw(mc_ctx *c, bool thermal = false)
{
    MetalSw m;
    m.gam = c->d_mgam;
    m.heat = c->d_mheat;
    m.ntr = c->ntr;
    m.metal_ne = c->p.metal_ne != 0;
    m.cx_two_way = c->p.charge_exchange == MC_CX_TWO_WAY;
    m.metal_heat = c->p.metal_heat != 0;
    m.metal_ff = c->p.metal_freefree != 0;
    m.ci_dere = c->p.ci_model == MC_CI_DERE_HYBRID;
    m.noise_cap = thermal && c->warm_mode != MC_WARM_NONE ? 2 : 0;
    m.nlt = thermal ? c->cool.nlt : nullptr;
    return m;
}

// the context's registry into c_mt, ordered before the kernels that read it
int metal_sync(mc_ctx *c)
{
    if (c->ntr > 0)
        MC_HIP(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_mt), &c->mtab, sizeof(DevMetalTab), 0,
                                      hipMemcpyHostToDevice, c->stream));
    return 0;
}

int sum_partials(mc_ctx *c, int nblk, double out[3])
{
    std::vector<double> h((size_t)nblk * 3);
    MC_HIP(hipMemcpyAsync(h.data(), c->st.partials, sizeof(double) * 3 * nblk,
                          hipMemcpyDeviceToHost, c->stream));
    MC_HIP(hipStreamSynchronize(c->stream));
    out[0] = 0.0; out[1] = 0.0; out[2] = 0.0;
    for (int b = 0; b < nblk; ++b) {      // fixed order: deterministic
        out[0] = std::fmax(out[0], h[(size_t)b * 3]);
        out[1] += h[(size_t)b * 3 + 1];
        out[2] += h[(size_t)b * 3 + 2];
    }
    return 0;
}

}  // namespace

// ---------------------------------------------------------------------------
// fractions of the device's atomic request rate at which the moment rows'
// issue form switches.  Measured on the 20.4 M-leaf multifractal model
// converging from the neutral start, both forms pass by pass: the per-lane
// form is faster up to 69% of the rate (its first pass, at 53%, by 6%),
// the transposed form from 70% on (by 10% at the converged state, where the
// per-lane form runs at 80%; by 26% with every leaf ionised, at 90%)
constexpr double kRowsUpFrac = 0.68, kRowsDownFrac = 0.4;

int mc_run_pass(mc_ctx *c)
{
    if (c->p.diffuse_field && c->p.hei_diffuse && !c->hei_ready) {
        mc_set_error("hei_diffuse needs the He I two-photon table (mc_diffuse_hei_setup)");
        return -1;
    }
    KArgs a = make_args(c);
    const size_t nl = (size_t)c->nleaf;
    hipStream_t st = c->stream;
    if (metal_sync(c)) return -100;
    MC_HIP(hipEventRecord(c->ev[0], st));
    const bool metals = c->ntr > 0;
    // energy-moment tallies in the mixed-form metal walk
    int use_mom = 0;
    if (c->p.precision == MC_MIXED) {
        use_mom = mc_mom_ready(c);
        if (use_mom < 0) return -100;
    }
    if (use_mom) {
        a.mom_dust = c->mom.ndust;
        a.mom = c->d_mom;
        a.mom_rows = (int64_t)nl + (int64_t)c->st.nrep * c->st.nhot;
        a.mom_tab = c->d_mom_tab;
        a.mom_bseg = c->d_mom_bseg;
        a.mom_ns = c->mom.nsub;
        k_zero<<<2048, kBlock, 0, st>>>(
            c->d_mom, ((size_t)nl + (size_t)c->st.nrep * c->st.nhot) * c->mom.nsub * MC_MOM_NM);
    }
    if (metals) k_zero<<<1024, kBlock, 0, st>>>(c->d_mraw, nl * c->ntr * 2);
    k_zero<<<1024, kBlock, 0, st>>>(c->st.hhe, nl * 8);
    k_zero<<<1024, kBlock, 0, st>>>(c->st.jt, nl * c->nnu);
    k_zero_counters<<<1, 64, 0, st>>>(c->st);
    if (c->st.nhot > 0) {
        k_zero<<<256, kBlock, 0, st>>>(c->st.hot_hhe, (size_t)c->st.nrep * c->st.nhot * 8);
        k_zero<<<256, kBlock, 0, st>>>(c->st.hot_jt, (size_t)c->st.nrep * c->st.nhot * c->nnu);
    }
    float t_launch = 0.0f, t_walk = 0.0f;
    const int64_t N = c->p.nphotons;
    const bool face = !c->car || c->p.car_walk_amr;
    const bool mixed = c->p.precision == MC_MIXED;
    // walk kernel: precision form x walk x metal path
    const void *kw;
    if (metals)
        kw = face ? (mixed ? (const void *)k_walk<MC_MIXED, kFaceExit, 1>
                           : (const void *)k_walk<MC_FP64, kFaceExit, 1>)
                  : (mixed ? (const void *)k_walk<MC_MIXED, kCarDDA, 1>
                           : (const void *)k_walk<MC_FP64, kCarDDA, 1>);
    else
        kw = face ? (mixed ? (const void *)k_walk<MC_MIXED, kFaceExit, 0>
                           : (const void *)k_walk<MC_FP64, kFaceExit, 0>)
                  : (mixed ? (const void *)k_walk<MC_MIXED, kCarDDA, 0>
                           : (const void *)k_walk<MC_FP64, kCarDDA, 0>);
    // the moment rows' issue form for this pass (mixed form with moments)
    const bool rows_lane = mixed && use_mom && !c->rows_together;
    if (rows_lane) {
        if (metals)
            kw = face ? (const void *)k_walk<MC_MIXED, kFaceExit, 1, kRowsLane>
                      : (const void *)k_walk<MC_MIXED, kCarDDA, 1, kRowsLane>;
        else
            kw = face ? (const void *)k_walk<MC_MIXED, kFaceExit, 0, kRowsLane>
                      : (const void *)k_walk<MC_MIXED, kCarDDA, 0, kRowsLane>;
    }
    const size_t dlds = (size_t)(kCoefRows + (metals ? c->ntr : 0)) * kBlock * sizeof(double);
    const int nblk = persistent_blocks(kw, dlds);
    const int nlds = c->car ? 0 : c->n_lds_nodes;
    // one batch of packet records: launch kernel, then the persistent walk
    // stellar batches walk in direction order (records sorted by source and
    // direction); diffuse packets, spread over the volume, in launch order
    const bool sorted = c->packet_order == MC_ORDER_DIRECTION;
    auto batch = [&](bool diffuse, int64_t first, int64_t n) -> int {
        a.first_id = first;
        a.npk = n;
        a.stream = diffuse ? 2 : 1;
        const bool coh = sorted && !diffuse;
        a.pk.key = coh ? c->pk.key : nullptr;
        a.perm = coh ? c->d_perm : nullptr;
        a.coherent = coh ? 1 : 0;
        MC_HIP(hipEventRecord(c->ev[1], st));
        if (diffuse)
            k_launch_diffuse<<<(unsigned)((n + kBlock - 1) / kBlock), kBlock, 0, st>>>(a);
        else
            k_launch<<<(unsigned)((n + kBlock - 1) / kBlock), kBlock,
                       (size_t)nlds * (sizeof(double4) + sizeof(int32_t)), st>>>(a, nlds);
        MC_HIP(hipGetLastError());
        if (coh) {
            const int rc = mc_sort_records(c, n);
            if (rc) return rc;
        }
        k_reset_counter<<<1, 1, 0, st>>>(c->st.counter);
        MC_HIP(hipEventRecord(c->ev[2], st));
        void *kargs[] = {&a};
        MC_HIP(hipLaunchKernel(kw, dim3(nblk), dim3(kBlock), kargs, dlds, st));
        MC_HIP(hipGetLastError());
        MC_HIP(hipEventRecord(c->ev[3], st));
        MC_HIP(hipEventSynchronize(c->ev[3]));
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, c->ev[1], c->ev[2]);
        t_launch += ms;
        (void)hipEventElapsedTime(&ms, c->ev[2], c->ev[3]);
        t_walk += ms;
        return 0;
    };
    for (int64_t first = 1; first <= N; first += c->pk.cap) {
        const int rc = batch(false, first, std::min<int64_t>(c->pk.cap, N - first + 1));
        if (rc) return rc;
    }
    c->diffuse_lum = 0.0;
    c->diffuse_nphot = 0;
    if (c->p.diffuse_field) {
        // emission from the current state (diffuse_build), leaf CDF, count
        const int nl1 = c->nleaf;
        MC_HIP(hipEventRecord(c->ev[1], st));
        k_diffuse_build<<<(nl1 + kBlock - 1) / kBlock, kBlock, 0, st>>>(
            c->grid, c->st, c->milne, c->p.He_abund, c->p.distance2cm, c->d_dif);
        MC_HIP(hipGetLastError());
        const int nb = (nl1 + kScanTile - 1) / kScanTile;
        k_scan_tiles<<<nb, kBlock, 0, st>>>(c->d_dif + 4, 5, nl1, c->d_dcdf, c->d_bsum);
        k_scan_sums<<<1, 1024, 0, st>>>(c->d_bsum, nb);
        k_scan_apply<<<(nl1 + kBlock - 1) / kBlock, kBlock, 0, st>>>(c->d_dcdf, nl1, c->d_bsum, nb);
        MC_HIP(hipGetLastError());
        double Ld = 0.0;
        MC_HIP(hipMemcpyAsync(&Ld, c->d_bsum + nb, sizeof(double), hipMemcpyDeviceToHost, st));
        MC_HIP(hipEventRecord(c->ev[2], st));
        MC_HIP(hipStreamSynchronize(st));
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, c->ev[1], c->ev[2]);
        t_launch += ms;
        const double lpk = c->Ltot / (double)N;
        c->diffuse_lum = Ld;
        c->diffuse_nphot = Ld > 0.0 ? (int64_t)std::llround(Ld / lpk) : 0;
        for (int64_t first = 1; first <= c->diffuse_nphot; first += c->pk.cap) {
            const int rc = batch(true, first, std::min<int64_t>(c->pk.cap, c->diffuse_nphot - first + 1));
            if (rc) return rc;
        }
    }
    if (c->st.nhot > 0) {
        const int64_t nf = (int64_t)c->st.nhot * (8 + c->nnu);
        k_fold_hot<<<(unsigned)((nf + kBlock - 1) / kBlock), kBlock, 0, st>>>(c->st, c->nnu);
        MC_HIP(hipGetLastError());
    }
    if (use_mom) {
        const int64_t mrows = (int64_t)c->nleaf + (int64_t)c->st.nrep * c->st.nhot;
        if (c->st.nhot > 0) {
            const int64_t nh = (int64_t)c->mom.nsub * c->st.nhot;
            k_mom_fold_hot<<<(unsigned)((nh + kBlock - 1) / kBlock), kBlock, 0, st>>>(
                c->st, c->d_mom, mrows, c->mom.nsub, c->nleaf);
            MC_HIP(hipGetLastError());
        }
        const unsigned nbf = (unsigned)((c->nleaf + kBlock - 1) / kBlock);
        if (c->ntr > MC_GPU_MT_FAST)
            k_mom_fold<7 + 2 * MC_GPU_MAXTR><<<nbf, kBlock, 0, st>>>(
                c->st, c->d_mom, mrows, c->d_mom_coef, c->d_mom_band, c->mom.nsub, c->ntr,
                c->mom.ndust, c->nnu, c->d_mraw, c->nleaf);
        else
            k_mom_fold<7 + 2 * MC_GPU_MT_FAST><<<nbf, kBlock, 0, st>>>(
                c->st, c->d_mom, mrows, c->d_mom_coef, c->d_mom_band, c->mom.nsub, c->ntr,
                c->mom.ndust, c->nnu, c->d_mraw, c->nleaf);
        MC_HIP(hipGetLastError());
    }
    double led[3];
    unsigned long long sg[4];
    MC_HIP(hipMemcpyAsync(led, c->st.ledger, sizeof(led), hipMemcpyDeviceToHost, st));
    MC_HIP(hipMemcpyAsync(sg, c->st.segs, sizeof(sg), hipMemcpyDeviceToHost, st));
    MC_HIP(hipStreamSynchronize(st));
    for (int k = 0; k < 3; ++k) c->ledger[k] = led[k];
    for (int k = 0; k < 4; ++k) c->segs[k] = (int64_t)sg[k];
    c->times[0] = t_launch;
    c->times[1] = t_walk;
    c->rows_together_last = mixed && use_mom ? (rows_lane ? 0 : 1) : -1;
    if (mixed && use_mom && c->row_issue == MC_ROWS_AUTO && t_walk > 0.0f) {
        // the next pass's issue form, from this pass's request rate: per-lane
        // rows until the walk drives the atomic unit near its ceiling, then
        // transposed rows (about MC_MOM_NM lane atomics per request) until
        // the per-lane form would run well below it; the gap between the two
        // thresholds keeps the choice from flapping
        const double rate = (double)sg[3] / (t_walk * 1e-3);
        if (!c->rows_together && rate > kRowsUpFrac * MC_ATOMIC_REQ_PER_S)
            c->rows_together = true;
        else if (c->rows_together && MC_MOM_NM * rate < kRowsDownFrac * MC_ATOMIC_REQ_PER_S)
            c->rows_together = false;
    }
    return 0;
}

int mc_run_rates(mc_ctx *c)
{
    hipStream_t st = c->stream;
    MC_HIP(hipEventRecord(c->ev[4], st));
    k_rates<<<(c->nleaf + kBlock - 1) / kBlock, kBlock, 0, st>>>(c->grid, c->st, c->p.distance2cm);
    MC_HIP(hipGetLastError());
    if (c->ntr > 0) {
        if (metal_sync(c)) return -100;
        k_metal_rates<<<(c->nleaf + kBlock - 1) / kBlock, kBlock, 0, st>>>(
            c->grid, c->d_mraw, c->d_mgam, c->d_mheat, c->ntr, c->p.distance2cm);
        MC_HIP(hipGetLastError());
    }
    MC_HIP(hipEventRecord(c->ev[5], st));
    MC_HIP(hipEventSynchronize(c->ev[5]));
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, c->ev[4], c->ev[5]);
    c->times[2] = ms;
    return 0;
}

int mc_run_equilibrium(mc_ctx *c, double *max_dx, double *dx_vol)
{
    const int nblk = (c->nleaf + kBlock - 1) / kBlock;
    hipStream_t st = c->stream;
    if (metal_sync(c)) return -100;
    MC_HIP(hipEventRecord(c->ev[6], st));
#define MC_EQ_ARGS c->grid, c->st, c->p.te_fixed, c->p.case_A != 0, c->p.He_abund, \
                   c->p.ion_relax, c->p.distance2cm, make_sw(c)
    if (c->ntr > MC_GPU_MT_FAST) {
        k_equilibrium<MC_GPU_MAXTR, MC_LAYOUT_GENERIC><<<nblk, kBlock, 0, st>>>(MC_EQ_ARGS);
    } else if (c->ntr > 0) {
        if (c->metal_layout == MC_LAYOUT_CNONeSAr)
            k_equilibrium<MC_GPU_MT_FAST, MC_LAYOUT_CNONeSAr><<<nblk, kBlock, 0, st>>>(MC_EQ_ARGS);
        else if (c->metal_layout == MC_LAYOUT_CNONeS)
            k_equilibrium<MC_GPU_MT_FAST, MC_LAYOUT_CNONeS><<<nblk, kBlock, 0, st>>>(MC_EQ_ARGS);
        else
            k_equilibrium<MC_GPU_MT_FAST, MC_LAYOUT_GENERIC><<<nblk, kBlock, 0, st>>>(MC_EQ_ARGS);
    } else {
        k_equilibrium<0, MC_LAYOUT_GENERIC><<<nblk, kBlock, 0, st>>>(MC_EQ_ARGS);
    }
#undef MC_EQ_ARGS
    MC_HIP(hipGetLastError());
    MC_HIP(hipEventRecord(c->ev[7], st));
    double m[3];
    if (sum_partials(c, nblk, m)) return -100;
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, c->ev[6], c->ev[7]);
    c->times[3] = ms;
    *max_dx = m[0];
    *dx_vol = m[1] / std::fmax(m[2], DBL_MIN);
    return 0;
}

int mc_run_thermal(mc_ctx *c, double out[4])
{
    const int nblk = (c->nleaf + kBlock - 1) / kBlock;
    hipStream_t st = c->stream;
    MC_HIP(hipEventRecord(c->ev[6], st));
    const int mode = c->warm_mode;
    if (mode == MC_WARM_MLP) {
        if (!c->mlp.ready) {
            mc_set_error("gas_thermal_update: warm start MC_WARM_MLP needs a network (mc_mlp_set)");
            return -1;
        }
        if (!c->d_feat) {
            MC_HIP(hipMalloc(&c->d_feat, sizeof(float) * (size_t)c->nleaf * MC_MLP_IN));
            MC_HIP(hipMalloc(&c->d_pred, sizeof(float) * (size_t)c->nleaf * MC_MLP_OUT));
        }
        int rc = mc_run_mlp_features(c, c->d_feat);
        if (!rc) rc = mc_run_mlp_infer(c, c->d_feat, c->nleaf, c->d_pred);
        if (rc) return rc;
    }
    const bool caseA = c->p.case_A != 0;
    if (mc_nl_ready(c)) return -1;        // local_ne tables, slots into cool / mtab
    if (metal_sync(c)) return -100;
#define MC_THERMAL_ARGS c->grid, c->st, c->cool, caseA, c->p.He_abund, c->p.ion_relax, \
                        c->p.te_min, c->p.te_max, c->p.distance2cm, c->d_pred, make_sw(c, true), \
                        c->warm_levels
#define MC_THERMAL_W(MT_, LAY_)                                                               \
    do {                                                                                    \
        if (mode == MC_WARM_MLP)                                                            \
            k_thermal<MC_WARM_MLP, MT_, LAY_><<<nblk, kBlock, 0, st>>>(MC_THERMAL_ARGS);    \
        else if (mode == MC_WARM_NEWTON)                                                    \
            k_thermal<MC_WARM_NEWTON, MT_, LAY_><<<nblk, kBlock, 0, st>>>(MC_THERMAL_ARGS); \
        else                                                                                \
            k_thermal<MC_WARM_NONE, MT_, LAY_><<<nblk, kBlock, 0, st>>>(MC_THERMAL_ARGS);   \
    } while (0)
    if (c->ntr > MC_GPU_MT_FAST) {
        MC_THERMAL_W(MC_GPU_MAXTR, MC_LAYOUT_GENERIC);
    } else if (c->ntr > 0) {
        if (c->metal_layout == MC_LAYOUT_CNONeSAr) MC_THERMAL_W(MC_GPU_MT_FAST, MC_LAYOUT_CNONeSAr);
        else if (c->metal_layout == MC_LAYOUT_CNONeS) MC_THERMAL_W(MC_GPU_MT_FAST, MC_LAYOUT_CNONeS);
        else MC_THERMAL_W(MC_GPU_MT_FAST, MC_LAYOUT_GENERIC);
    } else {
        MC_THERMAL_W(0, MC_LAYOUT_GENERIC);
    }
#undef MC_THERMAL_W
#undef MC_THERMAL_ARGS
    MC_HIP(hipGetLastError());
    MC_HIP(hipEventRecord(c->ev[7], st));
    std::vector<double> h((size_t)nblk * 9);
    MC_HIP(hipMemcpyAsync(h.data(), c->st.partials, sizeof(double) * 9 * nblk,
                          hipMemcpyDeviceToHost, st));
    MC_HIP(hipStreamSynchronize(st));
    double m[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int b = 0; b < nblk; ++b) {
        m[0] = std::fmax(m[0], h[(size_t)b * 9]);
        m[1] = std::fmax(m[1], h[(size_t)b * 9 + 1]);
        for (int k = 2; k < 9; ++k) m[k] += h[(size_t)b * 9 + k];
    }
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, c->ev[6], c->ev[7]);
    c->times[3] = ms;
    c->thermal_stats[0] = (int64_t)m[6];
    c->thermal_stats[1] = (int64_t)m[7];
    c->thermal_stats[2] = (int64_t)m[8];
    c->thermal_stats[3] = c->nleaf;
    out[0] = m[0];
    out[1] = m[1];
    out[2] = m[2] / std::fmax(m[3], DBL_MIN);
    out[3] = m[4] / std::fmax(m[5], DBL_MIN);
    return 0;
}

int mc_run_residual(mc_ctx *c, double *eps_max, double *eps_vol)
{
    const int nblk = (c->nleaf + kBlock - 1) / kBlock;
    if (metal_sync(c)) return -100;
    if (c->ntr > MC_GPU_MT_FAST) {
        k_residual<MC_GPU_MAXTR, MC_LAYOUT_GENERIC><<<nblk, kBlock, 0, c->stream>>>(
            c->grid, c->st, c->p.case_A != 0, make_sw(c));
    } else if (c->ntr > 0) {
        if (c->metal_layout == MC_LAYOUT_CNONeSAr)
            k_residual<MC_GPU_MT_FAST, MC_LAYOUT_CNONeSAr><<<nblk, kBlock, 0, c->stream>>>(
                c->grid, c->st, c->p.case_A != 0, make_sw(c));
        else if (c->metal_layout == MC_LAYOUT_CNONeS)
            k_residual<MC_GPU_MT_FAST, MC_LAYOUT_CNONeS><<<nblk, kBlock, 0, c->stream>>>(
                c->grid, c->st, c->p.case_A != 0, make_sw(c));
        else
            k_residual<MC_GPU_MT_FAST, MC_LAYOUT_GENERIC><<<nblk, kBlock, 0, c->stream>>>(
                c->grid, c->st, c->p.case_A != 0, make_sw(c));
    } else {
        k_residual<0, MC_LAYOUT_GENERIC><<<nblk, kBlock, 0, c->stream>>>(c->grid, c->st,
                                                                         c->p.case_A != 0, make_sw(c));
    }
    MC_HIP(hipGetLastError());
    double m[3];
    if (sum_partials(c, nblk, m)) return -100;
    *eps_max = m[0];
    *eps_vol = m[1] / std::fmax(m[2], DBL_MIN);
    return 0;
}

int mc_run_opacity_fill(mc_ctx *c)
{
    k_opacity_fill<<<(c->nleaf + kBlock - 1) / kBlock, kBlock, 0, c->stream>>>(
        c->grid, c->st, c->p.He_abund, c->p.distance2cm);
    MC_HIP(hipGetLastError());
    if (c->ntr > 0) return mc_run_stage_refresh(c);
    MC_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

int mc_run_stage_refresh(mc_ctx *c)
{
    if (c->ntr <= 0) return 0;
    if (metal_sync(c)) return -100;
    const unsigned nb = (unsigned)((c->nleaf + kBlock - 1) / kBlock);
    if (c->ntr > MC_GPU_MT_FAST)
        k_stage_refresh<MC_GPU_MAXTR><<<nb, kBlock, 0, c->stream>>>(
            c->grid, c->st, make_sw(c), c->d_mfrac, c->d_mop, c->d_mopf, c->nst, c->p.distance2cm);
    else
        k_stage_refresh<MC_GPU_MT_FAST><<<nb, kBlock, 0, c->stream>>>(
            c->grid, c->st, make_sw(c), c->d_mfrac, c->d_mop, c->d_mopf, c->nst, c->p.distance2cm);
    MC_HIP(hipGetLastError());
    MC_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

int mc_run_trace(mc_ctx *c, int precision, const double *pos, const double *dir, int32_t *cells,
                 int maxn, int32_t *n)
{
    KArgs a = make_args(c);
    int32_t *dcells = nullptr, *dn = nullptr;
    MC_HIP(hipMalloc(&dcells, sizeof(int32_t) * (size_t)maxn));
    MC_HIP(hipMalloc(&dn, sizeof(int32_t)));
    const bool face = !c->car || c->p.car_walk_amr;
    auto kern = face ? (precision == MC_MIXED ? k_trace<MC_MIXED, kFaceExit> : k_trace<MC_FP64, kFaceExit>)
                     : (precision == MC_MIXED ? k_trace<MC_MIXED, kCarDDA> : k_trace<MC_FP64, kCarDDA>);
    kern<<<1, 1, 0, c->stream>>>(a, pos[0], pos[1], pos[2], dir[0], dir[1], dir[2], dcells, maxn, dn);
    MC_HIP(hipGetLastError());
    MC_HIP(hipMemcpyAsync(n, dn, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    MC_HIP(hipStreamSynchronize(c->stream));
    MC_HIP(hipMemcpy(cells, dcells, sizeof(int32_t) * (size_t)(*n), hipMemcpyDeviceToHost));
    (void)hipFree(dcells);
    (void)hipFree(dn);
    return 0;
}

// ---- mc_trace_ray without device allocation -------------------------------
// The cell list and its count go to a file-scope device array of fixed size
// (each GPU holds its own copy until the process ends); host code reaches it
// only through the address hipGetSymbolAddress returns for the current
// device, looked up once per device.  Nothing in it must start at zero: the
// kernel writes the count and the cells before they are read.  Calls that
// overlap on one GPU share the array and race.
static __device__ __align__(256) int32_t scratch[MC_TRACE_NOALLOC_MAX + 64];

namespace {
constexpr int kMaxDevices = 64;
int32_t *g_scratch[kMaxDevices];
std::mutex g_scratch_mu;

int scratch_ptr(int32_t **out)
{
    int dev = 0;
    MC_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        mc_set_error("mc_trace_ray_noalloc: device index out of range");
        return -1;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            mc_set_error(std::string("mc_trace_ray_noalloc: hipGetSymbolAddress: ") +
                         hipGetErrorString(e));
            return -1;
        }
        g_scratch[dev] = static_cast<int32_t *>(p);
    }
    *out = g_scratch[dev];
    return 0;
}
}  // namespace

int mc_run_trace_noalloc(mc_ctx *c, int precision, const double *pos, const double *dir,
                         int32_t *cells, int maxn, int32_t *n)
{
    int32_t *buf = nullptr;
    if (scratch_ptr(&buf)) return -1;
    int32_t *dcells = buf, *dn = buf + MC_TRACE_NOALLOC_MAX;     // count after the cells
    KArgs a = make_args(c);
    const bool face = !c->car || c->p.car_walk_amr;
    auto kern = face ? (precision == MC_MIXED ? k_trace<MC_MIXED, kFaceExit> : k_trace<MC_FP64, kFaceExit>)
                     : (precision == MC_MIXED ? k_trace<MC_MIXED, kCarDDA> : k_trace<MC_FP64, kCarDDA>);
    kern<<<1, 1, 0, c->stream>>>(a, pos[0], pos[1], pos[2], dir[0], dir[1], dir[2], dcells, maxn, dn);
    MC_HIP(hipGetLastError());
    MC_HIP(hipMemcpyAsync(n, dn, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    MC_HIP(hipStreamSynchronize(c->stream));
    if (*n > 0)
        MC_HIP(hipMemcpyAsync(cells, dcells, sizeof(int32_t) * (size_t)(*n), hipMemcpyDeviceToHost,
                              c->stream));
    MC_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

// dust_temp_compute over the grid from a device Heat_dust array
int mc_run_dust_temp(mc_ctx *c, const double *d_heat)
{
    const IrBand &b = c->ir;
    k_dust_temp<<<(c->nleaf + kBlock - 1) / kBlock, kBlock, 0, c->stream>>>(
        c->nleaf, c->st.rhokap, d_heat, b.d_tdust, b.d_dt, c->p.distance2cm);
    MC_HIP(hipGetLastError());
    MC_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

// transport_dust_emission (src/dust_emis_transport_mod.f90:208-366) on the
// device-resident Heat_dust.  Per pass: leaf luminosities and their scan,
// stratified counts, then batches of launch records (one wave per emitting
// leaf) walked by the infrared walk variant; then heat_ir, the new
// Heat_dust, T_dust and the ledger / luminosity sums that decide the
// iteration.
int mc_run_dust_emission(mc_ctx *c, double rep[MC_DUSTRT_NREPORT])
{
    IrBand &b = c->ir;
    hipStream_t st = c->stream;
    const int nl = c->nleaf;
    for (int k = 0; k < MC_DUSTRT_NREPORT; ++k) rep[k] = 0.0;
    const int64_t N = (int64_t)std::fmax(c->p.dust_no_photons, 1.0);
    const int niter = std::max(c->p.dust_niter, 1);
    MC_HIP(hipMemcpyAsync(b.d_heat_euv, c->st.heat_dust, sizeof(double) * nl,
                          hipMemcpyDeviceToDevice, st));
    KArgs a = make_args(c);
    a.stream = 3;
    a.perm = nullptr;
    a.coherent = 0;
    a.pk.key = nullptr;
    a.ntr = 0;
    a.irabs = b.d_irabs;
    a.ir_lesc = b.d_lesc;
    a.ir_edge = b.d_edge;
    a.ir_nl = b.nl;
    a.s.ledger = b.d_led;
    a.s.segs = b.d_segs;
    const bool face = !c->car || c->p.car_walk_amr;
    const bool mixed = c->p.precision == MC_MIXED;
    const void *kw = face ? (mixed ? (const void *)k_walk<MC_MIXED, kFaceExit, 2>
                                   : (const void *)k_walk<MC_FP64, kFaceExit, 2>)
                          : (mixed ? (const void *)k_walk<MC_MIXED, kCarDDA, 2>
                                   : (const void *)k_walk<MC_FP64, kCarDDA, 2>);
    const int nblk_walk = persistent_blocks(kw, 0);
    const size_t nslot = (size_t)nblk_walk * (kBlock / 64) * 5;
    if (b.n_wave < nslot) {
        if (b.d_wave) (void)hipFree(b.d_wave);
        b.d_wave = nullptr;
        MC_HIP(hipMalloc(&b.d_wave, sizeof(double) * nslot));
        b.n_wave = nslot;
    }
    a.ir_wave = b.d_wave;
    std::vector<double> hw(nslot);
    const int nblk = (nl + kBlock - 1) / kBlock;
    const int ntile = (nl + kScanTile - 1) / kScanTile;
    IrLaunch q;
    q.lam = b.d_lam;
    q.sabs = b.d_sabs;
    q.tdust = b.d_tdust;
    q.khi = b.d_khi;
    q.nl = b.nl;
    q.nzero = b.nzero;
    q.fdisc = b.d_segs + 3;
    float t_launch = 0.0f, t_walk = 0.0f, t_rest = 0.0f;
    double L_prev = 0.0;
    int done = 0, converged = 0, rc = 0;
    b.segs[0] = b.segs[1] = 0;
    for (int iter = 1; iter <= niter; ++iter) {
        done = iter;
        MC_HIP(hipEventRecord(c->ev[4], st));
        k_ir_lum<<<nblk, kBlock, 0, st>>>(c->grid, c->st.heat_dust, c->p.distance2cm, b.d_lum);
        k_scan_tiles<<<ntile, kBlock, 0, st>>>(b.d_lum, 1, nl, b.d_lcum, b.d_bsum);
        k_scan_sums<<<1, 1024, 0, st>>>(b.d_bsum, ntile);
        MC_HIP(hipGetLastError());
        double tail[2];
        MC_HIP(hipMemcpyAsync(&tail[0], b.d_lcum + (nl - 1), sizeof(double), hipMemcpyDeviceToHost, st));
        MC_HIP(hipMemcpyAsync(&tail[1], b.d_bsum + (nl - 1) / kScanTile, sizeof(double),
                              hipMemcpyDeviceToHost, st));
        MC_HIP(hipStreamSynchronize(st));
        const double L_tot = tail[0] + tail[1];
        if (iter == 1) L_prev = L_tot;
        if (!(L_tot > 0.0)) break;
        // one shared stratification offset per pass
        double xi = rng_u((uint32_t)c->p.qmc_seed, 3, (uint64_t)iter << 40, 0);
        xi = std::fmin(std::fmax(xi, 1.0e-12), 1.0 - 1.0e-12);
        a.Lpk = L_tot / (double)N;
        k_ir_counts<<<nblk, kBlock, 0, st>>>(nl, b.d_lcum, b.d_bsum, N, xi, b.d_khi);
        k_imax_tiles<<<ntile, kBlock, 0, st>>>(b.d_khi, nl, b.d_kmax);
        k_imax_sums<<<1, 1024, 0, st>>>(b.d_kmax, ntile);
        k_imax_apply<<<nblk, kBlock, 0, st>>>(b.d_khi, nl, b.d_kmax);
        k_zero<<<64, kBlock, 0, st>>>(b.d_irabs, (size_t)nl);
        k_zero<<<1, kBlock, 0, st>>>(b.d_lesc, (size_t)b.nl);
        k_zero_counters<<<1, 64, 0, st>>>(a.s);
        MC_HIP(hipGetLastError());
        MC_HIP(hipEventRecord(c->ev[5], st));
        MC_HIP(hipEventSynchronize(c->ev[5]));
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, c->ev[4], c->ev[5]);
        t_rest += ms;
        q.pass = iter;
        long double n_emit = 0.0L, e_abs = 0.0L, e_esc = 0.0L;
        for (int64_t b0 = 1; b0 <= N; b0 += c->pk.cap) {
            const int64_t b1 = std::min<int64_t>(N, b0 + c->pk.cap - 1);
            q.b0 = b0;
            q.b1 = b1;
            a.first_id = ((int64_t)iter << 40) + b0;
            a.npk = b1 - b0 + 1;
            MC_HIP(hipEventRecord(c->ev[1], st));
            const int64_t nwave = (b1 - b0 + kIrChunk) / kIrChunk;
            k_ir_launch<<<(unsigned)((nwave + kBlock / 64 - 1) / (kBlock / 64)), kBlock, 0, st>>>(a, q);
            MC_HIP(hipGetLastError());
            k_reset_counter<<<1, 1, 0, st>>>(c->st.counter);
            MC_HIP(hipEventRecord(c->ev[2], st));
            void *kargs[] = {&a};
            MC_HIP(hipLaunchKernel(kw, dim3(nblk_walk), dim3(kBlock), kargs, 0, st));
            MC_HIP(hipGetLastError());
            MC_HIP(hipEventRecord(c->ev[3], st));
            MC_HIP(hipEventSynchronize(c->ev[3]));
            (void)hipEventElapsedTime(&ms, c->ev[1], c->ev[2]);
            t_launch += ms;
            (void)hipEventElapsedTime(&ms, c->ev[2], c->ev[3]);
            t_walk += ms;
            MC_HIP(hipMemcpy(hw.data(), b.d_wave, sizeof(double) * nslot, hipMemcpyDeviceToHost));
            for (size_t w = 0; w < nslot; w += 5) {
                n_emit += (long double)hw[w];
                e_abs += (long double)hw[w + 1] + (long double)hw[w + 2];
                e_esc += (long double)hw[w + 3] + (long double)hw[w + 4];
            }
        }
        MC_HIP(hipEventRecord(c->ev[4], st));
        k_ir_heat<<<nblk, kBlock, 0, st>>>(c->grid, b.d_irabs, c->st.rhokap, b.d_heat_euv,
                                           b.d_heat_ir, c->st.heat_dust, c->p.distance2cm,
                                           c->st.partials);
        k_dust_temp<<<nblk, kBlock, 0, st>>>(nl, c->st.rhokap, c->st.heat_dust, b.d_tdust, b.d_dt,
                                             c->p.distance2cm);
        MC_HIP(hipGetLastError());
        const double led[3] = {(double)(n_emit * (long double)a.Lpk), (double)e_abs, (double)e_esc};
        unsigned long long sg[4];
        MC_HIP(hipMemcpyAsync(sg, b.d_segs, sizeof(sg), hipMemcpyDeviceToHost, st));
        double m[3];
        if (sum_partials(c, nblk, m)) return -100;
        MC_HIP(hipEventRecord(c->ev[5], st));
        MC_HIP(hipEventSynchronize(c->ev[5]));
        (void)hipEventElapsedTime(&ms, c->ev[4], c->ev[5]);
        t_rest += ms;
        for (int k = 0; k < 3; ++k) b.led[k] = led[k];
        b.segs[0] = (int64_t)sg[0];
        b.segs[1] = (int64_t)sg[1];
        double fd;
        std::memcpy(&fd, &sg[3], sizeof(fd));
        b.frac_discarded = fd;
        const double emitted = led[0], extinct = led[1], escaped = led[2];
        const double closure = std::fabs(extinct + escaped - emitted) / std::fmax(emitted, DBL_MIN);
        const double L_abs = m[1], L_new = m[2];
        const double drel = std::fabs(L_new - L_prev) / std::fmax(L_new, DBL_MIN);
        rep[MC_DUSTRT_L_EMIT] = L_tot;
        rep[MC_DUSTRT_EMITTED] = emitted;
        rep[MC_DUSTRT_EXTINCT] = extinct;
        rep[MC_DUSTRT_ESCAPED] = escaped;
        rep[MC_DUSTRT_CLOSURE] = closure;
        rep[MC_DUSTRT_L_ABS_PATH] = L_abs;
        rep[MC_DUSTRT_DREL] = drel;
        if (closure > 5.0e-13) {
            mc_set_error("mc_transport_dust_emission: the transported infrared does not conserve "
                         "energy (closure " + std::to_string(closure) + ")");
            rc = -3;
            break;
        }
        L_prev = L_new;
        if (drel < c->p.dust_tol) {
            converged = 1;
            break;
        }
    }
    rep[MC_DUSTRT_PASSES] = done;
    rep[MC_DUSTRT_CONVERGED] = converged;
    rep[MC_DUSTRT_FDISC] = b.nzero > 0 ? b.frac_discarded : 0.0;
    std::memcpy(b.report, rep, sizeof(double) * MC_DUSTRT_NREPORT);
    b.times[0] = t_launch;
    b.times[1] = t_walk;
    b.times[2] = t_rest;
    return rc;
}

// Sobol/Owen uniforms of the launch (k_qmc_dump) into host memory.
int mc_run_qmc_dump(mc_ctx *c, int stream, int64_t first, int n, int ndim, double *out)
{
    double *d = nullptr;
    const size_t cnt = (size_t)n * (size_t)ndim;
    MC_HIP(hipMalloc(&d, sizeof(double) * cnt));
    k_qmc_dump<<<(unsigned)((cnt + kBlock - 1) / kBlock), kBlock, 0, c->stream>>>(
        c->d_tab, (uint32_t)first, n, ndim, stream - 1, d);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(out, d, sizeof(double) * cnt, hipMemcpyDeviceToHost,
                                            c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(d);
    MC_HIP(e);
    return 0;
}

// Launch records of stellar packets first_id .. first_id + n - 1 (k_launch
// into the context's packet buffers, launch order): rec n x 7 {x, y, z, kx,
// ky, kz, E}, bin_leaf n x 2 {0-based band bin, 0-based leaf or -1}.
int mc_run_launch_dump(mc_ctx *c, int64_t first_id, int n, double *rec, int32_t *bin_leaf)
{
    KArgs a = make_args(c);
    a.first_id = first_id;
    a.npk = n;
    a.stream = 1;
    a.pk.key = nullptr;
    const int nlds = c->car ? 0 : c->n_lds_nodes;
    k_launch<<<(unsigned)((n + kBlock - 1) / kBlock), kBlock,
               (size_t)nlds * (sizeof(double4) + sizeof(int32_t)), c->stream>>>(a, nlds);
    MC_HIP(hipGetLastError());
    std::vector<double> col((size_t)n);
    const double *src[7] = {c->pk.x, c->pk.y, c->pk.z, c->pk.kx, c->pk.ky, c->pk.kz, c->pk.E};
    for (int k = 0; k < 7; ++k) {
        MC_HIP(hipMemcpyAsync(col.data(), src[k], sizeof(double) * n, hipMemcpyDeviceToHost,
                              c->stream));
        MC_HIP(hipStreamSynchronize(c->stream));
        for (int i = 0; i < n; ++i) rec[(size_t)i * 7 + k] = col[(size_t)i];
    }
    std::vector<int32_t> ib((size_t)n);
    const int32_t *isrc[2] = {c->pk.inu, c->pk.leaf};
    for (int k = 0; k < 2; ++k) {
        MC_HIP(hipMemcpyAsync(ib.data(), isrc[k], sizeof(int32_t) * n, hipMemcpyDeviceToHost,
                              c->stream));
        MC_HIP(hipStreamSynchronize(c->stream));
        for (int i = 0; i < n; ++i) bin_leaf[(size_t)i * 2 + k] = ib[(size_t)i];
    }
    return 0;
}

// lines_write's luminosities (mc_lines_dev.h): recombination tables, then
// the metal ions of the registry with an n-level atom
int mc_run_lines(mc_ctx *c, std::vector<double> &L, std::vector<double> &wl,
                 std::vector<int32_t> &code)
{
    hipStream_t st = c->stream;
    std::vector<DevLineTab> tabs;
    static const int kCode[3] = {101, 202, 201};      // H I, He II, He I
    int nline = 0;
    for (int k = 0; k < 3; ++k) {
        if (!c->lt_on[k] || c->lt[k].nlin == 0) continue;
        const mc_line_table &t = c->lt[k];
        DevLineTab d;
        std::memset(&d, 0, sizeof(d));
        d.nT = t.nT; d.nD = t.nD; d.nlin = t.nlin; d.out0 = nline;
        for (int i = 0; i < t.nT; ++i) d.tgrid[i] = t.tgrid[i];
        for (int j = 0; j < t.nD; ++j) d.dgrid[j] = t.dgrid[j];
        for (int l = 0; l < t.nlin; ++l) {
            for (int i = 0; i < MC_LT_MAXG; ++i)
                for (int j = 0; j < MC_LT_MAXG; ++j)
                    d.etab[((size_t)l * MC_LT_MAXG + i) * MC_LT_MAXG + j] = t.etab[l][i][j];
            wl.push_back(t.wl[l]);
            code.push_back(kCode[k]);
        }
        nline += t.nlin;
        tabs.push_back(d);
    }
    std::vector<DevLineIon> ions;
    if (c->ntr > 0) {
        const int rc = mc_lines_ions(c, nline, ions, wl, code);
        if (rc < 0) return -1;
        nline = rc;
    }
    if (nline > kLnMaxLines) {
        mc_set_error("mc_lines_compute: more than 512 lines");
        return -1;
    }
    L.assign((size_t)nline, 0.0);
    if (nline == 0) return 0;
    double *dL = nullptr;
    DevLineTab *dT = nullptr;
    DevLineIon *dI = nullptr;
    MC_HIP(hipMalloc(&dL, sizeof(double) * nline));
    int rc = 0;
    hipError_t e = hipMemsetAsync(dL, 0, sizeof(double) * nline, st);
    if (e == hipSuccess && !tabs.empty()) e = hipMalloc(&dT, sizeof(DevLineTab) * tabs.size());
    if (e == hipSuccess && !tabs.empty())
        e = hipMemcpy(dT, tabs.data(), sizeof(DevLineTab) * tabs.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess && !ions.empty()) e = hipMalloc(&dI, sizeof(DevLineIon) * ions.size());
    if (e == hipSuccess && !ions.empty())
        e = hipMemcpy(dI, ions.data(), sizeof(DevLineIon) * ions.size(), hipMemcpyHostToDevice);
    const unsigned nb = (unsigned)((c->nleaf + kBlock - 1) / kBlock);
    if (e == hipSuccess) {
        int it = 0;
        for (int k = 0; k < 3; ++k) {
            if (!c->lt_on[k] || c->lt[k].nlin == 0) continue;
            k_lines_rec<<<nb, kBlock, 0, st>>>(c->grid, c->st, dT + it, k, c->p.He_abund,
                                               c->p.distance2cm, dL);
            ++it;
        }
        e = hipGetLastError();
    }
    if (e == hipSuccess && !ions.empty()) {
        rc = metal_sync(c);
        int nlmax = 0;
        for (const DevLineIon &d : ions) nlmax = std::max(nlmax, d.nlev);
        if (nlmax <= 6)
            k_lines_metal<6><<<nb, kBlock, 0, st>>>(c->grid, c->st, make_sw(c), dI,
                                                    (int)ions.size(), c->p.distance2cm, dL);
        else
            k_lines_metal<kLnLev><<<nb, kBlock, 0, st>>>(c->grid, c->st, make_sw(c), dI,
                                                         (int)ions.size(), c->p.distance2cm, dL);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(L.data(), dL, sizeof(double) * nline, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(dL);
    if (dT) (void)hipFree(dT);
    if (dI) (void)hipFree(dI);
    MC_HIP(e);
    return rc ? -100 : 0;
}
