// C-ABI of the transport / cell-balance library (declared in
// include/mochii_gpu.h).  Owns the device allocations of a context.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include "mc_internal.h"

static thread_local std::string g_err;

void mc_set_error(const std::string &msg) { g_err = msg; }

extern "C" const char *mc_last_error(void) { return g_err.c_str(); }

#define CHECK_CTX(c)                                 \
    do {                                             \
        if (!(c)) {                                  \
            mc_set_error("null context");            \
            return -1;                               \
        }                                            \
    } while (0)

template <class T>
static int dalloc(T **p, size_t n)
{
    *p = nullptr;
    if (n == 0) n = 1;
    MC_HIP(hipMalloc((void **)p, sizeof(T) * n));
    return 0;
}

static int upload_tables(mc_ctx *c)
{
    DevTables t;
    std::memset(&t, 0, sizeof(t));
    mc_qmc_tables(c->p.qmc_seed, t.V, t.key);
    for (int i = 0; i <= c->nnu; ++i) t.eedge[i] = c->eedge[i];
    for (int i = 0; i < c->ncomp; ++i) {
        t.src_cdf[i] = c->src_cdf[i];
        t.comp_kind[i] = c->comp_kind[i];
    }
    for (int i = 0; i < c->p.nsource; ++i) {
        t.src_pos[i][0] = c->p.src_x[i];
        t.src_pos[i][1] = c->p.src_y[i];
        t.src_pos[i][2] = c->p.src_z[i];
    }
    for (size_t i = 0; i < c->samp_off.size(); ++i) t.samp_off[i] = c->samp_off[i];
    t.nnu = c->nnu;
    t.ncomp = c->ncomp;
    t.multi = c->multi;
    t.nsource = c->p.nsource;
    t.qmc_ndim = c->qmc_ndim;
    t.ext_geometry = c->p.ext_geometry;
    t.Lpacket = c->Ltot / (double)c->p.nphotons;
    t.rmax = c->p.rmax;
    MC_HIP(hipMemcpy(c->d_tab, &t, sizeof(t), hipMemcpyHostToDevice));
    const size_t nk = c->samp_e.size();
    std::vector<double> s(3 * nk);
    std::copy(c->samp_e.begin(), c->samp_e.end(), s.begin());
    std::copy(c->samp_f.begin(), c->samp_f.end(), s.begin() + nk);
    std::copy(c->samp_c.begin(), c->samp_c.end(), s.begin() + 2 * nk);
    if (c->d_samp) (void)hipFree(c->d_samp);
    if (dalloc(&c->d_samp, 3 * nk)) return -100;
    MC_HIP(hipMemcpy(c->d_samp, s.data(), sizeof(double) * 3 * nk, hipMemcpyHostToDevice));
    return 0;
}

// Hot leaves: leaves whose centre lies within kHotRadius leaf widths of a
// point source.  Every packet of that source crosses them, so their tally
// updates are spread over kHotReplicas copies during the walk.
static const double kHotRadius = 6.0;
static const int kHotReplicas = 16;
static const int kMaxHot = 1 << 16;

static int setup_hot(mc_ctx *c)
{
    DevState &s = c->st;
    std::vector<int32_t> slot(c->nleaf, -1), leaves;
    for (int il = 0; il < c->nleaf && (int)leaves.size() < kMaxHot; ++il) {
        double x, y, z, h;
        if (c->car) {
            const int nxy = c->p.nx * c->p.ny, iz = il / nxy, rem = il - iz * nxy;
            const int iy = rem / c->p.nx, ix = rem - iy * c->p.nx;
            x = c->xmin + (ix + 0.5) * c->dcell;
            y = c->ymin + (iy + 0.5) * c->dcell;
            z = c->zmin + (iz + 0.5) * c->dcell;
            h = 0.5 * c->dcell;
        } else {
            const int k = c->hcell_of_leaf[il];
            x = c->hcx[k]; y = c->hcy[k]; z = c->hcz[k]; h = c->hch[k];
        }
        for (int is = 0; is < c->p.nsource; ++is) {
            const double dx = x - c->p.src_x[is], dy = y - c->p.src_y[is], dz = z - c->p.src_z[is];
            const double r = kHotRadius * 2.0 * h;
            if (dx * dx + dy * dy + dz * dz < r * r) {
                slot[il] = (int32_t)leaves.size();
                leaves.push_back(il);
                break;
            }
        }
    }
    s.nhot = (int32_t)leaves.size();
    s.nrep = kHotReplicas;
    if (s.nhot == 0) return 0;
    if (dalloc(&s.hot_slot, (size_t)c->nleaf) || dalloc(&s.hot_leaf, (size_t)s.nhot) ||
        dalloc(&s.hot_hhe, (size_t)s.nrep * s.nhot * 8) ||
        dalloc(&s.hot_jt, (size_t)s.nrep * s.nhot * c->nnu))
        return -100;
    MC_HIP(hipMemcpy(s.hot_slot, slot.data(), sizeof(int32_t) * c->nleaf, hipMemcpyHostToDevice));
    MC_HIP(hipMemcpy(s.hot_leaf, leaves.data(), sizeof(int32_t) * s.nhot, hipMemcpyHostToDevice));
    return 0;
}

// Milne ground-continuum tables for the diffuse field (milne_setup, built on
// the host after the band is fixed) -> device
static int upload_milne(mc_ctx *c)
{
    const size_t NT = MC_MILNE_NT, NX = MC_MILNE_NX;
    const size_t ncdf = 3 * NT * NX, nrow = 3 * NT;
    std::vector<double> h(ncdf + 2 * nrow + NX);
    std::vector<double> lgT(NT);
    mc_milne_tables(c->p.eion_max, h.data(), h.data() + ncdf, h.data() + ncdf + nrow,
                    h.data() + ncdf + 2 * nrow, lgT.data());
    if (dalloc(&c->d_milne, h.size())) return -100;
    MC_HIP(hipMemcpy(c->d_milne, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
    DevMilne &M = c->milne;
    M.cdf = c->d_milne;
    M.xbar = c->d_milne + ncdf;
    M.fin = c->d_milne + ncdf + nrow;
    M.xg = c->d_milne + ncdf + 2 * nrow;
    M.lgT0 = lgT[0];
    M.lgT1 = lgT[1];
    M.eth[0] = MC_ETH_HI; M.eth[1] = MC_ETH_HEI; M.eth[2] = MC_ETH_HEII;
    M.eion_max = c->p.eion_max;
    M.model = c->p.diffuse_model;
    return 0;
}

static void free_metal_arrays(mc_ctx *c)
{
    for (void *p : {(void *)c->d_mraw, (void *)c->d_mgam, (void *)c->d_mheat, (void *)c->d_mfrac,
                    (void *)c->d_mop, (void *)c->d_mopf})
        if (p) (void)hipFree(p);
    c->d_mraw = c->d_mgam = c->d_mheat = c->d_mfrac = c->d_mop = nullptr;
    c->d_mopf = nullptr;
}

// per-leaf metal arrays (species_setup / species_resize,
// src/species_mod.f90:143-245), zeroed
static int alloc_metal_arrays(mc_ctx *c)
{
    free_metal_arrays(c);
    if (c->ntr <= 0 || c->nleaf <= 0) return 0;
    const size_t n = (size_t)c->nleaf, nt = (size_t)c->ntr, ns = (size_t)c->nst;
    if (dalloc(&c->d_mraw, n * nt * 2) || dalloc(&c->d_mgam, n * nt) || dalloc(&c->d_mheat, n * nt) ||
        dalloc(&c->d_mfrac, n * ns) || dalloc(&c->d_mop, n * nt) || dalloc(&c->d_mopf, n * nt))
        return -100;
    MC_HIP(hipMemset(c->d_mraw, 0, sizeof(double) * n * nt * 2));
    MC_HIP(hipMemset(c->d_mgam, 0, sizeof(double) * n * nt));
    MC_HIP(hipMemset(c->d_mheat, 0, sizeof(double) * n * nt));
    MC_HIP(hipMemset(c->d_mfrac, 0, sizeof(double) * n * ns));
    MC_HIP(hipMemset(c->d_mop, 0, sizeof(double) * n * nt));
    MC_HIP(hipMemset(c->d_mopf, 0, sizeof(float) * n * nt));
    return 0;
}

static void free_leaf_arrays(mc_ctx *c)
{
    free_metal_arrays(c);
    mc_ir_free(c);
    mc_mom_free(c);
    DevState &s = c->st;
    for (void *p : {(void *)s.hot_slot, (void *)s.hot_leaf, (void *)s.hot_hhe, (void *)s.hot_jt})
        if (p) (void)hipFree(p);
    s.hot_slot = s.hot_leaf = nullptr;
    s.hot_hhe = s.hot_jt = nullptr;
    s.nhot = 0;
    for (void *p : {(void *)s.nH, (void *)s.xHI, (void *)s.xHeI, (void *)s.xHeII, (void *)s.ne,
                    (void *)s.Te, (void *)s.rhokap, (void *)s.opac, (void *)s.hhe, (void *)s.jt,
                    (void *)s.rates, (void *)s.partials, (void *)s.heat_dust})
        if (p) (void)hipFree(p);
    s.nH = s.xHI = s.xHeI = s.xHeII = s.ne = s.Te = s.rhokap = nullptr;
    s.opac = nullptr;
    s.hhe = s.jt = s.rates = s.partials = s.heat_dust = nullptr;
    for (void *p : {(void *)c->d_dif, (void *)c->d_dcdf, (void *)c->d_bsum})
        if (p) (void)hipFree(p);
    c->d_dif = c->d_dcdf = c->d_bsum = nullptr;
    for (void *p : {(void *)c->d_feat, (void *)c->d_pred})
        if (p) (void)hipFree(p);
    c->d_feat = c->d_pred = nullptr;
}

static int alloc_leaf_arrays(mc_ctx *c)
{
    free_leaf_arrays(c);
    const size_t n = (size_t)c->nleaf;
    DevState &s = c->st;
    if (dalloc(&s.nH, n) || dalloc(&s.xHI, n) || dalloc(&s.xHeI, n) || dalloc(&s.xHeII, n) ||
        dalloc(&s.ne, n) || dalloc(&s.Te, n) || dalloc(&s.rhokap, n) || dalloc(&s.opac, n) ||
        dalloc(&s.hhe, n * 8) || dalloc(&s.jt, n * (size_t)c->nnu) || dalloc(&s.rates, n * 6) ||
        dalloc(&s.partials, ((n + 255) / 256) * 10 + 10))
        return -100;
    if (c->p.diffuse_field) {
        const size_t nb = (n + 1023) / 1024;        // scan tiles of 1024 leaves
        if (dalloc(&c->d_dif, 5 * n) || dalloc(&c->d_dcdf, n) || dalloc(&c->d_bsum, nb + 1))
            return -100;
    }
    if (c->p.ion_add_dust) {
        if (dalloc(&s.heat_dust, n)) return -100;
        MC_HIP(hipMemset(s.heat_dust, 0, sizeof(double) * n));
    }
    MC_HIP(hipMemset(s.rhokap, 0, sizeof(double) * n));
    MC_HIP(hipMemset(s.hhe, 0, sizeof(double) * n * 8));
    MC_HIP(hipMemset(s.jt, 0, sizeof(double) * n * c->nnu));
    MC_HIP(hipMemset(s.rates, 0, sizeof(double) * n * 6));
    c->grid.nleaf = c->nleaf;
    c->state_ready = false;
    if (alloc_metal_arrays(c)) return -100;
    return setup_hot(c);
}

// launch-record buffers and the direction-sort workspace for batches of
// min(n, batch_max) packets (124 B of records each; the default depends on
// the precision form, mc_set_launch_batch); grown (never shrunk) when a
// band needs larger batches.  Large batches keep the direction-sorted waves
// coherent when many emitters share a pass: a batch holds every emitter's
// share of its packet ids, so the packets of one emitter and direction cell
// per batch grow with the batch.  Halved down to 1M when the device cannot
// hold them.
static void free_packet_buffers(mc_ctx *c)
{
    PacketBuf &pk = c->pk;
    for (void *q : {(void *)pk.x, (void *)pk.y, (void *)pk.z, (void *)pk.kx, (void *)pk.ky,
                    (void *)pk.kz, (void *)pk.E, (void *)pk.sig, (void *)pk.inu, (void *)pk.leaf,
                    (void *)pk.key, (void *)c->d_key_sorted, (void *)c->d_perm, c->d_sort_tmp})
        if (q) (void)hipFree(q);
    pk.x = pk.y = pk.z = pk.kx = pk.ky = pk.kz = pk.E = pk.sig = nullptr;
    pk.inu = pk.leaf = nullptr;
    pk.key = c->d_key_sorted = nullptr;
    c->d_perm = nullptr;
    c->d_sort_tmp = nullptr;
    pk.cap = 0;
}

static int alloc_packet_buffers(mc_ctx *c, int64_t cap)
{
    PacketBuf &pk = c->pk;
    pk.cap = cap;
    if (dalloc(&pk.x, cap) || dalloc(&pk.y, cap) || dalloc(&pk.z, cap) || dalloc(&pk.kx, cap) ||
        dalloc(&pk.ky, cap) || dalloc(&pk.kz, cap) || dalloc(&pk.E, cap) ||
        dalloc(&pk.sig, 6 * cap) || dalloc(&pk.inu, cap) || dalloc(&pk.leaf, cap) ||
        dalloc(&pk.key, cap) || dalloc(&c->d_key_sorted, cap) || dalloc(&c->d_perm, cap))
        return -100;
    c->sort_tmp_bytes = mc_sort_temp_bytes(cap);
    if (hipMalloc(&c->d_sort_tmp, std::max<size_t>(c->sort_tmp_bytes, 16)) != hipSuccess) {
        c->d_sort_tmp = nullptr;
        mc_set_error("sort workspace allocation failed");
        return -100;
    }
    return 0;
}

int mc_packet_capacity(mc_ctx *c, int64_t n)
{
    int64_t cap = std::min<int64_t>(std::max<int64_t>(n, 1), c->batch_max);
    if (c->pk.x && c->pk.cap >= cap) return 0;
    for (;;) {
        free_packet_buffers(c);
        if (alloc_packet_buffers(c, cap) == 0) return 0;
        (void)hipGetLastError();
        if (cap <= ((int64_t)1 << 20)) {
            free_packet_buffers(c);
            return -100;
        }
        cap = std::max<int64_t>(cap / 2, (int64_t)1 << 20);
    }
}

extern "C" int mc_init(const mc_params *p, mc_ctx **out)
{
    if (!p || !out) {
        mc_set_error("mc_init: null argument");
        return -1;
    }
    *out = nullptr;
    if (p->nphotons < 1) {
        mc_set_error("nphotons must be >= 1");
        return -1;
    }
    if (p->precision != MC_FP64 && p->precision != MC_MIXED) {
        mc_set_error("precision must be MC_FP64 or MC_MIXED");
        return -1;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        mc_set_error("no HIP device available");
        return -2;
    }
    if (p->diffuse_field && (p->diffuse_model < 0 || p->diffuse_model > 2)) {
        mc_set_error("diffuse_model must be 0 (threshold), 1 (exponential) or 2 (milne)");
        return -1;
    }
    if (p->cooling_model != MC_COOL_LOW_DENSITY && p->cooling_model != MC_COOL_LOCAL_NE) {
        mc_set_error("cooling_model: MC_COOL_LOW_DENSITY or MC_COOL_LOCAL_NE");
        return -1;
    }
    if (p->ci_model != MC_CI_VORONOV && p->ci_model != MC_CI_DERE_HYBRID) {
        mc_set_error("ci_model must be MC_CI_VORONOV or MC_CI_DERE_HYBRID");
        return -1;
    }
    if (p->charge_exchange != MC_CX_TWO_WAY && p->charge_exchange != MC_CX_METAL_ONLY) {
        mc_set_error("charge_exchange must be MC_CX_TWO_WAY or MC_CX_METAL_ONLY");
        return -1;
    }
    mc_ctx *c = new mc_ctx();
    c->p = *p;
    // the explicit diffuse field replaces on-the-spot absorption: case A
    // (src/setup.f90:580-583)
    if (c->p.diffuse_field) c->p.case_A = 1;
    if (mc_setup_band(c) || mc_setup_samplers(c)) {
        delete c;
        return -1;
    }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        mc_set_error("hipStreamCreate failed");
        delete c;
        return -100;
    }
    for (auto &e : c->ev) (void)hipEventCreate(&e);
    if (dalloc(&c->d_tab, 1) || upload_tables(c)) {
        mc_finalize(c);
        return -100;
    }
    if (dalloc(&c->st.ledger, 3) || dalloc(&c->st.segs, 4) || dalloc(&c->st.counter, 1)) {
        mc_finalize(c);
        return -100;
    }
    if (c->p.diffuse_field && upload_milne(c)) {
        mc_finalize(c);
        return -100;
    }
    // launch-record buffers: batches of up to MC_LAUNCH_BATCH_MIXED / _FP64 packets
    c->packet_order = MC_ORDER_DIRECTION;
    c->batch_max = p->precision == MC_MIXED ? MC_LAUNCH_BATCH_MIXED : MC_LAUNCH_BATCH_FP64;
    c->mom_on = 1;
    c->row_issue = MC_ROWS_AUTO;
    c->rows_together = false;
    c->rows_together_last = -1;
    c->warm_levels = MC_WARM_LEVELS_DEFAULT;
    if (mc_packet_capacity(c, p->nphotons)) {
        mc_finalize(c);
        return -100;
    }
    std::memset(&c->grid, 0, sizeof(c->grid));
    if (p->grid_type == MC_GRID_CAR) {
        const double dx = 2.0 * p->xmax / p->nx, dy = 2.0 * p->ymax / p->ny,
                     dz = 2.0 * p->zmax / p->nz;
        if (p->nx < 1 || p->ny < 1 || p->nz < 1 || std::fabs(dy - dx) > 1e-9 * dx ||
            std::fabs(dz - dx) > 1e-9 * dx) {
            mc_set_error("car grid needs cubic cells: 2*xmax/nx = 2*ymax/ny = 2*zmax/nz");
            mc_finalize(c);
            return -1;
        }
        c->car = 1;
        c->xmin = -p->xmax; c->xmaxb = p->xmax;
        c->ymin = -p->ymax; c->ymaxb = p->ymax;
        c->zmin = -p->zmax; c->zmaxb = p->zmax;
        c->dcell = (c->xmaxb - c->xmin) / p->nx;
        c->nleaf = c->ncells = p->nx * p->ny * p->nz;
        DevGrid &g = c->grid;
        g.car = 1;
        g.car_walk_amr = p->car_walk_amr;
        g.nx = p->nx; g.ny = p->ny; g.nz = p->nz;
        g.xmin = c->xmin; g.ymin = c->ymin; g.zmin = c->zmin;
        g.xmax = c->xmaxb; g.ymax = c->ymaxb; g.zmax = c->zmaxb;
        g.dcell = c->dcell;
        g.half = p->xmax;
        if (alloc_leaf_arrays(c)) {
            mc_finalize(c);
            return -100;
        }
    } else if (p->grid_type != MC_GRID_AMR) {
        mc_set_error("unknown grid_type");
        mc_finalize(c);
        return -1;
    }
    *out = c;
    return 0;
}

extern "C" int mc_finalize(mc_ctx *c)
{
    if (!c) return 0;
    free_leaf_arrays(c);
    for (void *p : {(void *)c->d_tab, (void *)c->d_samp, (void *)c->d_dust, (void *)c->st.ledger,
                    (void *)c->st.segs, (void *)c->st.counter, (void *)c->pk.x, (void *)c->pk.y,
                    (void *)c->pk.z, (void *)c->pk.kx, (void *)c->pk.ky, (void *)c->pk.kz,
                    (void *)c->pk.E, (void *)c->pk.sig, (void *)c->pk.inu, (void *)c->pk.leaf,
                    (void *)c->d_node_gd, (void *)c->d_node_gf, (void *)c->d_node_link,
                    (void *)c->d_node_nbr, (void *)c->d_leaf_node, (void *)c->d_milne,
                    (void *)c->d_hei, (void *)c->mlp.P, (void *)c->d_feat, (void *)c->d_pred, (void *)c->pk.key,
                    (void *)c->d_key_sorted, (void *)c->d_perm, c->d_sort_tmp, (void *)c->d_nlt})
        if (p) (void)hipFree(p);
    for (auto &e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return 0;
}

extern "C" int mc_grid_octree(mc_ctx *c, int nleaf, const double *xl, const double *yl,
                              const double *zl, const int32_t *lev, double half)
{
    CHECK_CTX(c);
    if (c->p.grid_type != MC_GRID_AMR) {
        mc_set_error("mc_grid_octree: context was created for a car grid");
        return -1;
    }
    if (nleaf < 1 || !xl || !yl || !zl || !lev || !(half > 0.0)) {
        mc_set_error("mc_grid_octree: bad arguments");
        return -1;
    }
    if (mc_build_octree_host(c, nleaf, xl, yl, zl, lev, half)) return -1;
    // breadth-first renumbering; every internal node must have 8 children
    const int nc = c->ncells;
    std::vector<int> bfs;
    bfs.reserve(nc);
    bfs.push_back(0);
    for (size_t q = 0; q < bfs.size(); ++q) {
        const int k = bfs[q];
        if (c->hileaf[k] >= 0) continue;
        for (int o = 0; o < 8; ++o) {
            const int ch = c->hchildren[(size_t)k * 8 + o];
            if (ch < 0) {
                mc_set_error("octree is not space-filling (an internal cell lacks a child)");
                return -1;
            }
            bfs.push_back(ch);
        }
    }
    if ((int)bfs.size() != nc) {
        mc_set_error("octree has unreachable cells");
        return -1;
    }
    std::vector<int> nid(nc);
    for (int i = 0; i < nc; ++i) nid[bfs[i]] = i;
    std::vector<double4> gd(nc);
    std::vector<float4> gf(nc);
    std::vector<int32_t> link(nc), nbr((size_t)nc * 6), lnode(nleaf);
    for (int i = 0; i < nc; ++i) {
        const int k = bfs[i];
        gd[i] = make_double4(c->hcx[k], c->hcy[k], c->hcz[k], c->hch[k]);
        gf[i] = make_float4((float)(c->hcx[k] / half), (float)(c->hcy[k] / half),
                            (float)(c->hcz[k] / half), (float)(c->hch[k] / half));
        link[i] = c->hileaf[k] >= 0 ? -(c->hileaf[k] + 1) : nid[c->hchildren[(size_t)k * 8]];
        for (int f = 0; f < 6; ++f) {
            const int nb = c->hneighbor[(size_t)k * 6 + f];
            nbr[(size_t)i * 6 + f] = nb >= 0 ? nid[nb] : -1;
        }
    }
    for (int l = 0; l < nleaf; ++l) lnode[l] = nid[c->hcell_of_leaf[l]];
    // prefix of whole levels that fits the launch kernel's LDS budget
    int nl = 0;
    for (int lv = 0, cnt = 1; nl + cnt <= 2048; ++lv, cnt *= 8) {
        int m = 0;
        for (int i = 0; i < nc; ++i) m += c->hlevel[bfs[i]] == lv;
        if (m == 0) break;
        nl += m;
        if (m < cnt) break;
    }
    c->n_lds_nodes = std::min(nl, nc);
    for (void *p : {(void *)c->d_node_gd, (void *)c->d_node_gf, (void *)c->d_node_link,
                    (void *)c->d_node_nbr, (void *)c->d_leaf_node})
        if (p) (void)hipFree(p);
    if (dalloc(&c->d_node_gd, nc) || dalloc(&c->d_node_gf, nc) || dalloc(&c->d_node_link, nc) ||
        dalloc(&c->d_node_nbr, (size_t)nc * 6) || dalloc(&c->d_leaf_node, (size_t)nleaf))
        return -100;
    MC_HIP(hipMemcpy(c->d_node_gd, gd.data(), sizeof(double4) * nc, hipMemcpyHostToDevice));
    MC_HIP(hipMemcpy(c->d_node_gf, gf.data(), sizeof(float4) * nc, hipMemcpyHostToDevice));
    MC_HIP(hipMemcpy(c->d_node_link, link.data(), sizeof(int32_t) * nc, hipMemcpyHostToDevice));
    MC_HIP(hipMemcpy(c->d_node_nbr, nbr.data(), sizeof(int32_t) * nc * 6, hipMemcpyHostToDevice));
    MC_HIP(hipMemcpy(c->d_leaf_node, lnode.data(), sizeof(int32_t) * nleaf, hipMemcpyHostToDevice));
    DevGrid &g = c->grid;
    g.car = 0;
    g.car_walk_amr = 0;
    g.nleaf = nleaf;
    g.nnode = nc;
    g.xmin = c->xmin; g.ymin = c->ymin; g.zmin = c->zmin;
    g.xmax = c->xmaxb; g.ymax = c->ymaxb; g.zmax = c->zmaxb;
    g.dcell = 0.0;
    g.half = half;
    g.node_gd = c->d_node_gd;
    g.node_gf = c->d_node_gf;
    g.node_link = c->d_node_link;
    g.node_nbr = c->d_node_nbr;
    g.leaf_node = c->d_leaf_node;
    return alloc_leaf_arrays(c);
}

static double *field_ptr(mc_ctx *c, int field)
{
    switch (field) {
    case MC_NH: return c->st.nH;
    case MC_XHI: return c->st.xHI;
    case MC_XHEI: return c->st.xHeI;
    case MC_XHEII: return c->st.xHeII;
    case MC_NE: return c->st.ne;
    case MC_TE: return c->st.Te;
    case MC_RHOKAP: return c->st.rhokap;
    default: return nullptr;
    }
}

extern "C" int mc_gas_state_setup(mc_ctx *c, const double *nH, const double *xHI)
{
    CHECK_CTX(c);
    if (!c->st.nH) {
        mc_set_error("mc_gas_state_setup: no grid yet");
        return -1;
    }
    if (!nH) {
        mc_set_error("mc_gas_state_setup: nH is required");
        return -1;
    }
    const int n = c->nleaf;
    const mc_params &p = c->p;
    std::vector<double> xh(n), x1(n, p.xHeI_init), x2(n, p.xHeII_init), ne(n), te(n, p.te_fixed);
    for (int i = 0; i < n; ++i) {
        xh[i] = xHI ? xHI[i] : p.xHI_init;
        const double x3 = std::max(0.0, 1.0 - x1[i] - x2[i]);
        ne[i] = nH[i] * ((1.0 - xh[i]) + p.He_abund * (x2[i] + 2.0 * x3));
    }
    MC_HIP(hipMemcpy(c->st.nH, nH, sizeof(double) * n, hipMemcpyHostToDevice));
    MC_HIP(hipMemcpy(c->st.xHI, xh.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    MC_HIP(hipMemcpy(c->st.xHeI, x1.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    MC_HIP(hipMemcpy(c->st.xHeII, x2.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    MC_HIP(hipMemcpy(c->st.ne, ne.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    MC_HIP(hipMemcpy(c->st.Te, te.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    c->state_ready = true;
    return mc_run_opacity_fill(c);
}

extern "C" int mc_state_set(mc_ctx *c, int field, const double *src)
{
    CHECK_CTX(c);
    double *d = field_ptr(c, field);
    if (!d || !src) {
        mc_set_error("mc_state_set: bad field or null source");
        return -1;
    }
    MC_HIP(hipMemcpy(d, src, sizeof(double) * c->nleaf, hipMemcpyHostToDevice));
    return mc_run_opacity_fill(c);
}

extern "C" int mc_state_get(mc_ctx *c, int field, double *dst)
{
    CHECK_CTX(c);
    double *d = field_ptr(c, field);
    if (!d || !dst) {
        mc_set_error("mc_state_get: bad field or null destination");
        return -1;
    }
    MC_HIP(hipMemcpy(dst, d, sizeof(double) * c->nleaf, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int mc_ion_dust_table(mc_ctx *c, int n, const double *lam, const double *alb,
                                 const double *g, const double *cext)
{
    CHECK_CTX(c);
    if (n < 2 || !lam || !alb || !g || !cext) {
        mc_set_error("mc_ion_dust_table: need n >= 2 rows");
        return -1;
    }
    // stable sort by wavelength (ion_dust_read's insertion sort keeps equal
    // wavelengths in input order)
    std::vector<int> ord(n);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return lam[a] < lam[b]; });
    c->dlam.resize(n); c->dalb.resize(n); c->dg.resize(n); c->dcext.resize(n);
    for (int i = 0; i < n; ++i) {
        c->dlam[i] = lam[ord[i]]; c->dalb[i] = alb[ord[i]];
        c->dg[i] = g[ord[i]]; c->dcext[i] = cext[ord[i]];
    }
    // C_ext at lambda_ref, log-log interpolation clamped at the table ends
    const double x0 = c->p.lambda_ref;
    if (x0 <= c->dlam[0]) c->dcref = c->dcext[0];
    else if (x0 >= c->dlam[n - 1]) c->dcref = c->dcext[n - 1];
    else {
        int i = 0;
        while (i < n - 1 && !(x0 < c->dlam[i + 1])) ++i;
        const double w = std::log(x0 / c->dlam[i]) / std::log(c->dlam[i + 1] / c->dlam[i]);
        c->dcref = std::exp(std::log(std::max(c->dcext[i], 2.2250738585072014e-308)) * (1.0 - w) +
                            std::log(std::max(c->dcext[i + 1], 2.2250738585072014e-308)) * w);
    }
    std::vector<double> h(4 * (size_t)n);
    std::copy(c->dlam.begin(), c->dlam.end(), h.begin());
    std::copy(c->dalb.begin(), c->dalb.end(), h.begin() + n);
    std::copy(c->dg.begin(), c->dg.end(), h.begin() + 2 * n);
    std::copy(c->dcext.begin(), c->dcext.end(), h.begin() + 3 * n);
    if (c->d_dust) (void)hipFree(c->d_dust);
    if (dalloc(&c->d_dust, 4 * (size_t)n)) return -100;
    MC_HIP(hipMemcpy(c->d_dust, h.data(), sizeof(double) * 4 * n, hipMemcpyHostToDevice));
    c->dn = n;
    return 0;
}

extern "C" int mc_gas_opacity_fill(mc_ctx *c)
{
    CHECK_CTX(c);
    return mc_run_opacity_fill(c);
}

extern "C" int mc_ionizing_field_and_rates(mc_ctx *c, int with_peel)
{
    CHECK_CTX(c);
    if (with_peel) {
        mc_set_error("peel-off imaging is not part of this library");
        return -1;
    }
    if (!c->state_ready) {
        mc_set_error("mc_ionizing_field_and_rates: gas state not set up");
        return -1;
    }
    if (c->p.ion_add_dust && c->dn < 2) {
        mc_set_error("ion_add_dust needs mc_ion_dust_table first");
        return -1;
    }
    int rc = mc_run_pass(c);
    if (rc) return rc;
    rc = mc_run_rates(c);
    if (rc) return rc;
    const double closure = std::fabs(c->ledger[1] + c->ledger[2] - c->ledger[0]) /
                           std::max(c->ledger[0], 2.2250738585072014e-308);
    if (closure > 5.0e-13) {
        mc_set_error("energy closure " + std::to_string(closure) + " exceeds 5e-13");
        return -4;
    }
    return 0;
}

extern "C" int mc_gas_equilibrium_update(mc_ctx *c, double *max_dx, double *dx_vol)
{
    CHECK_CTX(c);
    if (!max_dx || !dx_vol) {
        mc_set_error("mc_gas_equilibrium_update: null output");
        return -1;
    }
    const int rc = mc_run_equilibrium(c, max_dx, dx_vol);
    if (rc) return rc;
    return mc_run_stage_refresh(c);       // gas_opacity_fill after the update
}

extern "C" int mc_cooling_setup(mc_ctx *c, int nfit, const double *A, const double *Ti)
{
    CHECK_CTX(c);
    if (nfit < 0 || nfit > 8 || (nfit > 0 && (!A || !Ti))) {
        mc_set_error("mc_cooling_setup: need 0..8 fit terms");
        return -1;
    }
    double logT[41], gff[4][41];
    mc_gff_table(logT, gff);
    std::memcpy(c->cool.gff, gff, sizeof(gff));
    c->cool.logT0 = logT[0];
    c->cool.dlogT = logT[1] - logT[0];
    c->cool.nfit = nfit;
    for (int i = 0; i < nfit; ++i) {
        c->cool.A[i] = A[i];
        c->cool.Ti[i] = Ti[i];
    }
    c->cool_ready = true;
    return 0;
}

extern "C" int mc_gas_thermal_update(mc_ctx *c, double *max_dx, double *max_dte, double *dx_vol,
                                     double *dte_vol)
{
    CHECK_CTX(c);
    if (!max_dx || !max_dte || !dx_vol || !dte_vol) {
        mc_set_error("mc_gas_thermal_update: null output");
        return -1;
    }
    if (!c->cool_ready) {
        mc_set_error("mc_gas_thermal_update: call mc_cooling_setup first");
        return -1;
    }
    double m[4];
    int rc = mc_run_thermal(c, m);
    if (!rc) rc = mc_run_stage_refresh(c); // gas_opacity_fill after the update
    if (rc) return rc;
    *max_dx = m[0];
    *max_dte = m[1];
    *dx_vol = m[2];
    *dte_vol = m[3];
    return 0;
}

extern "C" int mc_set_packet_order(mc_ctx *c, int order)
{
    CHECK_CTX(c);
    if (order != MC_ORDER_LAUNCH && order != MC_ORDER_DIRECTION) {
        mc_set_error("mc_set_packet_order: order must be MC_ORDER_LAUNCH or MC_ORDER_DIRECTION");
        return -1;
    }
    c->packet_order = order;
    return 0;
}

extern "C" int mc_set_launch_batch(mc_ctx *c, int64_t max_records)
{
    CHECK_CTX(c);
    if (max_records < 1024) {
        mc_set_error("mc_set_launch_batch: max_records must be >= 1024");
        return -1;
    }
    c->batch_max = max_records;
    free_packet_buffers(c);
    return mc_packet_capacity(c, c->p.nphotons);
}

extern "C" int mc_set_row_issue(mc_ctx *c, int mode)
{
    CHECK_CTX(c);
    if (mode != MC_ROWS_AUTO && mode != MC_ROWS_LANE && mode != MC_ROWS_TOGETHER) {
        mc_set_error("mc_set_row_issue: mode must be MC_ROWS_AUTO, MC_ROWS_LANE or MC_ROWS_TOGETHER");
        return -1;
    }
    c->row_issue = mode;
    if (mode != MC_ROWS_AUTO) c->rows_together = mode == MC_ROWS_TOGETHER;
    return 0;
}

extern "C" int mc_row_issue(mc_ctx *c, int *mode, int *last)
{
    CHECK_CTX(c);
    if (mode) *mode = c->row_issue;
    if (last) *last = c->rows_together_last;
    return 0;
}

extern "C" int mc_set_tally_moments(mc_ctx *c, int on)
{
    CHECK_CTX(c);
    c->mom_on = on != 0;
    return 0;
}

extern "C" int mc_tally_moment_info(mc_ctx *c, int32_t *nsub, double *fit_err)
{
    CHECK_CTX(c);
    const int ndust = (c->p.ion_add_dust && c->dn >= 2) ? 1 : 0;
    if (c->mom.built_ntr != c->ntr || c->mom.built_dust != ndust) {
        mc_mom_free(c);
        if (mc_mom_build(c)) return -1;
    }
    if (nsub) *nsub = c->mom.nsub;
    if (fit_err) *fit_err = c->mom.fit_err;
    return 0;
}

extern "C" int mc_set_warm_start(mc_ctx *c, int mode)
{
    CHECK_CTX(c);
    if (mode != MC_WARM_NONE && mode != MC_WARM_NEWTON && mode != MC_WARM_MLP) {
        mc_set_error("mc_set_warm_start: mode must be MC_WARM_NONE, MC_WARM_NEWTON or MC_WARM_MLP");
        return -1;
    }
    c->warm_mode = mode;
    return 0;
}

extern "C" int mc_set_warm_levels(mc_ctx *c, int levels)
{
    CHECK_CTX(c);
    if (levels < 0 || levels > 24) {
        mc_set_error("mc_set_warm_levels: levels must be in 0..24");
        return -1;
    }
    c->warm_levels = levels;
    return 0;
}

extern "C" int mc_thermal_stats(mc_ctx *c, int64_t *out4)
{
    CHECK_CTX(c);
    if (!out4) {
        mc_set_error("mc_thermal_stats: null output");
        return -1;
    }
    for (int k = 0; k < 4; ++k) out4[k] = c->thermal_stats[k];
    return 0;
}

extern "C" int mc_diffuse_info(mc_ctx *c, double *lum, int64_t *nphot)
{
    CHECK_CTX(c);
    if (lum) *lum = c->diffuse_lum;
    if (nphot) *nphot = c->diffuse_nphot;
    return 0;
}

extern "C" int mc_diffuse_hei_setup(mc_ctx *c, int n, const double *y, const double *A)
{
    CHECK_CTX(c);
    if (!y || !A) {
        mc_set_error("mc_diffuse_hei_setup: null table");
        return -1;
    }
    std::vector<double> h(2 * MC_HEI2PH_NCDF);
    double nper = 0.0, eper = 0.0;
    const int rc = mc_hei2ph_tables(n, y, A, c->p.eion_max, h.data(), h.data() + MC_HEI2PH_NCDF,
                                    &nper, &eper);
    if (rc == -1) {
        mc_set_error("mc_diffuse_hei_setup: the two-photon table needs 4 to 128 rows, got " +
                     std::to_string(n));
        return -1;
    }
    if (rc == -2) {
        mc_set_error("mc_diffuse_hei_setup: the two-photon table's y must ascend");
        return -1;
    }
    if (rc) {
        mc_set_error("mc_diffuse_hei_setup: the two-photon table integrates to zero");
        return -1;
    }
    if (!c->d_hei && dalloc(&c->d_hei, h.size())) return -100;
    MC_HIP(hipMemcpy(c->d_hei, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
    c->hei_nper = nper;
    c->hei_eper = eper;
    c->hei_ready = true;
    DevMilne &M = c->milne;
    M.hei_y = c->d_hei;
    M.hei_cdf = c->d_hei + MC_HEI2PH_NCDF;
    M.hei_eper = eper;
    M.hei = c->p.diffuse_field && c->p.hei_diffuse;
    return 0;
}

extern "C" int mc_diffuse_hei_info(mc_ctx *c, double *n_per_decay, double *e_per_decay,
                                   double *grid_y, double *cdf_e)
{
    CHECK_CTX(c);
    if (!c->hei_ready) {
        mc_set_error("mc_diffuse_hei_info: no two-photon table (mc_diffuse_hei_setup)");
        return -1;
    }
    if (n_per_decay) *n_per_decay = c->hei_nper;
    if (e_per_decay) *e_per_decay = c->hei_eper;
    if (grid_y) MC_HIP(hipMemcpy(grid_y, c->d_hei, sizeof(double) * MC_HEI2PH_NCDF,
                                 hipMemcpyDeviceToHost));
    if (cdf_e) MC_HIP(hipMemcpy(cdf_e, c->d_hei + MC_HEI2PH_NCDF, sizeof(double) * MC_HEI2PH_NCDF,
                                hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int mc_hydrogen_balance_residual_grid(mc_ctx *c, double *eps_max, double *eps_vol)
{
    CHECK_CTX(c);
    if (!eps_max || !eps_vol) {
        mc_set_error("mc_hydrogen_balance_residual_grid: null output");
        return -1;
    }
    return mc_run_residual(c, eps_max, eps_vol);
}

extern "C" int mc_tallies_get(mc_ctx *c, double *hhe, double *jt, double *dust, int64_t *segs)
{
    CHECK_CTX(c);
    const size_t n = (size_t)c->nleaf;
    if (hhe || dust) {
        std::vector<double> h(n * 8);
        MC_HIP(hipMemcpy(h.data(), c->st.hhe, sizeof(double) * n * 8, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; ++i) {
            if (hhe)
                for (int k = 0; k < 6; ++k) hhe[i * 6 + k] = h[i * 8 + k];
            if (dust) dust[i] = h[i * 8 + 6];
        }
    }
    if (jt) MC_HIP(hipMemcpy(jt, c->st.jt, sizeof(double) * n * c->nnu, hipMemcpyDeviceToHost));
    if (segs) {
        segs[0] = c->segs[0];
        segs[1] = c->segs[1];
    }
    return 0;
}

extern "C" int mc_rates_get(mc_ctx *c, double *rates6)
{
    CHECK_CTX(c);
    if (!rates6) {
        mc_set_error("mc_rates_get: null output");
        return -1;
    }
    MC_HIP(hipMemcpy(rates6, c->st.rates, sizeof(double) * 6 * c->nleaf, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int mc_energy_ledger(mc_ctx *c, double *out3)
{
    CHECK_CTX(c);
    for (int k = 0; k < 3; ++k) out3[k] = c->ledger[k];
    return 0;
}

extern "C" int mc_info(mc_ctx *c, int32_t *nleaf, int32_t *nnu, double *Ltot, double *eedge)
{
    CHECK_CTX(c);
    if (nleaf) *nleaf = c->nleaf;
    if (nnu) *nnu = c->nnu;
    if (Ltot) *Ltot = c->Ltot;
    if (eedge)
        for (int i = 0; i <= c->nnu; ++i) eedge[i] = c->eedge[i];
    return 0;
}

extern "C" int mc_set_precision(mc_ctx *c, int precision)
{
    CHECK_CTX(c);
    if (precision != MC_FP64 && precision != MC_MIXED) {
        mc_set_error("precision must be MC_FP64 or MC_MIXED");
        return -1;
    }
    c->p.precision = precision;
    return 0;
}

extern "C" int mc_timing(mc_ctx *c, double *out4)
{
    CHECK_CTX(c);
    for (int k = 0; k < 4; ++k) out4[k] = c->times[k];
    return 0;
}

extern "C" int mc_trace_ray_noalloc(mc_ctx *c, int precision, const double *pos,
                                    const double *dir, int32_t *cells, int maxn, int32_t *n)
{
    CHECK_CTX(c);
    if (!pos || !dir || !cells || !n || maxn < 1) {
        mc_set_error("mc_trace_ray_noalloc: bad arguments");
        return -1;
    }
    if (maxn > MC_TRACE_NOALLOC_MAX) {
        mc_set_error("mc_trace_ray_noalloc: maxn " + std::to_string(maxn) + " exceeds the "
                     "device scratch of " + std::to_string(MC_TRACE_NOALLOC_MAX) + " cells");
        return -1;
    }
    return mc_run_trace_noalloc(c, precision, pos, dir, cells, maxn, n);
}

extern "C" int mc_qmc_uniforms(mc_ctx *c, int stream, int64_t first_index, int n, int ndim,
                               double *out)
{
    CHECK_CTX(c);
    if ((stream != 1 && stream != 2) || first_index < 0 || n < 1 || ndim < 1 ||
        ndim > MC_QMC_DIM || !out || (int64_t)n * ndim > (int64_t)1 << 26) {
        mc_set_error("mc_qmc_uniforms: stream 1 or 2, 1 <= ndim <= 12, 1 <= n, n ndim <= 2^26");
        return -1;
    }
    return mc_run_qmc_dump(c, stream, first_index, n, ndim, out);
}

extern "C" int mc_launch_records(mc_ctx *c, int64_t first_id, int n, double *rec,
                                 int32_t *bin_leaf)
{
    CHECK_CTX(c);
    if (first_id < 1 || n < 1 || n > c->pk.cap || !rec || !bin_leaf) {
        mc_set_error("mc_launch_records: need first_id >= 1, 1 <= n <= the launch batch, outputs");
        return -1;
    }
    if (!c->state_ready) {
        mc_set_error("mc_launch_records: gas state not set up");
        return -1;
    }
    if (c->p.ion_add_dust && c->dn < 2) {
        mc_set_error("ion_add_dust needs mc_ion_dust_table first");
        return -1;
    }
    return mc_run_launch_dump(c, first_id, n, rec, bin_leaf);
}

extern "C" int mc_trace_ray(mc_ctx *c, int precision, const double *pos, const double *dir,
                            int32_t *cells, int maxn, int32_t *n)
{
    CHECK_CTX(c);
    if (!pos || !dir || !cells || !n || maxn < 1) {
        mc_set_error("mc_trace_ray: bad arguments");
        return -1;
    }
    return mc_run_trace(c, precision, pos, dir, cells, maxn, n);
}

// ---- trace metals -----------------------------------------------------------
namespace {
// ci_dere_ratio (src/recomb_mod.f90:288-311): the Dere (2007) / Voronov
// ratio of stage `stage` (1 = X0 -> X+) of element Z, 1 where not tabulated
double dere_ratio(int Z, int stage)
{
    struct Row { int Z, n; double r[4]; };
    static const Row tab[] = {
        {1, 1, {0.9063}}, {2, 2, {1.0389, 1.0686}}, {6, 3, {1.0499, 0.913, 1.0377}},
        {7, 2, {1.0421, 1.1966}}, {8, 2, {1.041, 1.1181}}, {10, 2, {0.8089, 1.1395}},
        {12, 2, {0.3793, 0.9857}}, {14, 4, {0.7328, 0.8798, 0.4492, 0.8221}},
        {16, 3, {1.3572, 0.8925, 0.8119}}, {17, 4, {0.5412, 0.8428, 0.9237, 0.819}},
        {18, 3, {0.9242, 0.8644, 0.9752}}, {20, 4, {0.7652, 1.1668, 1.0422, 0.8705}},
        {26, 3, {0.9904, 1.0568, 1.824}}};
    for (const Row &row : tab)
        if (row.Z == Z) return stage >= 1 && stage <= row.n ? row.r[stage - 1] : 1.0;
    return 1.0;
}
}  // namespace

extern "C" int mc_species_setup(mc_ctx *c, int n, const mc_element *els)
{
    CHECK_CTX(c);
    if (n < 0 || n > MC_MAX_EL || (n > 0 && !els)) {
        mc_set_error("mc_species_setup: need 0..11 elements");
        return -1;
    }
    int ntr = 0, nst = 0;
    for (int ie = 0; ie < n; ++ie) {
        const mc_element &e = els[ie];
        if (e.nstage < 2 || e.nstage > MC_MAX_ST || !(e.abund > 0.0)) {
            mc_set_error("mc_species_setup: element needs 2..6 stages and abund > 0");
            return -1;
        }
        for (int it = 0; it < e.nstage - 1; ++it) {
            if ((e.photo_form[it] != 1 && e.photo_form[it] != 2) || e.ndr[it] < 0 ||
                e.ndr[it] > 9 || e.cxi_form[it] < 0 || e.cxi_form[it] > 7 || e.cxr_form[it] < 0 ||
                e.cxr_form[it] > 7) {
                mc_set_error("mc_species_setup: bad transition record");
                return -1;
            }
        }
        for (int i = 0; i < e.nstage; ++i)
            if (e.ncool[i] < 0 || e.ncool[i] > MC_MAX_COOLFIT) {
                mc_set_error("mc_species_setup: cooling fits have at most 8 terms");
                return -1;
            }
        ntr += e.nstage - 1;
        nst += e.nstage;
    }
    if (ntr > MC_GPU_MAXTR) {
        mc_set_error("mc_species_setup: " + std::to_string(ntr) +
                     " metal transitions; the device tables hold at most " +
                     std::to_string(MC_GPU_MAXTR) +
                     " (the reference's eleven elements have 32)");
        return -1;
    }
    DevMetalTab &m = c->mtab;
    std::memset(&m, 0, sizeof(m));
    m.nel = n; m.ntr = ntr; m.nst = nst;
    int t = 0, st = 0;
    for (int ie = 0; ie < n; ++ie) {
        const mc_element &e = els[ie];
        const int t0 = t;
        const bool isC = std::strcmp(e.name, "c") == 0, isO = std::strcmp(e.name, "o") == 0;
        for (int i = 0; i < e.nstage; ++i) {
            m.s_ncool[st + i] = e.ncool[i];
            m.s_charge[st + i] = i;
            for (int k = 0; k < e.ncool[i]; ++k) {
                m.s_A[st + i][k] = e.cool_A[i][k];
                m.s_T[st + i][k] = e.cool_T[i][k];
            }
        }
        for (int it = 0; it < e.nstage - 1; ++it, ++t) {
            m.t_first[t] = t0;
            m.t_last[t] = it == e.nstage - 2;
            m.t_it[t] = it;
            m.t_st[t] = st + it;
            m.t_cII[t] = isC && it == 0;
            m.t_oI[t] = isO && it == 0;
            m.t_ab[t] = e.abund;
            m.eth[t] = e.eth[it];
            m.photo_form[t] = e.photo_form[it];
            for (int k = 0; k < 8; ++k) m.photo[t][k] = e.photo[it][k];
            for (int k = 0; k < 5; ++k) m.ci[t][k] = e.ci[it][k];
            m.ci_ratio[t] = c->p.ci_model == MC_CI_DERE_HYBRID ? dere_ratio(e.Z, it + 1) : 1.0;
            m.has_rr2[t] = e.has_rr2[it];
            for (int k = 0; k < 6; ++k) m.rr[t][k] = e.rr[it][k];
            for (int k = 0; k < 2; ++k) m.rr2[t][k] = e.rr2[it][k];
            m.ndr[t] = e.ndr[it];
            for (int k = 0; k < 9; ++k) { m.dr_c[t][k] = e.dr_c[it][k]; m.dr_e[t][k] = e.dr_e[it][k]; }
            m.has_dr2[t] = e.has_dr2[it];
            for (int k = 0; k < 4; ++k) m.dr2[t][k] = e.dr2[it][k];
            m.cxi_form[t] = e.cxi_form[it];
            m.cxr_form[t] = e.cxr_form[it];
            for (int k = 0; k < 6; ++k) { m.cxi[t][k] = e.cxi[it][k]; m.cxr[t][k] = e.cxr[it][k]; }
        }
        st += e.nstage;
    }
    // threshold order: stable insertion sort (src/species_mod.f90:205-226)
    for (int i = 0; i < ntr; ++i) { m.flat_t[i] = i; m.flat_eth[i] = m.eth[i]; }
    for (int i = 1; i < ntr; ++i) {
        const int ti = m.flat_t[i];
        const double e = m.flat_eth[i];
        int j = i - 1;
        while (j >= 0 && m.flat_eth[j] > e) {
            m.flat_t[j + 1] = m.flat_t[j];
            m.flat_eth[j + 1] = m.flat_eth[j];
            --j;
        }
        m.flat_t[j + 1] = ti;
        m.flat_eth[j + 1] = e;
    }
    for (int k = 0; k < ntr; ++k) m.t_flat[m.flat_t[k]] = k;
    for (int ie = 0; ie < n; ++ie) m.el_ab[ie] = els[ie].abund;
    // compiled-in element layouts of the cell-solve kernels
    static const char *kCNONeSAr[] = {"c", "n", "o", "ne", "s", "ar"};
    auto layout_is = [&](const char *const *names, int ne_, const int *nst) {
        if (n != ne_) return false;
        for (int ie = 0; ie < n; ++ie)
            if (std::strcmp(els[ie].name, names[ie]) != 0 || els[ie].nstage != nst[ie]) return false;
        return true;
    };
    static const int kStages[] = {4, 3, 3, 3, 4, 4};
    c->metal_layout = layout_is(kCNONeSAr, 6, kStages)   ? MC_LAYOUT_CNONeSAr
                      : layout_is(kCNONeSAr, 5, kStages) ? MC_LAYOUT_CNONeS
                                                         : MC_LAYOUT_GENERIC;
    c->elements.assign(els, els + n);
    c->nel = n; c->ntr = ntr; c->nst = nst;
    mc_mom_free(c);            // the moment tables follow the registry
    if (alloc_metal_arrays(c)) return -100;
    if (c->state_ready) return mc_run_stage_refresh(c);
    return 0;
}

// element ie's offsets: first transition and first stage (element order)
static int element_offsets(mc_ctx *c, int ie, int &t0, int &s0)
{
    if (ie < 0 || ie >= c->nel) {
        mc_set_error("element index out of range");
        return -1;
    }
    t0 = 0; s0 = 0;
    for (int k = 0; k < ie; ++k) {
        t0 += c->elements[k].nstage - 1;
        s0 += c->elements[k].nstage;
    }
    return 0;
}

extern "C" int mc_species_fractions_get(mc_ctx *c, int ie, double *frac)
{
    CHECK_CTX(c);
    int t0, s0;
    if (element_offsets(c, ie, t0, s0) || !frac) return -1;
    const int ns = c->elements[ie].nstage, nl = c->nleaf;
    std::vector<double> h((size_t)nl * c->nst);
    MC_HIP(hipMemcpy(h.data(), c->d_mfrac, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
    for (int il = 0; il < nl; ++il)
        for (int i = 0; i < ns; ++i) frac[(size_t)il * ns + i] = h[(size_t)il * c->nst + s0 + i];
    return 0;
}

extern "C" int mc_species_rates_get(mc_ctx *c, int ie, double *gam, double *heat)
{
    CHECK_CTX(c);
    int t0, s0;
    if (element_offsets(c, ie, t0, s0)) return -1;
    const int nt = c->elements[ie].nstage - 1, nl = c->nleaf;
    std::vector<double> h((size_t)nl * c->ntr);
    for (int q = 0; q < 2; ++q) {
        double *dst = q ? heat : gam;
        if (!dst) continue;
        MC_HIP(hipMemcpy(h.data(), q ? c->d_mheat : c->d_mgam, sizeof(double) * h.size(),
                         hipMemcpyDeviceToHost));
        for (int il = 0; il < nl; ++il)
            for (int it = 0; it < nt; ++it)
                dst[(size_t)il * nt + it] = h[(size_t)il * c->ntr + t0 + it];
    }
    return 0;
}

extern "C" int mc_species_tallies_get(mc_ctx *c, int ie, double *raw)
{
    CHECK_CTX(c);
    int t0, s0;
    if (element_offsets(c, ie, t0, s0) || !raw) return -1;
    const int nt = c->elements[ie].nstage - 1, nl = c->nleaf, ntr = c->ntr;
    std::vector<double> h((size_t)nl * ntr * 2);
    MC_HIP(hipMemcpy(h.data(), c->d_mraw, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
    for (int il = 0; il < nl; ++il)
        for (int it = 0; it < nt; ++it) {
            const int k = c->mtab.t_flat[t0 + it];
            for (int q = 0; q < 2; ++q)
                raw[((size_t)il * nt + it) * 2 + q] = h[((size_t)il * ntr + k) * 2 + q];
        }
    return 0;
}

extern "C" int mc_walk_atomics(mc_ctx *c, int64_t *n)
{
    CHECK_CTX(c);
    if (!n) {
        mc_set_error("mc_walk_atomics: null output");
        return -1;
    }
    *n = c->segs[3];
    return 0;
}

extern "C" int mc_metal_segment_transitions(mc_ctx *c, int64_t *sum_ntr)
{
    CHECK_CTX(c);
    if (!sum_ntr) {
        mc_set_error("mc_metal_segment_transitions: null output");
        return -1;
    }
    *sum_ntr = c->segs[2];
    return 0;
}
