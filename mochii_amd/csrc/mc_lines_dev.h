// Line luminosities on the converged state (lines_write,
// src/lines_mod.f90:37-352): per leaf the recombination lines from the
// (T, n_e) tables (sh95_emis, src/sh95_mod.f90:146-176) and every metal
// ion's n-level emissivities (nlevel_emissivities, src/nlevel_mod.f90:
// 193-253) times the stage density and leaf volume, summed over leaves.
// One thread per leaf; each line's sum is reduced over the wave, then over
// the block in LDS, and added to the device total with one atomic per block.
// Included by mc_kernels.hip (it reads the registry in c_mt).
#pragma once

// Upsilon(T) of collision row k (ups_eval, src/nlevel_mod.f90:128-160)
__device__ double ln_ups(const DevLineIon &a, int k, double T)
{
#pragma clang fp contract(off)
    const double et = (1.380649e-16 / 2.1798723611035e-11) * T / a.dery[k];
    double st;
    if (a.ttype[k] == 1 || a.ttype[k] == 4)
        st = 1.0 - log(a.cups[k]) / log(et + a.cups[k]);
    else
        st = et / (et + a.cups[k]);
    st = fmin(fmax(st, a.st_lo[k]), a.st_hi[k]);
    const double u = a.st_hi[k] > a.st_lo[k]
                         ? 2.0 * (st - a.st_lo[k]) / (a.st_hi[k] - a.st_lo[k]) - 1.0
                         : 0.0;
    double b0 = 0.0, b1 = 0.0;
    for (int j = a.nc[k] - 1; j >= 1; --j) {
        const double b2 = b1;
        b1 = b0;
        b0 = 2.0 * u * b1 - b2 + a.coef[k][j];
    }
    double y = u * b0 - b1 + a.coef[k][0];
    if (a.logf[k] == 1) y = pow(10.0, y);
    double ups;
    switch (a.ttype[k]) {
    case 1: ups = y * log(et + 2.718281828459045); break;
    case 2: ups = y; break;
    case 3: ups = y / (et + 1.0); break;
    case 4: ups = y * log(et + a.cups[k]); break;
    case 5: ups = y / et; break;
    default: ups = pow(10.0, y); break;
    }
    return fmax(ups, 0.0);
}

// level populations at (T, n_e): the rate matrix, the closure row and
// Gaussian elimination with partial pivoting (src/nlevel_mod.f90:200-246)
template <int NL>
__device__ void ln_populations(const DevLineIon &a, double T, double ne, double *pop)
{
#pragma clang fp contract(off)
    constexpr int W = NL + 1;
    const int n = a.nlev;
    double R[NL * NL], M[NL * W];
    for (int i = 0; i < NL * NL; ++i) R[i] = 0.0;
    const double sqT = sqrt(T);
    for (int k = 0; k < a.nups; ++k) {
        const int l = a.clo[k], u = a.chi[k];
        const double q_ul = 8.629e-6 / (a.gu[k] * sqT) * ln_ups(a, k, T);
        const double q_lu = a.gr[k] * q_ul * exp(-a.cde[k] / (1.380649e-16 * T));
        R[u * NL + l] = R[u * NL + l] + ne * q_lu;
        R[l * NL + u] = R[l * NL + u] + ne * q_ul;
    }
    for (int k = 0; k < a.nrad; ++k) R[a.rlo[k] * NL + a.rhi[k]] = R[a.rlo[k] * NL + a.rhi[k]] + a.A[k];
    for (int i = 0; i < n; ++i)
        for (int j = 0; j <= n; ++j) M[i * W + j] = 0.0;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            if (i != j) {
                M[i * W + j] = R[i * NL + j];
                M[i * W + i] = M[i * W + i] - R[j * NL + i];
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

// the block's sum of v into the device total *dst: wave sum, LDS across
// the block's waves, one atomic (every thread of the block calls it)
__device__ __forceinline__ void ln_block_add(double v, double *dst, double *lds)
{
    v = wave_sum(v);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) lds[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int k = 0; k < kBlock / 64; ++k) s += lds[k];
        if (s != 0.0) unsafeAtomicAdd(dst, s);
    }
    __syncthreads();
}

__device__ __forceinline__ double ln_leaf_vol(const DevGrid &g, int il, double d2cm)
{
    const double w = 2.0 * leaf_half(g, il) * d2cm;
    return (w * w) * w;
}

// recombination lines: table tab weighted by n_e n_ion with
// kind 0: n_ion = n_H (1 - x_HI); 1: y n_H x_HeIII (skip x_HeIII <= 1e-30);
// 2: y n_H x_HeII (skip x_HeII <= 1e-30)
__global__ __launch_bounds__(kBlock) void k_lines_rec(DevGrid g, DevState s, const DevLineTab *tab,
                                                      int kind, double y, double d2cm, double *Lout)
{
#pragma clang fp contract(off)
    __shared__ double lds[kBlock / 64];
    const int il = blockIdx.x * kBlock + threadIdx.x;
    const DevLineTab &t = *tab;
    bool on = il < g.nleaf;
    double T = 0.0, ne = 0.0, w = 0.0;
    if (on) {
        const double nH = s.nH[il];
        on = nH > 0.0;
        if (on) {
            T = s.Te[il];
            ne = s.ne[il];
            double nion;
            // the ion's fraction: 1 - x_HI, x_HeIII or x_HeII
            if (kind == 0) {
                nion = 1.0 - s.xHI[il];
            } else if (kind == 1) {
                nion = fmax(0.0, 1.0 - s.xHeI[il] - s.xHeII[il]);
                on = nion > 1.0e-30;
            } else {
                nion = s.xHeII[il];
                on = nion > 1.0e-30;
            }
            // the reference's grouping: ne nH nion-fraction emis, then x vol
            w = kind == 0 ? (ne * nH) * nion : ((ne * nH) * y) * nion;
        }
    }
    int iT = 0, iD = 0;
    double wT = 0.0, wD = 0.0;
    if (on) {
        const double lt = fmin(fmax(T, t.tgrid[0]), t.tgrid[t.nT - 1]);
        const double ld = fmin(fmax(ne, t.dgrid[0]), t.dgrid[t.nD - 1]);
        while (iT < t.nT - 2 && !(lt < t.tgrid[iT + 1])) ++iT;
        while (iD < t.nD - 2 && !(ld < t.dgrid[iD + 1])) ++iD;
        wT = log(lt / t.tgrid[iT]) / log(t.tgrid[iT + 1] / t.tgrid[iT]);
        wD = log(ld / t.dgrid[iD]) / log(t.dgrid[iD + 1] / t.dgrid[iD]);
    }
    const double vol = on ? ln_leaf_vol(g, il, d2cm) : 0.0;
    for (int k = 0; k < t.nlin; ++k) {
        double v = 0.0;
        if (on) {
            const double *e = t.etab + (size_t)k * MC_LT_MAXG * MC_LT_MAXG;
            const double em = exp(log(e[iT * MC_LT_MAXG + iD]) * (1.0 - wT) * (1.0 - wD) +
                                  log(e[(iT + 1) * MC_LT_MAXG + iD]) * wT * (1.0 - wD) +
                                  log(e[iT * MC_LT_MAXG + iD + 1]) * (1.0 - wT) * wD +
                                  log(e[(iT + 1) * MC_LT_MAXG + iD + 1]) * wT * wD);
            v = (w * em) * vol;
        }
        ln_block_add(v, Lout + t.out0 + k, lds);
    }
}

// metal lines: the stage fractions of the leaf (species_fractions at
// (T_e, n_e, n_HI, n_H - n_HI) with the last metal rates), then for every
// ion its populations and rows; NL bounds the ions' level counts
template <int NL>
__global__ __launch_bounds__(kBlock) void k_lines_metal(DevGrid g, DevState s, MetalSw ms,
                                                        const DevLineIon *ions, int nion,
                                                        double d2cm, double *Lout)
{
#pragma clang fp contract(off)
    __shared__ double lds[kBlock / 64];
    const int il = blockIdx.x * kBlock + threadIdx.x;
    bool on = il < g.nleaf && s.nH[il] > 0.0;
    double lo[MC_GPU_MAXTR], up[MC_GPU_MAXTR];
    double T = 0.0, ne = 0.0, nH = 0.0, vol = 0.0;
    if (on) {
        T = s.Te[il];
        ne = s.ne[il];
        nH = s.nH[il];
        const double nHI = nH * s.xHI[il], nHII = nH - nHI;
        MCache<MC_GPU_MAXTR> cc;
        m_prepare<MC_GPU_MAXTR>((&c_mt), ms.gam + (size_t)il * ms.ntr, T, cc);
        m_fracs<MC_GPU_MAXTR>((&c_mt), cc, ne, nHI, nHII, lo, up);
        vol = ln_leaf_vol(g, il, d2cm);
    } else {
        for (int t = 0; t < MC_GPU_MAXTR; ++t) { lo[t] = 0.0; up[t] = 0.0; }
    }
    double pop[NL];
    for (int q = 0; q < nion; ++q) {
        const DevLineIon &a = ions[q];
        double frac = 0.0;
        for (int t = 0; t < MC_GPU_MAXTR; ++t)
            if (t == a.t) frac = a.up ? up[t] : lo[t];
        const double ni = a.abund * nH * frac;
        const bool lit = on && ni > 1.0e-30;
        if (lit) ln_populations<NL>(a, T, ne, pop);
        for (int k = 0; k < a.nrad; ++k) {
            const double v = lit ? ((pop[a.rhi[k]] * a.A[k] * a.rde[k]) * ni) * vol : 0.0;
            ln_block_add(v, Lout + a.out0 + k, lds);
        }
    }
}
